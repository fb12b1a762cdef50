"""alpacc_hip_lex_device_noalloc (the lexer without allocation, scratch in
a static device array) against alpacc_hip_lex_device on the same inputs:
tokens, count and verdict identical, on consecutive calls with the second
input smaller than the first (the scratch array keeps the first call's
contents), in both terminal numberings, a rejected input, and the size
limits refused before any GPU work."""

import ctypes as C

import numpy as np
import pytest

from alpacc_amd import datagen

pytestmark = pytest.mark.gpu


def noalloc_fn(lib):
    fn = lib.alpacc_hip_lex_device_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                   C.POINTER(C.c_uint64), C.c_void_p]
    return fn


def dev_bytes(ptr, nbytes):
    from alpacc_amd.hip import _take_dev
    return _take_dev(ptr, nbytes)


def run_both(be, fn, mode, data: bytes):
    import torch
    from alpacc_amd import hip
    lay = be.layouts[mode]
    n = len(data)
    d_in = torch.frombuffer(bytearray(data + b"\0"), dtype=torch.uint8).to("cuda")
    torch.cuda.synchronize()
    # the allocating entry point: results in the context's buffers
    ok_ref, out = be.lex_device(d_in.data_ptr(), n, mode)
    T = out.count
    ref = (dev_bytes(out.terminals, T * lay.terminal), dev_bytes(out.starts, T * 8),
           dev_bytes(out.lengths, T * lay.length)) if ok_ref else None
    # the allocation-free one: outputs in caller buffers, filled with junk first
    cap = max(n, 1)
    t = torch.full((cap * lay.terminal,), 0xAB, dtype=torch.uint8, device="cuda")
    s = torch.full((cap,), -7, dtype=torch.int64, device="cuda")
    ln = torch.full((cap * lay.length,), 0xCD, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    cnt = C.c_uint64()
    rc = fn(be.h, hip.MODES[mode], C.c_void_p(d_in.data_ptr()), n, C.c_void_p(t.data_ptr()),
            C.c_void_p(s.data_ptr()), C.c_void_p(ln.data_ptr()), cap, C.byref(cnt), None)
    assert rc >= 0, hip.load_library().alpacc_hip_last_error()
    assert (rc == 1) == ok_ref
    if not ok_ref:
        return
    assert cnt.value == T
    got = (t[:T * lay.terminal].cpu().numpy().tobytes(), s[:T].cpu().numpy().tobytes(),
           ln[:T * lay.length].cpu().numpy().tobytes())
    assert got == ref


def test_noalloc_matches_lex_device():
    from alpacc_amd import hip
    fn = noalloc_fn(hip.load_library())
    with hip.Backend("json") as be:
        big = datagen.generate("json", 8 << 20, seed=21, conforming=True).tobytes()
        small = datagen.generate("json", 100003, seed=22, conforming=True).tobytes()
        for mode in ("lexer", "both"):
            print(f"case noalloc json {mode}: 8 MiB then 100003 B", flush=True)
            run_both(be, fn, mode, big)
            run_both(be, fn, mode, small)
        run_both(be, fn, "lexer", small[:5000] + b"\x01" + small[5000:6000])  # a dead byte: rejected
        # the largest input it takes, then a tiny one, in the other lexer shape
        top = datagen.generate("json", 64 << 20, seed=23, conforming=True).tobytes()
        be.set_lex_shape(64, 0, 16)
        run_both(be, fn, "both", top)
        run_both(be, fn, "both", b"[1, 2]")


def test_noalloc_limits():
    from alpacc_amd import hip
    lib = hip.load_library()
    fn = noalloc_fn(lib)
    with hip.Backend("json") as be:
        cnt = C.c_uint64()
        assert fn(be.h, 0, C.c_void_p(16), (64 << 20) + 1, C.c_void_p(16), C.c_void_p(16), C.c_void_p(16),
                  (64 << 20) + 1, C.byref(cnt), None) == -4
        assert fn(be.h, 0, C.c_void_p(16), 100, C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), 99,
                  C.byref(cnt), None) == -1
        assert fn(be.h, 0, None, 0, None, None, None, 0, C.byref(cnt), None) == 1 and cnt.value == 0
