"""In-process access to the CPU oracle (test infrastructure only).

Loads oracle/build/lib<grammar>_<mode>.so (built by oracle/build.py) and
returns wire response records, so a test compares the HIP path and the oracle
byte for byte.  Only tests/, the smoke check and the benchmark's CPU-baseline
leg use this module.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "build")
_libs = {}


def _sfx(index32: bool, length: int) -> str:
    return ("_i32" if index32 else "") + (f"_l{length}" if length else "")


def ensure_built(grammar: str, index32: bool = False, length: int = 0, mode: str = None) -> None:
    sfx = _sfx(index32, length)
    modes = (mode,) if mode else (("lexer", "both") if sfx else ("lexer", "parser", "both"))
    if all(os.path.exists(os.path.join(BUILD, f"lib{grammar}_{m}{sfx}.so")) for m in modes):
        return
    if sfx:
        sys.path.insert(0, HERE)
        import build as oracle_build
        oracle_build.build_variant(grammar, index32, length)
    else:
        subprocess.check_call([sys.executable, os.path.join(HERE, "build.py"), grammar])


def _lib(grammar: str, mode: str, index32: bool = False, length: int = 0):
    key = (grammar, mode, index32, length)
    if key not in _libs:
        ensure_built(grammar, index32, length, mode)
        lib = C.CDLL(os.path.join(BUILD, f"lib{grammar}_{mode}{_sfx(index32, length)}.so"))
        lib.oracle_response.restype = C.c_int
        lib.oracle_response.argtypes = [C.c_uint64, C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]
        lib.oracle_free.argtypes = [C.c_void_p]
        lib.oracle_time_case.restype = C.c_int64
        lib.oracle_time_case.argtypes = [C.c_uint64, C.c_void_p, C.POINTER(C.c_double)]
        _libs[key] = lib
    return _libs[key]


def record(grammar: str, mode: str, payload: bytes, n: int = None, index32: bool = False, length: int = 0) -> bytes:
    """The oracle's response record for one request (of a width variant:
    index32, length bits)."""
    lib = _lib(grammar, mode, index32, length)
    if n is None:
        n = len(payload)
    rec = C.c_void_p()
    ln = C.c_uint64()
    buf = C.create_string_buffer(payload, len(payload) + 1)
    if lib.oracle_response(n, buf, C.byref(rec), C.byref(ln)) != 0:
        raise RuntimeError("oracle_response failed")
    # a buffer copy: string_at takes a C int size (records can pass 2 GiB)
    out = bytes((C.c_char * ln.value).from_address(rec.value)) if ln.value else b""
    lib.oracle_free(rec)
    return out


def time_case(grammar: str, mode: str, payload, n: int) -> float:
    """Seconds the sequential oracle spends on one request (output to
    /dev/null).  ``payload`` is any object exposing a buffer address via
    numpy (``.ctypes.data``) or bytes."""
    lib = _lib(grammar, mode)
    secs = C.c_double()
    if isinstance(payload, (bytes, bytearray)):
        buf = C.create_string_buffer(bytes(payload), len(payload) + 1)
        ptr = C.cast(buf, C.c_void_p)
    else:
        ptr = C.c_void_p(payload.ctypes.data)
    if lib.oracle_time_case(n, ptr, C.byref(secs)) != 0:
        raise RuntimeError("oracle_time_case failed")
    return secs.value


def program(grammar: str, mode: str) -> str:
    ensure_built(grammar, mode=mode)
    return os.path.join(BUILD, f"{grammar}_{mode}")
