"""Python mirror of the generated program's entry points over the HIP library.

``Backend(grammar)`` keeps one grammar resident on a GPU and exposes the
reference's operations (backends/c/lexer.c ``lex_string``, backends/c/parser.c
``parse_test`` / ``compute_parents`` / ``compact_tree``, and the generated
``run_case`` / ``print_layout``) with the reference's return convention:
``True`` = accepted, ``False`` = rejected, an exception for errors.

The library is alpacc_amd/lib/libalpacc_hip.so, built in-tree by
``python -m alpacc_amd.build``; there is no fallback, so a missing library is
an ImportError.
"""

from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import wire

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libalpacc_hip.so")

LEXER, PARSER, BOTH = 0, 1, 2
MODES = {"lexer": LEXER, "parser": PARSER, "both": BOTH}


class Lexemes(C.Structure):
    _fields_ = [("count", C.c_uint64), ("terminals", C.c_void_p), ("starts", C.c_void_p), ("lengths", C.c_void_p)]


class Tree(C.Structure):
    _fields_ = [("tokens", Lexemes), ("num_nodes", C.c_uint64), ("productions", C.c_void_p),
                ("parents", C.c_void_p), ("token_parents", C.c_void_p)]


class Batch(C.Structure):
    _fields_ = [("count", C.c_uint64), ("valid", C.c_void_p), ("token_offsets", C.c_void_p),
                ("node_offsets", C.c_void_p), ("node_counts", C.c_void_p), ("token_parent_offsets", C.c_void_p),
                ("tokens", Lexemes), ("productions", C.c_void_p), ("parents", C.c_void_p),
                ("token_parents", C.c_void_p)]


class LexShape(C.Structure):
    _fields_ = [("items_per_thread", C.c_int), ("workgroups_per_cu", C.c_int), ("lookback_window", C.c_int),
                ("ms", C.c_double)]


class ParseShape(C.Structure):
    _fields_ = [("positions_per_tile", C.c_int), ("positions_per_lane", C.c_int), ("ms", C.c_double)]


PARSE_SHAPES = [(1024, 4), (512, 4), (2048, 4), (1024, 2), (1024, 8)]


class HipError(RuntimeError):
    pass


_lib = None


def load_library(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: run `python -m alpacc_amd.build` first")
    # PyTorch ships its own HIP runtime under the same soname as the system
    # one, and a process keeps whichever is loaded first.  Loading torch's
    # first (when torch is installed) gives the library and torch one shared
    # runtime, so device pointers and streams pass between them and torch
    # still finds the GPU after the library has opened it.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    vp, u64, i32 = C.c_void_p, C.c_uint64, C.c_int
    sig = {
        "alpacc_hip_open": (i32, [C.c_char_p, i32, C.POINTER(vp)]),
        "alpacc_hip_close": (None, [vp]),
        "alpacc_hip_last_error": (C.c_char_p, []),
        "alpacc_hip_grammars": (C.POINTER(C.c_char_p), []),
        "alpacc_hip_layout": (i32, [vp, i32, C.c_char_p, C.c_size_t]),
        "alpacc_hip_payload_bytes": (u64, [vp, i32, u64]),
        "alpacc_hip_lex_string": (i32, [vp, i32, vp, u64, C.POINTER(Lexemes)]),
        "alpacc_hip_parse_test": (i32, [vp, vp, u64, C.POINTER(vp), C.POINTER(u64)]),
        "alpacc_hip_lex_parse": (i32, [vp, vp, u64, C.POINTER(Tree)]),
        "alpacc_hip_run_case": (i32, [vp, i32, u64, vp, C.POINTER(vp), C.POINTER(u64)]),
        "alpacc_hip_free": (None, [vp]),
        "alpacc_hip_lex_device": (i32, [vp, i32, vp, u64, C.POINTER(Lexemes), vp]),
        "alpacc_hip_parse_device": (i32, [vp, vp, u64, C.POINTER(vp), C.POINTER(u64), vp]),
        "alpacc_hip_lex_parse_device": (i32, [vp, vp, u64, C.POINTER(Tree), vp]),
        "alpacc_hip_grammar_layout": (i32, [C.c_char_p, i32, C.c_char_p, C.c_size_t]),
        "alpacc_hip_grammar_payload_bytes": (u64, [C.c_char_p, i32, u64]),
        "alpacc_hip_set_timing": (i32, [vp, i32]),
        "alpacc_hip_timing": (i32, [vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(u64)]),
        "alpacc_hip_timed_kernels": (C.c_char_p, [vp]),
        "alpacc_hip_run_batch": (i32, [vp, i32, vp, C.POINTER(u64), u64, C.POINTER(vp), C.POINTER(u64)]),
        "alpacc_hip_lex_parse_batch_device": (i32, [vp, vp, C.POINTER(u64), u64, C.POINTER(Batch), vp]),
        "alpacc_hip_server_open": (i32, [C.POINTER(C.c_char_p), i32, i32, i32, u64, u64, C.POINTER(vp)]),
        "alpacc_hip_server_close": (None, [vp]),
        "alpacc_hip_server_submit": (i32, [vp, i32, i32, vp, u64, C.POINTER(u64)]),
        "alpacc_hip_server_wait": (i32, [vp, u64, C.POINTER(vp), C.POINTER(u64)]),
        "alpacc_hip_server_replay": (i32, [vp, u64, C.POINTER(C.c_int32), C.POINTER(C.c_int32), vp,
                                           C.POINTER(u64), C.POINTER(u64), C.POINTER(u64), C.POINTER(vp),
                                           C.POINTER(u64)]),
        "alpacc_hip_server_stats": (i32, [vp, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64)]),
        "alpacc_hip_set_graphs": (i32, [vp, i32]),
        "alpacc_hip_set_bracket_width": (i32, [vp, i32]),
        "alpacc_hip_graph_stats": (i32, [vp, C.POINTER(u64), C.POINTER(u64)]),
        "alpacc_hip_server_set_graphs": (i32, [vp, i32]),
        "alpacc_hip_set_lex_shape": (i32, [vp, C.POINTER(LexShape)]),
        "alpacc_hip_lex_stream_begin": (i32, [vp, i32]),
        "alpacc_hip_lex_stream_push": (i32, [vp, vp, u64, i32, C.POINTER(Lexemes), vp]),
        "alpacc_hip_open_variant": (i32, [C.c_char_p, i32, i32, i32, C.POINTER(vp)]),
        "alpacc_hip_variant_layout": (i32, [C.c_char_p, i32, i32, i32, C.c_char_p, C.c_size_t]),
        "alpacc_hip_server_open_variant": (i32, [C.POINTER(C.c_char_p), i32, i32, i32, u64, u64, i32, i32,
                                                 C.POINTER(vp)]),
        "alpacc_hip_get_lex_shape": (i32, [vp, C.POINTER(LexShape)]),
        "alpacc_hip_autotune_lexer": (i32, [vp, i32, vp, u64, i32, C.POINTER(LexShape)]),
        "alpacc_hip_load_grammar": (i32, [C.c_char_p, C.c_char_p, C.c_size_t]),
        "alpacc_hip_set_parse_shape": (i32, [vp, C.POINTER(ParseShape)]),
        "alpacc_hip_get_parse_shape": (i32, [vp, C.POINTER(ParseShape)]),
        "alpacc_hip_autotune_parser": (i32, [vp, vp, u64, i32, C.POINTER(ParseShape)]),
        "alpacc_hip_grammar_limits": (i32, [C.c_char_p, C.POINTER(u64), C.POINTER(u64)]),
        "alpacc_hip_name_count": (i32, [C.c_char_p, i32]),
        "alpacc_hip_name": (i32, [C.c_char_p, i32, u64, C.c_char_p, C.c_size_t]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    lib = load_library()
    return [n for n in dir(lib) if n.startswith("alpacc_hip_")]


def grammars():
    lib = load_library()
    arr = lib.alpacc_hip_grammars()
    out = []
    i = 0
    while arr[i]:
        out.append(arr[i].decode())
        i += 1
    return out


def load_grammar(path: str) -> str:
    """Register a grammar's tables file (python -m alpacc_amd.gen hip) with
    the library; returns its name."""
    lib = load_library()
    buf = C.create_string_buffer(256)
    _check(min(lib.alpacc_hip_load_grammar(path.encode(), buf, 256), 0))
    return buf.value.decode()


NAME_KINDS = {"terminal_lexer": 0, "terminal": 1, "production": 2}


def names(grammar: str, kind: str):
    """Names of every id of a kind: terminal_lexer (lexer-only numbering),
    terminal (parser / combined numbering) or production."""
    lib = load_library()
    n = lib.alpacc_hip_name_count(grammar.encode(), NAME_KINDS[kind])
    _check(min(n, 0))
    buf = C.create_string_buffer(512)
    out = []
    for i in range(n):
        _check(min(lib.alpacc_hip_name(grammar.encode(), NAME_KINDS[kind], i, buf, 512), 0))
        out.append(buf.value.decode())
    return out


def limits(grammar: str):
    """(input bytes one lexer launch takes, framed positions one parser
    launch takes) for a grammar."""
    lib = load_library()
    a, b = C.c_uint64(), C.c_uint64()
    _check(min(lib.alpacc_hip_grammar_limits(grammar.encode(), C.byref(a), C.byref(b)), 0))
    return a.value, b.value


def _check(rc: int) -> bool:
    if rc < 0:
        raise HipError(f"alpacc_hip error {rc}: {load_library().alpacc_hip_last_error().decode()}")
    return rc == 1


def _take(ptr, nbytes: int, dtype) -> np.ndarray:
    if nbytes == 0:
        if ptr:
            load_library().alpacc_hip_free(ptr)
        return np.zeros(0, dtype=dtype)
    buf = (C.c_uint8 * nbytes).from_address(ptr)
    arr = np.frombuffer(bytes(buf), dtype=dtype).copy()
    load_library().alpacc_hip_free(ptr)
    return arr


class _DevBytes:
    """A device byte range as a __cuda_array_interface__ object."""

    def __init__(self, ptr: int, n: int):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": "|u1", "data": (ptr, False), "version": 2,
                                         "strides": None}


def _take_dev(ptr: int, nbytes: int) -> bytes:
    """Bytes of device memory at ptr (copied to the host)."""
    import torch
    if nbytes == 0:
        return b""
    return torch.as_tensor(_DevBytes(ptr, nbytes), device="cuda").cpu().numpy().tobytes()


class Backend:
    """One grammar resident on one device."""

    def __init__(self, grammar: str, device: int = 0, index32: bool = False, length: int = 0):
        """index32 / length (bits): the width variant of alpacc_hip_open_variant."""
        self.lib = load_library()
        self.grammar = grammar
        h = C.c_void_p()
        if index32 or length:
            _check(self.lib.alpacc_hip_open_variant(grammar.encode(), device, int(index32), int(length), C.byref(h)))
        else:
            _check(self.lib.alpacc_hip_open(grammar.encode(), device, C.byref(h)))
        self.h = h
        # the modes the grammar has (a lexer-only grammar: the lexer alone)
        self.layouts = {}
        for m in MODES:
            buf = C.create_string_buffer(512)
            if self.lib.alpacc_hip_layout(self.h, MODES[m], buf, 512) >= 0:
                self.layouts[m] = self._layout(m)

    def close(self):
        if self.h:
            self.lib.alpacc_hip_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- layout ----------------------------------------------------------------
    def layout_text(self, mode: str) -> str:
        buf = C.create_string_buffer(512)
        _check(self.lib.alpacc_hip_layout(self.h, MODES[mode], buf, 512) >= 0 and 1 or -1)
        return buf.value.decode()

    def _layout(self, mode: str) -> wire.Layout:
        kv = dict(l.split("=") for l in self.layout_text(mode).strip().splitlines())
        return wire.Layout(mode, int(kv["terminal_t"]), int(kv["index_t"]), int(kv.get("length_t", 0)),
                           int(kv.get("production_t", 0)))

    # -- reference entry points --------------------------------------------------
    def run_case(self, mode: str, payload: bytes, n: Optional[int] = None) -> bytes:
        """run_case: one request -> one wire response record."""
        lay = self.layouts[mode]
        if n is None:
            n = len(payload) // (lay.terminal if mode == "parser" else 1)
        rec = C.c_void_p()
        ln = C.c_uint64()
        src = C.c_char_p(payload)
        _check(self.lib.alpacc_hip_run_case(self.h, MODES[mode], n, src, C.byref(rec), C.byref(ln)))
        return _take(rec.value, ln.value, np.uint8).tobytes()

    def run_batch(self, mode: str, payloads) -> bytes:
        """One launch sequence for many requests: the concatenated response
        records, each identical to run_case's record for that payload."""
        offs = np.zeros(len(payloads) + 1, dtype=np.uint64)
        offs[1:] = np.cumsum([len(x) for x in payloads], dtype=np.uint64)
        packed = b"".join(payloads)
        rec = C.c_void_p()
        ln = C.c_uint64()
        rc = self.lib.alpacc_hip_run_batch(self.h, MODES[mode], C.c_char_p(packed),
                                           offs.ctypes.data_as(C.POINTER(C.c_uint64)), len(payloads),
                                           C.byref(rec), C.byref(ln))
        _check(min(rc, 0))
        return _take(rec.value, ln.value, np.uint8).tobytes() if ln.value else b""

    def lex_parse_batch_device(self, d_ptr: int, offsets: np.ndarray, stream: int = 0) -> Batch:
        """Combined mode over inputs packed in device memory; offsets is a
        host uint64 array of count + 1 entries.  Results stay on the device."""
        offs = np.ascontiguousarray(offsets, dtype=np.uint64)
        out = Batch()
        rc = self.lib.alpacc_hip_lex_parse_batch_device(self.h, C.c_void_p(d_ptr),
                                                        offs.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                        len(offs) - 1, C.byref(out), C.c_void_p(stream or None))
        _check(min(rc, 0))
        return out

    def lex_string(self, data: bytes, mode: str = "lexer"):
        """lex_string: (accepted, terminals, starts, lengths)."""
        out = Lexemes()
        ok = _check(self.lib.alpacc_hip_lex_string(self.h, MODES[mode], C.c_char_p(data), len(data), C.byref(out)))
        if not ok:
            return False, None, None, None
        lay = self.layouts[mode]
        n = out.count
        return (True, _take(out.terminals, n * lay.terminal, f"<u{lay.terminal}"),
                _take(out.starts, n * 8, "<i8"), _take(out.lengths, n * lay.length, f"<u{lay.length}"))

    def parse_test(self, tokens: np.ndarray):
        """parse_test: (accepted, productions)."""
        tok = np.ascontiguousarray(tokens, dtype=np.uint8)
        prods = C.c_void_p()
        npr = C.c_uint64()
        ok = _check(self.lib.alpacc_hip_parse_test(self.h, tok.ctypes.data, len(tok), C.byref(prods), C.byref(npr)))
        if not ok:
            return False, None
        pb = self.layouts["both"].production
        return True, _take(prods.value, npr.value * pb, f"<u{pb}")

    def lex_parse(self, data: bytes):
        """lex_string + parse_test + compute_parents + compact_tree."""
        out = Tree()
        ok = _check(self.lib.alpacc_hip_lex_parse(self.h, C.c_char_p(data), len(data), C.byref(out)))
        if not ok:
            return None
        lay = self.layouts["both"]
        T, N = out.tokens.count, out.num_nodes
        return dict(
            terminals=_take(out.tokens.terminals, T * lay.terminal, np.uint8),
            starts=_take(out.tokens.starts, T * 8, "<i8"),
            lengths=_take(out.tokens.lengths, T * lay.length, f"<u{lay.length}"),
            productions=_take(out.productions, N * lay.production, f"<u{lay.production}"),
            parents=_take(out.parents, N * 8, "<i8"),
            token_parents=_take(out.token_parents, T * 8, "<i8"),
        )

    # -- lexing a stream of chunks -------------------------------------------
    def lex_stream(self, chunks, mode: str = "lexer") -> bytes:
        """Lex one input given as chunks (staged one at a time in device
        memory) with alpacc_hip_lex_stream_*; returns the lexer-only wire
        record of the whole input built from the pushes' tokens."""
        import torch
        lay = self.layouts["lexer" if mode == "lexer" else "both"]
        _check(self.lib.alpacc_hip_lex_stream_begin(self.h, MODES[mode]) >= 0 and 1 or -1)
        terms, starts, lens = [], [], []
        ok = True
        chunks = list(chunks)
        for i, ch in enumerate(chunks):
            last = i == len(chunks) - 1
            dev = torch.frombuffer(bytearray(ch + b"\0"), dtype=torch.uint8).to("cuda")
            torch.cuda.synchronize()
            out = Lexemes()
            rc = self.lib.alpacc_hip_lex_stream_push(self.h, C.c_void_p(dev.data_ptr()), len(ch), int(last),
                                                     C.byref(out), None)
            ok = _check(rc)
            T = out.count
            if T:
                terms.append(_take_dev(out.terminals, T * lay.terminal))
                starts.append(_take_dev(out.starts, T * 8))
                lens.append(_take_dev(out.lengths, T * lay.length))
            del dev
        if not ok:
            return b"\x00"
        t = np.frombuffer(b"".join(terms), dtype=np.uint8)
        s = np.frombuffer(b"".join(starts), dtype="<i8")
        l = np.frombuffer(b"".join(lens), dtype=f"<u{lay.length}")
        rec = wire.Record(True, terminals=t, starts=s, lengths=l)
        return wire.encode_record(rec, wire.Layout("lexer", lay.terminal, lay.index, lay.length))

    # -- lexer launch shape ----------------------------------------------------
    def set_lex_shape(self, items_per_thread: int = 128, workgroups_per_cu: int = 0, lookback_window: int = 16):
        sh = LexShape(items_per_thread, workgroups_per_cu, lookback_window, 0.0)
        _check(self.lib.alpacc_hip_set_lex_shape(self.h, C.byref(sh)) >= 0 and 1 or -1)

    def lex_shape(self) -> dict:
        sh = LexShape()
        _check(self.lib.alpacc_hip_get_lex_shape(self.h, C.byref(sh)) >= 0 and 1 or -1)
        return {"items_per_thread": sh.items_per_thread, "workgroups_per_cu": sh.workgroups_per_cu,
                "lookback_window": sh.lookback_window}

    def autotune_lexer(self, d_ptr: int, n: int, mode: str = "both", reps: int = 3) -> dict:
        """Time every lexer shape on a device-resident sample, keep the
        fastest; returns it with its time (ms)."""
        sh = LexShape()
        _check(self.lib.alpacc_hip_autotune_lexer(self.h, MODES[mode], C.c_void_p(d_ptr), n, reps,
                                                  C.byref(sh)) >= 0 and 1 or -1)
        return {"items_per_thread": sh.items_per_thread, "workgroups_per_cu": sh.workgroups_per_cu,
                "lookback_window": sh.lookback_window, "ms": round(sh.ms, 4)}

    # -- parser tile shape --------------------------------------------------------
    def set_parse_shape(self, positions_per_tile: int = 1024, positions_per_lane: int = 4):
        sh = ParseShape(positions_per_tile, positions_per_lane, 0.0)
        _check(min(self.lib.alpacc_hip_set_parse_shape(self.h, C.byref(sh)), 0))

    def parse_shape(self) -> dict:
        sh = ParseShape()
        _check(min(self.lib.alpacc_hip_get_parse_shape(self.h, C.byref(sh)), 0))
        return {"positions_per_tile": sh.positions_per_tile, "positions_per_lane": sh.positions_per_lane}

    def autotune_parser(self, d_ptr: int, n: int, reps: int = 3) -> dict:
        """Time the parser in every tile shape on a device-resident sample
        (lexed in combined mode), keep the fastest; returns it with its time."""
        sh = ParseShape()
        _check(min(self.lib.alpacc_hip_autotune_parser(self.h, C.c_void_p(d_ptr), n, reps, C.byref(sh)), 0))
        return {"positions_per_tile": sh.positions_per_tile, "positions_per_lane": sh.positions_per_lane,
                "ms": round(sh.ms, 4)}

    # -- graphs for combined-mode batches -----------------------------------------
    def set_graphs(self, on: bool = True):
        """Run combined-mode batches through a captured HIP graph."""
        _check(self.lib.alpacc_hip_set_graphs(self.h, int(on)) >= 0 and 1 or -1)

    def set_bracket_width(self, nbytes: int = 0):
        """Element width of the parser's LDS bracket stacks: 0 automatic, 1 or 2 bytes."""
        _check(self.lib.alpacc_hip_set_bracket_width(self.h, int(nbytes)))

    def graph_stats(self):
        """(captures, replays) of the batch graph so far."""
        c, r = C.c_uint64(), C.c_uint64()
        _check(self.lib.alpacc_hip_graph_stats(self.h, C.byref(c), C.byref(r)) >= 0 and 1 or -1)
        return c.value, r.value

    # -- stage timing (HIP events on the launch stream) ------------------------------
    def set_timing(self, on: bool = True):
        _check(self.lib.alpacc_hip_set_timing(self.h, int(on)) >= 0 and 1 or -1)

    def timed_kernels(self):
        """Names of the kernels timed since set_timing(True)."""
        names = self.lib.alpacc_hip_timed_kernels(self.h)
        return [x for x in (names or b"").decode().split(",") if x]

    def timing(self, stage: str):
        """(total milliseconds, runs) of a stage since set_timing."""
        ms = C.c_double()
        runs = C.c_uint64()
        if self.lib.alpacc_hip_timing(self.h, stage.encode(), C.byref(ms), C.byref(runs)) < 0:
            _check(-1)
        return ms.value, runs.value

    # -- device-resident forms ---------------------------------------------------
    def lex_device(self, d_ptr: int, n: int, mode: str = "lexer", stream: int = 0):
        out = Lexemes()
        ok = _check(self.lib.alpacc_hip_lex_device(self.h, MODES[mode], C.c_void_p(d_ptr), n, C.byref(out),
                                                   C.c_void_p(stream or None)))
        return ok, out

    def lex_parse_device(self, d_ptr: int, n: int, stream: int = 0):
        out = Tree()
        ok = _check(self.lib.alpacc_hip_lex_parse_device(self.h, C.c_void_p(d_ptr), n, C.byref(out),
                                                         C.c_void_p(stream or None)))
        return ok, out


class Server:
    """Continuous-batching request loop over resident grammars (see
    alpacc_hip_server_open)."""

    def __init__(self, grammars, device: int = 0, lanes: int = 2, hbm_budget: int = 64 << 30, max_batch: int = 0):
        self.lib = load_library()
        self.grammars = list(grammars)
        arr = (C.c_char_p * len(self.grammars))(*[g.encode() for g in self.grammars])
        h = C.c_void_p()
        _check(self.lib.alpacc_hip_server_open(arr, len(self.grammars), device, lanes, hbm_budget, max_batch,
                                               C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.alpacc_hip_server_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def submit(self, grammar: str, mode: str, payload: bytes) -> int:
        t = C.c_uint64()
        _check(self.lib.alpacc_hip_server_submit(self.h, self.grammars.index(grammar), MODES[mode],
                                                 C.c_char_p(payload), len(payload), C.byref(t)))
        return t.value

    def wait(self, ticket: int) -> bytes:
        rec = C.c_void_p()
        ln = C.c_uint64()
        _check(self.lib.alpacc_hip_server_wait(self.h, ticket, C.byref(rec), C.byref(ln)))
        return _take(rec.value, ln.value, np.uint8).tobytes()

    def replay(self, grammar_idx, modes, payloads, arrival_ns):
        """Open-loop replay: (records concatenated, latencies in ns)."""
        count = len(payloads)
        g = np.ascontiguousarray(grammar_idx, dtype=np.int32)
        m = np.ascontiguousarray([MODES[x] if isinstance(x, str) else x for x in modes], dtype=np.int32)
        offs = np.zeros(count + 1, dtype=np.uint64)
        offs[1:] = np.cumsum([len(x) for x in payloads])
        packed = b"".join(payloads)
        arr = np.ascontiguousarray(arrival_ns, dtype=np.uint64)
        lat = np.zeros(count, dtype=np.uint64)
        rec = C.c_void_p()
        ln = C.c_uint64()
        u64p = C.POINTER(C.c_uint64)
        _check(self.lib.alpacc_hip_server_replay(
            self.h, count, g.ctypes.data_as(C.POINTER(C.c_int32)), m.ctypes.data_as(C.POINTER(C.c_int32)),
            C.c_char_p(packed), offs.ctypes.data_as(u64p), arr.ctypes.data_as(u64p), lat.ctypes.data_as(u64p),
            C.byref(rec), C.byref(ln)))
        return _take(rec.value, ln.value, np.uint8).tobytes(), lat

    def set_graphs(self, on: bool = True):
        """Combined-mode batches of every lane through its captured graph."""
        _check(self.lib.alpacc_hip_server_set_graphs(self.h, int(on)) >= 0 and 1 or -1)

    def stats(self):
        b, r, m = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(self.lib.alpacc_hip_server_stats(self.h, C.byref(b), C.byref(r), C.byref(m)))
        return {"batches": b.value, "requests": r.value, "max_batch": m.value}
