"""Seeded synthetic inputs of the benchmark workloads (wraps datagen.c)."""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libdatagen.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m alpacc_amd.build` first")
        lib = C.CDLL(LIB_PATH)
        for name, extra in (("datagen_arith", 1), ("datagen_json", 2), ("datagen_sexp", 2)):
            fn = getattr(lib, name)
            fn.restype = C.c_uint64
            fn.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64] + [C.c_int] * extra
        _lib = lib
    return _lib


def generate(grammar: str, n: int, seed: int, conforming: bool = True, wrap: bool = True,
             out: np.ndarray = None) -> np.ndarray:
    """n bytes of `grammar` text.  For json/sexp, ``wrap`` puts all values in
    one top-level container so the input is a single parseable value."""
    lib = _load()
    buf = out if out is not None else np.empty(n, dtype=np.uint8)
    assert buf.dtype == np.uint8 and buf.size >= n and buf.flags.c_contiguous
    ptr = buf.ctypes.data
    if grammar == "arithmetic":
        lib.datagen_arith(ptr, n, seed, int(conforming))
    elif grammar == "json":
        lib.datagen_json(ptr, n, seed, int(conforming), int(wrap))
    elif grammar == "sexp":
        lib.datagen_sexp(ptr, n, seed, int(conforming), int(wrap))
    else:
        raise ValueError(f"no generator for grammar {grammar!r}")
    return buf[:n]
