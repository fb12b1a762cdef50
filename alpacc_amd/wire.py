"""Byte-level wire protocol of generated alpacc programs (docs/wire-protocols.md).

Batch input:  u64 count, then per request ``u64 frame_len, u64 n, payload``.
Batch output: u64 count, then per request a response record: ``u8 valid``
and, if valid, the mode's fields (lexer: u64 count + AoS lexemes; parser:
u64 count + productions; combined: SoA tokens, compact tree, token parents).
All integers are little-endian with the grammar's native widths
(``--layout``).
"""

from __future__ import annotations

import dataclasses
import io
import struct
from typing import List, Optional, Sequence

import numpy as np

_NP = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}
_NPI = {4: np.int32, 8: np.int64}


@dataclasses.dataclass(frozen=True)
class Layout:
    mode: str                 # "lexer" | "parser" | "both"
    terminal: int
    index: int
    length: int = 0           # absent in parser-only builds
    production: int = 0       # absent in lexer-only builds

    def lines(self) -> str:
        out = [f"terminal_t={self.terminal}"]
        if self.mode != "lexer":
            out.append(f"production_t={self.production}")
        out.append(f"index_t={self.index}")
        if self.mode != "parser":
            out.append(f"length_t={self.length}")
        return "\n".join(out) + "\n"

    def payload_bytes(self, n: int) -> int:
        return n * self.terminal if self.mode == "parser" else n


def encode_batch(payloads: Sequence[bytes], layout: Layout, counts: Optional[Sequence[int]] = None) -> bytes:
    """One counted batch.  ``counts`` gives n per request (defaults to the
    payload length divided by the element width)."""
    out = io.BytesIO()
    out.write(struct.pack("<Q", len(payloads)))
    for i, p in enumerate(payloads):
        n = counts[i] if counts is not None else len(p) // (layout.terminal if layout.mode == "parser" else 1)
        out.write(struct.pack("<QQ", 8 + len(p), n))
        out.write(p)
    return out.getvalue()


@dataclasses.dataclass
class Record:
    valid: bool
    # lexer / combined
    terminals: Optional[np.ndarray] = None
    starts: Optional[np.ndarray] = None
    lengths: Optional[np.ndarray] = None
    # parser
    productions: Optional[np.ndarray] = None
    # combined tree
    nodes: Optional[np.ndarray] = None
    parents: Optional[np.ndarray] = None
    token_parents: Optional[np.ndarray] = None
    raw: bytes = b""


def _take(buf: memoryview, off: int, width: int, count: int, signed=False):
    dt = (_NPI if signed else _NP)[width]
    nbytes = width * count
    arr = np.frombuffer(buf[off:off + nbytes], dtype=np.dtype(dt).newbyteorder("<"), count=count)
    return arr, off + nbytes


def decode_record(buf: bytes, off: int, layout: Layout):
    """Decode one response record at ``off``; returns (Record, new offset)."""
    mv = memoryview(buf)
    start = off
    valid = buf[off]
    off += 1
    if valid == 0:
        return Record(False, raw=bytes(mv[start:off])), off
    (cnt,) = struct.unpack_from("<Q", buf, off)
    off += 8
    rec = Record(True)
    if layout.mode == "lexer":
        dt = np.dtype([("t", f"<u{layout.terminal}"), ("s", f"<i{layout.index}"), ("l", f"<u{layout.length}")])
        arr = np.frombuffer(mv[off:off + dt.itemsize * cnt], dtype=dt, count=cnt)
        off += dt.itemsize * cnt
        rec.terminals, rec.starts, rec.lengths = arr["t"].copy(), arr["s"].copy(), arr["l"].copy()
    elif layout.mode == "parser":
        rec.productions, off = _take(mv, off, layout.production, cnt)
    else:
        rec.terminals, off = _take(mv, off, layout.terminal, cnt)
        rec.starts, off = _take(mv, off, layout.index, cnt, signed=True)
        rec.lengths, off = _take(mv, off, layout.length, cnt)
        (nn,) = struct.unpack_from("<Q", buf, off)
        off += 8
        rec.nodes, off = _take(mv, off, layout.production, nn)
        rec.parents, off = _take(mv, off, layout.index, nn, signed=True)
        rec.token_parents, off = _take(mv, off, layout.index, cnt, signed=True)
    rec.raw = bytes(mv[start:off])
    return rec, off


def encode_record(rec: Record, layout: Layout) -> bytes:
    """The response record of ``rec`` in ``layout`` (the inverse of
    decode_record; integer fields are narrowed or widened to its widths)."""
    if not rec.valid:
        return b"\x00"
    out = io.BytesIO()
    out.write(b"\x01")
    if layout.mode == "parser":
        out.write(struct.pack("<Q", len(rec.productions)))
        out.write(np.asarray(rec.productions, dtype=f"<u{layout.production}").tobytes())
        return out.getvalue()
    cnt = len(rec.terminals)
    out.write(struct.pack("<Q", cnt))
    if layout.mode == "lexer":
        dt = np.dtype([("t", f"<u{layout.terminal}"), ("s", f"<i{layout.index}"), ("l", f"<u{layout.length}")])
        arr = np.empty(cnt, dtype=dt)
        arr["t"], arr["s"], arr["l"] = rec.terminals, rec.starts, rec.lengths
        out.write(arr.tobytes())
        return out.getvalue()
    out.write(np.asarray(rec.terminals, dtype=f"<u{layout.terminal}").tobytes())
    out.write(np.asarray(rec.starts, dtype=f"<i{layout.index}").tobytes())
    out.write(np.asarray(rec.lengths, dtype=f"<u{layout.length}").tobytes())
    out.write(struct.pack("<Q", len(rec.nodes)))
    out.write(np.asarray(rec.nodes, dtype=f"<u{layout.production}").tobytes())
    out.write(np.asarray(rec.parents, dtype=f"<i{layout.index}").tobytes())
    out.write(np.asarray(rec.token_parents, dtype=f"<i{layout.index}").tobytes())
    return out.getvalue()


def decode_batch(buf: bytes, layout: Layout) -> List[Record]:
    (count,) = struct.unpack_from("<Q", buf, 0)
    off = 8
    out = []
    for _ in range(count):
        r, off = decode_record(buf, off, layout)
        out.append(r)
    if off != len(buf):
        raise ValueError(f"{len(buf) - off} trailing bytes after {count} records")
    return out


def split_records(buf: bytes, layout: Layout) -> List[bytes]:
    return [r.raw for r in decode_batch(buf, layout)]
