"""HIP path vs the CPU oracle beyond the four benchmark grammars, and the
known answers on the GPU.

* Known answers, hard-coded (not the oracle's output): the JSON lexer-only
  record of `"a" 1` (number = 0, string = 1 in lexer mode), the combined
  record of `["a",1]`, and arithmetic `3-5` (num num: a lexer record, a
  combined reject).
* Every other grammar compiled into the library (lexer.alp has terminal
  rules only; llp01 has q = 0; list; numfmt has 457 endomorphisms) and the
  test grammars of tests/golden/grammars loaded from their tables files at
  run time (random grammars at (q, k) = (0, 0) lexer-only, (0, 1), (1, 1),
  (2, 2), (3, 3), as the reference's differential suite runs them,
  Makefile:80-88; a grammar whose windows hold up to 7 productions, which
  takes the resolve kernel's widest instantiation; one with bracket symbols
  up to 307, which needs two-byte bracket stacks; one of 268 productions,
  whose ids take two bytes): the test-vector tool's
  exhaustive, DFA-walk and derivation batches in every mode the grammar has,
  one batched call each, every record byte-identical to the oracle's; the
  combined batches also through the captured graph, and every bracket width
  the grammar allows.
* Inputs beyond one launch: a counted batch of two 1.2 GiB JSON frames
  through the wire program (answered in two launches), a 2.2 GiB JSON
  lexer-only input (lexed through the stream path in chunks) and a 2.05 GiB
  JSON combined input (the same, then one parser launch).
"""

import hashlib
import os
import struct
import subprocess

import numpy as np
import pytest

import pyoracle
from alpacc_amd import datagen, testvec, wire
from alpacc_amd.gen import emit_hip, tables

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TEST_GRAMMARS = os.path.join(ROOT, "tests", "golden", "grammars")
CLI = os.path.join(ROOT, "alpacc_amd", "lib", "alpacc-hip")
BUILTIN_EXTRA = ["lexer", "list", "llp01", "numfmt"]
TEST_NAMES = sorted(f[:-4] for f in os.listdir(TEST_GRAMMARS) if f.endswith(".alp"))


def grammar_of(name):
    p = os.path.join(TEST_GRAMMARS, name + ".alp")
    return tables.load(p if os.path.exists(p) else os.path.join(ROOT, "grammars", name + ".alp"))


def load_tables(name):
    from alpacc_amd import hip
    path = os.path.join(TEST_GRAMMARS, name + ".alphip")
    if not os.path.exists(path):  # normally written by the build
        with open(path, "wb") as f:
            f.write(emit_hip.emit_tables(grammar_of(name)))
    return hip.load_grammar(path)


def first_diff(a: bytes, b: bytes):
    """Offset of the first differing byte (None when equal), found block by
    block so that records of several GB compare quickly."""
    if a == b:
        return None
    n = min(len(a), len(b))
    va, vb = memoryview(a), memoryview(b)
    step = 1 << 26
    for s in range(0, n, step):
        e = min(n, s + step)
        if va[s:e] != vb[s:e]:
            return s + int(np.argmax(np.frombuffer(va[s:e], np.uint8) != np.frombuffer(vb[s:e], np.uint8)))
    return n


# ------------------------------------------------------------- known answers

def test_known_answer_records():
    from alpacc_amd import hip
    with hip.Backend("json") as js, hip.Backend("arithmetic") as ar:
        # lexer mode: number 0, string 1; u8 terminal, i64 start, u16 length
        exp = b"\x01" + struct.pack("<Q", 2) + struct.pack("<BqH", 1, 0, 3) + struct.pack("<BqH", 0, 4, 1)
        assert js.run_case("lexer", b'"a" 1') == exp
        r = js.lex_parse(b'["a",1]')
        assert r["terminals"].tolist() == [4, 0, 2, 1, 5]
        assert r["productions"].tolist() == [1, 13, 15, 2, 17, 3, 16]
        assert r["parents"].tolist() == [0, 0, 1, 2, 2, 4, 4]
        assert r["token_parents"].tolist() == [1, 3, 4, 5, 1]
        exp = b"\x01" + struct.pack("<Q", 2) + struct.pack("<BqQ", 0, 0, 1) + struct.pack("<BqQ", 0, 1, 2)
        assert ar.run_case("lexer", b"3-5") == exp
        assert ar.run_case("both", b"3-5") == b"\x00"


# ------------------------------------------------------- grammar by grammar

def batches(g, mode):
    """The test-vector tool's batches for a grammar and mode."""
    out = []
    alpha = testvec.alphabet(g) if mode != "parser" else \
        sorted(set(g.lex_both.term_id.values()) - {g.lex_both.ignore_id})
    k = 4 if len(alpha) <= 6 else (3 if len(alpha) <= 24 else 2)
    out.append(("exhaustive", testvec.generate(g, mode, exhaustive=k)))
    if mode != "parser":
        out.append(("walks", testvec.generate(g, mode, long=3000, count=12, seed=1)))
    if g.has_parser:
        out.append(("derivations", testvec.generate(g, mode, derivations=300, seed=2)))
    return out


def check_grammar(name, graphs=True):
    from alpacc_amd import hip
    g = grammar_of(name)
    with hip.Backend(name) as be:
        for mode in g.modes():
            for kind, items in batches(g, mode):
                print(f"case {name}/{mode}/{kind}: {len(items)} inputs", flush=True)
                got = be.run_batch(mode, items)
                exp = b"".join(pyoracle.record(name, mode, x) for x in items)
                if got != exp:
                    lay = testvec.layout(g, mode)
                    hdr = struct.pack("<Q", len(items))
                    msgs = testvec.compare(hdr + exp, hdr + got, lay)
                    pytest.fail(f"{name}/{mode}/{kind}: {msgs}")
                if mode == "both" and graphs:
                    be.set_graphs(True)
                    for rep in range(2):  # capture, then replay
                        assert be.run_batch(mode, items) == exp, (name, kind, "graph", rep)
                    be.set_graphs(False)
        # a few single calls through run_case
        mode = g.modes()[-1]
        for x in batches(g, mode)[0][1][:50]:
            assert be.run_case(mode, x) == pyoracle.record(name, mode, x), (name, mode, x)


@pytest.mark.parametrize("name", BUILTIN_EXTRA)
def test_builtin_grammar(name):
    check_grammar(name)


@pytest.mark.parametrize("name", TEST_NAMES)
def test_tables_file_grammar(name):
    assert load_tables(name) == name
    check_grammar(name)


@pytest.mark.parametrize("name", ["wide_windows", "many_symbols", "many_productions"])
def test_shaped_grammar_bracket_widths(name):
    """Every bracket stack width the grammar allows gives the oracle's
    records, eager and through the graph; one-byte stacks are refused when a
    bracket symbol does not fit."""
    from alpacc_amd import hip
    load_tables(name)
    g = grammar_of(name)
    items = testvec.generate(g, "both", derivations=300, seed=3)
    exp = b"".join(pyoracle.record(name, "both", x) for x in items)
    with hip.Backend(name) as be:
        for width in (0, 1, 2):
            if name in ("many_symbols", "many_productions") and width == 1:
                with pytest.raises(hip.HipError):
                    be.set_bracket_width(1)
                continue
            be.set_bracket_width(width)
            print(f"case {name} width {width}", flush=True)
            assert be.run_batch("both", items) == exp, (name, width)
            be.set_graphs(True)
            assert be.run_batch("both", items) == exp, (name, width, "graph")
            assert be.run_batch("both", items) == exp, (name, width, "graph replay")
            be.set_graphs(False)


@pytest.mark.parametrize("name", ["json", "arithmetic", "sexp"])
def test_hashed_window_table(name, tmp_path):
    """A grammar whose windows are looked up in the open-addressing table
    (the form grammars take when their dense window index would be too
    large), loaded under another name: every record equals the oracle's for
    the grammar itself, in the parser and combined modes, eager and through
    the graph, on the test-vector batches and on generated inputs."""
    from alpacc_amd import hip
    g = grammar_of(name)
    alias = name + "_hashed"
    f = tmp_path / (alias + ".alphip")
    f.write_bytes(emit_hip.emit_tables(g, hashed=True, name=alias))
    assert hip.load_grammar(str(f)) == alias
    gen = [datagen.generate(name, n, seed=40 + n % 7, conforming=True).tobytes() for n in (1, 100, 5000, 70000)]
    with hip.Backend(alias) as be:
        for mode in ("parser", "both"):
            for kind, items in batches(g, mode) + ([("generated", gen)] if mode == "both" else []):
                print(f"case {alias}/{mode}/{kind}: {len(items)} inputs", flush=True)
                exp = b"".join(pyoracle.record(name, mode, x) for x in items)
                assert be.run_batch(mode, items) == exp, (alias, mode, kind)
                if mode == "both":
                    be.set_graphs(True)
                    for rep in range(2):
                        assert be.run_batch(mode, items) == exp, (alias, kind, "graph", rep)
                    be.set_graphs(False)


def test_lexer_only_grammar_refuses_parser_modes():
    from alpacc_amd import hip
    with hip.Backend("lexer") as be:
        assert set(be.layouts) == {"lexer"}
        with pytest.raises(hip.HipError):
            be.lex_parse(b"1 2")
        with pytest.raises(hip.HipError):
            be.run_batch("both", [b"1"])


# ----------------------------------------------------- beyond one launch

def stream_hash(cmd, path):
    """sha256 and length of a program's stdout on the file at path."""
    h = hashlib.sha256()
    n = 0
    with open(path, "rb") as f:
        p = subprocess.Popen(cmd, stdin=f, stdout=subprocess.PIPE, stderr=subprocess.PIPE)
        while True:
            b = p.stdout.read(1 << 24)
            if not b:
                break
            h.update(b)
            n += len(b)
        err = p.stderr.read()
        rc = p.wait(timeout=600)
    return rc, h.hexdigest(), n, err


def test_cli_two_large_frames(tmp_path):
    """A counted batch of two 1.2 GiB JSON frames: more input than one lexer
    launch takes, answered frame by frame exactly as the oracle program."""
    n = int(1.2 * (1 << 30))
    path = tmp_path / "batch.bin"
    with open(path, "wb") as f:
        f.write(struct.pack("<Q", 2))
        for seed in (5, 6):
            data = datagen.generate("json", n, seed=seed, conforming=True)
            f.write(struct.pack("<QQ", 8 + n, n))
            f.write(memoryview(data))
            del data
    print("case cli two 1.2 GiB json frames", flush=True)
    ours = stream_hash([CLI, "-g", "json"], path)
    ref = stream_hash([pyoracle.program("json", "both")], path)
    assert ours[0] == 0, ours[3]
    assert ours[1:3] == ref[1:3]
    assert ref[2] > 8 + 2 * n  # both frames accepted: records larger than their inputs


def test_lexer_input_beyond_launch_limit():
    """A 2.2 GiB lexer-only input runs through the stream path in chunks."""
    from alpacc_amd import hip
    n = int(2.2 * (1 << 30))
    data = datagen.generate("json", n, seed=8, conforming=True).tobytes()
    print("case json lexer 2.2 GiB", flush=True)
    with hip.Backend("json") as be:
        got = be.run_case("lexer", data)
    exp = pyoracle.record("json", "lexer", data)
    assert got[0] == 1 and first_diff(got, exp) is None


def test_combined_input_beyond_launch_limit():
    """A combined-mode input past one lexer launch (2.05 GiB of JSON): lexed
    in chunks through the stream path, then parsed in one launch; the whole
    record (about 8 GB: tokens, then the tree) equals the oracle's."""
    from alpacc_amd import hip
    n = int(2.05 * (1 << 30))
    data = datagen.generate("json", n, seed=9, conforming=True).tobytes()
    print("case json combined 2.05 GiB", flush=True)
    with hip.Backend("json") as be:
        got = be.run_case("both", data)
    print("case json combined 2.05 GiB: oracle", flush=True)
    exp = pyoracle.record("json", "both", data)
    assert got[0] == 1 and first_diff(got, exp) is None
