"""The generator's outputs beyond the device tables, on the CPU: tables
files that the library loads at run time, id names, the random grammars of
the parity tests, the `dev` diagnostics, launch limits, and the wire
program's handling of a bogus frame length."""

import functools
import os
import struct
import subprocess
import sys

import pytest

from alpacc_amd.gen import emit_hip, tables
from alpacc_amd.gen.randgrammar import random_grammar

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TEST_GRAMMARS = os.path.join(ROOT, "tests", "golden", "grammars")
LIB = os.path.join(ROOT, "alpacc_amd", "lib", "libalpacc_hip.so")
CLI = os.path.join(ROOT, "alpacc_amd", "lib", "alpacc-hip")

needs_lib = pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")


@functools.lru_cache(maxsize=None)
def load(path):
    return tables.load(path)


def grammar_files():
    out = [os.path.join(ROOT, "grammars", f) for f in sorted(os.listdir(os.path.join(ROOT, "grammars")))
           if f.endswith(".alp")]
    return out + [os.path.join(TEST_GRAMMARS, f) for f in sorted(os.listdir(TEST_GRAMMARS)) if f.endswith(".alp")]


def tables_blob(g, d, name, **over):
    """A tables file of grammar g from its device tables d under another
    name, with fields replaced by `over` (to plant a malformed field)."""
    fields = emit_hip._fields(g, d)
    fields[0] = ("name", "str", name)
    fields = [(n, k, over.get(n, v)) for n, k, v in fields]
    blob = bytearray(emit_hip.TABLES_MAGIC) + struct.pack("<I", len(fields))
    blob += b"\0" * (-len(blob) % 8)
    for n, kind, v in fields:
        if kind == "str":
            blob += struct.pack("<I", len(v)) + v.encode()
        elif kind == "int":
            blob += b"I" + b"\0" * 7 + struct.pack("<q", int(v))
        else:
            fmt, size = {"u8": ("B", 1), "u16": ("H", 2), "i32": ("i", 4), "u64": ("Q", 8)}[kind]
            blob += b"A" + b"\0" * 3 + struct.pack("<IQ", size, len(v)) + struct.pack(f"<{len(v)}{fmt}", *v)
        blob += b"\0" * (-len(blob) % 8)
    return bytes(blob)


# ---------------------------------------------------------------- names

JSON_TERMINALS = ["string", "number"] + [f"literal_{i}" for i in range(2, 11)] + ["ignore", "empty_12"]


@needs_lib
def test_names_json():
    """Named outputs (Futhark/Generator.hs:91-118 names, Analyzer.hs:104-110,
    158-165): literals are literal_<id>, the dead terminal empty_<id>,
    productions their [Name] or <lhs>_<id>."""
    from alpacc_amd import hip
    assert hip.names("json", "terminal") == JSON_TERMINALS
    assert hip.names("json", "terminal_lexer") == ["number", "string"] + JSON_TERMINALS[2:]
    prods = hip.names("json", "production")
    assert prods[:3] == ["Object_production", "Array_production", "String"]
    assert prods[18:21] == ["string_18", "number_19", "ignore_20"] and prods[-1] == "start_30"
    assert hip.names("arithmetic", "production")[:5] == ["E0_0", "Add", "Sub", "Empty", "E1_4"]
    assert hip.names("lexer", "production") == []


@needs_lib
def test_names_match_generator_for_every_grammar():
    from alpacc_amd import hip
    for name in hip.grammars():
        g = load(os.path.join(ROOT, "grammars", name + ".alp"))
        assert hip.names(name, "terminal") == g.lex_both.names()
        assert hip.names(name, "terminal_lexer") == g.lex_only.names()
        assert hip.names(name, "production") == (g.parser.prod_names if g.has_parser else [])


def test_names_cli(tmp_path):
    if not os.path.exists(CLI):
        pytest.skip("cli not built")
    r = subprocess.run([CLI, "-g", "json", "--names"], capture_output=True, text=True)
    lines = r.stdout.splitlines()
    assert r.returncode == 0 and lines[0] == "terminal 0 string" and lines[13] == "production 0 Object_production"
    r = subprocess.run([CLI, "-g", "json", "--lexer", "--names"], capture_output=True, text=True)
    assert r.stdout.splitlines()[:2] == ["terminal 0 number", "terminal 1 string"]
    r = subprocess.run([sys.executable, "-m", "alpacc_amd.gen", "names", os.path.join(ROOT, "grammars", "json.alp")],
                       capture_output=True, text=True, cwd=ROOT)
    assert r.stdout.splitlines() == lines


# --------------------------------------------------------------- tables files

@needs_lib
def test_tables_files_load(tmp_path):
    """Every test grammar's tables file loads; the library then reports the
    generator's names, layout and limits for it.  Loading the same file
    twice is a no-op."""
    from alpacc_amd import hip
    for path in grammar_files():
        if os.path.dirname(path) != TEST_GRAMMARS:
            continue
        g = load(path)
        f = tmp_path / (g.name + ".alphip")
        f.write_bytes(emit_hip.emit_tables(g))
        assert hip.load_grammar(str(f)) == g.name
        assert hip.load_grammar(str(f)) == g.name
        assert g.name in hip.grammars()
        assert hip.names(g.name, "terminal") == g.lex_both.names()
        assert hip.names(g.name, "production") == (g.parser.prod_names if g.has_parser else [])
        lib = hip.load_library()
        import ctypes as C
        buf = C.create_string_buffer(512)
        assert lib.alpacc_hip_grammar_layout(g.name.encode(), 0, buf, 512) > 0
        kv = dict(x.split("=") for x in buf.value.decode().split())
        assert int(kv["terminal_t"]) == g.terminal_bytes_lex and int(kv["length_t"]) == g.length_bytes
        rc = lib.alpacc_hip_grammar_layout(g.name.encode(), 2, buf, 512)
        assert (rc > 0) == g.has_parser


@needs_lib
def test_tables_file_refusals(tmp_path):
    from alpacc_amd import hip
    g = load(os.path.join(TEST_GRAMMARS, "rand_q1k1_s3.alp"))
    blob = emit_hip.emit_tables(g)
    cases = {
        "magic": b"XXXXXXXX" + blob[8:],
        "truncated": blob[: len(blob) // 2],
        "fields": blob[:8] + struct.pack("<I", 3) + blob[12:],
    }
    for what, data in cases.items():
        f = tmp_path / f"{what}.alphip"
        f.write_bytes(data)
        with pytest.raises(hip.HipError):
            hip.load_grammar(str(f))
    # another grammar under a loaded name
    other = load(os.path.join(TEST_GRAMMARS, "rand_q1k1_s5.alp"))
    other.name = g.name
    f1, f2 = tmp_path / "a.alphip", tmp_path / "b.alphip"
    f1.write_bytes(blob)
    f2.write_bytes(emit_hip.emit_tables(other))
    hip.load_grammar(str(f1))
    with pytest.raises(hip.HipError):
        hip.load_grammar(str(f2))
    # the name of a compiled-in grammar
    j = load(os.path.join(ROOT, "grammars", "list.alp"))
    f3 = tmp_path / "list.alphip"
    f3.write_bytes(emit_hip.emit_tables(j))
    with pytest.raises(hip.HipError):
        hip.load_grammar(str(f3))


def test_gen_hip_command(tmp_path):
    out = tmp_path / "t.alphip"
    src = tmp_path / "t.cpp"
    r = subprocess.run([sys.executable, "-m", "alpacc_amd.gen", "hip", os.path.join(ROOT, "grammars", "sexp.alp"),
                        "-o", str(out), "--source", str(src)], cwd=ROOT, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert out.read_bytes()[:8] == emit_hip.TABLES_MAGIC and "GRAMMAR_SEXP" in src.read_text()
    bad = tmp_path / "bad.alp"
    bad.write_text('S -> "a" A | "b" B.\nA -> "c" A "d" A | "f".\nB -> "e" B "d" B | "f".\n')
    r = subprocess.run([sys.executable, "-m", "alpacc_amd.gen", "hip", str(bad)], cwd=ROOT, capture_output=True,
                       text=True)
    assert r.returncode == 1 and "not LLP" in r.stderr


# ------------------------------------------------------------ random grammars

def test_random_grammar_fixtures_regenerate():
    """The committed test grammars are the generator's output for their
    seeds (tests/golden/make_random_grammars.py)."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_random_grammars as M
    names = sorted(f[:-4] for f in os.listdir(TEST_GRAMMARS) if f.endswith(".alp"))
    assert len([n for n in names if n.startswith("rand_")]) == len(M.SETTINGS) * M.PER_SETTING
    for n in names:
        text = open(os.path.join(TEST_GRAMMARS, n + ".alp")).read()
        if n in M.SHAPED:
            assert text.split("\n", 1)[1] == M.SHAPED[n]
            continue
        seed = int(text.splitlines()[0].split()[-1])
        q, k = int(n[6]), int(n[8])
        lexer_only = (q, k) == (0, 0)
        body = random_grammar(seed, num_productions=0 if lexer_only else 3, q=None if lexer_only else q,
                              k=None if lexer_only else k)
        assert text.split("\n", 1)[1] == body


def test_random_grammar_shape():
    for seed in range(30):
        text = random_grammar(seed, q=1, k=1)
        cfg = __import__("alpacc_amd.gen.grammar", fromlist=["x"]).parse_cfg(text)
        assert [r.term[1] for r in cfg.t_rules] == ["a", "b", "c"]
        assert len(cfg.nt_rules) == 3
        for r in cfg.nt_rules:
            assert not r.rhs or r.rhs[0][0] == "T"
        assert random_grammar(seed, q=1, k=1) == text
    assert "->" not in random_grammar(3, num_productions=0)


# -------------------------------------------------------------------- dev

def test_dev_reports():
    from alpacc_amd.gen import dev
    g = load(os.path.join(ROOT, "grammars", "arithmetic.alp"))
    h = dev.composition_histogram(g)
    n = g.lex_both.endos.size
    assert f"N (endomorphisms)   : {n}" in h and f"Total cells (N*N)   : {n * n}" in h
    rows = [l for l in h.splitlines() if l.strip()[:1].isdigit()]
    assert sum(int(r.split()[2]) for r in rows) <= n * n and len(rows) <= 32
    ims = dev.image_sizes(g)
    # the closure of the byte maps is the generator's endomorphism closure
    # without the identity (the empty string's map) and, when no string of
    # the grammar's bytes dies, without the dead element the generator adds
    size = int(next(l for l in ims.splitlines() if l.startswith("Monoid closure size")).split(":")[1])
    assert size in (n - 2, n - 1)
    d = dev.dump_endo(g, ord("("))
    assert d.startswith("Char 40 (0x28):")
    assert "not in endo table" in dev.dump_endo(g, 0)
    for what in ("composition-histogram", "image-sizes", "raw-image-sizes"):
        r = subprocess.run([sys.executable, "-m", "alpacc_amd.gen", "dev", what,
                            os.path.join(ROOT, "grammars", "sexp.alp")], cwd=ROOT, capture_output=True, text=True)
        assert r.returncode == 0 and r.stdout


# ------------------------------------------------------------------ limits

@needs_lib
def test_launch_limits():
    """One parser launch takes positions p with p * per_position + inputs <
    2^31, per_position the largest push or pop count of a window;
    bigger batches are split, a bigger single input is refused."""
    from alpacc_amd import hip
    for name in ("json", "arithmetic", "sexp"):
        g = load(os.path.join(ROOT, "grammars", name + ".alp"))
        d = emit_hip.device_tables(g)
        per = max(1, d["max_bpush"], d["max_spush"], d["max_bpop"], d["max_spop"])
        lex_bytes, pos = hip.limits(name)
        assert lex_bytes == (1 << 31) - 1
        assert pos * per + pos < (1 << 31) <= (pos + 2) * (per + 1) + 1


def test_cli_bogus_frame_length():
    """A frame length far beyond the input is a truncated frame (the payload
    is read in bounded pieces), not an allocation failure."""
    if not os.path.exists(CLI):
        pytest.skip("cli not built")
    data = struct.pack("<QQQ", 1, 1 << 62, 5) + b"12345"
    r = subprocess.run([CLI, "-g", "arithmetic"], input=data, capture_output=True)
    assert r.returncode == 1 and r.stderr.decode().strip() == "error: truncated frame"
    assert r.stdout == struct.pack("<Q", 1)


# ------------------------------------------------------- hashed window tables

def probe(d, code):
    """The pair id a hashed window table gives a window code (the parser's
    probe: first slot from the code's Fibonacci hash, then linear)."""
    bits = d["window_hash"]
    slot, mask = emit_hip.window_slot(code, bits), (1 << bits) - 1
    while True:
        k = d["window_keys"][slot]
        if k == code:
            return d["window"][slot]
        if k == emit_hip.EMPTY_KEY:
            return 0xFFFF
        slot = (slot + 1) & mask


@pytest.mark.parametrize("name", ["json", "arithmetic", "sexp", "paper_grammar", "llp01", "list"])
def test_hashed_window_table_matches_dense_index(name):
    """The open-addressing window table (grammars whose dense index would be
    too large) gives every window code the pair id the dense index gives it,
    including the windows without one; the pair records are the same, the
    table is at most half full and at least one slot is empty."""
    g = load(os.path.join(ROOT, "grammars", name + ".alp"))
    dense, hashed = emit_hip.device_tables(g, hashed=False), emit_hip.device_tables(g, hashed=True)
    assert dense["window_hash"] == 0 and hashed["window_hash"] >= 4
    assert len(hashed["window"]) == len(hashed["window_keys"]) == 1 << hashed["window_hash"]
    live = sum(k != emit_hip.EMPTY_KEY for k in hashed["window_keys"])
    assert 0 < live <= len(hashed["window"]) // 2
    assert live == sum(v != 0xFFFF for v in dense["window"])
    assert hashed["pair_rec"] == dense["pair_rec"]
    assert all(probe(hashed, c) == v for c, v in enumerate(dense["window"]))
    # a code no window has (one past the last) finds nothing
    assert probe(hashed, len(dense["window"])) == 0xFFFF


def test_hashed_window_table_is_automatic_past_the_dense_limit(monkeypatch):
    """Past MAX_WINDOW_ENTRIES the generator emits the hashed table instead
    of refusing the grammar."""
    g = load(os.path.join(ROOT, "grammars", "json.alp"))
    monkeypatch.setattr(emit_hip, "MAX_WINDOW_ENTRIES", 15 ** 4 - 1)
    assert emit_hip.device_tables(g)["window_hash"] > 0
    monkeypatch.setattr(emit_hip, "MAX_WINDOW_ENTRIES", 15 ** 4)
    assert emit_hip.device_tables(g)["window_hash"] == 0


@needs_lib
def test_hashed_tables_file_loads_and_checks(tmp_path):
    """A hashed tables file loads under its own name; the loader refuses a
    dense index shorter than the window codes and a hashed table whose slot
    count is not 2^window_hash."""
    from alpacc_amd import hip
    g = load(os.path.join(ROOT, "grammars", "json.alp"))
    f = tmp_path / "json_hashed_cpu.alphip"
    f.write_bytes(emit_hip.emit_tables(g, hashed=True, name="json_hashed_cpu"))
    assert hip.load_grammar(str(f)) == "json_hashed_cpu"
    assert hip.names("json_hashed_cpu", "terminal") == g.lex_both.names()

    dense = emit_hip.device_tables(g, hashed=False)
    short = dense["window"][:-1]
    bad = tmp_path / "short.alphip"
    bad.write_bytes(tables_blob(g, dense, "json_short", nwindow=len(short), window=short))
    with pytest.raises(hip.HipError):
        hip.load_grammar(str(bad))
    hashed = emit_hip.device_tables(g, hashed=True)
    bad2 = tmp_path / "slots.alphip"
    bad2.write_bytes(tables_blob(g, hashed, "json_slots", window_hash=hashed["window_hash"] + 1))
    with pytest.raises(hip.HipError):
        hip.load_grammar(str(bad2))


def test_two_byte_production_ids():
    """A grammar of more than 255 productions (many_productions, 268) takes
    two-byte production ids: each pair record keeps the low byte of its
    productions (bytes 24..31), the side table the whole ids, 8 a pair, in
    the LLP table's order for that pair, zero past the pair's count and for
    the empty record after the last pair; some id passes 255."""
    g = load(os.path.join(TEST_GRAMMARS, "many_productions.alp"))
    d = emit_hip.device_tables(g)
    assert d["production_bytes"] == 2 and d["n_productions"] == 268
    npairs, rec, p16 = d["npairs"], d["pair_rec"], d["pair_prod16"]
    assert len(p16) == 8 * (npairs + 1) and len(rec) == 64 * (npairs + 1)
    for i in range(npairs):
        p0, pn = d["spans"][4 * i + 2], d["spans"][4 * i + 3]
        want = list(d["productions"][p0:p0 + pn])
        ids = p16[8 * i:8 * i + 8]
        assert rec[64 * i + 2] == pn
        assert ids == want + [0] * (8 - pn)
        assert rec[64 * i + 24:64 * i + 32] == [x & 0xFF for x in ids]
    assert p16[8 * npairs:] == [0] * 8
    assert max(p16) > 255 and max(p16) < d["n_productions"]
    # one-byte grammars carry no side table
    assert emit_hip.device_tables(load(os.path.join(ROOT, "grammars", "json.alp")))["pair_prod16"] == []


@needs_lib
def test_two_byte_production_tables_file(tmp_path):
    """The tables file of a two-byte-id grammar loads; the loader refuses a
    side table shorter than 8 ids a pair and production ids wider than two
    bytes."""
    from alpacc_amd import hip
    g = load(os.path.join(TEST_GRAMMARS, "many_productions.alp"))
    f = tmp_path / "many_productions_cpu.alphip"
    f.write_bytes(emit_hip.emit_tables(g, name="many_productions_cpu"))
    assert hip.load_grammar(str(f)) == "many_productions_cpu"
    assert len(hip.names("many_productions_cpu", "production")) == 268
    d = emit_hip.device_tables(g)
    for name, over in [("mp_short", dict(pair_prod16=d["pair_prod16"][:-1])), ("mp_wide", dict(production_bytes=3))]:
        bad = tmp_path / (name + ".alphip")
        bad.write_bytes(tables_blob(g, d, name, **over))
        with pytest.raises(hip.HipError):
            hip.load_grammar(str(bad))
