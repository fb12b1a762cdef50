"""The test-vector tool (alpacc_amd/testvec.py) against the oracle: its
generators build what they claim and its comparison finds a planted
difference.  The oracle program answers the generated batches."""

import os
import subprocess

import pytest

import pyoracle
from alpacc_amd import testvec, wire

GRAMMARS = ["json", "arithmetic", "sexp", "paper_grammar"]


def run_oracle(grammar, mode, batch: bytes) -> bytes:
    return subprocess.run([pyoracle.program(grammar, mode)], input=batch, capture_output=True, check=True,
                          timeout=120).stdout


@pytest.mark.parametrize("grammar", GRAMMARS)
def test_exhaustive_counts_and_compare(grammar):
    g = testvec.load(grammar)
    alpha = testvec.alphabet(g)
    items = testvec.generate(g, "lexer", exhaustive=2)
    assert len(items) == len(alpha) ** 2 and len(set(items)) == len(items)
    lay = testvec.layout(g, "lexer")
    out = run_oracle(grammar, "lexer", wire.encode_batch(items, lay))
    assert testvec.compare(out, out, lay) == []
    # a planted difference: flip the verdict byte of the first record
    bad = bytearray(out)
    bad[8] ^= 1
    assert testvec.compare(out, bytes(bad), lay)


@pytest.mark.parametrize("grammar", GRAMMARS)
def test_dfa_walks_are_accepted(grammar):
    g = testvec.load(grammar)
    lay = testvec.layout(g, "lexer")
    items = testvec.generate(g, "lexer", long=3000, count=4, seed=7)
    assert all(len(x) >= 3000 for x in items)
    recs = wire.decode_batch(run_oracle(grammar, "lexer", wire.encode_batch(items, lay)), lay)
    assert all(r.valid for r in recs)


@pytest.mark.parametrize("grammar", GRAMMARS)
def test_derivations_parse(grammar):
    g = testvec.load(grammar)
    items = testvec.generate(g, "parser", derivations=40, seed=3)
    lay = testvec.layout(g, "parser")
    recs = wire.decode_batch(run_oracle(grammar, "parser", wire.encode_batch(items, lay)), lay)
    # the unchanged derivations (even positions) are in the language
    assert all(r.valid for r in recs[0::2])
    if g.lex_both.ignore_id is not None:
        texts = testvec.generate(g, "both", derivations=20, seed=3)
        lay = testvec.layout(g, "both")
        recs = wire.decode_batch(run_oracle(grammar, "both", wire.encode_batch(texts, lay)), lay)
        assert all(r.valid for r in recs[0::2])


# A random grammar (alpacc_amd.gen.randgrammar) whose fewest-symbols
# alternative of C ("d h D B") recurses through B: a derivation that takes
# the shortest alternative past its budget grows without end.
RECURSIVE_SHORTEST = """params {
  lookback = 2.
  lookahead = 1.
}

a = /c/.
b = /c/.
c = /[beeacebcd]/.
d = /([ceeeb]|(([ae] e) [ebc]))/.
e = /c/.
f = /c/.
g = /([dbdbb]|([bdb] ()))/.
h = /a/.

B -> f b A.
A -> c C.
C -> e b g d a.
D -> b e.
C -> d h D B.
C -> g B C D B.
"""


def test_derivations_end_when_the_shortest_alternative_recurses():
    """Past its budget a derivation takes alternatives of least height, so it
    ends (C -> e b g d a here) instead of growing until it is dropped, and
    the batch generator finishes."""
    import random
    from alpacc_amd.gen import tables
    g = tables.analyse("recursive_shortest", RECURSIVE_SHORTEST)
    for seed in range(20):
        toks = testvec.derivation(g, random.Random(seed), budget=5)
        assert toks is not None and len(toks) < 1000
    items = testvec.generate(g, "parser", derivations=60, seed=1)
    assert len(items) == 60 and all(items[0::2])


def test_cli_round_trip(tmp_path):
    req, out = tmp_path / "req.bin", tmp_path / "out.bin"
    assert testvec.main(["generate", "json", "--mode", "both", "--derivations", "10", "-o", str(req)]) == 0
    out.write_bytes(run_oracle("json", "both", req.read_bytes()))
    assert testvec.main(["compare", "json", "--mode", "both", str(out), str(out)]) == 0


@pytest.mark.gpu
@pytest.mark.parametrize("grammar", ["json", "arithmetic", "sexp"])
def test_generated_vectors_hip_vs_oracle(grammar):
    """alpacc-hip answers the tool's batches (exhaustive, DFA walks,
    derivations; every mode) as the oracle program does."""
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "alpacc_amd", "lib", "alpacc-hip")
    g = testvec.load(grammar)
    for mode, kw in (("lexer", {"exhaustive": 2}), ("both", {"exhaustive": 2}), ("lexer", {"long": 50000, "count": 3}),
                     ("both", {"derivations": 60}), ("parser", {"derivations": 60}), ("parser", {"exhaustive": 3})):
        lay = testvec.layout(g, mode)
        batch = wire.encode_batch(testvec.generate(g, mode, seed=11, **kw), lay)
        exp = run_oracle(grammar, mode, batch)
        flag = {"lexer": ["--lexer"], "parser": ["--parser"], "both": []}[mode]
        got = subprocess.run([cli, "-g", grammar] + flag, input=batch, capture_output=True, timeout=120)
        assert got.returncode == 0, got.stderr
        assert testvec.compare(exp, got.stdout, lay) == [], (grammar, mode, kw)
