"""HIP path vs the CPU oracle: response records must be byte-identical.

Every case goes through the C ABI (alpacc_hip_run_case) and is compared with
the oracle's record for the same request.  Covered: the known answers, every
short input over grammar alphabets, generated text around tile boundaries in
both variants (grammar-conforming and as-written), single-byte flips, empty
and whitespace-only inputs, length_t overflow, and 10,000-deep nesting.
"""

import itertools
import os
import random
import sys

import numpy as np
import pytest

import pyoracle
from alpacc_amd import datagen

pytestmark = pytest.mark.gpu

GRAMMARS = ["arithmetic", "json", "sexp", "paper_grammar"]
MODES = ["lexer", "parser", "both"]


@pytest.fixture(scope="module")
def backends():
    from alpacc_amd import hip
    out = {g: hip.Backend(g) for g in GRAMMARS}
    yield out
    for b in out.values():
        b.close()


def lexed_tokens(grammar, data: bytes) -> bytes:
    """Combined-numbering token ids of `data` (parser-mode payload), from the
    oracle's combined record; None when the text is rejected."""
    from alpacc_amd import wire
    rec = pyoracle.record(grammar, "both", data)
    if rec[0] == 0:
        return None
    n = int.from_bytes(rec[1:9], "little")
    return rec[9:9 + n]


def check(backends, grammar, mode, payload: bytes, label=""):
    print(f"case {grammar}/{mode} n={len(payload)} {label}", flush=True)
    got = backends[grammar].run_case(mode, payload)
    exp = pyoracle.record(grammar, mode, payload)
    assert got == exp, f"{grammar}/{mode} {label}: record mismatch (gpu {len(got)} B, oracle {len(exp)} B)"


def test_layout_matches_oracle(backends):
    import subprocess
    for g in GRAMMARS:
        for m in MODES:
            ref = subprocess.run([pyoracle.program(g, m), "--layout"], capture_output=True, text=True).stdout
            ours = backends[g].layout_text(m)
            ref_kv = dict(l.split("=") for l in ref.split())
            ours_kv = dict(l.split("=") for l in ours.split())
            assert {k: ours_kv[k] for k in ref_kv} == ref_kv


def test_kat_arithmetic(backends):
    check(backends, "arithmetic", "lexer", b"12+(3*4)")
    check(backends, "arithmetic", "parser", bytes([0, 5, 2, 0, 4, 0, 3]))
    check(backends, "arithmetic", "both", b"12+(3*4)")
    rec = backends["arithmetic"].run_case("both", b"12+(3*4)")
    assert rec[0] == 1


ALPHABETS = {
    "arithmetic": b"1-a(+ )*",
    "json": b'"a1[,]{:}-. tn',
    "sexp": b"a1 ()-",
    "paper_grammar": b"1+[] ",
}


@pytest.mark.parametrize("grammar", GRAMMARS)
def test_exhaustive_short_inputs(backends, grammar):
    alpha = ALPHABETS[grammar]
    maxn = 3 if len(alpha) > 8 else 4
    for n in range(0, maxn + 1):
        for tup in itertools.product(alpha, repeat=n):
            data = bytes(tup)
            for mode in ("lexer", "both"):
                got = backends[grammar].run_case(mode, data)
                assert got == pyoracle.record(grammar, mode, data), (mode, data)


@pytest.mark.parametrize("grammar", GRAMMARS)
def test_exhaustive_short_token_strings(backends, grammar):
    from alpacc_amd.gen import tables
    g = tables.load(os.path.join(os.path.dirname(os.path.dirname(__file__)), "grammars", grammar + ".alp"))
    ids = sorted(set(g.lex_both.term_id.values()) - {g.lex_both.ignore_id})
    # include the sentinels and the dead id as hostile inputs
    ids += [g.parser.empty_t, g.parser.start_t, g.parser.end_t]
    maxn = 3 if len(ids) > 8 else 4
    for n in range(0, maxn + 1):
        for tup in itertools.product(ids, repeat=n):
            payload = bytes(tup)
            assert backends[grammar].run_case("parser", payload) == pyoracle.record(grammar, "parser", payload), tup


SIZES = [1, 2, 3, 31, 32, 33, 8191, 8192, 8193, 16385, 65536, 100003, 1 << 20]


@pytest.mark.parametrize("grammar", ["arithmetic", "json", "sexp"])
@pytest.mark.parametrize("conforming", [True, False])
def test_generated_sizes(backends, grammar, conforming):
    for n in SIZES:
        data = datagen.generate(grammar, n, seed=n, conforming=conforming).tobytes()
        for mode in ("lexer", "both"):
            check(backends, grammar, mode, data, f"conf={conforming}")
        toks = lexed_tokens(grammar, data)
        if toks is not None:
            check(backends, grammar, "parser", toks, "lexed tokens")


@pytest.mark.parametrize("grammar", ["arithmetic", "json", "sexp"])
def test_byte_flips(backends, grammar):
    rng = random.Random(7)
    base = datagen.generate(grammar, 4096, seed=11, conforming=True).tobytes()
    for i in range(60):
        b = bytearray(base)
        b[rng.randrange(len(b))] = rng.randrange(256)
        check(backends, grammar, "both", bytes(b), f"flip {i}")
        check(backends, grammar, "lexer", bytes(b), f"flip {i}")


def test_edge_inputs(backends):
    for g in ["arithmetic", "json", "sexp"]:
        for data in [b"", b" ", b" " * 100000, b"\n\t\r ", b"(", b")", b"\xff", b"1" * 9000]:
            for mode in ("lexer", "both"):
                check(backends, g, mode, data, "edge")
        check(backends, g, "parser", b"", "empty tokens")
    # length_t = u16: a 70,000-byte string token rejects, a long ignore run does not
    check(backends, "json", "lexer", b'"' + b"a" * 70000 + b'"', "overflow")
    check(backends, "json", "both", b"[1," + b" " * 70000 + b"2]", "long ignore")


def test_deep_nesting(backends):
    depth = 10000
    data = b"(" * depth + b"a" + b")" * depth
    check(backends, "sexp", "both", data, "depth 10000")
    check(backends, "sexp", "both", data[:-1], "unbalanced")
    arr = b"[" * 30 + b"1" + b"]" * 30
    check(backends, "json", "both", arr, "json depth 30")
    expr = b"(" * 64 + b"x" + b")" * 64
    check(backends, "arithmetic", "both", expr, "parens 64")


def nested_text(grammar, rng, budget, depth_cap):
    """Random nesting with long runs of opens and closes, so matches land at
    every distance: inside a lane, lanes apart, tiles apart."""
    out = []
    depth = 0
    opn, cls, sep, atom = {"json": ("[", "]", ",", "1"), "sexp": ("(", ")", " ", "a"),
                           "arithmetic": ("(", ")", "+", "1")}[grammar]
    need_item = True
    while len(out) < budget:
        r = rng.random()
        if r < 0.45 and depth < depth_cap:
            if not need_item:
                out.append(sep)
            for _ in range(rng.choice([1, 1, 2, 5, 40])):
                if depth >= depth_cap:
                    break
                out.append(opn)
                depth += 1
            need_item = True
        elif r < 0.75 or depth == 0:
            if not need_item:
                out.append(sep)
            out.append(atom)
            need_item = False
        else:
            if need_item:
                out.append(atom)
            for _ in range(min(depth, rng.choice([1, 1, 3, 40]))):
                out.append(cls)
                depth -= 1
            need_item = False
    if need_item:
        out.append(atom)
    out.append(cls * depth)
    return "".join(out).encode()


@pytest.mark.parametrize("grammar", ["json", "sexp", "arithmetic"])
def test_random_nesting(backends, grammar):
    rng = random.Random(5)
    for i, (budget, cap) in enumerate([(3000, 50), (20000, 300), (200000, 2000), (400000, 60)]):
        data = nested_text(grammar, rng, budget, cap)
        if grammar == "sexp":
            data = b"(" + data + b")"
        elif grammar == "json":
            data = b"[" + data + b"]"
        check(backends, grammar, "both", data, f"nesting {i}")
        rec = backends[grammar].run_case("both", data)
        assert rec[0] == 1, f"nesting {i} should be accepted"


@pytest.mark.parametrize("grammar,n", [("json", 64 << 20), ("sexp", 32 << 20), ("arithmetic", 16 << 20)])
def test_large_conforming(backends, grammar, n):
    data = datagen.generate(grammar, n, seed=3, conforming=True).tobytes()
    check(backends, grammar, "both", data, "large")
    check(backends, grammar, "lexer", data, "large")


def test_cli_batch_and_server_replay():
    """The wire program over the HIP library answers a mixed batch exactly as
    the oracle program does, and --server fed the batch twice answers twice
    (tests/testc.sh:102-112)."""
    import subprocess
    from alpacc_amd import wire
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "alpacc_amd", "lib", "alpacc-hip")
    for grammar in ("arithmetic", "json", "sexp"):
        payloads = [b"", b" "] + [datagen.generate(grammar, n, seed=n, conforming=n % 3 != 0).tobytes()
                                  for n in (1, 17, 300, 5000, 70000)]
        for mode in ("lexer", "both"):
            batch = wire.encode_batch(payloads, wire.Layout(mode, 1, 8, 2, 1))
            flag = ["--lexer"] if mode == "lexer" else []
            ours = subprocess.run([cli, "-g", grammar] + flag, input=batch, capture_output=True, timeout=120)
            ref = subprocess.run([pyoracle.program(grammar, mode)], input=batch, capture_output=True, timeout=120)
            assert ours.returncode == 0 and ours.stdout == ref.stdout, (grammar, mode, ours.stderr)
            srv = subprocess.run([cli, "-g", grammar, "--server"] + flag, input=batch + batch, capture_output=True,
                                 timeout=120)
            assert srv.returncode == 0 and srv.stdout == ref.stdout + ref.stdout, (grammar, mode)
