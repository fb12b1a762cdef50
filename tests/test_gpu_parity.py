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
import zlib

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


def first_diff(a: bytes, b: bytes):
    """None when a == b, else the first byte offset where they differ (keeps
    failure messages short for records of megabytes)."""
    if a == b:
        return None
    n = min(len(a), len(b))
    return next((i for i in range(n) if a[i] != b[i]), n)


def check(backends, grammar, mode, payload: bytes, label=""):
    print(f"case {grammar}/{mode} n={len(payload)} {label}", flush=True)
    got = backends[grammar].run_case(mode, payload)
    exp = pyoracle.record(grammar, mode, payload)
    d = first_diff(got, exp)
    assert d is None, f"{grammar}/{mode} {label}: record mismatch at byte {d} (gpu {len(got)} B, oracle {len(exp)} B)"


def test_parser_ids_past_the_terminals(backends):
    """Parser-mode strings holding ids past the grammar's terminals (the
    window codes they form lie past the window table, whose entries pass A
    reads with bounds-checked loads) are rejected exactly as the oracle
    rejects them, alone and in a batch beside valid strings."""
    for grammar in ("json", "arithmetic", "sexp"):
        strings = [bytes([200]), bytes([255] * 10), bytes([4, 1, 200, 1, 5, 0]), bytes(range(15, 60)),
                   bytes([0, 1, 2, 3] * 8 + [250])]
        for x in strings:
            check(backends, grammar, "parser", x, "ids past the terminals")
        toks = lexed_tokens(grammar, datagen.generate(grammar, 3000, seed=2, conforming=True).tobytes())
        batch = [toks, strings[2], toks[:40] + bytes([201]) + toks[40:], toks]
        exp = b"".join(pyoracle.record(grammar, "parser", x) for x in batch)
        assert first_diff(backends[grammar].run_batch("parser", batch), exp) is None, grammar


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


def test_dense_and_spanning_tokens(backends):
    """A token at every byte (tiles of 32 KiB hold up to 32768 tokens, far
    more than one staging round of the lexer) and tokens spanning tiles."""
    n = 1 << 20
    check(backends, "json", "both", b"[" + b"1," * (n // 2) + b"1]", "dense json")
    check(backends, "json", "lexer", b"[" * n, "open brackets")
    check(backends, "sexp", "both", b"(" + b"a " * (n // 2) + b")", "dense sexp")
    check(backends, "arithmetic", "both", b"1" + b"+1" * (n // 2), "dense arithmetic")
    check(backends, "arithmetic", "lexer", b"1+" * (n // 2), "dense arithmetic, trailing op")
    # dense and sparse tiles alternating
    parts = []
    for i in range(40):
        parts.append(b"1," * 20000 if i % 2 else b'"' + b"a" * 60000 + b'",')
    check(backends, "json", "both", b"[" + b"".join(parts) + b"1]", "alternating density")
    # strings of 65,535 bytes in total are the longest a u16 length holds
    for ln in (65533, 65534):
        data = b"[" + (b'"' + b"x" * ln + b'",') * 5 + b"1]"
        check(backends, "json", "lexer", data, f"string {ln + 2} bytes")


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
            assert ours.returncode == 0, (grammar, mode, ours.stderr)
            d = first_diff(ours.stdout, ref.stdout)
            assert d is None, f"{grammar}/{mode}: batch output differs at byte {d}"
            srv = subprocess.run([cli, "-g", grammar, "--server"] + flag, input=batch + batch, capture_output=True,
                                 timeout=120)
            assert srv.returncode == 0, (grammar, mode, srv.stderr)
            d = first_diff(srv.stdout, ref.stdout + ref.stdout)
            assert d is None, f"{grammar}/{mode}: server output differs at byte {d}"


def batch_payloads(grammar, mode, rng, count):
    """A mixed batch: empty, tiny, tile-boundary and large inputs, valid and
    rejected, at unaligned offsets when packed."""
    out = []
    for i in range(count):
        kind = rng.random()
        if kind < 0.05:
            n = 0
        elif kind < 0.5:
            n = rng.randrange(1, 200)
        elif kind < 0.9:
            n = rng.randrange(200, 20000)
        else:
            n = rng.choice([32767, 32768, 32769, 65537, 200000])
        data = datagen.generate(grammar, n, seed=rng.randrange(1 << 30), conforming=rng.random() < 0.8).tobytes() if n else b""
        if data and rng.random() < 0.1:
            b = bytearray(data)
            b[rng.randrange(len(b))] = rng.randrange(256)
            data = bytes(b)
        if mode == "parser":
            toks = lexed_tokens(grammar, data)
            data = toks if toks is not None else bytes(rng.randrange(8) for _ in range(rng.randrange(12)))
        out.append(data)
    return out


@pytest.mark.parametrize("grammar", ["json", "arithmetic", "sexp"])
@pytest.mark.parametrize("mode", MODES)
def test_batch_records(backends, grammar, mode):
    """alpacc_hip_run_batch answers every input of a batch exactly as the
    oracle answers it alone."""
    rng = random.Random(zlib.crc32(f"{grammar}/{mode}".encode()))
    for count in (1, 3, 57):
        payloads = batch_payloads(grammar, mode, rng, count)
        print(f"batch {grammar}/{mode} count={count} bytes={sum(map(len, payloads))}", flush=True)
        got = backends[grammar].run_batch(mode, payloads)
        exp = b"".join(pyoracle.record(grammar, mode, x) for x in payloads)
        d = first_diff(got, exp)
        assert d is None, f"{grammar}/{mode} batch of {count}: differs at byte {d}"


@pytest.mark.parametrize("grammar", ["json", "arithmetic"])
@pytest.mark.parametrize("mode", ["lexer", "both"])
def test_batch_many_small(backends, grammar, mode):
    """Hundreds of small inputs share each 256-chunk tile: sizes at the
    128-byte chunk edges, empty inputs in runs, rejected inputs (flips, a
    token longer than length_t) between accepted ones."""
    rng = random.Random(zlib.crc32(f"small {grammar}/{mode}".encode()))
    sizes = [0, 0, 1, 2, 127, 128, 129, 255, 256, 257, 383, 384, 385] + [rng.randrange(0, 400) for _ in range(900)]
    rng.shuffle(sizes)
    payloads = []
    for n in sizes:
        data = datagen.generate(grammar, n, seed=rng.randrange(1 << 30), conforming=rng.random() < 0.8).tobytes() if n else b""
        if data and rng.random() < 0.08:
            b = bytearray(data)
            b[rng.randrange(len(b))] = rng.randrange(256)
            data = bytes(b)
        payloads.append(data)
    if grammar == "json":  # a string token longer than 65535 bytes rejects its input only
        payloads.insert(300, b'"' + b"a" * 70000 + b'"')
    print(f"batch {grammar}/{mode} count={len(payloads)} bytes={sum(map(len, payloads))}", flush=True)
    got = backends[grammar].run_batch(mode, payloads)
    exp = b"".join(pyoracle.record(grammar, mode, x) for x in payloads)
    d = first_diff(got, exp)
    assert d is None, f"{grammar}/{mode} batch of {len(payloads)} small inputs: differs at byte {d}"


@pytest.mark.parametrize("mode", ["lexer", "both"])
def test_batch_token_closing_at_tile_start(backends, mode):
    """A token whose last byte ends a tile that the next inputs share: its
    input is the tile's first, found from the tile start."""
    cases = {"json": b"[" + b"1," * 20000 + b"1]", "arithmetic": b"1" + b"+1" * 20000}
    for grammar, dense in cases.items():
        small = [b"", b"1", b"[1]" if grammar == "json" else b"(2)", b""]
        payloads = [dense] + small + [dense[:33000]] + small
        got = backends[grammar].run_batch(mode, payloads)
        exp = b"".join(pyoracle.record(grammar, mode, x) for x in payloads)
        d = first_diff(got, exp)
        assert d is None, f"{grammar}/{mode}: differs at byte {d}"


def test_batch_graph_replays():
    """Combined batches through the captured graph: the first (largest)
    batch captures it, smaller ones replay it with eager calls on the same
    context between the replays, and every record of every replay equals
    the oracle's."""
    from alpacc_amd import hip
    for grammar in ("json", "arithmetic"):
        rng = random.Random(zlib.crc32(f"graph {grammar}".encode()))
        be = hip.Backend(grammar)
        try:
            be.set_graphs(True)
            for k, count in enumerate((60, 20, 45, 7, 60, 1)):
                payloads = batch_payloads(grammar, "both", rng, count)
                if k == 0:
                    payloads.append(datagen.generate(grammar, 300000, seed=5, conforming=True).tobytes())
                print(f"graph batch {grammar} #{k} count={len(payloads)} bytes={sum(map(len, payloads))}", flush=True)
                got = be.run_batch("both", payloads)
                exp = b"".join(pyoracle.record(grammar, "both", x) for x in payloads)
                d = first_diff(got, exp)
                assert d is None, f"{grammar} graph batch {k}: differs at byte {d}"
                x = payloads[-1][:5000]
                assert be.run_case("both", x) == pyoracle.record(grammar, "both", x)
            captures, replays = be.graph_stats()
            print(f"{grammar}: {captures} captures, {replays} replays", flush=True)
            assert captures >= 1 and replays >= 3
        finally:
            be.close()


def test_batch_device_graph_replays():
    """lex_parse_batch_device through the graph: batches read from one device
    buffer replay the captured sequence, and every replay's device arrays
    match the oracle."""
    import torch
    from alpacc_amd import hip
    rng = random.Random(21)
    be = hip.Backend("json")
    try:
        be.set_graphs(True)
        dev = torch.zeros(4 << 20, dtype=torch.uint8, device="cuda")
        for k, n in enumerate((40, 12, 33, 12, 40)):
            payloads = batch_payloads("json", "both", rng, n)
            if k == 0:  # the capacity covers the later batches
                payloads.append(datagen.generate("json", 1500000, seed=6, conforming=True).tobytes())
            buf = b"".join(payloads)
            assert len(buf) <= dev.numel()
            dev[:len(buf)] = torch.frombuffer(bytearray(buf), dtype=torch.uint8).to("cuda")
            torch.cuda.synchronize()
            print(f"device graph batch #{k} count={n} bytes={len(buf)}", flush=True)
            check_device_batch(be, dev, payloads)
        captures, replays = be.graph_stats()
        assert captures >= 1 and replays >= 2, (captures, replays)
    finally:
        be.close()


SHAPES = [(64, 2, 16), (64, 0, 64), (128, 3, 1), (128, 4, 32)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"ipt{s[0]}-wpc{s[1]}-win{s[2]}")
def test_lexer_shapes(shape):
    """Every lexer launch shape the tuner can pick gives the oracle's
    records: generated text around chunk and tile edges, flips, dense
    tokens and a batch of many small inputs."""
    from alpacc_amd import hip
    ipt, wpc, win = shape
    bes = {}
    try:
        for g in ("json", "arithmetic", "sexp"):
            bes[g] = hip.Backend(g)
            bes[g].set_lex_shape(ipt, wpc, win)
            assert bes[g].lex_shape() == {"items_per_thread": ipt, "workgroups_per_cu": wpc, "lookback_window": win}
        for g in ("json", "arithmetic", "sexp"):
            for n in (1, 63, 64, 65, 127, 128, 129, 16383, 16384, 16385, 100003, 1 << 20):
                data = datagen.generate(g, n, seed=n, conforming=n % 2 == 0).tobytes()
                for mode in ("lexer", "both"):
                    check(bes, g, mode, data, f"shape {shape}")
        check(bes, "json", "both", b"[" + b"1," * (1 << 19) + b"1]", f"dense, shape {shape}")
        rng = random.Random(zlib.crc32(f"shape {shape}".encode()))
        payloads = [datagen.generate("json", rng.randrange(0, 300), seed=i, conforming=True).tobytes() for i in range(700)]
        got = bes["json"].run_batch("both", payloads)
        exp = b"".join(pyoracle.record("json", "both", x) for x in payloads)
        assert first_diff(got, exp) is None, f"batch of small inputs, shape {shape}"
    finally:
        for b in bes.values():
            b.close()


def test_autotune_lexer():
    """The tuner picks a shape of its search space from a device-resident
    sample, keeps it in the context, and results stay the oracle's."""
    import torch
    from alpacc_amd import hip
    be = hip.Backend("json")
    try:
        data = datagen.generate("json", 4 << 20, seed=9, conforming=True).tobytes()
        dev = torch.frombuffer(bytearray(data), dtype=torch.uint8).to("cuda")
        torch.cuda.synchronize()
        best = be.autotune_lexer(dev.data_ptr(), len(data), "both", reps=2)
        print("tuned shape", best, flush=True)
        assert best["items_per_thread"] in (64, 128) and best["workgroups_per_cu"] in (0, 2, 3, 4)
        assert best["lookback_window"] in (4, 8, 16, 32, 64) and best["ms"] > 0
        assert be.lex_shape() == {k: best[k] for k in ("items_per_thread", "workgroups_per_cu", "lookback_window")}
        for n in (5000, 300000):
            x = data[:n]
            for mode in ("lexer", "both"):
                got = be.run_case(mode, x)
                assert first_diff(got, pyoracle.record("json", mode, x)) is None
        with pytest.raises(hip.HipError):
            be.set_lex_shape(96, 0, 64)
    finally:
        be.close()


def test_bracket_widths():
    """Every bracket stack width of the parser's resolve kernels gives the
    oracle's records: one-byte and two-byte stacks, eager single inputs and
    batches, and the graph path's striding kernels; a width the grammar's
    symbols do not fit, or not 0/1/2, is refused."""
    from alpacc_amd import hip
    for grammar in ("arithmetic", "json", "sexp"):  # the grammars with an input generator
        inputs = [datagen.generate(grammar, n, seed=n + 5, conforming=n % 3 != 1).tobytes()
                  for n in (0, 3, 1000, 4097, 70000, 1 << 20, 3 << 20)]
        exp = [pyoracle.record(grammar, "both", x) for x in inputs]
        be = hip.Backend(grammar)
        try:
            for width in (2, 1, 0):
                print(f"bracket width {grammar} {width}", flush=True)
                be.set_bracket_width(width)
                for x, e in zip(inputs, exp):
                    assert first_diff(be.run_case("both", x), e) is None, (grammar, width, len(x))
                be.set_graphs(False)
                assert first_diff(be.run_batch("both", inputs), b"".join(exp)) is None, (grammar, width, "batch")
                be.set_graphs(True)
                for _ in range(2):
                    assert first_diff(be.run_batch("both", inputs), b"".join(exp)) is None, (grammar, width, "graph")
                be.set_graphs(False)
            for bad in (3, -1):
                with pytest.raises(hip.HipError):
                    be.set_bracket_width(bad)
        finally:
            be.close()


@pytest.mark.parametrize("index32,length", [(True, 0), (False, 8), (True, 64)])
def test_width_variants(index32, length):
    """Contexts opened as width variants (--index32, length) answer every
    request as the oracle built with the same widths: single requests
    (including tokens too long for a narrow length_t), a batch, the layout."""
    import subprocess
    from alpacc_amd import hip
    for grammar in ("json", "arithmetic"):
        be = hip.Backend(grammar, index32=index32, length=length)
        try:
            kv = dict(l.split("=") for l in be.layout_text("both").split())
            assert int(kv["index_t"]) == (4 if index32 else 8)
            long_tok = b'"' + b"a" * 300 + b'"' if grammar == "json" else b"x" * 300 + b"+1"
            inputs = [datagen.generate(grammar, n, seed=n, conforming=n % 2 == 0).tobytes()
                      for n in (0, 1, 17, 300, 5000, 60000, 1 << 20)] + [long_tok]
            for mode in ("lexer", "both"):
                for x in inputs:
                    got = be.run_case(mode, x)
                    exp = pyoracle.record(grammar, mode, x, index32=index32, length=length)
                    assert first_diff(got, exp) is None, (grammar, mode, index32, length, len(x))
                got = be.run_batch(mode, inputs)
                exp = b"".join(pyoracle.record(grammar, mode, x, index32=index32, length=length) for x in inputs)
                assert first_diff(got, exp) is None, (grammar, mode, "batch")
        finally:
            be.close()
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "alpacc_amd", "lib", "alpacc-hip")
    flags = (["--index32"] if index32 else []) + (["--length", str(length)] if length else [])
    out = subprocess.run([cli, "-g", "json", "--layout"] + flags, capture_output=True, text=True, timeout=60).stdout
    assert f"index_t={4 if index32 else 8}" in out
    assert f"length_t={length // 8 if length else 2}" in out


@pytest.mark.parametrize("grammar", ["json", "arithmetic", "sexp"])
def test_lex_stream(backends, grammar):
    """One input pushed as chunks (cut at random points: inside tokens, at
    boundaries, 1-byte and empty chunks, an empty last chunk) gives the
    tokens of the whole input: the oracle's lexer record, also for rejected
    inputs and tokens spanning many chunks."""
    rng = random.Random(zlib.crc32(f"stream {grammar}".encode()))
    be = backends[grammar]
    texts = [datagen.generate(grammar, n, seed=n, conforming=n % 2 == 0).tobytes() for n in (1, 2, 100, 5000, 200000)]
    texts += [b"", b" " * 5000]
    if grammar == "json":
        texts += [b'"' + b"a" * 70000 + b'"', b"[" + b"1," * 30000 + b"1]", b'["' + b"x" * 40000 + b'", 1]']
    for x in texts:
        for trial in range(4):
            cuts = sorted(rng.randrange(0, len(x) + 1) for _ in range(rng.choice([0, 1, 3, 12])))
            if trial == 3 and len(x) <= 400:
                cuts = list(range(len(x) + 1))  # one byte per chunk
            pieces = [x[a:b] for a, b in zip([0] + cuts, cuts + [len(x)])]
            if trial % 2:
                pieces.append(b"")  # the end told by an empty last chunk
            for mode in ("lexer", "both"):
                got = be.lex_stream(pieces, mode)
                exp = pyoracle.record(grammar, mode, x) if mode == "lexer" else _lexer_part(grammar, x)
                if exp is None:  # lexed but not parsed: no combined-numbering tokens to compare
                    continue
                assert first_diff(got, exp) is None, (grammar, mode, len(x), [len(p) for p in pieces][:20])


def _lexer_part(grammar, x):
    """The combined-mode tokens of x as a lexer-only record (the combined
    oracle record's token arrays), or the reject record."""
    from alpacc_amd import wire
    g_lay = wire.Layout("both", 1, 8, 2 if grammar in ("json", "sexp") else 8, 1)
    rec = pyoracle.record(grammar, "both", x)
    r, _ = wire.decode_record(rec, 0, g_lay)
    if not r.valid:
        lex = pyoracle.record(grammar, "lexer", x)
        return b"\x00" if lex[0] == 0 else None
    return wire.encode_record(wire.Record(True, terminals=r.terminals, starts=r.starts, lengths=r.lengths),
                              wire.Layout("lexer", 1, 8, g_lay.length))


def test_sexp_256mib_and_flip_variants(backends):
    """The sexp workload at full size: one 256 MiB input (geometric nesting,
    1% of lists 10,000 deep) lexed and parsed as the oracle does, and 1,000
    variants of a 64 KiB sample with one random byte changed, as one batch:
    every verdict and tree as the oracle's."""
    n = 256 << 20
    data = datagen.generate("sexp", n, seed=2, conforming=True).tobytes()
    print(f"case sexp/both n={n} full size", flush=True)
    check(backends, "sexp", "both", data, "256 MiB")
    rng = random.Random(2)
    sample = data[:1 << 16]
    flips = []
    for _ in range(1000):
        b = bytearray(sample)
        b[rng.randrange(len(b))] = rng.randrange(256)
        flips.append(bytes(b))
    got = backends["sexp"].run_batch("both", flips)
    exp = b"".join(pyoracle.record("sexp", "both", x) for x in flips)
    assert first_diff(got, exp) is None, "sexp flip variants"


def test_json_1gib_lexer_only(backends):
    """The JSON lexer-only workload at full size: 1 GiB (seed 1) lexed with
    the lexer-only numbering, record identical to the oracle's (1.5 GB)."""
    n = 1 << 30
    data = datagen.generate("json", n, seed=1, conforming=True).tobytes()
    print(f"case json/lexer n={n} full size", flush=True)
    got = backends["json"].run_case("lexer", data)
    exp = pyoracle.record("json", "lexer", data)
    assert len(got) == len(exp) and got == exp, "json 1 GiB lexer-only record"


def test_json_1gib_literal_variant(backends):
    """The JSON workload as written (string escapes, exponents: the grammar
    rejects them) at 1 GiB, both modes: the verdicts equal the oracle's."""
    n = 1 << 30
    data = datagen.generate("json", n, seed=1, conforming=False, wrap=True).tobytes()
    for mode in ("lexer", "both"):
        print(f"case json/{mode} n={n} literal", flush=True)
        got = backends["json"].run_case(mode, data)
        exp = pyoracle.record("json", mode, data)
        assert got == exp, f"json 1 GiB literal {mode}"


def test_arithmetic_1mib_seed0(backends):
    """The plain-method workload exactly: arithmetic, 1 MiB, seed 0, both
    variants, every mode."""
    for conforming in (True, False):
        data = datagen.generate("arithmetic", 1 << 20, seed=0, conforming=conforming).tobytes()
        for mode in ("lexer", "both"):
            check(backends, "arithmetic", mode, data, f"seed 0 conf={conforming}")
        toks = lexed_tokens("arithmetic", data)
        if toks is not None:
            check(backends, "arithmetic", "parser", toks, "seed 0 lexed tokens")


def test_json_1gib_lex_parse(backends):
    """The measured workload at full size: 1 GiB of JSON (seed 1) lexed and
    parsed, tokens and tree identical to the oracle's (about 3.8 GB of
    record)."""
    n = 1 << 30
    data = datagen.generate("json", n, seed=1, conforming=True, wrap=True).tobytes()
    print(f"case json/both n={n} full size", flush=True)
    got = backends["json"].run_case("both", data)
    exp = pyoracle.record("json", "both", data)
    assert got[0] == 1 and len(got) == len(exp) and got == exp, "json 1 GiB lex+parse record"


def test_batching_workload_full(backends):
    """The batching workload at full size: 20,000 lognormal JSON inputs
    (bench.py's batch, seed 3) in one batched call, every record identical
    to the oracle's for that input alone."""
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    host, offs = bench.batch_inputs(20000, 3)
    payloads = [host[int(offs[i]):int(offs[i + 1])].tobytes() for i in range(len(offs) - 1)]
    print(f"batch json/both count={len(payloads)} bytes={int(offs[-1])} full size", flush=True)
    got = backends["json"].run_batch("both", payloads)
    exp = b"".join(pyoracle.record("json", "both", x) for x in payloads)
    d = first_diff(got, exp)
    assert d is None, f"batching workload differs at byte {d}"


def test_batch_alone_and_in_mix(backends):
    """Each result is the same alone and in a mix: every input of a batch
    against run_case on that input."""
    rng = random.Random(99)
    payloads = batch_payloads("json", "both", rng, 40)
    be = backends["json"]
    got = be.run_batch("both", payloads)
    alone = b"".join(be.run_case("both", x) for x in payloads)
    assert first_diff(got, alone) is None


def test_batch_device_arrays(backends):
    import torch
    rng = random.Random(3)
    payloads = batch_payloads("json", "both", rng, 25)
    dev = torch.from_numpy(np.frombuffer(b"".join(payloads) + b"\0", dtype=np.uint8).copy()).to("cuda")
    check_device_batch(backends["json"], dev, payloads)


def check_device_batch(be, dev, payloads):
    """lex_parse_batch_device over payloads packed at the start of the device
    tensor dev: every input's device arrays against the oracle's record."""
    import torch
    from alpacc_amd import wire
    offs = np.zeros(len(payloads) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum([len(x) for x in payloads])
    b = be.lex_parse_batch_device(dev.data_ptr(), offs)
    torch.cuda.synchronize()
    lay = wire.Layout("both", 1, 8, 2, 1)
    valid = np.ctypeslib.as_array((C_u8 * len(payloads)).from_address(b.valid))
    tok_off = np.ctypeslib.as_array((C_u64 * (len(payloads) + 1)).from_address(b.token_offsets))
    node_off = np.ctypeslib.as_array((C_u64 * len(payloads)).from_address(b.node_offsets))
    nodes = np.ctypeslib.as_array((C_u64 * len(payloads)).from_address(b.node_counts))
    tp_off = np.ctypeslib.as_array((C_u64 * len(payloads)).from_address(b.token_parent_offsets))
    T = int(tok_off[-1])
    N = int(max([node_off[i] + nodes[i] for i in range(len(payloads))] + [0]))
    for i, x in enumerate(payloads):
        rec, _ = wire.decode_record(pyoracle.record("json", "both", x), 0, lay)
        assert bool(valid[i]) == rec.valid, i
        if not rec.valid:
            continue
        a, n = int(tok_off[i]), int(tok_off[i + 1] - tok_off[i])
        assert n == len(rec.terminals)
        got_t = _dev_bytes(b.tokens.terminals + a, n)
        got_s = np.frombuffer(_dev_bytes(b.tokens.starts + 8 * a, 8 * n), dtype=np.int64)
        got_p = _dev_bytes(b.productions + int(node_off[i]), int(nodes[i]))
        got_par = np.frombuffer(_dev_bytes(b.parents + 8 * int(node_off[i]), 8 * int(nodes[i])), dtype=np.int64)
        got_tp = np.frombuffer(_dev_bytes(b.token_parents + 8 * int(tp_off[i]), 8 * n), dtype=np.int64)
        assert np.array_equal(np.frombuffer(got_t, np.uint8), rec.terminals), i
        assert np.array_equal(got_s, rec.starts), i
        assert np.array_equal(np.frombuffer(got_p, np.uint8), rec.nodes), i
        assert np.array_equal(got_par, rec.parents), i
        assert np.array_equal(got_tp, rec.token_parents), i


import ctypes as _ct

C_u8 = _ct.c_uint8
C_u64 = _ct.c_uint64


class _DevView:
    """A device byte range as a __cuda_array_interface__ object."""

    def __init__(self, ptr, n):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": "|u1", "data": (ptr, False), "version": 2,
                                         "strides": None}


def _dev_bytes(ptr, n):
    import torch
    if n == 0:
        return b""
    return torch.as_tensor(_DevView(ptr, n), device="cuda").cpu().numpy().tobytes()


def serve_trace(rng, count, rate, max_bytes=1 << 20):
    """A mixed request trace: 70% json / 30% arithmetic, lognormal sizes
    (median 2 KiB, sigma 2), 3% with one random byte flipped, Poisson
    arrivals at `rate` per second."""
    import math
    names = ["json", "arithmetic"]
    gidx, modes, payloads, arrival = [], [], [], []
    t = 0.0
    for i in range(count):
        g = 0 if rng.random() < 0.7 else 1
        n = int(min(max(rng.lognormvariate(math.log(2048), 2.0), 128), max_bytes))
        data = bytearray(datagen.generate(names[g], n, seed=rng.randrange(1 << 30), conforming=True).tobytes())
        if rng.random() < 0.03:
            data[rng.randrange(len(data))] = rng.randrange(256)
        t += rng.expovariate(rate)
        gidx.append(g)
        modes.append("both" if rng.random() < 0.8 else "lexer")
        payloads.append(bytes(data))
        arrival.append(int(t * 1e9))
    return names, gidx, modes, payloads, arrival


@pytest.mark.parametrize("graphs", [False, True])
def test_server_replay_matches_alone_and_oracle(backends, graphs):
    """Continuous batching: every answer of a mixed open-loop trace equals
    the oracle's and the answer to the request alone (also with each lane's
    combined batches replaying its captured graph, twice over the trace)."""
    from alpacc_amd import hip
    rng = random.Random(11)
    names, gidx, modes, payloads, arrival = serve_trace(rng, 400, 5000.0)
    srv = hip.Server(names, lanes=2)
    try:
        srv.set_graphs(graphs)
        recs, lat = srv.replay(gidx, modes, payloads, arrival)
        recs2, _ = srv.replay(gidx, modes, payloads, arrival) if graphs else (recs, None)
        stats = srv.stats()
    finally:
        srv.close()
    exp = b"".join(pyoracle.record(names[g], m, x) for g, m, x in zip(gidx, modes, payloads))
    assert first_diff(recs, exp) is None
    assert first_diff(recs2, exp) is None
    alone = b"".join(backends[names[g]].run_case(m, x) for g, m, x in zip(gidx, modes, payloads))
    assert first_diff(recs, alone) is None
    assert stats["requests"] == len(payloads) * (2 if graphs else 1)
    print("server stats", stats, "p50 ms", float(np.percentile(lat, 50)) / 1e6, flush=True)


def test_server_submit_wait_any_order():
    from alpacc_amd import hip
    srv = hip.Server(["sexp", "json"], lanes=2)
    try:
        items = [("sexp", "both", b"(a (b c))"), ("json", "lexer", b'[1, "x"]'), ("json", "both", b"[1,"),
                 ("sexp", "parser", bytes([1, 0, 2])), ("json", "both", b"")]
        tickets = [srv.submit(g, m, x) for g, m, x in items]
        got = {t: srv.wait(t) for t in reversed(tickets)}
        for t, (g, m, x) in zip(tickets, items):
            assert got[t] == pyoracle.record(g, m, x), (g, m, x)
    finally:
        srv.close()


def test_server_bench_trace_full_rate():
    """The benchmark's own serving trace at its highest rate (50,000 req/s
    for one second, sizes lognormal(ln 2048, 2) clipped to 16 MiB, 70% JSON /
    30% arithmetic, 3% byte flips), replayed open-loop against the
    continuous-batching server: every record equals the oracle's."""
    import bench
    from alpacc_amd import hip
    gidx, payloads, arrival = bench.serve_trace(50000, 50000.0, 4 + 50000)
    names = ["json", "arithmetic"]
    modes = ["both"] * len(payloads)
    print(f"case bench serving trace: {len(payloads)} requests, {sum(map(len, payloads))} bytes", flush=True)
    srv = hip.Server(names, lanes=2, hbm_budget=64 << 30)
    try:
        recs, lat = srv.replay(gidx, modes, payloads, arrival)
    finally:
        srv.close()
    exp = b"".join(pyoracle.record(names[g], "both", x) for g, x in zip(gidx, payloads))
    assert first_diff(recs, exp) is None
    print("p50 ms", float(np.percentile(lat, 50)) / 1e6, "p99 ms", float(np.percentile(lat, 99)) / 1e6, flush=True)


PARSE_SHAPES = [(1024, 4), (512, 4), (2048, 4), (1024, 2), (1024, 8)]


@pytest.mark.parametrize("shape", PARSE_SHAPES, ids=lambda s: f"pt{s[0]}-apl{s[1]}")
def test_parser_shapes(shape):
    """Every parser tile shape the tuner can pick gives the oracle's records:
    generated text around tile edges in all three grammars, deep nesting,
    parser-mode token strings with bad ids, a batch of many small inputs
    eagerly and through the graph, and the widest grammar windows."""
    from alpacc_amd import hip
    import test_gpu_grammars as TG
    pt, apl = shape
    bes = {}
    try:
        for g in ("json", "arithmetic", "sexp"):
            bes[g] = hip.Backend(g)
            bes[g].set_parse_shape(pt, apl)
            assert bes[g].parse_shape() == {"positions_per_tile": pt, "positions_per_lane": apl}
        for g in ("json", "arithmetic", "sexp"):
            for n in (1, 100, 4000, 8191, 8193, 16385, 100003, 1 << 20):
                data = datagen.generate(g, n, seed=n + pt, conforming=n % 3 != 0).tobytes()
                check(bes, g, "both", data, f"parser shape {shape}")
                toks = lexed_tokens(g, data)
                if toks is not None:
                    check(bes, g, "parser", toks, f"parser shape {shape}")
        check(bes, "sexp", "both", b"(" * 10000 + b"a" + b")" * 10000, f"deep, parser shape {shape}")
        check(bes, "json", "parser", bytes([4, 1, 2, 1, 14, 5, 0, 13]), f"bad ids, parser shape {shape}")
        rng = random.Random(zlib.crc32(f"parser shape {shape}".encode()))
        payloads = [datagen.generate("json", rng.randrange(0, 3000), seed=i, conforming=i % 5 != 0).tobytes()
                    for i in range(500)]
        exp = b"".join(pyoracle.record("json", "both", x) for x in payloads)
        assert first_diff(bes["json"].run_batch("both", payloads), exp) is None, f"batch, parser shape {shape}"
        bes["json"].set_graphs(True)
        for rep in range(2):
            assert first_diff(bes["json"].run_batch("both", payloads), exp) is None, f"graph {rep}, shape {shape}"
        bes["json"].set_graphs(False)
    finally:
        for b in bes.values():
            b.close()
    for name in ("wide_windows", "many_symbols", "many_productions"):
        TG.load_tables(name)
        g = TG.grammar_of(name)
        items = TG.testvec.generate(g, "both", derivations=200, seed=5)
        exp = b"".join(pyoracle.record(name, "both", x) for x in items)
        with hip.Backend(name) as be:
            be.set_parse_shape(pt, apl)
            assert first_diff(be.run_batch("both", items), exp) is None, (name, shape)


def test_autotune_parser():
    """The parser tuner picks one of the shapes on a device-resident sample,
    keeps it, and records stay the oracle's."""
    import torch
    from alpacc_amd import hip
    data = datagen.generate("json", 8 << 20, seed=9, conforming=True).tobytes()
    dev = torch.frombuffer(bytearray(data), dtype=torch.uint8).to("cuda")
    torch.cuda.synchronize()
    with hip.Backend("json") as be:
        best = be.autotune_parser(dev.data_ptr(), len(data), reps=2)
        assert (best["positions_per_tile"], best["positions_per_lane"]) in PARSE_SHAPES and best["ms"] > 0
        assert be.parse_shape() == {k: best[k] for k in ("positions_per_tile", "positions_per_lane")}
        small = datagen.generate("json", 200000, seed=10, conforming=True).tobytes()
        assert be.run_case("both", small) == pyoracle.record("json", "both", small)
        with pytest.raises(hip.HipError):
            be.set_parse_shape(4096, 4)


def test_graph_staging_overflow_falls_back():
    """The graph path stages seg_resolve's outputs for the largest tile of
    the batch it captured (plus a quarter).  A later batch within the same
    capacity whose tiles hold many more tree nodes (nested arrays: several
    nodes per token, against flat arrays of numbers) overflows that staging:
    its tiles are skipped, the batch runs again eagerly, and the next batch
    captures again.  Every record stays the oracle's."""
    from alpacc_amd import hip
    flat = [b"[" + b",".join(b"1" for _ in range(1500 + i)) + b"]" for i in range(64)]
    deep = [b"[" * (1200 + i) + b"1" + b"]" * (1200 + i) for i in range(64)]
    with hip.Backend("json") as be:
        be.set_graphs(True)
        for k, batch in enumerate([flat, flat, deep, deep, flat]):
            print(f"graph staging batch #{k}", flush=True)
            got = be.run_batch("both", batch)
            exp = b"".join(pyoracle.record("json", "both", x) for x in batch)
            assert first_diff(got, exp) is None, k
        captures, replays = be.graph_stats()
        assert captures >= 2, (captures, replays)


def test_graph_token_capacity_overflow_falls_back():
    """The graph path sizes its parser for half again the tokens of the
    batch it captured (to a power of two).  A later batch within the same
    input and byte capacity with far more tokens (flat arrays of one-digit
    numbers against one long string per input) overflows that capacity: the
    device plan leaves no tile, the batch runs again eagerly, and the next
    batch captures again.  Every record stays the oracle's."""
    from alpacc_amd import hip
    sparse = [b'"' + b"a" * (3000 + i) + b'"' for i in range(64)]
    dense = [b"[" + b",".join(b"1" for _ in range(1500 + i // 2)) + b"]" for i in range(64)]
    assert sum(map(len, sparse)) // 2 < sum(map(len, dense)) <= 2 * sum(map(len, sparse))
    with hip.Backend("json") as be:
        be.set_graphs(True)
        for k, batch in enumerate([sparse, sparse, dense, dense, sparse]):
            print(f"graph token capacity batch #{k}", flush=True)
            got = be.run_batch("both", batch)
            exp = b"".join(pyoracle.record("json", "both", x) for x in batch)
            assert first_diff(got, exp) is None, k
        captures, replays = be.graph_stats()
        assert captures >= 2 and replays >= 1, (captures, replays)
