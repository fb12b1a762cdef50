"""Benchmark: JSON-grammar lex + parse of a 1 GiB input on one MI355X.

One step = one alpacc_hip_lex_parse_device call over the whole input already
resident in HBM: the tile lexer, the LLP pair lookup, bracket matching,
parents and the compact tree (the combined mode of the reference).  The
input is synthetic JSON from the seeded generator (seed 1), grammar-
conforming so the full path runs (``--variant literal`` uses the workload as
written, which the grammar rejects early).

Prints one JSON line: input GB/s (value), the roofline of the dominant stage
measured with HIP events on the launch stream, and the sequential C oracle
timed on a bounded sample on this host (cpu_baseline).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

import ctypes

C_u8 = ctypes.c_uint8


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu():
    """The host's CPU model and logical CPU count (nproc)."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count()}


def cpu_baseline(grammar: str, mode: str, seed: int, conforming: bool, sample_bytes: int, budget_s: float = 20.0):
    """The sequential oracle on `sample_bytes` of the same generator, 1 core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from alpacc_amd import datagen

    data = datagen.generate(grammar, sample_bytes, seed=seed, conforming=conforming)
    t_total, runs = 0.0, 0
    while runs < 3 and t_total < budget_s:
        t_total += pyoracle.time_case(grammar, mode, data, sample_bytes)
        runs += 1
    secs = t_total / runs
    return {
        "value": round(sample_bytes / secs / 1e9, 4),
        "unit": "GB/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{grammar} {mode} {'conforming' if conforming else 'literal'} seed {seed}, "
                  f"{sample_bytes} bytes, mean of {runs} run(s), {secs:.3f} s each",
        **host_cpu(),
    }


# Algorithmic bytes of one launch of each kernel: input read plus final
# outputs written (JSON widths: terminal 1, index 8, length 2, production 1);
# intermediates (lane prefixes, pair ids, tile records) are not counted.
def kernel_bytes(name: str, n: int, T: int, N: int) -> int:
    return {
        "lex_fold": n,                 # input
        "lex_emit": n + 11 * T,        # input + (terminal, start, length) per token
        "seg_profile": T,              # tokens
        "seg_resolve": 8 * T + 9 * N,  # token parents + (production, parent) per node
    }.get(name, 0)


def load_traffic(path: str, kernel: str):
    """HBM bytes per launch of `kernel` from a summary written by
    --pmc-summary (FETCH_SIZE and WRITE_SIZE, each scaled by a copy kernel)."""
    if not path or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k["bytes_per_launch"]


def copy_calibration(nbytes: int, reps: int):
    """A plain copy of a known size, to scale the FETCH_SIZE / WRITE_SIZE
    counters: each dispatch reads and writes exactly nbytes."""
    import torch
    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    src.fill_(7)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "reps": reps}), flush=True)


def read_pmc(csv_dir: str, counter: str):
    """{kernel short name: mean counter value per dispatch} from a rocprofv3
    --pmc run written with --output-format csv."""
    import csv
    import glob
    files = glob.glob(os.path.join(csv_dir, "**", "*counter_collection.csv"), recursive=True)
    acc = {}
    for fn in files:
        with open(fn) as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter:
                    continue
                name = row["Kernel_Name"].replace("(anonymous namespace)::", "")
                short = name.split("(")[0].split("::")[-1].split("<")[0]  # template arguments dropped
                acc.setdefault(short, []).append(float(row["Counter_Value"]))
    return {k: (sum(v) / len(v), max(v)) for k, v in acc.items()}


def pmc_summary(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, copy_bytes, out_path):
    """Per-kernel HBM bytes per launch: FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch), each scaled so a plain copy of copy_bytes reads as copy_bytes."""
    fetch = {k: v[0] for k, v in read_pmc(fetch_dir, "FETCH_SIZE").items()}
    write = {k: v[0] for k, v in read_pmc(write_dir, "WRITE_SIZE").items()}
    cf, cw = read_pmc(copy_fetch_dir, "FETCH_SIZE"), read_pmc(copy_write_dir, "WRITE_SIZE")
    ck = "__amd_rocclr_copyBuffer"  # the device-to-device copy; its largest dispatch is the calibration copy
    f_scale = copy_bytes / cf[ck][1]
    w_scale = copy_bytes / cw[ck][1]
    out = {"copy_kernel": ck, "copy_bytes": copy_bytes, "fetch_scale_bytes_per_unit": f_scale,
           "write_scale_bytes_per_unit": w_scale, "kernels": {}}
    for k in sorted(set(fetch) | set(write)):
        fb = fetch.get(k, 0.0) * f_scale
        wb = write.get(k, 0.0) * w_scale
        out["kernels"][k] = {"fetch_bytes": round(fb), "write_bytes": round(wb), "bytes_per_launch": round(fb + wb)}
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out, sort_keys=True))


def batch_inputs(count: int, seed: int, median: int = 4096, sigma: float = 1.5, lo: int = 64, hi: int = 4 << 20):
    """`count` independent grammar-conforming JSON inputs, byte lengths
    lognormal(ln median, sigma) clipped to [lo, hi], packed into one buffer
    with a u64 offsets array (count + 1 entries)."""
    import numpy as np
    from alpacc_amd import datagen
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.round(rng.lognormal(np.log(median), sigma, count)), lo, hi).astype(np.int64)
    offs = np.zeros(count + 1, dtype=np.uint64)
    offs[1:] = np.cumsum(sizes)
    host = np.empty(int(offs[-1]) + 16, dtype=np.uint8)
    for i in range(count):
        datagen.generate("json", int(sizes[i]), seed=seed * 1000003 + i, conforming=True, wrap=True,
                         out=host[int(offs[i]):int(offs[i + 1])])
    return host, offs


def batch_workload(be, count: int, steps: int, seed: int = 3, plain: bool = True):
    """Batching: all inputs lexed and parsed by one batched call per step,
    against the plain path (one lex+parse call, i.e. one launch sequence,
    per input) over the same device-resident inputs (skipped without
    `plain`)."""
    import numpy as np
    import torch
    host, offs = batch_inputs(count, seed)
    total = int(offs[-1])
    dev = torch.from_numpy(host).to("cuda")
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream().cuda_stream
    base = dev.data_ptr()
    # the parser's tile shape for this batch: every shape timed on the batch
    # itself (its many partly filled per-input tiles favour other shapes
    # than one large input does); the context keeps the fastest
    from alpacc_amd import hip
    keep = be.parse_shape()
    best = None
    for sh in hip.PARSE_SHAPES:
        be.set_parse_shape(*sh)
        be.lex_parse_batch_device(base, offs, stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(2):
            be.lex_parse_batch_device(base, offs, stream=stream)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if best is None or dt < best[0]:
            best = (dt, sh)
    be.set_parse_shape(*best[1])
    log(f"batch parser shape: {best[1]}")
    b = be.lex_parse_batch_device(base, offs, stream=stream)  # warmup
    valid = int(np.ctypeslib.as_array((C_u8 * count).from_address(b.valid)).sum())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        be.lex_parse_batch_device(base, offs, stream=stream)
    torch.cuda.synchronize()
    batched_s = (time.perf_counter() - t0) / steps
    # the same batch through the captured graph (the first call captures it)
    be.set_graphs(True)
    bg = be.lex_parse_batch_device(base, offs, stream=stream)
    graph_valid = int(np.ctypeslib.as_array((C_u8 * count).from_address(bg.valid)).sum())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        be.lex_parse_batch_device(base, offs, stream=stream)
    torch.cuda.synchronize()
    graph_s = (time.perf_counter() - t0) / steps
    captures, replays = be.graph_stats()
    be.set_graphs(False)
    batch_shape = be.parse_shape()
    be.set_parse_shape(keep["positions_per_tile"], keep["positions_per_lane"])
    # plain: one call per input, in input order
    plain_valid, plain_s = None, None
    if plain:
        for i in range(min(count, 200)):  # warmup of the per-input path
            be.lex_parse_device(base + int(offs[i]), int(offs[i + 1] - offs[i]), stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plain_valid = 0
        for i in range(count):
            ok, _ = be.lex_parse_device(base + int(offs[i]), int(offs[i + 1] - offs[i]), stream=stream)
            plain_valid += bool(ok)
        torch.cuda.synchronize()
        plain_s = time.perf_counter() - t0
    del dev
    items = [("json", host[int(offs[i]):int(offs[i + 1])].tobytes()) for i in range(count)]
    return items, {
        "workload": f"{count} json inputs, lognormal(ln 4096, 1.5) bytes clipped to [64, 4 MiB], seed {seed}",
        "count": count, "bytes": total, "valid": valid, "plain_valid": plain_valid,
        "batched_ms": round(batched_s * 1e3, 3), "batched_GBps": round(total / batched_s / 1e9, 3),
        "graph_ms": round(graph_s * 1e3, 3), "graph_GBps": round(total / graph_s / 1e9, 3),
        "graph_valid": graph_valid, "graph_captures": captures, "graph_replays": replays,
        "parser_shape": batch_shape,
        "plain_ms": round(plain_s * 1e3, 3) if plain else None,
        "plain_GBps": round(total / plain_s / 1e9, 3) if plain else None,
        "speedup_vs_plain": round(plain_s / batched_s, 2) if plain else None,
    }


def serve_trace(count: int, rate: float, seed: int, max_bytes: int = 16 << 20):
    """Requests of the serving workload: 70% json / 30% arithmetic, sizes
    lognormal(ln 2048, 2) clipped to [128, max_bytes], 3% with one random
    byte flipped, open-loop Poisson arrivals at `rate` per second."""
    import numpy as np
    from alpacc_amd import datagen
    rng = np.random.default_rng(seed)
    gidx = (rng.random(count) >= 0.7).astype(np.int32)  # 0 json, 1 arithmetic
    sizes = np.clip(np.round(rng.lognormal(np.log(2048), 2.0, count)), 128, max_bytes).astype(np.int64)
    flips = rng.random(count) < 0.03
    arrival = np.cumsum(rng.exponential(1.0 / rate, count)) * 1e9
    payloads = []
    for i in range(count):
        data = datagen.generate(("json", "arithmetic")[gidx[i]], int(sizes[i]), seed=seed * 7919 + i,
                                conforming=True).copy()
        if flips[i]:
            data[rng.integers(0, len(data))] = rng.integers(0, 256)
        payloads.append(data.tobytes())
    return gidx, payloads, arrival.astype(np.uint64)


def serve_workload(rates, seconds: float, seed: int = 4, plain: bool = True):
    """Serving: the trace replayed open-loop against the continuous-batching
    server (two lanes per grammar), the same with each lane's combined
    batches through its captured graph, and the plain server (one request
    per launch sequence, one lane), all with json and arithmetic resident."""
    import numpy as np
    from alpacc_amd import hip
    names = ["json", "arithmetic"]
    out = []
    items = []
    for rate in rates:
        count = max(1, int(rate * seconds))
        gidx, payloads, arrival = serve_trace(count, rate, seed + int(rate))
        items = [(names[g], x) for g, x in zip(gidx, payloads)]
        total = sum(len(x) for x in payloads)
        modes = ["both"] * count
        row = {"rate_req_s": rate, "requests": count, "bytes": total}
        configs = [("served", 2, 0, False), ("served_graph", 2, 0, True)] + ([("plain", 1, 1, False)] if plain else [])
        for label, lanes, max_batch, graphs in configs:
            srv = hip.Server(names, lanes=lanes, hbm_budget=64 << 30, max_batch=max_batch)
            srv.set_graphs(graphs)
            # warmup: one untimed replay of the trace, so every lane's
            # workspaces have grown to its batches before the timed replay
            srv.replay(gidx, modes, payloads, arrival)
            recs, lat = srv.replay(gidx, modes, payloads, arrival)
            st = srv.stats()
            srv.close()
            span_s = float(np.max(arrival + lat)) / 1e9
            row[label] = {"p50_ms": round(float(np.percentile(lat, 50)) / 1e6, 3),
                          "p99_ms": round(float(np.percentile(lat, 99)) / 1e6, 3),
                          "GBps": round(total / span_s / 1e9, 3), "req_s": round(count / span_s, 1),
                          "batches": st["batches"], "max_batch": st["max_batch"]}
            log(f"serve {rate:.0f} req/s {label}: {row[label]}")
        if plain:
            row["throughput_vs_plain"] = round(row["served"]["GBps"] / max(row["plain"]["GBps"], 1e-9), 2)
        out.append(row)
    return out, items


def serve_capacity(rates, seconds: float, seed: int = 5, p99_limit_ms: float = 10.0, lanes: int = 4):
    """The continuous-batching server's capacity: the same trace shape at
    rising open-loop rates (each replayed once untimed, then timed) until p99
    latency reaches p99_limit_ms; the highest rate that stays below it is
    the sustainable one.  Four lanes per grammar: each lane's host side
    (packing a batch, copying its outputs back, cutting the records) is one
    thread, and at these rates that is the limit (2 lanes saturate near
    150k req/s, 8 contend for the device)."""
    import numpy as np
    from alpacc_amd import hip
    names = ["json", "arithmetic"]
    rows, best = [], None
    # one untimed replay on a throwaway server first: the process's first
    # pinned allocations and the grammars' first launches are not part of
    # any rate's latency
    if rates:
        count = max(1, int(rates[0] * seconds))
        gidx, payloads, arrival = serve_trace(count, rates[0], seed)
        srv = hip.Server(names, lanes=lanes, hbm_budget=64 << 30)
        srv.replay(gidx, ["both"] * count, payloads, arrival)
        srv.close()
    for rate in rates:
        count = max(1, int(rate * seconds))
        gidx, payloads, arrival = serve_trace(count, rate, seed + int(rate))
        total = sum(len(x) for x in payloads)
        modes = ["both"] * count
        srv = hip.Server(names, lanes=lanes, hbm_budget=64 << 30)
        srv.replay(gidx, modes, payloads, arrival)
        _, lat = srv.replay(gidx, modes, payloads, arrival)
        st = srv.stats()
        srv.close()
        span_s = float(np.max(arrival + lat)) / 1e9
        row = {"rate_req_s": rate, "lanes": lanes, "requests": count, "bytes": total,
               "p50_ms": round(float(np.percentile(lat, 50)) / 1e6, 3),
               "p99_ms": round(float(np.percentile(lat, 99)) / 1e6, 3),
               "GBps": round(total / span_s / 1e9, 3), "req_s": round(count / span_s, 1),
               "max_batch": st["max_batch"]}
        log(f"serve capacity {rate:.0f} req/s: {row}")
        rows.append(row)
        if row["p99_ms"] >= p99_limit_ms:
            break
        best = row
    return {"p99_limit_ms": p99_limit_ms, "trace_seconds": seconds, "lanes_per_grammar": lanes, "rates": rows,
            "sustained_req_s": best["rate_req_s"] if best else None,
            "sustained_GBps": best["GBps"] if best else None}


def _oracle_worker(args):
    """Seconds and bytes of the sequential oracle over a list of (grammar, payload)."""
    mode, items = args
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    secs, nbytes = 0.0, 0
    for grammar, payload in items:
        secs += pyoracle.time_case(grammar, mode, payload, len(payload))
        nbytes += len(payload)
    return secs, nbytes


def cpu_parallel_baseline(items, mode: str, label: str, cores: int = 0):
    """The oracle over `items` ((grammar, payload) pairs) in `cores` worker
    processes on this host; aggregate GB/s over the wall-clock of the
    parallel run.  The default is min(16, nproc): a process with one GPU
    gets a share of 16 CPUs on the machines this is measured on, whatever
    nproc reports (the line records both)."""
    import multiprocessing as mp
    cores = cores or max(1, min(16, os.cpu_count() or 1))
    chunks = [(mode, items[i::cores]) for i in range(cores)]
    with mp.get_context("spawn").Pool(cores) as pool:
        pool.map(_oracle_worker, [(mode, items[:1])] * cores)  # load the oracle in every worker
        t0 = time.perf_counter()
        res = pool.map(_oracle_worker, chunks)
        wall = time.perf_counter() - t0
    nbytes = sum(b for _, b in res)
    return {"value": round(nbytes / wall / 1e9, 4), "unit": "GB/s", "cores": cores, "kind": "port",
            "sample": f"{label}: {len(items)} inputs, {nbytes} bytes, {cores} processes, {wall:.2f} s wall",
            **host_cpu()}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=1 << 30)
    ap.add_argument("--variant", choices=["conforming", "literal"], default="conforming")
    ap.add_argument("--cpu-sample", type=int, default=256 << 20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default lexer shape instead of tuning it")
    ap.add_argument("--lexer-shape", metavar="IPT,WPC,WIN", default="",
                    help="use this lexer shape (bytes per lane, workgroups per CU, lookback window) instead of tuning")
    ap.add_argument("--parser-shape", metavar="PT,APL", default="",
                    help="use this parser tile shape (positions per tile, per lane) instead of tuning")
    ap.add_argument("--batch-count", type=int, default=20000, help="inputs of the batching workload (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batching workload's batched and graph calls (no 1 GiB step, plain path or "
                         "serving; for profiling the batch path), one JSON line of its results")
    ap.add_argument("--serve-rates", default="2000,10000,50000", help="request rates of the serving workload ('' = skip)")
    ap.add_argument("--serve-seconds", type=float, default=1.0)
    ap.add_argument("--serve-capacity", default="100000,125000,150000,200000,300000",
                    help="rising request rates of the capacity sweep, until p99 >= 10 ms ('' = skip)")
    ap.add_argument("--capacity-seconds", type=float, default=0.25)
    ap.add_argument("--no-literal", action="store_true", help="skip the step on the literal (as-written) input")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic_json_lexparse_1GiB.json"),
                    help="per-kernel HBM bytes per launch written by --pmc-summary")
    ap.add_argument("--copy-calibration", type=int, metavar="BYTES", default=0,
                    help="only run a plain device copy of BYTES (counter calibration)")
    ap.add_argument("--pmc-summary", nargs=5, metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR",
                                                       "COPY_WRITE_DIR", "OUT"),
                    help="summarise rocprofv3 FETCH_SIZE/WRITE_SIZE runs into OUT (no GPU)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported "
              f"(the lex/parse path runs on a single MI355X)", file=sys.stderr)
        return 2
    if a.pmc_summary:
        fd, wd, cfd, cwd, out = a.pmc_summary
        pmc_summary(fd, wd, cfd, cwd, a.copy_calibration or (1 << 30), out)
        return 0
    if a.copy_calibration:
        copy_calibration(a.copy_calibration, 3)
        return 0
    if a.batch_only:
        from alpacc_amd import hip
        be = hip.Backend("json", device=0)
        if a.lexer_shape:
            be.set_lex_shape(*[int(x) for x in a.lexer_shape.split(",")])
        _, batched = batch_workload(be, a.batch_count, a.batch_steps, plain=False)
        be.close()
        print(json.dumps(batched))
        return 0

    import numpy as np
    import torch
    from alpacc_amd import datagen, hip

    n = a.size
    conforming = a.variant == "conforming"
    t0 = time.time()
    host = datagen.generate("json", n, seed=1, conforming=conforming, wrap=True)
    log(f"generated {n} bytes in {time.time() - t0:.1f}s")
    dev = torch.from_numpy(host).to("cuda")
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream().cuda_stream
    be = hip.Backend("json", device=0)
    # the lexer's launch shape, tuned on this input before the timed region
    if a.lexer_shape:
        be.set_lex_shape(*[int(x) for x in a.lexer_shape.split(",")])
    shape = be.lex_shape() if a.no_autotune or a.lexer_shape else be.autotune_lexer(dev.data_ptr(), n, "both", reps=3)
    log(f"lexer shape: {shape}")
    # the parser's tile shape, tuned on the same input
    if a.parser_shape:
        be.set_parse_shape(*[int(x) for x in a.parser_shape.split(",")])
    pshape = be.parse_shape() if a.no_autotune or a.parser_shape else be.autotune_parser(dev.data_ptr(), n, reps=3)
    log(f"parser shape: {pshape}")

    def step():
        ok, tree = be.lex_parse_device(dev.data_ptr(), n, stream=stream)
        return ok, tree

    for i in range(a.warmup):
        ok, tree = step()
        log(f"warmup {i}: valid={ok}")
    # timed steps: no per-kernel events inside the timed region
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(a.steps):
        ok, tree = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / a.steps
    T = tree.tokens.count if ok else 0
    N = tree.num_nodes if ok else 0

    # per-kernel durations: the same steps again with HIP events around
    # every launch on the launch stream
    be.set_timing(True)
    for i in range(a.steps):
        step()
    kernels = {}
    for k in be.timed_kernels():
        tot, runs = be.timing(k)
        kernels[k] = tot / max(runs, 1)
    lex_ms, lex_runs = be.timing("lex")
    par_ms, par_runs = be.timing("parse")
    be.set_timing(False)
    log(f"step {ms:.3f} ms; lexer {lex_ms / max(lex_runs, 1):.3f} ms; parser {par_ms / max(par_runs, 1):.3f} ms; "
        f"T={T} N={N} valid={ok}")
    log("kernels (avg ms): " + ", ".join(f"{k} {v:.3f}" for k, v in sorted(kernels.items(), key=lambda x: -x[1])))

    dom = max(kernels, key=kernels.get)
    savg = kernels[dom]
    sbytes = kernel_bytes(dom, n, T, N)
    achieved = sbytes / (savg / 1e3) / 1e9 if savg > 0 else 0.0
    roofline = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": load_traffic(a.traffic, dom),
        "algorithmic_bytes": sbytes,
        "avg_ms": round(savg, 4),
    }
    batched = None
    batch_items = serve_items = None
    # the lexer-only workload on the same resident input (lexer numbering,
    # tokens only): the whole launch sequence of alpacc_hip_lex_device per step
    be.lex_device(dev.data_ptr(), n, "lexer", stream=stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(a.steps):
        xok, xl = be.lex_device(dev.data_ptr(), n, "lexer", stream=stream)
    torch.cuda.synchronize()
    xms = (time.perf_counter() - t0) * 1e3 / a.steps
    lexer_only = {"workload": f"json lexer-only, the same {n}-byte input", "ms_per_step": round(xms, 3),
                  "GBps": round(n / (xms / 1e3) / 1e9, 3), "tokens": int(xl.count), "valid": bool(xok)}
    log(f"lexer-only: {lexer_only}")
    if a.batch_count:
        log(f"batching workload: {a.batch_count} inputs")
        batch_items, batched = batch_workload(be, a.batch_count, a.batch_steps)
        log(f"batched {batched['batched_GBps']} GB/s (graph {batched['graph_GBps']} GB/s, "
            f"{batched['graph_captures']} capture, {batched['graph_replays']} replays), "
            f"plain {batched['plain_GBps']} GB/s, x{batched['speedup_vs_plain']}")
    literal = None
    if not a.no_literal and conforming:
        # the same step on the workload as written (escapes and exponents
        # the grammar has no rule for): rejected, but every kernel still runs
        # over the whole input
        lit = datagen.generate("json", n, seed=1, conforming=False, wrap=True)
        dev.copy_(torch.from_numpy(lit))
        del lit
        step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(a.steps):
            lok, _ = step()
        torch.cuda.synchronize()
        lms = (time.perf_counter() - t0) * 1e3 / a.steps
        literal = {"variant": "literal", "ms_per_step": round(lms, 3), "GBps": round(n / (lms / 1e3) / 1e9, 3),
                   "valid": bool(lok)}
        log(f"literal input: {literal}")
    served = None
    if a.serve_rates:
        rates = [float(x) for x in a.serve_rates.split(",") if x]
        served, serve_items = serve_workload(rates, a.serve_seconds)
    capacity = None
    if a.serve_capacity:
        capacity = serve_capacity([float(x) for x in a.serve_capacity.split(",") if x], a.capacity_seconds)
    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline("json", "both", 1, conforming, min(a.cpu_sample, n))
        if batched:
            batched["cpu_baseline"] = cpu_parallel_baseline(batch_items, "both", "batching workload")
        if served:
            served_items = serve_items[:max(1, len(serve_items) // 5)]  # a bounded sample of the trace
            cpu_srv = cpu_parallel_baseline(served_items, "both", "first fifth of the serving trace at the highest rate")
            for row in served:
                row["cpu_baseline"] = cpu_srv
    path_bytes = n + 19 * T + 9 * N  # input + tokens + tree (combined-mode outputs)
    # HBM bytes one step moves by the counters (FETCH_SIZE + WRITE_SIZE of
    # every kernel, from the same summary as roofline.traffic)
    measured = None
    if a.traffic and os.path.exists(a.traffic):
        with open(a.traffic) as f:
            measured = sum(k["bytes_per_launch"] for k in json.load(f).get("kernels", {}).values())
    out = {
        "metric": "lex+parse input GB/s (1 GiB, %HBM peak); batched GB/s; served p50/p99 ms",
        "value": round(n / (ms / 1e3) / 1e9, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": f"synthetic ({a.variant} JSON, seed 1)",
        "config": {"workload": f"json lex+parse, one {n}-byte input (combined mode, device-resident)",
                   "grammar": "grammars/json.alp", "bytes": n, "tokens": T, "tree_nodes": N,
                   "valid": bool(ok), "variant": a.variant,
                   "lexer_shape": {k: shape[k] for k in ("items_per_thread", "workgroups_per_cu", "lookback_window")},
                   "parser_shape": {k: pshape[k] for k in ("positions_per_tile", "positions_per_lane")}},
        "pct_hbm_peak_input": round(100 * n / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 2),
        "path_algorithmic_GBps": round(path_bytes / (ms / 1e3) / 1e9, 2),
        "path_frac": round(path_bytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "path_measured_bytes": measured,
        "path_measured_frac": round(measured / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4) if measured else None,
        "kernels_ms": {k: round(v, 4) for k, v in sorted(kernels.items(), key=lambda x: -x[1])},
        "roofline": roofline,
        "literal": literal,
        "lexer_only": lexer_only,
        "batched": batched,
        "served": served,
        "serve_capacity": capacity,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    be.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
