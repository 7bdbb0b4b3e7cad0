"""Benchmark of the GA* solve path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload plain-grid1000|plain-grid16384|deep-grid|plain-puzzle|plain-puzzle5|
                                bounded-puzzle5|batch-c4]
    python bench.py --copy-calib 1024     # one 1 GiB device-to-device copy (counter calibration)

Prints ONE JSON line (see the keys below). On the serving stream (the default workload)
a "step" is one batch of QUERIES_PER_STEP arrivals, so `--steps K` serves a stream of
K * QUERIES_PER_STEP queries; on the plain workloads a step is one query. `value` is
whole-job queries/s with every input already resident in HBM when the timed region
starts. The dominant kernel's roofline is computed from the algorithmic bytes of each
query (per-query counters x the per-operation bytes below) over the kernel's duration
measured with HIP events on the library's stream.

Algorithmic bytes (per operation, counted by the kernels):
  grid:   16 * pops + 18 * expanded + 8 * generated + 25 * inserted
  puzzle: 16 * pops + 16 * expanded + 18 * generated + 40 * inserted
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
QUERIES_PER_STEP = 10
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_teams_serve_traffic.md")


def measured_traffic(bench_args):
    """HBM bytes per serving launch from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy-kernel calibration), taken
    only when that summary profiled this same bench invocation (its `bench_args:` line);
    otherwise None. Returns (bytes per launch or None, what the figure comes from)."""
    vals = {}
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if ":" in line and not line.startswith((" ", "|", "#", "`")):
                    k, v = line.split(":", 1)
                    vals[k.strip()] = v.strip()
    except OSError:
        return None, "no counter profile"
    if vals.get("bench_args") != bench_args:
        return None, f"counter profile is of bench.py {vals.get('bench_args')!r}, not of this run"
    try:
        return float(vals["traffic_bytes_per_launch"]), f"rocprofv3 counters of bench.py {bench_args} " \
            f"({os.path.relpath(TRAFFIC_PROFILE, ROOT)})"
    except (KeyError, ValueError):
        return None, "counter profile without traffic_bytes_per_launch"


def copy_calib(mib):
    """One device-to-device copy of `mib` MiB through the HIP runtime's copy kernel, run
    under rocprofv3 counters to calibrate FETCH_SIZE / WRITE_SIZE against known bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(2):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    return {"copy_bytes": n, "copies": 2}


def grid_bytes_alg(r):
    return 16 * r["pops"] + 18 * r["expanded"] + 8 * r["generated"] + 25 * r["inserted"]


def puzzle_bytes_alg(r):
    return 16 * r["pops"] + 16 * r["expanded"] + 18 * r["generated"] + 40 * r["inserted"]


def say(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pct(xs, q):
    return float(np.percentile(np.asarray(xs, float), q)) if xs else None


# ------------------------------------------------------------------ workloads

def grid1000_specs(n, seed0=1):
    """1000x1000 grids, obstacles i.i.d. Bernoulli(0.30), start (0,0), goal (999,999)."""
    return [dict(kind="grid", H=1000, W=1000, seed=seed0 + i, p=0.30, s=(0, 0), g=(999, 999)) for i in range(n)]


def grid16384_specs(n, seed0=2):
    """16384x16384 grids (1 bit per cell), obstacles i.i.d. Bernoulli(0.35), opposite corners."""
    return [dict(kind="grid", H=16384, W=16384, seed=seed0 + i, p=0.35, s=(0, 0), g=(16383, 16383)) for i in range(n)]


def run_plain(ctx, specs, warm_specs):
    """One query per whole-GPU launch (ua_grid_solve), the reference's usage model.
    The warm-up queries (other seeds) run untimed first; every spec is then timed."""
    for s in warm_specs:
        g = ctx.grid_generate(s["H"], s["W"], s["seed"], s["p"], *s["s"], *s["g"])
        ctx.grid_solve(g, *s["s"], *s["g"])
        g.free()
    grids = [ctx.grid_generate(s["H"], s["W"], s["seed"], s["p"], *s["s"], *s["g"]) for s in specs]
    # every library call returns after synchronising the context's stream, so the
    # timed region starts and ends with the device idle
    lat, res = [], []
    say(f"timing {len(specs)} plain queries")
    t0 = time.perf_counter()
    for g, s in zip(grids, specs):
        t = time.perf_counter()
        r = ctx.grid_solve(g, *s["s"], *s["g"])
        lat.append((time.perf_counter() - t) * 1e3)
        res.append(r)
    wall = time.perf_counter() - t0
    for g in grids:
        g.free()
    return wall, lat, res


def cpu_baseline_grid(specs, budget_s=20.0, threads=None):
    """The oracle's restatement of the reference CPU A* (FP32 costs, reference
    containers), timed on this host over a bounded sample of the same stream."""
    from oracle import oracle as O
    threads = threads or min(16, os.cpu_count() or 1)
    # one query alone for the single-thread latency
    s = specs[0]
    g = O.grid_bernoulli(s["H"], s["W"], s["seed"], s["p"], *s["s"], *s["g"])
    t = time.perf_counter()
    O.grid_astar(g, s["H"], s["W"], *s["s"], *s["g"], "f32", want_path=False)
    one = time.perf_counter() - t
    n = max(threads, min(len(specs), int(budget_s * threads / max(one, 1e-3))))
    sample = [specs[i % len(specs)] for i in range(n)]
    inputs = [O.grid_bernoulli(s["H"], s["W"], s["seed"], s["p"], *s["s"], *s["g"]) for s in sample]

    def solve(k):
        s = sample[k]
        return O.grid_astar(inputs[k], s["H"], s["W"], *s["s"], *s["g"], "f32", want_path=False)

    t = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        list(ex.map(solve, range(n)))
    wall = time.perf_counter() - t
    model, nproc = host_cpu()
    return {"value": n / wall, "unit": "queries/s", "cores": threads, "kind": "port", "nproc": nproc,
            "cpu_model": model,
            "sample": f"{n} queries of the same stream on {threads} threads ({wall:.1f} s); "
                      f"one query alone on 1 thread: {one * 1e3:.0f} ms (oracle FP32 variant, reference containers)"}


# ------------------------------------------------------------------ mixed serving stream

GOLDEN = os.path.join(ROOT, "tests", "golden", "c3_puzzles.json")


def mixed_stream(n, seed=5, max_side=8192):
    """The serving stream: 70% grids (side log-uniform in [512, max_side], obstacles
    Bernoulli(p), p ~ U[0.2, 0.4], random open start and goal), 30% random solvable
    15-puzzles (boards of the seeded generator, seeds 3000..3255)."""
    import uastar_amd as U
    with open(GOLDEN) as f:
        boards = json.load(f)
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        if rng.random() < 0.7:
            side = int(math.floor(math.exp(rng.uniform(math.log(512), math.log(max_side)))))
            p = float(rng.uniform(0.2, 0.4))
            gseed = int(rng.integers(1, 1 << 62))
            pts = []
            while len(pts) < 2:
                x, y = int(rng.integers(side)), int(rng.integers(side))
                if U.cell_open(gseed, p, side, x, y):
                    pts.append((x, y))
            out.append(dict(kind="grid", H=side, W=side, seed=gseed, p=p, s=pts[0], g=pts[1]))
        else:
            b = boards[int(rng.integers(len(boards)))]
            out.append(dict(kind="puzzle", board=b["board"], steps=b["steps"], expanded=b["expanded"]))
    return out


def materialize(ctx, specs):
    for s in specs:
        if s["kind"] == "grid":
            s["grid"] = ctx.grid_generate(s["H"], s["W"], s["seed"], s["p"], *s["s"], *s["g"])


def answer(r, s):
    if not r["found"]:
        return None
    return (r["orth"], r["diag"]) if s["kind"] == "grid" else r["steps"]


def alg_bytes(r, s):
    return grid_bytes_alg(r) if s["kind"] == "grid" else puzzle_bytes_alg(r)


def plain_pass(ctx, specs):
    """One query per whole-GPU launch, in stream order; returns service times (ms)."""
    svc, res = [], []
    for s in specs:
        t = time.perf_counter()
        r = ctx.grid_solve(s["grid"], *s["s"], *s["g"]) if s["kind"] == "grid" else ctx.puzzle_solve(4, s["board"])
        svc.append((time.perf_counter() - t) * 1e3)
        res.append(r)
    return svc, res


def fifo_latency(arrivals_ms, service_ms):
    """Latency of a single FIFO server (the plain path) under the given arrivals."""
    t, lat = 0.0, []
    for a, d in zip(arrivals_ms, service_ms):
        t = max(t, a) + d
        lat.append(t - a)
    return lat


def serve_pass(ctx, specs, arrivals_ms):
    """Continuous batching: each query is submitted at its arrival time; the loop
    steps the device and retires queries as they finish."""
    import uastar_amd as U
    queries = [ctx.grid_query(s["grid"], *s["s"], *s["g"]) if s["kind"] == "grid"
               else ctx.puzzle_query(4, s["board"]) for s in specs]
    n = len(specs)
    ticket_of, done = {}, {}
    t0 = time.perf_counter()
    i = 0
    last_say = t0
    step_wall = 0.0  # host time inside serve_step: launch + host work around it
    while len(done) < n:
        now = (time.perf_counter() - t0) * 1e3
        while i < n and arrivals_ms[i] <= now:
            ticket_of[ctx.serve_submit(queries[i])] = i
            i += 1
        ts = time.perf_counter()
        active, pending = ctx.serve_step()
        step_wall += time.perf_counter() - ts
        tnow = (time.perf_counter() - t0) * 1e3
        for t, r in ctx.serve_poll():
            done[ticket_of[t]] = (r, tnow)
        if active == 0 and pending == 0 and i < n:
            time.sleep(max(0.0, (arrivals_ms[i] - tnow) / 1e3))
        if time.perf_counter() - last_say > 20:
            say(f"serve: {len(done)}/{n} done, {active} active, {pending} pending")
            last_say = time.perf_counter()
    res = [done[k][0] for k in range(n)]
    lat = [done[k][1] - arrivals_ms[k] for k in range(n)]
    span = max(done[k][1] for k in range(n)) - arrivals_ms[0]
    return res, lat, span, step_wall * 1e3


def host_cpu():
    """The host's CPU model and logical CPU count (as /proc/cpuinfo and os.cpu_count see them)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline_mixed(ctx, specs, budget_s=25.0, threads=None, seed=11, puzzle_cap=4_000_000):
    """The oracle (restatement of the reference CPU A*: FP32 grid costs and the reference's
    containers; puzzles with the 7-8 PDB) on a uniform random sample of the timed stream,
    one query per host thread. The sample is a prefix of a seeded random permutation of
    the stream, long enough that its predicted single-thread work fills about budget_s
    seconds on `threads` threads. A puzzle whose A* needs more than puzzle_cap expansions
    (its count is in the c3 fixture) runs to puzzle_cap, and its time is scaled up by its
    expansion count; such times are marked extrapolated. `value` is the stream's estimated
    CPU throughput: threads / (mean single-thread time per query of the sample)."""
    from oracle import oracle as O
    import hashlib
    threads = threads or min(16, os.cpu_count() or 1)
    # the 15-puzzle PDB: the device-built table, byte-equal to the oracle's BFS (checked
    # here against the golden digest; the CPU BFS takes minutes)
    pdb = ctx.puzzle_download_pdb(4)
    with open(os.path.join(ROOT, "tests", "golden", "pdb_digests.json")) as f:
        digests = json.load(f)["4"]
    off = 0
    for d in digests:
        assert hashlib.sha256(pdb[off: off + d["size"]].tobytes()).hexdigest() == d["sha256"], "PDB digest"
        off += d["size"]
    rng = np.random.default_rng(seed)
    order = [int(i) for i in rng.permutation(len(specs))]

    def predicted_s(s):  # rough single-thread seconds, only to size the sample
        if s["kind"] == "grid":
            return 1.2e-7 * s["H"] * s["W"] + 0.05
        return 3.5e-6 * min(s["expanded"], puzzle_cap)

    sample, work = [], 0.0
    for i in order:
        if len(sample) >= threads and work >= budget_s * threads:
            break
        sample.append(i)
        work += predicted_s(specs[i])

    def solve(i):
        s = specs[i]
        if s["kind"] == "grid":
            g = O.grid_bernoulli(s["H"], s["W"], s["seed"], s["p"], *s["s"], *s["g"])
            t = time.perf_counter()
            O.grid_astar(g, s["H"], s["W"], *s["s"], *s["g"], "f32", want_path=False)
            return time.perf_counter() - t, False
        t = time.perf_counter()
        capped = s["expanded"] > puzzle_cap
        O.puzzle_solve(4, np.array(s["board"], np.uint8), pdb, "astar", node_limit=puzzle_cap if capped else 0)
        dt = time.perf_counter() - t
        return (dt * s["expanded"] / puzzle_cap, True) if capped else (dt, False)

    t = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        out = list(ex.map(solve, sample))
    wall = time.perf_counter() - t
    times = [x[0] for x in out]
    n_ext = sum(x[1] for x in out)
    mean = float(np.mean(times))
    model, nproc = host_cpu()
    n_grid = sum(specs[i]["kind"] == "grid" for i in sample)
    return {"value": threads / mean, "unit": "queries/s", "cores": threads, "kind": "port",
            "nproc": nproc, "cpu_model": model,
            "sample": f"uniform random sample of {len(sample)} of the {len(specs)} timed stream queries "
                      f"({n_grid} grids, {len(sample) - n_grid} puzzles), solved by the oracle (FP32 grid costs, "
                      f"reference containers) one query per thread on {threads} threads in {wall:.1f} s; value = "
                      f"{threads} threads / mean single-thread time per query ({mean * 1e3:.0f} ms), i.e. the "
                      f"stream's CPU throughput with perfect packing; {n_ext} puzzle time(s) extrapolated from "
                      f"{puzzle_cap} expansions by their A* expansion counts; PDB build and input generation "
                      f"excluded",
            "extrapolated_stream_s": mean * len(specs) / threads,
            "measured_sample_qps": len(sample) / wall}


def run_serve_mixed(args):
    import uastar_amd as U
    ctx = U.Context(0, graphs=args.graph, grid_heaps=args.grid_heaps, xcd_shared=args.xcd_shared)
    n_warm, n_timed = args.warmup * QUERIES_PER_STEP, args.steps * QUERIES_PER_STEP
    specs = mixed_stream(n_warm + n_timed)
    materialize(ctx, specs)
    warm, specs = specs[:n_warm], specs[n_warm:]
    ctx.puzzle_build_pdb(4)
    tuning = None
    if not args.no_tune:
        # the autotuner measures on its own sample (another seed of the same stream)
        tspecs = mixed_stream(24, seed=77)
        materialize(ctx, tspecs)
        tq = [ctx.grid_query(s["grid"], *s["s"], *s["g"]) if s["kind"] == "grid" else ctx.puzzle_query(4, s["board"])
              for s in tspecs]
        t = time.perf_counter()
        tuning = ctx.autotune(tq)
        say(f"autotune ({time.perf_counter() - t:.1f} s): {tuning}")
        for s in tspecs:
            if s["kind"] == "grid":
                s["grid"].free()
    say(f"warm-up: {len(warm)} plain + served queries")
    plain_pass(ctx, warm)
    for s in warm:
        q = ctx.grid_query(s["grid"], *s["s"], *s["g"]) if s["kind"] == "grid" else ctx.puzzle_query(4, s["board"])
        ctx.serve_submit(q)
    ctx.serve_drain()
    ctx.serve_poll()
    say(f"plain pass: {len(specs)} queries, one per whole-GPU launch")
    svc, pres = plain_pass(ctx, specs)
    plain_qps = len(specs) / (sum(svc) / 1e3)
    say(f"plain: {plain_qps:.2f} queries/s")
    rate = args.load * plain_qps  # arrivals per second
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1e3 / rate, len(specs))).tolist()
    arrivals = [a - arrivals[0] for a in arrivals]
    st0, ls0 = ctx.serve_stats(), ctx.launch_stats()
    say(f"serve pass: Poisson arrivals at {rate:.1f} queries/s ({args.load:g} x plain capacity)")
    sres, slat, span, step_wall_ms = serve_pass(ctx, specs, arrivals)
    st1, ls1 = ctx.serve_stats(), ctx.launch_stats()
    serve_qps = len(specs) / (span / 1e3)
    plat = fifo_latency(arrivals, svc)
    mismatch = sum(answer(a, s) != answer(b, s) for a, b, s in zip(pres, sres, specs))
    errors = sum(r["status"] < 0 for r in pres + sres)
    kms = st1["kernel_ms"] - st0["kernel_ms"]
    launches = st1["launches"] - st0["launches"]
    alg = sum(alg_bytes(r, s) for r, s in zip(sres, specs))
    achieved = alg / (kms * 1e-3) / 1e9 if kms > 0 else 0.0  # = per-launch bytes / average launch time
    atomics = sum(r["atomics"] for r in sres)
    bench_args = f"--steps {args.steps} --warmup {args.warmup} --load {args.load:g}"
    traffic, traffic_src = measured_traffic(bench_args)
    out = {
        "metric": "queries/s + p50/p99 latency, batched serving vs plain GA*, 1 MI355X, exact",
        "value": serve_qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": span / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded generators)",
        "config": {"workload": f"serving stream of {len(specs)} queries ({args.steps} steps of "
                               f"{QUERIES_PER_STEP} arrivals, after {len(warm)} warm-up queries): 70% grids (side "
                               "log-uniform in [512, 8192], Bernoulli(p) obstacles with p ~ U[0.2,0.4], random open "
                               "start/goal), 30% random solvable 15-puzzles with the 7-8 PDB; Poisson arrivals at "
                               "load x plain capacity",
                   "queries": len(specs), "queries_per_step": QUERIES_PER_STEP,
                   "mode": "serve (continuous batching; team width per class from the autotuner)",
                   "xcds": "one class per team scope" if not args.xcd_shared else "shared",
                   "grid_open_list": "per-wave heaps" if args.grid_heaps else "bucket ring",
                   "load": args.load,
                   "arrival_rate_qps": rate},
        "p50_ms": pct(slat, 50),
        "p99_ms": pct(slat, 99),
        "plain": {"queries_per_s": plain_qps, "p50_ms": pct(plat, 50), "p99_ms": pct(plat, 99),
                  "service_p50_ms": pct(svc, 50), "service_p99_ms": pct(svc, 99)},
        "speedup_vs_plain": serve_qps / plain_qps,
        "exact_mismatches_serve_vs_plain": mismatch,
        "errors": errors,
        "serve_launches": launches,
        "serve_widened": st1["widened"] - st0["widened"],
        # host time per serving launch outside the kernel (admission, descriptor and
        # result copies, launch, retirement), and per plain query outside its kernel
        "host_overhead_ms": {"per_serve_launch": (step_wall_ms - kms) / max(1, launches),
                             "per_plain_query": (sum(svc) - sum(r["kernel_ms"] for r in pres)) / len(pres)},
        # the serving launches of the pass: replayed from a graph or issued call by call,
        # and the host time of each launch outside its kernel (staging, copies, launch, wait)
        "launch": {"graph_launches": ls1["graph_launches"] - ls0["graph_launches"],
                   "direct_launches": ls1["direct_launches"] - ls0["direct_launches"],
                   "host_ms_per_launch": ((ls1["wall_ms"] - ls0["wall_ms"]) - (ls1["kernel_ms"] - ls0["kernel_ms"]))
                   / max(1, launches),
                   "graph_error": ls1["graph_error"]},
        "tuning": tuning,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "k_teams<2, 0> (serving step, grids and 15-puzzles)", "launches": launches,
                     "kernel_ms_avg": kms / max(1, launches), "alg_bytes_per_launch": alg / max(1, launches),
                     "kernel_ms_total": kms, "alg_bytes_total": alg},
        # device-memory atomics on the search state (relaxing atomicMin on best words and
        # hash values, hash-claiming CAS), counted by the kernel, over the serving launches
        "atomics": {"total": atomics, "per_launch": atomics / max(1, launches),
                    "per_s": atomics / (kms * 1e-3) if kms > 0 else 0.0},
        # absolute search throughput: nodes expanded by the served queries over the
        # serving span (wall) and over the serving launches' kernel time, and by the
        # plain pass over its summed per-query service time
        "expanded": {"serve_total": sum(r["expanded"] for r in sres),
                     "serve_per_s_wall": sum(r["expanded"] for r in sres) / (span * 1e-3),
                     "serve_per_s_kernel": sum(r["expanded"] for r in sres) / (kms * 1e-3) if kms > 0 else 0.0,
                     "plain_per_s": sum(r["expanded"] for r in pres) / (sum(svc) * 1e-3)},
    }
    # the device's random-atomic throughput (ua_debug_atomic_rate, after the timed passes),
    # the ceiling of the relaxations
    peak = ctx.atomic_rate() * 1e9
    out["atomics"].update({"probe_peak_per_s": peak, "frac": out["atomics"]["per_s"] / peak if peak else None})
    def diag(res, lat=None):
        rows = []
        for k, (r, sp) in enumerate(zip(res, specs)):
            rows.append(dict(kind=sp["kind"], size=sp.get("H", 0), rounds=r["rounds"], expanded=r["expanded"],
                             dev_ms=r["device_ms"], us_per_round=1e3 * r["device_ms"] / max(1, r["rounds"]),
                             lat_ms=lat[k] if lat else None))
        by = {}
        for kind in ("grid", "puzzle"):
            rr = [x for x in rows if x["kind"] == kind]
            if rr:
                by[kind] = {"n": len(rr), "rounds_p50": pct([x["rounds"] for x in rr], 50),
                            "rounds_max": max(x["rounds"] for x in rr),
                            "us_per_round_p50": pct([x["us_per_round"] for x in rr], 50),
                            "expanded_p50": pct([x["expanded"] for x in rr], 50)}
        return rows, by
    srows, out["serve_diag"] = diag(sres, slat)
    done_at = [a + l for a, l in zip(arrivals, slat)]
    out["serve_tail"] = [dict(kind=specs[k]["kind"], size=specs[k].get("H", 0), arrival_ms=round(arrivals[k], 1),
                              done_ms=round(done_at[k], 1), rounds=sres[k]["rounds"],
                              device_ms=round(sres[k]["device_ms"], 1))
                         for k in sorted(range(len(specs)), key=lambda k: -done_at[k])[:5]]
    _, out["plain_diag"] = diag(pres)
    for x in sorted(srows, key=lambda x: -(x["lat_ms"] or 0))[:8]:
        say(f"slowest served: {x}")
    if not args.no_cpu:
        say("CPU baseline (oracle, uniform sample of the stream)")
        out["cpu_baseline"] = cpu_baseline_mixed(ctx, specs)
    for s in specs + warm:
        if s["kind"] == "grid":
            s["grid"].free()
    ctx.close()
    return out


def run_deep_grid(args):
    """The serving stream's deepest grid (largest start-goal Chebyshev distance: 7786^2,
    p = 0.2, 7149 steps apart) solved alone by one one-block team (ua_batch_solve of one
    query), `steps` times after `warmup` small grids. Its rounds are set by the path's
    depth, so the time per round is the figure that bounds the serving stream's tail."""
    import uastar_amd as U
    ctx = U.Context(0, grid_heaps=args.grid_heaps, team_waves=args.team_waves)
    specs = [s for s in mixed_stream(204) if s["kind"] == "grid"]
    deep = max(specs, key=lambda s: max(abs(s["s"][0] - s["g"][0]), abs(s["s"][1] - s["g"][1])))
    warm = [s for s in specs if s["H"] < 1000][: args.warmup]
    materialize(ctx, warm + [deep])
    for s in warm:
        ctx.batch_solve([ctx.grid_query(s["grid"], *s["s"], *s["g"])])
    q = ctx.grid_query(deep["grid"], *deep["s"], *deep["g"])
    res, lat = [], []
    for i in range(args.steps):
        t = time.perf_counter()
        res.append(ctx.batch_solve([q])[0])
        lat.append((time.perf_counter() - t) * 1e3)
        say(f"deep grid solve {i}: {lat[-1]:.1f} ms, {res[-1]['rounds']} rounds")
    r = res[-1]
    dev = [x["device_ms"] for x in res]
    alg = [grid_bytes_alg(x) for x in res]
    out = {
        "metric": "queries/s + p50/p99 latency, batched serving vs plain GA*, 1 MI355X, exact",
        "value": len(lat) / (sum(lat) / 1e3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": len(lat),
        "warmup": len(warm),
        "ms_per_step": float(np.mean(lat)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded generators)",
        "config": {"workload": f"the serving stream's deepest grid alone: {deep['H']}x{deep['W']}, "
                               f"p = {deep['p']:.3f}, start {tuple(deep['s'])}, goal {tuple(deep['g'])}",
                   "mode": f"batch of one query (a team of {r['heaps']} waves)"},
        "errors": sum(x["status"] < 0 for x in res),
        "found": bool(r["found"]), "orth": r["orth"], "diag": r["diag"],
        "rounds": r["rounds"], "expanded": r["expanded"], "pops": r["pops"], "inserted": r["inserted"],
        "device_ms": float(np.mean(dev)), "us_per_round": 1e3 * float(np.mean(dev)) / max(1, r["rounds"]),
        "roofline": {"bound": "hbm", "achieved": sum(alg) / (sum(dev) * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": sum(alg) / (sum(dev) * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_teams<1, 0> (batch launch)", "kernel_ms_avg": float(np.mean(dev))},
    }
    for s in warm + [deep]:
        s["grid"].free()
    ctx.close()
    return out


# ------------------------------------------------------------------ 15-puzzles plain, and the batch

def cpu_puzzle_times(pdb, rows, budget_s, threads, puzzle_cap=4_000_000, seed=13, n=4):
    """Single-thread oracle A* times (the PDB of side n) of a seeded random sample of
    fixture rows, one puzzle per host thread, sized to about budget_s seconds on `threads`
    threads. A puzzle beyond puzzle_cap expansions runs to puzzle_cap and its time is
    scaled by its fixture expansion count. Returns (times, extrapolated count, wall seconds)."""
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    order = [int(i) for i in rng.permutation(len(rows))]
    sample, work = [], 0.0
    for i in order:
        if len(sample) >= threads and work >= budget_s * threads:
            break
        sample.append(i)
        work += 3.5e-6 * min(rows[i]["expanded"], puzzle_cap)

    def solve(i):
        r = rows[i]
        capped = r["expanded"] > puzzle_cap
        t = time.perf_counter()
        O.puzzle_solve(n, np.array(r["board"], np.uint8), pdb, "astar", node_limit=puzzle_cap if capped else 0)
        dt = time.perf_counter() - t
        return (dt * r["expanded"] / puzzle_cap, True) if capped else (dt, False)

    t = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        out = list(ex.map(solve, sample))
    return [x[0] for x in out], sum(x[1] for x in out), time.perf_counter() - t


def device_pdb(ctx, n=4):
    """The device-built PDB of side n (7-8 for the 15-puzzle, 6-6-6-6 for the 24-puzzle),
    checked against the fixture digests (the CPU BFS takes minutes)."""
    import hashlib
    pdb = ctx.puzzle_download_pdb(n)
    with open(os.path.join(ROOT, "tests", "golden", "pdb_digests.json")) as f:
        digests = json.load(f)[str(n)]
    off = 0
    for d in digests:
        assert hashlib.sha256(pdb[off: off + d["size"]].tobytes()).hexdigest() == d["sha256"], "PDB digest"
        off += d["size"]
    return pdb


def run_plain_puzzle(args):
    """The 256 random solvable 15-puzzles of seeds 3000..3255 (the fixture's boards and
    CPU A* move counts), one plain whole-GPU launch each; `steps` of them are timed after
    `warmup` others."""
    import uastar_amd as U
    with open(GOLDEN) as f:
        rows = json.load(f)
    n = max(1, min(args.steps, len(rows)))
    timed, warm = rows[:n], rows[n:n + args.warmup] or rows[:1]
    ctx = U.Context(0, graphs=args.graph, plain_heaps=args.plain_heaps)
    ctx.puzzle_build_pdb(4)
    for r in warm:
        ctx.puzzle_solve(4, np.array(r["board"], np.uint8))
    say(f"timing {n} plain 15-puzzles")
    lat, res = [], []
    t0 = time.perf_counter()
    for r in timed:
        t = time.perf_counter()
        res.append(ctx.puzzle_solve(4, np.array(r["board"], np.uint8)))
        lat.append((time.perf_counter() - t) * 1e3)
    wall = time.perf_counter() - t0
    mismatch = sum(not x["found"] or x["steps"] != r["steps"] for x, r in zip(res, timed))
    alg = [puzzle_bytes_alg(x) for x in res]
    kms = [x["kernel_ms"] for x in res]
    achieved = sum(alg) / (sum(kms) * 1e-3) / 1e9
    out = {
        "metric": "queries/s + p50/p99 latency, batched serving vs plain GA*, 1 MI355X, exact",
        "value": n / wall, "unit": "queries/s", "n_gpus": 1, "steps": n, "warmup": len(warm),
        "ms_per_step": wall / n * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int64", "data": "synthetic (seeded generators)",
        "config": {"workload": f"plain GA*: {n} random solvable 15-puzzles (seeds 3000..{3000 + n - 1}), "
                               "7-8 PDB built on the device, one query per whole-GPU launch",
                   "mode": "plain", "heaps": sorted({x["heaps"] for x in res})},
        "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
        "exact_mismatches_vs_cpu_astar": mismatch,
        "errors": sum(x["status"] < 0 for x in res),
        "expanded_per_s": sum(x["expanded"] for x in res) / wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_teams<0, 0> (plain solve)",
                     "kernel_ms_avg": float(np.mean(kms)), "alg_bytes_avg": float(np.mean(alg))},
    }
    say(f"GPU: {out['value']:.2f} queries/s, p50 {out['p50_ms']:.2f} ms, {mismatch} mismatches")
    if not args.no_cpu:
        say("CPU baseline (oracle A*, bounded sample)")
        threads = min(16, os.cpu_count() or 1)
        times, n_ext, cwall = cpu_puzzle_times(device_pdb(ctx), timed, 20.0, threads)
        model, nproc = host_cpu()
        out["cpu_baseline"] = {
            "value": threads / float(np.mean(times)), "unit": "queries/s", "cores": threads, "kind": "port",
            "nproc": nproc, "cpu_model": model,
            "sample": f"seeded random sample of {len(times)} of the {n} timed puzzles, solved by the oracle's A* "
                      f"(7-8 PDB, reference containers) one per thread on {threads} threads in {cwall:.1f} s; "
                      f"value = {threads} threads / mean single-thread time ({np.mean(times) * 1e3:.0f} ms); "
                      f"{n_ext} time(s) extrapolated from 4M expansions by the fixture's A* expansion counts"}
    ctx.close()
    return out


def run_plain_puzzle5(args):
    """The 24-puzzle fixture boards (seeded walks off the goal, 22-68 moves, each solved
    by the oracle's A* with the 6-6-6-6 PDB and by IDA* with Manhattan distance), one plain
    whole-GPU launch each, after one warm-up solve. Every move count is checked against
    the fixture."""
    import uastar_amd as U
    with open(os.path.join(ROOT, "tests", "golden", "c5_puzzles.json")) as f:
        rows = sorted(json.load(f), key=lambda r: r["seed"])
    n = max(1, min(args.steps, len(rows)))
    timed = rows[:n]
    warm = min(rows, key=lambda r: r["expanded"])
    ctx = U.Context(0, graphs=args.graph, plain_heaps=args.plain_heaps)
    say("building the 5x5 PDB on the device")
    ctx.puzzle_build_pdb(5)
    for _ in range(max(1, args.warmup)):
        ctx.puzzle_solve(5, np.array(warm["board"], np.uint8))
    say(f"timing {n} plain 24-puzzles")
    lat, res = [], []
    t0 = time.perf_counter()
    for r in timed:
        t = time.perf_counter()
        res.append(ctx.puzzle_solve(5, np.array(r["board"], np.uint8)))
        lat.append((time.perf_counter() - t) * 1e3)
    wall = time.perf_counter() - t0
    mismatch = sum(not x["found"] or x["steps"] != r["steps"] for x, r in zip(res, timed))
    alg = [puzzle_bytes_alg(x) for x in res]
    kms = [x["kernel_ms"] for x in res]
    achieved = sum(alg) / (sum(kms) * 1e-3) / 1e9
    out = {
        "metric": "queries/s + p50/p99 latency, batched serving vs plain GA*, 1 MI355X, exact",
        "value": n / wall, "unit": "queries/s", "n_gpus": 1, "steps": n, "warmup": max(1, args.warmup),
        "ms_per_step": wall / n * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int64", "data": "fixture boards (seeded walks off the goal)",
        "config": {"workload": f"plain GA*: the {n} 24-puzzle fixture boards (seeds "
                               f"{', '.join(str(r['seed']) for r in timed)}), 6-6-6-6 PDB built on the "
                               "device, one query per whole-GPU launch",
                   "mode": "plain", "heaps": sorted({x["heaps"] for x in res})},
        "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
        "per_query": [{"seed": r["seed"], "steps": x["steps"], "expanded": x["expanded"], "ms": round(t, 3)}
                      for r, x, t in zip(timed, res, lat)],
        "exact_mismatches_vs_cpu_astar": mismatch,
        "errors": sum(x["status"] < 0 for x in res),
        "expanded_per_s": sum(x["expanded"] for x in res) / wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_teams<0, 1> (plain solve, 5x5)",
                     "kernel_ms_avg": float(np.mean(kms)), "alg_bytes_avg": float(np.mean(alg))},
    }
    say(f"GPU: {out['value']:.2f} queries/s, p50 {out['p50_ms']:.2f} ms, {mismatch} mismatches")
    if not args.no_cpu:
        say("CPU baseline (oracle A*, bounded sample)")
        threads = min(16, os.cpu_count() or 1)
        times, n_ext, cwall = cpu_puzzle_times(device_pdb(ctx, 5), timed, 20.0, threads, puzzle_cap=1_000_000, n=5)
        model, nproc = host_cpu()
        out["cpu_baseline"] = {
            "value": threads / float(np.mean(times)), "unit": "queries/s", "cores": threads, "kind": "port",
            "nproc": nproc, "cpu_model": model,
            "sample": f"seeded random sample of {len(times)} of the {n} timed boards, solved by the oracle's A* "
                      f"(6-6-6-6 PDB, reference containers) one per thread on {threads} threads in {cwall:.1f} s; "
                      f"value = {threads} threads / mean single-thread time ({np.mean(times) * 1e3:.0f} ms); "
                      f"{n_ext} time(s) extrapolated from 1M expansions by the fixture's A* expansion counts"}
    ctx.close()
    return out


def replay_board(n, board, moves):
    """The board after the blank moves (0..3 = down, up, right, left), or None if a move
    leaves the board."""
    b = [int(t) for t in board]
    z = b.index(0)
    for m in moves:
        x, y = z // n + (1, -1, 0, 0)[m], z % n + (0, 0, 1, -1)[m]
        if not (0 <= x < n and 0 <= y < n):
            return None
        b[z], b[x * n + y] = b[x * n + y], b[z]
        z = x * n + y
    return b


def run_bounded_puzzle5(args):
    """The memory-bounded scheme on the three hardest 24-puzzle fixture boards (56, 64 and
    68 moves; 3.0-7.0 M A* expansions on the CPU) under bounds from 4 GiB down to 64 MiB:
    plain A* where it fits (an optimal answer), else weighted searches in the same memory.
    Every returned move list is replayed to the goal and its length set against the
    fixture's optimum."""
    import uastar_amd as U
    with open(os.path.join(ROOT, "tests", "golden", "c5_puzzles.json")) as f:
        rows = sorted(json.load(f), key=lambda r: -r["expanded"])[:3]
    bounds = [4 << 30, 1 << 30, 256 << 20, 64 << 20]
    goal = list(range(1, 25)) + [0]
    ctx = U.Context(0, graphs=args.graph, plain_heaps=args.plain_heaps)
    say("building the 5x5 PDB on the device")
    ctx.puzzle_build_pdb(5)
    ctx.puzzle_solve_bounded(5, np.array(min(rows, key=lambda r: r["expanded"])["board"], np.uint8), bounds[0])
    say(f"timing {len(rows)} boards x {len(bounds)} memory bounds")
    solves, bad = [], 0
    t0 = time.perf_counter()
    for bound in bounds:
        for r in rows:
            t = time.perf_counter()
            x = ctx.puzzle_solve_bounded(5, np.array(r["board"], np.uint8), bound)
            ms = (time.perf_counter() - t) * 1e3
            mv = ctx.puzzle_moves() if x["found"] else []
            ok = bool(x["found"]) and len(mv) == x["steps"] and replay_board(5, r["board"], mv) == goal
            ok = ok and x["steps"] >= r["steps"] and (not x["optimal"] or x["steps"] == r["steps"])
            bad += not ok
            solves.append({"seed": r["seed"], "bound_mib": bound >> 20, "steps": x["steps"], "optimum": r["steps"],
                           "optimal": x["optimal"], "weight": x["weight"], "expanded": x["expanded"],
                           "kernel_ms": round(x["kernel_ms"], 3), "ms": round(ms, 3),
                           "alg_bytes": puzzle_bytes_alg(x)})
    wall = time.perf_counter() - t0
    kms = sum(s["kernel_ms"] for s in solves)
    achieved = sum(s["alg_bytes"] for s in solves) / (kms * 1e-3) / 1e9
    out = {
        "metric": "queries/s + p50/p99 latency, batched serving vs plain GA*, 1 MI355X, exact",
        "value": len(solves) / wall, "unit": "queries/s", "n_gpus": 1, "steps": len(solves), "warmup": 1,
        "ms_per_step": wall / len(solves) * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int64", "data": "fixture boards (seeded walks off the goal)",
        "config": {"workload": "memory-bounded scheme: the 24-puzzle fixture boards of seeds "
                               f"{', '.join(str(r['seed']) for r in rows)} under bounds of "
                               f"{', '.join(str(b >> 20) for b in bounds)} MiB, one solve each",
                   "mode": "plain (bounded)"},
        "p50_ms": pct([s["ms"] for s in solves], 50), "p99_ms": pct([s["ms"] for s in solves], 99),
        "solves": solves,
        "invalid_solutions": bad,
        "errors": 0,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_teams<0, 1> (plain and weighted solves, 5x5)",
                     "kernel_ms_avg": kms / len(solves),
                     "alg_bytes_avg": float(np.mean([s["alg_bytes"] for s in solves]))},
    }
    say(f"GPU: {out['value']:.2f} solves/s, {bad} invalid")
    if not args.no_cpu:
        say("CPU baseline (oracle A*, unbounded, bounded sample)")
        threads = min(16, os.cpu_count() or 1)
        times, n_ext, cwall = cpu_puzzle_times(device_pdb(ctx, 5), rows, 10.0, threads, puzzle_cap=1_000_000, n=5)
        model, nproc = host_cpu()
        out["cpu_baseline"] = {
            "value": threads / float(np.mean(times)), "unit": "queries/s", "cores": threads, "kind": "port",
            "nproc": nproc, "cpu_model": model,
            "sample": f"the oracle's plain A* (6-6-6-6 PDB, reference containers, no memory bound: the reference "
                      f"has no bounded scheme) on the {len(times)} boards, one per thread, in {cwall:.1f} s; "
                      f"value = {threads} threads / mean single-thread time ({np.mean(times) * 1e3:.0f} ms); "
                      f"{n_ext} time(s) extrapolated from 1M expansions by the fixture's A* expansion counts"}
    ctx.close()
    return out


def run_batch_c4(args):
    """The batched configuration: 4096 grids (side log-uniform in [256, 4096], p ~ U(0.1,
    0.4), random open endpoints) plus the 512 random solvable 15-puzzles of seeds
    41000..41511, solved by one ua_batch_solve call (persistent launches over
    memory-bounded sub-batches). The same batch runs once untimed first: its device
    arena (the budget's worth of HBM, allocated at about 40 GB/s) stays for the timed
    call; the first call's wall time is reported beside."""
    import uastar_amd as U
    from tests.helpers import c4_grid_specs
    gspecs = c4_grid_specs()
    with open(os.path.join(ROOT, "tests", "golden", "c4_puzzles.json")) as f:
        prows = json.load(f)
    ctx = U.Context(0, max_rounds=1_000_000, graphs=args.graph)
    ctx.puzzle_build_pdb(4)
    say("materialising the batch")
    grids = [ctx.grid_generate(s["H"], s["W"], s["seed"], s["p"], *s["s"], *s["g"]) for s in gspecs]
    queries = [ctx.grid_query(g, *s["s"], *s["g"]) for g, s in zip(grids, gspecs)]
    queries += [ctx.puzzle_query(4, r["board"]) for r in prows]
    say(f"first (untimed) call: {len(queries)} queries")
    t = time.perf_counter()
    ctx.batch_solve(queries)
    first = time.perf_counter() - t
    say(f"timing one batch of {len(queries)} queries")
    t = time.perf_counter()
    res = ctx.batch_solve(queries)
    wall = time.perf_counter() - t
    gres, pres = res[:len(gspecs)], res[len(gspecs):]
    mismatch = sum(not x["found"] or x["steps"] != r["steps"] for x, r in zip(pres, prows))
    cells = sum(s["H"] * s["W"] for s in gspecs)
    alg = sum(grid_bytes_alg(x) for x in gres) + sum(puzzle_bytes_alg(x) for x in pres)
    out = {
        "metric": "queries/s + p50/p99 latency, batched serving vs plain GA*, 1 MI355X, exact",
        "value": len(queries) / wall, "unit": "queries/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int64", "data": "synthetic (seeded generators)",
        "config": {"workload": f"one batch of {len(gspecs)} grids ({cells / 1e9:.2f} G cells, side log-uniform in "
                               f"[256, 4096], p ~ U(0.1, 0.4), random open endpoints, seed stream 4) plus "
                               f"{len(prows)} random solvable 15-puzzles (seeds 41000..41511)",
                   "mode": "batch (persistent launches over memory-bounded sub-batches)"},
        "puzzle_mismatches_vs_cpu_astar": mismatch,
        "grids_found": sum(x["found"] for x in gres),
        "errors": sum(x["status"] < 0 for x in res),
        "expanded_per_s": sum(x["expanded"] for x in res) / wall,
        "algorithmic_GBps": alg / wall / 1e9,
        "first_call_s": first,
    }
    say(f"GPU: {out['value']:.1f} queries/s ({wall:.2f} s for the batch), {mismatch} puzzle mismatches")
    if not args.no_cpu:
        from oracle import oracle as O
        say("CPU baseline (oracle, bounded sample, extrapolated to the batch)")
        threads = min(16, os.cpu_count() or 1)
        rng = np.random.default_rng(17)
        gi = [int(i) for i in rng.permutation(len(gspecs))[:2 * threads]]

        def gsolve(i):
            s = gspecs[i]
            g = O.grid_bernoulli(s["H"], s["W"], s["seed"], s["p"], *s["s"], *s["g"])
            t = time.perf_counter()
            O.grid_astar(g, s["H"], s["W"], *s["s"], *s["g"], "f32", want_path=False)
            return time.perf_counter() - t

        t = time.perf_counter()
        with cf.ThreadPoolExecutor(threads) as ex:
            gt = list(ex.map(gsolve, gi))
        gwall = time.perf_counter() - t
        pt, n_ext, pwall = cpu_puzzle_times(device_pdb(ctx), prows, 10.0, threads)
        est = float(np.mean(gt)) * len(gspecs) + float(np.mean(pt)) * len(prows)  # single-thread seconds
        model, nproc = host_cpu()
        out["cpu_baseline"] = {
            "value": len(queries) / (est / threads), "unit": "queries/s", "cores": threads, "kind": "port",
            "nproc": nproc, "cpu_model": model,
            "sample": f"seeded random samples of {len(gi)} grids ({gwall:.1f} s) and {len(pt)} puzzles ({pwall:.1f} s) "
                      f"of the batch, solved by the oracle (FP32 grid costs, reference containers; A* with the 7-8 PDB) "
                      f"one per thread on {threads} threads; the batch's single-thread time extrapolated from the "
                      f"samples' means ({est:.0f} s), value = queries / (that time / {threads} threads); {n_ext} puzzle "
                      f"time(s) extrapolated from 4M expansions",
            "extrapolated_batch_s": est / threads}
    for g in grids:
        g.free()
    ctx.close()
    return out


# ------------------------------------------------------------------ main

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="serving stream: batches of %d arrivals; plain workloads: queries" % QUERIES_PER_STEP)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="serve-mixed",
                    choices=["serve-mixed", "plain-grid1000", "plain-grid16384", "deep-grid", "plain-puzzle",
                             "plain-puzzle5", "bounded-puzzle5", "batch-c4"])
    ap.add_argument("--load", type=float, default=20.0, help="arrival rate as a multiple of plain capacity")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default settings)")
    ap.add_argument("--graph", action="store_true",
                    help="replay each search launch from a hipGraph instead of issuing it call by call")
    ap.add_argument("--plain-heaps", type=int, default=0,
                    help="heaps (waves) of the plain whole-GPU solve of the plain workloads (0 = default)")
    ap.add_argument("--grid-heaps", action="store_true",
                    help="grids keep one sorted heap per wave instead of the team's bucket ring (comparison)")
    ap.add_argument("--team-waves", type=int, default=0,
                    help="deep-grid: waves of the grid's team (multiple of 16; 0 = the default one block)")
    ap.add_argument("--xcd-shared", action="store_true",
                    help="serving: every team class on every XCD (no XCDs apart for one-block grid teams)")
    ap.add_argument("--atomic-probe", action="store_true",
                    help="only measure the device's random 32-bit atomicMin throughput (ua_debug_atomic_rate)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (rocprofv3 counter calibration)")
    args = ap.parse_args()
    if args.copy_calib:
        print(json.dumps(copy_calib(args.copy_calib)))
        return
    if args.atomic_probe:
        import uastar_amd as U
        with U.Context(0) as ctx:
            out = {f"blocks_{b}_gatomics_per_s": ctx.atomic_rate(26, b, 64) for b in (256, 1024, 4096)}
            kinds = ("load_wg", "atomicmin_wg", "atomicmin_agent", "load_agent")
            for lg, size in ((14, "128KiB"), (17, "1MiB"), (25, "256MiB")):
                for k, name in enumerate(kinds):
                    out[f"latency_ns_{name}_{size}"] = ctx.access_latency(lg, k)
            print(json.dumps(out))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU: every query is solved on a single MI355X "
              "(multi-GPU sharding is a design note, not measured).", file=sys.stderr)
        sys.exit(2)

    import uastar_amd as U
    if U.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.workload == "serve-mixed":
        print(json.dumps(run_serve_mixed(args)))
        return
    if args.workload == "plain-puzzle":
        print(json.dumps(run_plain_puzzle(args)))
        return
    if args.workload == "plain-puzzle5":
        print(json.dumps(run_plain_puzzle5(args)))
        return
    if args.workload == "bounded-puzzle5":
        print(json.dumps(run_bounded_puzzle5(args)))
        return
    if args.workload == "batch-c4":
        print(json.dumps(run_batch_c4(args)))
        return
    if args.workload == "deep-grid":
        print(json.dumps(run_deep_grid(args)))
        return
    ctx = U.Context(0, graphs=args.graph, grid_heaps=args.grid_heaps, plain_heaps=args.plain_heaps)

    big = args.workload == "plain-grid16384"
    specs = grid16384_specs(args.steps) if big else grid1000_specs(args.steps)
    warm = grid16384_specs(args.warmup, seed0=1_000_000) if big else grid1000_specs(args.warmup, seed0=1_000_000)
    wall, lat, res = run_plain(ctx, specs, warm)
    qps = len(lat) / wall
    bad = [r for r in res if r["status"] < 0]  # errors; a grid with no path is an answer
    no_path = sum(r["status"] == 1 for r in res)
    alg = [grid_bytes_alg(r) for r in res]
    kms = [r["kernel_ms"] for r in res]
    achieved = sum(alg) / (sum(kms) * 1e-3) / 1e9
    out = {
        "metric": "queries/s + p50/p99 latency, batched serving vs plain GA*, 1 MI355X, exact",
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": len(lat),
        "warmup": args.warmup,
        "ms_per_step": wall / len(lat) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded generators)",
        "config": {"workload": ("plain GA*: 16384x16384 grids, Bernoulli(0.35) obstacles from seeds 2.., "
                                "start (0,0), goal (16383,16383), one query per whole-GPU launch") if big else
                               ("plain GA*: 1000x1000 grids, Bernoulli(0.30) obstacles from seeds 1.., "
                                "start (0,0), goal (999,999), one query per whole-GPU launch"),
                   "mode": "plain", "heaps": sorted({r["heaps"] for r in res}),
                   "open_list": "per-wave heaps" if args.grid_heaps or res[0]["heaps"] > 16 else "bucket ring"},
        "rounds_p50": pct([r["rounds"] for r in res], 50),
        "answers": [(r["orth"], r["diag"]) for r in res[:4]],
        "p50_ms": pct(lat, 50),
        "p99_ms": pct(lat, 99),
        "errors": len(bad),
        "no_path": no_path,
        "expanded_per_s": sum(r["expanded"] for r in res) / wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_teams<0> (plain solve)", "kernel_ms_avg": float(np.mean(kms)),
                     "alg_bytes_avg": float(np.mean(alg))},
    }
    say(f"GPU: {qps:.2f} queries/s, p50 {out['p50_ms']:.2f} ms, {len(bad)} errors")
    if not args.no_cpu and not big:
        say("CPU baseline (oracle, bounded sample)")
        out["cpu_baseline"] = cpu_baseline_grid(specs)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
