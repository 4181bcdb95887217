"""Benchmark of the k-core engine on one MI355X: single-graph, batched and served peeling.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale 24] [--no-tune]
                    [--batch-graphs 4096] [--serve-seconds 10] [--serve-rate R] [--serve-sweep a,b,..]
                    [--only single|batch|serve] [--no-cpu-baseline]

Prints one JSON line. Its headline (metric/value) is the single-graph peel of an R-MAT scale-24
graph; "batched" and "served" hold the other two workloads.

single  A step is one full decomposition (kc_peel) of the resident graph: from "CSR + degrees in
        HBM" to "core numbers + kmax in HBM". Seeded R-MAT, edge factor 16, a=.57 b=.19 c=.19
        d=.05, symmetrised, deduplicated, loop-free; uploaded once; settings from the autotuner
        (kc_tune_peel); W untimed + K timed steps. value = undirected edges m / seconds per step.
        roofline: achieved = B_alg / average peel-kernel duration (HIP events on its stream),
        B_alg = 4 * (degree entries scanned) + 16 n + 20 m bytes per peel.
batched 4096 graphs, n log-uniform in [1K, 200K], mean degree uniform in [4, 32], even ones
        Erdos-Renyi, odd ones Chung-Lu (exponent 2.1), packed into one CSR and peeled by one
        persistent kernel launched as a captured HIP graph, block size from kc_tune_batch.
        Compared with a loop of the single-graph kernel, one launch per graph, on the same GPU;
        every graph's core numbers must agree bit for bit between the two.
served  kcore_serve: open-loop Poisson requests (edge counts lognormal, median 50K, sigma 1.0,
        truncated to [1K, 50M]; R-MAT or Chung-Lu by a seeded coin flip) answered by the
        resident serving kernel and by the plain per-request path (upload, peel, download,
        free). The served path's capacity is measured first (all requests at once); the
        open-loop rate is 0.7 x that capacity unless --serve-rate names one (--serve-sweep adds
        rates). Every served result is checked against core numbers computed by the oracle.
cpu_baseline  the CPU level-synchronous restatement of the reference's peel (oracle), run on
        the host threads on the same R-MAT graph; its core numbers are compared with the GPU's.

The autotuning and the batched leg run in child processes (this script with --leg), started before
this process touches the GPU, so that a failure there cannot take the headline line with it: the
headline then uses the default settings and says so (detail.settings_from).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBPS = 8000.0
PROFILE = os.path.join(HERE, "profiles", "peel_persistent_rmat24.md")
SERVE_BIN = os.path.join(HERE, "kcoregpu_amd", "bin", "kcore_serve")
PARAMS = ("variant", "block", "heavy_threshold", "chunk_arcs", "lds_depth", "compact_ratio", "poll_max")
KERNEL_OF = {1: "peel_persistent", 2: "peel_persistent", 3: "peel_persistent", 5: "peel_async"}


def profiled_traffic(args, tuned):
    """HBM bytes per launch of the peel kernel, from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this same workload and settings (None when they differ)."""
    if (args.scale, args.edge_factor, args.seed) != (24, 16, 3) or not os.path.exists(PROFILE):
        return None
    want = None
    vals = {}
    for line in open(PROFILE):
        if ":" in line and line.split(":")[0].strip() in ("traffic_bytes_per_launch", "settings"):
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
    if "traffic_bytes_per_launch" not in vals:
        return None
    want = vals.get("settings")
    have = " ".join(f"{k}={tuned[k]}" for k in PARAMS)
    if want is not None and want != have:
        return None
    return int(vals["traffic_bytes_per_launch"])


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_threads():
    threads = os.cpu_count() or 1
    aff = getattr(os, "sched_getaffinity", None)
    if aff:
        threads = len(aff(0))
    return min(threads, 16)


# ---------------------------------------------------------------------------------------------
DEFAULT_PARAMS = {"variant": 2, "block": 512, "heavy_threshold": 64, "chunk_arcs": 512, "lds_depth": 0,
                  "compact_ratio": 4, "poll_max": 2}


def bench_single(args, K, eng, tune_cache=None, tuned_by=None, tune_only=False):
    """tune_cache: settings tuned by a child process (answered from the cache here); None: the
    defaults. tune_only: tune, print the settings and return (the child's job)."""
    import numpy as np
    t0 = time.time()
    log(f"[single] generating R-MAT scale {args.scale} ef {args.edge_factor} seed {args.seed}")
    g = K.Graph.rmat(args.scale, args.seed, edge_factor=args.edge_factor)
    n, m = g.V, g.m
    log(f"[single] V={n} m={m} dmax={g.dmax} generated in {time.time() - t0:.1f}s")
    t1 = time.time()
    dg = eng.upload(g)
    h2d_s = time.time() - t1
    names = PARAMS
    if tune_cache is None:
        tuned = dict(DEFAULT_PARAMS, trials=0, ms=None)
    else:
        tt = time.time()
        tuned = dg.tune(reps=2, cache_path=tune_cache)
        if tuned_by:
            tuned["trials"] = tuned_by.get("trials", tuned["trials"])
        log(f"[single] tuned in {time.time() - tt:.1f}s ({tuned['trials']} trials): "
            + " ".join(f"{k}={tuned[k]}" for k in names))
        if tune_only:
            return tuned
    prm = {k: tuned[k] for k in names}

    def step():
        return dg.peel(download=False, **prm)

    for i in range(args.warmup):
        r = step()
        log(f"[single] warmup {i}: {r.stats['ms_peel']:.3f} ms kmax={r.kmax}")
    kernel_ms, peel_ms = [], []
    # kc_peel synchronises its stream before returning, so the host clock brackets finished work
    t_start = time.perf_counter()
    for _ in range(args.steps):
        r = step()
        kernel_ms.append(r.stats["ms_kernel"])
        peel_ms.append(r.stats["ms_peel"])
    t_total = time.perf_counter() - t_start
    ms_per_step = 1e3 * t_total / args.steps
    st = r.stats
    log(f"[single] {args.steps} steps: {ms_per_step:.3f} ms/step, kernel {np.mean(kernel_ms):.3f} ms")
    k_avg_s = float(np.mean(kernel_ms)) / 1e3
    achieved = st["bytes_alg"] / k_avg_s / 1e9
    core = dg.core()
    # the box's own HBM ceiling beside the published peak: the degree-copy kernel over 2 x 1 GiB
    try:
        copy_gbps = eng.copy_bandwidth(1 << 30, 10)
        log(f"[single] copy-kernel ceiling {copy_gbps:.0f} GB/s")
    except Exception as e:  # noqa: BLE001 - reported as null
        log(f"[single] copy-kernel ceiling failed: {e!r}")
        copy_gbps = None
    # the reference's own shape on the same GPU: host level loop over every level 0..kmax, one
    # select and per-cascade process launches per level, frontier size read back each time
    rs = dg.peel("stepwise", flags=K.FLAG_NO_LEVEL_SKIP, download=True)
    ref_shaped = {"ms": rs.stats["ms_peel"], "levels": rs.stats["levels"], "launches": rs.stats["subrounds"],
                  "identical": bool((rs.core == core).all()),
                  "speedup": rs.stats["ms_peel"] / (ms_per_step)}
    log(f"[single] reference-shaped loop: {rs.stats['ms_peel']:.1f} ms, {rs.stats['subrounds']} launches")

    cpu = None
    if not args.no_cpu_baseline:
        import oracle as O
        threads = host_threads()
        O.levelsync(g.neighbors_offset, g.neighbors, nthreads=threads)  # warm
        runs = []
        for _ in range(5):
            tc = time.perf_counter()
            res = O.levelsync(g.neighbors_offset, g.neighbors, nthreads=threads)
            runs.append(time.perf_counter() - tc)
        cpu_s = float(np.median(runs))
        same = bool((res["core"] == core).all())
        tb = time.perf_counter()
        bz_core, _ = O.bz(g.neighbors_offset, g.neighbors)
        bz_s = time.perf_counter() - tb
        cpu = {"value": m / cpu_s, "unit": "edges/s", "cores": threads, "kind": "port",
               "sample": f"the full R-MAT scale {args.scale} graph: level-synchronous restatement of the "
                         f"reference peel on {threads} threads, median of 5 after a warm-up "
                         f"({cpu_s:.2f} s, kmax {res['kmax']}); GPU core numbers "
                         f"{'identical' if same else 'DIFFERENT'}",
               "gpu_matches": same and bool((bz_core == core).all()),
               "bz_single_thread_s": bz_s, "cpu_model": cpu_model()}
        log(f"[single] cpu baseline {cpu_s:.2f}s on {threads} threads, BZ {bz_s:.2f}s; identical={same}")
    out = {
        "value": m / (ms_per_step / 1e3), "ms_per_step": ms_per_step,
        "config": {"workload": f"single-graph k-core peel, R-MAT scale {args.scale} edge factor "
                               f"{args.edge_factor} (a=.57 b=.19 c=.19 d=.05) seed {args.seed}, "
                               "uint32 CSR resident in HBM",
                   "n": n, "m": m, "arcs": g.arcs, "dmax": g.dmax, "kmax": r.kmax,
                   "levels_scanned": st["levels"], "subrounds": st["subrounds"],
                   "grid": st["grid"], "participants": st["participants"], "tuned": tuned},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": profiled_traffic(args, tuned),
                     "copy_ceiling": copy_gbps,
                     "frac_of_copy_ceiling": achieved / copy_gbps if copy_gbps else None,
                     "kernel": KERNEL_OF.get(tuned["variant"], "peel_persistent"),
                     "kernel_ms": float(np.mean(kernel_ms)),
                     "bytes_alg": st["bytes_alg"], "bytes_scan": st["bytes_scan"],
                     "bytes_edge": st["bytes_edge"]},
        "cpu_baseline": cpu,
        "detail": {"ms_peel_mean": float(np.mean(peel_ms)), "ms_kernel_min": float(np.min(kernel_ms)),
                   "h2d_s": h2d_s, "arcs_per_s": g.arcs / (ms_per_step / 1e3),
                   "reference_shaped_gpu": ref_shaped},
    }
    dg.free()
    return out


# ---------------------------------------------------------------------------------------------
def bench_batch(args, K, eng):
    import numpy as np
    G = args.batch_graphs
    t0 = time.time()
    b = K.Batch.generate(G, args.batch_seed)
    log(f"[batch] {G} graphs, {b.total_vertices} vertices, {b.total_arcs} arcs generated in "
        f"{time.time() - t0:.1f}s")
    db = eng.upload_batch(b)
    tuned = db.tune(reps=2, cache_path=args.tune_cache) if not args.no_tune else {"block": 512, "trials": 0}
    log(f"[batch] tuned block {tuned['block']} ({tuned['trials']} trials)")
    flags = K.FLAG_GRAPH
    db.peel(block=tuned["block"], flags=flags, download=False)  # capture + warm
    times = []
    t_start = time.perf_counter()
    for _ in range(args.batch_steps):
        _, _, st = db.peel(block=tuned["block"], flags=flags, download=False)
        times.append(st["ms_peel"])
    # B_alg summed over the graphs: per graph V_g entries scanned per level it ran, 16 n + 20 m
    b_alg = st["bytes_alg"]
    achieved = b_alg / (float(np.mean(times)) / 1e3) / 1e9
    wall = (time.perf_counter() - t_start) / args.batch_steps
    core, kmax, _ = db.peel(block=tuned["block"], flags=flags)
    log(f"[batch] {wall * 1e3:.2f} ms per batch peel ({G / wall:.0f} graphs/s)")

    # the plain path: one launch of the single-graph kernel per graph, graphs resident
    plain_s = 0.0
    mism = 0
    tl = time.time()
    for gi in range(G):
        gg = b.graph(gi)
        dg = eng.upload(gg)
        tp = time.perf_counter()
        r = dg.peel()
        plain_s += time.perf_counter() - tp
        if not (r.core == core[b.vertices(gi)]).all() or r.kmax != kmax[gi]:
            mism += 1
        dg.free()
    log(f"[batch] plain loop {plain_s:.2f}s of kc_peel ({G / plain_s:.0f} graphs/s), {mism} mismatches, "
        f"loop wall {time.time() - tl:.1f}s")
    out = {"graphs": G, "vertices": b.total_vertices, "arcs": b.total_arcs,
           "graphs_per_s": G / wall, "ms_per_batch": wall * 1e3,
           "edges_per_s": b.total_arcs / 2 / wall, "kernel_ms_mean": float(np.mean(times)),
           "bytes_alg": b_alg, "bytes_scan": st["bytes_scan"], "bytes_edge": st["bytes_edge"],
           "achieved_gbps": achieved, "frac_of_peak": achieved / HBM_PEAK_GBPS,
           "plain_graphs_per_s": G / plain_s, "speedup_vs_plain": (G / wall) / (G / plain_s),
           "bit_exact_vs_plain": mism == 0, "mismatched_graphs": mism,
           "block": tuned["block"], "hip_graph": True,
           "workload": f"{G} graphs, n log-uniform [1000, 200000], mean degree U[4, 32], ER / "
                       f"Chung-Lu 2.1 alternating, seed {args.batch_seed}"}
    db.free()
    return out


# ---------------------------------------------------------------------------------------------
def run_serve(argv, limit):
    p = subprocess.run([SERVE_BIN] + argv, capture_output=True, text=True, timeout=limit)
    if p.stderr.strip():
        log(p.stderr.strip())
    lines = [ln for ln in p.stdout.splitlines() if ln.startswith("{")]
    if not lines:
        raise RuntimeError(f"kcore_serve {' '.join(argv)} exited {p.returncode} without a result")
    return json.loads(lines[-1])


def bench_serve(args, K, expect_path):
    common = ["--pool", str(args.serve_pool), "--seed", str(args.serve_seed), "--seconds",
              str(args.serve_seconds)]
    if expect_path:
        common += ["--expect", expect_path]
    lim = int(args.serve_seconds * 3 + 120)
    res = {}
    keys = ("offered_rate", "seconds", "closed", "requests", "completed", "unfinished", "dropped", "overloaded",
            "graphs_per_s", "edges_per_s", "p50_ms", "p90_ms", "p99_ms", "max_ms", "mismatches", "checked_against")

    def leg(name, extra):
        r = run_serve(common + extra, lim)
        log(f"[serve] {name}: {r['graphs_per_s']:.0f} graphs/s p50 {r['p50_ms']:.2f} ms "
            f"p99 {r['p99_ms']:.2f} ms mismatches {r['mismatches']} unfinished {r['unfinished']}")
        res[name] = {k: r[k] for k in keys}
        return r

    # the served path's capacity (all requests at once), then the open-loop stream at 0.7 x that
    # rate unless --serve-rate names one
    cap = leg("served_capacity", ["--mode", "served", "--closed", str(args.serve_closed)])
    rate = args.serve_rate if args.serve_rate else round(0.7 * cap["graphs_per_s"])
    leg("served", ["--mode", "served", "--rate", str(rate)])
    leg("served_at_plain_rate", ["--mode", "served", "--rate", str(args.plain_rate)])
    leg("plain", ["--mode", "plain", "--rate", str(args.plain_rate), "--drain-limit", "30"])
    leg("plain_capacity", ["--mode", "plain", "--closed", str(args.plain_closed)])
    for r in args.serve_sweep:
        leg(f"served_at_{int(r)}", ["--mode", "served", "--rate", str(r)])
    res["rate_rule"] = ("--serve-rate" if args.serve_rate else "0.7 x served_capacity") + f": {rate}/s"
    res["graphs_per_s_vs_plain_capacity"] = (res["served"]["graphs_per_s"]
                                             / res["plain_capacity"]["graphs_per_s"])
    res["capacity_vs_plain_capacity"] = (res["served_capacity"]["graphs_per_s"]
                                         / res["plain_capacity"]["graphs_per_s"])
    res["workload"] = (f"open-loop Poisson, pool of {args.serve_pool} requests (seed "
                       f"{args.serve_seed}), edges lognormal median 50000 sigma 1.0 in [1000, 5e7], "
                       "R-MAT / Chung-Lu by coin flip")
    return res


def serve_expect(args, K, path):
    """Core numbers of every pool graph from the oracle's bucket peel (the checker for served
    results), and the CPU's own time for the pool."""
    import numpy as np
    import oracle as O
    t = 0.0
    edges = 0
    with open(path, "wb") as f:
        for j in range(args.serve_pool):
            g = K.Graph.request(args.serve_seed, j)
            tc = time.perf_counter()
            c, _ = O.bz(g.neighbors_offset, g.neighbors)
            t += time.perf_counter() - tc
            edges += g.m
            f.write(np.ascontiguousarray(c, np.uint32).tobytes())
    return {"value": args.serve_pool / t, "unit": "graphs/s", "cores": 1, "kind": "port",
            "sample": f"O(m) bucket peel of the {args.serve_pool} pool graphs ({edges} edges, {t:.2f} s)"}


def run_leg(args, leg, limit, tune_cache=None):
    """Runs this script as a child for one leg; returns (its JSON object, None) or (None, error)."""
    argv = [sys.executable, os.path.abspath(__file__), *sys.argv[1:], "--leg", leg]
    if tune_cache:
        argv += ["--tune-cache", tune_cache]
    try:
        p = subprocess.run(argv, stdout=subprocess.PIPE, text=True, timeout=limit)
    except subprocess.TimeoutExpired:
        log(f"[{leg}] child process exceeded {limit} s")
        return None, f"timed out after {limit} s"
    lines = [ln for ln in p.stdout.splitlines() if ln.startswith("{")]
    if p.returncode != 0 or not lines:
        log(f"[{leg}] child process exited {p.returncode}")
        return None, f"child process exited {p.returncode}"
    return json.loads(lines[-1]), None


# ---------------------------------------------------------------------------------------------
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--tune-cache", default=None,
                    help="autotuner cache file: settings tuned once are reused without timing")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=("single", "batch", "serve"), default=None)
    ap.add_argument("--batch-graphs", type=int, default=4096)
    ap.add_argument("--batch-seed", type=int, default=4)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--serve-rate", type=float, default=0.0,
                    help="open-loop arrival rate; 0 = 0.7 x the served path's measured capacity")
    ap.add_argument("--serve-closed", type=int, default=4000,
                    help="requests of the closed-loop run that measures the served capacity")
    ap.add_argument("--serve-sweep", type=lambda v: [float(x) for x in v.split(",") if x], default=[],
                    help="extra open-loop rates, comma-separated")
    ap.add_argument("--plain-rate", type=float, default=150.0)
    ap.add_argument("--plain-closed", type=int, default=600)
    ap.add_argument("--serve-pool", type=int, default=256)
    ap.add_argument("--serve-seed", type=int, default=5)
    ap.add_argument("--leg", choices=("tune", "batch"), default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this engine peels on a single MI355X (multi-GPU "
              "sharding is a design note, not a built path). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import kcoregpu_amd as K

    # Child processes: the tuning and the batched leg. Their failures (a device fault ends a
    # process) cannot take the headline line with them.
    if args.leg == "tune":
        eng = K.Engine(0)
        tuned = bench_single(args, K, eng, tune_cache=args.tune_cache, tune_only=True)
        eng.close()
        print(json.dumps({"tuned": tuned}), flush=True)
        return
    if args.leg == "batch":
        eng = K.Engine(0)
        res = bench_batch(args, K, eng)
        eng.close()
        print(json.dumps({"batched": res}), flush=True)
        return

    # Every child runs before this process touches the GPU itself.
    single = batch = served = None
    if args.only in (None, "serve") and args.serve_seconds > 0:
        try:
            with tempfile.TemporaryDirectory() as td:
                expect = None
                cpu_serve = None
                if not args.no_cpu_baseline:
                    expect = os.path.join(td, "expect.bin")
                    cpu_serve = serve_expect(args, K, expect)
                served = bench_serve(args, K, expect)
                served["cpu_baseline"] = cpu_serve
        except Exception as e:  # noqa: BLE001 - reported in the JSON line
            log(f"[serve] failed: {e!r}")
            served = {"error": repr(e)}
    if args.only in (None, "batch") and args.batch_graphs > 0:
        res, err = run_leg(args, "batch", 3000)
        batch = res["batched"] if res else {"error": err}
    with tempfile.TemporaryDirectory() as td:
        if args.only in (None, "single"):
            cache, tuned_by, tune_note = None, None, "defaults (--no-tune)"
            if not args.no_tune:
                cache = args.tune_cache or os.path.join(td, "tune.txt")
                res, err = run_leg(args, "tune", 1500, tune_cache=cache)
                if res:
                    tuned_by, tune_note = res["tuned"], "tuned in a child process"
                else:
                    cache, tune_note = None, f"defaults: tuning failed in its child process ({err})"
            eng = K.Engine(0)
            single = bench_single(args, K, eng, tune_cache=cache, tuned_by=tuned_by)
            single["detail"]["settings_from"] = tune_note
            eng.close()

    out = {"metric": "peeled edges/s", "unit": "edges/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "u32", "data": "synthetic"}
    if single:
        out.update(value=single["value"], ms_per_step=single["ms_per_step"], config=single["config"],
                   roofline=single["roofline"], cpu_baseline=single["cpu_baseline"],
                   detail=single["detail"])
    else:
        out.update(value=None, ms_per_step=None, config={"workload": f"--only {args.only}"},
                   roofline=None, cpu_baseline=None)
    out["batched"] = batch
    out["served"] = served
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
