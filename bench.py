"""Benchmark: exact K4 + K5 counts on a seeded R-MAT graph (scale 20, edge
factor 16, (a,b,c,d) = (.57,.19,.19,.05), seed 20) on one MI355X.

A step is one batched count of both patterns: the main persistent launch
and the launches after it (the workgroup tiers of the bitmap clique store).
`value` is undirected edges / median device time of a batched step (inputs
resident in HBM). The plain single-pattern engine (default config) is timed
beside it; `batched_vs_plain` = (t_K4 + t_K5) / t_batched at identical
counts. `roofline` is for whichever of the main and the second launch takes
longer: W_alg bytes of that launch (instrumented pass) / its hipEvent-timed
duration, against 8 TB/s; `hbm_achieved` is the same launch's measured HBM
bytes (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel) over
that duration. `cpu_baseline` times the CPU oracle's restatement
of the same algorithm on a bounded sample of top vertices and scales it by
algorithmic bytes.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "exact-count throughput (edges/s), batched+tuned vs plain, R-MAT s20 4/5-clique"
K4 = "0111101111011110"
K5 = "0111110111110111110111110"
HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def tuned_config(scale):
    """Kernel config picked by the autotuner for this workload, if recorded."""
    path = os.path.join(ROOT, "graphautotune_amd", "tuning", "tuned_configs.json")
    if os.path.exists(path):
        with open(path) as f:
            d = json.load(f)
        key = "rmat%d-k4k5" % scale
        if key in d:
            return d[key]["config"], key
    return {}, None


def timed(plan, cfg, steps):
    secs, mains, seconds_k, walls, counts = [], [], [], [], None
    for _ in range(steps):
        t0 = time.perf_counter()
        c, st = plan.count(cfg)
        walls.append(time.perf_counter() - t0)
        secs.append(st["seconds"])
        mains.append(st["main_kernel_ms"])
        seconds_k.append(st["second_kernel_ms"])
        if counts is None:
            counts = c
        elif c != counts:
            raise AssertionError("counts changed between steps: %s vs %s" % (c, counts))
    return counts, secs, (mains, seconds_k), walls, st


def cpu_baseline(g, total_walg, budget_s):
    """The oracle's restatement of the bitmap clique store (single thread) on
    every stride-th top vertex, K4 then K5 as the plain plans run them,
    scaled to the whole workload by algorithmic bytes (same definition on
    both sides)."""
    from oracle import api as orc
    off, adj = g.csr()
    t0 = time.perf_counter()
    ro, ra = orc.relabel(off, adj)
    log("cpu: relabel %.1fs" % (time.perf_counter() - t0))
    n = len(ro) - 1
    stride = 1 << 16
    while True:
        t0 = time.perf_counter()
        w = 0
        for k in (4, 5):
            w += orc.clique_bitmap_walg_sample(ro, ra, k, stride, 0)[1]
        dt = time.perf_counter() - t0
        log("cpu: stride %d took %.2fs (W %.3g B)" % (stride, dt, w))
        if dt >= budget_s / 4 or stride == 1:
            break
        stride = max(1, int(stride / min(64.0, (budget_s / 2) / max(dt, 0.01))))
    est = dt * total_walg / max(w, 1)
    return {
        "value": g.undirected_edges() / est,
        "unit": "edges/s",
        "cores": 1,
        "kind": "port",
        "sample": "oracle bitmap-store clique count (K4, then K5) over top vertices v0 = 0, %d, 2*%d, ... "
                  "(%d of %d vertices, %.1f s, %.3g of %.3g algorithmic bytes); time scaled by bytes"
                  % (stride, stride, (n + stride - 1) // stride, n, dt, w, total_walg),
    }


IEP_PATTERNS = {"house": "0100110110010110110010100", "c5": "0100110100010100010110010"}


def make_graph(gat, name):
    """rmatS: R-MAT scale S, edge factor 16, (.57,.19,.19,.05), seed 20;
    clN: Chung-Lu, N vertices, mean degree 24, exponent 2.1, degree cap
    50,000, seed 21."""
    if name.startswith("rmat"):
        return gat.Graph.rmat(int(name[4:]), 16, 0.57, 0.19, 0.19, 20), \
            "R-MAT scale %s, edge factor 16, (a,b,c,d)=(.57,.19,.19,.05), seed 20" % name[4:]
    if name.startswith("cl"):
        return gat.Graph.chung_lu(int(name[2:]), 24.0, 2.1, 50000, 21), \
            "Chung-Lu, %s vertices, mean degree 24, exponent 2.1, degree cap 50,000, seed 21" % name[2:]
    raise SystemExit("unknown graph %r (rmatS or clN)" % name)


def iep_line(gat, args):
    """Auxiliary line: one 5-vertex IEP pattern on the array store (default
    config, hub bitmaps on), unique mode. Roofline of the longer of the main
    launch (LDS) and the launches after it (split subtasks, HBM overflow),
    as for the headline."""
    g, desc = make_graph(gat, args.graph)
    E = g.undirected_edges()
    plan = gat.Plan(g, [IEP_PATTERNS[args.iep]], "unique")
    cfg = json.loads(args.config) if args.config else {}
    for _ in range(args.warmup):
        plan.count(cfg)
    _, acc = plan.count(cfg, account=True)
    counts, secs, (mains, seconds), walls, st = timed(plan, cfg, args.steps)
    t = statistics.median(secs)
    launches = [("gatk::engine_kernel<false>", statistics.median(mains), acc["alg_bytes_main"]),
                ("split and HBM launches", statistics.median(seconds), acc["alg_bytes_second"])]
    name, ms, walg = max(launches, key=lambda x: x[1])
    achieved = walg / (ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_iep.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f).get("%s-%s" % (args.graph, args.iep), {})
        if tr.get("config") == cfg:
            traffic = tr.get("bytes_per_launch", {}).get(name)
    print(json.dumps({
        "metric": "exact-count throughput (edges/s), 5-vertex IEP pattern, array store",
        "value": E / t, "unit": "edges/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sum(walls) / len(walls) * 1e3, "higher_is_better": True, "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "%s; exact unique %s count" % (desc, args.iep), "vertices": g.v_cnt(),
                   "undirected_edges": E, "pattern": IEP_PATTERNS[args.iep], "kernel_config": cfg},
        "counts": counts, "device_ms": t * 1e3, "overflow_tasks": st["overflow_tasks"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "hbm_achieved": traffic / (ms * 1e-3) / 1e9 if traffic else None,
                     "alg_bytes_per_launch": walg, "kernel": name, "kernel_ms": ms,
                     "launches_ms": {k: v for k, v, _ in launches},
                     "step_alg_bytes": acc["alg_bytes"], "step_achieved": acc["alg_bytes"] / t / 1e9},
    }), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--config", type=str, default=None, help="JSON kernel config for the batched run")
    ap.add_argument("--no-plain", action="store_true",
                    help="time only the batched plan (profiling runs: one kernel config per kernel name)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the device copy kernel over MIB MiB (3 timed launches) and exit; "
                         "calibrates the HBM byte counters")
    ap.add_argument("--iep", choices=sorted(IEP_PATTERNS), default=None,
                    help="auxiliary line: one 5-vertex IEP pattern on the array store instead of K4 + K5")
    ap.add_argument("--graph", default="rmat14", help="with --iep: rmatS or clN (see make_graph)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import graphautotune_amd as gat
    if args.copy_probe:
        import ctypes as C
        from graphautotune_amd._lib import check, lib
        gbps = C.c_double()
        check(lib().gat_copy_probe(args.copy_probe << 20, 3, C.byref(gbps)))
        print(json.dumps({"copy_probe_mib": args.copy_probe, "launches": 4, "gbps": gbps.value}), flush=True)
        return
    if args.iep:
        iep_line(gat, args)
        return
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing only
        sync = lambda: None  # noqa: E731

    t0 = time.perf_counter()
    g = gat.Graph.rmat(args.scale, 16, 0.57, 0.19, 0.19, 20)
    E = g.undirected_edges()
    log("graph: R-MAT s%d: %d vertices, %d edges, max degree %d (%.1fs)"
        % (args.scale, g.v_cnt(), E, g.max_degree(), time.perf_counter() - t0))
    cfg, tuned_key = tuned_config(args.scale)
    if args.config:
        cfg = json.loads(args.config)
    batched = gat.Plan(g, [K4, K5], "unique")
    plain4 = gat.Plan(g, [K4], "unique")
    plain5 = gat.Plan(g, [K5], "unique")

    plains = [] if args.no_plain else [plain4, plain5]
    for _ in range(args.warmup):
        batched.count(cfg)
        for p in plains:
            p.count({})
    # instrumented pass: algorithmic bytes (not timed)
    acc_counts, acc = batched.count(cfg, account=True)
    walg = acc["alg_bytes"]
    walg_main = acc["alg_bytes_main"]
    walg_second = acc["alg_bytes_second"]
    plain_walg = None
    if plains:
        plain_walg = sum(p.count({}, account=True)[1]["alg_bytes"] for p in plains)
        log("W_alg batched %.4g B, plain %.4g B" % (walg, plain_walg))

    sync()
    tw0 = time.perf_counter()
    counts, secs, mains, walls, st = timed(batched, cfg, args.steps)
    sync()
    wall = time.perf_counter() - tw0
    if counts != acc_counts:
        raise AssertionError("batched %s and instrumented %s counts differ" % (counts, acc_counts))
    t_b = statistics.median(secs)
    main_ms = statistics.median(mains[0])
    second_ms = statistics.median(mains[1])
    # the roofline is for the launch that dominates the step
    bitmap = cfg.get("vertex_store", 0) != 1
    launches = [("gatk::clique_kernel" if bitmap else "gatk::engine_kernel<false>", main_ms, walg_main),
                ("gatk::clique_block_kernel<16, true, 0, 8>" if bitmap else "gatk::engine_kernel<true>", second_ms,
                 walg_second)]
    kernel_name, kernel_ms, kernel_walg = max(launches, key=lambda x: x[1])
    achieved = kernel_walg / (kernel_ms * 1e-3) / 1e9
    t_plain, plain_ms = None, None
    if plains:
        c4, s4, _, _, _ = timed(plain4, {}, args.steps)
        c5, s5, _, _, _ = timed(plain5, {}, args.steps)
        if counts != [c4[0], c5[0]]:
            raise AssertionError("batched %s and plain %s counts differ" % (counts, c4 + c5))
        plain_ms = {"K4": statistics.median(s4) * 1e3, "K5": statistics.median(s5) * 1e3}
        t_plain = statistics.median(s4) + statistics.median(s5)
    log("counts K4 %d K5 %d | batched %.3f ms (main %.3f ms, second %.3f ms), plain %s ms | overflow tasks %d"
        % (counts[0], counts[1], t_b * 1e3, main_ms, second_ms, "%.3f" % (t_plain * 1e3) if t_plain else "-",
           st["overflow_tasks"]))

    traffic = None  # HBM bytes per launch of that kernel, from rocprofv3 FETCH_SIZE/WRITE_SIZE runs
    tpath = os.path.join(ROOT, "profiles", "traffic_rmat%d_k4k5.json" % args.scale)
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        if tr.get("config") == cfg:
            traffic = tr.get("bytes_per_launch", {}).get(kernel_name)

    out = {
        "metric": METRIC,
        "value": E / t_b,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": "R-MAT scale %d, edge factor 16, (a,b,c,d)=(.57,.19,.19,.05), seed 20; exact unique "
                        "K4 + K5 counts in one batched launch" % args.scale,
            "vertices": g.v_cnt(),
            "undirected_edges": E,
            "patterns": {"K4": K4, "K5": K5},
            "kernel_config": cfg,
            "tuned_from": tuned_key,
        },
        "counts": {"K4": counts[0], "K5": counts[1]},
        "reference_mode_counts": {"K4": 4 * counts[0], "K5": 5 * counts[1]},
        "device_ms_batched": t_b * 1e3,
        "device_ms_plain": plain_ms,
        "plain_value": E / t_plain if t_plain else None,
        "batched_vs_plain": t_plain / t_b if t_plain else None,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": traffic,
            # measured HBM rate of the same launch: far below `achieved` when
            # the operands are read from LDS (the bitmap rows of the
            # workgroup tiers), so `frac` can exceed 1 there
            "hbm_achieved": traffic / (kernel_ms * 1e-3) / 1e9 if traffic else None,
            "hbm_frac": traffic / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS if traffic else None,
            "alg_bytes_per_launch": kernel_walg,
            "kernel": kernel_name,
            "kernel_ms": kernel_ms,
            "launches_ms": {name: ms for name, ms, _ in launches},
            "step_alg_bytes": walg,
            "step_achieved": walg / t_b / 1e9,
        },
    }
    if not args.no_cpu:
        if plain_walg is None:
            plain_walg = sum(gat.Plan(g, [p], "unique").count({}, account=True)[1]["alg_bytes"] for p in (K4, K5))
        out["cpu_baseline"] = cpu_baseline(g, plain_walg, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
