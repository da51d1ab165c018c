"""Benchmark: exact K4 + K5 counts on a seeded R-MAT graph (scale 20, edge
factor 16, (a,b,c,d) = (.57,.19,.19,.05), seed 20) on one MI355X.

A step is one batched count of both patterns: the main persistent launch
and the launches after it (the workgroup tiers of the bitmap clique store).
`value` is undirected edges / median device time of a batched step (inputs
resident in HBM). The plain single-pattern engine (default config) is timed
beside it; `batched_vs_plain` = (t_K4 + t_K5) / t_batched at identical
counts. `roofline` is for the launch that takes longest (hipEvents inside
the timed steps). The bitmap clique store works on bit rows in LDS and is
charged against the vector ALU (rocprofv3 counters in
profiles/clique_rmat20_k4k5.json give its issue rate, lane use and LDS use):
`achieved` = W_alg / 4 algorithmic integer operations of that launch
(instrumented pass) / its duration, against the VALU's 78.6 T lane-ops/s;
`traffic` is the same launch's HBM bytes (rocprofv3 FETCH_SIZE +
WRITE_SIZE, calibrated on a copy kernel) and `hbm_frac` their rate against
8 TB/s. The array store's roofline is W_alg bytes against 8 TB/s.
`cpu_baseline` times the CPU oracle's whole-graph K4 + K5
count on the host's cores (at most 16) and checks its counts against the
GPU's; `naive_c1` times the reference's own CPU path (naive_pm, restated) on
its triangle workload, one thread. `hipgraph` times the same step captured
once into a hipGraph and replayed `--steps` times (gat_count_graph), its
counts checked against the eager steps'.

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
# MI355X vector ALU: 157.3 TFLOP/s FP32 vector peak (AMD) / 2 flops per FMA
# = 78.6 T lane-operations/s, i.e. 256 CUs x 4 SIMDs x 32 lanes x 2.4 GHz
# (two wave64 instructions per CU per cycle; rocprofv3 counts the clique
# tier-0 launch at 1.05 per CU-cycle, above what 16-lane SIMDs could issue).
VALU_PEAK_TOPS = 256 * 4 * 32 * 2.4e9 / 1e12
CLIQUE_KERNELS = ("gatk::clique_kernel<0>", "gatk::clique_block_kernel<16, true, 0, 8, 0>",
                  "gatk::clique_block_kernel<32, true, 1, 16, 0>")


def load_profile(name, cfg):
    """rocprofv3 summary in profiles/ (graphautotune_amd.profsum), when it was
    taken with this kernel config."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        prof = json.load(f)
    return prof if prof.get("config") == cfg else None


def clique_roofline(kernel, ms, walg, prof):
    """Bitmap clique store: its operands are bit rows in LDS, its work is
    integer bit operations on the vector ALU (rocprofv3 counters in the
    profile give the VALU issue rate, lane use and LDS use beside it).
    Algorithmic integer operations = W_alg / 4: every 4-byte operand word of
    an AND of rows is one AND (or one three-input bit op) or one
    popcount-accumulate."""
    ops = walg / 4.0
    achieved = ops / (ms * 1e-3) / 1e12
    out = {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TOPS, "unit": "Tops/s",
           "frac": achieved / VALU_PEAK_TOPS, "traffic": None, "kernel": kernel, "kernel_ms": ms,
           "alg_bytes_per_launch": walg, "alg_ops_per_launch": ops,
           "peak_source": "MI355X 157.3 TFLOP/s FP32 vector peak / 2 flops per FMA (256 CUs x 4 SIMDs x 32 "
                          "lanes x 2.4 GHz)"}
    if prof:
        hbm = prof.get("hbm_bytes_per_launch", {}).get(kernel)
        if hbm:
            out["traffic"] = hbm
            out["hbm_achieved_gbps"] = hbm / (ms * 1e-3) / 1e9
            out["hbm_frac"] = out["hbm_achieved_gbps"] / HBM_PEAK_GBPS
        low = prof.get("hbm_bytes_per_launch_unscaled_fetch", {}).get(kernel)
        if low:
            out["hbm_frac_unscaled_fetch"] = low / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS
        d = prof.get("derived", {}).get(kernel)
        if d:
            out["counters"] = {k: d[k] for k in ("valu_issue", "valu_lane_util", "lds_util", "lds_conflict_share",
                                                 "wait_share", "waves_per_cu") if k in d}
    return out


def hbm_roofline(kernel, ms, walg, prof):
    achieved = walg / (ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBPS, "traffic": None, "kernel": kernel, "kernel_ms": ms,
           "alg_bytes_per_launch": walg}
    if prof:
        hbm = prof.get("hbm_bytes_per_launch", {}).get(kernel)
        if hbm:
            out["traffic"] = hbm
            out["hbm_frac"] = hbm / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS
    return out


# launch groups of the array store's count: the main launch, and the split
# subtask levels with the HBM overflow launch and its subtasks
IEP_GROUPS = {"gatk::engine_kernel<false>": ("gatk::engine_kernel<false>",),
              "split and HBM launches": ("gatk::engine_sub_kernel<false>", "gatk::engine_kernel<true>",
                                         "gatk::engine_sub_kernel<true>")}


def iep_traffic(prof, group, key="hbm_bytes_per_launch"):
    """HBM bytes per count of a launch group (rocprofv3 FETCH_SIZE +
    WRITE_SIZE per dispatch, calibrated, times the group's dispatches per
    count; key "hbm_bytes_per_launch_unscaled_fetch" for the lower bound),
    the group's longest kernel and that kernel's counters."""
    calls = {k["kernel"]: k["calls"] for k in prof["kernel_stats"]}
    total = {k["kernel"]: k["total_ms"] for k in prof["kernel_stats"]}
    counts = calls.get("gatk::engine_kernel<false>")
    hb = prof.get(key, {})
    names = IEP_GROUPS[group]
    if not counts or not all(n in hb for n in names if calls.get(n)):
        return None, None, None
    traffic = sum(hb[n] * calls[n] for n in names if calls.get(n)) / counts
    longest = max(names, key=lambda n: total.get(n, 0.0))
    return traffic, longest, prof.get("derived", {}).get(longest)


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


def host_cpu():
    """CPU model and the host threads this process may use (at most 16: the
    share of the host that goes with one GPU)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return model, avail, max(1, min(16, avail))


def cpu_baseline(g, gpu_counts):
    """The oracle's whole-graph K4 + K5 count (a CPU port of the bitmap clique
    store, oracle/clique_mt.c) on the host's cores, timed, its counts checked
    against the GPU's. The degree relabelling is done before the timed span,
    as the GPU plan does it at plan creation."""
    from oracle import api as orc
    off, adj = g.csr()
    ro, ra = orc.relabel(off, adj)
    model, avail, threads = host_cpu()
    t0 = time.perf_counter()
    k3, k4, k5 = orc.clique_counts_mt(ro, ra, 5, threads)
    dt = time.perf_counter() - t0
    log("cpu: whole-graph K4 %d K5 %d on %d threads in %.2f s (%s)" % (k4, k5, threads, dt, model))
    if [k4, k5] != list(gpu_counts):
        raise AssertionError("CPU oracle counts %s differ from the GPU's %s" % ([k4, k5], gpu_counts))
    return {
        "value": g.undirected_edges() / dt,
        "unit": "edges/s",
        "cores": threads,
        "kind": "port",
        "sample": "whole graph: the oracle's K4 + K5 count over all %d top vertices (one pass for both), "
                  "%.2f s on %d threads; counts equal the GPU's" % (g.v_cnt(), dt, threads),
        "seconds": dt,
        "counts": {"K4": k4, "K5": k5},
        "cpu_model": model,
        "host_threads_available": avail,
    }


def naive_c1():
    """The reference's CPU path (target/naive_pm.cu:10-44, restated with a
    uint64 count) on its own workload: triangles of G(n = 10,000,
    p = 16/9,999), seed 1, one thread. It counts labeled embeddings, 6 per
    triangle."""
    import graphautotune_amd as gat
    from oracle import api as orc
    g = gat.Graph.erdos_renyi(10000, 16 / 9999, 1)
    off, adj = g.csr()
    t0 = time.perf_counter()
    labeled = orc.naive_count(off, adj, "011101110")
    dt = time.perf_counter() - t0
    t0 = time.perf_counter()
    gpu, st = gat.Plan(g, ["011101110"], "labeled").count()
    log("naive_pm restated: %d labeled triangles in %.1f s; GPU labeled count %d" % (labeled, dt, gpu[0]))
    if gpu[0] != labeled:
        raise AssertionError("naive_pm restatement %d and GPU labeled count %d differ" % (labeled, gpu[0]))
    model, _, _ = host_cpu()
    return {"workload": "Erdos-Renyi G(10000, 16/9999), seed 1, %d edges; labeled triangle embeddings"
                        % g.undirected_edges(),
            "labeled_count": labeled, "triangles": labeled // 6, "seconds": dt, "cores": 1,
            "value": g.undirected_edges() / dt, "unit": "edges/s", "cpu_model": model,
            "gpu_device_ms": st["seconds"] * 1e3, "kind": "reference algorithm, restated"}


def timeline_summary(tl, tpm, names):
    """Per launch: units, span, busy fraction (sum of unit durations over
    units x span), percentiles of unit end times, and the row-build share of
    the clique store's cycles."""
    import numpy as np
    out = {}
    for lid in sorted(set(int(x) for x in tl[:, 0])):
        r = tl[tl[:, 0] == lid]
        t0, t1 = r[:, 2] / tpm, r[:, 3] / tpm
        span = t1.max() - t0.min()
        busy = float((t1 - t0).sum() / (len(r) * span)) if span > 0 else 1.0
        ends = np.sort(t1 - t0.min())
        hist, edges = np.histogram(t1 - t0.min(), bins=10, range=(0, span if span > 0 else 1))
        cyc = r[:, 4].sum() + r[:, 5].sum()
        out[names.get(lid, str(lid))] = {
            "units": int(len(r)), "span_ms": float(span), "busy_frac": busy,
            "end_ms_p0_p50_p90_p99_p100": [float(np.percentile(ends, q)) for q in (0, 50, 90, 99, 100)],
            "end_hist_10bins": hist.tolist(),
            "items": int(r[:, 6].sum()), "pairs": int(r[:, 7].sum()),
            "build_cycle_share": float(r[:, 4].sum() / cyc) if cyc else None,
        }
    return out


IEP_PATTERNS = {"house": "0100110110010110110010100", "c5": "0100110100010100010110010"}


def graph_desc(name, roots=None):
    """The workload description of make_graph's graph (and of a root shard
    of it), as the IEP line and its profiles name it."""
    if name.startswith("rmat"):
        desc = "R-MAT scale %s, edge factor 16, (a,b,c,d)=(.57,.19,.19,.05), seed 20" % name[4:]
    elif name.startswith("cl"):
        desc = "Chung-Lu, %s vertices, mean degree 24, exponent 2.1, degree cap 50,000, seed 21" % name[2:]
    else:
        raise SystemExit("unknown graph %r (rmatS or clN)" % name)
    if roots:
        desc += "; roots [%d, %d) of the degree-ordered graph (raw sums)" % tuple(roots)
    return desc


def make_graph(gat, name):
    """rmatS: R-MAT scale S, edge factor 16, (.57,.19,.19,.05), seed 20;
    clN: Chung-Lu, N vertices, mean degree 24, exponent 2.1, degree cap
    50,000, seed 21."""
    desc = graph_desc(name)
    if name.startswith("rmat"):
        return gat.Graph.rmat(int(name[4:]), 16, 0.57, 0.19, 0.19, 20), desc
    return gat.Graph.chung_lu(int(name[2:]), 24.0, 2.1, 50000, 21), desc


class _Shard:
    """A plan restricted to a root range, with Plan.count's signature."""

    def __init__(self, plan, roots):
        self.plan, self.roots = plan, roots

    def count(self, cfg=None, account=False, timeline=False):
        return self.plan.count_shard(self.roots[0], self.roots[1], cfg, account=account, timeline=timeline)


def iep_cpu_reference(gat, g, args):
    """The reference's own algorithm on the CPU beside the IEP line: the
    oracle's restatement of the reference engine (engine_kernel.cuh with its
    IEP tail; reference mode, one thread) over the same roots of the
    degree-ordered graph, timed, and its value checked against the GPU's
    reference-mode count of the same roots of the same graph."""
    from oracle import api as orc
    off, adj = g.csr()
    ro, ra = orc.relabel(off, adj)
    lo, hi = (int(x) for x in args.roots.split(":")) if args.roots else (0, g.v_cnt())
    pat = IEP_PATTERNS[args.iep]
    gr = gat.Graph.from_csr(ro, ra)
    plan = gat.Plan(gr, [pat], "reference")
    plan.count_shard(lo, hi, {})
    got, st = plan.count_shard(lo, hi, {})
    from graphautotune_amd._lib import ScheduleData
    sd = orc.schedule(pat, ScheduleData)
    model, _, _ = host_cpu()
    t0 = time.perf_counter()
    want, ok = orc.reference_count_roots(ro, ra, sd, lo, hi)
    secs = time.perf_counter() - t0
    if not ok or got[0] != want:
        raise AssertionError("reference-mode roots [%d, %d): GPU %d, oracle %d (defined %s)" % (lo, hi, got[0], want, ok))
    edges = int(ro[hi] - ro[lo])
    log("cpu reference path: roots [%d, %d), %d edges, value %d in %.1f s on one thread (%s); GPU %.2f ms"
        % (lo, hi, edges, want, secs, model, st["seconds"] * 1e3))
    return {"value": edges / secs, "unit": "edges/s", "cores": 1, "kind": "port",
            "sample": "the oracle's restatement of the reference engine (reference mode, IEP tail) over roots "
                      "[%d, %d) of the degree-ordered graph (%d directed edges); value equal to the GPU's "
                      "reference-mode count of the same roots" % (lo, hi, edges),
            "seconds": secs, "reference_value": want, "gpu_reference_ms": st["seconds"] * 1e3, "cpu_model": model}


def iep_line(gat, args):
    """Auxiliary line: one 5-vertex IEP pattern on the array store (default
    config, hub bitmaps on), unique mode. Roofline of the longer of the main
    launch (LDS) and the launches after it (split subtasks, HBM overflow),
    as for the headline."""
    g, desc = make_graph(gat, args.graph)
    E = g.undirected_edges()
    plan = gat.Plan(g, [IEP_PATTERNS[args.iep]], "unique")
    cfg = json.loads(args.config) if args.config else {}
    roots = tuple(int(x) for x in args.roots.split(":")) if args.roots else None
    if roots:  # a root shard: raw sums of the roots [lo, hi) of the degree-ordered graph
        plan = _Shard(plan, roots)
        off = g.csr()[0]
        E = int(off[roots[1]] - off[roots[0]])  # directed edges leaving the shard's roots
        desc = graph_desc(args.graph, roots)
    for _ in range(args.warmup):
        plan.count(cfg)
    _, acc = plan.count(cfg, account=True)
    counts, secs, (mains, seconds), walls, st = timed(plan, cfg, args.steps)
    t = statistics.median(secs)
    if args.timeline:
        _, tst = plan.count(cfg, timeline=True)
        names = {0: "engine_kernel", 1: "subtasks level 1", 2: "subtasks level 2", 3: "subtasks level 3",
                 4: "HBM launch", 5: "HBM subtasks"}
        with open(args.timeline, "w") as f:
            json.dump({"workload": desc, "pattern": args.iep, "kernel_config": cfg, "device_ms": tst["seconds"] * 1e3,
                       "launches": timeline_summary(tst["timeline"], tst["ticks_per_ms"], names)}, f, indent=1)
        log("timeline summary written to %s" % args.timeline)
    launches = [("gatk::engine_kernel<false>", statistics.median(mains), acc["alg_bytes_main"]),
                ("split and HBM launches", statistics.median(seconds), acc["alg_bytes_second"])]
    name, ms, walg = max(launches, key=lambda x: x[1])
    achieved = walg / (ms * 1e-3) / 1e9
    traffic, traffic_low, counters = None, None, None
    prof = load_profile("engine_iep_%s_%s.json" % (args.graph, args.iep), cfg)
    if prof and prof.get("workload") == desc:
        traffic, longest, d = iep_traffic(prof, name)
        traffic_low = iep_traffic(prof, name, "hbm_bytes_per_launch_unscaled_fetch")[0]
        if d:
            counters = {k: d[k] for k in ("valu_issue", "valu_lane_util", "lds_util", "lds_conflict_share",
                                          "wait_share", "waves_per_cu") if k in d}
            counters["kernel"] = longest
    cpu = iep_cpu_reference(gat, g, args) if args.cpu_reference else None
    print(json.dumps({
        "metric": "exact-count throughput (edges/s), 5-vertex IEP pattern, array store" +
                  ("; edges = directed edges of the shard's roots" if roots else ""),
        "value": E / t, "unit": "edges/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sum(walls) / len(walls) * 1e3, "higher_is_better": True, "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "%s; exact unique %s count" % (desc, args.iep), "vertices": g.v_cnt(),
                   "undirected_edges": E, "pattern": IEP_PATTERNS[args.iep], "kernel_config": cfg},
        "counts": counts, "device_ms": t * 1e3, "overflow_tasks": st["overflow_tasks"], "cpu_baseline": cpu,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "hbm_achieved": traffic / (ms * 1e-3) / 1e9 if traffic else None,
                     "hbm_frac": traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS if traffic else None,
                     "hbm_frac_unscaled_fetch": traffic_low / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS if traffic_low
                     else None,
                     "counters": counters,
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
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--timeline", default=None, metavar="PATH",
                    help="also run one batched count with per-wave/per-workgroup timeline records and write "
                         "their summary (load-imbalance histograms) to PATH as JSON")
    ap.add_argument("--config", type=str, default=None, help="JSON kernel config for the batched run")
    ap.add_argument("--no-plain", action="store_true",
                    help="time only the batched plan (profiling runs: one kernel config per kernel name)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the device copy kernel over MIB MiB (3 timed launches) and exit; "
                         "calibrates the HBM byte counters")
    ap.add_argument("--iep", choices=sorted(IEP_PATTERNS), default=None,
                    help="auxiliary line: one 5-vertex IEP pattern on the array store instead of K4 + K5")
    ap.add_argument("--graph", default="rmat14", help="with --iep: rmatS or clN (see make_graph)")
    ap.add_argument("--cpu-reference", action="store_true",
                    help="with --iep: also time the oracle's restatement of the reference engine on the same "
                         "roots (one thread; minutes per hub-adjacent root)")
    ap.add_argument("--roots", default=None, metavar="LO:HI",
                    help="with --iep: count only the roots [LO, HI) of the degree-ordered graph (a shard)")
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
    # the same step captured once into a hipGraph and replayed (gat_count_graph)
    tg0 = time.perf_counter()
    g_counts, g_st = batched.count(cfg, graph_replays=args.steps)
    g_wall = time.perf_counter() - tg0
    if g_counts != counts:
        raise AssertionError("hipGraph replay %s and eager %s counts differ" % (g_counts, counts))
    main_ms = statistics.median(mains[0])
    second_ms = statistics.median(mains[1])
    # the roofline is for the launch that dominates the step
    bitmap = cfg.get("vertex_store", 0) != 1
    walg_rest = walg - walg_main - walg_second
    rest_ms = t_b * 1e3 - main_ms - second_ms
    if bitmap:
        launches = [(CLIQUE_KERNELS[0], main_ms, walg_main), (CLIQUE_KERNELS[1], second_ms, walg_second),
                    (CLIQUE_KERNELS[2] + " and tier 2", rest_ms, walg_rest)]
    else:
        launches = [("gatk::engine_kernel<false>", main_ms, walg_main),
                    ("split and HBM launches", second_ms, walg_second)]
    kernel_name, kernel_ms, kernel_walg = max(launches, key=lambda x: x[1])
    t_plain, plain_ms, t_plain_same, plain_same_ms = None, None, None, None
    if plains:
        c4, s4, _, _, _ = timed(plain4, {}, args.steps)
        c5, s5, _, _, _ = timed(plain5, {}, args.steps)
        if counts != [c4[0], c5[0]]:
            raise AssertionError("batched %s and plain %s counts differ" % (counts, c4 + c5))
        plain_ms = {"K4": statistics.median(s4) * 1e3, "K5": statistics.median(s5) * 1e3}
        t_plain = statistics.median(s4) + statistics.median(s5)
        # the same single-pattern plans under the batched run's (tuned)
        # config: separates the gain of batching from that of tuning
        for p in plains:
            p.count(cfg)
        c4, s4, _, _, _ = timed(plain4, cfg, args.steps)
        c5, s5, _, _, _ = timed(plain5, cfg, args.steps)
        if counts != [c4[0], c5[0]]:
            raise AssertionError("batched %s and plain (tuned config) %s counts differ" % (counts, c4 + c5))
        plain_same_ms = {"K4": statistics.median(s4) * 1e3, "K5": statistics.median(s5) * 1e3}
        t_plain_same = statistics.median(s4) + statistics.median(s5)
    log("counts K4 %d K5 %d | batched %.3f ms (main %.3f ms, second %.3f ms), plain %s ms | overflow tasks %d"
        % (counts[0], counts[1], t_b * 1e3, main_ms, second_ms, "%.3f" % (t_plain * 1e3) if t_plain else "-",
           st["overflow_tasks"]))

    prof = load_profile("clique_rmat%d_k4k5.json" % args.scale, cfg)
    roofline = clique_roofline(kernel_name, kernel_ms, kernel_walg, prof) if bitmap else \
        hbm_roofline(kernel_name, kernel_ms, kernel_walg, prof)
    roofline.update({"launches_ms": {name: ms for name, ms, _ in launches}, "step_alg_bytes": walg,
                     "step_alg_bytes_per_s": walg / t_b})

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
        "device_ms_plain_same_config": plain_same_ms,
        "batched_vs_plain_same_config": t_plain_same / t_b if t_plain_same else None,
        "roofline": roofline,
        "hipgraph": {"replays": args.steps, "device_ms_last_replay": g_st["seconds"] * 1e3,
                     "wall_ms_per_replay_with_capture": g_wall / args.steps * 1e3},
    }
    if args.timeline:
        _, tst = batched.count(cfg, timeline=True)
        names = {0: "clique_kernel", 1: "clique_block_kernel tier 0", 2: "clique_block_kernel tier 1",
                 3: "clique_block_kernel tier 2"} if bitmap else \
                {0: "engine_kernel", 1: "subtasks level 1", 2: "subtasks level 2", 3: "subtasks level 3",
                 4: "HBM launch", 5: "HBM subtasks"}
        summ = {"workload": out["config"]["workload"], "kernel_config": cfg, "device_ms": tst["seconds"] * 1e3,
                "launches_ms": {"main": tst["main_kernel_ms"], "second": tst["second_kernel_ms"]},
                "launches": timeline_summary(tst["timeline"], tst["ticks_per_ms"], names)}
        with open(args.timeline, "w") as f:
            json.dump(summ, f, indent=1)
        log("timeline summary written to %s" % args.timeline)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(g, counts)
        out["naive_c1"] = naive_c1()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
