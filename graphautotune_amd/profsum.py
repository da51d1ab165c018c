"""Summaries of rocprofv3 output for profiles/: per-kernel durations from a
--kernel-trace --stats run, and per-kernel counter means with the derived
utilisations from --pmc runs, as JSON.

    python -m graphautotune_amd.profsum --stats DIR [--pmc DIR ...] [--fetch DIR --write DIR
        --copy-fetch DIR --copy-write DIR --copy-mib N] --config JSON --out FILE

Derived quantities (MI355X: 256 CUs in 8 XCDs, 4 SIMDs per CU; the SQ and
GRBM counters come back summed over the XCDs, so a kernel's cycle count is
GRBM_GUI_ACTIVE / 8):
  valu_issue         SQ_ACTIVE_INST_VALU / (2 x CUs x cycles): wave64 VALU
                     instructions against two per CU per cycle (32 lanes per
                     SIMD per cycle, the rate of the 157.3 TFLOP/s FP32 vector
                     peak; the clique kernels issue above one per CU-cycle)
  valu_lane_util     SQ_THREAD_CYCLES_VALU / (64 x SQ_ACTIVE_INST_VALU): active
                     lanes per VALU instruction
  lds_util           SQ_LDS_IDX_ACTIVE / (CUs x cycles)
  lds_conflict_share SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE
  waves_per_cu       4 x SQ_WAVE_CYCLES / (CUs x cycles) (quad-cycle counter)
  wait_share         SQ_WAIT_ANY / SQ_WAVE_CYCLES: share of resident wave time
                     spent waiting (dependencies, memory, barriers)
  lds_load_bytes     64 x SQ_INSTS_LDS_LOAD_BANDWIDTH
HBM bytes per dispatch: FETCH_SIZE and WRITE_SIZE (KiB) times the factor
that makes a plain copy of known size read and write exactly its size
(FETCH_SIZE counts 64 B per read request, and the copy's requests are
128 B, so its factor is 2). A kernel of scattered small reads issues
smaller requests, for which FETCH_SIZE is already right: its scaled figure
is an upper bound, and the unscaled one (also written) a lower bound.
"""
import argparse
import collections
import csv
import glob
import json
import os

CUS, XCDS = 256, 8


def _one(d, suffix):
    hits = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not hits:
        raise SystemExit("no %s under %s" % (suffix, d))
    return hits[0]


def short(name):
    """Kernel name without its argument list."""
    name = name.replace("void ", "", 1)
    depth = 0
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return name[:i]
    return name


def kernel_stats(d):
    rows = list(csv.DictReader(open(_one(d, "kernel_stats.csv"))))
    return [{"kernel": short(r["Name"]), "calls": int(r["Calls"]), "total_ms": float(r["TotalDurationNs"]) / 1e6,
             "average_ms": float(r["AverageNs"]) / 1e6, "percent": float(r["Percentage"])} for r in rows]


def counter_means(d):
    """{kernel: {counter: mean over dispatches}}."""
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(_one(d, "counter_collection.csv"))):
        agg[short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in agg.items()}


def derived(c):
    out = {}
    cyc = c.get("GRBM_GUI_ACTIVE", 0) / XCDS
    if not cyc:
        return out
    out["cycles"] = cyc
    if "SQ_ACTIVE_INST_VALU" in c:
        out["valu_issue"] = c["SQ_ACTIVE_INST_VALU"] / (2 * CUS * cyc)
        if "SQ_THREAD_CYCLES_VALU" in c and c["SQ_ACTIVE_INST_VALU"]:
            out["valu_lane_util"] = c["SQ_THREAD_CYCLES_VALU"] / (64 * c["SQ_ACTIVE_INST_VALU"])
    if "SQ_LDS_IDX_ACTIVE" in c:
        out["lds_util"] = c["SQ_LDS_IDX_ACTIVE"] / (CUS * cyc)
        if "SQ_LDS_BANK_CONFLICT" in c and c["SQ_LDS_IDX_ACTIVE"]:
            out["lds_conflict_share"] = c["SQ_LDS_BANK_CONFLICT"] / c["SQ_LDS_IDX_ACTIVE"]
    if "SQ_WAVE_CYCLES" in c:
        out["waves_per_cu"] = 4 * c["SQ_WAVE_CYCLES"] / (CUS * cyc)
    if "SQ_WAIT_ANY" in c and c.get("SQ_WAVE_CYCLES"):
        out["wait_share"] = c["SQ_WAIT_ANY"] / c["SQ_WAVE_CYCLES"]
    if "SQ_INSTS_LDS_LOAD_BANDWIDTH" in c:
        out["lds_load_bytes"] = 64 * c["SQ_INSTS_LDS_LOAD_BANDWIDTH"]
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--stats", required=True, help="output directory of a --kernel-trace --stats run")
    ap.add_argument("--pmc", action="append", default=[], help="output directory of a --pmc run (repeatable)")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-fetch")
    ap.add_argument("--copy-write")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--config", default="{}", help="JSON kernel config of the profiled run")
    ap.add_argument("--workload", default="")
    ap.add_argument("--commands", default="", help="the rocprofv3 commands, for the record")
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    out = {"workload": args.workload, "config": json.loads(args.config), "commands": args.commands,
           "kernel_stats": kernel_stats(args.stats), "counters": {}, "derived": {}}
    merged = collections.defaultdict(dict)
    for d in args.pmc:
        for k, cs in counter_means(d).items():
            merged[k].update(cs)
    for k, cs in merged.items():
        out["counters"][k] = cs
        out["derived"][k] = derived(cs)
    if args.fetch and args.write and args.copy_fetch and args.copy_write:
        kib = args.copy_mib * 1024.0
        cf = counter_means(args.copy_fetch)
        cw = counter_means(args.copy_write)
        copy = [k for k in cf if "copy_kernel" in k][0]
        fscale = kib / cf[copy]["FETCH_SIZE"]
        wscale = kib / cw[copy]["WRITE_SIZE"]
        f, w = counter_means(args.fetch), counter_means(args.write)
        out["fetch_scale"], out["write_scale"] = fscale, wscale
        out["hbm_bytes_per_launch"] = {k: 1024.0 * (f[k]["FETCH_SIZE"] * fscale + w.get(k, {}).get("WRITE_SIZE", 0)
                                                   * wscale) for k in f if "FETCH_SIZE" in f[k]}
        # the same without the fetch scale: a lower bound for kernels whose
        # scattered reads issue requests smaller than the copy kernel's
        out["hbm_bytes_per_launch_unscaled_fetch"] = {
            k: 1024.0 * (f[k]["FETCH_SIZE"] + w.get(k, {}).get("WRITE_SIZE", 0) * wscale)
            for k in f if "FETCH_SIZE" in f[k]}
    with open(args.out, "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
