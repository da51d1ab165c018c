"""Measurement tooling on the CPU: rocprofv3 summaries (graphautotune_amd.profsum)
from synthetic output files of known values, and the bench's roofline
helpers on them."""
import csv
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _write(path, header, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def _stats(d, kernels):
    """kernels: [(name, calls, average ns)]"""
    tot = sum(c * a for _, c, a in kernels)
    _write(os.path.join(d, "run", "kt_kernel_stats.csv"), ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage"],
           [[n, c, c * a, a, 100.0 * c * a / tot] for n, c, a in kernels])


def _counters(d, rows):
    """rows: [(kernel, counter, value)], one dispatch each"""
    _write(os.path.join(d, "run", "x_counter_collection.csv"), ["Kernel_Name", "Counter_Name", "Counter_Value"],
           [[k, c, v] for k, c, v in rows])


@pytest.fixture
def profile(tmp_path):
    """A profiled array-store count: 2 counts, each one main launch and three
    split launches; the copy probe reads 1 GiB reported as half (128 B
    requests counted as 64 B) and writes 1 GiB."""
    main, sub = "void gatk::engine_kernel<false>(gatk::KArgs const*, int)", \
        "void gatk::engine_sub_kernel<false>(gatk::KArgs const*, int)"
    _stats(str(tmp_path / "kt"), [(main, 2, 1e6), (sub, 6, 3e6), ("gatk::fill_u64(unsigned long long*, long, "
                                                                  "unsigned long long)", 4, 1e3)])
    cyc = 1e6  # per XCD
    _counters(str(tmp_path / "pmc"), [
        (sub, "GRBM_GUI_ACTIVE", 8 * cyc), (sub, "SQ_ACTIVE_INST_VALU", 0.5 * 2 * 256 * cyc),
        (sub, "SQ_THREAD_CYCLES_VALU", 0.25 * 64 * 0.5 * 2 * 256 * cyc), (sub, "SQ_LDS_IDX_ACTIVE", 0.2 * 256 * cyc),
        (sub, "SQ_LDS_BANK_CONFLICT", 0.1 * 0.2 * 256 * cyc), (sub, "SQ_WAVE_CYCLES", 16 * 256 * cyc / 4),
        (sub, "SQ_WAIT_ANY", 0.6 * 16 * 256 * cyc / 4)])
    _counters(str(tmp_path / "fetch"), [(main, "FETCH_SIZE", 1000.0), (sub, "FETCH_SIZE", 4000.0)])
    _counters(str(tmp_path / "write"), [(main, "WRITE_SIZE", 100.0), (sub, "WRITE_SIZE", 200.0)])
    _counters(str(tmp_path / "cf"), [("gatk::copy_kernel(int4 const*, int4*, long)", "FETCH_SIZE", 512.0 * 1024)])
    _counters(str(tmp_path / "cw"), [("gatk::copy_kernel(int4 const*, int4*, long)", "WRITE_SIZE", 1024.0 * 1024)])
    out = tmp_path / "prof.json"
    from graphautotune_amd import profsum
    profsum.main(["--stats", str(tmp_path / "kt"), "--pmc", str(tmp_path / "pmc"), "--fetch", str(tmp_path / "fetch"),
                  "--write", str(tmp_path / "write"), "--copy-fetch", str(tmp_path / "cf"), "--copy-write",
                  str(tmp_path / "cw"), "--config", "{}", "--workload", "w", "--out", str(out)])
    return json.loads(out.read_text())


def test_profsum_derived_and_calibration(profile):
    sub = "gatk::engine_sub_kernel<false>"
    d = profile["derived"][sub]
    assert d["cycles"] == pytest.approx(1e6)
    assert d["valu_issue"] == pytest.approx(0.5) and d["valu_lane_util"] == pytest.approx(0.25)
    assert d["lds_util"] == pytest.approx(0.2) and d["lds_conflict_share"] == pytest.approx(0.1)
    assert d["waves_per_cu"] == pytest.approx(16) and d["wait_share"] == pytest.approx(0.6)
    assert profile["fetch_scale"] == pytest.approx(2.0) and profile["write_scale"] == pytest.approx(1.0)
    assert profile["hbm_bytes_per_launch"][sub] == pytest.approx(1024.0 * (2 * 4000 + 200))
    assert profile["hbm_bytes_per_launch_unscaled_fetch"][sub] == pytest.approx(1024.0 * (4000 + 200))
    st = {k["kernel"]: k for k in profile["kernel_stats"]}  # names without argument lists
    assert st[sub]["calls"] == 6 and st[sub]["average_ms"] == pytest.approx(3.0)
    assert st["gatk::engine_kernel<false>"]["total_ms"] == pytest.approx(2.0) and "gatk::fill_u64" in st


def test_bench_iep_traffic(profile):
    """Per count: the split group's bytes per launch times its launches per
    count (6 split launches over 2 counts), the lower bound without the fetch
    scale, and the counters of the group's longest kernel."""
    sys.path.insert(0, ROOT)
    import bench
    t, longest, d = bench.iep_traffic(profile, "split and HBM launches")
    assert longest == "gatk::engine_sub_kernel<false>" and d["valu_issue"] == pytest.approx(0.5)
    assert t == pytest.approx(3 * 1024.0 * (2 * 4000 + 200))
    low = bench.iep_traffic(profile, "split and HBM launches", "hbm_bytes_per_launch_unscaled_fetch")[0]
    assert low == pytest.approx(3 * 1024.0 * (4000 + 200))
    t_main = bench.iep_traffic(profile, "gatk::engine_kernel<false>")[0]
    assert t_main == pytest.approx(1024.0 * (2 * 1000 + 100))


def test_bench_clique_roofline(profile):
    """VALU-charged roofline: W_alg / 4 operations over the launch time against
    the 78.6 T lane-ops/s peak; HBM fractions from the profile."""
    sys.path.insert(0, ROOT)
    import bench
    k = "gatk::engine_sub_kernel<false>"
    r = bench.clique_roofline(k, 10.0, 4.0e12, profile)
    assert r["bound"] == "valu" and r["achieved"] == pytest.approx(1.0e12 / 0.01 / 1e12)
    assert r["frac"] == pytest.approx(r["achieved"] / bench.VALU_PEAK_TOPS)
    assert bench.VALU_PEAK_TOPS == pytest.approx(78.6432)
    hb = 1024.0 * (2 * 4000 + 200)
    assert r["traffic"] == pytest.approx(hb) and r["hbm_frac"] == pytest.approx(hb / 0.01 / 8e12)
    assert r["hbm_frac_unscaled_fetch"] == pytest.approx(1024.0 * 4200 / 0.01 / 8e12)
    assert r["counters"]["waves_per_cu"] == pytest.approx(16)


def test_committed_profiles_match_the_bench():
    """The profiles under profiles/ are the ones the bench lines use: the
    clique profile was taken with the tuned config of the headline workload
    and holds the tier kernels' counters; the array-store profiles name the
    bench's root-shard workloads and hold every kernel of their launch groups."""
    sys.path.insert(0, ROOT)
    import bench
    cfg, key = bench.tuned_config(20)
    prof = bench.load_profile("clique_rmat20_k4k5.json", cfg)
    assert key == "rmat20-k4k5" and prof is not None
    for k in bench.CLIQUE_KERNELS:
        assert k in prof["derived"] and k in prof["hbm_bytes_per_launch"], k
    for pat, roots in (("house", (8192, 8224)), ("c5", (65536, 65600))):
        p = bench.load_profile("engine_iep_cl4194304_%s.json" % pat, {})
        assert p is not None and p["workload"] == bench.graph_desc("cl4194304", roots)
        for group in bench.IEP_GROUPS:
            t, longest, d = bench.iep_traffic(p, group)
            assert t is not None and t > 0 and d, (pat, group)
