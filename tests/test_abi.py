"""The C ABI: libgat.so loads, exports every function include/gat.h declares,
and reports errors through return codes (CPU only: no compute calls)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "gat.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(gat_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported(gat):
    from graphautotune_amd import _lib
    lib = C.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    # and the Python binding covers them all
    assert set(names) == set(_lib.EXPORTED) | {"gat_plan_describe"} or set(names) <= set(_lib.EXPORTED)


def test_version_and_defaults(gat):
    from graphautotune_amd._lib import lib
    assert b"gfx950" in lib().gat_version()
    d = gat.default_config().as_dict()
    assert d["block_waves"] == 1 and d["waves_per_cu"] == 8 and d["queue_chunk"] == 16


def test_error_codes(gat):
    from graphautotune_amd._lib import (GAT_EINVAL, GAT_ELIMIT, GAT_EUNSUPPORTED, GatError, ScheduleData, lib)
    d = ScheduleData()
    assert lib().gat_schedule_build(b"012", C.byref(d), C.sizeof(d)) == GAT_EINVAL
    assert b"square" in lib().gat_last_error()
    assert lib().gat_schedule_build(b"011101110", C.byref(d), 8) == GAT_EINVAL  # buffer too small
    assert lib().gat_schedule_build(b"0" * 121, C.byref(d), C.sizeof(d)) == GAT_ELIMIT  # 11 vertices
    assert lib().gat_schedule_build(b"0110", C.byref(d), C.sizeof(d)) == GAT_EUNSUPPORTED  # an edge
    with pytest.raises(GatError):
        gat.Graph.load("/nonexistent/graph.bin")


def test_count_without_gpu_fails_loudly(gat):
    if gat.device_count() > 0:
        pytest.skip("a GPU is visible")
    from graphautotune_amd._lib import GAT_ENODEV, GatError
    g = gat.Graph.erdos_renyi(50, 0.2, 1)
    with pytest.raises(GatError) as e:
        gat.Plan(g, ["011101110"], "unique").count()
    assert e.value.code == GAT_ENODEV


def test_plan_compile_all_modes(gat):
    """Every connected 3..5-vertex pattern compiles in every mode; reference
    mode may refuse only the documented undefined-behaviour schedules."""
    from patterns import connected_patterns
    from graphautotune_amd._lib import GAT_EUNSUPPORTED, GatError
    g = gat.Graph.erdos_renyi(30, 0.3, 2)
    for n in (3, 4, 5):
        for p in connected_patterns(n):
            for mode in ("unique", "labeled"):
                gat.Plan(g, [p], mode)
            try:
                gat.Plan(g, [p], "reference")
            except GatError as e:
                assert e.code == GAT_EUNSUPPORTED
    plan = gat.Plan(g, connected_patterns(5), "unique")
    text = plan.describe()
    assert text.count("term pattern") == 21


def test_pm_cli_usage():
    pm = os.path.join(ROOT, "graphautotune_amd", "bin", "pm")
    assert os.path.exists(pm)
    r = subprocess.run([pm], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr
