"""The learned cost model against the method without it (GPU, through the
C ABI): trained on engine-measured runtimes of random kernel configs on small
R-MAT graphs, the MLP ranks a pool of configs for the batched set's unseen
6-vertex patterns on a larger R-MAT graph. Its top picks, re-measured, must
come closer to the pool's best (every config measured: the grid judge) than
the same number of random picks does on average (the exact expectation over
the pool), summed over the patterns."""
import pytest

pytestmark = pytest.mark.gpu


def test_model_picks_beat_random_picks(gat, gpu):
    from graphautotune_amd.tuning import config_study as cs
    res = cs.run_study([11, 12, 13], 14, patterns_per_scale=6, per_workload=14, pool=16, budget=3, target_ms=12.0,
                       steps=600, seed=0, log=lambda s: print(s, flush=True))
    print({k: v for k, v in res.items() if k not in ("results", "train_runs")}, flush=True)
    assert res["eval"]["patterns"] >= 6
    for r in res["results"]:
        assert r["grid_best_ms"] <= r["model_pick_ms"] and r["pool"] >= 10
    assert res["mean_model_regret"] < res["mean_random_regret"], res
