"""Tuner CLI (tuning/main.py:19-52 of the reference):

    python -m graphautotune_amd.tuning.main <data> <pattern> [-r tune|run_default|run_best] [-c]

<data> is a CSR .bin file (graph_backend.cuh layout) or rmat:<scale>[:<edge factor>[:<seed>]].
The history lives in tuning_results/record.json under --results (default
./tuning_results), as in tuning/tuning/common/const.py:18-22.
"""
import argparse
import json
import logging
import os
import shutil

from .. import core, workloads
from .driver import Driver
from .manipulator import Manipulator
from .modeling import Modeling
from .space import Config
from .study import featurizer, study_space
from .tuner import Tuner

logger = logging.getLogger("main")


def load_graph(spec):
    if spec.startswith("rmat:"):
        f = spec.split(":")[1:]
        scale = int(f[0])
        ef = int(f[1]) if len(f) > 1 else 16
        seed = int(f[2]) if len(f) > 2 else scale
        return core.Graph.rmat(scale, ef, 0.57, 0.19, 0.19, seed)
    return core.Graph.load(spec, binary=not spec.endswith(".txt"))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("data")
    ap.add_argument("pattern", help="pattern string, or several separated by commas (batched), or 'batched' "
                                    "for workloads.batched_patterns()")
    ap.add_argument("--schedule-ranks", default=None, metavar="JSON",
                    help="a result of tuning.batched_study: tune the plan with its per-pattern Schedule ranks")
    ap.add_argument("-r", "--run_mode", default="tune", choices=["tune", "run_default", "run_best"])
    ap.add_argument("-c", "--clear", action="store_true")
    ap.add_argument("--results", default="tuning_results")
    ap.add_argument("--rounds", type=int, default=20)
    ap.add_argument("--batch", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=20)
    args = ap.parse_args(argv)
    logging.basicConfig(level=logging.INFO)
    if args.clear and os.path.exists(args.results):
        shutil.rmtree(args.results)
    os.makedirs(args.results, exist_ok=True)
    g = load_graph(args.data)
    patterns = workloads.batched_patterns() if args.pattern == "batched" else args.pattern.split(",")
    opts = None
    if args.schedule_ranks:
        with open(args.schedule_ranks) as f:
            tuned = json.load(f)
        if [core.Pattern(p).adj for p in tuned["patterns"]] != [core.Pattern(p).adj for p in patterns]:
            ap.error("%s holds Schedule ranks for other patterns" % args.schedule_ranks)
        opts = {"schedule": "ranked", "schedule_rank": list(tuned["ranks"])}
    driver = Driver(g, patterns, "unique", reps=1, warmup=0, plan_options=opts, report=logger.info)
    feat = featurizer(g, patterns)
    model = Modeling(feat, os.path.join(args.results, "record.json"))
    if args.run_mode == "run_default":
        t = driver.run(Config())
        print(json.dumps({"config": Config().as_dict(), "seconds": t, "counts": driver.counts}))
        return
    if args.run_mode == "run_best":
        best, _ = model.get_best_config()
        t = driver.run(best)
        print(json.dumps({"config": best.as_dict(), "seconds": t, "counts": driver.counts}))
        return
    manip = Manipulator(model, driver, study_space())
    best, t = Tuner(manip, args.rounds, args.batch, args.warmup).tune(log=logger.info)
    out = {"config": best.as_dict(), "seconds": t, "counts": driver.counts}
    with open(os.path.join(args.results, "best_config.json"), "w") as f:
        json.dump(out, f)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
