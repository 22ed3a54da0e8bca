"""Choosing the candidate-list width k.

k changes the search itself (which edges the roulette can draw), so it is
chosen by solution quality at equal wall-clock time rather than by kernel
speed: for each candidate k a batched colony runs on sample instances for
the same time budget (launch shapes autotuned per k), and the k with the
lowest mean cost wins.  Launch shapes (aco_autotune) leave results unchanged
and are tuned separately.

    python -m cvrp_aco_ocl_amd.tuning --instances 16 --seconds 1.0
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import _lib
from .solver import GpuACO


def _run_for(insts, k, seconds, n_ants, seed, ls_flags, chunk=5):
    g = GpuACO(list(insts), n_ants=n_ants, k=k, selection=_lib.SEL_KNN, rng=_lib.RNG_PHILOX,
               ls_flags=ls_flags, tau0_rule=_lib.TAU0_NN_TOUR, seed=seed)
    try:
        g.init_aco()
        g.autotune(reps=2)
        g.synchronize()
        t0 = time.perf_counter()
        its = 0
        while time.perf_counter() - t0 < seconds:
            g.run(chunk)
            g.synchronize()
            its += chunk
        costs = [g.best(b)[1] for b in range(len(insts))]
        return float(np.mean(costs)), its
    finally:
        g.close()


def tune_candidate_k(insts, ks=(8, 12, 16, 20, 30), seconds=1.0, n_ants=64, seed=1,
                     ls_flags=7):
    """Returns (best k, {k: {"mean_cost", "iterations"}}) for equal time budgets."""
    n_min = min(it.n for it in insts)
    res = {}
    for k in ks:
        if k < 1 or k > min(64, n_min - 2):
            continue
        cost, its = _run_for(insts, k, seconds, n_ants, seed, ls_flags)
        res[k] = {"mean_cost": cost, "iterations": its}
    best = min(res, key=lambda q: res[q]["mean_cost"])
    return best, res


def main(argv=None):
    from . import instances as I
    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=16)
    ap.add_argument("--seconds", type=float, default=1.0)
    ap.add_argument("--ks", default="8,12,16,20,30")
    a = ap.parse_args(argv)
    insts = I.evaluation(a.instances, first=2000)
    best, res = tune_candidate_k(insts, tuple(int(v) for v in a.ks.split(",")), a.seconds)
    print(json.dumps({"best_k": best, "per_k": res, "instances": len(insts),
                      "seconds_per_k": a.seconds}))


if __name__ == "__main__":
    main()
