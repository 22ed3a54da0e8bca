"""Caller-chosen construction candidates (aco_set_candidates_at) against the
oracle.

* Rows of k customers drawn from each node's 2k nearest, in shuffled order
  (not distance-sorted), set before aco_init and after it: the whole loop
  (roulette over the candidates, full local search with the untouched
  distance-sorted 2-opt rows, statistics, pheromone) is bit-equal to an
  oracle colony whose Problem takes the same rows, iteration by iteration.
* The device arrays hold the new ids and their distances; the 2-opt rows
  are unchanged.
* Rows with a repeated node, the node itself, the depot or an id out of
  range are refused.
"""
import numpy as np
import pytest

import oracle
import cvrp_aco_ocl_amd as ca
from cvrp_aco_ocl_amd import instances as I

from helpers import log

pytestmark = pytest.mark.gpu


def _shuffled_candidates(inst, k, seed):
    r = np.random.default_rng(seed)
    pool = inst.nn_lists(2 * k)
    return np.stack([r.permutation(row)[:k] for row in pool]).astype(np.int32)


@pytest.mark.parametrize("when", ["before_init", "after_init"])
@pytest.mark.parametrize("rng", [ca.RNG_PHILOX, ca.RNG_PARK_MILLER])
def test_candidates_match_oracle_colony(when, rng):
    inst = I.clustered(4, customers=240)
    k, m, iters, seed = 12, 48, 5, 21
    cand = _shuffled_candidates(inst, k, 3)
    log(f"candidates {when} rng={rng}")
    with ca.GpuACO(inst, n_ants=m, k=k, seed=seed, selection=ca.SEL_KNN, rng=rng, ls_flags=7,
                   tau0_rule=ca.TAU0_NN_TOUR, use_graph=True) as g:
        nnls0 = g.setup_arrays()["nnls"]
        if when == "before_init":
            g.set_candidates(cand)
            g.init_aco()
        else:
            g.init_aco()
            g.set_candidates(cand)
        arrs = g.setup_arrays()
        np.testing.assert_array_equal(arrs["nnk"], cand)
        np.testing.assert_array_equal(arrs["nnkd"], np.take_along_axis(np.asarray(inst.dist, np.float32), cand, 1))
        np.testing.assert_array_equal(arrs["nnls"], nnls0)
        P = oracle.Problem(inst.dist, inst.demand, inst.capacity, inst.max_dist, inst.service_time,
                           k=k, nnk=cand)
        col = oracle.Colony(P, m, selection=ca.SEL_KNN, rng=rng, ls_flags=7, tau0_rule=ca.TAU0_NN_TOUR,
                            seed=seed)
        col.init()
        for it in range(iters):
            g.run_aco_iteration()
            col.iteration()
            gt, gs, gl = g.ants()
            ot, osz, ol = col.ants()
            np.testing.assert_array_equal(gs, osz, err_msg=f"iteration {it}")
            for a in range(m + 1):
                np.testing.assert_array_equal(gt[a, : gs[a]], ot[a, : osz[a]], err_msg=f"iteration {it} ant {a}")
            np.testing.assert_array_equal(gl, ol)
        gtau, gti = g.pheromone()
        otau, oti = col.tau()
        np.testing.assert_array_equal(gtau, otau)
        np.testing.assert_array_equal(gti, oti)
        assert [(r[0], r[1]) for r in g.bsf_records()] == [(r[0], r[1]) for r in col.bsf()]


def test_bad_candidate_rows_are_refused():
    inst = I.uniform_100(2)
    k = 8
    with ca.GpuACO(inst, n_ants=8, k=k, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX) as g:
        good = inst.nn_lists(k).astype(np.int32)
        for i, h, v in ((3, 2, int(good[3, 1])), (5, 0, 5), (7, 4, 0), (9, 1, inst.n)):
            bad = good.copy()
            bad[i, h] = v
            with pytest.raises(ca.AcoError):
                g.set_candidates(bad)
        with pytest.raises(ca.AcoError):
            g.set_candidates(good[:, : k - 1])
        g.set_candidates(good)
        g.init_aco()
        g.run(2)
        t, L = g.best()
        assert inst.check_solution(t)
