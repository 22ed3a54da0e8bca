"""Batched colonies (aco_create_batch): instances of uneven size in one
context, every launch covering all of them.  Each instance must evolve
exactly as a single-instance oracle colony seeded with seed + b."""
import numpy as np
import pytest

import oracle
import cvrp_aco_ocl_amd as ca
from cvrp_aco_ocl_amd import instances as I

from helpers import log, oracle_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", [
    dict(selection=ca.SEL_KNN, rng=ca.RNG_PHILOX, k=16, ls=7, tau0=ca.TAU0_FIRST_COLONY, m=16),
    # at most one ant per CU: the four-wave I-Roulette kernel
    dict(selection=ca.SEL_IROULETTE, rng=ca.RNG_PARK_MILLER, k=0, ls=ca.LS_2OPT, tau0=ca.TAU0_NN_TOUR, m=16),
    # 320 ants, more than the device's CUs: the single-wave I-Roulette kernel
    dict(selection=ca.SEL_IROULETTE, rng=ca.RNG_PARK_MILLER, k=0, ls=ca.LS_2OPT, tau0=ca.TAU0_FIRST_COLONY, m=64),
])
def test_batch_matches_per_instance_oracle(case):
    insts = I.batch(3, count=5)
    m, iters, seed = case["m"], 3, 40
    log(f"batch of {[x.n for x in insts]} {case}")
    with ca.GpuACO(insts, n_ants=m, k=case["k"], seed=seed, selection=case["selection"],
                   rng=case["rng"], ls_flags=case["ls"], tau0_rule=case["tau0"]) as g:
        g.init_aco()
        g.run(iters)
        assert g.B == len(insts)
        for b, inst in enumerate(insts):
            P = oracle_problem(inst, k=case["k"])
            col = oracle.Colony(P, m, selection=case["selection"], rng=case["rng"],
                                ls_flags=case["ls"], tau0_rule=case["tau0"], seed=seed + b)
            col.init()
            for _ in range(iters):
                col.iteration()
            gt, gs, gl = g.ants(b)
            ot, osz, ol = col.ants()
            np.testing.assert_array_equal(gs, osz, err_msg=f"instance {b}")
            for a in range(m + 1):
                np.testing.assert_array_equal(gt[a, : gs[a]], ot[a, : osz[a]], err_msg=f"instance {b} ant {a}")
            np.testing.assert_array_equal(gl, ol)
            gtau, gti = g.pheromone(b)
            otau, oti = col.tau()
            np.testing.assert_array_equal(gtau, otau)
            np.testing.assert_array_equal(gti, oti)
            assert [r[:2] for r in g.bsf_records(b)] == [r[:2] for r in col.bsf()]


def test_batch_256_instances_feasible():
    """The 256-instance batched workload (n in [51,501], 64 ants each) at full
    size for a few iterations: every best tour feasible and exactly priced."""
    insts = I.batch(0, count=256)
    with ca.GpuACO(insts, n_ants=64, k=16, seed=1, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX,
                   ls_flags=7) as g:
        g.init_aco()
        g.run(3)
        for b in range(0, 256, 17):
            t, L = g.best(b)
            assert insts[b].check_solution(t), b
            assert np.float32(insts[b].tour_length(t)) == np.float32(L)
        cnt = g.counters()
        assert cnt["iterations"] == 4
