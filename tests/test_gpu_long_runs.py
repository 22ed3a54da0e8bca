"""Long runs of the HIP path through the C ABI.

* The 100-customer workload under the reference's rules (I-Roulette over
  the full row with Park-Miller lane streams, 2-opt, tau0 = 1/(n L_nn)) for
  1000 graph-replayed iterations: past about 800 iterations tau * eta^2
  underflows through the denormals to zero, and a depot step with no mass
  takes the full-row greedy pick.  Tours, tau, total_info, records and the
  fallback count stay bit-equal to the oracle colony at every checkpoint,
  and the fallback path is reached.
* The 1001-node mixed workload (256 ants, k = 32, 2-opt + swap +
  relocate) for 20 graph-replayed iterations under property checks.
* The same workload at full size against the oracle colony, under the
  candidate roulette (the launch shape the autotuner picks for it: 12
  local-search waves per ant, 32-row ring of 16-bit rows with prefetch)
  and under the reference's rules (I-Roulette with Park-Miller lane
  streams, four waves per ant since the ants fit one per CU): tours,
  lengths, tau, total_info, records and evaluated-move counts bit-equal
  after every iteration.
* An instance at the local-search LDS boundary: the largest n that
  aco_create accepts runs whole iterations; one more node is refused with
  the LDS message (static LDS of the kernels counted).
"""
import math
import os

import numpy as np
import pytest

import oracle
import cvrp_aco_ocl_amd as ca
from cvrp_aco_ocl_amd import instances as I

from helpers import log, oracle_problem

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def test_reference_rules_1000_iterations_into_denormals():
    inst = I.uniform_100(0)
    m, seed = 100, 3
    P = oracle_problem(inst)
    col = oracle.Colony(P, m, selection=ca.SEL_IROULETTE, rng=ca.RNG_PARK_MILLER, ls_flags=1,
                        tau0_rule=ca.TAU0_NN_TOUR, seed=seed, threads=THREADS)
    col.init()
    with ca.GpuACO(inst, n_ants=m, seed=seed, selection=ca.SEL_IROULETTE, rng=ca.RNG_PARK_MILLER,
                   ls_flags=ca.LS_2OPT, tau0_rule=ca.TAU0_NN_TOUR, use_graph=True) as g:
        g.init_aco()
        done = 0
        for stop in (250, 500, 750, 1000):
            g.run(stop - done)
            for _ in range(stop - done):
                col.iteration()
            done = stop
            gt, gs, gl = g.ants()
            ot, osz, ol = col.ants()
            np.testing.assert_array_equal(gs, osz, err_msg=f"sizes at {stop}")
            for a in range(m + 1):
                np.testing.assert_array_equal(gt[a, : gs[a]], ot[a, : osz[a]],
                                              err_msg=f"ant {a} at iteration {stop}")
            np.testing.assert_array_equal(gl, ol)
            gtau, gti = g.pheromone()
            otau, oti = col.tau()
            np.testing.assert_array_equal(gtau, otau, err_msg=f"tau at {stop}")
            np.testing.assert_array_equal(gti, oti, err_msg=f"total_info at {stop}")
            grec = [(r[0], r[1]) for r in g.bsf_records()]
            orec = [(r[0], r[1]) for r in col.bsf()]
            assert grec == orec
            fb = g.counters()["fallbacks"]
            assert fb == col.fallbacks(), (stop, fb, col.fallbacks())
            denorm = np.count_nonzero((gti > 0) & (gti < np.finfo(np.float32).tiny))
            log(f"iteration {stop}: fallbacks {fb}, zero total_info {np.mean(gti == 0):.3f}, "
                f"denormal {denorm}, best {gl[m]}")
        assert fb > 0, "the zero-mass depot fallback was never reached"
        assert np.mean(gti == 0) > 0.5


def test_mixed_1001_twenty_graph_iterations():
    inst = I.mixed_1000(0)
    m = 256
    with ca.GpuACO(inst, n_ants=m, k=32, seed=7, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX,
                   ls_flags=7, use_graph=True) as g:
        g.init_aco()
        prev_best = math.inf
        for chunk in range(4):
            g.run(5)
            tours, sizes, lens = g.ants()
            for a in range(m + 1):
                t = tours[a, : sizes[a]]
                assert inst.check_solution(t), (chunk, a)
                assert np.float32(inst.tour_length(t)) == lens[a], (chunk, a)
            assert lens[m] <= lens[:m].min() and lens[m] <= prev_best
            prev_best = lens[m]
            tau, ti = g.pheromone()
            assert np.isfinite(tau).all() and (tau > 0).all() and np.isfinite(ti).all()
            log(f"1001 nodes: {5 * (chunk + 1)} iterations, best {lens[m]}")
        rec = g.bsf_records()
        assert all(b[1] < a[1] for a, b in zip(rec, rec[1:]))
        assert rec[-1][1] == lens[m]
        cnt = g.counters()
        assert cnt["iterations"] == 21 and cnt["ls_moves"] > 0


@pytest.mark.parametrize("rules", ["candidate_roulette", "reference"])
def test_mixed_1001_matches_oracle_colony(rules):
    inst = I.mixed_1000(0)
    m, iters, seed = 256, 3, 11
    if rules == "reference":
        sel, rng, k = ca.SEL_IROULETTE, ca.RNG_PARK_MILLER, 32
    else:
        sel, rng, k = ca.SEL_KNN, ca.RNG_PHILOX, 32
    log(f"1001 nodes vs oracle colony: {rules}")
    P = oracle_problem(inst, k=k)
    col = oracle.Colony(P, m, selection=sel, rng=rng, ls_flags=7, seed=seed, threads=THREADS)
    col.init()
    with ca.GpuACO(inst, n_ants=m, k=k, seed=seed, selection=sel, rng=rng, ls_flags=7,
                   use_graph=True) as g:
        g.init_aco()
        if rules == "candidate_roulette":
            g.set_launch_shape(12, 32, 1, 0, 1)
        for it in range(iters):
            g.run_aco_iteration()
            col.iteration()
            gt, gs, gl = g.ants()
            ot, osz, ol = col.ants()
            np.testing.assert_array_equal(gs, osz, err_msg=f"sizes at {it}")
            for a in range(m + 1):
                np.testing.assert_array_equal(gt[a, : gs[a]], ot[a, : osz[a]],
                                              err_msg=f"iteration {it} ant {a}")
            np.testing.assert_array_equal(gl, ol)
            assert g.counters()["ls_evals"] == col.ls_evals(), it
            log(f"iteration {it}: best {gl[m]}, evaluated moves {col.ls_evals()}")
        gtau, gti = g.pheromone()
        otau, oti = col.tau()
        np.testing.assert_array_equal(gtau, otau)
        np.testing.assert_array_equal(gti, oti)
        assert [(r[0], r[1]) for r in g.bsf_records()] == [(r[0], r[1]) for r in col.bsf()]


def _uniform(customers, seed):
    r = np.random.default_rng(seed)
    x = r.integers(0, 1001, size=customers + 1).astype(np.float64)
    y = r.integers(0, 1001, size=customers + 1).astype(np.float64)
    x[0] = y[0] = 500.0
    dem = np.concatenate([[0], r.integers(1, 51, size=customers)]).astype(np.int32)
    return ca.Instance.from_coords(x, y, dem, int(math.ceil(dem.sum() / 80)), ca.DIST_NINT)


def _accepted(inst):
    try:
        g = ca.GpuACO(inst, n_ants=4, k=16, seed=2, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX,
                      ls_flags=7)
    except ca.AcoError as e:
        assert "LDS" in str(e), str(e)
        return None
    return g


def test_local_search_lds_boundary():
    base = _uniform(3000, 5)
    lo, hi = 2001, 3001          # accepted, refused
    assert _accepted(base.subset(hi)) is None
    while hi - lo > 1:
        mid = (lo + hi) // 2
        g = _accepted(base.subset(mid))
        if g is None:
            hi = mid
        else:
            g.close()
            lo = mid
    log(f"largest n accepted with local search: {lo}")
    inst = base.subset(lo)
    with _accepted(inst) as g:
        g.init_aco()
        g.run(2)
        tours, sizes, lens = g.ants()
        for a in range(5):
            assert inst.check_solution(tours[a, : sizes[a]])
