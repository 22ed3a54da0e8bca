"""The reference's CPU 2-opt order on the GPU (ACO_LS_2OPT_RANDOM) and draw
injection through the C ABI.

* Local search with the random scan order (LocalSearch::two_opt_single_route
  with generate_random_permutation, localSearch.cpp:97-119,180-294): every
  ant equals the oracle's random-order 2-opt run ant after ant on one
  Park-Miller stream (which the CPU tests pin against the reference build),
  the evaluated-move count and the stream's state afterwards included.
* Whole colonies with it, under both next-node rules.
* A draw of exactly 1.0 (a state >= 2147483584 before the draw): in the
  permutation (where the reference asserts; both sides take the last
  offset), and in the candidate roulette with Park-Miller draws, where
  u * total reaches the wheel's end and the last lane with mass is taken
  (set through aco_set_rng_state).
"""
import os

import numpy as np
import pytest

import oracle
import cvrp_aco_ocl_amd as ca
from cvrp_aco_ocl_amd import instances as I

from helpers import VRPLIB, log, oracle_construct_all, oracle_problem, pack_tours

pytestmark = pytest.mark.gpu

M31 = 2147483647
THREADS = min(16, os.cpu_count() or 1)


def _state_before_one():
    """A Park-Miller state whose next draw is exactly 1.0f."""
    inv = pow(16807, M31 - 2, M31)
    return (2147483600 * inv) % M31


def _instances():
    return {"u100": I.uniform_100(21),
            "x101": ca.Instance.from_vrplib(os.path.join(VRPLIB, "X-n101-k25.vrp"), ca.DIST_NINT),
            "cmt6": ca.Instance.from_vrplib(os.path.join(VRPLIB, "CMT6.vrp"))}


def _oracle_tours(inst, P, m, seed):
    tau, ti = oracle.pheromone_init(P, 0.5)
    pm = np.array([[oracle.pm_seed(seed, a, l) for l in range(64)] for a in range(m)], np.int32)
    tours, lens, _, _ = oracle_construct_all(P, ti, 0, 0, 0, 0, pm)
    return tours, lens


@pytest.mark.parametrize("flags", [8, 8 | 2 | 4])
@pytest.mark.parametrize("name", ["u100", "x101", "cmt6"])
def test_random_order_local_search_matches_oracle(name, flags):
    inst = _instances()[name]
    m, z0 = 48, 777
    P = oracle_problem(inst)
    otours, olens = _oracle_tours(inst, P, m, 3)
    packed, sizes = pack_tours(otours + [otours[0]], 2 * inst.n)
    lens = np.concatenate([olens, olens[:1]]).astype(np.float32)
    log(f"random-order local search {name} flags={flags}")
    with ca.GpuACO(inst, n_ants=m, seed=1, ls_flags=flags, use_graph=False) as g:
        g.init_aco()
        g.set_ls_seed(z0)
        g.set_ants(packed, sizes, lens)
        g.local_search()
        tours, gsizes, glens = g.ants()
        z1 = g.ls_seed()
        g.update_statistics()
        cnt = g.counters()
    z, evals = z0, 0
    for a in range(m):
        t, L, ev, z = P.local_search_seeded(otours[a], flags, z)
        evals += ev
        assert gsizes[a] == len(t), a
        np.testing.assert_array_equal(tours[a, : len(t)], t, err_msg=f"ant {a}")
        assert glens[a] == np.float32(L), (a, glens[a], L)
    assert z1 == z
    assert cnt["ls_evals"] == evals


@pytest.mark.parametrize("rule", ["knn", "iroulette"])
def test_random_order_colony_matches_oracle(rule):
    inst = I.uniform_100(23)
    m, seed, iters = 64, 5, 6
    sel = ca.SEL_KNN if rule == "knn" else ca.SEL_IROULETTE
    rng = ca.RNG_PHILOX if rule == "knn" else ca.RNG_PARK_MILLER
    k = 20 if rule == "knn" else 0
    flags = 8 | 2 | 4 if rule == "knn" else 8
    P = oracle_problem(inst, k=k)
    col = oracle.Colony(P, m, selection=sel, rng=rng, ls_flags=flags, tau0_rule=ca.TAU0_NN_TOUR,
                        seed=seed, threads=THREADS)
    col.init()
    with ca.GpuACO(inst, n_ants=m, k=k, seed=seed, selection=sel, rng=rng, ls_flags=flags,
                   tau0_rule=ca.TAU0_NN_TOUR) as g:
        g.init_aco()
        assert g.ls_seed() == col.ls_seed() == oracle.ls_seed(seed)
        for it in range(iters):
            g.run_aco_iteration()
            col.iteration()
            gt, gs, gl = g.ants()
            ot, osz, ol = col.ants()
            np.testing.assert_array_equal(gs, osz, err_msg=f"iteration {it}")
            for a in range(m + 1):
                np.testing.assert_array_equal(gt[a, : gs[a]], ot[a, : osz[a]],
                                              err_msg=f"iteration {it} ant {a}")
            np.testing.assert_array_equal(gl, ol)
            assert g.ls_seed() == col.ls_seed(), it
        gtau, gti = g.pheromone()
        otau, oti = col.tau()
        np.testing.assert_array_equal(gtau, otau)
        np.testing.assert_array_equal(gti, oti)
        assert [(r[0], r[1]) for r in g.bsf_records()] == [(r[0], r[1]) for r in col.bsf()]


def test_permutation_draw_of_one():
    """The first draw of ant 0's first route is exactly 1.0f."""
    inst = I.uniform_100(25)
    m = 8
    z0 = _state_before_one()
    s = oracle.C.c_int32(z0)
    assert oracle.lib.orc_ran01(s) == 1.0
    P = oracle_problem(inst)
    otours, olens = _oracle_tours(inst, P, m, 9)
    packed, sizes = pack_tours(otours + [otours[0]], 2 * inst.n)
    lens = np.concatenate([olens, olens[:1]]).astype(np.float32)
    with ca.GpuACO(inst, n_ants=m, seed=1, ls_flags=8, use_graph=False) as g:
        g.init_aco()
        g.set_ls_seed(z0)
        g.set_ants(packed, sizes, lens)
        g.local_search()
        tours, gsizes, _ = g.ants()
        z1 = g.ls_seed()
    z = z0
    for a in range(m):
        t, L, _, z = P.local_search_seeded(otours[a], 8, z)
        np.testing.assert_array_equal(tours[a, : len(t)], t, err_msg=f"ant {a}")
    assert z1 == z


@pytest.mark.parametrize("rule", ["knn", "iroulette"])
def test_roulette_draw_of_one_through_injected_state(rule):
    """aco_set_rng_state puts lane streams whose next draw is 1.0f into every
    ant.  Candidate roulette (Park-Miller): r = 1.0 * total, no prefix
    exceeds it, the last lane with mass is taken.  I-Roulette: the lane's
    probability is (1.0 + 0.001) * total_info * 100."""
    inst = I.uniform_100(27)
    m, k = 16, 20
    sel = ca.SEL_KNN if rule == "knn" else ca.SEL_IROULETTE
    P = oracle_problem(inst, k=k if rule == "knn" else 0)
    z1 = _state_before_one()
    with ca.GpuACO(inst, n_ants=m, k=k if rule == "knn" else 0, seed=3, selection=sel,
                   rng=ca.RNG_PARK_MILLER, tau0_rule=ca.TAU0_NN_TOUR, use_graph=False) as g:
        g.init_aco()
        pm = g.rng_state()
        pm[:, 0] = z1                      # lane 0: the candidate roulette's stream
        pm[::2, 5] = z1                    # I-Roulette: lane 5 of every other ant
        g.set_rng_state(pm)
        np.testing.assert_array_equal(g.rng_state(), pm)
        _, ti = g.pheromone()
        it = g.iteration
        g.construct_solutions()
        tours, sizes, lens = g.ants()
        pm_after = g.rng_state()
    otours, olens, opm, _ = oracle_construct_all(P, ti, 1 if rule == "knn" else 0, 0, 3, it, pm)
    for a, t in enumerate(otours):
        assert sizes[a] == len(t), a
        np.testing.assert_array_equal(tours[a, : sizes[a]], t, err_msg=f"ant {a}")
        assert lens[a] == olens[a]
    if rule == "knn":
        np.testing.assert_array_equal(pm_after[:, 0], opm[:, 0])
    else:
        np.testing.assert_array_equal(pm_after, opm)
    with ca.GpuACO(inst, n_ants=m, seed=3) as g:
        bad = np.ones((m, 64), np.int32)
        bad[3, 7] = 0
        with pytest.raises(ca.AcoError):
            g.set_rng_state(bad)
