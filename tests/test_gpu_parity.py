"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bar: integer/index results (tours, sizes, elites, record iterations) are
bit-exact; float results on integer distances (tour lengths, deltas) are
bit-exact; pheromone and total_info are compared bit-exactly too (the
fused pass adds deposits in elite order exactly as the oracle does) and,
as the reference's own bar, within 1e-5 relative.
"""
import math
import os

import numpy as np
import pytest

import oracle
import cvrp_aco_ocl_amd as ca
from cvrp_aco_ocl_amd import instances as I

from helpers import VRPLIB, log, oracle_construct_all, oracle_problem, pack_tours, random_tau

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5  # pheromone / total_info relative tolerance (FP32)


def _instances():
    return {
        "u100": I.uniform_100(11),
        "x101": ca.Instance.from_vrplib(os.path.join(VRPLIB, "X-n101-k25.vrp"), ca.DIST_NINT),
        "cmt6": ca.Instance.from_vrplib(os.path.join(VRPLIB, "CMT6.vrp")),  # float dist, D, svc
    }


INSTANCES = _instances()


def _set_state(g, P, seed):
    tau = random_tau(P.n, seed)
    ti = (tau * P.heur).astype(np.float32)
    g.set_pheromone(tau, ti)
    return tau, ti


@pytest.mark.parametrize("name", list(INSTANCES))
def test_construct_iroulette_matches_oracle(name):
    inst = INSTANCES[name]
    log(f"construct iroulette {name} n={inst.n}")
    with ca.GpuACO(inst, n_ants=inst.n - 1, seed=5, selection=ca.SEL_IROULETTE,
                   rng=ca.RNG_PARK_MILLER, use_graph=False) as g:
        g.init_aco()
        P = oracle_problem(inst)
        _, ti = _set_state(g, P, 3)
        pm0 = g.rng_state()
        g.construct_solutions()
        tours, sizes, lens = g.ants()
        pm1 = g.rng_state()
    otours, olens, opm, _ = oracle_construct_all(P, ti, 0, 0, 5, 0, pm0)
    for a, t in enumerate(otours):
        assert sizes[a] == len(t), (a, sizes[a], len(t))
        np.testing.assert_array_equal(tours[a, : sizes[a]], t, err_msg=f"ant {a}")
        assert lens[a] == olens[a], (a, lens[a], olens[a])
        assert P.check(t)
    np.testing.assert_array_equal(pm1, opm)


@pytest.mark.parametrize("rng", [ca.RNG_PHILOX, ca.RNG_PARK_MILLER])
@pytest.mark.parametrize("name", list(INSTANCES))
def test_construct_knn_matches_oracle(name, rng):
    inst = INSTANCES[name]
    k = 20
    log(f"construct knn {name} rng={rng}")
    with ca.GpuACO(inst, n_ants=inst.n, k=k, seed=9, selection=ca.SEL_KNN, rng=rng,
                   use_graph=False) as g:
        g.init_aco()
        P = oracle_problem(inst, k=k)
        _, ti = _set_state(g, P, 4)
        pm0 = g.rng_state()
        it = g.iteration
        g.construct_solutions()
        tours, sizes, lens = g.ants()
        pm1 = g.rng_state()
    otours, olens, opm, fb = oracle_construct_all(P, ti, 1, rng, 9, it, pm0)
    for a, t in enumerate(otours):
        assert sizes[a] == len(t), (a, sizes[a], len(t))
        np.testing.assert_array_equal(tours[a, : sizes[a]], t, err_msg=f"ant {a}")
        assert lens[a] == olens[a]
        assert P.check(t)
    if rng == ca.RNG_PARK_MILLER:
        np.testing.assert_array_equal(pm1[:, 0], opm[:, 0])
    assert fb > 0  # the full-row fallback was exercised


@pytest.mark.parametrize("apb", [1, 2, 4])
@pytest.mark.parametrize("rng", [ca.RNG_PHILOX, ca.RNG_PARK_MILLER])
def test_partially_staged_construction_matches_oracle(apb, rng):
    """1001 nodes with k = 32: the candidate table (256 KB) does not fit in
    LDS, so the staged kernel holds its first rows there and reads the rest
    from the global candidate arrays; batched with a 101-node instance whose
    whole table fits.  Every ant equals the oracle's."""
    insts = [I.mixed_1000(0), I.uniform_100(6)]
    k, m = 32, 8
    log(f"partially staged construction apb={apb} rng={rng}")
    with ca.GpuACO(insts, n_ants=m, k=k, seed=9, selection=ca.SEL_KNN, rng=rng,
                   tau0_rule=ca.TAU0_NN_TOUR, use_graph=False) as g:
        g.init_aco()     # nearest-neighbour tau0: no colony, lane states still the seeds
        g.set_launch_shape(4, 16, 1, apb, 0)
        Ps, tis = [], []
        for b, inst in enumerate(insts):
            P = oracle_problem(inst, k=k)
            tau = random_tau(P.n, 40 + b, lo=1e-6, hi=1e-3)
            ti = (tau * P.heur).astype(np.float32)
            g.set_pheromone(tau, ti, b)
            Ps.append(P)
            tis.append(ti)
        it = g.iteration
        g.construct_solutions()
        got = [g.ants(b) for b in range(len(insts))]
    fbs = 0
    for b, inst in enumerate(insts):
        tours, sizes, lens = got[b]
        pm0 = np.array([[oracle.pm_seed(9 + b, a, l) for l in range(64)] for a in range(m)],
                       np.int32)
        otours, olens, _, fb = oracle_construct_all(Ps[b], tis[b], 1, rng, 9 + b, it, pm0)
        fbs += fb
        for a, t in enumerate(otours):
            assert sizes[a] == len(t), (b, a, sizes[a], len(t))
            np.testing.assert_array_equal(tours[a, : sizes[a]], t, err_msg=f"instance {b} ant {a}")
            assert lens[a] == olens[a]


def test_construct_knn_widest_row_feasible():
    """k = 64 (the widest candidate row one wavefront holds)."""
    inst = I.uniform_100(2)
    n = inst.n
    with ca.GpuACO(inst, n_ants=64, k=n - 2 if n - 2 <= 64 else 64, seed=1,
                   selection=ca.SEL_KNN, rng=ca.RNG_PHILOX, use_graph=False) as g:
        g.init_aco()
        tours, sizes, lens = g.ants()
        for a in range(64):
            assert inst.check_solution(tours[a, : sizes[a]])


@pytest.mark.parametrize("flags", [ca.LS_2OPT, ca.LS_SWAP, ca.LS_RELOCATE,
                                   ca.LS_2OPT | ca.LS_SWAP | ca.LS_RELOCATE])
@pytest.mark.parametrize("name", list(INSTANCES))
def test_local_search_matches_oracle(name, flags):
    inst = INSTANCES[name]
    m = 64
    log(f"local search {name} flags={flags}")
    P = oracle_problem(inst)
    tau, ti = oracle.pheromone_init(P, 0.5)
    pm = np.array([[oracle.pm_seed(21, a, l) for l in range(64)] for a in range(m)], np.int32)
    otours, olens, _, _ = oracle_construct_all(P, ti, 0, 0, 0, 0, pm)
    T = 2 * inst.n
    packed, sizes = pack_tours(otours + [otours[0]], T)
    lens = np.concatenate([olens, olens[:1]]).astype(np.float32)
    with ca.GpuACO(inst, n_ants=m, seed=1, ls_flags=flags, use_graph=False) as g:
        g.init_aco()
        g.set_ants(packed, sizes, lens)
        g.local_search()
        tours, gsizes, glens = g.ants()
        g.update_statistics()  # folds the per-ant counters into the totals
        cnt = g.counters()
    changed = 0
    evals = 0
    for a in range(m):
        t, L, ev = P.local_search(otours[a], flags)
        evals += ev
        assert gsizes[a] == len(t)
        np.testing.assert_array_equal(tours[a, : len(t)], t, err_msg=f"ant {a}")
        assert glens[a] == np.float32(L), (a, glens[a], L)
        assert P.check(t)
        changed += not np.array_equal(t, otours[a])
    assert changed > 0
    # the moves the sequential search evaluates (the 4-wave speculation and
    # the lane-parallel 2-opt scans count only those)
    assert cnt["ls_evals"] == evals, (cnt["ls_evals"], evals)


def test_statistics_and_pheromone_update_match_oracle():
    inst = INSTANCES["x101"]
    m = 100
    P = oracle_problem(inst)
    tau, ti = oracle.pheromone_init(P, 0.5)
    pm = np.array([[oracle.pm_seed(33, a, l) for l in range(64)] for a in range(m)], np.int32)
    otours, olens, _, _ = oracle_construct_all(P, ti, 0, 0, 0, 0, pm)
    # force ties: several ants share the minimum length, in several groups
    order = np.argsort(olens, kind="stable")
    for a in (5, 64, 70, 99):
        otours[a] = otours[order[0]].copy()
        olens[a] = olens[order[0]]
    T = 2 * inst.n
    bsf_tour = otours[order[-1]]
    packed, sizes = pack_tours(otours + [bsf_tour], T)
    lens = np.concatenate([olens, [olens[order[-1]]]]).astype(np.float32)
    tau0 = random_tau(inst.n, 8)
    ti0 = (tau0 * P.heur).astype(np.float32)
    with ca.GpuACO(inst, n_ants=m, seed=1, use_graph=False) as g:
        g.init_aco()
        it0 = g.iteration
        nrec0 = len(g.bsf_records())
        g.set_ants(packed, sizes, lens)
        g.set_pheromone(tau0, ti0)
        g.update_statistics()
        g.pheromone_update()
        gt, gs, gl = g.ants()
        gtau, gti = g.pheromone()
        rec = g.bsf_records()
        it1 = g.iteration
    best, flag = oracle.iteration_best(lens[:m], lens[m])
    assert best == oracle.minreduce_phase0(lens[:m])
    assert flag
    np.testing.assert_array_equal(gt[m, : gs[m]], otours[best])
    assert gl[m] == lens[best]
    assert len(rec) == nrec0 + 1 and rec[-1][0] == it0 and rec[-1][1] == lens[best]
    assert it1 == it0 + 1
    lens_after = lens.copy()
    lens_after[m] = lens[best]
    el = oracle.elites(lens_after[:m], 6)
    el[0] = m
    all_tours = otours + [otours[best]]
    etours = [all_tours[e] for e in el]
    elens = [lens_after[e] for e in el]
    otau, oti = oracle.pheromone_update(P, tau0, etours, elens)
    np.testing.assert_array_equal(gtau, otau)
    np.testing.assert_array_equal(gti, oti)
    rel = np.abs(gtau - otau) / np.maximum(np.abs(otau), 1e-30)
    assert rel.max() <= REL_TOL


@pytest.mark.parametrize("case", [
    dict(name="x101", selection=ca.SEL_IROULETTE, rng=ca.RNG_PARK_MILLER, k=0, ls=0),
    dict(name="x101", selection=ca.SEL_IROULETTE, rng=ca.RNG_PARK_MILLER, k=0, ls=ca.LS_2OPT),
    dict(name="u100", selection=ca.SEL_KNN, rng=ca.RNG_PHILOX, k=20, ls=ca.LS_2OPT),
    dict(name="cmt6", selection=ca.SEL_KNN, rng=ca.RNG_PHILOX, k=20, ls=7),
])
@pytest.mark.parametrize("tau0_rule", [ca.TAU0_FIRST_COLONY, ca.TAU0_NN_TOUR])
def test_iterations_match_oracle_colony(case, tau0_rule):
    """init + 6 iterations of the whole loop, GPU (graph replay) vs oracle."""
    inst = INSTANCES[case["name"]]
    m, iters = 48, 6
    log(f"colony {case} tau0={tau0_rule}")
    P = oracle_problem(inst, k=case["k"])
    col = oracle.Colony(P, m, selection=case["selection"], rng=case["rng"], ls_flags=case["ls"],
                        tau0_rule=tau0_rule, seed=77)
    col.init()
    with ca.GpuACO(inst, n_ants=m, k=case["k"], seed=77, selection=case["selection"],
                   rng=case["rng"], ls_flags=case["ls"], tau0_rule=tau0_rule,
                   use_graph=True) as g:
        g.init_aco()
        for it in range(iters):
            g.run_aco_iteration()
            col.iteration()
            gt, gs, gl = g.ants()
            ot, osz, ol = col.ants()
            np.testing.assert_array_equal(gs, osz, err_msg=f"iter {it}")
            for a in range(m + 1):
                np.testing.assert_array_equal(gt[a, : gs[a]], ot[a, : osz[a]], err_msg=f"iter {it} ant {a}")
            np.testing.assert_array_equal(gl, ol)
        gtau, gti = g.pheromone()
        otau, oti = col.tau()
        np.testing.assert_array_equal(gtau, otau)
        np.testing.assert_array_equal(gti, oti)
        grec = [(r[0], r[1]) for r in g.bsf_records()]
        orec = [(r[0], r[1]) for r in col.bsf()]
        assert grec == orec
        assert g.iteration == col.iter


def test_graph_replay_equals_eager():
    inst = INSTANCES["u100"]
    out = []
    for use_graph in (True, False):
        with ca.GpuACO(inst, n_ants=64, k=20, seed=3, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX,
                       ls_flags=7, use_graph=use_graph) as g:
            g.init_aco()
            g.run(10)
            out.append((g.ants(), g.pheromone(), g.bsf_records()))
    (a1, p1, r1), (a2, p2, r2) = out
    for x, y in zip(a1 + p1, a2 + p2):
        np.testing.assert_array_equal(x, y)
    assert [r[:2] for r in r1] == [r[:2] for r in r2]


def test_clustered_481_properties():
    """The 481-node workload at full size: every tour feasible, stored
    lengths equal recomputed lengths exactly, best-so-far monotone, runs
    repeat bit for bit."""
    inst = I.clustered(0)
    res = []
    for _ in range(2):
        with ca.GpuACO(inst, n_ants=480, k=30, seed=12, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX,
                       ls_flags=7) as g:
            g.init_aco()
            g.run(5)
            tours, sizes, lens = g.ants()
            rec = g.bsf_records()
            cnt = g.counters()
            res.append((tours, sizes, lens))
    t1, s1, l1 = res[0]
    t2, s2, l2 = res[1]
    np.testing.assert_array_equal(t1, t2)
    np.testing.assert_array_equal(l1, l2)
    for a in range(481):
        t = t1[a, : s1[a]]
        assert inst.check_solution(t), a
        assert np.float32(inst.tour_length(t)) == l1[a]
    lens = [r[1] for r in rec]
    assert all(b < a for a, b in zip(lens, lens[1:]))
    assert lens[-1] == l1[480]
    assert cnt["ls_evals"] > 0 and cnt["iterations"] == 6


def test_mixed_1001_properties():
    inst = I.mixed_1000(0)
    with ca.GpuACO(inst, n_ants=256, k=32, seed=4, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX,
                   ls_flags=7) as g:
        g.init_aco()
        g.run(2)
        tours, sizes, lens = g.ants()
    for a in range(257):
        t = tours[a, : sizes[a]]
        assert inst.check_solution(t), a
        assert np.float32(inst.tour_length(t)) == lens[a]


def test_small_and_degenerate_instances():
    # n = 3 (two customers), few ants (elites repeat index 0), one route each
    x = np.array([0.0, 3.0, 0.0])
    y = np.array([0.0, 0.0, 4.0])
    inst = ca.Instance.from_coords(x, y, [0, 5, 5], 10)
    with ca.GpuACO(inst, n_ants=2, seed=1, ls_flags=7) as g:
        g.init_aco()
        g.run(3)
        t, L = g.best()
        assert inst.check_solution(t) and L == 12.0
    # capacity forces one customer per route
    inst = I.uniform_100(5)
    inst.capacity = 30
    with ca.GpuACO(inst, n_ants=8, k=10, seed=2, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX,
                   ls_flags=7) as g:
        g.init_aco()
        g.run(2)
        tours, sizes, lens = g.ants()
        for a in range(9):
            assert inst.check_solution(tours[a, : sizes[a]])


def test_errors_from_device_calls():
    inst = INSTANCES["u100"]
    with ca.GpuACO(inst, n_ants=8) as g:
        with pytest.raises(ca.AcoError) as e:
            g.construct_solutions()  # before init_aco
        assert e.value.code == -3
        g.init_aco()
        with pytest.raises(ca.AcoError):
            bad = np.zeros((9, 5), np.int32)
            bad[:, 2] = 9999  # node id out of range
            g.set_ants(bad, np.full(9, 5, np.int32), np.zeros(9, np.float32))


def _uniform(n_customers, seed, span=1000):
    r = np.random.default_rng(seed)
    x = r.integers(0, span + 1, size=n_customers + 1).astype(np.float64)
    y = r.integers(0, span + 1, size=n_customers + 1).astype(np.float64)
    x[0] = y[0] = span / 2
    dem = np.concatenate([[0], r.integers(1, 51, size=n_customers)]).astype(np.int32)
    return ca.Instance.from_coords(x, y, dem, int(math.ceil(dem.sum() / 80)), ca.DIST_NINT)


def test_large_instance_matches_oracle():
    """n = 2001: the default local-search shape falls back to the 8-row fp32
    ring to fit LDS; one whole iteration equals the oracle's."""
    inst = _uniform(2000, 31)
    m = 4
    P = oracle_problem(inst, k=24)
    col = oracle.Colony(P, m, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX, ls_flags=7, seed=13)
    col.init()
    with ca.GpuACO(inst, n_ants=m, k=24, seed=13, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX,
                   ls_flags=7) as g:
        g.init_aco()
        log(f"n={inst.n} shape {g.launch_shape()}")
        g.run_aco_iteration()
        col.iteration()
        gt, gs, gl = g.ants()
        ot, osz, ol = col.ants()
        np.testing.assert_array_equal(gs, osz)
        for a in range(m + 1):
            np.testing.assert_array_equal(gt[a, : gs[a]], ot[a, : osz[a]], err_msg=f"ant {a}")
            assert inst.check_solution(gt[a, : gs[a]])
        np.testing.assert_array_equal(gl, ol)


def test_too_large_for_local_search_is_a_clean_error():
    """Past what LDS holds, aco_create fails with a message; without local
    search the same instance runs."""
    inst = _uniform(3000, 5)
    with pytest.raises(ca.AcoError) as e:
        ca.GpuACO(inst, n_ants=4, k=16, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX, ls_flags=7)
    assert "LDS" in str(e.value)
    with ca.GpuACO(inst, n_ants=4, k=16, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX, ls_flags=0) as g:
        g.init_aco()
        g.run(1)
        tours, sizes, lens = g.ants()
        for a in range(4):
            assert inst.check_solution(tours[a, : sizes[a]])
