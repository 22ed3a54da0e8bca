"""Sub-problem decomposition on the GPU (include/cvrp_decompose.h) against
the oracle.

* aco_set_ant_at: the ant's tour and its step-order length equal the
  oracle's length of the same tour; bad tours are refused.
* aco_pheromone_from_subs: the master's pheromone and total_info after the
  write-back equal the oracle's scatter of the sub-colonies' trails
  (best_pheromone, problem.h:138-139) with compute_total_info's formula.
* aco_decompose_round: the merged tour equals a replay on the oracle from
  the same best-so-far: oracle partition, one oracle colony per solved
  sub-problem (same parameters and seeds), oracle merge.
* solve_decomposed, the outer loop: best tour, length and final pheromone
  equal an oracle replay of the same loop (master iterations, rounds,
  merged tours put in as the last ant, with and without the pheromone
  write-back).
* An instance past the local-search LDS limit (3001 nodes, which a colony
  with local search refuses): the decomposition still runs, every merged
  tour is a valid solution and never longer than the best it started from.
  The same at the library's 4096-node limit.
"""
import numpy as np
import pytest

import oracle
import cvrp_aco_ocl_amd as ca
from cvrp_aco_ocl_amd import decompose as dc
from cvrp_aco_ocl_amd import instances as I

from helpers import log, oracle_problem

pytestmark = pytest.mark.gpu

THREADS = 16


def _sub_colonies(inst, parts, *, m, k, ls_flags, seed, iters):
    """Oracle colonies of the sub-problems with at least two customers, as
    aco_decompose_round sets them up; returns (bests per part, colonies by
    solved index, solved part indices)."""
    solved = [s for s, (r, _) in enumerate(parts) if len(r) >= 3]
    subs = [dc.sub_problem(inst, parts[s][0]) for s in solved]
    kk = max(1, min(k, min(sb.n for sb in subs) - 2))
    bests = [np.zeros(0, np.int32) for _ in parts]
    cols = []
    for i, (s, sb) in enumerate(zip(solved, subs)):
        P = oracle_problem(sb, k=kk)
        col = oracle.Colony(P, m, selection=oracle.SEL_KNN, rng=oracle.RNG_PHILOX, ls_flags=ls_flags,
                            seed=seed + i, threads=THREADS)
        col.init()
        for _ in range(iters):
            col.iteration()
        bests[s] = np.asarray(col.best()[0], np.int32)
        cols.append(col)
    return bests, cols, solved


def test_set_ant_matches_oracle_length():
    inst = I.clustered(5, customers=200)
    tour = oracle.Problem(inst.dist, inst.demand, inst.capacity).nn_tour()[0]
    with ca.GpuACO(inst, n_ants=16, k=16, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX, ls_flags=7) as g:
        g.init_aco()
        g.construct_solutions()
        g.set_ant(5, tour)
        tours, sizes, lens = g.ants()
        assert sizes[5] == len(tour)
        np.testing.assert_array_equal(tours[5, : sizes[5]], tour)
        assert lens[5] == np.float32(oracle.tour_length(inst.dist, tour))
        dup = tour.copy()
        dup[2] = dup[1]                      # a customer twice, another missing
        empty = np.concatenate([tour[:1], tour])  # an empty route at the start
        for bad in ([1, 2, 0], [0, 5, 9999, 0], tour[:-2].tolist() + [0], dup, empty):
            with pytest.raises(ca.AcoError):
                g.set_ant(1, np.array(bad, np.int32))
        with pytest.raises(ca.AcoError):
            g.set_ant(16, tour)


def test_pheromone_from_subs_equals_oracle():
    inst = I.clustered(6, customers=240)
    with ca.GpuACO(inst, n_ants=24, k=12, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX, ls_flags=7, seed=3) as g:
        g.init_aco()
        g.run(3)
        best, _ = g.best()
        parts = dc.partition_routes(inst, best, 5, 2)
        tau, ti = g.pheromone()
        subs = [dc.sub_problem(inst, r) for r, _ in parts]
        mapping = [3, 0, 4, 1, 2]
        with ca.GpuACO([subs[s] for s in mapping], n_ants=16, k=8, selection=ca.SEL_KNN,
                       rng=ca.RNG_PHILOX, ls_flags=7, seed=9) as sc:
            sc.init_aco()
            sc.run(4)
            sub_tau = [sc.pheromone(i)[0] for i in range(len(mapping))]
            g.pheromone_from_subs(sc, parts, mapping)
        gtau, gti = g.pheromone()
    heur = oracle.heuristic(inst.dist, 2.0)
    otau, oti = tau.copy(), ti.copy()
    for i, s in enumerate(mapping):
        oracle.pheromone_from_sub(1.0, heur, otau, oti, parts[s][0], sub_tau[i])
    np.testing.assert_array_equal(gtau, otau)
    np.testing.assert_array_equal(gti, oti)
    assert not np.array_equal(gtau, tau)


@pytest.mark.parametrize("subs,rotation", [(0, 1), (3, 4)])
def test_decompose_round_equals_oracle_replay(subs, rotation):
    inst = I.clustered(7, customers=300)
    m_sub, k_sub, iters, seed = 16, 12, 8, 7
    with ca.GpuACO(inst, n_ants=32, k=16, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX, ls_flags=7) as g:
        g.init_aco()
        g.run(4)
        bsf, blen = g.best()
        params = dc.sub_params(n_ants=m_sub, k=k_sub, seed=seed)
        tour, length, nsub, rep = dc.decompose_round(g, inst, subs=subs, sub_iterations=iters,
                                                     rotation=rotation, params=params)
    parts = oracle.partition_routes(inst.x, inst.y, inst.demand, bsf, subs, rotation)
    bests, _, _ = _sub_colonies(inst, parts, m=m_sub, k=k_sub, ls_flags=7, seed=seed, iters=iters)
    otour, olen, orep = oracle.merge_sub_tours(inst.dist, parts, bests)
    log(f"round: {nsub} sub-problems, {rep} replaced, {blen} -> {length}")
    assert nsub == len(parts)
    np.testing.assert_array_equal(tour, otour)
    assert length == olen
    assert rep == int(orep.sum())
    assert inst.check_solution(tour)
    assert length <= blen


def test_decompose_round_with_route_length_limit_equals_oracle():
    """CMT6: float distances, a route-length limit and service times; the
    sub-problems inherit both, every merged route stays within the limit."""
    import os
    from helpers import VRPLIB
    inst = ca.Instance.from_vrplib(os.path.join(VRPLIB, "CMT6.vrp"))
    assert np.isfinite(inst.max_dist) and inst.service_time > 0
    m_sub, k_sub, iters, seed = 16, 10, 6, 3
    with ca.GpuACO(inst, n_ants=24, k=12, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX, ls_flags=7) as g:
        g.init_aco()
        g.run(3)
        bsf, blen = g.best()
        tour, length, nsub, rep = dc.decompose_round(g, inst, subs=2, sub_iterations=iters, rotation=1,
                                                     params=dc.sub_params(n_ants=m_sub, k=k_sub, seed=seed))
    parts = oracle.partition_routes(inst.x, inst.y, inst.demand, bsf, 2, 1)
    bests, _, _ = _sub_colonies(inst, parts, m=m_sub, k=k_sub, ls_flags=7, seed=seed, iters=iters)
    otour, olen, orep = oracle.merge_sub_tours(inst.dist, parts, bests)
    log(f"CMT6 round: {nsub} sub-problems, {rep} replaced, {blen} -> {length}")
    np.testing.assert_array_equal(tour, otour)
    assert length == olen and rep == int(orep.sum())
    assert inst.check_solution(tour)


def _oracle_solve_decomposed(inst, iterations, *, master, sub, subs, sub_iterations, pheromone_back):
    P = oracle_problem(inst, k=master["k"])
    col = oracle.Colony(P, master["n_ants"], selection=oracle.SEL_KNN, rng=oracle.RNG_PHILOX,
                        ls_flags=master["ls_flags"], seed=master["seed"], threads=THREADS)
    col.init()
    m = master["n_ants"]
    heur = oracle.heuristic(inst.dist, 2.0)
    pending, rounds = None, []
    for it in range(iterations):
        if pending is not None:
            col.construct()
            if master["ls_flags"]:
                col.local_search()
            tours, sizes, lens = col.ants()
            tours = tours.copy()
            sizes = sizes.copy()
            lens = lens.copy()
            tours[m - 1, :] = 0
            tours[m - 1, : len(pending)] = pending
            sizes[m - 1] = len(pending)
            lens[m - 1] = np.float32(oracle.tour_length(inst.dist, pending))
            col.set_ants(tours, sizes, lens)
            col.statistics()
            col.pheromone_update()
            col.next_iteration()
            pending = None
        else:
            col.iteration()
        r = len(rounds)
        bsf, before = col.best()
        parts = oracle.partition_routes(inst.x, inst.y, inst.demand, bsf, subs, r)
        bests, cols, solved = _sub_colonies(inst, parts, m=sub["n_ants"], k=sub["k"], ls_flags=7,
                                            seed=sub["seed"] + 1000003 * r, iters=sub_iterations)
        merged, mlen, _ = oracle.merge_sub_tours(inst.dist, parts, bests)
        if pheromone_back:
            tau, ti = col.tau()
            tau, ti = tau.copy(), ti.copy()
            for i, s in enumerate(solved):
                oracle.pheromone_from_sub(1.0, heur, tau, ti, parts[s][0], cols[i].tau()[0])
            col.set_tau(tau, ti)
        rounds.append((before, mlen))
        if mlen < before:
            pending = merged
    best, blen = col.best()
    if pending is not None and rounds[-1][1] < blen:
        best, blen = pending, rounds[-1][1]
    return np.asarray(best), blen, col.tau(), rounds


@pytest.mark.parametrize("pheromone_back", [False, True])
def test_solve_decomposed_equals_oracle_replay(pheromone_back):
    inst = I.clustered(8, customers=200, routes=12)
    master = dict(n_ants=24, k=12, ls_flags=0, seed=2)
    sub = dict(n_ants=16, k=10, seed=5)
    iters, sub_iters = 4, 6
    res = dc.solve_decomposed(inst, iters, subs=4, sub_iterations=sub_iters, pheromone_back=pheromone_back,
                              master=dict(master, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX),
                              sub=dict(sub), pheromone=True)
    otour, olen, (otau, oti), orounds = _oracle_solve_decomposed(
        inst, iters, master=master, sub=sub, subs=4, sub_iterations=sub_iters, pheromone_back=pheromone_back)
    log(f"rounds gpu {[(b, l) for b, l, _, _ in res['rounds']]} oracle {orounds}")
    assert [(b, l) for b, l, _, _ in res["rounds"]] == [(np.float32(b), np.float32(l)) for b, l in orounds]
    np.testing.assert_array_equal(res["tour"], otour)
    assert res["length"] == olen
    assert inst.check_solution(res["tour"])
    gtau, gti = res["pheromone"]
    np.testing.assert_array_equal(gtau, otau)
    np.testing.assert_array_equal(gti, oti)


def test_decomposition_past_the_local_search_limit():
    inst = I.clustered(11, customers=3000, nclust=12, sigma=40.0, routes=40)
    with pytest.raises(ca.AcoError):
        ca.GpuACO(inst, n_ants=64, k=16, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX, ls_flags=7)
    res = dc.solve_decomposed(inst, 3, sub_iterations=20,
                              master=dict(n_ants=256, k=16, ls_flags=0, seed=4),
                              sub=dict(n_ants=64, k=16, seed=6))
    log(f"n=3001: rounds {res['rounds']}, {res['wall_s']:.2f}s")
    assert inst.check_solution(res["tour"])
    assert res["length"] == np.float32(oracle.tour_length(inst.dist, res["tour"]))
    for before, after, nsub, rep in res["rounds"]:
        assert after <= before and nsub >= 1
    assert sum(rep for *_, rep in res["rounds"]) > 0
    assert res["length"] < res["rounds"][0][0]


def test_decomposition_at_the_node_limit():
    """4096 nodes (the library's maximum): two rounds over min(n/50, R/2)
    sub-problems; every merged tour valid, the length never higher."""
    inst = I.clustered(12, customers=4095, nclust=16, sigma=30.0, routes=60)
    res = dc.solve_decomposed(inst, 2, sub_iterations=10,
                              master=dict(n_ants=128, k=16, ls_flags=0, seed=2),
                              sub=dict(n_ants=32, k=16, seed=3))
    log(f"n=4096: rounds {res['rounds']}, {res['wall_s']:.2f}s")
    assert inst.check_solution(res["tour"])
    for before, after, nsub, rep in res["rounds"]:
        assert after <= before and nsub >= 1
    assert res["length"] <= res["rounds"][0][0]
