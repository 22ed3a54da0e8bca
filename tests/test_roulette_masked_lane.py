"""The candidate roulette takes only lanes with mass.

tests/golden/roulette_masked_lane.npz is a state captured from a GPU run (a
67-node single-route sub-problem of the 3001-node decomposition workload,
its pheromone and total_info before iteration 75, ant 26, Philox seed):
at step 26 the draw is u = 0.9999873, so r = u * total lies within 4e-8 of
the wheel's end.  The wheel's prefix is the DPP scan (Hillis-Steele inside
rows of 16 lanes), whose additions are grouped differently per lane, so the
prefix is not monotone in floating point: lane 6 (node 22, already visited,
mass 0) rounds above lane 5 and above r.  Taking the first lane with
prefix > r regardless of mass visited node 22 twice.  The rule takes the
first lane with mass and prefix > r.

* CPU: the scan of that step, restated in numpy float32, shows the
  non-monotone prefix at the massless lane with r between; the oracle's
  ant is a valid solution.
* GPU: every ant of that iteration equals the oracle's and is valid.
"""
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
CASE = np.load(os.path.join(HERE, "golden", "roulette_masked_lane.npz"))


def _problem():
    return oracle.Problem(CASE["dist"], CASE["demand"], int(CASE["capacity"]), k=int(CASE["k"]))


def _dpp_scan(pv):
    """Inclusive scan in the DPP order (offsets 1, 2, 4, 8 inside rows of
    16 lanes, then the row totals), float32."""
    x = np.array(pv, np.float32).copy()
    for d in (1, 2, 4, 8):
        y = x.copy()
        for c in range(64):
            if (c & 15) >= d:
                y[c] = np.float32(x[c] + x[c - d])
        x = y
    t15, t47 = x[15], x[47]
    x[16:32] = x[16:32] + t15
    x[48:64] = x[48:64] + t47
    x[32:64] = x[32:64] + x[31]
    return x


def test_scan_is_not_monotone_at_the_visited_lane():
    P = _problem()
    n, k = P.n, P.k
    tiknn = oracle.gather_knn(P, CASE["total_info"]).reshape(n, k)
    nnk = np.asarray(P.nnk).reshape(n, k)
    bad = CASE["bad_tour"]
    vis = np.zeros(n, bool)
    vis[0] = True
    load = 0
    for step in range(1, 26):
        vis[bad[step]] = True
        load += int(CASE["demand"][bad[step]])
    cur = bad[25]
    pv = np.zeros(64, np.float32)
    for c in range(k):
        j = nnk[cur, c]
        if not vis[j] and load + CASE["demand"][j] <= int(CASE["capacity"]):
            pv[c] = tiknn[cur, c]
    x = _dpp_scan(pv)
    u = oracle.lib.orc_philox_u01(int(CASE["seed"]), int(CASE["iteration"]), int(CASE["ant"]), 26)
    r = np.float32(np.float32(u) * x[k - 1])
    unmasked = int(np.flatnonzero(x[:k] > r)[0])
    assert pv[unmasked] == 0.0 and vis[nnk[cur, unmasked]]          # the old rule's pick
    assert x[unmasked] > x[unmasked - 1]                             # non-monotone prefix
    masked = int(np.flatnonzero((pv[:k] > 0) & (x[:k] > r))[0])
    assert pv[masked] > 0.0 and not vis[nnk[cur, masked]]


def test_oracle_ant_is_valid():
    P = _problem()
    ti = CASE["total_info"]
    t, L, _ = P.construct_ant(ti, oracle.gather_knn(P, ti), 1, 1, int(CASE["seed"]), int(CASE["iteration"]),
                              int(CASE["ant"]), np.ones(64, np.int32))
    assert P.check(t)
    assert not np.array_equal(np.asarray(t), CASE["bad_tour"])


@pytest.mark.gpu
def test_gpu_iteration_equals_oracle():
    import cvrp_aco_ocl_amd as ca
    from helpers import oracle_construct_all
    n = CASE["dist"].shape[0]
    x = np.zeros(n)
    inst = ca.Instance(n=n, x=x, y=x, demand=CASE["demand"].astype(np.int32), capacity=int(CASE["capacity"]),
                       dist=CASE["dist"].astype(np.float32))
    seed, it, m = int(CASE["seed"]), int(CASE["iteration"]), 64
    with ca.GpuACO(inst, n_ants=m, k=int(CASE["k"]), seed=seed, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX,
                   use_graph=False) as g:
        g.init_aco()
        g.run(it - g.iteration)
        assert g.iteration == it
        g.set_pheromone(CASE["tau"], CASE["total_info"])
        g.construct_solutions()
        tours, sizes, lens = g.ants()
    P = _problem()
    otours, olens, _, _ = oracle_construct_all(P, CASE["total_info"], 1, 1, seed, it, np.ones((m, 64), np.int32))
    for a in range(m):
        assert sizes[a] == len(otours[a]), a
        np.testing.assert_array_equal(tours[a, : sizes[a]], otours[a], err_msg=f"ant {a}")
        assert inst.check_solution(tours[a, : sizes[a]])
    np.testing.assert_array_equal(lens[:m], olens)
