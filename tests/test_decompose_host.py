"""Host steps of the sub-problem decomposition (include/cvrp_decompose.h)
against the oracle and the reference build, without a GPU.

* aco_route_centers equals compute_route_centers of the reference build
  (vrpHelper.cpp:274-294) bit for bit, and the oracle's restatement; a
  one-customer route's centre is that customer, a route of equal demands
  has the mean of its nodes' coordinates weighted as the formula says.
* aco_sub_problem equals init_sub_problem (problem.cpp:84-130) of the
  reference build: distances and demands, and the sub-problem's neighbour
  lists from aco_compute_nn_lists equal the ones init_problem builds.
* aco_partition_routes equals the oracle's partition for several sub-problem
  counts and rotations, and the rule's invariants hold: every customer in
  exactly one sub-problem, each route whole in one, the routes of a
  sub-problem consecutive in the circular angle order (checked with an
  independent numpy atan2), balanced customer counts, n/50 by default.
* aco_merge_sub_tours equals the oracle's merge; a sub tour is taken iff it
  is strictly shorter; the merged tour is a valid solution whose length is
  the original minus the gains; brute force on tiny sub-problems.
"""
import numpy as np
import pytest

import oracle
import cvrp_aco_ocl_amd as ca
from cvrp_aco_ocl_amd import decompose as dc
from cvrp_aco_ocl_amd import instances as I

REF = oracle.load_ref()


def _instances():
    return [I.clustered(3, customers=300), I.uniform_100(5), I.mixed_1000(1), I.batch(7, count=3)[2]]


def _tour(inst, seed=0):
    """A valid solution: the nearest-neighbour tour, or with seed > 0 a
    random capacity-feasible split of a random permutation."""
    if seed == 0:
        return oracle.Problem(inst.dist, inst.demand, inst.capacity).nn_tour()[0]
    r = np.random.default_rng(seed)
    t, load = [0], 0
    for v in r.permutation(np.arange(1, inst.n)):
        if load + inst.demand[v] > inst.capacity:
            t.append(0)
            load = 0
        t.append(int(v))
        load += int(inst.demand[v])
    t.append(0)
    return np.array(t, np.int32)


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


@pytest.mark.parametrize("case", range(4))
@pytest.mark.parametrize("seed", [0, 1])
def test_route_centers_equal_reference(case, seed):
    inst = _instances()[case]
    t = _tour(inst, seed)
    beg, end, cx, cy, rd = dc.route_centers(inst, t)
    ob, oe, ocx, ocy, ord_ = oracle.route_centers(inst.x, inst.y, inst.demand, t)
    for a, b in [(beg, ob), (end, oe), (cx, ocx), (cy, ocy), (rd, ord_)]:
        np.testing.assert_array_equal(a, b)
    if REF is None:
        pytest.skip("reference build not present")
    R = len(beg)
    rcx, rcy, rrd = np.empty(R, np.float32), np.empty(R, np.float32), np.empty(R, np.int32)
    REF.ref_route_centers(inst.n, _f32(inst.x), _f32(inst.y), np.ascontiguousarray(inst.demand, np.int32),
                          np.ascontiguousarray(t, np.int32), R, beg, end, rcx, rcy, rrd)
    np.testing.assert_array_equal(cx, rcx)
    np.testing.assert_array_equal(cy, rcy)
    np.testing.assert_array_equal(rd, rrd)


def test_route_center_closed_forms():
    x = np.array([0, 10, 20, 30, 7], np.float64)
    y = np.array([0, 5, -5, 8, 3], np.float64)
    inst = ca.Instance.from_coords(x, y, [0, 4, 4, 4, 9], 100)
    t = np.array([0, 4, 0, 1, 2, 3, 0], np.int32)
    _, _, cx, cy, rd = dc.route_centers(inst, t)
    assert (cx[0], cy[0], rd[0]) == (7.0, 3.0, 9)            # one customer: itself
    assert rd[1] == 12 and abs(cx[1] - 20.0) < 1e-5 and abs(cy[1] - 8.0 / 3.0) < 1e-5


@pytest.mark.parametrize("case", range(4))
def test_sub_problem_equals_init_sub_problem(case):
    inst = _instances()[case]
    parts = oracle.partition_routes(inst.x, inst.y, inst.demand, _tour(inst), 0, 1)
    for real, _ in parts[:3]:
        sub = dc.sub_problem(inst, real)
        od, odem = oracle.sub_problem(inst.dist, inst.demand, real)
        np.testing.assert_array_equal(sub.dist, od)
        np.testing.assert_array_equal(sub.demand, odem)
        if REF is None or len(real) < 4:
            continue
        sn = len(real)
        rd, rdem = np.empty((sn, sn), np.float32), np.empty(sn, np.int32)
        rnn = np.empty((sn, sn - 2), np.int32)
        REF.ref_sub_problem(inst.n, _f32(inst.x), _f32(inst.y), np.ascontiguousarray(inst.demand, np.int32),
                            inst.capacity, np.float32(inst.max_dist), np.float32(inst.service_time),
                            _f32(inst.dist), sn, np.ascontiguousarray(real, np.int32), rd, rdem,
                            rnn.ctypes.data)
        np.testing.assert_array_equal(sub.dist, rd)
        np.testing.assert_array_equal(sub.demand, rdem)
        np.testing.assert_array_equal(sub.nn_lists(sn - 2, 0), rnn)


def _angles(inst, t):
    """Independent centre angles: demand-weighted mean in float64."""
    out = []
    rb = np.flatnonzero(np.asarray(t) == 0)
    for a, b in zip(rb[:-1], rb[1:]):
        if b - a < 2:
            continue
        nodes = np.asarray(t[a + 1:b])
        w = inst.demand[nodes].astype(np.float64)
        cx = (inst.x[nodes] * w).sum() / w.sum()
        cy = (inst.y[nodes] * w).sum() / w.sum()
        out.append((np.arctan2(cy - inst.y[0], cx - inst.x[0]), tuple(nodes)))
    return out


@pytest.mark.parametrize("case", range(4))
@pytest.mark.parametrize("subs,rotation", [(0, 0), (0, 5), (4, 2), (1, 0), (7, -3), (10 ** 6, 0)])
def test_partition_equals_oracle_and_invariants(case, subs, rotation):
    inst = _instances()[case]
    t = _tour(inst, 1 + case)
    parts = dc.partition_routes(inst, t, subs, rotation)
    oparts = oracle.partition_routes(inst.x, inst.y, inst.demand, t, subs, rotation)
    assert len(parts) == len(oparts)
    for (r, st), (orr, ost) in zip(parts, oparts):
        np.testing.assert_array_equal(r, orr)
        np.testing.assert_array_equal(st, ost)
    routes = _angles(inst, t)
    R = len(routes)
    G = max(1, min(subs if subs > 0 else min(inst.n // 50, R // 2), R))
    assert len(parts) == G
    allc = np.concatenate([r[1:] for r, _ in parts])
    np.testing.assert_array_equal(np.sort(allc), np.arange(1, inst.n))
    total = inst.n - 1
    route_of = {}
    for i, (_, nodes) in enumerate(routes):
        for v in nodes:
            route_of[v] = i
    order = sorted(range(R), key=lambda i: (routes[i][0], i))
    pos = {r: q for q, r in enumerate(order)}
    for real, st in parts:
        assert real[0] == 0 and st[0] == 0 and st[-1] == 0
        # the sub tour is its routes in local numbering: each node once
        np.testing.assert_array_equal(np.sort(st[st > 0]), np.arange(1, len(real)))
        rs = sorted({route_of[v] for v in real[1:]})
        for i in rs:   # each route whole
            assert set(routes[i][1]) <= set(real[1:].tolist())
        qs = sorted(pos[i] for i in rs)
        # consecutive in the circular order
        gaps = [(qs[(j + 1) % len(qs)] - qs[j]) % R for j in range(len(qs))]
        assert len(qs) == 1 or sorted(gaps)[:-1] == [1] * (len(qs) - 1)
        # no run takes more than its share plus one route
        assert len(real) - 1 <= total / G + max(len(nodes) for _, nodes in routes)


def _subs_of(inst, t, subs, rotation):
    parts = dc.partition_routes(inst, t, subs, rotation)
    return parts, [dc.sub_problem(inst, r) for r, _ in parts]


@pytest.mark.parametrize("case", range(4))
def test_merge_equals_oracle_and_keeps_the_shorter(case):
    inst = _instances()[case]
    t = _tour(inst, 3)
    parts, subs = _subs_of(inst, t, 0, 0)
    bests = []
    for s, ((real, st), sub) in enumerate(zip(parts, subs)):
        if s % 3 == 2:
            bests.append(np.zeros(0, np.int32))          # not solved
        elif s % 3 == 1:
            bests.append(_tour(sub, 11 + s))             # random split: usually longer
        else:
            bests.append(oracle.Problem(sub.dist, sub.demand, sub.capacity).nn_tour()[0])
    m, L, rep = dc.merge_sub_tours(inst, parts, bests)
    om, oL, orep = oracle.merge_sub_tours(inst.dist, parts, bests)
    np.testing.assert_array_equal(m, om)
    assert L == oL
    np.testing.assert_array_equal(rep, orep)
    assert inst.check_solution(m)
    assert L == np.float32(oracle.tour_length(inst.dist, m))
    gain = 0.0
    for s, ((real, st), b) in enumerate(zip(parts, bests)):
        lo = oracle.tour_length(inst.dist, real[st])
        if len(b) == 0:
            assert rep[s] == 0
            continue
        lb = oracle.tour_length(inst.dist, real[b])
        assert rep[s] == (lb < lo)
        gain += (lo - lb) if rep[s] else 0.0
    # integer distances: exact
    assert L == oracle.tour_length(inst.dist, t) - gain


def test_merge_brute_force_tiny():
    """Two 3-customer sub-problems: the best of all single-route orders
    replaces its routes exactly when strictly shorter."""
    from itertools import permutations
    inst = I.uniform_100(9).subset(7)
    t = np.array([0, 1, 2, 3, 0, 4, 5, 6, 0], np.int32)
    if not inst.check_solution(t):
        pytest.skip("capacity does not admit the fixed tour")
    parts, subs = _subs_of(inst, t, 2, 0)
    bests = []
    for sub in subs:
        cands = [np.array((0,) + p + (0,), np.int32) for p in permutations(range(1, sub.n))]
        cands = [c for c in cands if sub.check_solution(c)]
        bests.append(min(cands, key=lambda c: oracle.tour_length(sub.dist, c)))
    m, L, rep = dc.merge_sub_tours(inst, parts, bests)
    om, oL, _ = oracle.merge_sub_tours(inst.dist, parts, bests)
    np.testing.assert_array_equal(m, om)
    assert L == oL <= oracle.tour_length(inst.dist, t)


def test_decomposition_rejects_bad_input():
    inst = I.uniform_100(5)
    with pytest.raises(ca.AcoError):
        dc.partition_routes(inst, np.array([1, 2, 0], np.int32))
    with pytest.raises(ca.AcoError):
        dc.route_centers(inst, np.array([0, 5, 500, 0], np.int32))
    with pytest.raises(ca.AcoError):
        dc.sub_problem(inst, np.array([3, 0, 5], np.int32))      # depot not first
    with pytest.raises(ca.AcoError):
        dc.sub_problem(inst, np.array([0, 5, 5], np.int32))      # repeated node
    t = oracle.Problem(inst.dist, inst.demand, inst.capacity).nn_tour()[0]
    parts = dc.partition_routes(inst, t, 2)
    r0, t0 = parts[0]
    for bad in (np.r_[r0[:-1], inst.n], np.r_[r0[:-1], -1], np.r_[5, r0[1:]]):
        with pytest.raises(ca.AcoError):                        # node list entry out of range
            dc.merge_sub_tours(inst, [(bad.astype(np.int32), t0)] + parts[1:], [[]] * len(parts))


def test_pheromone_from_sub_round_trip_and_locality():
    """orc_pheromone_from_sub is the inverse of the sub-problem gather: the
    master's own trail gathered through real_nodes and written back leaves
    tau and total_info unchanged; a changed sub trail lands exactly on the
    sub-problem's ordered pairs (i != j) with total_info = tau * heur."""
    inst = I.clustered(4, customers=120)
    parts = oracle.partition_routes(inst.x, inst.y, inst.demand, _tour(inst), 3, 0)
    real = parts[1][0]
    heur = oracle.heuristic(inst.dist, 2.0)
    r = np.random.default_rng(1)
    tau = r.uniform(1e-4, 1e-2, (inst.n, inst.n)).astype(np.float32)
    ti = (tau * heur).astype(np.float32)
    t2, i2 = tau.copy(), ti.copy()
    oracle.pheromone_from_sub(1.0, heur, t2, i2, real, tau[np.ix_(real, real)])
    np.testing.assert_array_equal(t2, tau)
    np.testing.assert_array_equal(i2, ti)
    sub = r.uniform(1.0, 2.0, (len(real), len(real))).astype(np.float32)
    oracle.pheromone_from_sub(1.0, heur, t2, i2, real, sub)
    mask = np.zeros_like(tau, bool)
    mask[np.ix_(real, real)] = True
    mask[real, real] = False
    np.testing.assert_array_equal(t2[~mask], tau[~mask])
    np.testing.assert_array_equal(t2[np.ix_(real, real)][~np.eye(len(real), dtype=bool)],
                                  sub[~np.eye(len(real), dtype=bool)])
    np.testing.assert_array_equal(i2[mask], (t2 * heur)[mask])


def test_merge_keeps_route_length_limit_and_service_time():
    """CMT6 (float distances, DISTANCE and SERVICE_TIME): sub-problems carry
    the limit and the service time; tours valid in the sub-problem merge
    into a tour valid in the master, equal to the oracle's merge."""
    import os
    from helpers import VRPLIB
    inst = ca.Instance.from_vrplib(os.path.join(VRPLIB, "CMT6.vrp"))
    P = oracle.Problem(inst.dist, inst.demand, inst.capacity, inst.max_dist, inst.service_time)
    t = P.nn_tour()[0]
    parts = dc.partition_routes(inst, t, 3, 2)
    bests = []
    for real, st in parts:
        sub = dc.sub_problem(inst, real)
        assert sub.max_dist == inst.max_dist and sub.service_time == inst.service_time
        sp = oracle.Problem(sub.dist, sub.demand, sub.capacity, sub.max_dist, sub.service_time)
        b = sp.nn_tour()[0]
        assert sp.check(b) and sp.check(st)
        bests.append(b)
    m, L, rep = dc.merge_sub_tours(inst, parts, bests)
    om, oL, orep = oracle.merge_sub_tours(inst.dist, parts, bests)
    np.testing.assert_array_equal(m, om)
    assert L == oL
    np.testing.assert_array_equal(rep, orep)
    assert inst.check_solution(m)
