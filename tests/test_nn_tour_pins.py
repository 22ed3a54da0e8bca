"""Pins of the tau0 = 1/(n * L_nn) rule's greedy tour, on the CPU.

The rule has no reference counterpart (the reference derives tau0 from a
first colony, g_aco.cpp:79-90), so the oracle's orc_nn_tour and the
product's aco_nn_tour are pinned by what the rule's statement fixes, each
independently of the other:

* hand-worked instances whose tour and L_nn are written out literally here
  (collinear customers, capacity forcing the skip of a nearer customer,
  capacity forcing one customer per route, a route-length limit with
  service time);
* closed forms: customers on a ray from the depot give L = 2 * x_max; one
  customer per route gives L = sum 2 d(0, i), visited by (d(0, i), index);
* a check of the OUTPUT tour on random instances with many ties: at every
  step the next node is the nearest feasible unvisited customer, ties to
  the lowest index, and the depot comes only when no customer fits.

A wrong tie order, a capacity or route-length misreading, or a dropped edge
fails one of them.
"""
import math

import numpy as np
import pytest

import oracle
import cvrp_aco_ocl_amd as ca


def _line(xs, dem, Q, D=math.inf, svc=0.0):
    """Depot at x = 0, customers at xs on the x axis (nint distances)."""
    x = np.array([0.0] + list(xs))
    y = np.zeros_like(x)
    return ca.Instance.from_coords(x, y, [0] + list(dem), Q, ca.DIST_NINT, max_dist=D,
                                   service_time=svc)


def _both(inst):
    """(tour, L) from the oracle and from the product library."""
    P = oracle.Problem(inst.dist, inst.demand, inst.capacity, inst.max_dist, inst.service_time)
    ot, ol = P.nn_tour()
    pt, pl = inst.nn_tour()
    return (list(ot), ol), (list(pt), pl)


HAND_WORKED = [
    # customers at x = 3, 1, 4, 2, 5 (indices 1..5), ample capacity:
    # visited by position, then back from x = 5
    dict(xs=[3, 1, 4, 2, 5], dem=[1] * 5, Q=100, tour=[0, 2, 4, 1, 3, 5, 0], L=10.0),
    # Q = 7, demands 3, 3, 5, 1 at x = 1..4: after 1 and 2 (load 6) customer 3
    # does not fit but customer 4 does, so it is skipped, not a stopping point
    dict(xs=[1, 2, 3, 4], dem=[3, 3, 5, 1], Q=7, tour=[0, 1, 2, 4, 0, 3, 0], L=14.0),
    # Q = 7, demands 3 each: two customers per route
    dict(xs=[1, 2, 3, 4], dem=[3, 3, 3, 3], Q=7, tour=[0, 1, 2, 0, 3, 4, 0], L=12.0),
    # length limit 17, no service time: 0-1-2-3-8 needs 3 + 5 + 8 = 16 <= 17
    dict(xs=[1, 2, 3, 8], dem=[1] * 4, Q=100, D=17.0, tour=[0, 1, 2, 3, 4, 0], L=16.0),
    # the same with service time 1: from 3, rdist 6 + (5 + 1) + 8 = 20 > 17,
    # so back to the depot; then 0 + (8 + 1) + 8 = 17 <= 17
    dict(xs=[1, 2, 3, 8], dem=[1] * 4, Q=100, D=17.0, svc=1.0, tour=[0, 1, 2, 3, 0, 4, 0],
         L=22.0),
    # tie at the depot: x = -1 (index 1) and +1 (index 2) both at distance 1;
    # the lowest index goes first, then +1 (d 2) before +3 (d 4)
    dict(xs=[-1, 1, 3], dem=[1] * 3, Q=100, tour=[0, 1, 2, 3, 0], L=1.0 + 2.0 + 2.0 + 3.0),
]


@pytest.mark.parametrize("case", HAND_WORKED, ids=lambda c: f"{c['xs']}-Q{c['Q']}")
def test_hand_worked_tours(case):
    inst = _line(case["xs"], case["dem"], case["Q"], case.get("D", math.inf), case.get("svc", 0.0))
    for tour, L in _both(inst):
        assert tour == case["tour"]
        assert L == case["L"]


@pytest.mark.parametrize("seed", range(4))
def test_ray_closed_form(seed):
    """Customers at the distinct positions 1..N of a ray, ample capacity:
    visited outwards, L = 2 N."""
    r = np.random.default_rng(seed)
    N = int(r.integers(5, 40))
    pos = r.permutation(np.arange(1, N + 1))
    inst = _line(pos.tolist(), [1] * N, 10 * N)
    expect = [0] + [int(i) + 1 for i in np.argsort(pos)] + [0]
    for tour, L in _both(inst):
        assert tour == expect
        assert L == 2.0 * N


@pytest.mark.parametrize("seed", range(4))
def test_one_customer_per_route_closed_form(seed):
    """Every demand above Q/2: each route serves one customer, so
    L = sum 2 d(0, i) and customers come in (d(0, i), index) order."""
    r = np.random.default_rng(100 + seed)
    n = 30
    x = r.integers(0, 12, n).astype(float)   # small grid: many equal d(0, i)
    y = r.integers(0, 12, n).astype(float)
    x[0] = y[0] = 6.0
    dem = np.full(n, 6, np.int32)
    dem[0] = 0
    inst = ca.Instance.from_coords(x, y, dem, 10, ca.DIST_NINT)
    d0 = inst.dist[0]
    order = sorted(range(1, n), key=lambda i: (d0[i], i))
    expect = [0]
    for i in order:
        expect += [i, 0]
    total = float(np.sum(2 * d0[1:].astype(np.float64)))
    for tour, L in _both(inst):
        assert tour == expect
        assert L == total


def _circle_instance():
    """12 customers at integer points at distance 5 from the depot, indices
    in a scrambled order."""
    pts = [(5, 0), (0, 5), (-5, 0), (0, -5), (3, 4), (4, 3), (-3, 4), (-4, 3), (3, -4), (4, -3),
           (-3, -4), (-4, -3)]
    perm = [7, 2, 11, 0, 5, 9, 1, 4, 10, 3, 8, 6]
    x = [0.0] + [float(pts[p][0]) for p in perm]
    y = [0.0] + [float(pts[p][1]) for p in perm]
    return x, y


def test_equidistant_ties_single_routes():
    x, y = _circle_instance()
    inst = ca.Instance.from_coords(x, y, [0] + [6] * 12, 10, ca.DIST_NINT)
    expect = [0]
    for i in range(1, 13):
        expect += [i, 0]
    for tour, L in _both(inst):
        assert tour == expect          # every d(0, i) = 5: index order
        assert L == 120.0


def _assert_greedy_output(inst, tour, L):
    """Every step of `tour` is the rule's choice, checked on the output."""
    n, d = inst.n, inst.dist.astype(np.float32)
    Q, D, svc = inst.capacity, np.float32(inst.max_dist), np.float32(inst.service_time)
    assert tour[0] == 0 and tour[-1] == 0
    vis = np.zeros(n, bool)
    vis[0] = True
    cur, load, rdist = 0, 0, np.float32(0.0)
    length = np.float32(0.0)
    for w in tour[1:-1]:
        feas = []
        for i in range(1, n):
            if vis[i] or load + inst.demand[i] > Q:
                continue
            t = np.float32(rdist + np.float32(d[cur, i] + svc))
            t = np.float32(t + d[i, 0])
            if t <= D:
                feas.append(i)
        if w == 0:
            assert not feas, f"returned to the depot from {cur} while {feas} fit"
            assert cur != 0
            load, rdist = 0, np.float32(0.0)
        else:
            assert w in feas, f"{w} is not a feasible unvisited customer at {cur}"
            best = min(feas, key=lambda i: (d[cur, i], i))
            assert w == best, f"from {cur}: took {w} (d {d[cur, w]}), nearest is {best}"
            vis[w] = True
            load += int(inst.demand[w])
            rdist = np.float32(rdist + np.float32(d[cur, w] + svc))
        length = np.float32(length + d[cur, w])
        cur = w
    length = np.float32(length + d[cur, 0])
    assert vis.all()
    assert L == float(length)
    assert inst.check_solution(np.array(tour, np.int32))


def test_equidistant_ties_one_route():
    x, y = _circle_instance()
    inst = ca.Instance.from_coords(x, y, [0] + [1] * 12, 100, ca.DIST_NINT)
    (ot, ol), (pt, pl) = _both(inst)
    assert ot[1] == 1                  # all at distance 5 from the depot: lowest index
    _assert_greedy_output(inst, ot, ol)
    _assert_greedy_output(inst, pt, pl)


@pytest.mark.parametrize("seed", range(12))
def test_random_instances_every_step_is_the_rule(seed):
    r = np.random.default_rng(seed)
    n = int(r.integers(8, 60))
    grid = int(r.choice([6, 15, 100]))          # small grids give many ties
    x = r.integers(0, grid, n).astype(float)
    y = r.integers(0, grid, n).astype(float)
    dem = r.integers(1, 10, n).astype(np.int32)
    dem[0] = 0
    Q = int(max(dem.max(), r.integers(10, 40)))
    mode = ca.DIST_NINT if seed % 3 else ca.DIST_FLOAT
    D, svc = math.inf, 0.0
    if seed % 4 == 1:
        svc = float(r.integers(0, 3))
        far = max(math.hypot(x[i] - x[0], y[i] - y[0]) for i in range(1, n))
        D = float(2 * far + svc + r.integers(2, 3 * grid))
    inst = ca.Instance.from_coords(x, y, dem, Q, mode, max_dist=D, service_time=svc)
    (ot, ol), (pt, pl) = _both(inst)
    _assert_greedy_output(inst, ot, ol)
    _assert_greedy_output(inst, pt, pl)
    assert ot == pt and ol == pl


def test_unservable_customer_is_an_error():
    inst = _line([1, 30], [1, 1], 10, D=40.0)   # 0 -> 30 -> 0 needs 60 > 40
    with pytest.raises(ca.AcoError):
        inst.nn_tour()
    P = oracle.Problem(inst.dist, inst.demand, inst.capacity, inst.max_dist, inst.service_time)
    t = np.empty(2 * inst.n, np.int32)
    L = oracle.C.c_float()
    assert oracle.lib.orc_nn_tour(oracle.C.byref(P.c), t, oracle.C.byref(L)) == -1
