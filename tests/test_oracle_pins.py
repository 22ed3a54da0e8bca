"""Pins of the CPU oracle (no GPU): known-answer vectors, the reference's own
CPU build (oracle/_ref, where the reference sources were present at build
time), brute force on small inputs, closed forms and special cases."""
import ctypes as C
import itertools

import numpy as np
import pytest

import oracle
import cvrp_aco_ocl_amd as ca
from cvrp_aco_ocl_amd import instances as I

from helpers import oracle_construct_all, oracle_problem

REF = oracle.load_ref()
needs_ref = pytest.mark.skipif(REF is None, reason="reference CPU build (oracle/_ref) not built")


# ---------------------------------------------------------------- RNG ----

def test_park_miller_known_answer():
    # Park & Miller (CACM 31(10), 1988): z_10000 = 1043618065 from z_0 = 1
    _, z = oracle.ran01_stream(1, 10000)
    assert z == 1043618065


def test_ran01_closed_form():
    vals, _ = oracle.ran01_stream(42, 500)
    z = 42
    for v in vals:
        z = (16807 * z) % 2147483647
        assert v == np.float32(np.float32(z) * np.float32(2.0 ** -31))


@needs_ref
def test_ran01_matches_reference_build():
    s = C.c_int(987654)
    out = np.empty(2000, np.float32)
    REF.ref_ran01(C.byref(s), 2000, out)
    vals, z = oracle.ran01_stream(987654, 2000)
    np.testing.assert_array_equal(out, vals)
    assert z == s.value


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    assert oracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert oracle.philox([0xffffffff] * 4, [0xffffffff] * 2) == \
        [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert oracle.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                         [0xa4093822, 0x299f31d0]) == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_pm_seeds_in_range_and_distinct():
    s = {oracle.pm_seed(5, a, l) for a in range(64) for l in range(64)}
    assert len(s) == 64 * 64
    assert all(1 <= v <= 2147483646 for v in s)


# ------------------------------------------------------- reductions ----

@pytest.mark.parametrize("N", list(range(1, 80)) + [101, 127, 128, 129, 481, 1001])
def test_tree_rank_equals_literal_reduction(N):
    """argmax by (value, -rank) == the reference's reduction loop, on inputs
    with many ties (kernel.cl:340-366)."""
    rng = np.random.default_rng(N)
    rank = oracle.tree_rank(N)
    assert sorted(rank) == list(range(N))
    for _ in range(40):
        vals = rng.integers(0, 3, N).astype(np.float32)
        a = oracle.tree_argmax(vals)
        b = max(range(N), key=lambda i: (vals[i], -rank[i]))
        assert a == b


def test_phase0_min_reduction_tie_order():
    """best_solution_phase_0/1 (kernel.cl:594-665) == min by (value,
    group, rank-in-group of the 64-wide halving tree)."""
    r64 = oracle.tree_rank(64)
    rng = np.random.default_rng(3)
    for m in (1, 5, 63, 64, 65, 100, 130, 480):
        for _ in range(30):
            vals = rng.integers(0, 3, m).astype(np.float32)
            a = oracle.minreduce_phase0(vals)
            b = min(range(m), key=lambda i: (vals[i], i // 64, r64[i % 64]))
            assert a == b


def test_elites_lowest_index_ties():
    lens = np.array([5, 3, 3, 9, 1, 3, 1], np.float32)
    e = oracle.elites(lens, 6)
    assert list(e) == [7, 4, 6, 1, 2, 5]
    e = oracle.elites(np.array([4, 2], np.float32), 6)  # fewer ants than ranks
    assert list(e) == [2, 1, 0, 0, 0, 0]


# ------------------------------------------------------- preprocessing ----

@needs_ref
def test_sort2_matches_reference_build():
    rng = np.random.default_rng(0)
    for n in (2, 3, 10, 57, 200):
        v = rng.integers(0, 20, n).astype(np.float32)
        v2 = np.arange(n, dtype=np.int32)
        a, a2 = v.copy(), v2.copy()
        b, b2 = v.copy(), v2.copy()
        REF.ref_sort2(a, a2, 0, n - 1)
        oracle.lib.orc_sort2(b, b2, 0, n - 1)
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(a2, b2)


@needs_ref
@pytest.mark.parametrize("seed", range(4))
def test_nn_lists_match_reference_build(seed):
    inst = I.uniform_100(seed)  # integer distances: many ties
    n = inst.n
    x = inst.x.astype(np.float32)
    y = inst.y.astype(np.float32)
    nn = np.empty(n * (n - 2), np.int32)
    REF.ref_nn_lists(n, x, y, inst.dist.ravel().ctypes.data, None, nn)
    np.testing.assert_array_equal(nn.reshape(n, n - 2), oracle.nn_lists(inst.dist, n - 2, 0))
    # the product's own builder (host side of the C ABI) too
    np.testing.assert_array_equal(nn.reshape(n, n - 2), inst.nn_lists(n - 2, 0))


def test_nn_lists_stable_tie_rule():
    inst = I.uniform_100(1)
    n = inst.n
    got = oracle.nn_lists(inst.dist, n - 2, 1)
    for i in range(n):
        keys = [(inst.dist[i, j], j) for j in range(1, n) if j != i]
        want = [j for _, j in sorted(keys)][: n - 2]
        assert list(got[i]) == want


def test_heuristic_closed_form_is_correctly_rounded_pow():
    d = np.arange(0, 5000, dtype=np.float32)
    y = (np.float32(1.0) / (d + np.float32(0.1))).astype(np.float32)
    want = (y.astype(np.float64) ** 2).astype(np.float32)  # y*y is exact in double
    dm = np.zeros((len(d), len(d)), np.float32)
    dm[0] = d
    got = oracle.heuristic(dm, 2.0)[0]
    np.testing.assert_array_equal(got, want)


# -------------------------------------------------------- construction ----

def _small_problem(seed, n=15, Q=None, k=0):
    r = np.random.default_rng(seed)
    x = r.integers(0, 50, n).astype(np.float64)
    y = r.integers(0, 50, n).astype(np.float64)
    dem = np.concatenate([[0], r.integers(1, 10, n - 1)]).astype(np.int32)
    inst = ca.Instance.from_coords(x, y, dem, Q or 25)
    return inst, oracle_problem(inst, k=k)


@pytest.mark.parametrize("selection,rng,k", [(0, 0, 0), (1, 1, 8), (1, 0, 8), (1, 1, 13)])
def test_construction_invariants(selection, rng, k):
    for seed in range(10):
        inst, P = _small_problem(seed, k=k)
        tau = np.random.default_rng(seed).uniform(1e-3, 1.0, (inst.n, inst.n)).astype(np.float32)
        ti = (tau * P.heur).astype(np.float32)
        pm = np.array([[oracle.pm_seed(seed, a, l) for l in range(64)] for a in range(8)], np.int32)
        tours, lens, _, _ = oracle_construct_all(P, ti, selection, rng, seed, 3, pm)
        for t, L in zip(tours, lens):
            assert P.check(t)
            assert t[0] == 0 and t[-1] == 0 and sorted(set(t[t > 0])) == list(range(1, inst.n))
            assert L == P.length(t)
            assert len(t) <= 2 * inst.n - 1


def _pm_lane_draws(seed_states, n):
    """draws r_i + 0.001 for nodes 0..n-1 from the 64 lane streams."""
    st = [int(s) for s in seed_states]
    out = np.zeros(n, np.float32)
    for lane in range(64):
        for i in range(lane, n, 64):
            st[lane] = (16807 * st[lane]) % 2147483647
            out[i] = np.float32(np.float32(np.float32(st[lane]) * np.float32(2.0 ** -31)) + np.float32(0.001))
    return out, st


def test_iroulette_equal_info_picks_largest_draw():
    """alpha = beta = 0 (total_info constant): I-Roulette reduces to the
    argmax of the masked draws, by the reduction's tie order."""
    inst, P = _small_problem(4, n=20, Q=1000)
    ti = np.ones((inst.n, inst.n), np.float32)
    pm = np.array([oracle.pm_seed(9, 0, l) for l in range(64)], np.int32)
    tour, L, _ = P.construct_ant(ti, None, 0, 0, 0, 0, 0, pm.copy())
    # replay: all customers fit in one route, so each step picks the
    # unvisited node with the largest draw
    st = pm.copy()
    visited = {0}
    cur_tour = [0]
    for _ in range(inst.n - 1):
        draws, st = _pm_lane_draws(st, inst.n)
        probs = np.array([draws[i] * np.float32(100.0) if i not in visited else 0.0
                          for i in range(inst.n)], np.float32)
        w = oracle.tree_argmax(probs)
        visited.add(w)
        cur_tour.append(w)
    assert list(tour[: inst.n]) == cur_tour


def test_single_feasible_candidate_is_forced():
    # customer 1 has demand Q, the others Q-... only one fits at a time:
    # every route holds exactly one customer
    x = np.array([0, 1, 2, 3, 4, 5], np.float64)
    y = np.zeros(6)
    inst = ca.Instance.from_coords(x, y, [0, 10, 10, 10, 10, 10], 10)
    P = oracle_problem(inst, k=3)
    ti = np.ones((6, 6), np.float32)
    for sel, rng in ((0, 0), (1, 1)):
        tiknn = oracle.gather_knn(P, ti)
        pm = np.array([oracle.pm_seed(1, 0, l) for l in range(64)], np.int32)
        t, L, _ = P.construct_ant(ti, tiknn, sel, rng, 1, 0, 0, pm)
        assert len(t) == 11 and all(t[::2] == 0)
        assert L == 2 * (1 + 2 + 3 + 4 + 5)


def test_knn_full_row_roulette_with_exact_sums():
    """k >= n-2: the candidate row is the whole customer row.  With
    integer total_info the wheel sums are exact, so the pick equals a plain
    cumulative-sum roulette in candidate order."""
    inst, P = _small_problem(7, n=12, Q=1000, k=10)
    rng = np.random.default_rng(0)
    ti = rng.integers(1, 50, (inst.n, inst.n)).astype(np.float32)
    tiknn = oracle.gather_knn(P, ti)
    seed, it, ant = 11, 2, 0
    t, L, fb = P.construct_ant(ti, tiknn, 1, 1, seed, it, ant, np.zeros(64, np.int32))
    visited = {0}
    cur = 0
    for step in range(1, inst.n):
        cand = P.nnk[cur]
        p = np.array([ti[cur, c] if c not in visited else 0 for c in cand], np.float64)
        u = np.float32(oracle.lib.orc_philox_u01(seed, it, ant, step))
        r = float(np.float32(u * np.float32(p.sum())))
        cs = np.cumsum(p)
        w = cand[int(np.argmax(cs > r))] if (cs > r).any() else cand[np.nonzero(p)[0][-1]]
        assert t[step] == w
        visited.add(w)
        cur = w
    assert fb == 0


def test_zero_mass_at_depot_takes_greedy_pick():
    inst, P = _small_problem(2, n=10, Q=1000)
    ti = np.zeros((inst.n, inst.n), np.float32)
    ti[0, 7] = 0.0
    pm = np.array([oracle.pm_seed(1, 0, l) for l in range(64)], np.int32)
    t, L, fb = P.construct_ant(ti, None, 0, 0, 0, 0, 0, pm)
    # all mass zero: at the depot the greedy pick takes the lowest feasible
    # index; at a customer zero mass sends the ant home (kernel.cl:461-465)
    want = [0]
    for c in range(1, inst.n):
        want += [c, 0]
    assert list(t) == want
    assert fb == inst.n - 1


# -------------------------------------------------------- local search ----

def _tours(inst, P, count, seed):
    tau, ti = oracle.pheromone_init(P, 0.5)
    pm = np.array([[oracle.pm_seed(seed, a, l) for l in range(64)] for a in range(count)], np.int32)
    tours, _, _, _ = oracle_construct_all(P, ti, 0, 0, 0, 0, pm)
    return tours


def _ref_args(inst):
    return [inst.n, inst.x.astype(np.float32), inst.y.astype(np.float32), inst.demand,
            inst.capacity, inst.max_dist, inst.service_time, inst.dist.ravel().ctypes.data]


@needs_ref
@pytest.mark.parametrize("which", ["u100", "cmt6"])
def test_two_opt_and_swap_match_reference_localsearch(which):
    import os
    from helpers import VRPLIB
    inst = I.uniform_100(8) if which == "u100" else ca.Instance.from_vrplib(os.path.join(VRPLIB, "CMT6.vrp"))
    P = oracle_problem(inst)
    moves = 0
    for a, tour in enumerate(_tours(inst, P, 16, 3)):
        # random scan order 2-opt (LocalSearch::do_local_search(ant))
        t_ref = tour.copy()
        s = C.c_int(500 + a)
        REF.ref_two_opt(*_ref_args(inst), C.byref(s), t_ref, len(t_ref))
        t_or, mv, _, s2 = P.two_opt(tour, random_order=True, seed=500 + a)
        np.testing.assert_array_equal(t_ref, t_or)
        assert s.value == s2
        # 2-opt then swap (LocalSearch::do_local_search())
        t_ref = tour.copy()
        s = C.c_int(900 + a)
        L_ref = REF.ref_two_opt_swap(*_ref_args(inst), C.byref(s), t_ref, len(t_ref))
        t1, _, _, _ = P.two_opt(tour, random_order=True, seed=900 + a)
        t2, mv2, _ = P.swap(t1)
        np.testing.assert_array_equal(t_ref, t2)
        assert np.float32(L_ref) == np.float32(P.length(t2))
        moves += mv + mv2
    assert moves > 0


def test_park_miller_jump_ahead():
    """orc_pm_jump(z, t) = the state after t draws (16807^t z mod 2^31-1)."""
    for z in (1, 12345, 2147483646, 987654321):
        s = C.c_int32(z)
        for t in range(1, 300):
            oracle.lib.orc_ran01(s)
            assert oracle.pm_jump(z, t) == s.value
    assert oracle.pm_jump(1, 10000) == 1043618065      # Park & Miller's known answer
    assert oracle.pm_jump(5, 2147483646) == 5          # the generator's full period


@needs_ref
@pytest.mark.parametrize("which", ["u100", "cmt6"])
def test_random_order_two_opt_stream_over_ants(which):
    """The reference's CPU 2-opt (LocalSearch::do_local_search(ant)) draws one
    value per route position: size - 1 per ant.  Called ant after ant on one
    stream, it equals the oracle's random-order 2-opt carrying the seed, and
    equals each ant started at the jump-ahead of the draws before it (how
    the device runs the ants in parallel)."""
    import os
    from helpers import VRPLIB
    inst = I.uniform_100(12) if which == "u100" else ca.Instance.from_vrplib(os.path.join(VRPLIB, "CMT6.vrp"))
    P = oracle_problem(inst)
    tours = _tours(inst, P, 12, 5)
    s_ref = C.c_int(4242)
    s_or = 4242
    before = 0
    for a, tour in enumerate(tours):
        t_ref = tour.copy()
        start = s_ref.value
        REF.ref_two_opt(*_ref_args(inst), C.byref(s_ref), t_ref, len(t_ref))
        assert s_ref.value == oracle.pm_jump(start, len(tour) - 1), a
        t_or, L_or, _, s_or = P.local_search_seeded(tour, 8, s_or)
        np.testing.assert_array_equal(t_ref, t_or, err_msg=f"ant {a}")
        assert s_or == s_ref.value
        t_jump, _, _, _ = P.local_search_seeded(tour, 8, oracle.pm_jump(4242, before))
        np.testing.assert_array_equal(t_jump, t_or)
        before += len(tour) - 1


@needs_ref
def test_insertion_and_exchange_apply_match_reference():
    rng = np.random.default_rng(1)
    for _ in range(200):
        t = rng.permutation(20).astype(np.int32)
        i, j = rng.choice(np.arange(1, 19), 2, replace=False)
        # relocate: oracle apply semantics restated here from move.cpp:32-51
        r = t.copy()
        REF.ref_apply_move(0, r, len(r), int(i), int(j), 0.0, 0.0)
        lst = list(t)
        v = lst.pop(i)
        lst.insert(j, v)
        assert list(r) == lst
        r = t.copy()
        REF.ref_apply_move(1, r, len(r), int(i), int(j), 0.0, 0.0)
        s = t.copy()
        s[i], s[j] = s[j], s[i]
        np.testing.assert_array_equal(r, s)


def _relocate_brute(inst, tour):
    """First-improvement relocate with deltas measured as cost(apply) -
    cost, the reference's InsertionMove::apply, and the load/route-length
    rules; no delta formula involved."""
    t = list(tour)
    dist, dem, Q = inst.dist, inst.demand, inst.capacity

    def routes_of(tt):
        rid, r = [0] * len(tt), 0
        for x in range(1, len(tt)):
            rid[x] = r
            if tt[x] == 0:
                r += 1
        return rid

    def cost(tt):
        return sum(int(dist[a, b]) for a, b in zip(tt[:-1], tt[1:]))

    def route_loads(tt):
        loads, cur = [], 0
        for x in tt[1:]:
            cur += dem[x]
            if x == 0:
                loads.append(cur)
                cur = 0
        return loads

    i = 1
    while i < len(t) - 1:
        if t[i] == 0:
            i += 1
            continue
        rid = routes_of(t)
        count = sum(1 for x in range(1, len(t) - 1) if t[x] != 0 and rid[x] == rid[i])
        if count < 2:
            i += 1
            continue
        base = cost(t)
        found = False
        for j in range(1, len(t) - 1):
            if j == i or t[j] == 0:
                continue
            lst = t.copy()
            v = lst.pop(i)
            lst.insert(j, v)
            delta = cost(lst) - base
            if not delta < -0.001:
                continue
            if rid[i] != rid[j]:
                if route_loads(t)[rid[j]] + dem[t[i]] > Q:
                    continue
            t = lst
            found = True
            break
        if not found:
            i += 1
    return np.array(t, np.int32)


def test_relocate_equals_brute_force_first_improvement():
    for seed in range(8):
        inst, P = _small_problem(seed, n=14, Q=20)
        for tour in _tours(inst, P, 4, seed):
            got, mv, _ = P.relocate(tour)
            want = _relocate_brute(inst, tour)
            np.testing.assert_array_equal(got, want)
            assert P.check(got)


def _swap_brute(inst, tour):
    """LocalSearch::swap's scan (localSearch.cpp:339-407) with every delta
    measured as cost(after) - cost(before) and loads recounted; no delta
    formula involved (unlimited route length)."""
    t = list(tour)
    dist, dem, Q = inst.dist, inst.demand, inst.capacity

    def cost(tt):
        return sum(int(dist[a, b]) for a, b in zip(tt[:-1], tt[1:]))

    def rid(tt):
        out, r = [0] * len(tt), 0
        for x in range(1, len(tt)):
            out[x] = r
            if tt[x] == 0:
                r += 1
        return out

    def loads(tt):
        out, cur = [], 0
        for x in tt[1:]:
            cur += dem[x]
            if x == 0:
                out.append(cur)
                cur = 0
        return out

    i = 1
    while i < len(t):
        if t[i] != 0:
            base = cost(t)
            r = rid(t)
            for j in range(i + 1, len(t)):
                if t[j] == 0:
                    continue
                s2 = t.copy()
                s2[i], s2[j] = s2[j], s2[i]
                if not cost(s2) - base < -0.001:
                    continue
                if r[i] != r[j] and max(loads(s2)) > Q:
                    continue
                t = s2
                i -= 1
                break
        i += 1
    return np.array(t, np.int32)


def test_swap_equals_brute_force_first_improvement():
    for seed in range(8):
        inst, P = _small_problem(seed, n=14, Q=20)
        for tour in _tours(inst, P, 4, seed):
            got, mv, _ = P.swap(tour)
            np.testing.assert_array_equal(got, _swap_brute(inst, tour))
            assert P.check(got)


def test_two_opt_moves_decrease_cost_exactly():
    inst = I.uniform_100(4)
    P = oracle_problem(inst)
    for tour in _tours(inst, P, 10, 1):
        before = P.length(tour)
        t, mv, _, _ = P.two_opt(tour)
        after = P.length(t)
        assert after <= before - mv  # integer distances: each move gains >= 1
        assert P.check(t)


# ----------------------------------------------------------- pheromone ----

def test_pheromone_update_closed_form():
    inst = I.uniform_100(6)
    P = oracle_problem(inst)
    rng = np.random.default_rng(1)
    tau0 = rng.uniform(1e-4, 1e-2, (inst.n, inst.n)).astype(np.float32)
    tours = _tours(inst, P, 6, 4)
    lens = np.array([P.length(t) for t in tours], np.float32)
    tau, ti = oracle.pheromone_update(P, tau0, tours, lens)
    want = 0.9 * tau0.astype(np.float64)
    for g, t in enumerate(tours):
        for a, b in zip(t[:-1], t[1:]):
            want[a, b] += (6 - g) / float(lens[g])
    np.testing.assert_allclose(tau, want, rtol=1e-6)
    np.testing.assert_array_equal(ti, (tau * P.heur).astype(np.float32))
    # directional deposit: reversed edges untouched unless travelled
    a, b = tours[0][1], tours[0][2]
    if not any(((t[:-1] == b) & (t[1:] == a)).any() for t in tours):
        assert tau[b, a] == np.float32(np.float32(0.9) * tau0[b, a])


# ---------------------------------------------------------- solutions ----

@needs_ref
def test_check_solution_matches_reference_build():
    import os
    from helpers import VRPLIB
    inst = ca.Instance.from_vrplib(os.path.join(VRPLIB, "CMT6.vrp"))
    P = oracle_problem(inst)
    tours = _tours(inst, P, 4, 9)
    rng = np.random.default_rng(0)
    cases = []
    for t in tours:
        cases.append(t)
        bad = t.copy()
        bad[3] = bad[5]  # duplicate
        cases.append(bad)
        cases.append(np.concatenate([[0], rng.permutation(np.arange(1, inst.n)), [0]]).astype(np.int32))
    for t in cases:
        ln = C.c_float()
        want = REF.ref_check_solution(*_ref_args(inst), t, len(t), C.byref(ln))
        assert bool(want) == P.check(t) == inst.check_solution(t)


def test_colony_records_are_monotone_and_consistent():
    inst = I.uniform_100(2)
    P = oracle_problem(inst, k=20)
    col = oracle.Colony(P, 20, selection=1, rng=1, ls_flags=7, seed=4)
    col.init()
    for _ in range(8):
        col.iteration()
    rec = col.bsf()
    assert [r[0] for r in rec] == sorted(r[0] for r in rec)
    assert all(b[1] < a[1] for a, b in zip(rec, rec[1:]))
    t, L = col.best()
    assert P.check(t) and L == P.length(t) == rec[-1][1]
