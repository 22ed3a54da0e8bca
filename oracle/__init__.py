"""CPU oracle for the CVRP ant-colony hot path -- TEST INFRASTRUCTURE ONLY.

ctypes binding of ``oracle/liboracle.so`` (plain C restatement of the
reference's per-iteration semantics, see ``oracle/oracle.h``) and of the
reference CPU build ``oracle/_ref/libcvrpref.so`` when it exists.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libcvrpref.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


class _Problem(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("dist", C.c_void_p), ("dem", C.c_void_p), ("Q", C.c_int),
        ("D", C.c_float), ("svc", C.c_float), ("heur", C.c_void_p), ("nnk", C.c_void_p),
        ("k", C.c_int), ("nnls", C.c_void_p), ("nn_ls", C.c_int),
    ]


class _DrawCfg(C.Structure):
    _fields_ = [("selection", C.c_int), ("rng", C.c_int), ("seed", C.c_uint64),
                ("iteration", C.c_uint32)]


class _Params(C.Structure):
    _fields_ = [
        ("m", C.c_int), ("alpha", C.c_float), ("beta", C.c_float), ("rho", C.c_float),
        ("ras", C.c_int), ("selection", C.c_int), ("rng", C.c_int), ("ls_flags", C.c_int),
        ("tau0_rule", C.c_int), ("seed", C.c_uint64), ("threads", C.c_int),
    ]


class BsfRecord(C.Structure):
    _fields_ = [("iter", C.c_int), ("length", C.c_float), ("time", C.c_float)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "orc_ran01": (C.c_float, [C.POINTER(C.c_int32)]),
        "orc_splitmix64": (C.c_uint64, [C.c_uint64]),
        "orc_pm_seed": (C.c_int32, [C.c_uint64, C.c_int, C.c_int]),
        "orc_ls_seed": (C.c_int32, [C.c_uint64]),
        "orc_pm_jump": (C.c_int32, [C.c_int32, C.c_uint64]),
        "orc_philox4x32_10": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                     C.POINTER(C.c_uint32)]),
        "orc_philox_u01": (C.c_float, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]),
        "orc_distances": (None, [C.c_int, _f64p, _f64p, C.c_int, _f32p]),
        "orc_sort2": (None, [_f32p, _i32p, C.c_int, C.c_int]),
        "orc_nn_lists": (None, [C.c_int, _f32p, C.c_int, C.c_int, _i32p]),
        "orc_heuristic": (None, [C.c_int, _f32p, C.c_float, C.c_void_p, _f32p]),
        "orc_tree_rank": (None, [C.c_int, _i32p]),
        "orc_tree_argmax": (C.c_int, [C.c_int, _f32p]),
        "orc_minreduce_phase0": (C.c_int, [C.c_int, C.c_int, _f32p]),
        "orc_construct_ant": (C.c_int, [C.POINTER(_Problem), _f32p, C.c_void_p,
                                        C.POINTER(_DrawCfg), C.c_int, _i32p, _i32p,
                                        C.POINTER(C.c_float), C.POINTER(C.c_int)]),
        "orc_iteration_best": (C.c_int, [C.c_int, _f32p, C.c_float, C.POINTER(C.c_int)]),
        "orc_elites": (None, [C.c_int, C.c_int, _f32p, _i32p]),
        "orc_pheromone_init": (None, [C.c_int, C.c_float, _f32p, C.c_float, _f32p, _f32p]),
        "orc_gather_knn": (None, [C.c_int, C.c_int, _i32p, _f32p, _f32p]),
        "orc_two_opt": (C.c_int, [C.POINTER(_Problem), _i32p, C.c_int, C.c_int,
                                  C.POINTER(C.c_int32), C.POINTER(C.c_longlong)]),
        "orc_swap": (C.c_int, [C.POINTER(_Problem), _i32p, C.c_int, C.POINTER(C.c_longlong)]),
        "orc_relocate": (C.c_int, [C.POINTER(_Problem), _i32p, C.c_int,
                                   C.POINTER(C.c_longlong)]),
        "orc_local_search": (C.c_float, [C.POINTER(_Problem), C.c_int, _i32p, C.c_int,
                                         C.POINTER(C.c_longlong)]),
        "orc_local_search_seeded": (C.c_float, [C.POINTER(_Problem), C.c_int, _i32p, C.c_int,
                                                C.POINTER(C.c_longlong), C.POINTER(C.c_int32)]),
        "orc_tour_length": (C.c_float, [C.c_int, _f32p, _i32p, C.c_int]),
        "orc_check_solution": (C.c_int, [C.POINTER(_Problem), _i32p, C.c_int]),
        "orc_nn_tour": (C.c_int, [C.POINTER(_Problem), _i32p, C.POINTER(C.c_float)]),
        "orc_colony_create": (C.c_void_p, [C.POINTER(_Problem), C.POINTER(_Params)]),
        "orc_colony_destroy": (None, [C.c_void_p]),
        "orc_colony_init": (None, [C.c_void_p]),
        "orc_colony_construct": (None, [C.c_void_p]),
        "orc_colony_local_search": (None, [C.c_void_p]),
        "orc_colony_statistics": (None, [C.c_void_p]),
        "orc_colony_pheromone_update": (None, [C.c_void_p]),
        "orc_colony_iteration": (None, [C.c_void_p]),
        "orc_colony_iter": (C.c_int, [C.c_void_p]),
        "orc_colony_next_iteration": (None, [C.c_void_p]),
        "orc_colony_best": (C.c_float, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]),
        "orc_colony_ants": (C.c_int, [C.c_void_p, _i32p, C.c_int, _i32p, _f32p]),
        "orc_colony_bsf": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
        "orc_colony_tau": (None, [C.c_void_p, _f32p, _f32p]),
        "orc_colony_set_tau": (None, [C.c_void_p, _f32p, _f32p]),
        "orc_colony_set_ants": (None, [C.c_void_p, _i32p, C.c_int, _i32p, _f32p]),
        "orc_colony_ls_evals": (C.c_longlong, [C.c_void_p]),
        "orc_colony_fallbacks": (C.c_longlong, [C.c_void_p]),
        "orc_colony_ls_seed": (C.c_int32, [C.c_void_p]),
        "orc_colony_set_ls_seed": (None, [C.c_void_p, C.c_int32]),
        "orc_colony_set_pm_state": (None, [C.c_void_p, _i32p]),
        "orc_colony_pm_state": (None, [C.c_void_p, _i32p]),
        "orc_route_centers": (C.c_int, [C.c_int, _f32p, _f32p, _i32p, _i32p, C.c_int, C.c_int,
                                        _i32p, _i32p, _f32p, _f32p, _i32p]),
        "orc_partition_routes": (C.c_int, [C.c_int, _f32p, _f32p, _i32p, _i32p, C.c_int, C.c_int,
                                           C.c_int, _i32p, _i32p, _i32p, _i32p]),
        "orc_sub_problem": (None, [C.c_int, _f32p, _i32p, C.c_int, _i32p, _f32p, _i32p]),
        "orc_merge_sub_tours": (C.c_int, [C.c_int, _f32p, C.c_int, _i32p, _i32p, _i32p, _i32p,
                                          _i32p, _i32p, _i32p, C.POINTER(C.c_float), _i32p]),
        "orc_pheromone_from_sub": (None, [C.c_int, C.c_float, _f32p, _f32p, _f32p, C.c_int,
                                          _i32p, _f32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    lib.orc_pheromone_update.restype = None
    return lib


lib = _load()

SEL_IROULETTE, SEL_KNN = 0, 1
RNG_PM, RNG_PHILOX = 0, 1
LS_2OPT, LS_SWAP, LS_RELOCATE = 1, 2, 4
EPSILON = np.float32(0.001)


# ----------------------------------------------------------------------------
# thin wrappers
# ----------------------------------------------------------------------------

def ran01_stream(seed: int, count: int):
    s = C.c_int32(seed)
    out = np.empty(count, np.float32)
    for i in range(count):
        out[i] = lib.orc_ran01(C.byref(s))
    return out, s.value


def ls_seed(seed: int) -> int:
    """Seed of the random-order 2-opt's scan-order stream."""
    return int(lib.orc_ls_seed(seed))


def pm_jump(z: int, t: int) -> int:
    """Park-Miller state after t draws from state z."""
    return int(lib.orc_pm_jump(z, t))


def pm_seed(seed: int, ant: int, lane: int) -> int:
    return int(lib.orc_pm_seed(seed, ant, lane))


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.orc_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def distances(x, y, mode: int) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    n = len(x)
    d = np.empty(n * n, np.float32)
    lib.orc_distances(n, x, y, mode, d)
    return d.reshape(n, n)


def nn_lists(dist: np.ndarray, nn: int, tie: int = 0) -> np.ndarray:
    n = dist.shape[0]
    out = np.empty(n * nn, np.int32)
    lib.orc_nn_lists(n, np.ascontiguousarray(dist, np.float32).ravel(), nn, tie, out)
    return out.reshape(n, nn)


def heuristic(dist: np.ndarray, beta: float = 2.0, prior=None) -> np.ndarray:
    n = dist.shape[0]
    h = np.empty(n * n, np.float32)
    pr = None
    if prior is not None:
        pr = np.ascontiguousarray(prior, np.float32).ravel()
    lib.orc_heuristic(n, np.ascontiguousarray(dist, np.float32).ravel(), beta,
                      pr.ctypes.data if pr is not None else None, h)
    return h.reshape(n, n)


def tree_rank(N: int) -> np.ndarray:
    r = np.empty(N, np.int32)
    lib.orc_tree_rank(N, r)
    return r


def tree_argmax(vals) -> int:
    v = np.ascontiguousarray(vals, np.float32)
    return int(lib.orc_tree_argmax(len(v), v))


def minreduce_phase0(vals, grp: int = 64) -> int:
    v = np.ascontiguousarray(vals, np.float32)
    return int(lib.orc_minreduce_phase0(len(v), grp, v))


def tour_length(dist: np.ndarray, tour) -> float:
    t = np.ascontiguousarray(tour, np.int32)
    return float(lib.orc_tour_length(dist.shape[0], np.ascontiguousarray(dist, np.float32).ravel(),
                                     t, len(t)))


class Problem:
    """Arrays of one CVRP instance, held alive for the C structure."""

    def __init__(self, dist, dem, Q, D=np.inf, svc=0.0, beta=2.0, k=0, nn_ls=0, tie=0,
                 prior=None, nnk=None, nnls=None):
        self.dist = np.ascontiguousarray(dist, np.float32)
        self.n = self.dist.shape[0]
        self.dem = np.ascontiguousarray(dem, np.int32)
        self.Q, self.D, self.svc = int(Q), float(D), float(svc)
        self.heur = heuristic(self.dist, beta, prior)
        self.k = int(k)
        self.nn_ls = int(nn_ls) if nn_ls else self.n - 2
        if nnls is None:
            nnls = nn_lists(self.dist, self.nn_ls, tie)
        self.nnls = np.ascontiguousarray(nnls, np.int32)
        if self.k:
            self.nnk = np.ascontiguousarray(self.nnls[:, : self.k] if nnk is None else nnk,
                                            np.int32)
        else:
            self.nnk = np.zeros((self.n, 1), np.int32)
        self.c = _Problem(self.n, self.dist.ctypes.data, self.dem.ctypes.data, self.Q, self.D,
                          self.svc, self.heur.ctypes.data, self.nnk.ctypes.data, self.k,
                          self.nnls.ctypes.data, self.nn_ls)

    def check(self, tour) -> bool:
        t = np.ascontiguousarray(tour, np.int32)
        return bool(lib.orc_check_solution(C.byref(self.c), t, len(t)))

    def length(self, tour) -> float:
        return tour_length(self.dist, tour)

    def nn_tour(self):
        t = np.empty(2 * self.n, np.int32)
        L = C.c_float()
        sz = lib.orc_nn_tour(C.byref(self.c), t, C.byref(L))
        return t[:sz], L.value

    def construct_ant(self, tinfo, tiknn, selection, rng, seed, iteration, ant, pm):
        t = np.empty(2 * self.n, np.int32)
        L = C.c_float()
        fb = C.c_int(0)
        cfg = _DrawCfg(selection, rng, seed, iteration)
        tk = None if tiknn is None else np.ascontiguousarray(tiknn, np.float32).ravel()
        sz = lib.orc_construct_ant(C.byref(self.c), np.ascontiguousarray(tinfo, np.float32).ravel(),
                                   tk.ctypes.data if tk is not None else None, C.byref(cfg), ant,
                                   pm, t, C.byref(L), C.byref(fb))
        return t[:sz].copy(), L.value, fb.value

    def two_opt(self, tour, random_order=False, seed=1):
        t = np.ascontiguousarray(tour, np.int32).copy()
        s = C.c_int32(seed)
        ev = C.c_longlong(0)
        mv = lib.orc_two_opt(C.byref(self.c), t, len(t), int(random_order), C.byref(s), C.byref(ev))
        return t, mv, ev.value, s.value

    def swap(self, tour):
        t = np.ascontiguousarray(tour, np.int32).copy()
        ev = C.c_longlong(0)
        mv = lib.orc_swap(C.byref(self.c), t, len(t), C.byref(ev))
        return t, mv, ev.value

    def relocate(self, tour):
        t = np.ascontiguousarray(tour, np.int32).copy()
        ev = C.c_longlong(0)
        mv = lib.orc_relocate(C.byref(self.c), t, len(t), C.byref(ev))
        return t, mv, ev.value

    def local_search_seeded(self, tour, flags, seed):
        """(tour, length, evaluated moves, seed after): flags may hold 8,
        the random-order 2-opt drawing from the stream at `seed`."""
        t = np.ascontiguousarray(tour, np.int32).copy()
        ev = C.c_longlong(0)
        s = C.c_int32(seed)
        L = lib.orc_local_search_seeded(C.byref(self.c), flags, t, len(t), C.byref(ev), C.byref(s))
        return t, L, ev.value, s.value

    def local_search(self, tour, flags):
        t = np.ascontiguousarray(tour, np.int32).copy()
        ev = C.c_longlong(0)
        L = lib.orc_local_search(C.byref(self.c), flags, t, len(t), C.byref(ev))
        return t, L, ev.value


def pheromone_init(prob: Problem, tau0: float, alpha: float = 1.0):
    n = prob.n
    tau = np.empty(n * n, np.float32)
    ti = np.empty(n * n, np.float32)
    lib.orc_pheromone_init(n, tau0, prob.heur.ravel(), alpha, tau, ti)
    return tau.reshape(n, n), ti.reshape(n, n)


def pheromone_update(prob: Problem, tau, tours, lens, rho=0.1, alpha=1.0):
    """tours: list of elite tours in elite order (best-so-far first)."""
    n = prob.n
    ras = len(tours)
    tau = np.ascontiguousarray(tau, np.float32).copy().ravel()
    ti = np.empty(n * n, np.float32)
    arrs = [np.ascontiguousarray(t, np.int32) for t in tours]
    ptrs = (C.POINTER(C.c_int) * ras)(*[a.ctypes.data_as(C.POINTER(C.c_int)) for a in arrs])
    sizes = np.array([len(a) for a in arrs], np.int32)
    lens = np.ascontiguousarray(lens, np.float32)
    lib.orc_pheromone_update.argtypes = [C.c_int, C.c_float, C.c_float, C.c_int,
                                         C.POINTER(C.POINTER(C.c_int)), _i32p, _f32p, _f32p,
                                         _f32p, _f32p]
    lib.orc_pheromone_update(n, rho, alpha, ras, ptrs, sizes, lens, prob.heur.ravel(), tau, ti)
    return tau.reshape(n, n), ti.reshape(n, n)


def gather_knn(prob: Problem, tinfo):
    out = np.empty(prob.n * prob.k, np.float32)
    lib.orc_gather_knn(prob.n, prob.k, prob.nnk.ravel(), np.ascontiguousarray(tinfo, np.float32).ravel(),
                       out)
    return out.reshape(prob.n, prob.k)


def elites(lens, ras: int):
    lens = np.ascontiguousarray(lens, np.float32)
    e = np.empty(ras, np.int32)
    lib.orc_elites(len(lens), ras, lens, e)
    return e


def iteration_best(lens, bsf_len: float):
    lens = np.ascontiguousarray(lens, np.float32)
    flag = C.c_int()
    b = lib.orc_iteration_best(len(lens), lens, bsf_len, C.byref(flag))
    return int(b), bool(flag.value)


class Colony:
    """The whole loop (g_aco.cpp:63-134) on the CPU oracle."""

    def __init__(self, prob: Problem, m, alpha=1.0, beta=2.0, rho=0.1, ras=6, selection=SEL_KNN,
                 rng=RNG_PHILOX, ls_flags=0, tau0_rule=0, seed=1, threads=1):
        self.prob = prob
        self.m = m
        self.c_params = _Params(m, alpha, beta, rho, ras, selection, rng, ls_flags, tau0_rule,
                                seed, threads)
        self.h = lib.orc_colony_create(C.byref(prob.c), C.byref(self.c_params))

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_colony_destroy(self.h)
            self.h = None

    def init(self):
        lib.orc_colony_init(self.h)

    def construct(self):
        lib.orc_colony_construct(self.h)

    def local_search(self):
        lib.orc_colony_local_search(self.h)

    def statistics(self):
        lib.orc_colony_statistics(self.h)

    def pheromone_update(self):
        lib.orc_colony_pheromone_update(self.h)

    def iteration(self):
        lib.orc_colony_iteration(self.h)

    def next_iteration(self):
        """The caller's iteration++ after the four phases called one by one."""
        lib.orc_colony_next_iteration(self.h)

    @property
    def iter(self):
        return lib.orc_colony_iter(self.h)

    def best(self):
        t = np.empty(2 * self.prob.n, np.int32)
        sz = C.c_int()
        L = lib.orc_colony_best(self.h, t.ctypes.data, C.byref(sz))
        return t[: sz.value].copy(), L

    def ants(self):
        T = 2 * self.prob.n
        tours = np.zeros((self.m + 1) * T, np.int32)
        sizes = np.zeros(self.m + 1, np.int32)
        lens = np.zeros(self.m + 1, np.float32)
        lib.orc_colony_ants(self.h, tours, T, sizes, lens)
        return tours.reshape(self.m + 1, T), sizes, lens

    def set_ants(self, tours, sizes, lens):
        T = tours.shape[1]
        lib.orc_colony_set_ants(self.h, np.ascontiguousarray(tours, np.int32).ravel(), T,
                                np.ascontiguousarray(sizes, np.int32),
                                np.ascontiguousarray(lens, np.float32))

    def bsf(self):
        cnt = lib.orc_colony_bsf(self.h, None, 0)
        rec = (BsfRecord * max(cnt, 1))()
        lib.orc_colony_bsf(self.h, C.cast(rec, C.c_void_p), cnt)
        return [(r.iter, r.length, r.time) for r in rec[:cnt]]

    def tau(self):
        n = self.prob.n
        tau = np.empty(n * n, np.float32)
        ti = np.empty(n * n, np.float32)
        lib.orc_colony_tau(self.h, tau, ti)
        return tau.reshape(n, n), ti.reshape(n, n)

    def set_tau(self, tau, ti):
        lib.orc_colony_set_tau(self.h, np.ascontiguousarray(tau, np.float32).ravel(),
                               np.ascontiguousarray(ti, np.float32).ravel())

    def ls_evals(self):
        return int(lib.orc_colony_ls_evals(self.h))

    def fallbacks(self):
        return int(lib.orc_colony_fallbacks(self.h))

    def ls_seed(self):
        return int(lib.orc_colony_ls_seed(self.h))

    def set_ls_seed(self, seed):
        lib.orc_colony_set_ls_seed(self.h, int(seed))

    def set_pm_state(self, state):
        lib.orc_colony_set_pm_state(self.h, np.ascontiguousarray(state, np.int32).ravel())

    def pm_state(self):
        out = np.empty(self.m * 64, np.int32)
        lib.orc_colony_pm_state(self.h, out)
        return out.reshape(self.m, 64)


# ----------------------------------------------------------------------------
# sub-problem decomposition
# ----------------------------------------------------------------------------

def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def route_centers(x, y, dem, tour):
    """orc_route_centers: (beg, end, cx, cy, demand) per route."""
    x, y, dem, tour = _f32(x), _f32(y), _i32(dem), _i32(tour)
    cap = len(tour)
    beg, end, rdem = (np.empty(cap, np.int32) for _ in range(3))
    cx, cy = np.empty(cap, np.float32), np.empty(cap, np.float32)
    R = lib.orc_route_centers(len(x), x, y, dem, tour, len(tour), cap, beg, end, cx, cy, rdem)
    assert R >= 0
    return beg[:R], end[:R], cx[:R], cy[:R], rdem[:R]


def partition_routes(x, y, dem, tour, subs=0, rotation=0):
    """orc_partition_routes: list of (real_nodes, sub_tour) per sub-problem."""
    x, y, dem, tour = _f32(x), _f32(y), _i32(dem), _i32(tour)
    n, size = len(x), len(tour)
    off = np.empty(size + 2, np.int32)
    toff = np.empty(size + 2, np.int32)
    real = np.empty(n + size, np.int32)
    st = np.empty(2 * size + 2, np.int32)
    G = lib.orc_partition_routes(n, x, y, dem, tour, size, subs, rotation, off, real, toff, st)
    assert G >= 1
    return [(real[off[g]:off[g + 1]].copy(), st[toff[g]:toff[g + 1]].copy()) for g in range(G)]


def sub_problem(dist, dem, real_nodes):
    """orc_sub_problem: (sub_dist, sub_dem)."""
    dist, dem, real = _f32(dist), _i32(dem), _i32(real_nodes)
    sn = len(real)
    sd = np.empty((sn, sn), np.float32)
    sdem = np.empty(sn, np.int32)
    lib.orc_sub_problem(len(dem), dist, dem, sn, real, sd, sdem)
    return sd, sdem


def merge_sub_tours(dist, parts, bests):
    """orc_merge_sub_tours: parts = [(real_nodes, sub_tour)], bests = [best
    tour of each sub (local numbering) or an empty array]; returns (merged
    tour, length, replaced flags)."""
    dist = _f32(dist)
    n = dist.shape[0]
    off = np.cumsum([0] + [len(r) for r, _ in parts]).astype(np.int32)
    toff = np.cumsum([0] + [len(t) for _, t in parts]).astype(np.int32)
    boff = np.cumsum([0] + [len(b) for b in bests]).astype(np.int32)
    real = _i32(np.concatenate([r for r, _ in parts]))
    st = _i32(np.concatenate([t for _, t in parts]))
    bt = _i32(np.concatenate([np.asarray(b, np.int32) for b in bests] + [np.zeros(1, np.int32)]))
    out = np.empty(2 * n + 2, np.int32)
    ln = C.c_float()
    rep = np.empty(len(parts), np.int32)
    sz = lib.orc_merge_sub_tours(n, dist, len(parts), off, real, toff, st, boff, bt, out, C.byref(ln), rep)
    return out[:sz].copy(), ln.value, rep


def pheromone_from_sub(alpha, heur, tau, tinfo, real_nodes, sub_tau):
    """orc_pheromone_from_sub, in place on tau / tinfo (float32 n x n)."""
    assert tau.dtype == np.float32 and tau.flags.c_contiguous
    assert tinfo.dtype == np.float32 and tinfo.flags.c_contiguous
    real = _i32(real_nodes)
    lib.orc_pheromone_from_sub(tau.shape[0], alpha, _f32(heur), tau, tinfo, len(real), real,
                               _f32(sub_tau))


# ----------------------------------------------------------------------------
# reference CPU build (oracle/_ref), present where /root/reference was built
# ----------------------------------------------------------------------------

def load_ref():
    """Return the reference CPU build, or None when it was not built."""
    if not os.path.exists(REF_PATH):
        return None
    r = C.CDLL(REF_PATH)
    r.ref_ran01.argtypes = [C.POINTER(C.c_int), C.c_int, _f32p]
    r.ref_sort2.argtypes = [_f32p, _i32p, C.c_int, C.c_int]
    r.ref_nn_lists.argtypes = [C.c_int, _f32p, _f32p, C.c_void_p, C.c_void_p, _i32p]
    common = [C.c_int, _f32p, _f32p, _i32p, C.c_int, C.c_float, C.c_float, C.c_void_p]
    r.ref_two_opt.argtypes = common + [C.POINTER(C.c_int), _i32p, C.c_int]
    r.ref_two_opt.restype = C.c_float
    r.ref_two_opt_swap.argtypes = common + [C.POINTER(C.c_int), _i32p, C.c_int]
    r.ref_two_opt_swap.restype = C.c_float
    r.ref_check_solution.argtypes = common + [_i32p, C.c_int, C.POINTER(C.c_float)]
    r.ref_apply_move.argtypes = [C.c_int, _i32p, C.c_int, C.c_int, C.c_int, C.c_float,
                                 C.c_float]
    r.ref_apply_move.restype = C.c_float
    r.ref_read_instance.argtypes = [C.c_char_p, C.c_int, _f32p, _f32p, _i32p,
                                    C.POINTER(C.c_int), C.POINTER(C.c_float),
                                    C.POINTER(C.c_float), C.POINTER(C.c_float)]
    r.ref_route_centers.argtypes = [C.c_int, _f32p, _f32p, _i32p, _i32p, C.c_int, _i32p, _i32p,
                                    _f32p, _f32p, _i32p]
    r.ref_sub_problem.argtypes = [C.c_int, _f32p, _f32p, _i32p, C.c_int, C.c_float, C.c_float, _f32p,
                                  C.c_int, _i32p, _f32p, _i32p, C.c_void_p]
    return r
