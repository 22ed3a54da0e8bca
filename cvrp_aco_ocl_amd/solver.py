"""Python mirror of the reference's host structure over the C ABI.

``Instance`` holds what ``Problem`` (cvrp_ocl/problem.h:82-140) carries into
the loop; ``GpuACO`` mirrors ``g_ACO`` (cvrp_ocl/gpu/g_aco.h:16-37) method
for method and drives libcvrpaco.so.  Every computation of the loop runs in
the HIP library; this module only marshals arrays.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import lib, check


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


@dataclass
class Instance:
    """One CVRP instance: node 0 is the depot."""

    n: int
    x: np.ndarray
    y: np.ndarray
    demand: np.ndarray
    capacity: int
    dist: np.ndarray
    max_dist: float = math.inf
    service_time: float = 0.0
    optimum: float = 0.0
    name: str = ""
    dist_mode: int = _lib.DIST_NINT
    meta: dict = field(default_factory=dict)

    @classmethod
    def from_coords(cls, x, y, demand, capacity, dist_mode=_lib.DIST_NINT, max_dist=math.inf,
                    service_time=0.0, name="", optimum=0.0, meta=None):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        n = len(x)
        dist = np.empty((n, n), np.float32)
        check(lib.aco_compute_distances(n, _ptr(x), _ptr(y), dist_mode, _ptr(dist)),
              "aco_compute_distances")
        return cls(n=n, x=x, y=y, demand=np.ascontiguousarray(demand, np.int32),
                   capacity=int(capacity), dist=dist, max_dist=float(max_dist),
                   service_time=float(service_time), optimum=float(optimum), name=name,
                   dist_mode=dist_mode, meta=dict(meta or {}))

    @classmethod
    def from_vrplib(cls, path: str, dist_mode: int | None = None):
        """read_instance_file (io.cpp:334-517).  dist_mode None keeps the
        file's EDGE_WEIGHT_TYPE with EUC_2D as the reference's float sqrt."""
        n = C.c_int32()
        check(lib.aco_read_vrplib(path.encode(), 0, C.byref(n), None, None, None, None, None,
                                  None, None, None), f"aco_read_vrplib({path})")
        n = n.value
        x = np.empty(n, np.float64)
        y = np.empty(n, np.float64)
        dem = np.empty(n, np.int32)
        Q = C.c_int32()
        D = C.c_float()
        svc = C.c_float()
        opt = C.c_float()
        ew = C.c_int32()
        nn = C.c_int32()
        check(lib.aco_read_vrplib(path.encode(), n, C.byref(nn), _ptr(x), _ptr(y), _ptr(dem),
                                  C.byref(Q), C.byref(D), C.byref(svc), C.byref(opt),
                                  C.byref(ew)), f"aco_read_vrplib({path})")
        if dist_mode is None:
            dist_mode = {0: _lib.DIST_FLOAT, 1: _lib.DIST_CEIL, 2: _lib.DIST_GEO,
                         3: _lib.DIST_ATT}[ew.value]
        import os
        return cls.from_coords(x, y, dem, Q.value, dist_mode, D.value, svc.value,
                               name=os.path.basename(path), optimum=opt.value)

    def nn_lists(self, nn: int, tie: int = 0) -> np.ndarray:
        out = np.empty((self.n, nn), np.int32)
        d = np.ascontiguousarray(self.dist, np.float32)
        check(lib.aco_compute_nn_lists(self.n, _ptr(d), nn, tie, _ptr(out)), "aco_compute_nn_lists")
        return out

    @staticmethod
    def distances_batch(xs, ys, mode=_lib.DIST_NINT) -> list:
        """Distance matrices of many coordinate sets in one pass on the GPU
        (aco_compute_distances_batch)."""
        xs = [np.ascontiguousarray(x, np.float64) for x in xs]
        ys = [np.ascontiguousarray(y, np.float64) for y in ys]
        outs = [np.empty((len(x), len(x)), np.float32) for x in xs]
        ns = np.array([len(x) for x in xs], np.int32)
        xp = (C.c_void_p * len(xs))(*[a.ctypes.data for a in xs])
        yp = (C.c_void_p * len(ys))(*[a.ctypes.data for a in ys])
        op = (C.c_void_p * len(outs))(*[a.ctypes.data for a in outs])
        check(lib.aco_compute_distances_batch(len(xs), _ptr(ns), C.cast(xp, C.c_void_p), C.cast(yp, C.c_void_p),
                                              mode, C.cast(op, C.c_void_p)), "aco_compute_distances_batch")
        return outs

    @staticmethod
    def nn_lists_batch(insts, nn, tie: int = 0) -> list:
        """nn_lists of many instances in one pass on the GPU
        (aco_compute_nn_lists_batch); nn is one width or one per instance."""
        insts = list(insts)
        widths = [nn] * len(insts) if np.isscalar(nn) else list(nn)
        outs = [np.empty((it.n, w), np.int32) for it, w in zip(insts, widths)]
        ds = [np.ascontiguousarray(it.dist, np.float32) for it in insts]
        ns = np.array([it.n for it in insts], np.int32)
        ws = np.array(widths, np.int32)
        dptr = (C.c_void_p * len(insts))(*[d.ctypes.data for d in ds])
        optr = (C.c_void_p * len(insts))(*[o.ctypes.data for o in outs])
        check(lib.aco_compute_nn_lists_batch(len(insts), _ptr(ns), C.cast(dptr, C.c_void_p), _ptr(ws), tie,
                                             C.cast(optr, C.c_void_p)), "aco_compute_nn_lists_batch")
        return outs

    def check_solution(self, tour) -> bool:
        t = np.ascontiguousarray(tour, np.int32)
        d = np.ascontiguousarray(self.dist, np.float32)
        rc = lib.aco_check_solution(self.n, _ptr(d), _ptr(self.demand), self.capacity,
                                    self.max_dist, self.service_time, _ptr(t), len(t))
        if rc < 0:
            check(rc, "aco_check_solution")
        return bool(rc)

    def subset(self, count: int) -> "Instance":
        """The depot and the first count-1 customers (same capacity and
        limits)."""
        return Instance(n=count, x=self.x[:count].copy(), y=self.y[:count].copy(),
                        demand=self.demand[:count].copy(), capacity=self.capacity,
                        dist=np.ascontiguousarray(self.dist[:count, :count]),
                        max_dist=self.max_dist, service_time=self.service_time,
                        name=f"{self.name}[:{count}]", dist_mode=self.dist_mode)

    def nn_tour(self):
        """The tau0 = 1/(n * L_nn) rule's greedy tour (aco_nn_tour): returns
        (tour, L_nn)."""
        d = np.ascontiguousarray(self.dist, np.float32)
        t = np.empty(2 * self.n, np.int32)
        size = C.c_int32()
        length = C.c_float()
        check(lib.aco_nn_tour(self.n, _ptr(d), _ptr(self.demand), self.capacity, self.max_dist,
                              self.service_time, _ptr(t), len(t), C.byref(size), C.byref(length)),
              "aco_nn_tour")
        return t[: size.value].copy(), float(length.value)

    def tour_length(self, tour) -> float:
        t = np.asarray(tour, np.int64)
        length = np.float32(0.0)
        for a, b in zip(t[:-1], t[1:]):
            length = np.float32(length + self.dist[a, b])
        return float(length)


class GpuACO:
    """g_ACO (cvrp_ocl/gpu/g_aco.h:16-37) on libcvrpaco.so.

    ``inst`` is one Instance, or a list of Instances solved together in one
    context (aco_create_batch: every launch covers all of them, instance b
    seeded with seed + b).  Accessors take the instance index ``b``."""

    def __init__(self, inst, *, alpha=1.0, beta=2.0, rho=0.1, n_ants=0, ras_ranks=6,
                 k=0, nn_ls=0, selection=_lib.SEL_IROULETTE, rng=_lib.RNG_PARK_MILLER,
                 ls_flags=0, tau0_rule=_lib.TAU0_FIRST_COLONY, seed=1, use_graph=True,
                 bsf_capacity=4096, prior=None, nn_list=None):
        insts = list(inst) if isinstance(inst, (list, tuple)) else [inst]
        priors = prior if isinstance(prior, (list, tuple)) else [prior] * len(insts)
        nnls = nn_list if isinstance(nn_list, (list, tuple)) else [nn_list] * len(insts)
        self.insts = insts
        self.inst = insts[0]
        p = _lib.AcoParams()
        lib.aco_default_params(C.byref(p))
        p.alpha, p.beta, p.rho = alpha, beta, rho
        p.n_ants, p.ras_ranks, p.k, p.nn_ls = n_ants, ras_ranks, k, nn_ls
        p.selection, p.rng, p.ls_flags, p.tau0_rule = selection, rng, ls_flags, tau0_rule
        p.seed, p.use_graph, p.bsf_capacity = seed, int(bool(use_graph)), bsf_capacity
        self.params = p
        self._keep = []
        probs = (_lib.AcoProblem * len(insts))()
        for b, it in enumerate(insts):
            d = np.ascontiguousarray(it.dist, np.float32)
            dem = np.ascontiguousarray(it.demand, np.int32)
            pr = None if priors[b] is None else np.ascontiguousarray(priors[b], np.float32)
            nn = None if nnls[b] is None else np.ascontiguousarray(nnls[b], np.int32)
            self._keep += [d, dem, pr, nn]
            probs[b] = _lib.AcoProblem(it.n, _ptr(d), _ptr(dem), it.capacity, it.max_dist,
                                       it.service_time, _ptr(nn), 0 if nn is None else nn.shape[1],
                                       _ptr(pr))
        h = C.c_void_p()
        check(lib.aco_create_batch(probs, len(insts), C.byref(p), C.byref(h)), "aco_create")
        self._keep = None  # inputs are copied at creation
        self.h = h
        self.B = len(insts)
        self.n = insts[0].n
        self.m = n_ants if n_ants > 0 else insts[0].n
        self.T = 2 * insts[0].n

    # -- sub-problem decomposition (include/cvrp_decompose.h) ---------------
    def set_ant(self, ant: int, tour, b: int = 0):
        """aco_set_ant_at: ant `ant` becomes `tour` (length in step order)."""
        t = np.ascontiguousarray(tour, np.int32)
        check(lib.aco_set_ant_at(self.h, b, ant, _ptr(t), len(t)), "aco_set_ant_at")

    def pheromone_from_subs(self, subs: "GpuACO", parts, mapping, b: int = 0):
        """aco_pheromone_from_subs: instance s of ``subs`` holds the nodes
        parts[mapping[s]][0] of instance b."""
        off = np.cumsum([0] + [len(r) for r, _ in parts]).astype(np.int32)
        real = np.ascontiguousarray(np.concatenate([r for r, _ in parts]), np.int32)
        mp = np.ascontiguousarray(mapping, np.int32)
        check(lib.aco_pheromone_from_subs(self.h, b, subs.h, _ptr(off), _ptr(real), _ptr(mp)),
              "aco_pheromone_from_subs")

    # -- lifetime ----------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            lib.aco_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- the reference loop (g_aco.cpp:63-134) -------------------------------
    def init_aco(self):
        check(lib.aco_init(self.h), "aco_init")

    def run_aco_iteration(self):
        check(lib.aco_iteration(self.h), "aco_iteration")

    def run(self, iterations: int):
        check(lib.aco_run(self.h, iterations), "aco_run")

    def exit_aco(self, b: int = 0):
        """g_aco.cpp:95-117: best-so-far ant and its record trajectory."""
        tour, length = self.best(b)
        return tour, length, self.bsf_records(b)

    def construct_solutions(self):
        check(lib.aco_construct(self.h), "aco_construct")

    def local_search(self):
        check(lib.aco_local_search(self.h), "aco_local_search")

    def update_statistics(self):
        check(lib.aco_update_statistics(self.h), "aco_update_statistics")

    def pheromone_update(self):
        check(lib.aco_pheromone_update(self.h), "aco_pheromone_update")

    def pheromone_init(self, tau0: float):
        check(lib.aco_pheromone_init(self.h, tau0), "aco_pheromone_init")

    def synchronize(self):
        check(lib.aco_synchronize(self.h), "aco_synchronize")

    # -- state readback ----------------------------------------------------
    @property
    def iteration(self) -> int:
        it = C.c_int32()
        check(lib.aco_get_iteration(self.h, C.byref(it)), "aco_get_iteration")
        return it.value

    def _T(self, b):
        return 2 * self.insts[b].n

    def best(self, b: int = 0):
        T = self._T(b)
        t = np.empty(T, np.int32)
        sz = C.c_int32()
        L = C.c_float()
        check(lib.aco_get_best_at(self.h, b, _ptr(t), T, C.byref(sz), C.byref(L)), "aco_get_best")
        return t[: sz.value].copy(), L.value

    def ants(self, b: int = 0):
        T = self._T(b)
        tours = np.zeros((self.m + 1, T), np.int32)
        sizes = np.zeros(self.m + 1, np.int32)
        lens = np.zeros(self.m + 1, np.float32)
        check(lib.aco_get_ants_at(self.h, b, _ptr(tours), T, _ptr(sizes), _ptr(lens)),
              "aco_get_ants")
        return tours, sizes, lens

    def set_ants(self, tours, sizes, lens, b: int = 0):
        tours = np.ascontiguousarray(tours, np.int32)
        sizes = np.ascontiguousarray(sizes, np.int32)
        lens = np.ascontiguousarray(lens, np.float32)
        check(lib.aco_set_ants_at(self.h, b, _ptr(tours), tours.shape[1], _ptr(sizes), _ptr(lens)),
              "aco_set_ants")

    def bsf_records(self, b: int = 0):
        cnt = C.c_int32()
        check(lib.aco_get_bsf_at(self.h, b, None, 0, C.byref(cnt)), "aco_get_bsf")
        rec = (_lib.AcoBsfRecord * max(cnt.value, 1))()
        check(lib.aco_get_bsf_at(self.h, b, C.cast(rec, C.c_void_p), cnt.value, C.byref(cnt)),
              "aco_get_bsf")
        return [(r.iter, r.length, r.time) for r in rec[: cnt.value]]

    def pheromone(self, b: int = 0):
        n = self.insts[b].n
        tau = np.empty((n, n), np.float32)
        ti = np.empty((n, n), np.float32)
        check(lib.aco_get_pheromone_at(self.h, b, _ptr(tau), _ptr(ti)), "aco_get_pheromone")
        return tau, ti

    def set_pheromone(self, tau, ti, b: int = 0):
        tau = np.ascontiguousarray(tau, np.float32)
        ti = np.ascontiguousarray(ti, np.float32)
        check(lib.aco_set_pheromone_at(self.h, b, _ptr(tau), _ptr(ti)), "aco_set_pheromone")

    def rng_state(self):
        out = np.empty((self.m, 64), np.int32)
        check(lib.aco_get_rng_state(self.h, _ptr(out)), "aco_get_rng_state")
        return out

    def setup_arrays(self, b: int = 0) -> dict:
        """The arrays aco_create derived on the device (aco_get_setup_at)."""
        it = self.insts[b]
        n, k = it.n, self.params.k
        nl = self.params.nn_ls if self.params.nn_ls > 0 else n - 2
        out = {"heur": np.empty((n, n), np.float32), "d0": np.empty(n, np.float32),
               "nnk": np.empty((n, max(k, 1)), np.int32), "nnkd": np.empty((n, max(k, 1)), np.float32),
               "nnls": np.empty((n, nl), np.int32), "nnd": np.empty((n, nl), np.float32),
               "nnpos": np.empty((n, n), np.uint16), "dist16": np.zeros((n, n), np.uint16)}
        check(lib.aco_get_setup_at(self.h, b, *[_ptr(out[key]) for key in
                                               ("heur", "d0", "nnk", "nnkd", "nnls", "nnd", "nnpos",
                                                "dist16")]), "aco_get_setup_at")
        if k == 0:
            out["nnk"] = out["nnk"][:, :0]
            out["nnkd"] = out["nnkd"][:, :0]
        return out

    def set_candidates(self, cand, b: int = 0):
        """Construction candidate rows of instance b (n x >= k ids; the first
        k of each row are used), aco_set_candidates_at."""
        cand = np.ascontiguousarray(cand, np.int32)
        check(lib.aco_set_candidates_at(self.h, b, _ptr(cand), cand.shape[1]), "aco_set_candidates_at")

    def set_rng_state(self, state):
        st = np.ascontiguousarray(state, np.int32)
        check(lib.aco_set_rng_state(self.h, _ptr(st)), "aco_set_rng_state")

    def ls_seed(self) -> int:
        z = C.c_int32()
        check(lib.aco_get_ls_seed(self.h, C.byref(z)), "aco_get_ls_seed")
        return z.value

    def set_ls_seed(self, seed: int):
        check(lib.aco_set_ls_seed(self.h, int(seed)), "aco_set_ls_seed")

    def counters(self) -> dict:
        c = _lib.AcoCounters()
        check(lib.aco_get_counters(self.h, C.byref(c)), "aco_get_counters")
        return {"ls_evals": c.ls_evals, "ls_moves": c.ls_moves, "fallbacks": c.fallbacks,
                "iterations": c.iterations, "bsf_overflow": c.bsf_overflow}

    # -- launch shapes and the autotuner ---------------------------------
    def launch_shape(self) -> tuple:
        """(local-search waves, ring rows, prefetch, construction ants per
        block, 16-bit local-search rows, fused iteration)."""
        sh = _lib.AcoLaunchShape()
        check(lib.aco_get_launch_shape(self.h, C.byref(sh)), "aco_get_launch_shape")
        return (sh.ls_waves, sh.ls_ring, sh.ls_prefetch, sh.construct_apb, sh.ls_row16, sh.fused_iteration)

    def set_launch_shape(self, waves: int, ring: int, prefetch: int, construct_apb: int = 0,
                         row16: int = 0, fused: int = 0):
        sh = _lib.AcoLaunchShape(waves, ring, prefetch, construct_apb, row16, fused)
        check(lib.aco_set_launch_shape(self.h, C.byref(sh)), "aco_set_launch_shape")

    def autotune(self, reps: int = 3):
        """Times every construction and local-search candidate on the current
        state and keeps the fastest of each (state and results unchanged).
        Returns [(shape, stage, ms or -1)], stage 0 = construction,
        1 = local search, 2 = whole iteration (per-phase or fused)."""
        res = (_lib.AcoTuneResult * 32)()
        cnt = C.c_int32()
        check(lib.aco_autotune(self.h, reps, C.cast(res, C.c_void_p), 32, C.byref(cnt)),
              "aco_autotune")
        return [((r.shape.ls_waves, r.shape.ls_ring, r.shape.ls_prefetch, r.shape.construct_apb,
                  r.shape.ls_row16, r.shape.fused_iteration), r.stage, r.ms) for r in res[: cnt.value]]

    def profile_enable(self, on: bool = True):
        check(lib.aco_profile_enable(self.h, int(on)), "aco_profile_enable")

    def profile(self, name: str):
        ms = C.c_double()
        cnt = C.c_int64()
        check(lib.aco_profile_get(self.h, name.encode(), C.byref(ms), C.byref(cnt)),
              "aco_profile_get")
        return ms.value, cnt.value
