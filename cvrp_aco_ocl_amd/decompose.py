"""Sub-problem decomposition of a colony's best solution (the C ABI of
include/cvrp_decompose.h).

The reference holds the parts of a route-based decomposition without a
driver: route centres (compute_route_centers, cvrp_ocl/vrpHelper.cpp:274-294),
sub-problem set-up (init_sub_problem, problem.cpp:84-130), num_subs =
num_node/50, 75 sub-colony iterations and 1 master iteration per outer loop
(problem.cpp:216-218), real_nodes and best_pheromone (problem.h:136-139) and
the report print_problem_decompositon (io.cpp:83-97).  Here:

* ``route_centers``, ``partition_routes``, ``sub_problem``,
  ``merge_sub_tours``: the host steps, one C call each;
* ``decompose_round``: one round on the GPU (aco_decompose_round): the
  master's best-so-far is cut into sub-problems, all of them are solved in
  one batched context, and the merged tour is returned;
* ``solve_decomposed``: the outer loop -- master iterations, a round every
  ``period`` iterations, the merged tour put into the colony as an ant of
  the next iteration when it beats the best-so-far.  The master runs
  without local search by default, so instances past the local-search LDS
  limit are improved by the sub-colonies' local search.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib
from ._lib import check, lib
from .solver import GpuACO, Instance, _ptr


def _coords(inst: Instance):
    """Node coordinates as the reference's Point floats (problem.h:40-44)."""
    return np.ascontiguousarray(inst.x, np.float32), np.ascontiguousarray(inst.y, np.float32)


def route_centers(inst: Instance, tour):
    """aco_route_centers: (beg, end, cx, cy, demand) of each route."""
    x, y = _coords(inst)
    t = np.ascontiguousarray(tour, np.int32)
    dem = np.ascontiguousarray(inst.demand, np.int32)
    cap = len(t)
    beg, end, rdem = (np.empty(cap, np.int32) for _ in range(3))
    cx, cy = np.empty(cap, np.float32), np.empty(cap, np.float32)
    cnt = C.c_int32()
    check(lib.aco_route_centers(inst.n, _ptr(x), _ptr(y), _ptr(dem), _ptr(t), len(t), cap, _ptr(beg),
                                _ptr(end), _ptr(cx), _ptr(cy), _ptr(rdem), C.byref(cnt)), "aco_route_centers")
    R = cnt.value
    return beg[:R], end[:R], cx[:R], cy[:R], rdem[:R]


def partition_routes(inst: Instance, tour, subs: int = 0, rotation: int = 0):
    """aco_partition_routes: [(real_nodes, sub_tour)] per sub-problem."""
    x, y = _coords(inst)
    t = np.ascontiguousarray(tour, np.int32)
    dem = np.ascontiguousarray(inst.demand, np.int32)
    size = len(t)
    off, toff = np.empty(size + 1, np.int32), np.empty(size + 1, np.int32)
    real, st = np.empty(inst.n + size, np.int32), np.empty(2 * size, np.int32)
    cnt = C.c_int32()
    check(lib.aco_partition_routes(inst.n, _ptr(x), _ptr(y), _ptr(dem), _ptr(t), size, subs, rotation,
                                   _ptr(off), _ptr(real), _ptr(toff), _ptr(st), C.byref(cnt)),
          "aco_partition_routes")
    return [(real[off[g]:off[g + 1]].copy(), st[toff[g]:toff[g + 1]].copy()) for g in range(cnt.value)]


def sub_problem(inst: Instance, real_nodes) -> Instance:
    """aco_sub_problem: the sub-instance on real_nodes (depot first)."""
    real = np.ascontiguousarray(real_nodes, np.int32)
    sn = len(real)
    sd = np.empty((sn, sn), np.float32)
    sdem = np.empty(sn, np.int32)
    d = np.ascontiguousarray(inst.dist, np.float32)
    dem = np.ascontiguousarray(inst.demand, np.int32)
    check(lib.aco_sub_problem(inst.n, _ptr(d), _ptr(dem), sn, _ptr(real), _ptr(sd), _ptr(sdem)),
          "aco_sub_problem")
    return Instance(n=sn, x=inst.x[real].copy(), y=inst.y[real].copy(), demand=sdem,
                    capacity=inst.capacity, dist=sd, max_dist=inst.max_dist,
                    service_time=inst.service_time, name=f"{inst.name}-sub", dist_mode=inst.dist_mode,
                    meta={"real_nodes": real})


def merge_sub_tours(inst: Instance, parts, bests):
    """aco_merge_sub_tours: bests[s] = best tour of sub s (local numbering)
    or an empty array; returns (merged tour, length, replaced flags)."""
    d = np.ascontiguousarray(inst.dist, np.float32)
    off = np.cumsum([0] + [len(r) for r, _ in parts]).astype(np.int32)
    toff = np.cumsum([0] + [len(t) for _, t in parts]).astype(np.int32)
    boff = np.cumsum([0] + [len(b) for b in bests]).astype(np.int32)
    real = np.ascontiguousarray(np.concatenate([r for r, _ in parts]), np.int32)
    st = np.ascontiguousarray(np.concatenate([t for _, t in parts]), np.int32)
    bt = np.ascontiguousarray(np.concatenate([np.asarray(b, np.int32) for b in bests] + [np.zeros(1, np.int32)]),
                              np.int32)
    cap = 2 * inst.n + 2
    out = np.empty(cap, np.int32)
    size, length = C.c_int32(), C.c_float()
    rep = np.empty(len(parts), np.int32)
    check(lib.aco_merge_sub_tours(inst.n, _ptr(d), len(parts), _ptr(off), _ptr(real), _ptr(toff), _ptr(st),
                                  _ptr(boff), _ptr(bt), _ptr(out), cap, C.byref(size), C.byref(length),
                                  _ptr(rep)), "aco_merge_sub_tours")
    return out[:size.value].copy(), length.value, rep


def sub_params(*, n_ants=64, alpha=1.0, beta=2.0, rho=0.1, ras_ranks=6, k=16, nn_ls=0,
               selection=_lib.SEL_KNN, rng=_lib.RNG_PHILOX, ls_flags=7,
               tau0_rule=_lib.TAU0_FIRST_COLONY, seed=1) -> _lib.AcoParams:
    """Parameters of the sub-colonies (aco_params)."""
    p = _lib.AcoParams()
    lib.aco_default_params(C.byref(p))
    p.alpha, p.beta, p.rho = alpha, beta, rho
    p.n_ants, p.ras_ranks, p.k, p.nn_ls = n_ants, ras_ranks, k, nn_ls
    p.selection, p.rng, p.ls_flags, p.tau0_rule = selection, rng, ls_flags, tau0_rule
    p.seed, p.use_graph = seed, 1
    return p


def decompose_round(g: GpuACO, inst: Instance, b: int = 0, *, subs: int = 0, sub_iterations: int = 75,
                    rotation: int = 0, pheromone_back: bool = False, params: _lib.AcoParams | None = None):
    """aco_decompose_round on instance b of ``g`` (whose instance is
    ``inst``): returns (merged tour, length, sub-problems, replaced)."""
    x, y = _coords(inst)
    d = np.ascontiguousarray(inst.dist, np.float32)
    dem = np.ascontiguousarray(inst.demand, np.int32)
    pb = _lib.AcoProblem(inst.n, _ptr(d), _ptr(dem), inst.capacity, inst.max_dist, inst.service_time,
                         None, 0, None)
    dp = _lib.AcoDecomposeParams(subs, sub_iterations, rotation, int(bool(pheromone_back)),
                                 params if params is not None else sub_params())
    cap = 2 * inst.n + 2
    out = np.empty(cap, np.int32)
    size, length = C.c_int32(), C.c_float()
    nsub, rep = C.c_int32(), C.c_int32()
    check(lib.aco_decompose_round(g.h, b, C.byref(pb), _ptr(x), _ptr(y), C.byref(dp), _ptr(out), cap,
                                  C.byref(size), C.byref(length), C.byref(nsub), C.byref(rep)),
          "aco_decompose_round")
    return out[:size.value].copy(), length.value, nsub.value, rep.value


def solve_decomposed(inst: Instance, iterations: int, *, period: int = 1, subs: int = 0,
                     sub_iterations: int = 75, pheromone_back: bool = False, rotation_step: int = 1,
                     master: dict | None = None, sub: dict | None = None, pheromone: bool = False):
    """The outer loop: ``iterations`` master iterations; after every
    ``period``-th a decomposition round (rotation = round x rotation_step,
    sub-colony seed = sub seed + 1000003 x round); a merged tour shorter
    than the best-so-far becomes the last ant of the next iteration.
    Returns a dict with the best tour, its length, the master's best-so-far
    records, per round (length before, merged length, sub-problems,
    replaced) and, with ``pheromone``, the master's final (tau, total_info)."""
    mk = dict(n_ants=inst.n, k=16, selection=_lib.SEL_KNN, rng=_lib.RNG_PHILOX, ls_flags=0, seed=1)
    mk.update(master or {})
    sk = dict(sub or {})
    base_seed = sk.pop("seed", 1)
    rounds = []
    t0 = time.perf_counter()
    with GpuACO(inst, **mk) as g:
        g.init_aco()
        pending = None
        for it in range(iterations):
            if pending is not None:
                g.construct_solutions()
                if mk.get("ls_flags", 0):
                    g.local_search()
                g.set_ant(g.m - 1, pending)
                g.update_statistics()
                g.pheromone_update()
                pending = None
            else:
                g.run_aco_iteration()
            if (it + 1) % period == 0:
                r = len(rounds)
                _, before = g.best()
                tour, length, nsub, rep = decompose_round(
                    g, inst, subs=subs, sub_iterations=sub_iterations, rotation=r * rotation_step,
                    pheromone_back=pheromone_back, params=sub_params(seed=base_seed + 1000003 * r, **sk))
                rounds.append((before, length, nsub, rep))
                if length < before:
                    pending = tour
        best, blen = g.best()
        if pending is not None and len(rounds) and rounds[-1][1] < blen:
            best, blen = pending, rounds[-1][1]
        records = g.bsf_records()
        trail = g.pheromone() if pheromone else None
    return {"tour": best, "length": float(blen), "records": records, "rounds": rounds,
            "wall_s": time.perf_counter() - t0, "pheromone": trail}
