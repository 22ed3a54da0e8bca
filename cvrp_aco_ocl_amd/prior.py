"""Learned edge prior over the C ABI of cvrp_prior.h.

``PriorNet`` wraps one network on the GPU (features, forward, gradient,
Adam, minibatch training, dense prior for ``GpuACO(prior=...)``).
``solve_plain`` / ``edge_labels`` make the training labels (the edges of
the plain colony's best tours); ``run_colony`` runs a batched colony with
or without the prior and returns its best-so-far trajectories.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import lib, check

NPARAM = _lib.PRIOR_PARAMS
NF = _lib.PRIOR_FEATURES
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                            "prior_weights.f32")


def _ptr(a):
    return None if a is None else a.ctypes.data


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def _instance_arrays(inst):
    """Contiguous fp32 / int32 copies that stay alive across the C call."""
    return _f32(inst.x), _f32(inst.y), _i32(inst.demand), _f32(inst.dist)


def features(inst, nn, k):
    """GPU features of the n*k candidate edges of ``inst`` ([n*k, 12])."""
    nn = _i32(nn)
    x, y, dem, dist = _instance_arrays(inst)
    out = np.empty((inst.n * k, NF), np.float32)
    check(lib.aco_prior_features(inst.n, _ptr(x), _ptr(y), _ptr(dem), inst.capacity, _ptr(dist),
                                 _ptr(nn), nn.shape[1], k, _ptr(out)), "aco_prior_features")
    return out


def edge_labels(nn, k, tour):
    """1 where candidate edge (i, nn[i][r]) joins consecutive tour nodes."""
    nn = np.asarray(nn, np.int64)[:, :k]
    n = nn.shape[0]
    adj = np.zeros((n, n), bool)
    t = np.asarray(tour, np.int64)
    adj[t[:-1], t[1:]] = True
    adj[t[1:], t[:-1]] = True
    return adj[np.repeat(np.arange(n), k), nn.ravel()].astype(np.float32)


class PriorNet:
    def __init__(self, seed: int = 1, params=None):
        h = C.c_void_p()
        check(lib.aco_prior_create(seed, C.byref(h)), "aco_prior_create")
        self.h = h
        if params is not None:
            self.set_params(params)

    def close(self):
        if getattr(self, "h", None):
            lib.aco_prior_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- parameters --------------------------------------------------------
    def params(self) -> np.ndarray:
        out = np.empty(NPARAM, np.float32)
        check(lib.aco_prior_get_params(self.h, _ptr(out), NPARAM), "aco_prior_get_params")
        return out

    def set_params(self, p):
        p = _f32(p)
        check(lib.aco_prior_set_params(self.h, _ptr(p), NPARAM), "aco_prior_set_params")

    def save(self, path: str = WEIGHTS_PATH):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        self.params().tofile(path)

    @classmethod
    def load(cls, path: str = WEIGHTS_PATH):
        p = np.fromfile(path, np.float32)
        if p.size != NPARAM:
            raise ValueError(f"{path}: {p.size} floats, expected {NPARAM}")
        return cls(params=p)

    # -- compute -----------------------------------------------------------
    def forward(self, feat) -> np.ndarray:
        feat = _f32(feat)
        out = np.empty(len(feat), np.float32)
        check(lib.aco_prior_forward(self.h, _ptr(feat), len(feat), _ptr(out)), "aco_prior_forward")
        return out

    def gradient(self, feat, label, pos_weight: float):
        feat, label = _f32(feat), _f32(label)
        g = np.empty(NPARAM, np.float32)
        loss = C.c_float()
        check(lib.aco_prior_gradient(self.h, _ptr(feat), _ptr(label), len(feat), pos_weight,
                                     _ptr(g), C.byref(loss)), "aco_prior_gradient")
        return g, loss.value

    def adam_step(self, grad, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        grad = _f32(grad)
        check(lib.aco_prior_adam_step(self.h, _ptr(grad), lr, beta1, beta2, eps),
              "aco_prior_adam_step")

    def train(self, feat, label, **kw):
        feat, label = _f32(feat), _f32(label)
        tp = _lib.AcoPriorTrainParams()
        lib.aco_prior_default_train_params(C.byref(tp))
        for key, val in kw.items():
            setattr(tp, key, val)
        losses = np.empty(tp.epochs, np.float32)
        check(lib.aco_prior_train(self.h, _ptr(feat), _ptr(label), len(feat), C.byref(tp),
                                  _ptr(losses)), "aco_prior_train")
        return losses

    def build(self, inst, nn, k, exponent=2.0, floor=0.05) -> np.ndarray:
        """Dense n x n multiplier for GpuACO(prior=...): sigmoid(logit)^exponent
        on candidate edges, floor^exponent elsewhere."""
        nn = _i32(nn)
        x, y, dem, dist = _instance_arrays(inst)
        out = np.empty((inst.n, inst.n), np.float32)
        check(lib.aco_prior_build(self.h, inst.n, _ptr(x), _ptr(y), _ptr(dem), inst.capacity,
                                  _ptr(dist), _ptr(nn), nn.shape[1], k, exponent, floor,
                                  _ptr(out)), "aco_prior_build")
        return out


    def apply(self, g, k=16, exponent=2.0, floor=0.05):
        """Multiply eta of every instance of the GpuACO context ``g`` (before
        init_aco) by the prior, on the device (aco_prior_apply)."""
        xs = [_f32(it.x) for it in g.insts]
        ys = [_f32(it.y) for it in g.insts]
        xp = (C.c_void_p * len(xs))(*[a.ctypes.data for a in xs])
        yp = (C.c_void_p * len(ys))(*[a.ctypes.data for a in ys])
        check(lib.aco_prior_apply(g.h, self.h, C.cast(xp, C.c_void_p), C.cast(yp, C.c_void_p), k,
                                  exponent, floor), "aco_prior_apply")

    def rank_candidates(self, inst, k, pool):
        """Construction candidate rows chosen by the network: for each node,
        the k of its ``pool`` nearest neighbours (distance-sorted, ties to the
        lower index) with the highest logits, ties to the nearer, in order of
        decreasing logit.  Features are computed over the pool (rank r/pool,
        kd the distance to the pool-th neighbour), so a network trained on
        k-wide lists sees them rescaled.  For GpuACO.set_candidates."""
        nn = np.ascontiguousarray(inst.nn_lists(pool), np.int32)
        z = self.forward(features(inst, nn, pool)).reshape(inst.n, pool)
        order = np.argsort(-z, axis=1, kind="stable")[:, :k]
        return np.ascontiguousarray(np.take_along_axis(nn, order, axis=1), np.int32)


def solve_plain(insts, iterations, n_ants=64, k=16, seed=1, ls_flags=7):
    """Plain colony (roulette over k candidates, Philox, local search) on a
    batch of instances in one context; returns the best tours."""
    from .solver import GpuACO
    g = GpuACO(list(insts), n_ants=n_ants, k=k, selection=_lib.SEL_KNN, rng=_lib.RNG_PHILOX,
               ls_flags=ls_flags, tau0_rule=_lib.TAU0_NN_TOUR, seed=seed)
    try:
        g.init_aco()
        g.run(iterations)
        return [g.best(b) for b in range(len(insts))]
    finally:
        g.close()


def run_colony(insts, iterations, net=None, n_ants=64, k=16, seed=1, ls_flags=7, beta=2.0,
               floor=0.05, targets=None, chunk=25, prior_exponent=None, autotune=True, rerank_pool=0):
    """Batched colony over ``insts`` (the prior from ``net`` multiplies eta
    when given).  Stops early once every instance reached its target cost
    (``targets``).  The prior is applied on the device (aco_prior_apply)
    with exponent ``prior_exponent`` (default beta: the probability
    multiplies eta; 0: eta unchanged).  With ``rerank_pool`` > k the network
    also chooses the construction candidates among each node's rerank_pool
    nearest (rank_candidates), and eta is scaled on those pool edges.
    Returns dict with per-instance best-so-far records
    [(iteration, length, seconds of search)], the prior's application time
    (seconds, host wall clock), the iterations run and the search's wall
    time."""
    import time
    from .solver import GpuACO
    g = GpuACO(list(insts), n_ants=n_ants, k=k, beta=beta, selection=_lib.SEL_KNN,
               rng=_lib.RNG_PHILOX, ls_flags=ls_flags, tau0_rule=_lib.TAU0_NN_TOUR, seed=seed)
    try:
        t0 = time.perf_counter()
        if net is not None:
            if rerank_pool > k:
                for b, inst in enumerate(insts):
                    g.set_candidates(net.rank_candidates(inst, k, rerank_pool), b)
            expo = prior_exponent if prior_exponent is not None else beta
            if expo != 0:
                net.apply(g, k=rerank_pool if rerank_pool > k else k, exponent=expo, floor=floor)
        t_build = time.perf_counter() - t0
        g.init_aco()
        if autotune:
            g.autotune(reps=2)   # launch shape only; results and record clock unaffected
        g.synchronize()
        t0 = time.perf_counter()
        done = 0
        while done < iterations:
            step = min(chunk, iterations - done)
            g.run(step)
            done += step
            if targets is not None:
                if all(g.best(b)[1] <= targets[b] for b in range(len(insts))):
                    break
        g.synchronize()
        wall = time.perf_counter() - t0
        recs = [g.bsf_records(b) for b in range(len(insts))]
    finally:
        g.close()
    return {"records": recs, "build_s": t_build, "iterations": done, "wall_s": wall,
            "ant_solutions": len(insts) * n_ants * done}


def time_to_target(records, target, offset=0.0):
    for it, length, t in records:
        if length <= target:
            return t + offset
    return float("inf")


def best_at(records, budget, offset=0.0):
    best = float("inf")
    for it, length, t in records:
        if t + offset <= budget:
            best = length
    return best
