"""Shared test helpers: build the oracle's view of an instance and run the
oracle on the same inputs the HIP path sees."""
import os

import numpy as np

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
VRPLIB = os.path.join(GOLDEN, "vrplib")


def oracle_problem(inst, k=0, nn_ls=0, beta=2.0, prior=None):
    return oracle.Problem(inst.dist, inst.demand, inst.capacity, inst.max_dist,
                          inst.service_time, beta=beta, k=k, nn_ls=nn_ls, tie=0, prior=prior)


def oracle_construct_all(P, tinfo, selection, rng, seed, iteration, pm_state):
    """All ants on the oracle; pm_state (m,64) is consumed like the device."""
    tiknn = oracle.gather_knn(P, tinfo) if P.k else None
    pm = np.ascontiguousarray(pm_state, np.int32).copy()
    tours, lens, fbs = [], [], 0
    for a in range(pm.shape[0]):
        row = pm[a].copy()
        t, L, fb = P.construct_ant(tinfo, tiknn, selection, rng, seed, iteration, a, row)
        pm[a] = row
        tours.append(t)
        lens.append(L)
        fbs += fb
    return tours, np.array(lens, np.float32), pm, fbs


def random_tau(n, seed, lo=1e-4, hi=1e-2):
    r = np.random.default_rng(seed)
    return r.uniform(lo, hi, size=(n, n)).astype(np.float32)


def pack_tours(tours, T):
    out = np.zeros((len(tours), T), np.int32)
    sizes = np.zeros(len(tours), np.int32)
    for i, t in enumerate(tours):
        out[i, : len(t)] = t
        sizes[i] = len(t)
    return out, sizes


def log(msg):
    print(msg, flush=True)
