"""Instance set-up on the device (init_problem -> compute_distances /
compute_nn_lists, vrpHelper.cpp:160-236) through the C ABI.

* aco_compute_distances_batch equals the host routine aco_compute_distances
  element for element (EUC_2D float, nint, CEIL_2D, ATT) on random and
  dataset coordinates.
* The arrays aco_create derives on the device (eta^beta times the prior,
  depot column, candidate and neighbour rows with their distances, nnpos,
  16-bit rows) equal the same arrays built on the host from the oracle's
  heuristic and the host neighbour lists: uniform, clustered and dataset
  instances, float and integer distances, a prior, beta 1, 2 and 3 (beta 3
  is computed on the host), given lists and device-built lists, batched
  instances of uneven size.
"""
import math
import os

import numpy as np
import pytest

import oracle
import cvrp_aco_ocl_amd as ca
from cvrp_aco_ocl_amd import instances as I
from cvrp_aco_ocl_amd.solver import Instance

from helpers import VRPLIB, log

pytestmark = pytest.mark.gpu


def _coords(seed):
    r = np.random.default_rng(seed)
    n = int(r.integers(3, 400))
    if seed % 3 == 0:
        x, y = r.integers(0, 1001, n).astype(float), r.integers(0, 1001, n).astype(float)
    elif seed % 3 == 1:
        x, y = r.uniform(-500, 500, n), r.uniform(-500, 500, n)
    else:
        x, y = r.normal(0, 40, n).round(2), r.normal(0, 40, n).round(2)
    return x, y


@pytest.mark.parametrize("mode", [ca.DIST_FLOAT, ca.DIST_NINT, ca.DIST_CEIL, ca.DIST_ATT])
def test_gpu_distances_equal_host(mode):
    xs, ys = zip(*[_coords(s) for s in range(12)])
    files = [os.path.join(VRPLIB, f) for f in ("X-n101-k25.vrp", "CMT6.vrp", "Golden_1.vrp")]
    fx = [Instance.from_vrplib(f) for f in files]
    xs = list(xs) + [it.x for it in fx]
    ys = list(ys) + [it.y for it in fx]
    got = Instance.distances_batch(xs, ys, mode)
    for b, (x, y) in enumerate(zip(xs, ys)):
        ref = np.empty((len(x), len(x)), np.float32)
        ca._lib.check(ca._lib.lib.aco_compute_distances(len(x), np.ascontiguousarray(x).ctypes.data,
                                                       np.ascontiguousarray(y).ctypes.data, mode,
                                                       ref.ctypes.data), "aco_compute_distances")
        np.testing.assert_array_equal(got[b], ref, err_msg=f"coordinate set {b} mode {mode}")
    with pytest.raises(ca.AcoError):
        Instance.distances_batch(xs[:1], ys[:1], ca.DIST_GEO)


def _expected(inst, beta, k, nl, prior=None, nn=None):
    heur = oracle.heuristic(inst.dist, beta, prior)
    width = max(k, nl)
    lists = inst.nn_lists(width) if nn is None else np.ascontiguousarray(nn[:, :width])
    n = inst.n
    rows = np.repeat(np.arange(n), nl)
    nnpos = np.full((n, n), 0xFFFF, np.uint16)
    nnpos[rows, lists[:, :nl].ravel()] = np.tile(np.arange(nl, dtype=np.uint16), n)
    exp = {"heur": heur, "d0": inst.dist[:, 0].copy(), "nnk": lists[:, :k],
           "nnkd": np.take_along_axis(inst.dist, lists[:, :k].astype(np.int64), 1),
           "nnls": lists[:, :nl], "nnd": np.take_along_axis(inst.dist, lists[:, :nl].astype(np.int64), 1),
           "nnpos": nnpos}
    if np.all(inst.dist == np.round(inst.dist)) and inst.dist.max() <= 65535:
        exp["dist16"] = inst.dist.astype(np.uint16)
    return exp


CASES = [
    dict(name="u100", beta=2.0, k=20, nl=0),
    dict(name="x101", beta=2.0, k=16, nl=30),
    dict(name="cmt6", beta=2.0, k=12, nl=0),       # float distances: no 16-bit rows
    dict(name="u100", beta=1.0, k=8, nl=0),
    dict(name="u100", beta=3.0, k=8, nl=40),       # eta^3 on the host
    dict(name="clustered", beta=2.0, k=30, nl=0, prior=True),
    dict(name="u100", beta=2.0, k=20, nl=0, given=True),
    dict(name="u100", beta=2.0, k=0, nl=0),        # I-Roulette: no candidate rows
]


def _instance(name):
    if name == "u100":
        return I.uniform_100(31)
    if name == "x101":
        return Instance.from_vrplib(os.path.join(VRPLIB, "X-n101-k25.vrp"), ca.DIST_NINT)
    if name == "cmt6":
        return Instance.from_vrplib(os.path.join(VRPLIB, "CMT6.vrp"))
    return I.clustered(7, customers=300)


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(f"{k}{v}" for k, v in c.items()))
def test_device_setup_equals_host_arrays(case):
    inst = _instance(case["name"])
    k, nl, beta = case["k"], case["nl"], case["beta"]
    prior = None
    if case.get("prior"):
        prior = np.random.default_rng(3).uniform(0.05, 1.0, (inst.n, inst.n)).astype(np.float32)
    nn = inst.nn_lists(inst.n - 2, tie=1) if case.get("given") else None
    sel = ca.SEL_KNN if k else ca.SEL_IROULETTE
    rng = ca.RNG_PHILOX if k else ca.RNG_PARK_MILLER
    with ca.GpuACO(inst, n_ants=8, k=k, nn_ls=nl, beta=beta, selection=sel, rng=rng, ls_flags=7,
                   prior=prior, nn_list=nn) as g:
        got = g.setup_arrays()
    exp = _expected(inst, beta, k, nl if nl else inst.n - 2, prior, nn)
    for key, ref in exp.items():
        np.testing.assert_array_equal(got[key], ref, err_msg=key)
    log(f"{case}: {', '.join(exp)} equal")


def test_batched_setup_equals_host_arrays():
    insts = I.batch(11, count=12)
    k = 16
    with ca.GpuACO(insts, n_ants=8, k=k, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX, ls_flags=7) as g:
        for b, inst in enumerate(insts):
            got = g.setup_arrays(b)
            exp = _expected(inst, 2.0, k, inst.n - 2)
            for key, ref in exp.items():
                np.testing.assert_array_equal(got[key], ref, err_msg=f"instance {b} {key}")
