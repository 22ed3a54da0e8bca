"""C ABI, host side (no GPU): the library loads and exports every symbol
include/cvrp_aco.h declares; host-side preparation equals the oracle and
the reference build; invalid arguments are rejected before any device
work."""
import ctypes as C
import glob
import math
import os
import re

import numpy as np
import pytest

import oracle
import cvrp_aco_ocl_amd as ca
from cvrp_aco_ocl_amd import _lib, instances as I

from helpers import VRPLIB

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def _declared():
    names = set()
    for h in HEADERS:
        txt = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names.update(re.findall(r"\b(aco_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.EXPORTED)


def test_probe_library_exports_the_roof_microbenchmarks():
    """libcvrpprobe.so (bench.py's measured roofs) loads and exports its
    three entry points; they are only called on a GPU."""
    lib = C.CDLL(os.path.join(ROOT, "cvrp_aco_ocl_amd", "libcvrpprobe.so"))
    for name in ("probe_lds_bandwidth", "probe_hbm_copy", "probe_mfma_bf16"):
        assert hasattr(lib, name), name


def test_default_params_are_the_references():
    p = _lib.AcoParams()
    _lib.lib.aco_default_params(C.byref(p))
    assert (p.alpha, p.beta, p.rho, p.ras_ranks) == (1.0, 2.0, np.float32(0.1), 6)
    assert p.selection == ca.SEL_IROULETTE and p.tau0_rule == ca.TAU0_FIRST_COLONY


@pytest.mark.parametrize("mode", [ca.DIST_FLOAT, ca.DIST_NINT])
def test_distances_match_oracle(mode):
    r = np.random.default_rng(mode)
    x = r.uniform(0, 1000, 60)
    y = r.uniform(0, 1000, 60)
    inst = ca.Instance.from_coords(x, y, np.ones(60, np.int32), 10, mode)
    np.testing.assert_array_equal(inst.dist, oracle.distances(x, y, mode))


def test_special_distance_types():
    x = np.array([0.0, 3.0, 10.5])
    y = np.array([0.0, 4.0, 2.25])
    ceil = ca.Instance.from_coords(x, y, [0, 1, 1], 5, ca.DIST_CEIL).dist
    assert ceil[0, 1] == 5.0 and ceil[0, 2] == math.ceil(math.hypot(10.5, 2.25))
    att = ca.Instance.from_coords(x, y, [0, 1, 1], 5, ca.DIST_ATT).dist
    r = math.sqrt(25 / 10.0)
    assert att[0, 1] == (int(r) + 1 if int(r) < r else int(r))


@pytest.mark.parametrize("tie", [0, 1])
def test_nn_lists_match_oracle(tie):
    inst = I.uniform_100(9)
    np.testing.assert_array_equal(inst.nn_lists(40, tie), oracle.nn_lists(inst.dist, 40, tie))


def test_nn_lists_batch_rejects_bad_arguments_before_device_work():
    insts = [I.uniform_100(9), I.uniform_100(10)]
    with pytest.raises(ca.AcoError):
        ca.Instance.nn_lists_batch(insts, [40, insts[1].n - 1])      # wider than n-2
    with pytest.raises(ca.AcoError):
        ca.Instance.nn_lists_batch(insts, 40, tie=2)
    assert _lib.lib.aco_compute_nn_lists_batch(0, None, None, None, 0, None) == 0


def test_read_vrplib_fields():
    inst = ca.Instance.from_vrplib(os.path.join(VRPLIB, "CMT6.vrp"))
    assert inst.n == 51 and inst.capacity == 160
    assert inst.max_dist == 200.0 and inst.service_time == 10.0
    assert inst.optimum == np.float32(555.43)
    x = ca.Instance.from_vrplib(os.path.join(VRPLIB, "X-n101-k25.vrp"))
    assert x.n == 101 and x.capacity == 206 and x.max_dist == math.inf and x.demand[0] == 0


def test_read_vrplib_matches_reference_reader():
    ref = oracle.load_ref()
    if ref is None:
        pytest.skip("reference CPU build not present")
    for name in ("CMT1.vrp", "Golden_1.vrp", "X-n148-k46.vrp"):
        path = os.path.join(VRPLIB, name)
        inst = ca.Instance.from_vrplib(path)
        cap = 2000
        x = np.empty(cap, np.float32)
        y = np.empty(cap, np.float32)
        d = np.empty(cap, np.int32)
        Q, D, s, o = C.c_int(), C.c_float(), C.c_float(), C.c_float()
        n = ref.ref_read_instance(path.encode(), cap, x, y, d, C.byref(Q), C.byref(D), C.byref(s),
                                  C.byref(o))
        assert n == inst.n
        np.testing.assert_array_equal(x[:n], inst.x.astype(np.float32))
        np.testing.assert_array_equal(d[:n], inst.demand)
        assert (Q.value, D.value, s.value) == (inst.capacity, inst.max_dist, inst.service_time)


def test_read_vrplib_errors(tmp_path):
    with pytest.raises(ca.AcoError) as e:
        ca.Instance.from_vrplib(str(tmp_path / "missing.vrp"))
    assert e.value.code == -6
    bad = tmp_path / "bad.vrp"
    bad.write_text("NAME : t\nTYPE : TSP\nDIMENSION : 3\n")
    with pytest.raises(ca.AcoError) as e:
        ca.Instance.from_vrplib(str(bad))
    assert e.value.code == -6


def test_check_solution_matches_oracle():
    inst = I.uniform_100(3)
    P = oracle.Problem(inst.dist, inst.demand, inst.capacity)
    t, _ = P.nn_tour()
    assert inst.check_solution(t) and P.check(t)
    bad = t.copy()
    bad[2] = bad[1]
    assert not inst.check_solution(bad) and not P.check(bad)


def _create(inst, **kw):
    prob = _lib.AcoProblem(inst.n, inst.dist.ctypes.data, inst.demand.ctypes.data, inst.capacity,
                           inst.max_dist, inst.service_time, None, 0, None)
    p = _lib.AcoParams()
    _lib.lib.aco_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    h = C.c_void_p()
    return _lib.lib.aco_create(C.byref(prob), C.byref(p), C.byref(h)), h


@pytest.mark.parametrize("kw", [
    dict(ras_ranks=0), dict(ras_ranks=17), dict(rho=0.0), dict(rho=1.5),
    dict(selection=1, k=0), dict(selection=1, k=65), dict(selection=0, rng=1),
    dict(ls_flags=16), dict(ls_flags=9), dict(tau0_rule=3), dict(nn_ls=200), dict(selection=7),
])
def test_create_rejects_bad_params_before_device_work(kw):
    rc, h = _create(I.uniform_100(0), **kw)
    assert rc == -1 and not h.value
    assert _lib.lib.aco_last_error()


def test_create_rejects_unservable_customer():
    inst = I.uniform_100(0)
    inst.demand = inst.demand.copy()
    inst.demand[5] = inst.capacity + 1
    rc, h = _create(inst)
    assert rc == -4 and b"demand" in _lib.lib.aco_last_error()
    inst = ca.Instance.from_vrplib(os.path.join(VRPLIB, "CMT6.vrp"))
    inst.max_dist = 20.0
    rc, h = _create(inst)
    assert rc == -4 and b"max_dist" in _lib.lib.aco_last_error()


def test_create_rejects_asymmetric_or_negative_distances():
    inst = I.uniform_100(0)
    inst.dist = inst.dist.copy()
    inst.dist[1, 2] += 1
    assert _create(inst)[0] == -1
    inst.dist[1, 2] = -1
    inst.dist[2, 1] = -1
    assert _create(inst)[0] == -1


def test_generators_are_deterministic_and_sized():
    a, b = I.clustered(3), I.clustered(3)
    np.testing.assert_array_equal(a.dist, b.dist)
    assert a.n == 481 and a.capacity >= a.demand.max()
    assert abs(a.demand.sum() / a.capacity - 30) < 1
    m = I.mixed_1000(1)
    assert m.n == 1001 and m.x[0] == 500 and m.y[0] == 500
    u = I.uniform_100(0)
    assert u.n == 101 and u.capacity == 200 and 1 <= u.demand[1:].min() and u.demand.max() <= 30
    bs = I.batch(0, count=12)
    assert all(51 <= x.n <= 501 and x.capacity >= x.demand.max() for x in bs)
