"""aco_compute_nn_lists_batch_noalloc (device arrays in and out, library
scratch, no allocation) equals aco_compute_nn_lists_batch, on two calls in a
row -- the second with a smaller input -- through the C ABI."""
import ctypes as C

import numpy as np
import pytest

import cvrp_aco_ocl_amd as ca
from cvrp_aco_ocl_amd import _lib, instances as I

from helpers import log

pytestmark = pytest.mark.gpu


def _uniform(customers, seed):
    r = np.random.default_rng(seed)
    x = np.concatenate([[500.0], r.integers(0, 1001, customers).astype(float)])
    y = np.concatenate([[500.0], r.integers(0, 1001, customers).astype(float)])
    dem = np.concatenate([[0], r.integers(1, 30, customers)]).astype(np.int32)
    return ca.Instance.from_coords(x, y, dem, 200, ca.DIST_NINT)


class _Hip:
    """The HIP runtime the library itself is linked against (symbols looked
    up through the library's handle, which searches its dependencies):
    device buffers and a stream for the caller's side of the no-alloc
    call."""

    def __init__(self):
        h = C.CDLL(_lib.LIB_PATH)
        for name, args in (("hipMalloc", [C.POINTER(C.c_void_p), C.c_size_t]),
                           ("hipFree", [C.c_void_p]),
                           ("hipMemcpy", [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
                           ("hipMemset", [C.c_void_p, C.c_int, C.c_size_t]),
                           ("hipStreamCreate", [C.POINTER(C.c_void_p)]),
                           ("hipStreamDestroy", [C.c_void_p]),
                           ("hipStreamSynchronize", [C.c_void_p])):
            getattr(h, name).argtypes = args
            getattr(h, name).restype = C.c_int
        self.h = h

    def ok(self, rc, what):
        assert rc == 0, f"{what}: hip error {rc}"

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.ok(self.h.hipMalloc(C.byref(p), a.nbytes), "hipMalloc")
        self.ok(self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p

    def filled(self, nbytes):
        p = C.c_void_p()
        self.ok(self.h.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        self.ok(self.h.hipMemset(p, 0xFF, nbytes), "hipMemset")
        return p

    def download(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        self.ok(self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
        return out


def _noalloc(insts, widths, tie, use_stream):
    hip = _Hip()
    dd = [hip.upload(np.ascontiguousarray(it.dist, np.float32)) for it in insts]
    oo = [hip.filled(it.n * w * 4) for it, w in zip(insts, widths)]
    ns = np.array([it.n for it in insts], np.int32)
    ws = np.array(widths, np.int32)
    dptr = (C.c_void_p * len(insts))(*[d.value for d in dd])
    optr = (C.c_void_p * len(insts))(*[o.value for o in oo])
    stream = C.c_void_p()
    if use_stream:
        hip.ok(hip.h.hipStreamCreate(C.byref(stream)), "hipStreamCreate")
    try:
        rc = _lib.lib.aco_compute_nn_lists_batch_noalloc(
            len(insts), ns.ctypes.data, C.cast(dptr, C.c_void_p), ws.ctypes.data, tie,
            C.cast(optr, C.c_void_p), stream)
        _lib.check(rc, "aco_compute_nn_lists_batch_noalloc")
        hip.ok(hip.h.hipStreamSynchronize(stream), "hipStreamSynchronize")
        return [hip.download(o, (it.n, w), np.int32) for o, it, w in zip(oo, insts, widths)]
    finally:
        if use_stream:
            hip.h.hipStreamDestroy(stream)
        for p in dd + oo:
            hip.h.hipFree(p)


@pytest.mark.parametrize("tie", [0, 1])
def test_noalloc_equals_allocating_builder_on_two_calls(tie):
    # call 1: 41 instances (more than one launch's job table) including one
    # whose rows need more than the whole scratch (n = 3001: 72 MB)
    first = I.batch(3, count=40) + [_uniform(3000, 11)]
    w1 = [min(24, it.n - 2) if b % 2 else it.n - 2 for b, it in enumerate(first)]
    log(f"call 1: {len(first)} instances, n {min(i.n for i in first)}..{max(i.n for i in first)}, tie {tie}")
    got = _noalloc(first, w1, tie, use_stream=True)
    ref = ca.Instance.nn_lists_batch(first, w1, tie)
    for b in range(len(first)):
        np.testing.assert_array_equal(got[b], ref[b], err_msg=f"call 1, instance {b}")
    # call 2: smaller -- 3 instances, the scratch keeps call 1's contents
    second = [I.uniform_100(5), I.clustered(2, customers=150), _uniform(60, 3)]
    w2 = [20, 40, 58]
    log(f"call 2: {len(second)} instances")
    got = _noalloc(second, w2, tie, use_stream=False)
    for b, it in enumerate(second):
        np.testing.assert_array_equal(got[b], it.nn_lists(w2[b], tie), err_msg=f"call 2, instance {b}")


def test_noalloc_rejects_bad_sizes():
    it = I.uniform_100(1)
    ns = np.array([it.n], np.int32)
    ws = np.array([it.n - 1], np.int32)      # wider than n - 2
    ptrs = (C.c_void_p * 1)(C.c_void_p(1))
    rc = _lib.lib.aco_compute_nn_lists_batch_noalloc(1, ns.ctypes.data, C.cast(ptrs, C.c_void_p),
                                                     ws.ctypes.data, 0, C.cast(ptrs, C.c_void_p), None)
    assert rc == -1 and b"nn <= n-2" in _lib.lib.aco_last_error()
    ns = np.array([5000], np.int32)
    ws = np.array([10], np.int32)
    rc = _lib.lib.aco_compute_nn_lists_batch_noalloc(1, ns.ctypes.data, C.cast(ptrs, C.c_void_p),
                                                     ws.ctypes.data, 0, C.cast(ptrs, C.c_void_p), None)
    assert rc == -1
