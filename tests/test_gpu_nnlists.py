"""Candidate lists built on the GPU (aco_compute_nn_lists_batch) equal the
host builder aco_compute_nn_lists -- itself equal to the reference build's
compute_nn_lists (vrpHelper.cpp:188-236) -- row for row, in both tie orders,
on instances with many equal distances."""
import numpy as np
import pytest

import cvrp_aco_ocl_amd as ca
from cvrp_aco_ocl_amd import instances as I

from helpers import log

pytestmark = pytest.mark.gpu


def _tied(seed, customers, span):
    """Integer coordinates on a small grid: many equal nint distances."""
    r = np.random.default_rng(seed)
    x = r.integers(0, span, size=customers + 1).astype(np.float64)
    y = r.integers(0, span, size=customers + 1).astype(np.float64)
    dem = np.concatenate([[0], r.integers(1, 10, size=customers)]).astype(np.int32)
    return ca.Instance.from_coords(x, y, dem, 100, ca.DIST_NINT)


def _cases():
    return ([_tied(s, c, span) for s, (c, span) in enumerate([(30, 6), (120, 12), (400, 20), (999, 40)])]
            + [I.uniform_100(3), I.clustered(1, customers=300)] + I.batch(11, count=6))


@pytest.mark.parametrize("tie", [0, 1])
def test_batch_lists_equal_host_lists(tie):
    insts = _cases()
    widths = [max(1, min(it.n - 2, w)) for it, w in zip(insts, [it.n - 2 for it in insts])]
    got = ca.Instance.nn_lists_batch(insts, widths, tie)
    for b, (it, w) in enumerate(zip(insts, widths)):
        ref = it.nn_lists(w, tie)
        np.testing.assert_array_equal(got[b], ref, err_msg=f"instance {b} n={it.n} tie={tie}")
    log(f"tie {tie}: {len(insts)} instances, n {min(i.n for i in insts)}..{max(i.n for i in insts)}")


def test_batch_lists_narrow_widths_and_errors():
    insts = _cases()[:5]
    got = ca.Instance.nn_lists_batch(insts, 7)
    for b, it in enumerate(insts):
        np.testing.assert_array_equal(got[b], it.nn_lists(7))
    with pytest.raises(ca.AcoError):
        ca.Instance.nn_lists_batch(insts[:1], insts[0].n - 1)   # wider than n-2


def test_context_without_lists_equals_context_with_host_lists():
    """aco_create builds missing lists on the device: same colony either way."""
    insts = [_tied(7, 60, 8), I.uniform_100(4)]
    runs = []
    for given in (False, True):
        lists = [it.nn_lists(it.n - 2) for it in insts] if given else None
        with ca.GpuACO(insts, n_ants=16, k=12, seed=5, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX,
                       ls_flags=7, nn_list=lists) as g:
            g.init_aco()
            g.run(5)
            runs.append([g.ants(b) for b in range(len(insts))])
    for b in range(len(insts)):
        for x, y in zip(runs[0][b], runs[1][b]):
            np.testing.assert_array_equal(x, y)


def test_batch_lists_at_the_node_limit():
    """4096 nodes (the library's maximum; two rows per block of the LDS
    sort) with many equal distances, next to a small instance in the same
    launch."""
    insts = [_tied(21, 4095, 64), _tied(22, 40, 6)]
    widths = [it.n - 2 for it in insts]
    got = ca.Instance.nn_lists_batch(insts, widths, 0)
    for b, (it, w) in enumerate(zip(insts, widths)):
        np.testing.assert_array_equal(got[b], it.nn_lists(w, 0), err_msg=f"instance {b} n={it.n}")
