"""The fused iteration (one launch per iteration: work queue of W ants, the
block finishing an instance runs its statistics and pheromone pass) gives
the same colony as the per-phase launches: tours, pheromone, best-so-far
records and counters, on batches of uneven instances."""
import numpy as np
import pytest

import cvrp_aco_ocl_amd as ca
from cvrp_aco_ocl_amd import instances as I

from helpers import log

pytestmark = pytest.mark.gpu


def _run(insts, shape, iters, graph, rng=ca.RNG_PHILOX, ants=16):
    with ca.GpuACO(insts, n_ants=ants, k=12, seed=9, selection=ca.SEL_KNN, rng=rng, ls_flags=7,
                   use_graph=graph) as g:
        g.init_aco()
        if shape is not None:
            g.set_launch_shape(*shape)
        g.run(iters)
        out = [(g.ants(b), g.pheromone(b), g.bsf_records(b)) for b in range(len(insts))]
        return out, g.counters()


@pytest.mark.parametrize("shape,rng,graph", [
    ((8, 16, 0, 0, 1, 1), ca.RNG_PHILOX, True),
    ((4, 16, 1, 0, 0, 1), ca.RNG_PHILOX, False),
    ((8, 32, 1, 0, 0, 1), ca.RNG_PARK_MILLER, True),
])
def test_fused_equals_per_phase(shape, rng, graph):
    insts = I.batch(21, count=5) + [I.uniform_100(5)]
    ref, c_ref = _run(insts, shape[:5] + (0,), 6, graph, rng)
    got, c_got = _run(insts, shape, 6, graph, rng)
    log(f"shape {shape}: evals {c_got['ls_evals']} fallbacks {c_got['fallbacks']}")
    assert c_got == c_ref
    for b in range(len(insts)):
        (a0, p0, r0), (a1, p1, r1) = ref[b], got[b]
        for x, y in zip(a0, a1):
            np.testing.assert_array_equal(x, y, err_msg=f"instance {b}")
        np.testing.assert_array_equal(p0[0], p1[0])
        np.testing.assert_array_equal(p0[1], p1[1])
        assert [r[:2] for r in r0] == [r[:2] for r in r1]


def test_fused_rejected_where_it_does_not_apply():
    inst = I.uniform_100(1)
    with ca.GpuACO(inst, n_ants=12, k=12, selection=ca.SEL_KNN, ls_flags=7) as g:
        with pytest.raises(ca.AcoError):          # 12 ants: not a multiple of 8
            g.set_launch_shape(8, 16, 0, 0, 0, 1)
        g.set_launch_shape(4, 16, 1, 0, 0, 1)     # 12 is a multiple of 4
    with ca.GpuACO(inst, n_ants=8, ls_flags=7) as g:   # I-Roulette
        with pytest.raises(ca.AcoError):
            g.set_launch_shape(4, 16, 1, 0, 0, 1)
    with ca.GpuACO(inst, n_ants=8, k=12, selection=ca.SEL_KNN, ls_flags=0) as g:
        with pytest.raises(ca.AcoError):          # no local search
            g.set_launch_shape(4, 16, 1, 0, 0, 1)
