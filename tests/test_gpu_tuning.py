"""Launch shapes and the autotuner: every construction and local-search
shape gives the same tours, lengths, fallback and evaluated-move counts,
and autotuning leaves the colony's trajectory unchanged."""
import numpy as np
import pytest

import cvrp_aco_ocl_amd as ca
from cvrp_aco_ocl_amd import instances as I

from helpers import log

pytestmark = pytest.mark.gpu

SHAPES = [(4, 16, 1, 0, 0), (4, 16, 0, 1, 0), (4, 8, 0, 2, 0), (8, 32, 1, 4, 0), (8, 16, 0, 0, 0),
          (8, 16, 0, 2, 1), (4, 16, 1, 0, 1), (8, 32, 1, 2, 1), (12, 16, 0, 0, 0), (12, 32, 1, 4, 1),
          (16, 32, 0, 0, 1), (16, 32, 0, 2, 0)]


def _small():
    return [I.uniform_100(1), I.clustered(2, customers=480), I.batch(4, count=2)[1]]


def _large():
    # n = 1101: past the 4-wave prefetch limit (n <= 1024), so those ants
    # run without prefetch while the 100-node instance prefetches
    return [I.clustered(3, customers=1100, routes=80), I.uniform_100(2)]


def _ls_once(insts, shape, k=16, rng=ca.RNG_PHILOX):
    with ca.GpuACO(insts, n_ants=12, k=k, seed=7, selection=ca.SEL_KNN, rng=rng,
                   ls_flags=7, use_graph=False) as g:
        g.init_aco()
        if shape is not None:
            g.set_launch_shape(*shape)
        g.construct_solutions()
        g.local_search()
        out = [g.ants(b) for b in range(len(insts))]
        g.update_statistics()
        c = g.counters()
        return out, (c["ls_evals"], c["fallbacks"])


@pytest.mark.parametrize("which", ["small", "large"])
def test_every_shape_gives_identical_results(which):
    insts = _small() if which == "small" else _large()
    ref, ref_ev = _ls_once(insts, SHAPES[0])
    assert ref_ev[0] > 0
    for sh in SHAPES[1:]:
        if which == "large" and sh[1] == 32 and not sh[4]:
            # 32 rows of 1101 floats do not fit in 160 KB of LDS
            with pytest.raises(ca.AcoError):
                _ls_once(insts, sh)
            continue
        log(f"{which} shape {sh}")
        got, ev = _ls_once(insts, sh)
        assert ev == ref_ev, (sh, ev, ref_ev)
        for b in range(len(insts)):
            for x, y in zip(got[b], ref[b]):
                np.testing.assert_array_equal(x, y, err_msg=f"shape {sh} instance {b}")


@pytest.mark.parametrize("rng", [ca.RNG_PHILOX, ca.RNG_PARK_MILLER])
def test_staged_construction_with_fallbacks(rng):
    """k = 4 candidates leave many steps without mass (full-row fallback)."""
    insts = [I.uniform_100(3), I.batch(6, count=1)[0]]
    ref, ref_c = _ls_once(insts, (4, 16, 1, 0), k=4, rng=rng)
    assert ref_c[1] > 0
    for apb in (1, 2, 4):
        got, c = _ls_once(insts, (4, 16, 1, apb), k=4, rng=rng)
        assert c == ref_c, (apb, c, ref_c)
        for b in range(len(insts)):
            for x, y in zip(got[b], ref[b]):
                np.testing.assert_array_equal(x, y, err_msg=f"apb {apb} instance {b}")


def test_invalid_shapes_rejected():
    inst = I.uniform_100(1)
    with ca.GpuACO(inst, n_ants=8, ls_flags=7) as g:   # I-Roulette: no staged construction
        assert g.launch_shape() == (4, 16, 1, 0, 0, 0)
        for bad in [(2, 16, 0), (4, 4, 0), (4, 8, 1), (8, 8, 0), (8, 16, 1), (4, 12, 0),
                    (6, 16, 0), (12, 8, 0), (12, 16, 1), (16, 16, 0), (16, 32, 1),
                    (4, 16, 1, 1), (4, 16, 1, 3)]:
            with pytest.raises(ca.AcoError):
                g.set_launch_shape(*bad)
        assert g.launch_shape() == (4, 16, 1, 0, 0, 0)
        g.set_launch_shape(4, 16, 1, 0, 1)       # nint distances: 16-bit rows allowed
    fl = I.uniform_100(1, dist_mode=ca.DIST_FLOAT)
    with ca.GpuACO(fl, n_ants=8, ls_flags=7) as g:
        with pytest.raises(ca.AcoError):         # float distances: no 16-bit rows
            g.set_launch_shape(4, 16, 1, 0, 1)
    with ca.GpuACO(inst, n_ants=6, k=16, selection=ca.SEL_KNN, ls_flags=7) as g:
        g.set_launch_shape(4, 16, 1, 2)
        with pytest.raises(ca.AcoError):     # 6 ants are not a multiple of 4
            g.set_launch_shape(4, 16, 1, 4)


def test_autotune_keeps_the_trajectory():
    insts = I.batch(5, count=6)
    runs = []
    for tune in (False, True):
        with ca.GpuACO(insts, n_ants=16, k=16, seed=3, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX,
                       ls_flags=7) as g:
            g.init_aco()
            if tune:
                res = g.autotune(reps=2)
                log(f"autotune: {res} -> {g.launch_shape()}")
                assert any(ms > 0 for _, st, ms in res if st == 0)
                assert any(ms > 0 for _, st, ms in res if st == 1)
                chosen = g.launch_shape()
                if chosen[5]:
                    assert chosen in [sh for sh, st, ms in res if st == 2 and ms > 0]
                else:
                    assert (chosen[:3] + chosen[4:5]) in \
                        [sh[:3] + sh[4:5] for sh, st, ms in res if st == 1 and ms > 0]
                    assert chosen[3] in [sh[3] for sh, st, ms in res if st == 0 and ms > 0]
                assert any(ms > 0 for _, st, ms in res if st == 2)
            g.run(4)
            runs.append([(g.ants(b), g.pheromone(b), g.bsf_records(b)) for b in range(len(insts))])
    for b in range(len(insts)):
        (a0, p0, r0), (a1, p1, r1) = runs[0][b], runs[1][b]
        for x, y in zip(a0, a1):
            np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(p0[0], p1[0])
        assert [r[:2] for r in r0] == [r[:2] for r in r1]


def test_candidate_k_tuning_by_quality():
    from cvrp_aco_ocl_amd.tuning import tune_candidate_k
    insts = I.batch(8, count=4)
    best, res = tune_candidate_k(insts, ks=(8, 16), seconds=0.2, n_ants=16)
    log(f"k tuning: {res} -> {best}")
    assert set(res) == {8, 16} and best in res
    assert all(v["iterations"] > 0 and v["mean_cost"] > 0 for v in res.values())
    assert res[best]["mean_cost"] == min(v["mean_cost"] for v in res.values())
