"""The whole iteration (construction, local search, statistics, pheromone
update) on random instances against the oracle colony: both next-node rules
and random-number generators, candidate widths, tau0 rules, search flags,
route-length limits with service times, fp32 and 16-bit local-search rows
and random launch shapes (staged construction where it applies).  Tours,
lengths, tau, total_info and the best-so-far records must equal the
oracle's after every iteration (g_aco.cpp:122-134 as restated there)."""
import math

import numpy as np
import pytest

import oracle
import cvrp_aco_ocl_amd as ca
from cvrp_aco_ocl_amd.solver import Instance

from helpers import log, oracle_problem

pytestmark = pytest.mark.gpu

LS_SHAPES = [(4, 16, 1), (4, 16, 0), (4, 8, 0), (8, 32, 1), (8, 16, 0), (12, 32, 1), (16, 32, 0)]
RULES = [(ca.SEL_IROULETTE, ca.RNG_PARK_MILLER), (ca.SEL_KNN, ca.RNG_PHILOX), (ca.SEL_KNN, ca.RNG_PARK_MILLER)]


def _case(seed):
    r = np.random.default_rng(5000 + seed)
    n = int(r.integers(12, 400))
    if r.random() < 0.5:
        xy = r.integers(0, 1001, size=(n, 2)).astype(np.float64)
    else:
        c = r.uniform(0, 1000, size=(int(r.integers(2, 10)), 2))
        xy = np.clip(np.rint(c[r.integers(0, len(c), size=n)] +
                             r.normal(0, r.uniform(3, 40), size=(n, 2))), 0, 1000)
    dem = np.concatenate([[0], r.integers(1, 101, size=n - 1)]).astype(np.int32)
    Q = max(int(math.ceil(dem.sum() / int(r.integers(2, 40)))), int(dem.max()))
    mode = ca.DIST_NINT if r.random() < 0.7 else ca.DIST_FLOAT
    base = Instance.from_coords(xy[:, 0], xy[:, 1], dem, Q, mode)
    svc = float(r.integers(0, 10)) if r.random() < 0.3 else 0.0
    D = math.inf
    if r.random() < 0.3:
        alone = max(float(base.dist[0, i]) + svc + float(base.dist[i, 0]) for i in range(1, n))
        D = alone * float(r.uniform(1.05, 3.0))
    inst = Instance.from_coords(xy[:, 0], xy[:, 1], dem, Q, mode, max_dist=D, service_time=svc)
    sel, rng = RULES[int(r.integers(0, len(RULES)))]
    k = int(r.integers(5, min(40, n - 2) + 1)) if sel == ca.SEL_KNN else 0
    m = int(r.choice([16, 32, 48]))
    ls = int(r.choice([0, 1, 3, 7, 7]))
    tau0 = int(r.choice([ca.TAU0_FIRST_COLONY, ca.TAU0_NN_TOUR]))
    w, ring, pf = LS_SHAPES[int(r.integers(0, len(LS_SHAPES)))]
    int16 = bool(np.all(inst.dist == np.round(inst.dist)) and inst.dist.max() <= 65535)
    row16 = int(int16 and r.random() < 0.7)
    # staged candidate table: candidate roulette without a length limit
    apb = int(r.choice([0, 1, 2, 4])) if sel == ca.SEL_KNN and math.isinf(D) else 0
    return inst, dict(sel=sel, rng=rng, k=k, m=m, ls=ls, tau0=tau0, shape=(w, ring, pf, apb, row16))


@pytest.mark.parametrize("seed", range(16))
def test_random_colony_matches_oracle(seed):
    inst, c = _case(seed)
    log(f"random colony seed={seed} n={inst.n} Q={inst.capacity} D={inst.max_dist} {c}")
    iters = 4
    P = oracle_problem(inst, k=c["k"])
    col = oracle.Colony(P, c["m"], selection=c["sel"], rng=c["rng"], ls_flags=c["ls"],
                        tau0_rule=c["tau0"], seed=11 + seed)
    col.init()
    with ca.GpuACO(inst, n_ants=c["m"], k=c["k"], seed=11 + seed, selection=c["sel"], rng=c["rng"],
                   ls_flags=c["ls"], tau0_rule=c["tau0"], use_graph=True) as g:
        g.set_launch_shape(*c["shape"])
        g.init_aco()
        for it in range(iters):
            g.run_aco_iteration()
            col.iteration()
            gt, gs, gl = g.ants()
            ot, osz, ol = col.ants()
            np.testing.assert_array_equal(gs, osz, err_msg=f"iter {it}")
            for a in range(c["m"] + 1):
                np.testing.assert_array_equal(gt[a, : gs[a]], ot[a, : osz[a]], err_msg=f"iter {it} ant {a}")
            np.testing.assert_array_equal(gl, ol)
        gtau, gti = g.pheromone()
        otau, oti = col.tau()
        np.testing.assert_array_equal(gtau, otau)
        np.testing.assert_array_equal(gti, oti)
        assert [(r[0], r[1]) for r in g.bsf_records()] == [(r[0], r[1]) for r in col.bsf()]


def _batch(seed, fused):
    """3-6 random instances of one rule set (context-wide parameters)."""
    r = np.random.default_rng(9000 + seed)
    insts = []
    for b in range(int(r.integers(3, 7))):
        inst, _ = _case(100 * seed + b)
        if fused and not math.isinf(inst.max_dist):
            # the persistent kernel: no route-length limit
            inst = Instance.from_coords(inst.x, inst.y, inst.demand, inst.capacity, inst.dist_mode,
                                        service_time=inst.service_time)
        insts.append(inst)
    int16 = all(bool(np.all(x.dist == np.round(x.dist)) and x.dist.max() <= 65535) for x in insts)
    if fused:
        sel, rng = (ca.SEL_KNN, ca.RNG_PHILOX) if r.random() < 0.5 else (ca.SEL_KNN, ca.RNG_PARK_MILLER)
        w, ring, pf = [(4, 16, 1), (4, 16, 0), (8, 32, 1), (8, 16, 0)][int(r.integers(0, 4))]
        ls = int(r.choice([1, 3, 7]))
    else:
        sel, rng = RULES[int(r.integers(0, len(RULES)))]
        w, ring, pf = LS_SHAPES[int(r.integers(0, len(LS_SHAPES)))]
        ls = int(r.choice([0, 1, 7]))
    k = int(r.integers(5, min(30, min(x.n for x in insts) - 2) + 1)) if sel == ca.SEL_KNN else 0
    m = 16 if fused else int(r.choice([16, 32]))
    tau0 = int(r.choice([ca.TAU0_FIRST_COLONY, ca.TAU0_NN_TOUR]))
    row16 = int(int16 and r.random() < 0.7)
    return insts, dict(sel=sel, rng=rng, k=k, m=m, ls=ls, tau0=tau0,
                       shape=(w, ring, pf, 0, row16, 1 if fused else 0))


def _check_batch(insts, c, seed, iters, graph):
    with ca.GpuACO(insts, n_ants=c["m"], k=c["k"], seed=seed, selection=c["sel"], rng=c["rng"],
                   ls_flags=c["ls"], tau0_rule=c["tau0"], use_graph=graph) as g:
        g.set_launch_shape(*c["shape"])
        g.init_aco()
        g.run(iters)
        for b, inst in enumerate(insts):
            P = oracle_problem(inst, k=c["k"])
            col = oracle.Colony(P, c["m"], selection=c["sel"], rng=c["rng"], ls_flags=c["ls"],
                                tau0_rule=c["tau0"], seed=seed + b)
            col.init()
            for _ in range(iters):
                col.iteration()
            gt, gs, gl = g.ants(b)
            ot, osz, ol = col.ants()
            np.testing.assert_array_equal(gs, osz, err_msg=f"instance {b}")
            for a in range(c["m"] + 1):
                np.testing.assert_array_equal(gt[a, : gs[a]], ot[a, : osz[a]], err_msg=f"instance {b} ant {a}")
            np.testing.assert_array_equal(gl, ol)
            gtau, gti = g.pheromone(b)
            otau, oti = col.tau()
            np.testing.assert_array_equal(gtau, otau)
            np.testing.assert_array_equal(gti, oti)
            assert [r[:2] for r in g.bsf_records(b)] == [r[:2] for r in col.bsf()]


@pytest.mark.parametrize("seed", range(6))
def test_random_batch_matches_oracle(seed):
    """Instances of uneven size in one context, each equal to its own oracle
    colony seeded with seed + b."""
    insts, c = _batch(seed, fused=False)
    log(f"random batch seed={seed} n={[x.n for x in insts]} {c}")
    # eager launches and graph replay (size classes on forked streams in both)
    _check_batch(insts, c, 21 + seed, 3, graph=bool(seed % 2))


@pytest.mark.parametrize("seed", range(4))
def test_random_persistent_kernel_matches_oracle(seed):
    """The one-launch iteration (persistent kernel) on random batches, each
    instance equal to its own oracle colony."""
    insts, c = _batch(seed, fused=True)
    log(f"random persistent batch seed={seed} n={[x.n for x in insts]} {c}")
    _check_batch(insts, c, 31 + seed, 3, graph=bool(seed % 2))
