"""Local search on random instances against the oracle: sizes 12..700,
uniform and clustered points, capacities giving 2..90 routes (the
relocate capacity bits and their fallback past 64 routes), route-length
limits with service times, fp32 and 16-bit distance rows, and every
local-search launch shape.  Tours, lengths and the evaluated-move count
must equal the oracle's (localSearch.cpp, neighbourSearch.cpp,
kernel.cl:126-319 as restated there)."""
import math

import numpy as np
import pytest

import oracle
import cvrp_aco_ocl_amd as ca
from cvrp_aco_ocl_amd.solver import Instance

from helpers import log, oracle_construct_all, oracle_problem, pack_tours

pytestmark = pytest.mark.gpu

SHAPES = [(4, 16, 1), (4, 16, 0), (4, 8, 0), (8, 32, 1), (8, 16, 0), (12, 16, 0), (12, 32, 1),
          (16, 32, 0)]
ANTS = 32


def _case(seed):
    r = np.random.default_rng(1000 + seed)
    n = int(r.integers(12, 700))
    if r.random() < 0.5:
        xy = r.integers(0, 1001, size=(n, 2)).astype(np.float64)
    else:
        c = r.uniform(0, 1000, size=(int(r.integers(2, 12)), 2))
        xy = np.clip(np.rint(c[r.integers(0, len(c), size=n)] +
                             r.normal(0, r.uniform(3, 40), size=(n, 2))), 0, 1000)
    dem = np.concatenate([[0], r.integers(1, 101, size=n - 1)]).astype(np.int32)
    Q = max(int(math.ceil(dem.sum() / int(r.integers(2, 90)))), int(dem.max()))
    mode = ca.DIST_NINT if r.random() < 0.7 else ca.DIST_FLOAT
    base = Instance.from_coords(xy[:, 0], xy[:, 1], dem, Q, mode)
    svc = float(r.integers(0, 10)) if r.random() < 0.3 else 0.0
    D = math.inf
    if r.random() < 0.3:
        # every customer servable alone, then some slack
        alone = max(float(base.dist[0, i]) + svc + float(base.dist[i, 0]) for i in range(1, n))
        D = alone * float(r.uniform(1.05, 3.0))
    inst = Instance.from_coords(xy[:, 0], xy[:, 1], dem, Q, mode, max_dist=D, service_time=svc)
    flags = int(r.choice([1, 2, 4, 3, 6, 7, 7, 7]))
    shape = SHAPES[int(r.integers(0, len(SHAPES)))]
    int16 = bool(np.all(inst.dist == np.round(inst.dist)) and inst.dist.max() <= 65535)
    row16 = int(int16 and r.random() < 0.7)
    return inst, flags, shape + (row16,)


@pytest.mark.parametrize("seed", range(24))
def test_random_local_search_matches_oracle(seed):
    inst, flags, (w, ring, pf, row16) = _case(seed)
    log(f"random local search seed={seed} n={inst.n} Q={inst.capacity} D={inst.max_dist} "
        f"flags={flags} shape={(w, ring, pf, row16)}")
    P = oracle_problem(inst)
    tau, ti = oracle.pheromone_init(P, 0.5)
    pm = np.array([[oracle.pm_seed(7 + seed, a, l) for l in range(64)] for a in range(ANTS)], np.int32)
    otours, olens, _, _ = oracle_construct_all(P, ti, 0, 0, 0, 0, pm)
    packed, sizes = pack_tours(otours + [otours[0]], 2 * inst.n)
    lens = np.concatenate([olens, olens[:1]]).astype(np.float32)
    with ca.GpuACO(inst, n_ants=ANTS, seed=1, ls_flags=flags, use_graph=False) as g:
        g.init_aco()
        g.set_launch_shape(w, ring, pf, 0, row16)
        g.set_ants(packed, sizes, lens)
        g.local_search()
        tours, gsizes, glens = g.ants()
        g.update_statistics()   # folds the per-ant counters into the totals
        cnt = g.counters()
    evals = 0
    for a in range(ANTS):
        t, L, ev = P.local_search(otours[a], flags)
        evals += ev
        assert gsizes[a] == len(t), a
        np.testing.assert_array_equal(tours[a, : len(t)], t, err_msg=f"ant {a}")
        assert glens[a] == np.float32(L), (a, glens[a], L)
    assert cnt["ls_evals"] == evals, (cnt["ls_evals"], evals)
