"""Synthetic CVRP instance generators and VRPLIB solution files.

Generators for the benchmark workloads, all seeded (numpy PCG64) so every
run sees the same instances:

* ``uniform_100``   100 customers, integer coordinates U{0..100}^2, depot
                    (50,50), demand U{1..30}, Q = 200.
* ``clustered``     customers in Gaussian clusters (sigma 6) whose centres
                    are U[0,1000]^2, coordinates rounded and clipped to
                    [0,1000], depot U{0..1000}^2, demand U{1..100},
                    Q = ceil(sum(d) / routes).  480 customers, 8 clusters and
                    30 routes is the 481-node workload.
* ``mixed_1000``    500 uniform + 500 clustered customers (10 clusters,
                    sigma 25) in [0,1000]^2, depot (500,500), demand
                    U{1..100} with 5% of customers redrawn U{80..100},
                    Q = ceil(12 * mean(d)) (about 12 customers per route).
* ``batch``         ``count`` instances with n-1 ~ U{50..500}, uniform or
                    4-12 Gaussian clusters, demand U{1..50}, routes
                    U{5..40}, Q = max(ceil(sum(d)/routes), max(d)).
* ``training``      instance ``index`` of the learned prior's training set:
                    n-1 ~ U{100..500}; half Gaussian clusters (4-12
                    centres, sigma U[6,40]), a quarter uniform, a quarter
                    half/half; depot U{0..1000}^2; demand U{1..100}; routes
                    U{10..40}.
* ``evaluation``    the prior's evaluation set: ``clustered`` instances
                    (500 customers, 8 clusters, sigma 6, 30 routes) with
                    seeds 100000 + i, disjoint from every other generator's.

Distances are VRPLIB nint integers unless a caller asks otherwise.
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib
from .solver import Instance


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def _clusters(r, count, nclust, sigma, lo=0.0, hi=1000.0):
    centres = r.uniform(lo, hi, size=(nclust, 2))
    which = r.integers(0, nclust, size=count)
    pts = centres[which] + r.normal(0.0, sigma, size=(count, 2))
    return np.clip(np.rint(pts), lo, hi)


def uniform_100(seed: int = 0, dist_mode=_lib.DIST_NINT) -> Instance:
    r = _rng(seed)
    xy = r.integers(0, 101, size=(100, 2)).astype(np.float64)
    x = np.concatenate([[50.0], xy[:, 0]])
    y = np.concatenate([[50.0], xy[:, 1]])
    dem = np.concatenate([[0], r.integers(1, 31, size=100)]).astype(np.int32)
    return Instance.from_coords(x, y, dem, 200, dist_mode, name=f"uniform100-s{seed}")


def clustered(seed: int = 0, customers: int = 480, nclust: int = 8, sigma: float = 6.0,
              routes: int = 30, dist_mode=_lib.DIST_NINT) -> Instance:
    r = _rng(seed)
    pts = _clusters(r, customers, nclust, sigma)
    depot = r.integers(0, 1001, size=2).astype(np.float64)
    x = np.concatenate([[depot[0]], pts[:, 0]])
    y = np.concatenate([[depot[1]], pts[:, 1]])
    dem = np.concatenate([[0], r.integers(1, 101, size=customers)]).astype(np.int32)
    Q = int(math.ceil(dem.sum() / routes))
    Q = max(Q, int(dem.max()))
    return Instance.from_coords(x, y, dem, Q, dist_mode,
                                name=f"clustered{customers + 1}-s{seed}")


def mixed_1000(seed: int = 0, dist_mode=_lib.DIST_NINT) -> Instance:
    r = _rng(seed)
    uni = r.integers(0, 1001, size=(500, 2)).astype(np.float64)
    clu = _clusters(r, 500, 10, 25.0)
    pts = np.concatenate([uni, clu])
    x = np.concatenate([[500.0], pts[:, 0]])
    y = np.concatenate([[500.0], pts[:, 1]])
    dem = r.integers(1, 101, size=1000)
    heavy = r.random(1000) < 0.05
    dem[heavy] = r.integers(80, 101, size=int(heavy.sum()))
    dem = np.concatenate([[0], dem]).astype(np.int32)
    Q = int(math.ceil(12 * dem[1:].mean()))
    return Instance.from_coords(x, y, dem, Q, dist_mode, name=f"mixed1001-s{seed}")


def batch(seed: int = 0, count: int = 256, dist_mode=_lib.DIST_NINT) -> list[Instance]:
    r = _rng(seed)
    out = []
    for b in range(count):
        customers = int(r.integers(50, 501))
        if r.random() < 0.5:
            pts = r.integers(0, 1001, size=(customers, 2)).astype(np.float64)
            geom = "uniform"
        else:
            nc = int(r.integers(4, 13))
            pts = _clusters(r, customers, nc, float(r.uniform(10.0, 60.0)))
            geom = f"clusters{nc}"
        depot = r.integers(0, 1001, size=2).astype(np.float64)
        x = np.concatenate([[depot[0]], pts[:, 0]])
        y = np.concatenate([[depot[1]], pts[:, 1]])
        dem = np.concatenate([[0], r.integers(1, 51, size=customers)]).astype(np.int32)
        routes = int(r.integers(5, 41))
        Q = max(int(math.ceil(dem.sum() / routes)), int(dem.max()))
        out.append(Instance.from_coords(x, y, dem, Q, dist_mode, name=f"batch{b}-{geom}",
                                        meta={"routes": routes, "geometry": geom}))
    return out


def training(index: int, seed: int = 0, dist_mode=_lib.DIST_NINT) -> Instance:
    r = _rng((seed << 32) + 7919 * index + 17)
    customers = int(r.integers(100, 501))
    u = r.random()
    if u < 0.5:
        nc = int(r.integers(4, 13))
        pts = _clusters(r, customers, nc, float(r.uniform(6.0, 40.0)))
        geom = f"clusters{nc}"
    elif u < 0.75:
        pts = r.integers(0, 1001, size=(customers, 2)).astype(np.float64)
        geom = "uniform"
    else:
        half = customers // 2
        nc = int(r.integers(4, 13))
        pts = np.concatenate([r.integers(0, 1001, size=(half, 2)).astype(np.float64),
                              _clusters(r, customers - half, nc, float(r.uniform(6.0, 40.0)))])
        geom = f"mixed{nc}"
    depot = r.integers(0, 1001, size=2).astype(np.float64)
    x = np.concatenate([[depot[0]], pts[:, 0]])
    y = np.concatenate([[depot[1]], pts[:, 1]])
    dem = np.concatenate([[0], r.integers(1, 101, size=customers)]).astype(np.int32)
    routes = int(r.integers(10, 41))
    Q = max(int(math.ceil(dem.sum() / routes)), int(dem.max()))
    return Instance.from_coords(x, y, dem, Q, dist_mode, name=f"train{index}-{geom}",
                                meta={"routes": routes, "geometry": geom})


def evaluation(count: int = 64, first: int = 0) -> list[Instance]:
    return [clustered(100000 + i, customers=500, nclust=8, sigma=6.0, routes=30)
            for i in range(first, first + count)]


def read_solution(path: str):
    """A VRPLIB ``.sol`` file of the reference's dataset: total cost, route
    count, then one line per route ``... cost load stops 0 c1 c2 ... 0``.
    Returns (total, routes) with routes as lists of node ids (depot 0 at
    both ends)."""
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip()]
    total = float(lines[0][0])
    nroutes = int(lines[1][0])
    routes = []
    for parts in lines[4:]:
        vals = parts
        # 0 1 1 cost load stops 0 ... 0
        stops = int(vals[5])
        nodes = [int(v) for v in vals[6:6 + stops]]
        routes.append({"cost": float(vals[3]), "load": int(float(vals[4])), "nodes": nodes})
    assert len(routes) == nroutes, (path, len(routes), nroutes)
    return total, routes


def write_solution(path: str, inst: Instance, tour) -> float:
    """Write ``tour`` (the reference's layout, depot 0 between routes) in the
    ``.sol`` layout of the reference's dataset that ``read_solution`` reads:
    total cost, route count, two solver-metadata lines (0 here: no solver
    time or iteration is recorded), then per route
    ``0 1 1 cost load stops 0 c1 ... ck 0``.  A route's cost is its distances
    summed in order (no service time), numbers printed with 6 significant
    digits as in the dataset's files.  Returns the total."""
    t = [int(v) for v in np.asarray(tour).ravel()]
    if not t or t[0] != 0 or t[-1] != 0:
        raise ValueError("a tour starts and ends at the depot")
    d = np.asarray(inst.dist)
    dem = np.asarray(inst.demand)
    routes, cur = [], [0]
    for v in t[1:]:
        cur.append(v)
        if v == 0:
            if len(cur) > 2:
                routes.append(cur)
            cur = [0]
    lines, total = [], 0.0
    for r in routes:
        cost = float(sum(float(d[a, b]) for a, b in zip(r[:-1], r[1:])))
        total += cost
        load = int(sum(int(dem[v]) for v in r))
        lines.append(f" 0 1 1 {cost:.6g} {load}  {len(r)} " + " ".join(str(v) for v in r))
    with open(path, "w") as f:
        f.write(f"{total:.6g}\n{len(routes)}\n0\n0\n" + "\n".join(lines) + "\n")
    return total


def solution_tour(routes) -> np.ndarray:
    """Concatenate routes into the reference's tour layout [0, .., 0, .., 0]."""
    t = [0]
    for r in routes:
        nodes = r["nodes"]
        assert nodes[0] == 0 and nodes[-1] == 0
        t.extend(nodes[1:])
    return np.asarray(t, np.int32)
