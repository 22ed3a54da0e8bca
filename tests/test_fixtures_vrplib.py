"""Golden vectors from the reference's dataset (tests/golden/vrplib): the
best-known solutions pin route costs, loads and feasibility.

X-set (Uchoa et al.) solutions are priced with VRPLIB nint distances: every
route cost and load must match exactly.  CMT / Golden / Li solutions are
priced with the reference's unrounded float distances: totals within 5e-4
relative.
"""
import glob
import os

import numpy as np
import pytest

import oracle
import cvrp_aco_ocl_amd as ca
from cvrp_aco_ocl_amd.instances import read_solution, solution_tour, write_solution

from helpers import VRPLIB

X_FILES = sorted(glob.glob(os.path.join(VRPLIB, "X-*.vrp")))
FLOAT_FILES = sorted(glob.glob(os.path.join(VRPLIB, "CMT*.vrp")) +
                     glob.glob(os.path.join(VRPLIB, "Golden_*.vrp")) +
                     glob.glob(os.path.join(VRPLIB, "Li_*.vrp")))
REF_X = sorted(glob.glob("/root/reference/dataset/X/*.sol"))


def _check_x(vrp, sol):
    inst = ca.Instance.from_vrplib(vrp, ca.DIST_NINT)
    total, routes = read_solution(sol)
    P = oracle.Problem(inst.dist, inst.demand, inst.capacity)
    for r in routes:
        nodes = np.array(r["nodes"], np.int32)
        assert P.length(nodes) == r["cost"], (sol, r)
        assert int(inst.demand[nodes].sum()) == r["load"]
        assert r["load"] <= inst.capacity
    tour = solution_tour(routes)
    assert P.check(tour) and inst.check_solution(tour)
    assert P.length(tour) == total
    return inst, tour


@pytest.mark.parametrize("vrp", X_FILES, ids=os.path.basename)
def test_x_solutions_exact(vrp):
    _check_x(vrp, vrp[:-4] + ".sol")


@pytest.mark.skipif(not REF_X, reason="reference dataset not present")
def test_all_reference_x_solutions_exact():
    for sol in REF_X:
        vrp = sol[:-4] + ".vrp"
        if os.path.exists(vrp):
            _check_x(vrp, sol)


def test_x_opt_file_matches_sol():
    vrp = os.path.join(VRPLIB, "X-n242-k48.vrp")
    inst, tour = _check_x(vrp, vrp[:-4] + ".sol")
    routes = []
    with open(vrp[:-4] + ".opt") as f:
        for line in f:
            if line.startswith("Route"):
                routes.append([int(v) for v in line.split(":")[1].split()])
    cost = sum(inst.tour_length([0] + r + [0]) for r in routes)
    stated = [float(line.split()[1]) for line in open(vrp[:-4] + ".opt") if line.lower().startswith("cost")]
    assert stated and cost == stated[0]
    tour = np.array([0] + [v for r in routes for v in r + [0]], np.int32)
    assert inst.check_solution(tour)
    assert cost <= read_solution(vrp[:-4] + ".sol")[0]


@pytest.mark.parametrize("vrp", FLOAT_FILES, ids=os.path.basename)
def test_float_distance_solutions(vrp):
    inst = ca.Instance.from_vrplib(vrp)  # EUC_2D as the reference computes it
    total, routes = read_solution(vrp[:-4] + ".sol")
    tour = solution_tour(routes)
    assert inst.check_solution(tour)
    P = oracle.Problem(inst.dist, inst.demand, inst.capacity, inst.max_dist, inst.service_time)
    assert P.check(tour)
    L = P.length(tour)
    assert abs(L - total) <= 5e-4 * total
    if inst.optimum:  # COMMENT holds an older best-known value in some files
        assert abs(inst.optimum - total) <= 1e-3 * total


def test_oracle_distances_equal_product_distances():
    for vrp in X_FILES[:3] + FLOAT_FILES[:3]:
        inst = ca.Instance.from_vrplib(vrp)
        mode = 0 if inst.dist_mode == ca.DIST_FLOAT else 1
        np.testing.assert_array_equal(oracle.distances(inst.x, inst.y, mode), inst.dist)
        inst1 = ca.Instance.from_vrplib(vrp, ca.DIST_NINT)
        np.testing.assert_array_equal(oracle.distances(inst.x, inst.y, 1), inst1.dist)


@pytest.mark.parametrize("vrp", X_FILES[:3] + FLOAT_FILES[:2], ids=os.path.basename)
def test_write_solution_round_trip(vrp, tmp_path):
    """A fixture's solution written and read back: the same routes, loads
    and (to the layout's 6 significant digits) costs."""
    mode = ca.DIST_NINT if os.path.basename(vrp).startswith("X-") else ca.DIST_FLOAT
    inst = ca.Instance.from_vrplib(vrp, mode)
    total, routes = read_solution(vrp[:-4] + ".sol")
    tour = solution_tour(routes)
    out = tmp_path / "out.sol"
    written = write_solution(str(out), inst, tour)
    total2, routes2 = read_solution(str(out))
    assert [r["nodes"] for r in routes2] == [r["nodes"] for r in routes]
    assert [r["load"] for r in routes2] == [r["load"] for r in routes]
    assert total2 == pytest.approx(written, rel=1e-5)
    for r, r2 in zip(routes, routes2):
        assert r2["cost"] == pytest.approx(r["cost"], rel=1e-5)
    np.testing.assert_array_equal(solution_tour(routes2), tour)
    if mode == ca.DIST_NINT:
        assert total2 == total


def test_write_solution_rejects_open_tour(tmp_path):
    inst = ca.Instance.from_vrplib(X_FILES[0], ca.DIST_NINT)
    with pytest.raises(ValueError):
        write_solution(str(tmp_path / "x.sol"), inst, [1, 2, 0])
