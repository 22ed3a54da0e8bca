"""The cvrp_aco command-line driver (the reference's main.cpp trial loop) on
the device: valid best-so-far solutions, reports written, bad input
rejected with exit status 1."""
import os
import subprocess

import pytest

import cvrp_aco_ocl_amd as ca
from cvrp_aco_ocl_amd.instances import read_solution, solution_tour
from helpers import VRPLIB

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cvrp_aco_ocl_amd", "cvrp_aco")


def _run(args, timeout=120):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)


def test_cli_trials_on_cmt1(tmp_path):
    r = _run(["--tries", "2", "--iterations", "30", "--ls", "full", "--selection", "knn", "--k", "20",
              "--report", str(tmp_path), "--sol", str(tmp_path), os.path.join(VRPLIB, "CMT1.vrp")])
    assert r.returncode == 0, r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("try ")]
    assert len(lines) == 2 and all("valid yes" in l for l in lines)
    lengths = [float(l.split("Best Length ")[1].split()[0]) for l in lines]
    assert 524.6 <= lengths[0] < 1.1 * 524.61
    assert (tmp_path / "best.CMT1.vrp").exists()
    assert (tmp_path / "best_so_far.CMT1.vrp").read_text().count("iter") >= 1
    # the best try's tour in the dataset's .sol layout
    total, routes = read_solution(str(tmp_path / "CMT1.sol"))
    inst = ca.Instance.from_vrplib(os.path.join(VRPLIB, "CMT1.vrp"))
    tour = solution_tour(routes)
    assert inst.check_solution(tour)
    assert total == pytest.approx(min(lengths), rel=1e-4)
    for rt in routes:
        assert rt["load"] == int(inst.demand[rt["nodes"]].sum()) <= inst.capacity


def test_cli_reference_defaults_with_distance_limit():
    # CMT6 carries DISTANCE and SERVICE_TIME; reference rule set (I-Roulette)
    r = _run(["--tries", "1", "--iterations", "20", "--ants", "50", os.path.join(VRPLIB, "CMT6.vrp")])
    assert r.returncode == 0, r.stderr
    assert "valid yes" in r.stdout


def test_cli_rejects_missing_file():
    r = _run(["--tries", "1", "/nonexistent.vrp"])
    assert r.returncode == 1 and "read_instance_file" in r.stderr


def test_cli_decomposition_rounds(tmp_path):
    """--decompose on X-n1001 without local search on the master: rounds of
    sub-colonies after every iteration; the result is valid and shorter than
    the same run without rounds."""
    path = os.path.join(VRPLIB, "X-n1001-k43.vrp")
    if not os.path.exists(path):
        pytest.skip("X-n1001-k43 not in the fixtures")
    base = ["--tries", "1", "--iterations", "4", "--ants", "128", "--selection", "knn", "--k", "16",
            "--dist", "nint", "--seed", "3"]
    plain = _run(base + [path])
    dec = _run(base + ["--decompose", "--sub-iterations", "25", "--pheromone-back", path])
    assert plain.returncode == 0 and dec.returncode == 0, dec.stderr
    lp = float(plain.stdout.split("Best Length ")[1].split()[0])
    ld = float(dec.stdout.split("Best Length ")[1].split()[0])
    assert "valid yes" in dec.stdout
    assert ld < lp
