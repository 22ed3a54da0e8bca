"""The cvrp_aco command-line driver (the reference's main.cpp trial loop) on
the device: valid best-so-far solutions, reports written, bad input
rejected with exit status 1."""
import os
import subprocess

import pytest

from helpers import VRPLIB

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cvrp_aco_ocl_amd", "cvrp_aco")


def _run(args, timeout=120):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)


def test_cli_trials_on_cmt1(tmp_path):
    r = _run(["--tries", "2", "--iterations", "30", "--ls", "full", "--selection", "knn", "--k", "20",
              "--report", str(tmp_path), os.path.join(VRPLIB, "CMT1.vrp")])
    assert r.returncode == 0, r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("try ")]
    assert len(lines) == 2 and all("valid yes" in l for l in lines)
    best = float(lines[0].split("Best Length ")[1].split()[0])
    assert 524.6 <= best < 1.1 * 524.61
    assert (tmp_path / "best.CMT1.vrp").exists()
    assert (tmp_path / "best_so_far.CMT1.vrp").read_text().count("iter") >= 1


def test_cli_reference_defaults_with_distance_limit():
    # CMT6 carries DISTANCE and SERVICE_TIME; reference rule set (I-Roulette)
    r = _run(["--tries", "1", "--iterations", "20", "--ants", "50", os.path.join(VRPLIB, "CMT6.vrp")])
    assert r.returncode == 0, r.stderr
    assert "valid yes" in r.stdout


def test_cli_rejects_missing_file():
    r = _run(["--tries", "1", "/nonexistent.vrp"])
    assert r.returncode == 1 and "read_instance_file" in r.stderr
