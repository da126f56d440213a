"""The recon program's host-side behaviour (Program/ProgramOptions.cpp:
argp options, RECON.ini, error exits; ReconData.cpp:83-144 file checks).
Every case here exits before any device call, so it runs without a GPU."""
import os
import subprocess

import numpy as np
import pytest

from cli_files import write_acquisition

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "fastrecon_amd", "fastrecon")


def run(args, cwd=None):
    if not os.path.exists(CLI):
        pytest.fail("fastrecon program not built (run __graft_entry__.build())")
    return subprocess.run([CLI, *args], cwd=cwd, capture_output=True, text=True, timeout=60)


def test_help_lists_reference_options():
    r = run(["--help"])
    assert r.returncode == 0
    for opt in ("RECON.ini", "-g", "-w", "--pils", "--gpu", "--show"):
        assert opt in r.stdout


def test_missing_ini_exits_1(tmp_path):
    r = run([str(tmp_path / "nope.ini")])
    assert r.returncode == 1
    assert "does not exist" in r.stderr


def test_argument_count(tmp_path):
    r = run([])
    assert r.returncode == 1 and "Missing RECON.ini" in r.stderr
    r = run(["a.ini", "b.ini"])
    assert r.returncode == 1 and "too many arguments" in r.stderr


def test_ini_and_overrides_are_shown(tmp_path):
    """-g overrides overgridding_factor; the parameters block is printed
    before the data is read (ProgramOptions.cpp:41-42, 121-129)."""
    M = 8 * 16
    traj = np.zeros((M, 2), np.float32)
    data = np.zeros((1, M), np.complex64)
    ini = write_acquisition(str(tmp_path), traj, np.ones(M), data, 16, 8, (16, 16, 1), og=2.0)
    os.remove(os.path.join(str(tmp_path), "weights.float"))
    r = run(["-g", "1.5", "-w", "5", ini])
    assert r.returncode == 1
    assert "Overgridding factor:  1.5" in r.stdout
    assert "Kernel width:         5" in r.stdout
    assert "Samples:              16" in r.stdout
    assert "Error: wrong data size in" in r.stdout


@pytest.mark.parametrize("broken", ["traj", "data"])
def test_wrong_file_size_exits_1(tmp_path, broken):
    M = 8 * 16
    traj = np.zeros((M, 2), np.float32)
    data = np.zeros((2, M), np.complex64)
    ini = write_acquisition(str(tmp_path), traj, np.ones(M), data, 16, 8, (16, 16, 1))
    name = "traj_y.float" if broken == "traj" else "coil_001.c64"
    with open(os.path.join(str(tmp_path), name), "r+b") as f:
        f.truncate(100)
    r = run([ini])
    assert r.returncode == 1
    assert "Error: wrong data size in" in r.stdout and name in r.stdout


def test_files_are_name_sorted(tmp_path):
    """Coil and trajectory files are taken in name order (QDir::Name)."""
    M = 8 * 16
    traj = np.zeros((M, 2), np.float32)
    data = np.zeros((3, M), np.complex64)
    ini = write_acquisition(str(tmp_path), traj, np.ones(M), data, 16, 8, (16, 16, 1))
    with open(os.path.join(str(tmp_path), "coil_002.c64"), "r+b") as f:
        f.truncate(8)
    r = run([ini])
    out = r.stdout
    assert out.index("coil_000") < out.index("coil_001") < out.index("coil_002")
    assert out.index("traj_x") < out.index("traj_y")
