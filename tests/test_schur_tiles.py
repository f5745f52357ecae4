"""The reduced camera system's work decomposition (factor blocks, tile
visits, batches, slot and pair lists, segments and their thread parts,
maplab_amd/csrc/ba_tiles.cc) against a brute-force S on the CPU: the check
program replays the block kernel's staging and products with bounds checks
and compares S and rhs with the sum over incidence pairs."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "maplab_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not available")
    exe = str(tmp_path_factory.mktemp("tiles") / "schur_tiles_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", CSRC, "-o", exe, os.path.join(HERE, "native", "schur_tiles_check.cc"),
                    os.path.join(CSRC, "ba_tiles.cc")], check=True)
    return exe


@pytest.mark.parametrize("pvar, landmarks, seed", [(7, 200, 1), (16, 500, 2), (75, 3000, 3), (33, 20000, 4),
                                                   (200, 6000, 5)])
def test_decomposition_sums_every_pair_once(checker, pvar, landmarks, seed):
    out = subprocess.run([checker, str(pvar), str(landmarks), str(seed)], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
