"""Pieces of the device code compiled for the host (tests/native/*.cpp):

* the density moment sums (deorbit_amd/csrc/moments.h) against direct sums of
  the reference's kernel (sph.hpp:19-24) over random neighbour sets: the first
  iteration at h and the second at every h1 the moments cover, with candidates
  on the band edges; covers() refuses h1 outside the band;
* the tiled sort-column numbering (deorbit_amd/csrc/common.h)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    pytest.skip("hipcc not available")


def test_moment_sums_equal_direct_sums(tmp_path):
    """Moment sums against direct sums (tests/native/density_moments.cpp)."""
    exe = os.path.join(tmp_path, "density_moments")
    subprocess.run([hipcc(), "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17",
                    "-I", os.path.join(ROOT, "deorbit_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "density_moments.cpp"), "-o", exe],
                   check=True, capture_output=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    worst = float(out.stdout.split("max rel err")[1].split(",")[0])
    assert worst < 1e-13


def test_column_tiles_number_every_column_once(tmp_path):
    """The tiled sort-column numbering (deorbit_amd/csrc/common.h col_index,
    cell_key, block_key), compiled for the host: one to one onto the padded
    columns for every grid shape and tile width tried, keys within the segment's
    ranges (tests/native/column_tiles.cpp)."""
    exe = os.path.join(tmp_path, "column_tiles")
    subprocess.run([hipcc(), "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17",
                    "-I", os.path.join(ROOT, "deorbit_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "column_tiles.cpp"), "-o", exe],
                   check=True, capture_output=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
