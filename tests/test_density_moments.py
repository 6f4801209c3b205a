"""The density moment sums of the device code (deorbit_amd/csrc/moments.h),
compiled for the host, against direct sums of the reference's kernel
(sph.hpp:19-24) over random neighbour sets: the first iteration at h and the
second at every h1 the moments cover, with candidates on the band edges; and
covers() refuses h1 outside the band (tests/native/density_moments.cpp)."""
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
