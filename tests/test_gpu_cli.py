"""The gpu_ic-compatible CLI driven the way boundary_daemon.py:230-248 drives gpu_ic."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from deorbit_amd import formats, ics

pytestmark = pytest.mark.gpu
CLI = os.path.join(ROOT, "deorbit_amd", "bin", "dsph_gpu_ic")


def test_cli_runs_a_grazing_ic_like_the_daemon(tmp_path):
    ic = ics.grazing_ic(10.0, 8000.0, cppr=4, D=100.0, L=(1500.0, 400.0, 500.0))
    icf = os.path.join(tmp_path, "ic.bin")
    formats.write_ic(icf, ic)
    t_end = 0.01
    r = subprocess.run([CLI, icf, f"{ic['dx']}", f"{t_end}", "--snap", f"{0.6 * t_end}"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    # the daemon's melt= parser (boundary_daemon.py:240-248)
    melts = []
    for line in r.stdout.splitlines():
        k = line.find("melt=")
        if k >= 0:
            melts.append(float(line[k + 5:line.index("%", k)]) / 100.0)
    assert melts and all(0.0 <= m <= 1.0 for m in melts)
    assert "DONE:" in r.stdout
    snaps = sorted(f for f in os.listdir(tmp_path) if f.startswith("gpu_snap_") and f.endswith(".bin"))
    assert snaps
    s = formats.read_snapshot(os.path.join(tmp_path, snaps[-1]))
    assert len(s["u"]) == len(ic["mass"])
    assert np.array_equal(s["mat"], ic["mat"])                 # written in IC order
    fixed = ic["fixed"] == 1
    assert np.array_equal(s["pos"][fixed], ic["pos"][fixed])   # frozen wall particles never move
    imp = ic["mat"] == 1
    assert np.all(s["pos"][imp][:, 0] > ic["pos"][imp][:, 0])  # the impactor moved downrange
    assert np.all(np.isfinite(s["rho"])) and np.all(s["h"] > 0)


def test_cli_usage_error_exit_code(tmp_path):
    r = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "usage" in r.stdout
    r = subprocess.run([CLI, os.path.join(tmp_path, "missing.bin"), "10", "0.1"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stdout


def _cli(args, cwd):
    r = subprocess.run([CLI] + [str(a) for a in args], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    return r.stdout


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_cli_checkpoint_restart_continues_bit_for_bit(tmp_path, precision):
    """Eight steps straight equal four steps, a checkpoint, and a restarted run of
    four more: the two final checkpoints (every field, time, step count and next
    CFL step) are byte-identical."""
    ic = ics.tiny_impact()
    icf = os.path.join(tmp_path, "ic.bin")
    formats.write_ic(icf, ic)
    extra = ["--frac", "--materials", "granite,iron80", "--snap", "1e9"] + (["--fp64"] if precision == "fp64" else [])
    a, b, c = (os.path.join(tmp_path, f) for f in ("a.ckpt", "b.ckpt", "c.ckpt"))
    _cli([icf, ic["dx"], 10.0, "--max-steps", 8, "--checkpoint", a] + extra, tmp_path)
    _cli([icf, ic["dx"], 10.0, "--max-steps", 4, "--checkpoint", b] + extra, tmp_path)
    out = _cli([icf, ic["dx"], 10.0, "--max-steps", 8, "--restart", b, "--checkpoint", c] + extra, tmp_path)
    assert "restart at" in out and "extended checkpoint" in out
    ca, cb, cc = (formats.read_checkpoint(p) for p in (a, b, c))
    assert cb["steps"] == 4 and ca["steps"] == cc["steps"] == 8 and ca["extended"]
    with open(a, "rb") as fa, open(c, "rb") as fc:
        assert fa.read() == fc.read()
    assert np.array_equal(ca["mat"], ic["mat"])                   # IC order is kept through the files


def test_cli_restart_from_a_plain_run_ic_checkpoint(tmp_path):
    """A reference-format checkpoint (no extension) restarts by re-running init()."""
    ic = ics.tiny_impact()
    icf = os.path.join(tmp_path, "ic.bin")
    formats.write_ic(icf, ic)
    full = os.path.join(tmp_path, "full.ckpt")
    _cli([icf, ic["dx"], 10.0, "--max-steps", 3, "--checkpoint", full, "--snap", "1e9"], tmp_path)
    ck = formats.read_checkpoint(full)
    plain = os.path.join(tmp_path, "plain.ckpt")
    formats.write_checkpoint(plain, ck, ck["lo"], ck["hi"], ck["t"], ck["dx"])
    out = _cli([icf, ic["dx"], 10.0, "--max-steps", 3, "--restart", plain, "--snap", "1e9"], tmp_path)
    assert "plain checkpoint" in out and "DONE:" in out
