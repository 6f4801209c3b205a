"""Gravity and relaxation damping of the gpu_ic driver (sph_force.metal:232-237,
283-303; gpu_ic.cpp:38-39) against closed forms, and the CLI passing them through.

Particles far apart (no neighbour within 2h) feel no pair force, so one
kick-drift-kick step of length dt under gravity g (along -z) gives

    v_n = v_{n-1} - g dt,    z_n = z_{n-1} + (v_{n-1} - g dt / 2) dt,

hence after steps dt_1..dt_k totalling t:  v_z = v_0 - g t  and
z = z_0 + v_0 t - g t^2 / 2 (the sum of (t_n^2 - t_{n-1}^2)/2 telescopes, so this
holds for variable steps too).  Damping scales the velocity of non-fixed
material-0 particles by `damp` at the end of each step (after the closing half
kick): with g = 0, v_k = damp^k v_0 and x_k = x_0 + v_0 dt (1 - damp^k)/(1 - damp);
other materials keep v_0.  Fixed particles never move.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from deorbit_amd import dsph, formats
from deorbit_amd import materials as M

pytestmark = pytest.mark.gpu
CLI = os.path.join(ROOT, "deorbit_amd", "bin", "dsph_gpu_ic")


def lone_particles(v0, mat, fixed, dx=10.0, spacing=400.0):
    """Particles on a line `spacing` apart (more than 4 h), no neighbours."""
    n = len(mat)
    pos = np.zeros((n, 3))
    pos[:, 0] = spacing * np.arange(n)
    vel = np.tile(np.asarray(v0, float), (n, 1))
    mats = M.impact_materials(False)
    mass = np.array([mats[m]["rho0"] * dx ** 3 for m in mat])
    return pos, vel, mass, np.asarray(mat, np.int32), np.asarray(fixed, np.uint8), mats


@pytest.mark.parametrize("precision", [dsph.FP64, dsph.MIXED])
def test_gravity_gives_the_free_fall_closed_form(precision, gpu):
    g, dt, k = 3.71, 2.0e-3, 40
    pos, vel, mass, mat, fixed, mats = lone_particles((50.0, 0.0, 30.0), [0, 1, 0, 1], [0, 0, 1, 0])
    prm = M.default_params(10.0, pos.min(0) - 1e4, pos.max(0) + 1e4, False)
    prm["g"] = g
    s = dsph.System(prm, mats, precision)
    s.add(pos, vel, mass, np.zeros(len(mass)), mat, fixed)
    s.init()
    assert np.all(s.get("acc") == 0.0)                 # no pair within reach
    s.run(k, dt, True)
    t = k * dt
    p, v = s.get("pos"), s.get("vel")
    mov = fixed == 0
    assert np.allclose(v[mov, 2], 30.0 - g * t, rtol=0, atol=1e-12 * g * t)
    assert np.allclose(v[mov, 0], 50.0, rtol=0, atol=1e-12)
    assert np.allclose(p[mov, 2], pos[mov, 2] + 30.0 * t - 0.5 * g * t * t, rtol=0, atol=1e-12 * 30.0 * t)
    assert np.allclose(p[mov, 0], pos[mov, 0] + 50.0 * t, rtol=0, atol=1e-12 * 50.0 * t)
    assert np.array_equal(p[~mov], pos[~mov]) and np.array_equal(v[~mov], vel[~mov])
    assert abs(s.diagnostics()["t"] - t) <= 1e-15 * t
    s.close()


@pytest.mark.parametrize("precision", [dsph.FP64, dsph.MIXED])
def test_damping_scales_material_zero_only(precision, gpu):
    damp, dt, k = 0.9, 1.0e-3, 25
    v0 = np.array([40.0, -20.0, 10.0])
    pos, vel, mass, mat, fixed, mats = lone_particles(v0, [0, 1, 2, 0], [0, 0, 0, 1])
    prm = M.default_params(10.0, pos.min(0) - 1e4, pos.max(0) + 1e4, False)
    prm["damp"] = damp
    s = dsph.System(prm, mats, precision)
    s.add(pos, vel, mass, np.zeros(len(mass)), mat, fixed)
    s.init()
    for _ in range(k):
        s.step(dt)
    p, v = s.get("pos"), s.get("vel")
    # material 0, free: damped; materials 1 and 2: untouched; fixed material 0: never moves
    assert np.allclose(v[0], damp ** k * v0, rtol=1e-13, atol=0)
    assert np.allclose(p[0], pos[0] + v0 * dt * (1 - damp ** k) / (1 - damp), rtol=1e-13, atol=1e-12)
    for q in (1, 2):
        assert np.array_equal(v[q], v0)
        assert np.allclose(p[q], pos[q] + v0 * k * dt, rtol=1e-14, atol=1e-12)
    assert np.array_equal(p[3], pos[3]) and np.array_equal(v[3], vel[3])
    s.close()


def test_cli_passes_gravity_and_damping_through(tmp_path):
    """dsph_gpu_ic --g / --damp: the same closed forms through the process boundary,
    with the device's own CFL steps (so the clock, not a fixed step, is summed)."""
    g, damp, t_end = 3.71, 0.98, 0.02
    v0 = np.array([20.0, 0.0, 15.0])
    pos, vel, mass, mat, fixed, _ = lone_particles(v0, [0, 1, 1], [0, 0, 0])
    ic = dict(pos=pos, vel=vel, mass=mass, u=np.full(len(mass), 1e5), mat=mat, fixed=fixed,
              lo=tuple(pos.min(0) - 1e4), hi=tuple(pos.max(0) + 1e4))
    icf = os.path.join(tmp_path, "ic.bin")
    formats.write_ic(icf, ic)
    r = subprocess.run([CLI, icf, "10", str(t_end), "--g", str(g), "--damp", str(damp), "--snap", str(t_end),
                        "--fp64", "--materials", "basalt,iron"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    steps = int(re.search(r"DONE: (\d+) steps", r.stdout).group(1))
    snaps = sorted(f for f in os.listdir(tmp_path) if f.startswith("gpu_snap_"))
    sn = formats.read_snapshot(os.path.join(tmp_path, snaps[-1]))
    # materials 1: free fall over t_end
    for q in (1, 2):
        assert abs(sn["vel"][q, 2] - (15.0 - g * t_end)) <= 1e-10 * g * t_end
        assert abs(sn["pos"][q, 2] - (pos[q, 2] + 15.0 * t_end - 0.5 * g * t_end ** 2)) <= 1e-9
        assert abs(sn["vel"][q, 0] - 20.0) <= 1e-12
    # material 0: damped once per step, so |v_x| = 20 damp^steps
    assert abs(sn["vel"][0, 0] - 20.0 * damp ** steps) <= 1e-10 * 20.0


def test_cli_exits_nonzero_when_the_device_reports_trouble(tmp_path):
    """An internal energy that overflows FP32 (1e39 J/kg) makes the mixed path's
    pressure non-finite, the neighbours' accelerations NaN and their positions
    non-finite at the first drift: the library reports it and the driver exits 1
    with a message instead of running on (boundary_daemon.py:235 calls it with
    check=True)."""
    dx = 10.0
    i, j, k = np.meshgrid(np.arange(6), np.arange(6), np.arange(6), indexing="ij")
    pos = np.stack([i.ravel(), j.ravel(), -k.ravel()], 1).astype(float) * dx
    n = len(pos)
    u = np.full(n, 1e3)
    u[n // 2] = 1e39
    ic = dict(pos=pos, vel=np.zeros((n, 3)), mass=np.full(n, 2700.0 * dx ** 3), u=u, mat=np.zeros(n, np.int32),
              fixed=np.zeros(n, np.uint8), lo=tuple(pos.min(0) - 50), hi=tuple(pos.max(0) + 50))
    icf = os.path.join(tmp_path, "ic.bin")
    formats.write_ic(icf, ic)
    r = subprocess.run([CLI, icf, str(dx), "1.0", "--snap", "10"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    print(r.stdout[-1500:])
    assert r.returncode == 1, r.stdout[-2000:]
    assert "failed" in r.stdout and "DONE:" not in r.stdout
