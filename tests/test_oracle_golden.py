"""The oracle against the fixtures the reference solver wrote (tests/golden).

Every comparison is bitwise: the fixtures come from the reference's own code
(cpu_dump and the reference headers), and the oracle restates it operation for
operation.  These run on any CPU, with or without the reference tree.
"""
import numpy as np
import pytest

from conftest import golden
from deorbit_amd import ics, formats
from deorbit_amd import materials as M
from oracle import pyoracle as po


def force_ref_system(which="oracle"):
    """Oracle system holding the cpu_dump state (cpu_dump.cpp:11-40) before its last step."""
    g = golden("force_ref.npz")
    rec = g["rec"]
    ic = ics.sheared_cube()
    s = po.System(ic["params"], ic["materials"], which)
    s.add(rec[:, 0:3], rec[:, 3:6], rec[:, 6], rec[:, 7], rec[:, 10].astype(np.int32))
    s.set(po.RHO, rec[:, 8])
    s.set(po.HH, rec[:, 9])
    s.set(po.S, rec[:, 11:17][:, formats.FILE_TO_MEM])
    s.set(po.D, rec[:, 17])
    s.set(po.EPS_ACT, rec[:, 18])
    return s, g


def test_force_ref_header_matches_driver_constants():
    g = golden("force_ref.npz")
    assert g["n"] == 22 ** 3
    eta, alpha, beta, eps, alpha_u, eps_as = g["hdr"]
    p = ics.sheared_cube()["params"]
    assert (eta, alpha, beta, eps, alpha_u, eps_as) == (p["eta"], p["alpha"], p["beta"], p["eps"],
                                                        p["alpha_u"], p["eps_as"])


def test_force_ref_forces_bitwise():
    s, g = force_ref_system()
    rec = g["rec"]
    s.build_cells()
    s.compute_forces()
    assert np.array_equal(s.get(po.ACC), rec[:, 19:22])
    assert np.array_equal(s.get(po.DUDT), rec[:, 22])
    assert np.array_equal(s.get(po.DSDT), rec[:, 23:29][:, formats.FILE_TO_MEM])


def test_force_ref_step_bitwise():
    s, g = force_ref_system()
    s.build_cells()
    s.compute_forces()
    s.step(float(g["dt"]))
    post = g["post"]
    assert np.array_equal(s.get(po.POS), post[:, 0:3])
    assert np.array_equal(s.get(po.VEL), post[:, 3:6])
    assert np.array_equal(s.get(po.U), post[:, 6])
    assert np.array_equal(s.get(po.S), post[:, 7:13][:, formats.FILE_TO_MEM])


EOS_MATS = {"iron": M.iron_preset(), "basalt": M.basalt_preset(), "water_preset": M.water_preset(),
            "granite": M.granite(True), "water": M.water(), "iron_w": M.iron(True)}


@pytest.mark.parametrize("name", sorted(EOS_MATS))
def test_eos_sweep_bitwise(name):
    g = golden("eos_sweep.npz")
    m = EOS_MATS[name]
    R, U = g[name + "_rho"], g[name + "_u"]
    P = np.array([po.eos("oracle", m, r, u)[0] for r, u in zip(R, U)])
    cs = np.array([po.eos("oracle", m, r, u)[1] for r, u in zip(R, U)])
    assert np.array_equal(P, g[name + "_P"])
    assert np.array_equal(cs, g[name + "_cs"])
    assert po.emod("oracle", m) == float(g[name + "_emod"])


def test_eig_max_and_yield_bitwise():
    g = golden("eig_yield.npz")
    eig = np.array([po.eig_max("oracle", s) for s in g["S"]])
    assert np.array_equal(eig, g["eig"])
    mats = [M.basalt_preset(), M.water_preset()]
    prm = M.default_params(1.0, (-1, -1, -1), (1, 1, 1), False)
    n = len(g["S"])
    s = po.System(prm, mats, "oracle")
    s.add(np.zeros((n, 3)), np.zeros((n, 3)), np.ones(n), np.zeros(n), g["mat"])
    s.set(po.S, g["S"])
    for i in range(n):
        s.yield_limit(i)
    assert np.array_equal(s.get(po.S), g["S_yield"])


def test_weibull_flaws_bitwise():
    g = golden("weibull.npz")
    ic = ics.graze_1m()
    stride = int(g["stride"])
    assert np.array_equal(ic["mat"][::stride], g["mat"])
    for seed in (12345, 1, 4242):
        s = po.make_system(ic, ic["params"], ic["materials"], "oracle")
        s.seed_damage(ic["dx"], seed)
        assert np.array_equal(s.get(po.EPS_ACT)[::stride], g[f"seed{seed}"])


def test_tiny_impact_trajectory_bitwise():
    g = golden("tiny_ref.npz")
    ic = ics.tiny_impact()
    s = po.make_system(ic, ic["params"], ic["materials"], "oracle", seed_damage=12345, dx=ic["dx"])
    s.init()
    fields = ("pos", "vel", "u", "rho", "hh", "P", "csig", "acc", "dudt", "S", "dSdt", "D", "sigmax",
              "eps_act")
    st = s.state()
    for f in fields:
        assert np.array_equal(st[f], g[f"s0_{f}"]), f
    for k in range(20):
        dt = s.timestep()
        assert dt == g["dt"][k]
        s.step(dt)
        if k + 1 in (1, 5, 20):
            st = s.state()
            for f in fields:
                assert np.array_equal(st[f], g[f"s{k + 1}_{f}"]), (k + 1, f)


def test_impact_37k_trajectory_bitwise():
    g = golden("impact_37k_ref.npz")
    ic = ics.impact_37k()
    s = po.make_system(ic, ic["params"], ic["materials"], "oracle")
    s.init()
    for k in range(100):
        dt = s.timestep()
        assert dt == g["dt"][k], k
        s.step(dt)
        if k + 1 == 1:
            st = s.state()
            for f in ("pos", "vel", "u", "rho", "hh", "P", "csig", "acc", "dudt"):
                assert np.array_equal(st[f][::4], g[f"s1_{f}"]), f
        if k + 1 in (10, 100):
            st = s.state()
            for f in ("pos", "vel", "u", "rho", "hh", "P"):
                assert np.array_equal(st[f][::16], g[f"s{k + 1}_{f}"]), (k + 1, f)
