"""Write the golden fixtures of this directory from the reference solver.

Everything here is produced by the reference's own code, compiled in place
from the reference tree into oracle/_ref/ by `make -C oracle ref`:

* force_ref.npz   -- the reference's cpu_dump harness output (cpu_dump.cpp:42-64):
                     a sheared 22^3 basalt/iron cube after 12 steps, its force
                     evaluation, and the state after one more step.
* eos_sweep.npz   -- Material::pressure / sound_speed (eos.hpp:52-113) over the
                     eos_gpu_test.cpp:27-33 grid, extended across branch edges,
                     for iron, basalt, water and the workload granite/water.
* eig_yield.npz   -- eig_max (sph.hpp:41-54) and yield_limit (sph.hpp:320-331)
                     on random symmetric tensors.
* weibull.npz     -- seed_damage (sph.hpp:117-128) flaw strains.
* impact_37k_ref.npz -- the CPU-runnable impact workload: dt sequence, energy and
                     aggregates for 100 steps, every 4th particle after 1 step and
                     every 16th particle after 10 and 100 steps.
* tiny_ref.npz    -- a 1,217-particle impact with strength and damage: full
                     state after init and after 1, 5 and 20 steps.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from deorbit_amd import ics, materials as M  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

REF_DIR = os.path.join(ROOT, "oracle", "_ref")


def force_ref():
    exe = os.path.join(REF_DIR, "cpu_dump")
    with tempfile.TemporaryDirectory() as td:
        subprocess.run([exe], cwd=td, check=True, capture_output=True)
        raw = open(os.path.join(td, "force_ref.bin"), "rb").read()
    n = int(np.frombuffer(raw[:4], np.int32)[0])
    off = 4
    hdr = np.frombuffer(raw[off:off + 48], np.float64).copy(); off += 48
    rec = np.frombuffer(raw[off:off + 8 * 29 * n], np.float64).reshape(n, 29).copy(); off += 8 * 29 * n
    dt = float(np.frombuffer(raw[off:off + 8], np.float64)[0]); off += 8
    post = np.frombuffer(raw[off:off + 8 * 13 * n], np.float64).reshape(n, 13).copy(); off += 8 * 13 * n
    assert off == len(raw)
    np.savez_compressed(os.path.join(HERE, "force_ref.npz"), n=n, hdr=hdr, rec=rec, dt=dt, post=post)
    print("force_ref.npz", n)


def eos_sweep():
    mats = {"iron": M.iron_preset(), "basalt": M.basalt_preset(), "water_preset": M.water_preset(),
            "granite": M.granite(True), "water": M.water(), "iron_w": M.iron(True)}
    out = {}
    for name, m in mats.items():
        fr = np.concatenate([np.arange(0.5, 1.5001, 0.05), [0.999999, 1.0, 1.000001, 0.2, 2.0]])
        uu = np.concatenate([np.arange(0.0, 3.0e7 + 1, 1.0e6),
                             [m["uiv"] * (1 - 1e-9), m["uiv"], m["uiv"] * (1 + 1e-9),
                              m["ucv"] * (1 - 1e-9), m["ucv"], m["ucv"] * (1 + 1e-9)]])
        R, U = np.meshgrid(fr * m["rho0"], uu, indexing="ij")
        R, U = R.ravel(), U.ravel()
        P = np.array([po.eos("ref", m, r, u)[0] for r, u in zip(R, U)])
        cs = np.array([po.eos("ref", m, r, u)[1] for r, u in zip(R, U)])
        out[name + "_rho"], out[name + "_u"], out[name + "_P"], out[name + "_cs"] = R, U, P, cs
        out[name + "_emod"] = np.array(po.emod("ref", m))
    np.savez_compressed(os.path.join(HERE, "eos_sweep.npz"), **out)
    print("eos_sweep.npz", len(mats))


def eig_yield():
    rng = np.random.Generator(np.random.PCG64(77))
    S = rng.normal(0, 1e9, size=(400, 6))
    S[:50, 3:] = 0.0                              # diagonal tensors take the max branch
    S[50:60, 3:] = 1e-20                          # p1 < 1e-30 as well
    eig = np.array([po.eig_max("ref", s) for s in S])
    # yield: one particle per tensor, basalt Y = 3.5e8 and a Y <= 0 material
    mats = [M.basalt_preset(), M.water_preset()]
    prm = M.default_params(1.0, (-1, -1, -1), (1, 1, 1), False)
    n = len(S)
    sys_ = po.System(prm, mats, "ref")
    mat = np.zeros(n, np.int32)
    mat[::7] = 1
    sys_.add(np.zeros((n, 3)), np.zeros((n, 3)), np.ones(n), np.zeros(n), mat)
    sys_.set(po.S, S)
    for i in range(n):
        sys_.yield_limit(i)
    np.savez_compressed(os.path.join(HERE, "eig_yield.npz"), S=S, eig=eig, mat=mat, S_yield=sys_.get(po.S))
    print("eig_yield.npz")


def weibull():
    ic = ics.graze_1m()
    out = {}
    for seed in (12345, 1, 4242):
        s = po.make_system(ic, ic["params"], ic["materials"], "ref")
        s.seed_damage(ic["dx"], seed)
        out[f"seed{seed}"] = s.get(po.EPS_ACT)[::97].copy()
    out["stride"] = np.array(97)
    out["mat"] = ic["mat"][::97].copy()
    np.savez_compressed(os.path.join(HERE, "weibull.npz"), **out)
    print("weibull.npz")


AGG_FIELDS = ["pos", "vel", "u", "rho", "hh", "P", "acc"]


def aggregates(st):
    out = []
    for k in AGG_FIELDS:
        a = np.asarray(st[k], float).reshape(len(st["mass"]), -1)
        out.append(np.concatenate([a.min(0), a.max(0), a.mean(0), np.sqrt((a * a).mean(0))]))
    return np.concatenate(out)


def energy(st):
    v = st["vel"]
    return float(np.sum(0.5 * st["mass"] * np.sum(v * v, axis=1)) + np.sum(st["mass"] * st["u"]))


def impact_37k():
    ic = ics.impact_37k()
    s = po.make_system(ic, ic["params"], ic["materials"], "ref")
    s.init()
    dts, E, agg, mom = [], [], [], []
    keep = {}
    st = s.state()
    E.append(energy(st))
    for k in range(100):
        dt = s.timestep()
        dts.append(dt)
        s.step(dt)
        st = s.state()
        E.append(energy(st))
        agg.append(aggregates(st))
        mom.append(np.sum(st["mass"][:, None] * st["vel"], axis=0))
        if k + 1 == 1:
            for f in ("pos", "vel", "u", "rho", "hh", "P", "csig", "acc", "dudt"):
                keep[f"s1_{f}"] = st[f][::4].copy()
        if k + 1 in (10, 100):
            for f in ("pos", "vel", "u", "rho", "hh", "P"):
                keep[f"s{k + 1}_{f}"] = st[f][::16].copy()
        print("impact_37k step", k + 1, flush=True)
    np.savez_compressed(os.path.join(HERE, "impact_37k_ref.npz"), dt=np.array(dts), E=np.array(E),
                        agg=np.array(agg), mom=np.array(mom), **keep)


def tiny():
    ic = ics.tiny_impact()
    s = po.make_system(ic, ic["params"], ic["materials"], "ref", seed_damage=12345, dx=ic["dx"])
    s.init()
    out = {}
    fields = ("pos", "vel", "u", "rho", "hh", "P", "csig", "acc", "dudt", "S", "dSdt", "D", "sigmax",
              "eps_act")
    st = s.state()
    for f in fields:
        out[f"s0_{f}"] = st[f].copy()
    dts = []
    for k in range(20):
        dt = s.timestep()
        dts.append(dt)
        s.step(dt)
        if k + 1 in (1, 5, 20):
            st = s.state()
            for f in fields:
                out[f"s{k + 1}_{f}"] = st[f].copy()
    np.savez_compressed(os.path.join(HERE, "tiny_ref.npz"), dt=np.array(dts), **out)
    print("tiny_ref.npz")


if __name__ == "__main__":
    which = sys.argv[1:] or ["force_ref", "eos_sweep", "eig_yield", "weibull", "tiny", "impact_37k"]
    for w in which:
        globals()[w]()
