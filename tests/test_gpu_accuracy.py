"""Per-field accuracy of both device paths against the FP64 oracle over 100 steps.

The oracle's time-step sequence is fed to the device (dt lockstep), so the
errors are the kernels' own and not a divergence of dt.  For rho, P, Pf, |v|, u
and D at steps 1, 10 and 100 the test reports

  max and RMS of the scale-normalised error |g - c| / max|c|   (step_test.cpp:103-108)
  the largest per-particle relative error of rho, and of P where |P| > 10 MPa
  (eos_gpu_test.cpp:72-73)

and at step 100 the aggregates a campaign reads (impactor centroid and mean
speed, melt fraction u > 1 MJ/kg, u max) and the total-energy drift of both
runs.  Two workloads: the 36,937-particle hydrodynamic impact (iron sphere at
45 deg, 8 km/s, on a 64x32x16 granite slab) and a 22,601-particle impact with
strength, Weibull damage and artificial stress (R = 10 dx at 30 deg, 8 km/s, on
a 48x24x16 slab), where cracks grow within the 100 steps.

Three device paths: FP64, MIXED (FP32 force pass; FP64 density terms and
pressure) and MIXED_FAST (FP32 density sums and pressure as well).

Bars (asserted): both mixed paths keep density and pressure within 1e-4 of the
oracle (scale-normalised), density also per particle, through step 10, and
their energy drift within 1e-6 of the oracle's after 100 steps.  Pressure per
particle above 10 MPa: MIXED within 1e-4 through step 10; MIXED_FAST within the
reference's FP32 EOS gate of 5e-3 (eos_gpu_test.cpp:81), since P ~ A (rho/rho0 - 1)
turns its ~1e-7 density error into ~1e-3 of P where rho/rho0 - 1 ~ 1e-4.  The
FP64 path stays within 1e-12 (scale-normalised) through step 10.  Later
horizons are reported with the looser bars written below (impacts amplify
round-off as shocks steepen and flaws activate).
"""
import math

import numpy as np
import pytest

from conftest import device_fields, field_report, format_report, impactor_aggregates, oracle_fields
from deorbit_amd import dsph, ics
from deorbit_amd import materials as M
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

HORIZONS = (1, 10, 100)


def strength_case():
    return ics.impact(10, 30.0, 8000.0, 48, 24, [(M.MAT_TARGET, 16)], seed=1100, strength=True, frac=True,
                      name="strength_damage_23k", steps=100)


CASES = {"impact_37k": ics.impact_37k, "strength_damage_23k": strength_case}


def energy(m, v, u):
    return math.fsum(0.5 * m * np.sum(v * v, axis=1)) + math.fsum(m * u)


@pytest.fixture(scope="module")
def trajectories():
    """Oracle states at the horizons and the dt sequence, per case (CPU, once)."""
    out = {}
    for name, make in CASES.items():
        ic = make()
        o = po.make_system(ic, ic["params"], ic["materials"], seed_damage=12345, dx=ic["dx"])
        o.init()
        m, v, u = o.get(po.MASS), o.get(po.VEL), o.get(po.U)
        E0 = energy(m, v, u)
        dts, states = [], {}
        for k in range(1, HORIZONS[-1] + 1):
            dt = o.timestep()
            o.step(dt)
            dts.append(dt)
            if k in HORIZONS:
                states[k] = oracle_fields(o, po)
        st = states[HORIZONS[-1]]
        drift = (energy(st["mass"], st["vel"], st["u"]) - E0) / E0
        out[name] = (ic, dts, states, drift)
    return out


TAGS = {dsph.FP64: "fp64", dsph.MIXED: "mixed", dsph.MIXED_FAST: "mixed_fast"}
# per-particle pressure bar above 10 MPa through step 10, by path
P_PARTICLE = {dsph.MIXED: 1e-4, dsph.MIXED_FAST: 5e-3}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("precision", list(TAGS), ids=list(TAGS.values()))
def test_per_field_errors_against_the_oracle(trajectories, case, precision, gpu):
    ic, dts, states, drift_ref = trajectories[case]
    tag = TAGS[precision]
    s = dsph.make_system(ic, precision)
    s.init()
    E0 = s.diagnostics()["e_tot"]
    for k in range(1, HORIZONS[-1] + 1):
        s.step(dts[k - 1])
        if k not in HORIZONS:
            continue
        dev, orc = device_fields(s), states[k]
        errs, rho_rel, p_rel = field_report(dev, orc)
        row = format_report(errs)
        print(f"{chr(10) if k == 1 else ''}| {case} | {tag} | {k} | {row} | rho rel {rho_rel:.1e} | "
              f"P rel {p_rel:.1e} |", flush=True)
        if precision == dsph.FP64:
            bar = {1: 1e-12, 10: 1e-12, 100: 1e-9}[k]
            assert max(mx for mx, _ in errs.values()) < bar, (k, errs)
        else:
            if k <= 10:
                # density and pressure within 1e-4 of the oracle (scale-normalised),
                # density also per particle; per-particle pressure above 10 MPa to
                # the path's bar (see the module docstring)
                assert errs["rho"][0] < 1e-4 and errs["P"][0] < 1e-4, (k, errs)
                assert rho_rel < 1e-4 and p_rel < P_PARTICLE[precision], (k, rho_rel, p_rel)
            bar = {1: 2e-5, 10: 2e-4, 100: 2e-2}[k]
            assert max(mx for f, (mx, _) in errs.items() if f != "D") < bar, (k, errs)
            # damage switches on where sigma_max / Emod crosses the flaw strain: an
            # FP32 sigma_max a few ulp away starts a crack one step early or late,
            # so single particles differ by up to one growth increment
            assert errs["D"][1] < 1e-2, (k, errs["D"])
    d = s.diagnostics()
    drift = (d["e_tot"] - E0) / E0
    a = impactor_aggregates(device_fields(s), ic["mat"])
    b = impactor_aggregates(states[HORIZONS[-1]], ic["mat"])
    s.close()
    cen = max(abs(a["x"] - b["x"]), abs(a["z"] - b["z"]))
    print(f"| {case} | {tag} | aggregates | centroid {cen:.1e} m | mean |v| "
          f"{abs(a['speed'] - b['speed']) / b['speed']:.1e} rel | melt {a['melt']:.4f} vs {b['melt']:.4f} | umax "
          f"{abs(a['umax'] - b['umax']) / b['umax']:.1e} rel | drift {drift:.6e} vs {drift_ref:.6e}, "
          f"difference {abs(drift - drift_ref):.1e} |", flush=True)
    assert abs(drift - drift_ref) <= 1e-6
    assert abs(a["speed"] - b["speed"]) <= 1e-3 * b["speed"] and abs(a["melt"] - b["melt"]) <= 1e-3


@pytest.mark.parametrize("case", list(CASES))
def test_mixed_pressure_per_evaluation_on_the_state_after_100_steps(trajectories, case, gpu):
    """The per-particle pressure bar of DSPH_MIXED (1e-4 where |P| > 10 MPa) holds
    per evaluation on the post-impact state after 100 steps.  The mixed path's own
    state after 100 dt-lockstep steps (positions, velocities, u, h, S, D, flaws) is
    handed to the oracle, and both evaluate the density rounds and the EOS on it
    (compute_density_h, compute_forces: sph.hpp:185-206, 260-318).  This separates
    the evaluation error from the drift of the trajectories, which after 100 steps
    sets the per-particle pressure difference of the two runs (~1.5e-4, reported by
    test_per_field_errors_against_the_oracle)."""
    ic, dts, _, _ = trajectories[case]
    s = dsph.make_system(ic, dsph.MIXED)
    s.init()
    for dt in dts[:100]:
        s.step(dt)
    st = s.state(["pos", "vel", "u", "hh", "S", "D", "eps_act"])
    o = po.make_system(ic, ic["params"], ic["materials"], seed_damage=12345, dx=ic["dx"])
    for name, field in (("pos", po.POS), ("vel", po.VEL), ("u", po.U), ("hh", po.HH), ("S", po.S),
                        ("D", po.D), ("eps_act", po.EPS_ACT)):
        o.set(field, st[name])
    o.compute_density_h()
    o.compute_forces()
    s.compute_density_h()
    s.compute_forces()
    rho_g, rho_c = s.get("rho"), o.get(po.RHO)
    p_g, p_c = s.get("P"), o.get(po.P)
    big = np.abs(p_c) > 1e7
    rho_rel = float(np.max(np.abs(rho_g - rho_c) / rho_c))
    p_rel = float(np.max(np.abs(p_g[big] - p_c[big]) / np.abs(p_c[big])))
    h_rel = float(np.max(np.abs(s.get("hh") - o.get(po.HH)) / o.get(po.HH)))
    print(f"| {case} | mixed, evaluated on its state after 100 steps | rho rel {rho_rel:.1e} | h rel {h_rel:.1e} | "
          f"P rel {p_rel:.1e} over {int(big.sum())} particles above 10 MPa |", flush=True)
    assert big.sum() > 1000
    assert rho_rel < 1e-4 and h_rel < 1e-4
    assert p_rel < 1e-4
    s.close()
