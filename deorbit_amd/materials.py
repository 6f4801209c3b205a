"""Material table and solver parameters for the SPH impact workloads.

A material is a plain dict with the fields of the reference ``struct Material``
(hydrocodes/common/eos.hpp:15-39): ``type`` ("tillotson" | "ideal"), ``gamma``,
Tillotson ``rho0, A, B, a, b, alpha, beta, u0 (=E0), uiv (=Eiv), ucv (=Ecv)``,
strength ``G, Y`` and Weibull flaw parameters ``wm`` (modulus) and ``wk``
(flaw density, m^-3; 0 = ductile).  Young's modulus is derived inside the
library exactly as eos.hpp:37 does (9AG/(3A+G), K ~ A).
"""
from __future__ import annotations

import copy


def tillotson(rho0, A, B, a, b, alpha, beta, u0, uiv, ucv, G=0.0, Y=0.0, wm=0.0, wk=0.0):
    """Argument order of ``Material::tillotson`` (eos.hpp:30-32)."""
    return dict(type="tillotson", gamma=1.4, rho0=rho0, A=A, B=B, a=a, b=b, alpha=alpha,
                beta=beta, u0=u0, uiv=uiv, ucv=ucv, G=G, Y=Y, wm=wm, wk=wk)


def ideal(gamma):
    return dict(type="ideal", gamma=gamma, rho0=0.0, A=0.0, B=0.0, a=0.0, b=0.0, alpha=0.0,
                beta=0.0, u0=0.0, uiv=0.0, ucv=0.0, G=0.0, Y=0.0, wm=0.0, wk=0.0)


# Reference presets (eos.hpp:42-50).
def iron_preset():
    return tillotson(7800, 1.28e11, 1.05e11, 0.5, 1.5, 5, 5, 9.5e6, 2.4e6, 8.67e6, 7.75e10, 1.0e9)


def basalt_preset():
    return tillotson(2700, 2.67e10, 2.67e10, 0.5, 1.5, 5, 5, 4.87e8, 4.72e6, 1.82e7, 2.27e10, 3.5e8,
                     16.0, 1.0e61)


def water_preset():
    return tillotson(998, 2.2e9, 9.94e9, 0.7, 0.15, 10, 5, 7.0e6, 4.19e5, 2.69e6, 0.0, 0.0)


def aluminum_preset():
    return tillotson(2700, 7.52e10, 6.5e10, 0.5, 1.63, 5, 5, 5.0e6, 3.0e6, 1.39e7, 0.0, 0.0)


# Workload materials (iron impactor, granite and water targets) with the
# Tillotson constants of the benchmark definitions.
def iron(strength: bool):
    return tillotson(7800, 128e9, 105e9, 0.5, 1.5, 5, 5, 9.5e6, 2.4e6, 8.67e6,
                     80e9 if strength else 0.0, 1e9 if strength else 0.0)


def granite(strength: bool):
    # Weibull k = 1e39 m^-3, m = 9 on the brittle target only (iron is ductile,
    # following the wk = 0 convention of eos.hpp:24-27).
    return tillotson(2680, 18e9, 18e9, 0.5, 1.3, 5, 5, 16e6, 3.5e6, 18e6,
                     30e9 if strength else 0.0, 2e9 if strength else 0.0,
                     9.0 if strength else 0.0, 1e39 if strength else 0.0)


def water():
    return tillotson(1000, 2.18e9, 13.25e9, 0.7, 0.15, 10, 5, 7e6, 0.419e6, 2.69e6)


# Material indices used by the impact generators: 0 target base, 1 impactor, 2 water layer.
MAT_TARGET, MAT_IMPACTOR, MAT_WATER = 0, 1, 2


def impact_materials(strength: bool):
    return [granite(strength), iron(strength), water()]


def default_params(dx: float, lo, hi, frac: bool, periodic=(False, False, False)) -> dict:
    """Driver constants of the reference (gpu_ic.cpp:54-55, run_ic.cpp:69-70,161):
    eta=1.3, h_init=eta*dx, alpha=1, beta=2, eps=0.01, alpha_u=0, cfl=0.2, and with
    fracture on eps_as=0.3 plus damage."""
    return dict(eta=1.3, h_init=1.3 * dx, alpha=1.0, beta=2.0, eps=0.01, alpha_u=0.0, cfl=0.2,
                eps_as=0.3 if frac else 0.0, damage_on=bool(frac),
                periodic=tuple(bool(p) for p in periodic),
                lo=tuple(float(v) for v in lo), hi=tuple(float(v) for v in hi),
                g=0.0, damp=1.0)


def copy_params(p: dict, **kw) -> dict:
    q = copy.deepcopy(p)
    q.update(kw)
    return q
