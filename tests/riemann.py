"""Exact solution of the 1D Riemann problem for an ideal gas (test helper).

Standard construction (E. F. Toro, Riemann Solvers and Numerical Methods for
Fluid Dynamics, ch. 4): the star-region pressure solves
    f_L(p) + f_R(p) + (u_R - u_L) = 0,
with f_K the shock (Rankine-Hugoniot) branch for p > p_K and the rarefaction
(isentropic) branch otherwise; the solution is self-similar in x/t.
"""
from __future__ import annotations

import math


class RiemannIdealGas:
    def __init__(self, rhoL, uL, pL, rhoR, uR, pR, gamma=1.4):
        self.g = gamma
        self.L = (rhoL, uL, pL, math.sqrt(gamma * pL / rhoL))
        self.R = (rhoR, uR, pR, math.sqrt(gamma * pR / rhoR))
        self.p_star, self.u_star = self._star()

    def _f(self, p, side):
        rho, _, pk, ck = side
        g = self.g
        if p > pk:
            a = 2.0 / ((g + 1.0) * rho)
            b = (g - 1.0) / (g + 1.0) * pk
            q = math.sqrt(a / (p + b))
            return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (p + b))
        e = (g - 1.0) / (2.0 * g)
        return 2.0 * ck / (g - 1.0) * ((p / pk) ** e - 1.0), (p / pk) ** (-(g + 1.0) / (2.0 * g)) / (rho * ck)

    def _star(self):
        du = self.R[1] - self.L[1]
        p = max(1e-12, 0.5 * (self.L[2] + self.R[2]))
        for _ in range(200):
            fl, dl = self._f(p, self.L)
            fr, dr = self._f(p, self.R)
            step = (fl + fr + du) / (dl + dr)
            p_new = max(1e-12, p - step)
            if abs(p_new - p) <= 1e-14 * p:
                p = p_new
                break
            p = p_new
        fl, _ = self._f(p, self.L)
        fr, _ = self._f(p, self.R)
        return p, 0.5 * (self.L[1] + self.R[1]) + 0.5 * (fr - fl)

    def shock_speed_right(self):
        rho, u, p, c = self.R
        g = self.g
        return u + c * math.sqrt((g + 1.0) / (2.0 * g) * self.p_star / p + (g - 1.0) / (2.0 * g))

    def sample(self, xi):
        """(rho, u, p) at similarity coordinate xi = x / t."""
        g, ps, us = self.g, self.p_star, self.u_star
        left = xi <= us
        rho, u, p, c = self.L if left else self.R
        s = -1.0 if left else 1.0                      # direction away from the contact
        if ps > p:                                      # shock on this side
            speed = u + s * c * math.sqrt((g + 1.0) / (2.0 * g) * ps / p + (g - 1.0) / (2.0 * g))
            if s * (xi - speed) >= 0.0:
                return rho, u, p
            r = ps / p
            gm = (g - 1.0) / (g + 1.0)
            return rho * (r + gm) / (gm * r + 1.0), us, ps
        head = u + s * c                                # rarefaction on this side
        c_star = c * (ps / p) ** ((g - 1.0) / (2.0 * g))
        tail = us + s * c_star
        if s * (xi - head) >= 0.0:
            return rho, u, p
        if s * (xi - tail) <= 0.0:
            return rho * (ps / p) ** (1.0 / g), us, ps
        # inside the fan
        uf = 2.0 / (g + 1.0) * (-s * c + 0.5 * (g - 1.0) * u + xi)
        cf = 2.0 / (g + 1.0) * (c - s * 0.5 * (g - 1.0) * (u - xi))
        return rho * (cf / c) ** (2.0 / (g - 1.0)), uf, p * (cf / c) ** (2.0 * g / (g - 1.0))
