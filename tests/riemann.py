"""Exact solution of the 1D Riemann problem for an ideal gas (Toro, Riemann Solvers and Numerical
Methods for Fluid Dynamics, 3rd ed., ch. 4): Newton iteration for the star pressure, then sampling
of the self-similar solution. Test infrastructure (the physics check of the Sod set-up, the
reference's share/spearhed/ValidationSetups/SodShockTube/validate_sod.py:55-163 role).
"""
from __future__ import annotations

import math

import numpy as np


def _fk(p, rho_k, p_k, g):
    """Toro eq. 4.6/4.7: pressure function of one side and its derivative."""
    if p > p_k:  # shock
        A = 2.0 / ((g + 1.0) * rho_k)
        B = (g - 1.0) / (g + 1.0) * p_k
        q = math.sqrt(A / (p + B))
        return (p - p_k) * q, q * (1.0 - 0.5 * (p - p_k) / (p + B))
    c_k = math.sqrt(g * p_k / rho_k)  # rarefaction
    f = 2.0 * c_k / (g - 1.0) * ((p / p_k) ** ((g - 1.0) / (2.0 * g)) - 1.0)
    df = (p / p_k) ** (-(g + 1.0) / (2.0 * g)) / (rho_k * c_k)
    return f, df


def star_state(rho_l, u_l, p_l, rho_r, u_r, p_r, g):
    p = max(1e-8, 0.5 * (p_l + p_r))
    for _ in range(200):
        fl, dfl = _fk(p, rho_l, p_l, g)
        fr, dfr = _fk(p, rho_r, p_r, g)
        dp = (fl + fr + u_r - u_l) / (dfl + dfr)
        p_new = max(1e-10, p - dp)
        if abs(p_new - p) < 1e-12 * (p_new + p):
            p = p_new
            break
        p = p_new
    fl, _ = _fk(p, rho_l, p_l, g)
    fr, _ = _fk(p, rho_r, p_r, g)
    u = 0.5 * (u_l + u_r) + 0.5 * (fr - fl)
    return p, u


def sample(x, t, x0=0.5, rho_l=1.0, u_l=0.0, p_l=1.0, rho_r=0.125, u_r=0.0, p_r=0.1, g=1.4):
    """rho, u, p at positions x and time t for a discontinuity at x0."""
    ps, us = star_state(rho_l, u_l, p_l, rho_r, u_r, p_r, g)
    c_l = math.sqrt(g * p_l / rho_l)
    c_r = math.sqrt(g * p_r / rho_r)
    gm = (g - 1.0) / (g + 1.0)
    out_rho = np.empty_like(x)
    out_u = np.empty_like(x)
    out_p = np.empty_like(x)
    for n, xv in enumerate(np.asarray(x, dtype=float)):
        s = (xv - x0) / t
        if s <= us:  # left of the contact
            if ps > p_l:  # left shock
                S = u_l - c_l * math.sqrt((g + 1) / (2 * g) * ps / p_l + (g - 1) / (2 * g))
                if s <= S:
                    r, u, p = rho_l, u_l, p_l
                else:
                    r, u, p = rho_l * (ps / p_l + gm) / (gm * ps / p_l + 1.0), us, ps
            else:  # left rarefaction
                head = u_l - c_l
                c_star = c_l * (ps / p_l) ** ((g - 1) / (2 * g))
                tail = us - c_star
                if s <= head:
                    r, u, p = rho_l, u_l, p_l
                elif s >= tail:
                    r, u, p = rho_l * (ps / p_l) ** (1.0 / g), us, ps
                else:
                    fac = 2.0 / (g + 1.0) + gm / c_l * (u_l - s)
                    r = rho_l * fac ** (2.0 / (g - 1.0))
                    u = 2.0 / (g + 1.0) * (c_l + (g - 1.0) / 2.0 * u_l + s)
                    p = p_l * fac ** (2.0 * g / (g - 1.0))
        else:  # right of the contact
            if ps > p_r:  # right shock
                S = u_r + c_r * math.sqrt((g + 1) / (2 * g) * ps / p_r + (g - 1) / (2 * g))
                if s >= S:
                    r, u, p = rho_r, u_r, p_r
                else:
                    r, u, p = rho_r * (ps / p_r + gm) / (gm * ps / p_r + 1.0), us, ps
            else:  # right rarefaction
                head = u_r + c_r
                c_star = c_r * (ps / p_r) ** ((g - 1) / (2 * g))
                tail = us + c_star
                if s >= head:
                    r, u, p = rho_r, u_r, p_r
                elif s <= tail:
                    r, u, p = rho_r * (ps / p_r) ** (1.0 / g), us, ps
                else:
                    fac = 2.0 / (g + 1.0) - gm / c_r * (u_r - s)
                    r = rho_r * fac ** (2.0 / (g - 1.0))
                    u = 2.0 / (g + 1.0) * (-c_r + (g - 1.0) / 2.0 * u_r + s)
                    p = p_r * fac ** (2.0 * g / (g - 1.0))
        out_rho[n], out_u[n], out_p[n] = r, u, p
    return out_rho, out_u, out_p
