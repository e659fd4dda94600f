"""Exact solution of the Riemann problem for the ideal-gas Euler equations
(Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics", 3rd ed.,
ch. 4: pressure function f_K, Newton iteration for p*, wave-pattern
sampling).  Test infrastructure: the known-answer fixture for the Sod shock
tube is written by make_fixtures.py next to this file.

The reference keeps its own numpy version of the same textbook solver
(caseUtils/compressible/sod/sodSolution.py:200-272); this file is an
independent restatement used to pin the CPU oracle's Sod run.
"""
from __future__ import annotations

import math

import numpy as np


def _fK(p, rho, pK, cK, g):
    """Pressure function f_K(p) and its derivative (Toro eq. 4.6, 4.7, 4.37)."""
    if p > pK:  # shock
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * pK
        s = math.sqrt(A / (p + B))
        return (p - pK) * s, s * (1.0 - 0.5 * (p - pK) / (p + B))
    # rarefaction
    e = (g - 1.0) / (2.0 * g)
    f = 2.0 * cK / (g - 1.0) * ((p / pK) ** e - 1.0)
    return f, 1.0 / (rho * cK) * (p / pK) ** (-(g + 1.0) / (2.0 * g))


def star_state(rl, ul, pl, rr, ur, pr, g):
    """p*, u* by Newton iteration on f_L + f_R + du = 0 (Toro sec. 4.3.2)."""
    cl = math.sqrt(g * pl / rl)
    cr = math.sqrt(g * pr / rr)
    p = max(1e-10, 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (cl + cr))  # PVRS guess
    for _ in range(100):
        fl, dl = _fK(p, rl, pl, cl, g)
        fr, dr = _fK(p, rr, pr, cr, g)
        dp = (fl + fr + (ur - ul)) / (dl + dr)
        pn = max(1e-10, p - dp)
        if abs(pn - p) / (0.5 * (pn + p)) < 1e-14:
            p = pn
            break
        p = pn
    fl, _ = _fK(p, rl, pl, cl, g)
    fr, _ = _fK(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(xi, rl, ul, pl, rr, ur, pr, g):
    """(rho, u, p) at similarity coordinate xi = x / t (Toro sec. 4.5)."""
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl = math.sqrt(g * pl / rl)
    cr = math.sqrt(g * pr / rr)
    gm = (g - 1.0) / (g + 1.0)
    out = np.empty((len(xi), 3))
    for k, s in enumerate(xi):
        if s <= us:  # left of the contact
            if ps > pl:  # left shock
                sl = ul - cl * math.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                out[k] = (rl, ul, pl) if s <= sl else (rl * (ps / pl + gm) / (gm * ps / pl + 1), us, ps)
            else:  # left rarefaction
                csl = cl * (ps / pl) ** ((g - 1) / (2 * g))
                if s <= ul - cl:
                    out[k] = (rl, ul, pl)
                elif s >= us - csl:
                    out[k] = (rl * (ps / pl) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    r = rl * (c / cl) ** (2 / (g - 1))
                    out[k] = (r, u, pl * (c / cl) ** (2 * g / (g - 1)))
        else:  # right of the contact
            if ps > pr:  # right shock
                sr = ur + cr * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                out[k] = (rr, ur, pr) if s >= sr else (rr * (ps / pr + gm) / (gm * ps / pr + 1), us, ps)
            else:  # right rarefaction
                csr = cr * (ps / pr) ** ((g - 1) / (2 * g))
                if s >= ur + cr:
                    out[k] = (rr, ur, pr)
                elif s <= us + csr:
                    out[k] = (rr * (ps / pr) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    r = rr * (c / cr) ** (2 / (g - 1))
                    out[k] = (r, u, pr * (c / cr) ** (2 * g / (g - 1)))
    return out


def doubled_tube(x, t, g=1.4):
    """The periodic doubled Sod tube on [-1, 1] (high state on [-0.5, 0.5]):
    two mirrored Riemann problems at x = +-0.5, exact while their waves have
    not met (t <= 0.2)."""
    x = np.asarray(x, dtype=np.float64)
    out = np.empty((len(x), 3))
    right = x >= 0
    out[right] = sample((x[right] - 0.5) / t, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1, g)
    left = ~right
    mir = sample((-x[left] - 0.5) / t, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1, g)
    mir[:, 1] = -mir[:, 1]
    out[left] = mir
    return out
