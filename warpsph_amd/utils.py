"""Support-radius helpers (and the export timestamp) with the names and signatures of
src/warpSPH/utils/support.py:6-44.

A support radius H around a particle of volume V holds the particles of the
ball of volume omega_d H^d, omega_d the volume of the unit ball (2, pi,
4 pi / 3 for d = 1, 2, 3).  So n_H = omega_d H^d / V neighbours, and with
n_h = H / dx particles per radius on a lattice of spacing dx (V = dx^d),
n_H = omega_d n_h^d.
"""
from __future__ import annotations

import math

import torch


def _unit_ball(dim: int) -> float:
    """omega_d = pi^(d/2) / Gamma(d/2 + 1)."""
    return math.pi ** (dim / 2) / math.gamma(dim / 2 + 1)


def _root(x, dim: int):
    """x^(1/d) for floats and tensors (the square root exactly where d = 2)."""
    if dim == 1:
        return x
    if isinstance(x, torch.Tensor):
        return torch.sqrt(x) if dim == 2 else torch.pow(x, 1.0 / dim)
    return math.sqrt(x) if dim == 2 else x ** (1.0 / dim)


def n_h_to_nH(n_h, dim):
    """Neighbours per support volume from particles per support radius."""
    return _unit_ball(dim) * n_h ** dim


def nH_to_n_h(nH, dim):
    """Particles per support radius from neighbours per support volume."""
    return _root(nH / _unit_ball(dim), dim)


def volumeToSupportHelper(volume, targetNeighbors, dim):
    """Support radius H holding ``targetNeighbors`` particles of ``volume``:
    H = (n_H V / omega_d)^(1/d)."""
    return _root(targetNeighbors * volume / _unit_ball(dim), dim)


def getCurrentTimestamp():
    """Local time as YYYY-MM-DD_HH-MM-SS (src/warpSPH/utils/util.py:3-4)."""
    import datetime
    return datetime.datetime.now().strftime("%Y-%m-%d_%H-%M-%S")
