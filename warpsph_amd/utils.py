"""Support-radius helpers, src/warpSPH/utils/support.py:6-44."""
from __future__ import annotations

import numpy as np
import torch


def n_h_to_nH(n_h, dim):
    """Particles per support volume from particles per support radius."""
    spacing = 1 / n_h
    v = spacing ** dim
    vH = 2.0 if dim == 1 else (np.pi if dim == 2 else (4 / 3) * np.pi)
    return vH / v


def nH_to_n_h(nH, dim):
    vH = 2.0 if dim == 1 else (np.pi if dim == 2 else (4 / 3) * np.pi)
    v = vH / nH
    return (1 / v) ** (1 / dim)


def volumeToSupportHelper(volume, targetNeighbors, dim):
    """Support radius H holding ``targetNeighbors`` particles of ``volume``."""
    if dim == 1:
        return targetNeighbors * volume / 2
    elif dim == 2:
        if isinstance(volume, torch.Tensor):
            return torch.sqrt(targetNeighbors * volume / np.pi)
        return np.sqrt(targetNeighbors * volume / np.pi)
    else:
        if isinstance(volume, torch.Tensor):
            return torch.pow(targetNeighbors * volume / np.pi * 3 / 4, 1 / 3)
        return np.power(targetNeighbors * volume / np.pi * 3 / 4, 1 / 3)
