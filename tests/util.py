"""Shared helpers of the parity tests."""
from __future__ import annotations

import numpy as np
import torch

# Per-step FP32 agreement bar: max |gpu - oracle| <= RTOL * max |oracle| per
# field, and per component for a vector field (an (n, 2) or (n, 3) array is
# measured column by column, so a small component such as the transverse
# velocity of a shear is held to its own scale, not to the largest one).
RTOL = 1e-5


def relerr(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if b.ndim == 2 and 1 < b.shape[1] <= 3 and a.shape == b.shape:
        return max(relerr(a[:, c], b[:, c]) for c in range(b.shape[1]))
    scale = np.max(np.abs(b)) if b.size else 0.0
    if scale == 0.0:
        return float(np.max(np.abs(a))) if a.size else 0.0
    return float(np.max(np.abs(a - b)) / scale)


def cpu_copy(system, config, scheme):
    """The same system with CPU tensors (oracle input)."""
    from warpsph_amd.systems import CompSPHSystem
    st = system.state.initializeNewState()
    for f in ("positions", "velocities", "supports", "masses", "densities", "internalEnergies", "entropies",
              "pressures", "soundspeeds", "divergence", "alpha0s", "alphas"):
        t = getattr(st, f)
        if t is not None:
            setattr(st, f, t.detach().cpu())
    return CompSPHSystem(state=st, adjacency=None, domain=system.domain, t=system.t)


def edge_order(adj):
    """Indices that put the GPU edge storage (sorted rows) into user-row order."""
    nn = adj.numNeighbors.cpu()
    no = adj.neighborOffsets.cpu()
    E = int(nn.sum())
    user_off = torch.zeros(adj.n + 1, dtype=torch.int64)
    user_off[1:] = torch.cumsum(nn, 0)
    within = torch.arange(E, dtype=torch.int64) - torch.repeat_interleave(user_off[:-1], nn)
    return (torch.repeat_interleave(no, nn) + within).numpy(), user_off.numpy()


def user_rows(adj):
    """(offsets, neighbour user ids) of a GPU adjacency in user-row order."""
    ep, off = edge_order(adj)
    j = adj.j.cpu().numpy()[ep]
    return off, j, ep
