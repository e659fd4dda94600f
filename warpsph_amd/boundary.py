"""Boundary-condition hooks of the CRKSPH step (modules/boundaryConditions/bcs.py:16-92).

A boundary condition carries a signed distance function ``sdf(positions) ->
(d, normals)`` and user callbacks ``fn(state, config, compParams, positions,
d, normals, t, dt) -> values``; wherever ``d < 0`` the callback's values
replace

* a state field (``dirichletFunctions``: ``{field name: fn}``), applied after
  the CRK factors and before the equation of state (crkSPH.py:94);
* a field of the update (``updateFunctions``), applied once the update is
  complete (crkSPH.py:339);

and ``forcingFunctions`` (a list) add accelerations, summed over the
conditions and divided by the masses (crkSPH.py:327-328).  Positions are
wrapped into the periodic domain before ``sdf`` is evaluated
(utils/math.py:5-11).  The callbacks run on the caller's user-order torch
tensors, as in the reference; ``crkSPH_step`` hands their changes back to the
library (``crk_state_refresh``, ``crk_update_refresh``).
"""
from __future__ import annotations

from typing import Any

import torch


def getPeriodicPositions(x: torch.Tensor, domain) -> torch.Tensor:
    """Positions wrapped into [min, max) along the periodic dimensions."""
    lo = domain.min.detach().to(x.device)
    hi = domain.max.detach().to(x.device)
    cols = []
    for d in range(int(domain.dim)):
        c = x[:, d]
        if bool(domain.periodic[d]):
            c = torch.remainder(c - lo[d], hi[d] - lo[d]) + lo[d]
        cols.append(c)
    return torch.stack(cols, dim=-1)


def _state_of(system: Any):
    return system.state if hasattr(system, "state") else system


def enforceDirichlet(system: Any, t: float, dt: float, config, compParams) -> bool:
    """bcs.py:16-36.  Returns whether any state field was set."""
    st = _state_of(system)
    changed = False
    for bc in compParams.boundaryConditions:
        fns = getattr(bc, "dirichletFunctions", None)
        if not fns:
            continue
        for name, fn in fns.items():
            var = getattr(st, name)
            pos = getPeriodicPositions(st.positions, config.domain)
            d, nrm = bc.sdf(pos)
            values = fn(st, config, compParams, pos, d, nrm, t, dt)
            inside = d < 0
            var[inside] = values[inside]
            changed = True
    return changed


def computeForcing(system: Any, dt: float, t: float, config, compParams) -> torch.Tensor:
    """bcs.py:38-62: the summed forcing accelerations times mass ([N, d])."""
    st = _state_of(system)
    total = torch.zeros_like(st.positions)
    for bc in compParams.boundaryConditions:
        fns = getattr(bc, "forcingFunctions", None)
        if not fns:
            continue
        for fn in fns:
            pos = getPeriodicPositions(st.positions, config.domain)
            d, nrm = bc.sdf(pos)
            values = fn(st, config, compParams, pos, d, nrm, t, dt)
            inside = d < 0
            total[inside] += values[inside]
    return total


def enforceUpdates(updates: Any, system: Any, dt: float, t: float, config, compParams) -> bool:
    """bcs.py:64-92.  Returns whether any update field was set."""
    st = _state_of(system)
    changed = False
    for bc in compParams.boundaryConditions:
        fns = getattr(bc, "updateFunctions", None)
        if not fns:
            continue
        for name, fn in fns.items():
            var = getattr(updates, name)
            pos = getPeriodicPositions(st.positions, config.domain)
            d, nrm = bc.sdf(pos)
            values = fn(st, config, compParams, pos, d, nrm, t, dt)
            inside = d < 0
            var[inside] = values[inside]
            changed = True
    return changed
