"""Two-stage Runge-Kutta driver with the warpSPHIntegrators call protocol.

The reference obtains it from ``getIntegrator(rungeKutta2)``
(src/warpSPH/configurations/simulationConfig.py:121) and calls
``integrator.function(state, f, dt, config, compParams, verbose)``
(examples/compressible/01-sod-shock-tube-1d.py:196-204); the result exposes
``.state`` and ``.stages`` (StageResult(aux=[adjacency, state], update),
src/warpSPH/io.py:174-193) and the system's ``finalize`` is called with the
stage results and the Butcher weights (systems/compSPH.py:87).

warpSPHIntegrators is not available, so the tableau is this library's
choice: Heun's method, c = (0, 1), b = (1/2, 1/2).  Non-integrated fields of
stage 2 start from stage 1's values.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Any, List

import numpy as np

from . import abi
from .enumTypes import IntegrationSchemeType


@dataclass
class StageResult:
    aux: List[Any]
    update: Any


@dataclass
class IntegrationResult:
    state: Any
    stages: List[StageResult] = field(default_factory=list)


class RungeKutta2:
    nodes = (0.0, 1.0)
    weights = (0.5, 0.5)

    def function(self, state, f, dt, config, compParams, verbose=False, **kwargs) -> IntegrationResult:
        from .runtime import runtime
        initial = state
        st0 = initial.state
        runs = getattr(initial, "runs", None)
        rt = runtime(st0.positions.device)
        s = rt.stream()
        nD = st0.positions.numel()
        n = st0.positions.shape[0]
        D = nD // n
        device_dt = isinstance(dt, str) and dt == "device"
        if runs is not None:
            # an ensemble: every run advances by its own dt (each particle's dt)
            if not device_dt:
                dt = runs.set_dt(dt.detach().cpu().numpy() if hasattr(dt, "detach") else dt)
            dtp = runs.dtp

            def axpy(cnt, y0, d, y):
                abi.call("crk_axpy_p", n, cnt // n, abi.ptr(y0), abi.ptr(d), abi.ptr(dtp), abi.ptr(y), s)

            def combine(cnt, y0, d0, d1, b0, b1, y):
                abi.call("crk_rk_combine_p", n, cnt // n, abi.ptr(y0), abi.ptr(d0), abi.ptr(d1), b0, b1,
                         abi.ptr(dtp), abi.ptr(y), s)
        else:
            dt = float(dt)

            def axpy(cnt, y0, d, y):
                abi.call("crk_axpy", cnt, abi.ptr(y0), abi.ptr(d), dt, abi.ptr(y), s)

            def combine(cnt, y0, d0, d1, b0, b1, y):
                abi.call("crk_rk_combine", cnt, abi.ptr(y0), abi.ptr(d0), abi.ptr(d1), b0, b1, dt, abi.ptr(y), s)
        # stage 1 reads the initial positions, velocities and energies
        sys1 = initial.stageSystem("share")
        upd1, adj1, st1 = f(sys1, dt, config, compParams, verbose=verbose)
        # stage 2 starts from stage 1's state with y0 + dt k1
        sys2 = sys1.stageSystem("alloc")
        sys2.adjacency = adj1
        if not device_dt:
            sys2.t = initial.t + self.nodes[1] * dt  # the stage's time t + c_2 dt (read by boundary callbacks)
        s2 = sys2.state
        axpy(nD, st0.positions, upd1.dxdt, s2.positions)
        axpy(nD, st0.velocities, upd1.dvdt, s2.velocities)
        axpy(n, st0.internalEnergies, upd1.dudt, s2.internalEnergies)
        upd2, adj2, stt2 = f(sys2, dt, config, compParams, verbose=verbose)
        # final state
        b0, b1 = self.weights
        fin = sys2
        x_new = st0.positions.new_empty(st0.positions.shape)
        v_new = st0.velocities.new_empty(st0.velocities.shape)
        combine(nD, st0.positions, upd1.dxdt, upd2.dxdt, b0, b1, x_new)
        combine(nD, st0.velocities, upd1.dvdt, upd2.dvdt, b0, b1, v_new)
        if not compParams.compatibleEnergy:
            u_new = st0.internalEnergies.new_empty(st0.internalEnergies.shape)
            combine(n, st0.internalEnergies, upd1.dudt, upd2.dudt, b0, b1, u_new)
            fin.state.internalEnergies = u_new
        stages = [(adj1, st1), (adj2, stt2)]
        fin.finalize(initial, 1.0 if device_dt else (float(np.max(dt)) if runs is not None else dt), stages,
                     [upd1, upd2], [b0, b1], config, compParams)
        fin.state.positions = x_new
        fin.state.velocities = v_new
        from .configurations import toCrkConfig
        cfg = toCrkConfig(config, compParams)
        abi.call("crk_wrap_positions", C.byref(cfg), n, abi.ptr(x_new), s)
        fin.t = initial.t if device_dt else initial.t + dt  # (device dt: the caller keeps the times)
        if runs is not None:
            rt.bind_runs(None)  # direct library calls after the step see one simulation
        return IntegrationResult(state=fin, stages=[StageResult(aux=[adj1, st1], update=upd1),
                                                    StageResult(aux=[adj2, stt2], update=upd2)])


def getIntegrator(scheme=IntegrationSchemeType.rungeKutta2):
    if scheme != IntegrationSchemeType.rungeKutta2:
        raise ValueError(f"integration scheme {scheme} is not built (rungeKutta2)")
    return RungeKutta2()
