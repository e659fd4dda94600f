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
    weights = (0.5, 0.5)

    def function(self, state, f, dt, config, compParams, verbose=False, **kwargs) -> IntegrationResult:
        from .runtime import runtime
        initial = state
        st0 = initial.state
        dt = float(dt)
        rt = runtime(st0.positions.device)
        s = rt.stream()
        # stage 1 reads the initial positions, velocities and energies
        sys1 = initial.stageSystem("share")
        upd1, adj1, st1 = f(sys1, dt, config, compParams, verbose=verbose)
        # stage 2 starts from stage 1's state with y0 + dt k1
        sys2 = sys1.stageSystem("alloc")
        sys2.adjacency = adj1
        s2 = sys2.state
        nD = st0.positions.numel()
        n = st0.positions.shape[0]
        abi.call("crk_axpy", nD, abi.ptr(st0.positions), abi.ptr(upd1.dxdt), dt, abi.ptr(s2.positions), s)
        abi.call("crk_axpy", nD, abi.ptr(st0.velocities), abi.ptr(upd1.dvdt), dt, abi.ptr(s2.velocities), s)
        abi.call("crk_axpy", n, abi.ptr(st0.internalEnergies), abi.ptr(upd1.dudt), dt, abi.ptr(s2.internalEnergies), s)
        upd2, adj2, stt2 = f(sys2, dt, config, compParams, verbose=verbose)
        # final state
        b0, b1 = self.weights
        fin = sys2
        x_new = st0.positions.new_empty(st0.positions.shape)
        v_new = st0.velocities.new_empty(st0.velocities.shape)
        abi.call("crk_rk_combine", nD, abi.ptr(st0.positions), abi.ptr(upd1.dxdt), abi.ptr(upd2.dxdt), b0, b1, dt,
                 abi.ptr(x_new), s)
        abi.call("crk_rk_combine", nD, abi.ptr(st0.velocities), abi.ptr(upd1.dvdt), abi.ptr(upd2.dvdt), b0, b1, dt,
                 abi.ptr(v_new), s)
        if not compParams.compatibleEnergy:
            u_new = st0.internalEnergies.new_empty(st0.internalEnergies.shape)
            abi.call("crk_rk_combine", n, abi.ptr(st0.internalEnergies), abi.ptr(upd1.dudt), abi.ptr(upd2.dudt),
                     b0, b1, dt, abi.ptr(u_new), s)
            fin.state.internalEnergies = u_new
        stages = [(adj1, st1), (adj2, stt2)]
        fin.finalize(initial, dt, stages, [upd1, upd2], [b0, b1], config, compParams)
        fin.state.positions = x_new
        fin.state.velocities = v_new
        from .configurations import toCrkConfig
        cfg = toCrkConfig(config, compParams)
        abi.call("crk_wrap_positions", C.byref(cfg), n, abi.ptr(x_new), s)
        fin.t = initial.t + dt
        return IntegrationResult(state=fin, stages=[StageResult(aux=[adj1, st1], update=upd1),
                                                    StageResult(aux=[adj2, stt2], update=upd2)])


def getIntegrator(scheme=IntegrationSchemeType.rungeKutta2):
    if scheme != IntegrationSchemeType.rungeKutta2:
        raise ValueError(f"integration scheme {scheme} is not built (rungeKutta2)")
    return RungeKutta2()
