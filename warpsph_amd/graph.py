"""RK2 steps replayed as HIP graphs.

One step of a system (a single run or a batched ensemble) is captured once
per adjacency with torch's graph capture: the per-run CFL dt on the device
(``crk_cfl_dt_device``), both derivative evaluations with their Verlet tests
deferred to the context's device flag, the stage update, the
compatible-energy finalize, and a commit that copies the new state over the
input state (and advances each run's time) only while the flag is clear
(``crk_copy_unless_flag``, ``crk_runs_advance_time``).  A replay therefore
never needs the host; afterwards the host reads the flag once.  When some
run's Verlet list failed, the input state is untouched: the step is redone
eagerly (which rebuilds the lists) and the graph is captured again for the
new adjacency.  Replayed and eager steps give the same bits.  While the
recent lists live fewer than ``min_life`` steps on average, steps run
eagerly instead: a capture used once or twice costs more than it saves.

No counterpart in the reference (it launches every pass from Python).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np
import torch

from . import abi
from .runtime import runtime

# fields of the state a step writes (the rest are read only)
_COMMIT = ("positions", "velocities", "internalEnergies", "supports", "densities", "entropies", "pressures",
           "soundspeeds", "divergence", "alpha0s", "alphas")


class GraphStepper:
    """Steps ``system`` in place with graph replays (see the module text)."""

    def __init__(self, system, config, scheme, integrator=None):
        from .configurations import toCrkConfig
        from .ensemble import Runs
        from .integrator import getIntegrator
        st = system.state
        n = st.numParticles
        self.dev = st.positions.device
        if system.runs is None:  # one run: a run table of one
            system.runs = Runs([0, n], self.dev)
            system.t = np.array([float(system.t)])
        self.system = system
        self.config, self.scheme = config, scheme
        self.cfg = toCrkConfig(config, scheme)
        self.integ = integrator or getIntegrator()
        runs = system.runs
        R = runs.n_runs
        self.dt_prev = torch.zeros(R, dtype=torch.float32, device=self.dev)  # 0: no growth limit yet
        self.dt_new = torch.empty(R, dtype=torch.float32, device=self.dev)
        self.dtp = torch.empty(n, dtype=torch.float32, device=self.dev)
        self.t = torch.tensor(np.asarray(system.t, dtype=np.float64).reshape(R), device=self.dev)
        self.graph: Optional[torch.cuda.CUDAGraph] = None
        self._graph_adj = None
        self._pool = torch.cuda.graph_pool_handle()  # one memory pool for every capture
        self.replays = self.captures = self.redos = 0
        # lists that live fewer steps than this (recent mean) are stepped
        # eagerly.  A list of L steps costs one capture and one wasted replay
        # and saves L-1 times (eager step - replay): on MI355X a 128^2 run
        # captures in 2.5 ms and replays a step in 0.28 ms against 0.53 ms
        # eager, which breaks even near L = 12 (larger runs later still)
        self.min_life = 12.0
        self._lives = []  # steps each recent adjacency lasted
        self._adj_seen, self._adj_steps = None, 0
        self.eager_steps = 0

    # ---------------------------------------------------------------- body
    def _body(self):
        from .scheme import DEVICE_DT, crkSPH_step
        S = self.system
        st = S.state
        runs = S.runs
        rt = runtime(self.dev)
        s = rt.stream()
        runs.dtp, runs.dt = self.dtp, None
        runs.bind(rt, self.dtp)
        abi.call("crk_flag_clear", rt.ctx, s)
        abi.call("crk_cfl_dt_device", rt.ctx, C.byref(self.cfg), runs.n, abi.ptr(st.supports),
                 abi.ptr(st.internalEnergies), abi.ptr(self.dt_prev), abi.ptr(self.dt_new), abi.ptr(self.dtp), s)
        rt.deferred = True
        try:
            out = self.integ.function(S, crkSPH_step, DEVICE_DT, self.config, self.scheme).state
        finally:
            rt.deferred = False
        runs.bind(rt, self.dtp)  # (the integrator unbinds at its end)
        for f in _COMMIT:
            src, dst = getattr(out.state, f), getattr(st, f)
            if src is not None and dst is not None and src.data_ptr() != dst.data_ptr():
                abi.call("crk_copy_unless_flag", rt.ctx, abi.ptr(src), abi.ptr(dst), src.numel() * 4, s)
        abi.call("crk_runs_advance_time", rt.ctx, abi.ptr(self.dt_new), abi.ptr(self.t), s)
        abi.call("crk_copy_unless_flag", rt.ctx, abi.ptr(self.dt_new), abi.ptr(self.dt_prev),
                 self.dt_new.numel() * 4, s)
        rt.bind_runs(None)
        return out

    def _flag(self) -> int:
        rt = runtime(self.dev)
        f = C.c_int32()
        abi.call("crk_flag_read", rt.ctx, C.byref(f), rt.stream())
        return f.value

    def _redo_eager(self):
        """The step with host decisions (rebuilds) from the untouched input."""
        from .ensemble import cfl_dt
        from .scheme import crkSPH_step
        S = self.system
        prev = self.dt_prev.double().cpu().numpy()  # (0: no growth limit)
        dts = cfl_dt(S, self.config, self.scheme, prev)
        out = self.integ.function(S, crkSPH_step, dts, self.config, self.scheme).state
        st = S.state
        with torch.no_grad():
            for f in _COMMIT:
                src, dst = getattr(out.state, f), getattr(st, f)
                if src is not None and dst is not None and src.data_ptr() != dst.data_ptr():
                    dst.copy_(src)
        S.adjacency = out.adjacency
        d32 = torch.tensor(dts.astype(np.float32), device=self.dev)
        self.dt_prev.copy_(d32)
        self.t += d32.double()

    def _capture(self):
        """The step as a graph for the current adjacency (no work runs)."""
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, pool=self._pool):
            self._body()
        self.graph, self._graph_adj = g, self.system.adjacency
        self.captures += 1

    # ---------------------------------------------------------------- steps
    def step(self):
        """One RK2 step of every run."""
        S = self.system
        if S.adjacency is not self._adj_seen:  # (a rebuild since the last step)
            if self._adj_seen is not None:
                self._lives = (self._lives + [self._adj_steps])[-4:]
            self._adj_seen, self._adj_steps = S.adjacency, 0
        self._adj_steps += 1
        if S.adjacency is None or (len(self._lives) >= 2 and sum(self._lives) / len(self._lives) < self.min_life):
            # the first step builds the lists; short-lived lists: eager steps
            self.graph = None
            self._redo_eager()
            self.eager_steps += 1
            return
        if self.graph is None or self._graph_adj is not S.adjacency:
            # a new adjacency: this step runs the deferred body eagerly (which
            # also allocates every buffer at its size), then the graph is
            # captured for the following steps
            self.graph = None  # (its memory returns to the pool)
            self._body()
            if self._flag():
                self._redo_eager()
                self.redos += 1
            else:
                self._capture()
            return
        self.graph.replay()
        self.replays += 1
        if self._flag():
            self._redo_eager()
            self.redos += 1

    def times(self) -> np.ndarray:
        return self.t.cpu().numpy()
