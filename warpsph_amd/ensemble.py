"""Batched ensembles: many independent CRKSPH runs stepped together.

The reference steps one simulation per ``crkSPH_step`` call
(src/warpSPH/schemes/crkSPH.py:19-366).  An ensemble concatenates the
particles of its runs (run r = user indices ``offsets[r] .. offsets[r+1]``)
and binds the run table to the library (``crk_bind_runs``), so that every
pass covers all runs in one launch while each run sees only its own
particles: per-run neighbour grids, per-run support range for the Owen
update, per-run CFL dt and per-run Verlet tests.  A run of an ensemble
evolves as it would alone with the same dt.

``sample_ensemble`` is the data generator: two-state splits at a random line
in the periodic unit square, run sizes log-uniform.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional, Sequence

import numpy as np
import torch

from . import abi
from .configurations import CRKSPHConfig, buildConfig
from .enumTypes import AdaptiveSupportScheme, KernelFunctions, ViscositySwitch
from .systems import CompSPHState, CompSPHSystem


class Runs:
    """Run table of an ensemble and the per-particle dt of its current step."""

    def __init__(self, offsets: Sequence[int], device):
        off = np.asarray(offsets, dtype=np.int64)
        if off.ndim != 1 or off.size < 2 or off[0] != 0 or np.any(np.diff(off) <= 0):
            raise ValueError("offsets must start at 0 and increase strictly")
        if off[-1] >= 2 ** 31:
            raise ValueError("an ensemble holds fewer than 2^31 particles")
        self.offsets = off.astype(np.int32)
        self.n_runs = int(off.size - 1)
        self.n = int(off[-1])
        self.device = torch.device(device)
        self.dt: Optional[np.ndarray] = None   # per-run dt of the current step (float64)
        self.dtp: Optional[torch.Tensor] = None  # per-particle dt (float32, device)
        self.hmm = torch.empty(2 * self.n_runs, dtype=torch.float32, device=self.device)  # Owen support ranges

    @property
    def counts(self) -> np.ndarray:
        return np.diff(self.offsets.astype(np.int64))

    def bind(self, rt, dtp: Optional[torch.Tensor] = None):
        """crk_bind_runs with this table and ``dtp`` (default: the current step's)."""
        dtp = self.dtp if dtp is None else dtp
        abi.call("crk_bind_runs", rt.ctx, self.n_runs, C.c_void_p(self.offsets.ctypes.data), abi.ptr(dtp),
                 rt.stream())

    def set_dt(self, dt) -> np.ndarray:
        """Per-run dt (a scalar applies to every run); fills a new per-particle
        dt tensor (each step keeps its own, for the pair data exported later)."""
        from .runtime import runtime
        d = np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.n_runs,)).copy()
        if np.any(~(d > 0)):
            raise ValueError("every run's dt must be positive")
        if self.dt is not None and self.dtp is not None and np.array_equal(d, self.dt):
            return self.dt
        rt = runtime(self.device)
        dtp = torch.empty(self.n, dtype=torch.float32, device=self.device)
        self.bind(rt, dtp)
        abi.call("crk_runs_dt", rt.ctx, self.n, C.c_void_p(d.ctypes.data), abi.ptr(dtp), rt.stream())
        self.dt, self.dtp = d, dtp
        return d

    def run_ids(self) -> torch.Tensor:
        return torch.repeat_interleave(torch.arange(self.n_runs, dtype=torch.int32, device=self.device),
                                       torch.as_tensor(self.counts, device=self.device))


def cfl_dt(system: CompSPHSystem, config, scheme, prev: Optional[np.ndarray] = None) -> np.ndarray:
    """computeTimestep (timestep/compressible.py:12-66) of every run; sets the
    ensemble's per-particle dt and returns the per-run dt."""
    from .configurations import toCrkConfig
    from .runtime import runtime
    runs: Runs = system.runs
    st = system.state
    cfg = toCrkConfig(config, scheme)
    rt = runtime(st.positions.device)
    dtp = torch.empty(runs.n, dtype=torch.float32, device=runs.device)
    runs.bind(rt, dtp)
    out = np.empty(runs.n_runs, dtype=np.float64)
    pv = None
    if prev is not None:
        pv = np.ascontiguousarray(np.broadcast_to(np.asarray(prev, dtype=np.float64), (runs.n_runs,)))
    abi.call("crk_cfl_dt_runs", rt.ctx, C.byref(cfg), runs.n, abi.ptr(st.supports), abi.ptr(st.internalEnergies),
             C.c_void_p(pv.ctypes.data) if pv is not None else None, C.c_void_p(out.ctypes.data), abi.ptr(dtp),
             rt.stream())
    runs.dt, runs.dtp = out, dtp
    return out


_FIELDS = ("positions", "velocities", "supports", "masses", "densities", "kinds", "materials", "UIDs",
           "internalEnergies", "totalEnergies", "entropies", "pressures", "soundspeeds", "divergence", "alpha0s",
           "alphas")


def concatenate(systems: Sequence[CompSPHSystem]) -> CompSPHSystem:
    """One ensemble system from independent systems of the same domain and
    configuration (materials = run index, UIDs renumbered)."""
    if not systems:
        raise ValueError("no systems")
    dev = systems[0].state.positions.device
    counts = [s.state.numParticles for s in systems]
    offsets = np.concatenate([[0], np.cumsum(counts)])
    kw = {}
    for f in _FIELDS:
        vals = [getattr(s.state, f) for s in systems]
        if any(v is None for v in vals):
            kw[f] = None
            continue
        kw[f] = torch.cat([v.to(dev) for v in vals]).contiguous()
    n = int(offsets[-1])
    kw["materials"] = torch.repeat_interleave(torch.arange(len(systems), dtype=torch.int32),
                                              torch.tensor(counts)).to(dev)
    kw["UIDs"] = torch.arange(n, dtype=torch.int32, device=dev)
    st = CompSPHState(UIDcounter=n, **kw)
    sys = CompSPHSystem(state=st, adjacency=None, domain=systems[0].domain,
                        t=np.array([float(s.t) for s in systems]))
    sys.runs = Runs(offsets, dev)
    return sys


def run_system(system: CompSPHSystem, r: int) -> CompSPHSystem:
    """Run r of an ensemble as a standalone system (copies of its fields)."""
    runs: Runs = system.runs
    b, e = int(runs.offsets[r]), int(runs.offsets[r + 1])
    kw = {}
    for f in _FIELDS:
        v = getattr(system.state, f)
        kw[f] = None if v is None else v[b:e].clone()
    kw["UIDs"] = torch.arange(e - b, dtype=torch.int32, device=system.state.positions.device)
    t = system.t[r] if isinstance(system.t, np.ndarray) else system.t
    return CompSPHSystem(state=CompSPHState(UIDcounter=e - b, **kw), adjacency=None, domain=system.domain, t=float(t))


def two_state_run(nx: int, seed: int, device='cuda', jitter: float = 0.1, targetNeighbors: float = 32.0,
                  kernel=KernelFunctions.Wendland2, switch=ViscositySwitch.CullenDehnen2010,
                  gamma: float = 5.0 / 3.0):
    """One run of the generator: nx^2 particles on a jittered lattice of the
    periodic unit square, two states split at a random line (a point uniform in
    the square, a direction uniform in [0, pi)): rho, P ~ U[0.1, 1],
    v ~ U[-0.5, 0.5]^2 on each side.  The states depend on the seed only, so
    runs of one seed at different nx solve the same problem.  Returns
    (system, config, scheme)."""
    from .sample import _domain, _state_from, lattice
    g = torch.Generator().manual_seed(int(seed))
    # the states first: the same seed gives the same problem at every resolution
    u = torch.rand(11, generator=g, dtype=torch.float64)
    pos, dx = lattice((nx, nx), (0.0, 0.0), (1.0, 1.0), jitter=jitter, generator=g)
    p0 = u[0:2]
    th = math.pi * float(u[2])
    normal = torch.tensor([math.cos(th), math.sin(th)], dtype=torch.float64)
    side = ((pos.to(torch.float64) - p0) @ normal) > 0
    rho_ab = 0.1 + 0.9 * u[3:5]
    p_ab = 0.1 + 0.9 * u[5:7]
    v_ab = -0.5 + u[7:11].reshape(2, 2)
    pick = side.long()
    rho = rho_ab[pick].to(torch.float32)
    P = p_ab[pick].to(torch.float32)
    vel = v_ab[pick].to(torch.float32)
    dom = _domain([0.0, 0.0], [1.0, 1.0], [True, True], device)
    config, _ = buildConfig(device=torch.device(device), domain=dom, dim=2, kernel=kernel,
                            targetNeighbors=targetNeighbors, cflFactor=0.3)
    scheme = CRKSPHConfig(gamma=gamma, adaptiveSupportScheme=AdaptiveSupportScheme.Owen)
    scheme.viscositySwitchParams.scheme = switch
    st = _state_from(pos, rho, vel, P, gamma, dx, 2, targetNeighbors, device)
    config.dx = dx
    config.nx = nx
    return CompSPHSystem(state=st, adjacency=None, domain=dom, t=0.0), config, scheme


def run_sizes(n_runs: int, nmin: int = 4096, nmax: int = 65536, seed: int = 0) -> np.ndarray:
    """Lattice sides of the runs: N log-uniform in [nmin, nmax], nx = round(sqrt(N))
    kept within [ceil(sqrt(nmin)), floor(sqrt(nmax))]; run r has nx[r]^2 particles."""
    rng = np.random.default_rng(seed)
    N = np.exp(rng.uniform(math.log(nmin), math.log(nmax), n_runs))
    lo, hi = math.ceil(math.sqrt(nmin)), math.floor(math.sqrt(nmax))
    return np.clip(np.rint(np.sqrt(N)), lo, hi).astype(np.int64)


def sample_ensemble(n_runs: int = 256, nmin: int = 4096, nmax: int = 65536, seed: int = 0, device='cuda',
                    **kw):
    """The batched data generator: ``n_runs`` two-state runs (``two_state_run``,
    seeds ``seed * 100003 + r``) with log-uniform sizes, as one ensemble system.
    Returns (system, config, scheme) with ``system.runs`` set."""
    sides = run_sizes(n_runs, nmin, nmax, seed)
    systems = []
    config = scheme = None
    for r, nx in enumerate(sides):
        s, c, sc = two_state_run(int(nx), seed * 100003 + r, device=device, **kw)
        systems.append(s)
        if config is None:
            config, scheme = c, sc
    ens = concatenate(systems)
    ens.t = np.zeros(n_runs)
    return ens, config, scheme
