"""Initial conditions of the CRKSPH workloads.

The lattice follows sampleRegularParticles / buildPointCloud
(src/warpSPH/sample/regular.py:6-56): cell-centred in periodic dimensions,
``meshgrid(indexing='xy')`` ordering (x fastest), one-sided jitter
``rand * dx * jitter``, support from volumeToSupportHelper, mass dx^d rho.
All random numbers come from a seeded CPU ``torch.Generator``; tensors are
generated on the CPU and then moved to the target device.
"""
from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy as np
import torch

from .configurations import CRKSPHConfig, DomainDescription, SimulationConfig, buildConfig
from .enumTypes import AdaptiveSupportScheme, KernelFunctions, ViscositySwitch
from .systems import CompSPHState, CompSPHSystem
from .utils import volumeToSupportHelper


def _domain(lo, hi, periodic, device):
    lo = torch.tensor(lo, dtype=torch.float32)
    hi = torch.tensor(hi, dtype=torch.float32)
    return DomainDescription(lo.to(device), hi.to(device),
                             torch.tensor(periodic, dtype=torch.bool, device=device), len(lo))


def lattice(ns, lo, hi, jitter=0.0, generator=None):
    """Cell-centred lattice of prod(ns) points in [lo, hi) (x fastest)."""
    dim = len(ns)
    axes = []
    for d in range(dim):
        dx = (hi[d] - lo[d]) / ns[d]
        axes.append(torch.linspace(lo[d] + dx / 2, hi[d] - dx / 2, ns[d], dtype=torch.float32))
    # x varies fastest, as meshgrid(..., indexing='xy') does in 2D
    rev = torch.meshgrid(*reversed(axes), indexing='ij')
    pos = torch.stack([g.flatten() for g in reversed(rev)], dim=1)
    dx = (hi[0] - lo[0]) / ns[0]
    if jitter > 0:
        pos = pos + torch.rand(pos.shape, generator=generator, dtype=torch.float32) * dx * jitter
    return pos, dx


def _state_from(pos, rho, vel, P, gamma, dx, dim, targetNeighbors, device, masses=None):
    n = pos.shape[0]
    h = float(volumeToSupportHelper(dx ** dim, targetNeighbors, dim))
    u = P / (rho * (gamma - 1.0))
    cs = torch.sqrt(u.abs() * gamma * (gamma - 1.0))
    A = u * (gamma - 1.0) * rho ** (1.0 - gamma)
    m = masses if masses is not None else rho * dx ** dim
    st = CompSPHState(
        positions=pos.to(device).contiguous(),
        velocities=vel.to(device).contiguous(),
        supports=torch.full((n,), h, dtype=torch.float32).to(device),
        masses=m.to(torch.float32).to(device).contiguous(),
        densities=rho.to(torch.float32).to(device).contiguous(),
        kinds=torch.zeros(n, dtype=torch.int32, device=device),
        materials=torch.zeros(n, dtype=torch.int32, device=device),
        UIDs=torch.arange(n, dtype=torch.int32, device=device),
        UIDcounter=n,
        internalEnergies=u.to(torch.float32).to(device).contiguous(),
        totalEnergies=((u + 0.5 * (vel ** 2).sum(-1)) * m).to(torch.float32).to(device),
        entropies=A.to(torch.float32).to(device).contiguous(),
        pressures=P.to(torch.float32).to(device).contiguous(),
        soundspeeds=cs.to(torch.float32).to(device).contiguous(),
        divergence=torch.zeros(n, dtype=torch.float32, device=device),
        alpha0s=torch.ones(n, dtype=torch.float32, device=device),
        alphas=torch.ones(n, dtype=torch.float32, device=device),
    )
    return st


def cfl_dt(state: CompSPHState, config: SimulationConfig, scheme: CRKSPHConfig, prev: Optional[float] = None):
    """computeTimestep (timestep/compressible.py:12-66) on the GPU."""
    import ctypes as C
    from . import abi
    from .configurations import toCrkConfig
    from .runtime import runtime
    cfg = toCrkConfig(config, scheme)
    rt = runtime(state.positions.device)
    out = C.c_double()
    abi.call("crk_cfl_dt", rt.ctx, C.byref(cfg), state.numParticles, abi.ptr(state.supports),
             abi.ptr(state.internalEnergies), float(prev) if prev else -1.0, C.byref(out), rt.stream())
    return out.value


def conservation(state: CompSPHState) -> dict:
    """The driver's conservation sums (examples/compressible/01-sod-shock-tube-1d.py:158-160,
    212-214) in FP64 on the GPU: total mass, momentum, internal, kinetic and
    total energy."""
    import ctypes as C
    from . import abi
    from .runtime import runtime
    n, dim = state.positions.shape
    rt = runtime(state.positions.device)
    out = (C.c_double * (dim + 3))()
    abi.call("crk_conservation", rt.ctx, dim, n, abi.ptr(state.masses), abi.ptr(state.velocities),
             abi.ptr(state.internalEnergies), out, rt.stream())
    vals = list(out)
    return {"mass": vals[0], "momentum": np.array(vals[1:dim + 1]), "internal": vals[dim + 1],
            "kinetic": vals[dim + 2], "total": vals[dim + 1] + vals[dim + 2]}


def shear_layer_2d(nx: int = 1024, seed: int = 2, device='cuda', jitter: float = 0.1, targetNeighbors: float = 32.0,
                   kernel=KernelFunctions.Wendland2, switch=ViscositySwitch.CullenDehnen2010, gamma: float = 5.0 / 3.0):
    """2D periodic shear layer on [0,1]^2, nx^2 particles: band |y-0.5| < 0.25
    has rho 2, vx +0.5; outside rho 1, vx -0.5; vy = 0.01 sin(4 pi x); P = 2.5.
    (Kelvin-Helmholtz setup of caseUtils/compressible/kelvinHelmholtz/sample.py:17-75
    with sharp interfaces.)  Returns (system, config, schemeConfig) for CPU
    tensors when device='cpu' (oracle inputs) or GPU tensors."""
    g = torch.Generator().manual_seed(seed)
    pos, dx = lattice((nx, nx), (0.0, 0.0), (1.0, 1.0), jitter=jitter, generator=g)
    y = pos[:, 1]
    band = (y - 0.5).abs() < 0.25
    rho = torch.where(band, torch.tensor(2.0), torch.tensor(1.0)).to(torch.float32)
    vx = torch.where(band, torch.tensor(0.5), torch.tensor(-0.5)).to(torch.float32)
    vy = (0.01 * torch.sin(4 * math.pi * pos[:, 0].to(torch.float64))).to(torch.float32)
    vel = torch.stack([vx, vy], dim=1)
    P = torch.full_like(rho, 2.5)
    dom = _domain([0.0, 0.0], [1.0, 1.0], [True, True], device if device != 'cpu' else 'cpu')
    config, _ = buildConfig(device=torch.device(device), domain=dom, dim=2, kernel=kernel,
                            targetNeighbors=targetNeighbors, cflFactor=0.3)
    scheme = CRKSPHConfig(gamma=gamma, adaptiveSupportScheme=AdaptiveSupportScheme.Owen)
    scheme.viscositySwitchParams.scheme = switch
    st = _state_from(pos, rho, vel, P, gamma, dx, 2, targetNeighbors, device)
    config.dx = dx
    config.nx = nx
    return CompSPHSystem(state=st, adjacency=None, domain=dom, t=0.0), config, scheme


def sod_1d(n_left: int = 889, ratio: int = 8, device='cuda', targetNeighbors: float = 6.0,
           kernel=KernelFunctions.Wendland4, switch=ViscositySwitch.CullenDehnen2010, gamma: float = 1.4):
    """Sod shock tube as the reference's periodic doubled tube
    (caseUtils/compressible/sod/sod.py:21-80): high state (rho 1, P 1) on
    [-0.5, 0.5], low state (rho 0.125, P 0.1) on [-1, -0.5) u [0.5, 1), equal
    particle masses (spacing ratio 8), domain [-1, 1] periodic.  889 + 111 =
    1000 particles."""
    n_right = n_left // ratio
    dxl = 1.0 / n_left
    xl = -0.5 + dxl / 2 + dxl * torch.arange(n_left, dtype=torch.float64)
    dxr = 1.0 / n_right
    xr = -0.5 + dxr / 2 + dxr * torch.arange(n_right, dtype=torch.float64)
    xr = torch.where(xr < 0, xr - 0.5, xr + 0.5)
    pos = torch.cat([xl, xr]).to(torch.float32)[:, None]
    hi_state = torch.cat([torch.ones(n_left, dtype=torch.bool), torch.zeros(n_right, dtype=torch.bool)])
    rho = torch.where(hi_state, torch.tensor(1.0), torch.tensor(0.125)).to(torch.float32)
    P = torch.where(hi_state, torch.tensor(1.0), torch.tensor(0.1)).to(torch.float32)
    vel = torch.zeros((pos.shape[0], 1), dtype=torch.float32)
    m = torch.full_like(rho, dxl * 1.0)
    dom = _domain([-1.0], [1.0], [True], device)
    config, _ = buildConfig(device=torch.device(device), domain=dom, dim=1, kernel=kernel,
                            targetNeighbors=targetNeighbors, cflFactor=0.3)
    scheme = CRKSPHConfig(gamma=gamma, adaptiveSupportScheme=AdaptiveSupportScheme.Owen)
    scheme.viscositySwitchParams.scheme = switch
    st = _state_from(pos, rho, vel, P, gamma, dxl, 1, targetNeighbors, device, masses=m)
    # support from the local spacing (left: dxl, right: dxr)
    h = torch.where(hi_state, torch.tensor(float(volumeToSupportHelper(dxl, targetNeighbors, 1))),
                    torch.tensor(float(volumeToSupportHelper(dxr, targetNeighbors, 1)))).to(torch.float32)
    st.supports = h.to(device)
    config.dx = dxl
    config.nx = n_left
    return CompSPHSystem(state=st, adjacency=None, domain=dom, t=0.0), config, scheme


def blast_3d(nx: int = 128, seed: int = 3, device='cuda', jitter: float = 0.1, targetNeighbors: float = 64.0,
             kernel=KernelFunctions.Wendland4, switch=ViscositySwitch.CullenDehnen2010, gamma: float = 5.0 / 3.0,
             energy: float = 1.0):
    """3D point blast in a periodic unit cube: rho 1, P 1e-6, energy E
    deposited kernel-weighted on particles within 2h of the centre."""
    g = torch.Generator().manual_seed(seed)
    pos, dx = lattice((nx, nx, nx), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), jitter=jitter, generator=g)
    n = pos.shape[0]
    rho = torch.ones(n, dtype=torch.float32)
    P = torch.full((n,), 1e-6, dtype=torch.float32)
    vel = torch.zeros((n, 3), dtype=torch.float32)
    dom = _domain([0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [True, True, True], device)
    config, _ = buildConfig(device=torch.device(device), domain=dom, dim=3, kernel=kernel,
                            targetNeighbors=targetNeighbors, cflFactor=0.3)
    scheme = CRKSPHConfig(gamma=gamma, adaptiveSupportScheme=AdaptiveSupportScheme.Owen)
    scheme.viscositySwitchParams.scheme = switch
    st = _state_from(pos, rho, vel, P, gamma, dx, 3, targetNeighbors, 'cpu')
    h = float(st.supports[0])
    r = torch.linalg.norm(pos.to(torch.float64) - 0.5, dim=1)
    q = r / (2 * h)
    w = torch.clamp(1 - q, min=0) ** 6 * (1 + 6 * q + 35.0 / 3.0 * q * q)
    m = st.masses.to(torch.float64)
    u_add = energy * w / (w * m).sum()
    st.internalEnergies = (st.internalEnergies.to(torch.float64) + u_add).to(torch.float32)
    u = st.internalEnergies
    st.pressures = ((gamma - 1.0) * rho * u).to(torch.float32)
    st.soundspeeds = torch.sqrt(u.abs() * gamma * (gamma - 1.0)).to(torch.float32)
    st.entropies = (u * (gamma - 1.0) * rho ** (1.0 - gamma)).to(torch.float32)
    for f in ("positions", "velocities", "supports", "masses", "densities", "internalEnergies", "entropies",
              "pressures", "soundspeeds", "divergence", "alpha0s", "alphas", "kinds", "materials", "UIDs",
              "totalEnergies"):
        setattr(st, f, getattr(st, f).to(device).contiguous())
    config.dx = dx
    config.nx = nx
    return CompSPHSystem(state=st, adjacency=None, domain=dom, t=0.0), config, scheme
