"""Initial-condition generators on the GPU, with the reference's names and
argument meaning:

  buildPointCloud / sampleRegularParticles   sample/regular.py:6-56
  setupBasicCompressibleInitialState         sample/compressible.py:9-71
  sampleKHH (Kelvin-Helmholtz)               caseUtils/compressible/kelvinHelmholtz/sample.py:10-75
  buildSod1D (Sod shock tube)                caseUtils/compressible/sod/sod.py:21-80 (doubled periodic tube)
  buildSedov (Sedov blast)                   caseUtils/compressible/sedov/initial.py:24-218

Tensors are created on the configuration's device (random numbers, when
jitter is asked for, from a seeded generator on that device).  The adaptive
support and the gather density of buildSedov run on the GPU through the
CRKSPH library (Owen iterations, crk_density_gather).  The reference's
default kernel for buildSedov is B7, which this build does not have; the
Wendland kernels are available.
"""
from __future__ import annotations

import ctypes as C
from collections import namedtuple
from typing import Optional

import numpy as np
import torch

from . import abi
from .configurations import CRKSPHConfig, buildDomainDescription
from .enumTypes import AdaptiveSupportScheme, KernelFunctions
from .systems import CompSPHState, CompSPHSystem
from .utils import volumeToSupportHelper

PointCloud = namedtuple("PointCloud", ["positions", "supports"])
ParticleSet = namedtuple("ParticleSet", ["positions", "supports", "masses", "densities"])


def buildPointCloud(nx, domain, targetNeighbors=16, jitter=0.0, band=0, shortEdge=True, generator=None):
    """sample/regular.py:6-49: a lattice with spacing from the shorter (or
    longer) edge, cell-centred in periodic dimensions, ``band`` extra layers,
    one-sided jitter ``rand * dx * jitter``.  Returns (PointCloud, area, support)."""
    # the reference's arithmetic (0-dim tensors of the domain's dtype)
    dev = domain.min.device
    dt = domain.min.dtype
    per = [bool(p) for p in domain.periodic.tolist()]
    dxs = []
    for d in range(domain.dim):
        l = domain.max[d] - domain.min[d]
        dxs.append(l / (nx if per[d] else nx - 1))
    dx = torch.min(torch.stack(dxs)) if shortEdge else torch.max(torch.stack(dxs))
    spaces = []
    for d in range(domain.dim):
        l = domain.max[d] - domain.min[d]
        nd = int(torch.ceil(l / dx).to(torch.int32))
        off = dx / 2 if per[d] else torch.zeros_like(dx)
        off = off - dx * band
        spaces.append(torch.linspace(float(domain.min[d] + off), float(domain.max[d] - off), nd + band * 2,
                                     device=dev, dtype=dt))
    grid = torch.meshgrid(*spaces, indexing='xy')
    pos = torch.stack([g.flatten() for g in grid], dim=1)
    if jitter > 0:
        pos = pos + torch.rand(pos.shape, generator=generator, device=dev, dtype=dt) * dx * jitter
    area = float(dx ** domain.dim)
    support = volumeToSupportHelper(area, targetNeighbors, domain.dim)
    supports = torch.full_like(pos[:, 0], float(support))
    return PointCloud(positions=pos.contiguous(), supports=supports), area, support


def sampleRegularParticles(nx, domain, targetNeighbors, jitter=0.0, band=0, shortEdge=True, generator=None):
    """sample/regular.py:52-56: the point cloud with masses = area, densities = 1."""
    pc, area, support = buildPointCloud(nx, domain, targetNeighbors, jitter, band, shortEdge, generator)
    ones = torch.ones_like(pc.positions[:, 0])
    return ParticleSet(positions=pc.positions, supports=pc.supports, masses=ones * area, densities=ones.clone())


def idealGasEOS(A=None, u=None, P=None, rho=None, gamma=5.0 / 3.0):
    """eos/idealGas.py:6-38 (the host-side form used by the generators)."""
    P_, u_, A_ = P, u, A
    if u is not None:
        c_s = torch.sqrt(u.abs() * gamma * (gamma - 1))
    elif P is not None:
        c_s = torch.sqrt(gamma * P.abs() / rho)
    else:
        c_s = torch.sqrt(gamma * rho ** (gamma - 1) * A)
    if P is None and u is not None:
        P_ = (gamma - 1) * rho * u
    elif P is None and A is not None:
        P_ = A * rho ** gamma
    if u is None and A is not None:
        u_ = A * rho ** (gamma - 1) / (gamma - 1)
    elif u is None and P is not None:
        u_ = P / rho / (gamma - 1)
    if A is None and u is not None:
        A_ = u * (gamma - 1) * rho ** (1 - gamma)
    elif A is None and P is not None:
        A_ = P / rho ** gamma
    return A_, u_, P_, c_s


def _state(ps: ParticleSet, u, A, P, c, velocities=None) -> CompSPHState:
    n = ps.positions.shape[0]
    dev = ps.positions.device
    v = torch.zeros_like(ps.positions) if velocities is None else velocities
    return CompSPHState(
        positions=ps.positions, supports=ps.supports, masses=ps.masses, densities=ps.densities, velocities=v,
        kinds=torch.zeros(n, dtype=torch.int32, device=dev), materials=torch.zeros(n, dtype=torch.int32, device=dev),
        UIDs=torch.arange(n, dtype=torch.int32, device=dev), UIDcounter=n, internalEnergies=u,
        totalEnergies=(u + 0.5 * (v * v).sum(-1)) * ps.masses, entropies=A, pressures=P, soundspeeds=c,
        alphas=torch.ones_like(ps.densities), alpha0s=torch.ones_like(ps.densities),
        divergence=torch.zeros_like(ps.densities))


def setupBasicCompressibleInitialState(nx, config, schemeConfig):
    """sample/compressible.py:9-71: a regular lattice at rest, rho = 1, P = 0
    (the support scheme there is NoScheme, so supports and densities stay);
    sets config.dx = min(m)^(1/d)."""
    ps = sampleRegularParticles(nx, config.domain, config.targetNeighbors, jitter=0.0)
    A, u, P, c = idealGasEOS(P=torch.zeros_like(ps.densities), rho=ps.densities, gamma=schemeConfig.gamma)
    st = _state(ps, u, A, P, c)
    config.dx = st.masses.min() ** (1.0 / config.dim)
    return CompSPHSystem(state=st, adjacency=None, domain=config.domain, t=0.0)


def sampleKHH(rho1, rho2, v1, v2, delta, sigma, freq, w0, nx, config, schemeConfig):
    """kelvinHelmholtz/sample.py:10-75: bands y < 1/4 and y >= 3/4 at rho1, the
    middle at rho2 (sharp), v_x = v1 / v2 with exponential transitions of width
    delta at y = 1/4, 3/4, the v_y seed w0 sin(freq pi x) (Gaussians of width
    sigma around the interfaces) |1/2 - y|, P = 2.5, m = dx^d rho.  Sets
    config.dt = computeTimestep * 2/3."""
    from .sample import cfl_dt
    system = setupBasicCompressibleInitialState(nx, config, schemeConfig)
    st = system.state
    x = st.positions
    thpt = 1.0 / (2.0 * sigma * sigma)
    v_y = (w0 * torch.sin(freq * np.pi * x[:, 0]) * (torch.exp(-((x[:, 1] - 0.25) ** 2 * thpt)) +
                                                     torch.exp(-((x[:, 1] - 0.75) ** 2 * thpt)))) * \
        torch.abs(0.5 - x[:, 1])
    y = x[:, 1]
    r1 = (y >= 0) & (y < 0.25)
    r2 = (y >= 0.25) & (y < 0.5)
    r3 = (y >= 0.5) & (y < 0.75)
    r4 = (y >= 0.75) & (y <= 1)
    rho = torch.zeros_like(st.densities)
    rho[r1] = rho1
    rho[r2] = rho2
    rho[r3] = rho2
    rho[r4] = rho1
    vm = (v1 - v2) / 2
    v_x = torch.zeros_like(y)
    v_x[r1] = v1 - vm * torch.exp((y[r1] - 0.25) / delta)
    v_x[r2] = v2 + vm * torch.exp((0.25 - y[r2]) / delta)
    v_x[r3] = v2 + vm * torch.exp((y[r3] - 0.75) / delta)
    v_x[r4] = v1 - vm * torch.exp((0.75 - y[r4]) / delta)
    vel = torch.stack([v_x, v_y], dim=1).contiguous()
    A, u, P, c = idealGasEOS(P=torch.full_like(rho, 2.5), rho=rho, gamma=schemeConfig.gamma)
    st.internalEnergies = u.contiguous()
    st.pressures = P.contiguous()
    st.soundspeeds = c.contiguous()
    st.entropies = A.contiguous()
    st.velocities = vel
    st.densities = rho.contiguous()
    st.masses = (config.dx ** config.dim * rho).contiguous()
    st.totalEnergies = (u + (vel * vel).sum(-1) / 2) * st.masses
    # computeTimestep * 2/3 on the GPU (a CPU-built state leaves config.dt as it is)
    if st.positions.is_cuda:
        config.dt = cfl_dt(st, config, schemeConfig) * 2.0 / 3.0
    return system


def buildSod1D(config=None, device='cuda', **kw):
    """sod/sod.py:21-80 as the periodic doubled tube (the build's sod_1d)."""
    from .sample import sod_1d
    return sod_1d(device=device, **kw)


def optimal_support(system, config, scheme, iterations=16, threshold=1e-3):
    """evaluateOptimalSupport with Owen (optimalSupportOwen.py:35-127) on the
    GPU: Owen iterations until max |h'/h - 1| < threshold, then the gather
    density.  Returns (rho, h) and updates system.state.supports."""
    from .configurations import toCrkConfig
    from .runtime import ensure_adjacency, runtime
    st = system.state
    sc = CRKSPHConfig(gamma=scheme.gamma, adaptiveSupportScheme=AdaptiveSupportScheme.Owen,
                      adaptiveSupportIterations=iterations, adaptiveSupportThreshold=threshold)
    cfg = toCrkConfig(config, sc)
    rt = runtime(st.positions.device)
    rt.bind_runs(None)
    s = rt.stream()
    rt.ensure_lut(cfg.kernel, int(config.dim), cfg.targetNeighbors)
    n = st.numParticles
    hmm = rt.scratch[0:2]
    ratio = rt.scratch[2:3]
    cst = st.cabi()
    abi.call("crk_minmax", rt.ctx, n, abi.ptr(st.supports), abi.ptr(hmm), s)
    adj = system.adjacency
    for it in range(iterations):
        adj, _ = ensure_adjacency(cfg, adj, st.positions, st.supports)
        abi.call("crk_stage_begin", rt.ctx, C.byref(cfg), C.byref(adj.cabi()), C.byref(cst), s)
        abi.call("crk_owen_iteration", rt.ctx, C.byref(cfg), C.byref(cst), abi.ptr(hmm), abi.ptr(ratio), s)
        abi.call("crk_minmax", rt.ctx, n, abi.ptr(st.supports), abi.ptr(hmm), s)
        if float(ratio.item()) < threshold:
            break
    adj, _ = ensure_adjacency(cfg, adj, st.positions, st.supports)
    abi.call("crk_stage_begin", rt.ctx, C.byref(cfg), C.byref(adj.cabi()), C.byref(cst), s)
    rho = torch.empty_like(st.supports)
    abi.call("crk_density_gather", rt.ctx, C.byref(cfg), abi.ptr(rho), s)
    system.adjacency = adj
    return rho, st.supports


def gather_density(system, config, scheme) -> torch.Tensor:
    """warpOperation Density (gather, W(x_ij, h_i)) of the current state on the GPU."""
    from .configurations import toCrkConfig
    from .runtime import ensure_adjacency, runtime
    st = system.state
    cfg = toCrkConfig(config, scheme)
    rt = runtime(st.positions.device)
    rt.bind_runs(None)
    adj, _ = ensure_adjacency(cfg, system.adjacency, st.positions, st.supports)
    abi.call("crk_stage_begin", rt.ctx, C.byref(cfg), C.byref(adj.cabi()), C.byref(st.cabi()), rt.stream())
    rho = torch.empty_like(st.supports)
    abi.call("crk_density_gather", rt.ctx, C.byref(cfg), abi.ptr(rho), rt.stream())
    system.adjacency = adj
    return rho


def buildSedov(config=None, nx: int = 200, dim: int = 2, domainExtent: float = 2, periodicDomain: bool = True,
               rho0: float = 1, E0: float = 1, initialization: str = 'singular', gamma: float = 5 / 3,
               kernel=KernelFunctions.Wendland4, targetNeighbors=50, dtype=torch.float32, device='cuda'):
    """sedov/initial.py:24-218 with the 'singular' (E0 in the particle closest to
    the origin) and 'quadrant' (E0 shared by the 2^d closest) initialisations
    ('hat' is not implemented there either): a regular lattice of mass
    rho0 * area, the gather density, Owen supports (16 iterations, threshold
    1e-3) and their gather density, u from E0, EOS.  Returns (system, config, scheme)."""
    import warnings
    from .configurations import buildConfig
    from .sample import cfl_dt
    if initialization not in ('singular', 'quadrant'):
        raise ValueError(f"Unknown initialization {initialization}" if initialization != 'hat' else
                         "Hat initialization not implemented yet")
    if initialization == 'singular' and nx % 2 == 0:
        warnings.warn('nx should be odd for singular initialization, setting to nx + 1')
        nx += 1
    elif initialization == 'quadrant' and nx % 2 == 1:
        warnings.warn('nx should be even for quadrant initialization, setting to nx - 1')
        nx -= 1
    domain = buildDomainDescription(domainExtent, dim, periodic=periodicDomain, device=device, dtype=dtype)
    if config is None:
        config, _ = buildConfig(device=torch.device(device), domain=domain, dim=dim, kernel=kernel,
                                targetNeighbors=targetNeighbors, cflFactor=0.3)
    else:
        config.domain = domain
        config.dim = dim
    scheme = CRKSPHConfig(gamma=gamma, adaptiveSupportScheme=AdaptiveSupportScheme.Owen)
    ps = sampleRegularParticles(nx, domain, targetNeighbors)
    ps = ps._replace(masses=(ps.masses * rho0).contiguous())
    zero = torch.zeros_like(ps.densities)
    system = CompSPHSystem(state=_state(ps, zero.clone(), zero.clone(), zero.clone(), zero.clone()), adjacency=None,
                           domain=domain, t=0.0)
    system.state.densities = gather_density(system, config, scheme)
    rho, h = optimal_support(system, config, scheme, iterations=16, threshold=1e-3)
    st = system.state
    st.densities = rho
    st.supports = h
    dist = torch.linalg.norm(st.positions, dim=-1)
    _, idx = torch.sort(dist)
    u = torch.zeros_like(st.masses)
    if initialization == 'singular':
        u[idx[0]] = E0 / st.masses[idx[0]]
    else:
        k = 4 if dim == 2 else 2
        u[idx[:k]] = E0 / st.masses[idx[:k]] / k
    A, u, P, c = idealGasEOS(u=u, rho=rho, gamma=gamma)
    st.internalEnergies, st.entropies, st.pressures, st.soundspeeds = u, A, P, c
    st.totalEnergies = u * st.masses
    config.dx = float(st.masses.min())
    config.dt = cfl_dt(st, config, scheme)
    return system, config, scheme
