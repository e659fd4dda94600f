"""Initial-condition generators with the reference's names, signatures and
argument meaning:

  buildPointCloud / sampleRegularParticles   sample/regular.py:6-56
  setupBasicCompressibleInitialState         sample/compressible.py:9-71
  sampleKHH (Kelvin-Helmholtz)               caseUtils/compressible/kelvinHelmholtz/sample.py:10-75
  buildSod1D (Sod shock tube)                caseUtils/compressible/sod/sod.py:12-225
  buildSedov (Sedov blast)                   caseUtils/compressible/sedov/initial.py:24-218
  evaluateOptimalSupport                     modules/adaptiveSupport/optimalSupport.py:12-43
  computeTimestep                            modules/timestep/compressible.py:12-66

The lattices and the field assembly are tensor arithmetic on the
configuration's device.  The adaptive supports (Owen or Monaghan), the gather
densities, the Verlet list and the CFL step run on the GPU through the CRKSPH
library; they raise on tensors that are not on a GPU.
"""
from __future__ import annotations

import ctypes as C
from collections import namedtuple
from typing import NamedTuple, Optional

import numpy as np
import torch

from . import abi
from .configurations import CompressibleSPHConfig, CRKSPHConfig, buildDomainDescription
from .enumTypes import AdaptiveSupportScheme, KernelFunctions, SupportScheme
from .systems import CompSPHState, CompSPHSystem
from .utils import volumeToSupportHelper

PointCloud = namedtuple("PointCloud", ["positions", "supports"])
ParticleSet = namedtuple("ParticleSet", ["positions", "supports", "masses", "densities"])


class sodInitialState(NamedTuple):
    """One side of the shock tube (sod/sod.py:12-15)."""
    p: float
    rho: float
    v: float


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


def _crk_params(compParams) -> CRKSPHConfig:
    """The library's parameter block for a CompressibleSPHConfig (the
    adaptive-support fields and gamma are the ones used here)."""
    if isinstance(compParams, CRKSPHConfig):
        return compParams
    return CRKSPHConfig(gamma=compParams.gamma, adaptiveSupportScheme=compParams.adaptiveSupportScheme,
                        adaptiveSupportIterations=compParams.adaptiveSupportIterations,
                        adaptiveSupportThreshold=compParams.adaptiveSupportThreshold,
                        adaptiveSupportCorrections=compParams.adaptiveSupportCorrections)


def _work_state(st) -> CompSPHState:
    """The fields the neighbour passes read, float32 and contiguous; supports
    copied (the support iterations update them in place), absent velocities
    and energies as zeros."""
    f = lambda t: t.to(torch.float32).contiguous()
    x = f(st.positions)
    n = x.shape[0]
    z = torch.zeros(n, dtype=torch.float32, device=x.device)
    return CompSPHState(positions=x, velocities=f(st.velocities) if st.velocities is not None else torch.zeros_like(x),
                        supports=f(st.supports).clone(), masses=f(st.masses),
                        densities=f(st.densities) if st.densities is not None else z.clone(),
                        internalEnergies=f(st.internalEnergies) if st.internalEnergies is not None else z)


def _density_gather(rt, cfg, ws, adj):
    from .runtime import ensure_adjacency
    adj, _ = ensure_adjacency(cfg, adj, ws.positions, ws.supports)
    abi.call("crk_stage_begin", rt.ctx, C.byref(cfg), C.byref(adj.cabi()), C.byref(ws.cabi()), rt.stream())
    rho = torch.empty_like(ws.supports)
    abi.call("crk_density_gather", rt.ctx, C.byref(cfg), abi.ptr(rho), rt.stream())
    return rho, adj


def evaluateOptimalSupport(particleState, config, compParams, supportScheme: SupportScheme = SupportScheme.Gather,
                           adjacency=None):
    """optimalSupport.py:12-43 on the GPU: Owen (optimalSupportOwen.py:35-127)
    or Monaghan (optimalSupportMonaghan.py:20-108) iterations until
    max |h'/h - 1| < adaptiveSupportThreshold or adaptiveSupportIterations,
    then the gather density with the final supports; NoScheme returns the
    state's densities and supports.  Returns (densities, supports, adjacency,
    rhos, supports per iteration) and leaves the state's tensors as they
    are.  Every caller in the reference asks for the gather form, the only
    one provided."""
    from .configurations import toCrkConfig
    from .runtime import ensure_adjacency, runtime
    scheme_h = compParams.adaptiveSupportScheme
    if scheme_h == AdaptiveSupportScheme.NoScheme:
        return particleState.densities, particleState.supports, adjacency, None, None
    if scheme_h not in (AdaptiveSupportScheme.Owen, AdaptiveSupportScheme.Monaghan):
        raise ValueError(f"Unsupported adaptive support scheme: {scheme_h}")
    if supportScheme != SupportScheme.Gather:
        raise ValueError(f"evaluateOptimalSupport: only the gather form is provided (got {supportScheme})")
    cfg = toCrkConfig(config, _crk_params(compParams))
    ws = _work_state(particleState)
    rt = runtime(ws.positions.device)
    rt.bind_runs(None)
    s = rt.stream()
    monaghan = scheme_h == AdaptiveSupportScheme.Monaghan
    if not monaghan:
        rt.ensure_lut(cfg.kernel, int(config.dim), cfg.targetNeighbors)
    update = "crk_monaghan_iteration" if monaghan else "crk_owen_iteration"
    n = ws.numParticles
    hmm = torch.empty(2, dtype=torch.float32, device=ws.positions.device)
    ratio = torch.empty(1, dtype=torch.float32, device=ws.positions.device)
    cst = ws.cabi()
    abi.call("crk_minmax", rt.ctx, n, abi.ptr(ws.supports), abi.ptr(hmm), s)
    hs = [particleState.supports]
    adj = adjacency
    for _ in range(int(compParams.adaptiveSupportIterations)):
        adj, _ = ensure_adjacency(cfg, adj, ws.positions, ws.supports)
        abi.call("crk_stage_begin", rt.ctx, C.byref(cfg), C.byref(adj.cabi()), C.byref(cst), s)
        abi.call(update, rt.ctx, C.byref(cfg), C.byref(cst), abi.ptr(hmm), abi.ptr(ratio), s)
        hs.append(ws.supports.clone())
        abi.call("crk_minmax", rt.ctx, n, abi.ptr(ws.supports), abi.ptr(hmm), s)
        if float(ratio.item()) < compParams.adaptiveSupportThreshold:
            break
    rho, adj = _density_gather(rt, cfg, ws, adj)
    h = ws.supports.to(particleState.supports.dtype)
    return rho.to(particleState.supports.dtype), h, adj, [particleState.densities, rho], hs


def buildVerletList(state, config):
    """buildVerletList(state, domain, verletScale, SuperSymmetric) on the GPU
    with the configuration's Verlet scale (the generators' 2^(1/d) where the
    configuration has buildConfig's default)."""
    from .configurations import toCrkConfig
    from .runtime import build_adjacency
    ws = _work_state(state)
    return build_adjacency(toCrkConfig(config, CRKSPHConfig()), ws.positions, ws.supports)


def computeTimestep(system, config, compParams, dt: Optional[float] = None) -> float:
    """timestep/compressible.py:12-66: with adaptiveDt, cflFactor h_min /
    (c_s xi) over the particles (c_s from the internal energies), clamped to
    [minDt, maxDt] and, when dt is given, to dtGrowthFactor * config.dt;
    otherwise dt (which must be given)."""
    from .sample import cfl_dt
    if config.adaptiveDt:
        prev = config.dt if dt is not None else None
        return cfl_dt(system.state, config, _crk_params(compParams), prev)
    if dt is not None:
        return dt
    raise ValueError("dt must be provided if adaptiveDt is False")


def _finish_state(st, u, rho, gamma):
    A, u, P, c = idealGasEOS(u=u, rho=rho, gamma=gamma)
    st.internalEnergies = u.contiguous()
    st.entropies = A.contiguous()
    st.pressures = P.contiguous()
    st.soundspeeds = c.contiguous()
    st.totalEnergies = ((u + torch.linalg.norm(st.velocities, dim=-1) ** 2 / 2) * st.masses).contiguous()


def setupBasicCompressibleInitialState(nx, config, schemeConfig, SimulationState=CompSPHState,
                                       SimulationSystem=CompSPHSystem):
    """sample/compressible.py:9-71: a regular lattice at rest, rho = 1, P = 0
    (the support scheme there is NoScheme, so supports and densities stay);
    config.dx = min(m)^(1/d), config.dt = computeTimestep(..., dt=config.dt)."""
    ps = sampleRegularParticles(nx, config.domain, config.targetNeighbors, jitter=0.0)
    v = torch.zeros_like(ps.positions)
    A, u, P, c = idealGasEOS(P=torch.zeros_like(ps.densities), rho=ps.densities, gamma=schemeConfig.gamma)
    n = ps.positions.shape[0]
    dev = ps.positions.device
    st = SimulationState(
        positions=ps.positions, supports=ps.supports, masses=ps.masses, densities=ps.densities, velocities=v,
        kinds=torch.zeros(n, dtype=torch.int32, device=dev), materials=torch.zeros(n, dtype=torch.int32, device=dev),
        UIDs=torch.arange(n, dtype=torch.int32, device=dev), UIDcounter=n, internalEnergies=u,
        totalEnergies=(u + torch.linalg.norm(v, dim=-1) ** 2 / 2) * ps.masses, entropies=A, pressures=P,
        soundspeeds=c, alphas=torch.ones_like(ps.densities), alpha0s=torch.ones_like(ps.densities),
        divergence=torch.zeros_like(ps.densities))
    system = SimulationSystem(state=st, adjacency=None, domain=config.domain)
    config.dx = float(st.masses.min()) ** (1.0 / config.dim)
    config.dt = computeTimestep(system, config, schemeConfig, dt=config.dt)
    return system


def sampleKHH(rho1, rho2, v1, v2, delta, sigma, freq, w0, nx, config, schemeConfig, SimulationState=CompSPHState,
              SimulationSystem=CompSPHSystem):
    """kelvinHelmholtz/sample.py:10-75: bands y < 1/4 and y >= 3/4 at rho1, the
    middle at rho2 (sharp), v_x = v1 / v2 with exponential transitions of width
    delta at y = 1/4, 3/4, the v_y seed w0 sin(freq pi x) (Gaussians of width
    sigma around the interfaces) |1/2 - y|, P = 2.5, m = dx^d rho.  Sets
    config.dt = computeTimestep(..., dt=config.dt) * 2/3."""
    system = setupBasicCompressibleInitialState(nx, config, schemeConfig, SimulationState, SimulationSystem)
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
    st.velocities = vel
    st.densities = rho.contiguous()
    # (the entropies keep the lattice's P = 0 values, as in the reference)
    st.totalEnergies = ((u + torch.linalg.norm(vel, dim=-1) ** 2 / 2) * st.masses).contiguous()
    st.masses = (config.dx ** config.dim * rho).contiguous()
    config.dt = computeTimestep(system, config, schemeConfig, dt=config.dt) * 2 / 3
    return system


def buildSod1D(SimulationSystem, SimulationState, samplingRatio: int, leftState: sodInitialState,
               rightState: sodInitialState, gamma: float, config, smoothIC: bool = False,
               adaptiveSupportScheme: AdaptiveSupportScheme = AdaptiveSupportScheme.Monaghan):
    """sod/sod.py:21-225: the periodic doubled tube on config.domain
    ([-1, 1]): config.nx particles of the left state on [-1/2, 1/2] and
    config.nx // samplingRatio of the right state on [-1, -1/2) u [1/2, 1),
    each a cell-centred lattice of the unit interval (the right one shifted
    out by 1/2); left masses m_l rho_l, right m_l rho_r nx / (nx //
    samplingRatio), m_l the left lattice's spacing; supports from the two
    lattices; then the adaptive supports (16 iterations, threshold 1e-3,
    ``adaptiveSupportScheme``), u = P / ((gamma - 1) rho) with rho the
    sampled densities or, with smoothIC, the adapted gather densities and u
    blended across both interfaces by a logistic ramp of width dx_l / 2; the
    EOS, the Verlet list, config.dx = min m and config.dt = computeTimestep."""
    nx = config.nx
    nr = nx // samplingRatio
    actualRatio = nx / nr
    unit = lambda: buildDomainDescription(1, config.dim, periodic=True, device=config.device, dtype=config.dtype)
    pl = sampleRegularParticles(nx, unit(), config.targetNeighbors, jitter=0.0)
    pr = sampleRegularParticles(nr, unit(), config.targetNeighbors, jitter=0.0)
    pos_r = pr.positions.clone()
    pos_r[pos_r[:, 0] < 0, 0] -= 0.5
    pos_r[pos_r[:, 0] > 0, 0] += 0.5
    positions = torch.cat([pl.positions, pos_r], dim=0).contiguous()
    dev = positions.device
    nl = pl.positions.shape[0]
    tags = torch.cat([torch.zeros(nl, dtype=torch.int32, device=dev),
                      torch.ones(pos_r.shape[0], dtype=torch.int32, device=dev)])
    left = tags == 0
    m_l = pl.masses.min()
    masses = torch.where(left, m_l * leftState.rho, m_l * rightState.rho * actualRatio).to(positions.dtype)
    supports = torch.where(left, pl.supports.min(), pr.supports.min()).to(positions.dtype)
    densities = torch.where(left, torch.tensor(float(leftState.rho), device=dev),
                            torch.tensor(float(rightState.rho), device=dev)).to(positions.dtype)
    velocities = torch.where(left[:, None], float(leftState.v), float(rightState.v)).to(positions.dtype)
    velocities = velocities.expand(-1, positions.shape[1]).contiguous()
    UIDs = torch.cat([torch.arange(nl, dtype=torch.int32, device=dev),
                      torch.arange(pos_r.shape[0], dtype=torch.int32, device=dev)])
    st = SimulationState(
        positions=positions, velocities=velocities, supports=supports.contiguous(), masses=masses.contiguous(),
        densities=densities.contiguous(), kinds=torch.zeros_like(tags), materials=tags, UIDs=UIDs,
        UIDcounter=int(UIDs.max()) + 1, internalEnergies=None, totalEnergies=None, entropies=None, pressures=None,
        soundspeeds=None, divergence=torch.zeros_like(densities), alpha0s=torch.ones_like(densities),
        alphas=torch.ones_like(densities))
    compParams = CompressibleSPHConfig(gamma=gamma, adaptiveSupportCorrections=True, adaptiveSupportIterations=16,
                                       adaptiveSupportThreshold=1e-3, adaptiveSupportScheme=adaptiveSupportScheme)
    rho_optimal, h_optimal, adjacency, _, _ = evaluateOptimalSupport(st, config, compParams, SupportScheme.Gather)
    st.supports = h_optimal.contiguous()
    eosRho = rho_optimal if smoothIC else st.densities
    P_initial = torch.where(st.materials == 0, torch.tensor(float(leftState.p), device=dev),
                            torch.tensor(float(rightState.p), device=dev)).to(positions.dtype)
    u = 1 / (gamma - 1) * (P_initial / eosRho)
    if smoothIC:
        # the logistic blend of u across x = +-1/2 (sod.py:172-190)
        dx = 0.5 * (pl.positions[1, 0] - pl.positions[0, 0])
        ramped = lambda a, b, x: (a - b) / (torch.exp(x / dx) + 1) + b
        left_u = 1 / (gamma - 1) * (leftState.p / leftState.rho)
        right_u = 1 / (gamma - 1) * (rightState.p / rightState.rho)
        x = st.positions[:, 0]
        u = torch.where(x > 0., ramped(left_u, right_u, x - 0.5), u)
        u = torch.where(x < 0., ramped(right_u, left_u, x + 0.5), u).to(st.densities.dtype)
    st.densities = eosRho.contiguous()
    _finish_state(st, u, eosRho, gamma)
    system = SimulationSystem(state=st, adjacency=None, domain=config.domain)
    system.adjacency = buildVerletList(st, config)
    compParams.gamma = gamma
    config.dx = float(st.masses.min())
    config.dt = computeTimestep(system, config, compParams, dt=None)
    return system


def buildSedov(SimulationSystem, SimulationState, config, nx: int = 200, dim: int = 2, domainExtent: float = 2,
               periodicDomain: bool = True, rho0: float = 1, E0: float = 1, initialization: str = 'hat',
               gamma: float = 5 / 3, kernel: KernelFunctions = None, targetNeighbors=50, dtype=torch.float32,
               device=None, shellSampling: bool = False):
    """sedov/initial.py:24-218: a regular lattice of the domain of extent
    ``domainExtent`` with masses rho0 dx^d, Owen supports (16 iterations,
    threshold 1e-3, with config's kernel and domain) and their gather
    density, E0 deposited as internal energy in the particle closest to the
    origin ('singular', nx made odd) or shared by the 2^d closest
    ('quadrant', nx made even), the EOS, the Verlet list, config.dx = min m
    and config.dt = computeTimestep(..., dt=config.dt).  'hat' raises
    NotImplementedError, as there (here before any work).  ``kernel`` sets only the lattice's
    first density estimate, which the Owen pass replaces before any use, so
    it does not enter the result and is not evaluated.  Shell sampling
    (sample/shell.py) is not provided."""
    import warnings
    if device is None:
        device = torch.device('cuda') if torch.cuda.is_available() else torch.device('cpu')
    domain = buildDomainDescription(domainExtent, dim, periodic=periodicDomain, device=device, dtype=dtype)
    if initialization == 'singular':
        if nx % 2 == 0:
            warnings.warn('nx should be odd for singular initialization, setting to nx + 1')
            nx += 1
    elif initialization == 'quadrant':
        if nx % 2 == 1:
            warnings.warn('nx should be even for quadrant initialization, setting to nx - 1')
            nx -= 1
    if shellSampling:
        raise NotImplementedError('shell sampling (sample/shell.py) is not provided')
    if initialization == 'hat':
        raise NotImplementedError('Hat initialization not implemented yet')
    if initialization not in ('singular', 'quadrant'):
        raise ValueError(f'Unknown initialization {initialization}')
    ps = sampleRegularParticles(nx, domain, targetNeighbors)
    ps = ps._replace(masses=(ps.masses * rho0).contiguous())
    n = ps.positions.shape[0]
    dev = ps.positions.device
    st = SimulationState(
        positions=ps.positions, supports=ps.supports, masses=ps.masses, densities=ps.densities,
        velocities=torch.zeros_like(ps.positions), kinds=torch.zeros(n, dtype=torch.int32, device=dev),
        materials=torch.zeros(n, dtype=torch.int32, device=dev), UIDs=torch.arange(n, dtype=torch.int32, device=dev),
        UIDcounter=n, internalEnergies=None, totalEnergies=None, entropies=None, pressures=None, soundspeeds=None,
        divergence=torch.zeros_like(ps.densities), alpha0s=torch.ones_like(ps.densities),
        alphas=torch.ones_like(ps.densities))
    compParams = CompressibleSPHConfig(adaptiveSupportIterations=16, adaptiveSupportThreshold=1e-3,
                                       adaptiveSupportScheme=AdaptiveSupportScheme.Owen, gamma=gamma)
    rho_optimal, h_optimal, _, _, _ = evaluateOptimalSupport(st, config, compParams, SupportScheme.Gather)
    st.densities = rho_optimal.contiguous()
    st.supports = h_optimal.contiguous()
    dist = torch.linalg.norm(st.positions, dim=-1)
    _, idx = torch.sort(dist)
    u = torch.zeros_like(st.masses)
    if initialization == 'singular':
        u[idx[0]] = E0 / st.masses[idx[0]]
    else:
        k = 4 if dim == 2 else 2
        u[idx[:k]] = E0 / st.masses[idx[:k]] / k
    _finish_state(st, u, rho_optimal, gamma)
    system = SimulationSystem(state=st, adjacency=None, domain=config.domain)
    system.adjacency = buildVerletList(st, config)
    config.dx = float(st.masses.min())
    config.dt = computeTimestep(system, config, compParams, dt=config.dt)
    return system
