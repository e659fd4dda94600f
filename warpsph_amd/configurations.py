"""Configuration objects of the CRKSPH step.

Mirrors the reference's dataclasses so that callers keep working:
  SimulationConfig / buildConfig     src/warpSPH/configurations/simulationConfig.py:16-121
  DomainDescription / buildDomain... src/warpSPH/utils/domain.py:7-22
  DiffusionParameters                .../moduleConfigurations/diffusionParameters.py:23-36
  CRKViscosity                       src/warpSPH/configurations/crkSPH.py:17-36
  ViscositySwitchConfig              .../moduleConfigurations/viscositySwitchParameters.py:12-24
  CompressibleSPHConfig              src/warpSPH/configurations/compressibleConfig.py:23-40
  BoundaryCondition & enums          .../moduleConfigurations/boundaryConditions.py:8-41
  CRKSPHConfig, crkSPHConfigToDict, dictToCRKSPHConfig   crkSPH.py:52-111
The reference's ``wp.struct`` parameter blocks become plain dataclasses here;
``toCrkConfig`` packs them into the C-ABI ``CrkConfig``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from enum import Enum
from typing import Any, Dict, List, Optional

import torch

from . import abi
from .enumTypes import (AdaptiveSupportScheme, EnergyScheme, GradientScheme, IntegrationSchemeType,
                        KernelFunctions, LaplacianScheme, SamplingScheme, SupportScheme,
                        ViscositySwitch)
from .utils import n_h_to_nH


# ------------------------------------------------------------------ domain
@dataclass
class DomainDescription:
    min: torch.Tensor
    max: torch.Tensor
    periodic: torch.Tensor
    dim: int


def buildDomainDescription(l, dim, periodic=False, device='cpu', dtype=torch.float32):
    """utils/domain.py:19-22: a cube [-l/2, l/2]^dim."""
    mn = [-l / 2] * dim
    mx = [l / 2] * dim
    per = torch.tensor([periodic] * dim, dtype=torch.bool, device=device) if isinstance(periodic, bool) else periodic
    return DomainDescription(torch.tensor(mn, device=device, dtype=dtype),
                             torch.tensor(mx, device=device, dtype=dtype), per, dim)


# ------------------------------------------------------------------ simulation
@dataclass
class SimulationConfig:
    device: torch.device = field(default_factory=lambda: torch.device('cuda:0') if torch.cuda.is_available() else torch.device('cpu'))
    dtype: torch.dtype = torch.float32
    domain: DomainDescription = field(default_factory=lambda: buildDomainDescription(l=2, dim=2, periodic=True))
    dim: int = 2
    verletScale: float = 2.0 ** 0.5
    kernel: KernelFunctions = KernelFunctions.Wendland2
    integrationScheme: IntegrationSchemeType = IntegrationSchemeType.rungeKutta2
    cflFactor: float = 0.3
    dt: Optional[float] = None
    minDt: Optional[float] = 1e-6
    maxDt: Optional[float] = 1e-2
    dtGrowthFactor: Optional[float] = 1.1
    adaptiveDt: bool = True
    dx: Optional[float] = None
    nx: Optional[int] = None
    targetNeighbors: float = field(default_factory=lambda: n_h_to_nH(4, 2))
    supportMode: SupportScheme = SupportScheme.SuperSymmetric
    gradientMode: GradientScheme = GradientScheme.Difference
    laplacianMode: LaplacianScheme = LaplacianScheme.Brookshaw
    samplingScheme: SamplingScheme = SamplingScheme.regular


def buildConfig(device=None, dtype=None, domain=None, dim=None, verletScale=None, kernel=None,
                integrationScheme=None, cflFactor=None, dt=None, minDt=None, maxDt=None,
                dtGrowthFactor=None, adaptiveDt=None, targetNeighbors=None, supportMode=None,
                gradientMode=None, laplacianMode=None, samplingScheme=None):
    """simulationConfig.py:47-121.  Returns (config, integrator)."""
    from .integrator import getIntegrator
    if device is None:
        device = torch.device('cuda:0') if torch.cuda.is_available() else torch.device('cpu')
    dtype = torch.float32 if dtype is None else dtype
    dim = 2 if dim is None else dim
    domain = buildDomainDescription(l=2, dim=dim, periodic=True) if domain is None else domain
    verletScale = 2.0 ** (1 / dim) if verletScale is None else verletScale
    kernel = KernelFunctions.Wendland2 if kernel is None else kernel
    integrationScheme = IntegrationSchemeType.rungeKutta2 if integrationScheme is None else integrationScheme
    cfg = SimulationConfig(
        device=device, dtype=dtype, domain=domain, dim=dim, verletScale=verletScale, kernel=kernel,
        integrationScheme=integrationScheme, cflFactor=0.3 if cflFactor is None else cflFactor, dt=dt,
        minDt=1e-6 if minDt is None else minDt, maxDt=1e-2 if maxDt is None else maxDt,
        dtGrowthFactor=1.1 if dtGrowthFactor is None else dtGrowthFactor,
        adaptiveDt=True if adaptiveDt is None else adaptiveDt,
        targetNeighbors=n_h_to_nH(4, dim) if targetNeighbors is None else targetNeighbors,
        supportMode=SupportScheme.SuperSymmetric if supportMode is None else supportMode,
        gradientMode=GradientScheme.Difference if gradientMode is None else gradientMode,
        laplacianMode=LaplacianScheme.Brookshaw if laplacianMode is None else laplacianMode,
        samplingScheme=SamplingScheme.regular if samplingScheme is None else samplingScheme)
    return cfg, getIntegrator(integrationScheme)


# ------------------------------------------------------------------ parameter blocks
def configurationToDict(config: SimulationConfig) -> Dict[str, Any]:
    """simulationConfig.py:125-155: the configuration as plain JSON values."""
    dt = config.dt.detach().cpu().item() if isinstance(config.dt, torch.Tensor) else config.dt
    return {
        'device': str(config.device), 'dtype': str(config.dtype),
        'domain': {'min': _host_values(config.domain.min), 'max': _host_values(config.domain.max),
                   'periodic': [bool(b) for b in _host_values(config.domain.periodic)], 'dim': config.domain.dim},
        'dim': config.dim, 'verletScale': config.verletScale, 'kernel': config.kernel.name,
        'integrationScheme': config.integrationScheme.name, 'cflFactor': config.cflFactor, 'dt': dt,
        'minDt': config.minDt, 'maxDt': config.maxDt, 'dtGrowthFactor': config.dtGrowthFactor,
        'adaptiveDt': config.adaptiveDt, 'targetNeighbors': config.targetNeighbors,
        'supportMode': config.supportMode.name, 'gradientMode': config.gradientMode.name,
        'laplacianMode': config.laplacianMode.name, 'samplingScheme': config.samplingScheme.name,
    }


def dictToConfig(configDict: Dict[str, Any]) -> SimulationConfig:
    """simulationConfig.py:157-200: the inverse of configurationToDict."""
    device = torch.device(configDict['device'])
    dtype = getattr(torch, configDict['dtype'].split('.')[-1])
    dd = configDict['domain']
    domain = DomainDescription(torch.tensor(dd['min'], dtype=dtype, device=device),
                               torch.tensor(dd['max'], dtype=dtype, device=device),
                               torch.tensor(dd['periodic'], dtype=torch.bool, device=device), int(dd['dim']))
    dt = configDict.get('dt')
    return SimulationConfig(
        device=device, dtype=dtype, domain=domain, dim=int(configDict['dim']),
        verletScale=float(configDict['verletScale']), kernel=KernelFunctions[configDict['kernel']],
        integrationScheme=IntegrationSchemeType[configDict['integrationScheme']],
        cflFactor=float(configDict['cflFactor']), dt=None if dt is None else float(dt),
        minDt=configDict['minDt'], maxDt=configDict['maxDt'], dtGrowthFactor=configDict['dtGrowthFactor'],
        adaptiveDt=bool(configDict['adaptiveDt']), targetNeighbors=float(configDict['targetNeighbors']),
        supportMode=SupportScheme[configDict['supportMode']], gradientMode=GradientScheme[configDict['gradientMode']],
        laplacianMode=LaplacianScheme[configDict['laplacianMode']],
        samplingScheme=SamplingScheme[configDict['samplingScheme']])


class ViscosityTerms(Enum):  # diffusionParameters.py:9-21
    Default = 0
    MonaghanGingold1983 = 1
    Cleary1998 = 2
    Monaghan1992 = 3
    Monaghan1997a = 4
    Monaghan1997b = 5
    Dukowicz = 6
    Price2012_98 = 7
    Price2012 = 8
    Price2008 = 9
    Wadsley2008 = 10
    DeltaSPH = 11


@dataclass
class DiffusionParameters:
    c_s: float = 1.0
    C_l: float = 1.0
    C_q: float = 2.0
    Cu_l: float = 1.0
    Cu_q: float = 2.0
    K: float = 1.0
    thermalConductivity: float = 0.5
    viscosityTerm: ViscosityTerms = ViscosityTerms.Price2012_98
    thermalConducitiyTerm: ViscosityTerms = ViscosityTerms.Price2012_98
    scaleBeta: bool = False
    monaghanSwitch: bool = True
    correctXi: bool = True


def buildDefaultDiffusionParamsCRKSPH():
    """configurations/crkSPH.py:52-68."""
    return DiffusionParameters(c_s=1, C_l=1, C_q=1, Cu_l=1, Cu_q=1, K=1.0, thermalConductivity=0.5,
                               viscosityTerm=ViscosityTerms.Monaghan1992, thermalConducitiyTerm=ViscosityTerms.Price2012_98,
                               scaleBeta=False, monaghanSwitch=True, correctXi=True)


def buildDefaultDiffusionParamsCompressibleSPH():
    return DiffusionParameters(c_s=1, C_l=1, C_q=0, Cu_l=1, Cu_q=0, K=1.0, thermalConductivity=0.5,
                               viscosityTerm=ViscosityTerms.Price2012_98, thermalConducitiyTerm=ViscosityTerms.Price2008,
                               scaleBeta=False, monaghanSwitch=True, correctXi=True)


@dataclass
class CRKViscosity:
    eta_fold: float = 0.2
    eta_crit: float = 0.3333333
    enableCRKLimiter: bool = True
    enableVanLeerLimiter: bool = True
    forceVanLeerOff: bool = False
    forceVanLeerOn: bool = False


def buildDefaultCRKViscosityParams():
    return CRKViscosity()


@dataclass
class ViscositySwitchConfig:
    scheme: ViscositySwitch = ViscositySwitch.NoneSwitch
    correctVelocityGradient: bool = False
    divergenceScheme: Optional[str] = 'naive'
    alpha_min: float = 0.02
    alpha_max: float = 2.0
    beta_c: float = 0.7
    beta_d: float = 0.05
    beta_xi: float = 2.0
    limitXi: bool = True  # the reference declares limitXi twice; the later default (True) is effective
    # extension (no reference counterpart): the trained network of
    # ViscositySwitch.Learned (warpsph_amd.learned.LearnedSwitch)
    model: Optional[Any] = None


# ------------------------------------------------------------------ boundary conditions
# .../moduleConfigurations/boundaryConditions.py:8-41 (the fields the step's
# hooks read: sdf, dirichletFunctions, updateFunctions, forcingFunctions;
# warpsph_amd.boundary applies them)
class BCType(Enum):
    constant = 1
    noSlip = 2
    freeSlip = 3
    extended = 4
    zeros = 5


class BoundaryConditionType(Enum):
    constant = "constant"
    dynamic = "dynamic"
    reflective = "reflective"
    extending = "extending"
    forcing = "forcing"
    dynamic_forcing = "dynamic_forcing"


class VectorProjectionType(Enum):
    none = "none"
    normal = "normal"
    tangential = "tangential"


@dataclass
class BoundaryCondition:
    type: BoundaryConditionType
    sdf: Any  # positions [N, d] -> (signed distance [N], normals [N, d]); inside where < 0
    dirichletFunctions: Optional[Dict[str, Any]] = None  # state field -> fn(state, config, params, x, d, n, t, dt)
    updateFunctions: Optional[Dict[str, Any]] = None     # update field -> fn(...)
    forcingFunctions: Optional[List[Any]] = None         # [fn(...) -> force [N, d]]
    vectorProjectionType: VectorProjectionType = VectorProjectionType.none


@dataclass
class CompressibleSPHConfig:
    gamma: float = 1.4
    backgroundPressure: float = 0.0
    rho0: float = 1.0
    adaptiveSupportScheme: AdaptiveSupportScheme = AdaptiveSupportScheme.Monaghan
    adaptiveSupportIterations: int = 1
    adaptiveSupportThreshold: float = 1e-3
    adaptiveSupportCorrections: bool = True
    diffusionParams: DiffusionParameters = field(default_factory=buildDefaultDiffusionParamsCompressibleSPH)
    viscositySwitchParams: ViscositySwitchConfig = field(default_factory=ViscositySwitchConfig)
    schemeName: str = 'Compressible SPH'
    boundaryConditions: List[Any] = field(default_factory=list)


@dataclass
class CRKSPHConfig(CompressibleSPHConfig):
    energyScheme: EnergyScheme = EnergyScheme.CRK
    diffusionParams: DiffusionParameters = field(default_factory=buildDefaultDiffusionParamsCRKSPH)
    crkViscosityParams: CRKViscosity = field(default_factory=buildDefaultCRKViscosityParams)
    schemeName: str = 'CRKSPH'
    compatibleEnergy: bool = True
    # extensions (no reference counterpart); False keeps the reference behaviour
    alphaWeightedViscosity: bool = False
    evolveAlpha0: bool = False


# ------------------------------------------------------------------ dict round trips
def _enum_name(v):
    return v.name if isinstance(v, Enum) else v


def crkSPHConfigToDict(config: CRKSPHConfig) -> Dict[str, Any]:
    """compressibleConfig.py:43-56 + crkSPH.py:82-98."""
    d = {
        'gamma': config.gamma, 'backgroundPressure': config.backgroundPressure, 'rho0': config.rho0,
        'adaptiveSupportScheme': config.adaptiveSupportScheme.name,
        'adaptiveSupportIterations': config.adaptiveSupportIterations,
        'adaptiveSupportThreshold': config.adaptiveSupportThreshold,
        'adaptiveSupportCorrections': config.adaptiveSupportCorrections,
        'diffusionParams': {k: _enum_name(v) for k, v in config.diffusionParams.__dict__.items()},
        'viscositySwitchParams': {k: _enum_name(v) for k, v in config.viscositySwitchParams.__dict__.items()},
        'schemeName': config.schemeName,
        'boundaryConditions': [],
        'energyScheme': config.energyScheme.name,
        'compatibleEnergy': config.compatibleEnergy,
        'crkViscosityParams': dict(config.crkViscosityParams.__dict__),
        'alphaWeightedViscosity': config.alphaWeightedViscosity,
        'evolveAlpha0': config.evolveAlpha0,
    }
    return d


def dictToCRKSPHConfig(d: Dict[str, Any]) -> CRKSPHConfig:
    """crkSPH.py:98-111."""
    c = CRKSPHConfig()
    c.gamma = d['gamma']
    c.backgroundPressure = d['backgroundPressure']
    c.rho0 = d['rho0']
    c.adaptiveSupportScheme = AdaptiveSupportScheme[d['adaptiveSupportScheme']] if isinstance(d['adaptiveSupportScheme'], str) else d['adaptiveSupportScheme']
    c.adaptiveSupportIterations = d['adaptiveSupportIterations']
    c.adaptiveSupportThreshold = d['adaptiveSupportThreshold']
    c.adaptiveSupportCorrections = d['adaptiveSupportCorrections']
    dp = dict(d['diffusionParams'])
    for k in ('viscosityTerm', 'thermalConducitiyTerm'):
        if isinstance(dp.get(k), str):
            dp[k] = ViscosityTerms[dp[k]]
    c.diffusionParams = DiffusionParameters(**dp)
    vs = dict(d['viscositySwitchParams'])
    vs['scheme'] = ViscositySwitch[vs['scheme']] if isinstance(vs['scheme'], str) else vs['scheme']
    c.viscositySwitchParams = ViscositySwitchConfig(**vs)
    c.schemeName = d['schemeName']
    c.energyScheme = EnergyScheme[d['energyScheme']] if isinstance(d['energyScheme'], str) else d['energyScheme']
    c.compatibleEnergy = d['compatibleEnergy']
    c.crkViscosityParams = CRKViscosity(**d['crkViscosityParams'])
    c.alphaWeightedViscosity = d.get('alphaWeightedViscosity', False)
    c.evolveAlpha0 = d.get('evolveAlpha0', False)
    return c


# ------------------------------------------------------------------ C-ABI packing
def _host_values(t) -> list:
    """Host copy of a small (domain) tensor, cached on the tensor while its
    storage and version are unchanged, so that packing the configuration on
    every call does not wait for the GPU."""
    if not isinstance(t, torch.Tensor):
        return list(t)
    key = (t.data_ptr(), t._version, tuple(t.shape), t.dtype)
    cached = getattr(t, "_crk_host_values", None)
    if cached is not None and cached[0] == key:
        return cached[1]
    vals = t.detach().cpu().tolist()
    try:
        t._crk_host_values = (key, vals)
    except (AttributeError, RuntimeError):
        pass
    return vals


def toCrkConfig(config: SimulationConfig, scheme: CRKSPHConfig) -> abi.CrkConfig:
    """Pack the dataclasses into the POD parameter block of include/crksph.h."""
    c = abi.CrkConfig()
    dom = config.domain
    c.domain.dim = int(dom.dim)
    mn = _host_values(dom.min)
    mx = _host_values(dom.max)
    per = _host_values(dom.periodic)
    for d in range(dom.dim):
        c.domain.min[d] = float(mn[d])
        c.domain.max[d] = float(mx[d])
        c.domain.periodic[d] = int(bool(per[d]))
    kern = config.kernel
    c.kernel = int(kern.value if hasattr(kern, 'value') else kern)
    c.verletScale = float(config.verletScale)
    c.targetNeighbors = float(config.targetNeighbors)
    c.cflFactor = float(config.cflFactor)
    c.minDt = float(config.minDt if config.minDt is not None else 0.0)
    c.maxDt = float(config.maxDt if config.maxDt is not None else 1e30)
    c.dtGrowthFactor = float(config.dtGrowthFactor if config.dtGrowthFactor is not None else 1e30)
    c.gamma = float(scheme.gamma)
    c.energyScheme = int(scheme.energyScheme.value)
    if scheme.energyScheme != EnergyScheme.CRK:
        raise ValueError(f"energy scheme {scheme.energyScheme} is not built (only EnergyScheme.CRK)")
    ads = scheme.adaptiveSupportScheme
    if ads not in (AdaptiveSupportScheme.Owen, AdaptiveSupportScheme.Monaghan, AdaptiveSupportScheme.NoScheme):
        raise ValueError(f"adaptive support scheme {ads} is not built (Owen, Monaghan, NoScheme)")
    c.adaptiveSupportScheme = int(ads.value)
    c.adaptiveSupportIterations = int(scheme.adaptiveSupportIterations)
    c.adaptiveSupportThreshold = float(scheme.adaptiveSupportThreshold)
    c.compatibleEnergy = int(bool(getattr(scheme, 'compatibleEnergy', True)))
    dp = scheme.diffusionParams
    for k in ("c_s", "C_l", "C_q", "Cu_l", "Cu_q", "K", "thermalConductivity"):
        setattr(c.diffusion, k, float(getattr(dp, k)))
    c.diffusion.viscosityTerm = int(getattr(dp.viscosityTerm, 'value', dp.viscosityTerm))
    c.diffusion.thermalConductivityTerm = int(getattr(dp.thermalConducitiyTerm, 'value', dp.thermalConducitiyTerm))
    c.diffusion.scaleBeta = int(bool(dp.scaleBeta))
    c.diffusion.monaghanSwitch = int(bool(dp.monaghanSwitch))
    c.diffusion.correctXi = int(bool(dp.correctXi))
    cv = scheme.crkViscosityParams
    c.crkViscosity.eta_fold = float(cv.eta_fold)
    c.crkViscosity.eta_crit = float(cv.eta_crit)
    c.crkViscosity.enableCRKLimiter = int(bool(cv.enableCRKLimiter))
    c.crkViscosity.enableVanLeerLimiter = int(bool(cv.enableVanLeerLimiter))
    c.crkViscosity.forceVanLeerOff = int(bool(cv.forceVanLeerOff))
    c.crkViscosity.forceVanLeerOn = int(bool(cv.forceVanLeerOn))
    sp = scheme.viscositySwitchParams
    if sp.scheme not in (ViscositySwitch.NoneSwitch, ViscositySwitch.CullenDehnen2010, ViscositySwitch.Learned):
        raise ValueError(f"Unsupported viscosity switch scheme: {sp.scheme}")
    if sp.correctVelocityGradient:
        raise ValueError("correctVelocityGradient=True (M-matrix correction) is not built")
    c.switchParams.scheme = int(sp.scheme.value)
    c.switchParams.limitXi = int(bool(sp.limitXi))
    c.switchParams.correctVelocityGradient = 0
    c.switchParams.alpha_min = float(sp.alpha_min)
    c.switchParams.alpha_max = float(sp.alpha_max)
    c.switchParams.beta_c = float(sp.beta_c)
    c.switchParams.beta_d = float(sp.beta_d)
    c.switchParams.beta_xi = float(sp.beta_xi)
    c.alphaWeightedQ = int(bool(getattr(scheme, 'alphaWeightedViscosity', False)))
    c.evolveAlpha0 = int(bool(getattr(scheme, 'evolveAlpha0', False)))
    return c


def kernelConstants(kernel: KernelFunctions, dim: int):
    """(C_d, H/h) of a kernel (host call into the HIP library, no GPU needed)."""
    import ctypes as C
    cd, ks = C.c_double(), C.c_double()
    abi.call("crk_kernel_constants", int(kernel.value), int(dim), C.byref(cd), C.byref(ks))
    return cd.value, ks.value


def sphKernelScale(kernel: KernelFunctions, dim: int) -> float:
    return kernelConstants(kernel, dim)[1]


def sphKernel_xi(kernel: KernelFunctions, dim: int) -> float:
    """CFL / switch kernel factor (crk_kernel_xi): the kernel scale H/h for
    the Wendland kernels (warpSPHCore's values are not available), for B7
    the values the reference's example outputs fix."""
    import ctypes as C
    xi = C.c_double()
    abi.call("crk_kernel_xi", int(kernel.value if hasattr(kernel, 'value') else kernel), int(dim), C.byref(xi))
    return xi.value

