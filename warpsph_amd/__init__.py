"""MI355X-native CRKSPH compressible step (drop-in for warpSPH's
``buildScheme(CompressibleSPHScheme.CRKSPH)``).

The step runs in hand-written gfx950 HIP kernels (``warpsph_amd/csrc``) behind
the C-ABI of ``include/crksph.h``; this package is the host side that keeps
warpSPH's SimulationState / SimulationConfig / SimulationUpdate API.
"""
from .configurations import (BoundaryCondition, BoundaryConditionType, CRKSPHConfig, CRKViscosity,
                             CompressibleSPHConfig, DiffusionParameters,
                             DomainDescription, SimulationConfig, ViscositySwitchConfig, buildConfig,
                             buildDomainDescription, crkSPHConfigToDict, dictToCRKSPHConfig, sphKernel_xi,
                             sphKernelScale, toCrkConfig)
from .enumTypes import (AdaptiveSupportScheme, CompressibleSPHScheme, EnergyScheme, IntegrationSchemeType,
                        KernelFunctions, SupportScheme, ViscositySwitch)
from .integrator import getIntegrator
from .scheme import buildScheme, crkSPH_step
from .systems import AdjacencyList, CompressibleSystemUpdate, CompSPHState, CompSPHSystem
from .utils import n_h_to_nH, nH_to_n_h, volumeToSupportHelper

__all__ = [
    "AdaptiveSupportScheme", "AdjacencyList", "BoundaryCondition", "BoundaryConditionType", "CRKSPHConfig", "CRKViscosity", "CompSPHState", "CompSPHSystem",
    "CompressibleSPHConfig", "CompressibleSPHScheme", "CompressibleSystemUpdate", "DiffusionParameters",
    "DomainDescription", "EnergyScheme", "IntegrationSchemeType", "KernelFunctions", "SimulationConfig",
    "SupportScheme", "ViscositySwitch", "ViscositySwitchConfig", "buildConfig", "buildDomainDescription",
    "buildScheme", "crkSPHConfigToDict", "crkSPH_step", "dictToCRKSPHConfig", "getIntegrator", "n_h_to_nH",
    "nH_to_n_h", "sphKernelScale", "sphKernel_xi", "toCrkConfig", "volumeToSupportHelper",
]
