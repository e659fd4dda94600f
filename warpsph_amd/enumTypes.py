"""Enumerations of the CRKSPH step.

Values of the reference enums follow src/warpSPH/enumTypes.py:5-34.  The enums
that the reference imports from the un-vendored warpSPHCore /
warpSPHIntegrators packages (KernelFunctions, SupportScheme, GradientScheme,
LaplacianScheme, IntegrationSchemeType) are re-declared with the members the
CRKSPH step uses.
"""
from enum import Enum


class EnergyScheme(Enum):  # enumTypes.py:5-11
    equalWork = 0
    PdV = 1
    diminishing = 2
    monotonic = 3
    hybrid = 4
    CRK = 5


class AdaptiveSupportScheme(Enum):  # enumTypes.py:14-17
    NoScheme = 0
    Monaghan = 1
    Owen = 2


class ViscositySwitch(Enum):  # enumTypes.py:20-27 (+ Learned: per-particle MLP closure)
    Balsara1995 = 0
    Colagrossi2004 = 1
    CullenDehnen2010 = 2
    CullenHopkins = 3
    MorrisMonaghan1997 = 4
    Rosswog2000 = 5
    NoneSwitch = 6
    Learned = 100


class CompressibleSPHScheme(Enum):  # enumTypes.py:30-34
    Monaghan = 0
    CompSPH = 1
    CRKSPH = 2


class KernelFunctions(Enum):
    """Wendland kernels (Wendland 1995; Dehnen & Aly 2012) and B7, the
    centred degree-7 B-spline with knots at r/H = 1/4, 1/2, 3/4 (the kernel
    the reference's compressible examples select).  Values are the library's
    CRK_KERNEL_* codes."""
    Wendland2 = 0
    Wendland4 = 1
    Wendland6 = 2
    B7 = 3


class SupportScheme(Enum):
    Gather = 0
    Scatter = 1
    Symmetric = 2
    SuperSymmetric = 3
    KernelMeanSymmetric = 4


class GradientScheme(Enum):
    Naive = 0
    Difference = 1
    Summation = 2


class LaplacianScheme(Enum):
    Brookshaw = 0


class IntegrationSchemeType(Enum):
    rungeKutta2 = 2


class SamplingScheme(Enum):
    regular = 0
