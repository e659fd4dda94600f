"""Pins of the oracle's SPH kernel functions (Wendland 1995; Dehnen & Aly 2012).

normalisation by quadrature, derivative by finite differences, compact
support, and the kernel scale H/h (h = twice the standard deviation) against
quadrature and against the published Table 1 values of Dehnen & Aly (2012).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

SURFACE = {1: 2.0, 2: 2.0 * math.pi, 3: 4.0 * math.pi}
# Dehnen & Aly (2012), Table 1, H/h for Wendland C2, C4, C6 in 1D, 2D, 3D
DA_TABLE = {
    (0, 1): 1.620185, (0, 2): 1.897367, (0, 3): 1.936492,
    (1, 1): 1.936492, (1, 2): 2.171239, (1, 3): 2.207940,
    (2, 1): 2.207940, (2, 2): 2.415230, (2, 3): 2.449490,
}
CLOSED_C = {
    (0, 1): 5 / 4, (0, 2): 7 / math.pi, (0, 3): 21 / (2 * math.pi),
    (1, 1): 3 / 2, (1, 2): 9 / math.pi, (1, 3): 495 / (32 * math.pi),
    (2, 1): 55 / 32, (2, 2): 78 / (7 * math.pi), (2, 3): 1365 / (64 * math.pi),
}
ALL = [(k, d) for k in range(3) for d in (1, 2, 3)]


def quad(f, n=200001):
    q = np.linspace(0.0, 1.0, n)
    y = f(q)
    return float(np.trapezoid(y, q))


@pytest.mark.parametrize("kernel,dim", ALL)
def test_normalisation(kernel, dim):
    C, _ = O.kernel_constants(kernel, dim)
    assert C == pytest.approx(CLOSED_C[(kernel, dim)], rel=1e-14)
    integral = quad(lambda q: SURFACE[dim] * q ** (dim - 1) * C * O.kernel_eval(kernel, dim, q)[0])
    assert integral == pytest.approx(1.0, abs=1e-9)


@pytest.mark.parametrize("kernel,dim", ALL)
def test_derivative_matches_finite_difference(kernel, dim):
    q = np.linspace(0.01, 0.99, 97)
    eps = 1e-6
    kp, _ = O.kernel_eval(kernel, dim, q + eps)
    km, _ = O.kernel_eval(kernel, dim, q - eps)
    _, dk = O.kernel_eval(kernel, dim, q)
    fd = (kp - km) / (2 * eps)
    assert np.max(np.abs(fd - dk)) < 1e-7 * max(1.0, np.max(np.abs(dk)))


@pytest.mark.parametrize("kernel,dim", ALL)
def test_compact_support_and_smoothness(kernel, dim):
    k, dk = O.kernel_eval(kernel, dim, np.array([0.0, 1.0, 1.5, 1.0 - 1e-9]))
    assert k[0] == 1.0  # k(0) = 1 for every Wendland polynomial here
    assert k[1] == 0.0 and k[2] == 0.0 and dk[1] == 0.0 and dk[2] == 0.0
    assert dk[0] == 0.0  # smooth at the origin
    assert abs(k[3]) < 1e-20 and abs(dk[3]) < 1e-15
    q = np.linspace(0, 0.999, 500)
    kk, dd = O.kernel_eval(kernel, dim, q)
    assert np.all(np.diff(kk) <= 0) and np.all(dd <= 0)  # monotone decreasing


@pytest.mark.parametrize("kernel,dim", ALL)
def test_kernel_scale(kernel, dim):
    C, scale = O.kernel_constants(kernel, dim)
    m2 = quad(lambda q: SURFACE[dim] * q ** (dim + 1) * C * O.kernel_eval(kernel, dim, q)[0])
    sigma = math.sqrt(m2 / dim)
    assert scale == pytest.approx(1.0 / (2.0 * sigma), rel=1e-9)
    assert scale == pytest.approx(DA_TABLE[(kernel, dim)], abs=5e-6)
