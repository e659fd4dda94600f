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


# ------------------------------------------------------------------ B7
B7 = 3


def test_b7_is_the_degree_7_b_spline():
    """k(q) = 7!/4^7 M8(4 q): the centred B-spline of eight unit boxes
    (scipy's B-spline basis element on the knots -4 .. 4) on the support."""
    from scipy.interpolate import BSpline
    m8 = BSpline.basis_element(np.arange(-4.0, 5.0), extrapolate=False)
    q = np.linspace(0.0, 0.999, 1999)
    k, dk = O.kernel_eval(B7, 2, q)
    want = math.factorial(7) / 4.0 ** 7 * m8(4.0 * q)
    dwant = math.factorial(7) / 4.0 ** 6 * m8.derivative()(4.0 * q)
    assert np.max(np.abs(k - want)) < 1e-14
    assert np.max(np.abs(dk - dwant)) < 1e-13
    assert k[0] == pytest.approx(151.0 / 1024.0, rel=1e-15) and dk[0] == 0.0


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_b7_normalisation_scale_and_support(dim):
    """C_d from the closed-form moments, the quadrature integral 1, H/h from
    the second moment (1D: sqrt(6) exactly, eight unit boxes have variance
    2/3 on a support of 4), compact support and C^6 smoothness at the knots
    (the inner piece's polynomial and the truncated powers meet)."""
    C, scale = O.kernel_constants(B7, dim)
    I = {1: 315 / 8192, 2: 7435 / 1179648, 3: 105 / 65536}[dim]
    assert C == pytest.approx(1.0 / (SURFACE[dim] * I), rel=1e-14)
    integral = quad(lambda q: SURFACE[dim] * q ** (dim - 1) * C * O.kernel_eval(B7, dim, q)[0])
    assert integral == pytest.approx(1.0, abs=1e-9)
    m2 = quad(lambda q: SURFACE[dim] * q ** (dim + 1) * C * O.kernel_eval(B7, dim, q)[0])
    assert scale == pytest.approx(1.0 / (2.0 * math.sqrt(m2 / dim)), rel=1e-9)
    if dim == 1:
        assert scale == pytest.approx(math.sqrt(6.0), rel=1e-15)
    k, dk = O.kernel_eval(B7, dim, np.array([1.0, 1.2, 1.0 - 1e-9]))
    assert k[0] == 0.0 and k[1] == 0.0 and dk[0] == 0.0 and abs(k[2]) < 1e-50
    for knot in (0.25, 0.5, 0.75):
        a, da = O.kernel_eval(B7, dim, np.array([knot - 1e-12]))
        b, db = O.kernel_eval(B7, dim, np.array([knot + 1e-12]))
        assert abs(a[0] - b[0]) < 1e-11 and abs(da[0] - db[0]) < 1e-10
    q = np.linspace(0.01, 0.99, 97)
    eps = 1e-6
    fd = (O.kernel_eval(B7, dim, q + eps)[0] - O.kernel_eval(B7, dim, q - eps)[0]) / (2 * eps)
    assert np.max(np.abs(fd - O.kernel_eval(B7, dim, q)[1])) < 1e-8
    kk, dd = O.kernel_eval(B7, dim, np.linspace(0, 0.999, 500))
    assert np.all(np.diff(kk) <= 0) and np.all(dd <= 0)


@pytest.mark.parametrize("kernel,dim", [(k, d) for k in range(4) for d in (1, 2, 3)])
def test_xi_of_the_cfl_step(kernel, dim):
    """xi is the kernel scale for the Wendland kernels; for B7 the 1D and 2D
    values fixed by the reference's Sod and Kelvin-Helmholtz outputs (the GPU
    tests reproduce those steps), ~1.58 times the kernel scale."""
    _, scale = O.kernel_constants(kernel, dim)
    xi = O.kernel_xi(kernel, dim)
    if kernel < 3:
        assert xi == scale
    else:
        want = {1: 3.865665, 2: 3.932735}.get(dim, 3.932735 * scale / O.kernel_constants(3, 2)[1])
        assert xi == pytest.approx(want, rel=1e-15)
        assert 1.57 < xi / scale < 1.59
