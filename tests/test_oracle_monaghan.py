"""Pins of the oracle's Monaghan adaptive support (no GPU).

evaluateOptimalSupportMonaghan (optimalSupportMonaghan.py:20-108) is one
Newton step on F(h) = h - volumeToSupportHelper(m / rho(h), n_H, d) per
iteration.  The checks recompute everything by brute force over all particle
pairs in numpy:

  one iteration   h' = h - F / (Omega + 1e-6) with Omega = 1 - (dh/drho) drho/dh,
                  drho/dh by central finite differences of the brute-force
                  density (independent of the oracle's analytic dW/dh)
  fixed point     many iterations converge to F(h) = 0
"""
import numpy as np
import torch

from tests.test_oracle_pins import _kernel


def _system(nx=20, amp=0.25, seed=4):
    from warpsph_amd import sample
    from warpsph_amd.enumTypes import AdaptiveSupportScheme
    system, config, scheme = sample.shear_layer_2d(nx=nx, device="cpu", seed=seed)
    scheme.adaptiveSupportScheme = AdaptiveSupportScheme.Monaghan
    st = system.state
    x = st.positions.double()
    # supports off their equilibrium so that F is far from zero
    st.supports = (st.supports.double() * (1.0 + amp * torch.sin(2 * np.pi * x[:, 0]) *
                                           torch.cos(2 * np.pi * x[:, 1]))).float()
    return system, config, scheme


def _density(x, m, h, kernel, L):
    """rho_i = sum_j m_j W(|x_i - x_j|, h_i) over all j (minimum image)."""
    dx = x[:, None, :] - x[None, :, :]
    dx -= L * np.round(dx / L)
    r = np.linalg.norm(dx, axis=2)
    W, _ = _kernel(kernel, x.shape[1], r, h[:, None])
    return (W * m[None, :]).sum(axis=1)


def _H(V, nT, dim):
    if dim == 1:
        return nT * V / 2
    if dim == 2:
        return np.sqrt(nT * V / np.pi)
    return (nT * V / np.pi * 3 / 4) ** (1 / 3)


def _oracle_h(system, config, scheme):
    from oracle import oracle as O
    sim = O.from_system(system, config, scheme, 64)
    sim.evaluate(sim.cfl_dt())
    return sim.get("h")


def test_monaghan_newton_step_against_brute_force():
    from warpsph_amd.configurations import toCrkConfig
    system, config, scheme = _system()
    cfg = toCrkConfig(config, scheme)
    st = system.state
    x = st.positions.double().numpy()
    m = st.masses.double().numpy()
    h = st.supports.double().numpy()
    L = (config.domain.max - config.domain.min).double().numpy()
    nT, dim = float(np.float32(config.targetNeighbors)), 2
    rho = _density(x, m, h, cfg.kernel, L)
    eps = 1e-6 * h
    drho = (_density(x, m, h + eps, cfg.kernel, L) - _density(x, m, h - eps, cfg.kernel, L)) / (2 * eps)
    F = h - _H(m / rho, nT, dim)
    omega = 1.0 + h / (rho * dim) * drho
    expect = np.clip(h - F / (omega + 1e-6), 0.25 * h.min(), 4.0 * h.max())
    got = _oracle_h(system, config, scheme)
    assert np.max(np.abs(F)) > 1e-2 * h.mean()  # the step is not trivial
    assert np.max(np.abs(got - expect) / h) < 1e-7, np.max(np.abs(got - expect) / h)


def test_monaghan_iterations_reach_the_fixed_point():
    from warpsph_amd.configurations import toCrkConfig
    system, config, scheme = _system(amp=0.1)
    scheme.adaptiveSupportIterations = 40
    scheme.adaptiveSupportThreshold = 1e-13
    cfg = toCrkConfig(config, scheme)
    st = system.state
    x = st.positions.double().numpy()
    m = st.masses.double().numpy()
    L = (config.domain.max - config.domain.min).double().numpy()
    got = _oracle_h(system, config, scheme)
    rho = _density(x, m, got, cfg.kernel, L)
    F = got - _H(m / rho, float(np.float32(config.targetNeighbors)), 2)
    assert np.max(np.abs(F) / got) < 1e-9, np.max(np.abs(F) / got)
    # and it moved the supports: the start was off the fixed point
    assert np.max(np.abs(got - st.supports.double().numpy()) / got) > 1e-2
