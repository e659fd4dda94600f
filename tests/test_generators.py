"""Initial-condition generators (warpsph_amd/generators.py) on the CPU:
lattice geometry of sampleRegularParticles and the Kelvin-Helmholtz fields
of sampleKHH (sample/regular.py:6-56, kelvinHelmholtz/sample.py:10-75)."""
import math

import numpy as np
import torch

from warpsph_amd.configurations import CRKSPHConfig, buildConfig, buildDomainDescription
from warpsph_amd.generators import buildPointCloud, sampleKHH, sampleRegularParticles
from warpsph_amd.utils import volumeToSupportHelper


def test_periodic_lattice_is_cell_centred():
    dom = buildDomainDescription(2.0, 2, periodic=True)
    pc, area, h = buildPointCloud(8, dom, targetNeighbors=32)
    x = pc.positions.numpy()
    assert x.shape == (64, 2)
    assert np.allclose(np.unique(x[:, 0]), -1.0 + 0.125 + 0.25 * np.arange(8), atol=1e-6)
    # meshgrid(indexing='xy'): x varies fastest
    assert np.allclose(x[:8, 1], x[0, 1]) and np.allclose(np.diff(x[:8, 0]), 0.25, atol=1e-6)
    assert abs(area - 0.0625) < 1e-7
    assert abs(h - volumeToSupportHelper(0.0625, 32, 2)) < 1e-7
    assert np.allclose(pc.supports.numpy(), h)


def test_open_lattice_spans_the_ends():
    # regular.py:13, 24-30: dx = l / (nx - 1), then ceil(l / dx) points from min to max
    # (so nx = 11 on [-0.5, 0.5] gives 10 points, as in the reference)
    dom = buildDomainDescription(1.0, 1, periodic=False)
    ps = sampleRegularParticles(11, dom, targetNeighbors=6)
    x = ps.positions.numpy()[:, 0]
    nd = int(np.ceil(np.float32(1.0) / np.float32(np.float32(1.0) / np.float32(10))))
    assert np.allclose(x, np.linspace(-0.5, 0.5, nd), atol=1e-6)
    assert np.allclose(ps.masses.numpy(), 0.1) and np.allclose(ps.densities.numpy(), 1.0)


def test_jitter_is_one_sided_and_seeded():
    dom = buildDomainDescription(1.0, 2, periodic=True)
    g = torch.Generator().manual_seed(3)
    a = sampleRegularParticles(16, dom, 32, jitter=0.1, generator=g).positions
    b = sampleRegularParticles(16, dom, 32).positions
    d = (a - b).numpy()
    assert np.all(d >= 0) and np.all(d <= 0.1 / 16 + 1e-7)


def test_kelvin_helmholtz_fields():
    from warpsph_amd.configurations import DomainDescription
    dom = DomainDescription(torch.tensor([0.0, 0.0]), torch.tensor([1.0, 1.0]), torch.tensor([True, True]), 2)
    config, _ = buildConfig(device=torch.device("cpu"), domain=dom, dim=2, targetNeighbors=32)
    scheme = CRKSPHConfig(gamma=5.0 / 3.0)
    rho1, rho2, v1, v2, delta, sigma, freq, w0 = 1.0, 2.0, 0.5, -0.5, 0.025, 0.05, 4, 0.01
    s = sampleKHH(rho1, rho2, v1, v2, delta, sigma, freq, w0, 32, config, scheme).state
    x = s.positions.double().numpy()
    y = x[:, 1]
    band = (y >= 0.25) & (y < 0.75)
    assert np.allclose(s.densities.numpy(), np.where(band, rho2, rho1))
    vm = (v1 - v2) / 2
    vx = np.where(y < 0.25, v1 - vm * np.exp((y - 0.25) / delta),
                  np.where(y < 0.5, v2 + vm * np.exp((0.25 - y) / delta),
                           np.where(y < 0.75, v2 + vm * np.exp((y - 0.75) / delta),
                                    v1 - vm * np.exp((0.75 - y) / delta))))
    assert np.allclose(s.velocities[:, 0].numpy(), vx, atol=1e-6)
    th = 1 / (2 * sigma * sigma)
    vy = w0 * np.sin(freq * math.pi * x[:, 0]) * (np.exp(-(y - 0.25) ** 2 * th) + np.exp(-(y - 0.75) ** 2 * th)) \
        * np.abs(0.5 - y)
    assert np.allclose(s.velocities[:, 1].numpy(), vy, atol=1e-7)
    assert np.allclose(s.pressures.numpy(), 2.5)
    assert np.allclose(s.internalEnergies.numpy(), 2.5 / (s.densities.numpy() * (2.0 / 3.0)), rtol=1e-6)
    assert np.allclose(s.masses.numpy(), (1.0 / 32) ** 2 * s.densities.numpy(), rtol=1e-5)
