"""Initial-condition generators (warpsph_amd/generators.py) on the CPU:
lattice geometry of sampleRegularParticles, and the field assembly of
sampleKHH, buildSod1D and buildSedov called with the reference's signatures
the way examples/compressible/01-sod-shock-tube-1d.py:136-157 calls them
(sample/regular.py:6-56, kelvinHelmholtz/sample.py:17-75, sod/sod.py:21-225,
sedov/initial.py:24-218).  The GPU steps inside the generators (adaptive
supports, Verlet list, CFL step) are replaced here by recording stand-ins;
tests/test_gpu_generators.py runs them for real."""
import math
import warnings

import numpy as np
import pytest
import torch

from warpsph_amd import generators as G
from warpsph_amd.configurations import CRKSPHConfig, buildConfig, buildDomainDescription
from warpsph_amd.enumTypes import AdaptiveSupportScheme, CompressibleSPHScheme, KernelFunctions, SupportScheme
from warpsph_amd.generators import buildPointCloud, sampleKHH, sampleRegularParticles
from warpsph_amd.scheme import buildScheme
from warpsph_amd.utils import volumeToSupportHelper


class _GPUSteps:
    """Recording stand-ins for the generators' GPU steps."""

    def __init__(self, monkeypatch, dt=2.5e-4, rho_scale=1.0):
        self.support_calls, self.dt_calls, self.verlet_calls = [], [], []
        self.dt, self.rho_scale = dt, rho_scale
        monkeypatch.setattr(G, "evaluateOptimalSupport", self._support)
        monkeypatch.setattr(G, "computeTimestep", self._timestep)
        monkeypatch.setattr(G, "buildVerletList", self._verlet)

    def _support(self, state, config, compParams, supportScheme=SupportScheme.Gather, adjacency=None):
        self.support_calls.append(dict(h=state.supports.clone(), rho=state.densities.clone(),
                                       m=state.masses.clone(), compParams=compParams, scheme=supportScheme))
        # "adapted" supports 10 % larger, densities scaled
        return state.densities * self.rho_scale, state.supports * 1.1, None, None, None

    def _timestep(self, system, config, compParams, dt=None):
        self.dt_calls.append(dict(dt=dt, gamma=compParams.gamma))
        return self.dt

    def _verlet(self, state, config):
        self.verlet_calls.append(state)
        return "verlet-list"


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


def test_kelvin_helmholtz_fields(monkeypatch):
    from warpsph_amd.configurations import DomainDescription
    steps = _GPUSteps(monkeypatch)
    dom = DomainDescription(torch.tensor([0.0, 0.0]), torch.tensor([1.0, 1.0]), torch.tensor([True, True]), 2)
    config, _ = buildConfig(device=torch.device("cpu"), domain=dom, dim=2, targetNeighbors=32)
    scheme = CRKSPHConfig(gamma=5.0 / 3.0)
    rho1, rho2, v1, v2, delta, sigma, freq, w0 = 1.0, 2.0, 0.5, -0.5, 0.025, 0.05, 4, 0.01
    SimulationSystem, SimulationState, *_ = buildScheme(CompressibleSPHScheme.CRKSPH)
    config.dt = 1e-3
    system = sampleKHH(rho1, rho2, v1, v2, delta, sigma, freq, w0, 32, config, scheme, SimulationState,
                       SimulationSystem)
    assert isinstance(system, SimulationSystem)
    s = system.state
    # computeTimestep(dt=config.dt) in the lattice set-up, then again * 2/3
    assert [c["dt"] for c in steps.dt_calls] == [1e-3, steps.dt]
    assert config.dt == pytest.approx(steps.dt * 2 / 3)
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


def _sod_setup(nx=200, ratio=8, device="cpu"):
    """The example's configuration (01-sod-shock-tube-1d.py:109-135)."""
    from warpsph_amd.configurations import DomainDescription
    dom = DomainDescription(torch.tensor([-1.0]), torch.tensor([1.0]), torch.tensor([True]), 1)
    config, _ = buildConfig(device=torch.device(device), domain=dom, dim=1, kernel=KernelFunctions.Wendland4,
                            targetNeighbors=6.0, dt=1e-3, adaptiveDt=True, cflFactor=0.3)
    config.nx = nx
    SimulationSystem, SimulationState, SimulationConfig, SimulationUpdate, fn, export_fn, import_fn = \
        buildScheme(CompressibleSPHScheme.CRKSPH)
    schemeConfig = SimulationConfig()
    schemeConfig.gamma = 1.4
    schemeConfig.adaptiveSupportScheme = AdaptiveSupportScheme.Owen
    left = G.sodInitialState(p=1.0, rho=1.0, v=0.0)
    right = G.sodInitialState(p=0.1, rho=0.125, v=0.0)
    return config, schemeConfig, SimulationSystem, SimulationState, left, right


@pytest.mark.parametrize("smoothIC", [False, True])
def test_sod_tube_assembly(monkeypatch, smoothIC):
    steps = _GPUSteps(monkeypatch, rho_scale=1.03)
    nx, ratio, gamma = 256, 8, 1.4  # nx // ratio even: no right-lattice point at x = 0
    config, schemeConfig, SimulationSystem, SimulationState, left, right = _sod_setup(nx, ratio)
    compSystem = G.buildSod1D(SimulationSystem, SimulationState, ratio, left, right, gamma, config, smoothIC,
                              adaptiveSupportScheme=schemeConfig.adaptiveSupportScheme)
    assert isinstance(compSystem, SimulationSystem) and compSystem.adjacency == "verlet-list"
    st = compSystem.state
    nr = nx // ratio
    x = st.positions[:, 0].double().numpy()
    lt = st.materials.numpy() == 0
    assert st.positions.shape == (nx + nr, 1) and int(lt.sum()) == nx
    # the left lattice is cell-centred on [-1/2, 1/2], the right one on the outer halves
    assert np.allclose(np.sort(x[lt]), -0.5 + (np.arange(nx) + 0.5) / nx, atol=1e-6)
    xr = np.sort(x[~lt])
    base = -0.5 + (np.arange(nr) + 0.5) / nr
    assert np.allclose(xr, np.sort(np.where(base < 0, base - 0.5, base + 0.5)), atol=1e-6)
    # masses: m_l rho_l and m_l rho_r nx / nr
    m = st.masses.double().numpy()
    assert np.allclose(m[lt], 1.0 / nx, rtol=1e-6) and np.allclose(m[~lt], 0.125 / nx * (nx / nr), rtol=1e-6)
    # the adaptive-support call: lattice supports, 16 iterations, 1e-3, the gather form
    call = steps.support_calls[0]
    h0 = call["h"].double().numpy()
    assert np.allclose(h0[lt], volumeToSupportHelper(1.0 / nx, 6.0, 1), rtol=1e-6)
    assert np.allclose(h0[~lt], volumeToSupportHelper(1.0 / nr, 6.0, 1), rtol=1e-6)
    cp = call["compParams"]
    assert (cp.gamma, cp.adaptiveSupportIterations, cp.adaptiveSupportThreshold, cp.adaptiveSupportScheme) == \
        (gamma, 16, 1e-3, AdaptiveSupportScheme.Owen) and call["scheme"] == SupportScheme.Gather
    assert torch.allclose(st.supports, call["h"] * 1.1)
    rho = st.densities.double().numpy()
    u = st.internalEnergies.double().numpy()
    P = st.pressures.double().numpy()
    u_l, u_r = 1.0 / (0.4 * 1.0), 0.1 / (0.4 * 0.125)
    if not smoothIC:
        # sampled densities, u = P / ((gamma - 1) rho): the EOS gives back P
        assert np.allclose(rho, np.where(lt, 1.0, 0.125))
        assert np.allclose(u, np.where(lt, u_l, u_r), rtol=1e-6)
        assert np.allclose(P, np.where(lt, 1.0, 0.1), rtol=1e-6)
    else:
        # the adapted densities; u blended across x = +-1/2 by a logistic ramp
        # of width dx_l / 2: the sides keep u_l, u_r, the blend is monotone
        assert np.allclose(rho, 1.03 * np.where(lt, 1.0, 0.125), rtol=1e-6)
        dxh = 0.5 / nx
        far = np.minimum(np.abs(x - 0.5), np.abs(x + 0.5)) > 40 * dxh
        assert np.allclose(u[far], np.where(lt, u_l, u_r)[far], rtol=1e-6)
        inner = (x > 0) & (x < 0.5)
        o = np.argsort(x[inner])
        assert np.all(np.diff(u[inner][o]) <= 1e-6 * u_l)
        # the particles next to the interface sit half-way up the ramp's logistic
        k = np.argmin(np.abs(x - (0.5 - 0.5 / nx)))
        s = 1 / (np.exp((x[k] - 0.5) / dxh) + 1)
        assert u[k] == pytest.approx((u_l - u_r) * s + u_r, rel=1e-5)
        assert np.allclose(P, 0.4 * rho * u, rtol=1e-6)
    assert np.allclose(st.soundspeeds.double().numpy(), np.sqrt(1.4 * 0.4 * np.abs(u)), rtol=1e-6)
    assert np.allclose(st.totalEnergies.double().numpy(), u * m, rtol=1e-6)
    assert config.dx == pytest.approx(float(m.min())) and config.dt == steps.dt
    assert steps.dt_calls == [dict(dt=None, gamma=gamma)]
    assert st.UIDcounter == nx and torch.equal(st.alphas, torch.ones_like(st.densities))


def test_sedov_assembly(monkeypatch):
    from warpsph_amd.configurations import DomainDescription
    steps = _GPUSteps(monkeypatch)
    SimulationSystem, SimulationState, *_ = buildScheme(CompressibleSPHScheme.CRKSPH)
    dom = buildDomainDescription(2.0, 2, periodic=True)
    config, _ = buildConfig(device=torch.device("cpu"), domain=dom, dim=2, targetNeighbors=32, dt=1e-3)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        system = G.buildSedov(SimulationSystem, SimulationState, config, 20, 2, 2.0, True, 1.0, 2.0, 'singular',
                              5 / 3, None, 32, torch.float32, 'cpu')
    assert any("nx + 1" in str(x.message) for x in w)
    st = system.state
    assert st.positions.shape == (21 * 21, 2)
    E = (st.internalEnergies.double() * st.masses.double())
    assert int((E > 0).sum()) == 1 and float(E.sum()) == pytest.approx(2.0, rel=1e-6)
    k = int(torch.argmax(E))
    assert float(torch.linalg.norm(st.positions[k])) < 1e-6
    assert steps.dt_calls == [dict(dt=1e-3, gamma=5 / 3)] and config.dt == steps.dt
    cp = steps.support_calls[0]["compParams"]
    assert cp.adaptiveSupportScheme == AdaptiveSupportScheme.Owen and cp.adaptiveSupportIterations == 16
    assert torch.allclose(st.masses, torch.full_like(st.masses, (2.0 / 21) ** 2))
    # quadrant: 2^d particles share E0 (nx made even)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        q = G.buildSedov(SimulationSystem, SimulationState, config, nx=21, initialization='quadrant',
                         device='cpu', targetNeighbors=32).state
    Eq = (q.internalEnergies.double() * q.masses.double())
    assert q.positions.shape[0] == 400 and int((Eq > 0).sum()) == 4
    assert np.allclose(Eq[Eq > 0].numpy(), 0.25)
    with pytest.raises(NotImplementedError):
        G.buildSedov(SimulationSystem, SimulationState, config, nx=20, device='cpu')  # 'hat', the default
    with pytest.raises(NotImplementedError):
        G.buildSedov(SimulationSystem, SimulationState, config, nx=21, initialization='singular', device='cpu',
                     shellSampling=True)


def test_generator_gpu_steps_refuse_host_tensors():
    """No host path: the adaptive supports need the GPU library."""
    config, schemeConfig, SimulationSystem, SimulationState, left, right = _sod_setup(64, 4)
    with pytest.raises(Exception):
        G.buildSod1D(SimulationSystem, SimulationState, 4, left, right, 1.4, config, False,
                     adaptiveSupportScheme=AdaptiveSupportScheme.Owen)
