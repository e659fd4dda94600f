"""crk_cfl_dt (computeTimestep, timestep/compressible.py:12-66) on the GPU
against the oracle's FP32 CFL step, with the growth limit and the clamps."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _system(nx, seed):
    from warpsph_amd import sample
    system, config, scheme = sample.shear_layer_2d(nx=nx, device="cuda")
    g = torch.Generator().manual_seed(seed)
    n = system.state.numParticles
    st = system.state
    st.internalEnergies = (0.2 + 3.0 * torch.rand(n, generator=g)).to("cuda")
    st.supports = (st.supports.cpu() * (0.8 + 0.4 * torch.rand(n, generator=g))).to("cuda")
    return system, config, scheme


@pytest.mark.parametrize("nx,seed", [(32, 1), (200, 2)])
def test_cfl_dt_matches_oracle(nx, seed):
    from oracle import oracle as O
    from warpsph_amd import sample
    system, config, scheme = _system(nx, seed)
    print(f"cfl: nx {nx} seed {seed}", flush=True)
    sim = O.from_system(system, config, scheme, 32)
    want = sim.cfl_dt()
    got = sample.cfl_dt(system.state, config, scheme)
    assert got == pytest.approx(want, rel=2e-7)
    # growth limit from a smaller previous step, none from a larger one
    for prev in (0.5 * want, 0.95 * want, 3.0 * want):
        assert sample.cfl_dt(system.state, config, scheme, prev=prev) == pytest.approx(sim.cfl_dt(prev), rel=2e-7)
    assert sample.cfl_dt(system.state, config, scheme, prev=0.5 * want) == \
        pytest.approx(float(np.float32(config.dtGrowthFactor)) * 0.5 * want, rel=2e-7)


def test_cfl_dt_clamps():
    from oracle import oracle as O
    from warpsph_amd import sample
    system, config, scheme = _system(32, 3)
    raw = sample.cfl_dt(system.state, config, scheme)
    config.maxDt = 0.25 * raw
    assert sample.cfl_dt(system.state, config, scheme) == pytest.approx(float(np.float32(config.maxDt)), rel=1e-7)
    config.maxDt = 1.0
    config.minDt = 4.0 * raw
    got = sample.cfl_dt(system.state, config, scheme)
    assert got == pytest.approx(float(np.float32(config.minDt)), rel=1e-7)
    assert got == pytest.approx(O.from_system(system, config, scheme, 32).cfl_dt(), rel=1e-7)
