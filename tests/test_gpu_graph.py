"""Steps replayed as HIP graphs (warpsph_amd/graph.py) against eager steps:
the same bits for an ensemble and for a single run, with Verlet-list
failures redone eagerly."""
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def unbind():
    yield
    from warpsph_amd.runtime import runtime
    runtime("cuda:0").bind_runs(None)


FIELDS = ("positions", "velocities", "internalEnergies", "supports", "densities", "pressures", "alphas")


def _eager(system, config, scheme, nsteps, ensemble):
    from warpsph_amd import sample
    from warpsph_amd.ensemble import cfl_dt
    from warpsph_amd.integrator import getIntegrator
    from warpsph_amd.scheme import crkSPH_step
    integ = getIntegrator()
    s = system
    prev = None
    for _ in range(nsteps):
        # the growth limit from the previous dt as the stepper keeps it (FP32)
        if ensemble:
            dt = cfl_dt(s, config, scheme, prev)
            prev = dt.astype(np.float32).astype(np.float64)
        else:
            dt = sample.cfl_dt(s.state, config, scheme, prev)
            prev = float(np.float32(dt))
        s = integ.function(s, crkSPH_step, dt, config, scheme).state
    return s


@pytest.mark.parametrize("min_life", [0.0, 4.0, 1e9])
def test_graph_steps_equal_eager_steps_ensemble(min_life):
    """min_life 0: every list is captured; 1e9: eager steps once two list
    lifetimes are known; 4: a switch between the two within the run."""
    from warpsph_amd.ensemble import concatenate, two_state_run
    from warpsph_amd.graph import GraphStepper
    announce(f"graph ensemble min_life={min_life}")
    sides = (24, 32, 40)
    mk = lambda: concatenate([two_state_run(nx, 77 + r, device="cuda")[0] for r, nx in enumerate(sides)])
    _, config, scheme = two_state_run(24, 77, device="cuda")
    nsteps = 14
    ref = _eager(mk(), config, scheme, nsteps, True)
    sysg = mk()
    gs = GraphStepper(sysg, config, scheme)
    gs.min_life = min_life
    for _ in range(nsteps):
        gs.step()
    torch.cuda.synchronize()
    print(f"replays {gs.replays}, captures {gs.captures}, eager redos {gs.redos}, eager steps {gs.eager_steps}")
    if min_life == 0.0:
        assert gs.replays >= 4 and gs.redos >= 1
    if min_life == 1e9:
        assert gs.eager_steps >= nsteps // 2
    for f in FIELDS:
        assert torch.equal(getattr(sysg.state, f), getattr(ref.state, f)), f
    assert np.allclose(gs.times(), ref.t, rtol=1e-6, atol=0)


def test_graph_steps_equal_eager_steps_single_run():
    from warpsph_amd import sample
    from warpsph_amd.graph import GraphStepper
    announce("graph single run")
    nsteps = 10
    s0, config, scheme = sample.shear_layer_2d(nx=48, device="cuda")
    ref = _eager(s0, config, scheme, nsteps, False)
    s1, _, _ = sample.shear_layer_2d(nx=48, device="cuda")
    gs = GraphStepper(s1, config, scheme)
    gs.min_life = 0.0
    for _ in range(nsteps):
        gs.step()
    torch.cuda.synchronize()
    print(f"replays {gs.replays}, captures {gs.captures}, eager redos {gs.redos}")
    assert gs.replays >= 4
    for f in FIELDS:
        assert torch.equal(getattr(s1.state, f), getattr(ref.state, f)), f
    assert abs(gs.times()[0] - ref.t) <= 1e-6 * ref.t
