"""Host-side contracts that need no GPU: bench.py's command line, refusals of unsupported options."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2
    line = [x for x in r.stdout.splitlines() if x.startswith("{")][-1]
    assert "one GPU" in json.loads(line)["error"]


def test_boundary_hooks_follow_the_reference_semantics():
    """bcs.py:16-92 on CPU tensors: Dirichlet values replace a state field
    where sdf < 0 (positions wrapped into the periodic domain first), forcing
    sums over conditions, update functions replace update fields."""
    import torch
    from types import SimpleNamespace
    from warpsph_amd import boundary, sample
    system, config, scheme = sample.shear_layer_2d(nx=8, device="cpu")
    st = system.state
    st.positions[0] = torch.tensor([1.25, 0.5])  # outside [0, 1): wraps to x = 0.25
    calls = []

    def sdf(p):  # left half of the box
        return p[:, 0] - 0.5, torch.zeros_like(p)

    def zero_v(state, cfg, comp, p, d, nrm, t, dt):
        calls.append((t, dt))
        return torch.zeros_like(state.velocities)

    def push(state, cfg, comp, p, d, nrm, t, dt):
        return torch.ones_like(state.positions) * state.masses.view(-1, 1)

    def no_heat(state, cfg, comp, p, d, nrm, t, dt):
        return torch.full_like(state.internalEnergies, -1.0)

    bc = SimpleNamespace(sdf=sdf, dirichletFunctions={"velocities": zero_v}, forcingFunctions=[push, push],
                         updateFunctions={"dudt": no_heat})
    scheme.boundaryConditions = [bc]
    left = boundary.getPeriodicPositions(st.positions, config.domain)[:, 0] < 0.5
    assert bool(left[0])
    v_before = st.velocities.clone()
    assert boundary.enforceDirichlet(system, 0.3, 0.01, config, scheme)
    assert calls[-1] == (0.3, 0.01)
    assert torch.all(st.velocities[left] == 0) and torch.equal(st.velocities[~left], v_before[~left])
    f = boundary.computeForcing(system, 0.01, 0.3, config, scheme)
    assert torch.allclose(f[left], 2 * st.masses[left].view(-1, 1).expand(-1, 2)) and torch.all(f[~left] == 0)
    upd = SimpleNamespace(dudt=torch.zeros(st.numParticles))
    assert boundary.enforceUpdates(upd, system, 0.01, 0.3, config, scheme)
    assert torch.all(upd.dudt[left] == -1) and torch.all(upd.dudt[~left] == 0)


def test_boundary_conditions_with_an_ensemble_are_refused():
    import numpy as np
    import pytest
    from types import SimpleNamespace
    from warpsph_amd import sample
    from warpsph_amd.ensemble import Runs
    from warpsph_amd.scheme import crkSPH_step
    system, config, scheme = sample.shear_layer_2d(nx=8, device="cpu")
    system.runs = Runs([0, 32, 64], "cpu")
    scheme.boundaryConditions = [SimpleNamespace(sdf=None)]
    with pytest.raises(NotImplementedError):
        crkSPH_step(system, np.array([1e-4, 1e-4]), config, scheme)


def test_a_state_without_particles_is_refused():
    """An empty state raises before any device work."""
    import pytest
    import torch
    from warpsph_amd import sample
    from warpsph_amd.scheme import crkSPH_step
    system, config, scheme = sample.shear_layer_2d(nx=8, device="cpu")
    st = system.state
    for f in ("positions", "velocities", "supports", "masses", "internalEnergies", "densities"):
        v = getattr(st, f)
        if v is not None:
            setattr(st, f, v[:0].clone())
    assert st.numParticles == 0
    with pytest.raises(ValueError, match="no particles"):
        crkSPH_step(system, 1e-4, config, scheme)


def test_parity_bar_is_per_component():
    """The per-step bar holds each column of a vector field to its own scale:
    an error of 1e-4 of a component whose values are 50x smaller than the
    other's fails, although it is 2e-6 of the field's largest value."""
    import numpy as np
    from tests.util import RTOL, relerr
    rng = np.random.default_rng(3)
    ref = np.stack([rng.uniform(-0.5, 0.5, 1000), rng.uniform(-0.01, 0.01, 1000)], axis=1)
    ref[0] = (0.5, 0.01)
    got = ref.copy()
    got[7, 1] += 1e-4 * 0.01
    assert relerr(got, ref) == pytest.approx(1e-4, rel=1e-6)
    assert relerr(got, ref) > RTOL
    # a scalar field and a single column are measured as before
    assert relerr(got[:, 0], ref[:, 0]) == 0.0
    assert relerr(ref[:, 0] + 1e-6, ref[:, 0]) == pytest.approx(2e-6, rel=1e-6)
