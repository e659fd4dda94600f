"""The batched data generator on the CPU: run sizes, per-run states and the
ensemble's run table."""
import math

import numpy as np
import torch

from warpsph_amd.ensemble import Runs, concatenate, run_sizes, run_system, sample_ensemble, two_state_run


def test_run_sizes_log_uniform_and_seeded():
    s = run_sizes(4096, 4096, 65536, seed=1)
    assert s.min() >= 64 and s.max() <= 256
    assert np.array_equal(s, run_sizes(4096, 4096, 65536, seed=1))
    # log-uniform N = nx^2: the median of log N sits near the middle of [log 4096, log 65536]
    mid = 0.5 * (math.log(4096) + math.log(65536))
    assert abs(np.median(np.log(s.astype(np.float64) ** 2)) - mid) < 0.1


def test_two_state_run_states_and_split():
    sysr, config, scheme = two_state_run(32, seed=5, device="cpu")
    st = sysr.state
    assert st.positions.shape == (1024, 2)
    rho = st.densities.numpy()
    vals = np.unique(rho)
    assert len(vals) == 2 and vals.min() >= 0.1 and vals.max() <= 1.0
    P = st.pressures.numpy()
    assert np.all((P >= 0.1 - 1e-6) & (P <= 1.0 + 1e-6))
    v = st.velocities.numpy()
    assert np.all(np.abs(v) <= 0.5)
    # the two states are separated by a straight line: a line with
    # y_i (w . x_i + b) >= 1 for every particle exists (linear program)
    from scipy.optimize import linprog
    x = st.positions.numpy().astype(np.float64)
    y = np.where(rho == vals[0], -1.0, 1.0)
    A = -y[:, None] * np.c_[x, np.ones(len(x))]
    res = linprog(np.zeros(3), A_ub=A, b_ub=-np.ones(len(x)), bounds=[(None, None)] * 3, method="highs")
    assert res.status == 0, "the two states are not separated by a line"
    assert bool(config.domain.periodic.all())
    # internal energy from the ideal gas law, u = P / (rho (gamma - 1))
    u = st.internalEnergies.numpy()
    assert np.allclose(u, P / (rho * (scheme.gamma - 1.0)), rtol=1e-6)


def test_concatenate_and_split_round_trip():
    parts = [two_state_run(nx, seed=r, device="cpu") for r, nx in enumerate((16, 20, 24))]
    ens = concatenate([p[0] for p in parts])
    assert list(ens.runs.offsets) == [0, 256, 656, 1232]
    assert np.array_equal(ens.state.materials.numpy(), np.repeat(np.arange(3), [256, 400, 576]))
    for r, (s, _, _) in enumerate(parts):
        back = run_system(ens, r)
        for f in ("positions", "velocities", "supports", "masses", "internalEnergies", "densities"):
            assert torch.equal(getattr(back.state, f), getattr(s.state, f))


def test_runs_table_rejects_bad_offsets():
    import pytest
    for bad in ([1, 5], [0, 5, 5], [0], [0, 3, 2]):
        with pytest.raises(ValueError):
            Runs(bad, "cpu")


def test_sample_ensemble_small():
    ens, config, scheme = sample_ensemble(n_runs=4, nmin=256, nmax=1024, seed=2, device="cpu")
    counts = ens.runs.counts
    assert len(counts) == 4 and counts.min() >= 256 and counts.max() <= 1024
    assert ens.state.positions.shape[0] == counts.sum()
    assert np.all(ens.t == 0)


def test_same_seed_same_problem_at_every_resolution():
    a, _, _ = two_state_run(16, seed=9, device="cpu")
    b, _, _ = two_state_run(32, seed=9, device="cpu")
    for f in ("densities", "pressures", "velocities"):
        va = np.unique(getattr(a.state, f).numpy(), axis=0)
        vb = np.unique(getattr(b.state, f).numpy(), axis=0)
        assert np.array_equal(va, vb), f
