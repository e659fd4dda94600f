"""Speculative stages (Runtime.speculate, runtime.begin_speculative): the
adaptive-support pass and the support sweep queued before the host reads
the Verlet tests in front of them give the same bits as blocking tests,
through steps that reuse the list, steps whose stage test fails (a
particle moved past the list's bound) and steps that rebuild after the
supports' test."""
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

FIELDS = ("positions", "velocities", "internalEnergies", "supports", "densities", "pressures", "alphas")


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture
def blocking_after():
    from warpsph_amd.runtime import runtime
    yield
    runtime("cuda:0").speculate = True


def _run(speculate, nsteps, kick, monkeypatch):
    """nsteps RK2 steps of a 96^2 shear layer; with kick, one particle is
    moved past the list's displacement bound before step 3 (its stage test
    then fails and the list is rebuilt).  Returns the fields, the number of
    distinct adjacencies the stages used and the speculative stages'
    outcomes ("fallback": the stage test failed, "rebuild": the supports'
    test failed, "kept")."""
    from warpsph_amd import sample
    from warpsph_amd import scheme as S
    from warpsph_amd.configurations import toCrkConfig
    from warpsph_amd.integrator import getIntegrator
    from warpsph_amd.runtime import runtime
    from warpsph_amd.scheme import crkSPH_step
    rt = runtime("cuda:0")
    rt.speculate = speculate
    outcomes = []
    inner = S.begin_speculative

    def recording(*a, **k):
        r = inner(*a, **k)
        outcomes.append("fallback" if r is None or r[0] is None else ("rebuild" if r[2] else "kept"))
        return r
    monkeypatch.setattr(S, "begin_speculative", recording)
    system, config, scheme = sample.shear_layer_2d(nx=96, device="cuda:0")
    dt = sample.cfl_dt(system.state, config, scheme)
    integ = getIntegrator()
    s = system
    lists = []
    for it in range(nsteps):
        if kick and it == 3:
            adj = s.adjacency
            sv = float(np.float32(toCrkConfig(config, scheme).verletScale))
            bound = 0.5 * (sv - 1.0) * float(adj.hb.min())
            p = 1234
            with torch.no_grad():
                s.state.positions[p, 0] += 1.5 * bound
        r = integ.function(s, crkSPH_step, dt, config, scheme)
        lists += [st.aux[0] for st in r.stages]
        s = r.state
    torch.cuda.synchronize()
    out = {f: getattr(s.state, f).detach().cpu().numpy().copy() for f in FIELDS}
    distinct = len({id(a) for a in lists})
    return out, distinct, outcomes


@pytest.mark.parametrize("kick", [False, True])
def test_speculative_steps_equal_blocking_steps(kick, blocking_after, monkeypatch):
    announce(f"speculative vs blocking, kick={kick}")
    spec, n_spec, outcomes = _run(True, 24, kick, monkeypatch)
    block, n_block, none = _run(False, 24, kick, monkeypatch)
    print({o: outcomes.count(o) for o in set(outcomes)}, flush=True)
    assert not none  # (blocking tests: never speculative)
    # every stage tried to speculate; the first has no list yet (it builds one)
    assert len(outcomes) == 2 * 24 and outcomes[0] == "fallback"
    if kick:
        assert "fallback" in outcomes[1:]
    # the same lists at the same stages: the speculation changes nothing but
    # when the host learns the tests' outcome
    assert n_spec == n_block
    assert n_spec >= 3  # (the shear rebuilds within 24 steps; the kick forces one more)
    for f in FIELDS:
        assert np.array_equal(spec[f], block[f]), f
