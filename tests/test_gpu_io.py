"""Export after a GPU step (io.py:63-127 layout): the adjacency in the
reference's user-index form, the state with its pair data, both stages."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_export_after_a_step(tmp_path):
    from warpsph_amd import io as wio
    from warpsph_amd import sample
    from warpsph_amd.enumTypes import CompressibleSPHScheme
    from warpsph_amd.integrator import getIntegrator
    from warpsph_amd.scheme import crkSPH_step
    print("[gpu-case] export after a step", flush=True)
    system, config, scheme = sample.shear_layer_2d(nx=24, device="cuda")
    dt = sample.cfl_dt(system.state, config, scheme)
    res = getIntegrator().function(system, crkSPH_step, dt, config, scheme)
    s = res.state
    path = wio.exportSimulationSystem(s, CompressibleSPHScheme.CRKSPH, str(tmp_path), "t1", stages=res.stages)
    back, name, f = wio.importSimulationSystem(path, device="cuda")
    assert name == "CRKSPH" and abs(back.t - s.t) < 1e-12
    assert torch.equal(back.state.positions, s.state.positions)
    adj = f["adjacency"]
    E = int(np.asarray(adj["numNeighbors"]).sum())
    assert np.asarray(adj["i"]).shape == (E,) and adj.attrs["numRows"] == s.state.numParticles
    st1 = f["stages"]["stage_1"]
    assert st1.attrs["index"] == 1 and "dvdt" in st1.keys()
    aux_state = st1["aux"]["1_state"]
    assert np.asarray(aux_state["ap_ij"]).shape[0] == np.asarray(st1["aux"]["0_adjacency"]["j"]).shape[0]
