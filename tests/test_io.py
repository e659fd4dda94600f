"""Trajectory export / import with the reference's layout (io.py:8-284),
through the fallback container when h5py is absent: a state (and the edge
data) survive the round trip, and the group/attribute layout is the
reference's."""
import dataclasses

import numpy as np
import torch

from warpsph_amd import io as wio
from warpsph_amd.enumTypes import CompressibleSPHScheme


def _system():
    from warpsph_amd import sample
    system, config, scheme = sample.shear_layer_2d(nx=8, device="cpu")
    system.t = 0.125
    return system


def test_round_trip_of_a_state(tmp_path):
    system = _system()
    path = wio.exportSimulationSystem(system, CompressibleSPHScheme.CRKSPH, str(tmp_path), "step0",
                                      extras={"note": 3, "meta": {"a": torch.arange(4), "b": "x"}})
    back, scheme, f = wio.importSimulationSystem(path, device="cpu")
    assert scheme == "CRKSPH" and back.t == 0.125
    for fld in dataclasses.fields(system.state):
        v = getattr(system.state, fld.name)
        if isinstance(v, torch.Tensor):
            assert torch.equal(getattr(back.state, fld.name), v), fld.name
    assert back.state.UIDcounter == system.state.UIDcounter
    # the reference's groups and attributes
    assert "state" in f.keys() and f.attrs["note"] == 3 and f.attrs["meta"] == "dict"
    assert np.array_equal(f["dict_meta"]["a"], np.arange(4)) and f["dict_meta"].attrs["b"] == "x"
