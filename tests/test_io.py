"""Trajectory export / import with the reference's API and layout
(src/warpSPH/io.py:8-312), through the fallback container when h5py is
absent.  The calls are made positionally and by keyword the way
examples/compressible/01-sod-shock-tube-1d.py:136-171,236-248 makes them:
prepExport, then exportSimulationSystem(exportPath, tag, scheme, system,
exportAdjacency, stages, exportStagesAdjacency, extraData), and
importSimulationSystem(path, device, dtype) returns (system, stages, scheme,
extraData) with the adjacency and the stages restored."""
import dataclasses
import json
import os

import numpy as np
import pytest
import torch

from warpsph_amd import io as wio
from warpsph_amd.enumTypes import CompressibleSPHScheme
from warpsph_amd.integrator import StageResult
from warpsph_amd.scheme import buildScheme
from warpsph_amd.systems import AdjacencyList, CompressibleSystemUpdate, CompSPHState, StoredPairs

SCHEME = CompressibleSPHScheme.CRKSPH


def _case(seed=0):
    from warpsph_amd import sample
    system, config, scheme = sample.shear_layer_2d(nx=8, device="cpu")
    system.t = 0.125
    return system, config, scheme


def _export_form(x, h, order):
    """A symmetric list in the reference's export form, rows laid out in
    ``order``: every pair within max(h_i, h_j) (brute force, no periodic
    images: the file format does not care)."""
    n = x.shape[0]
    d = torch.cdist(x.double(), x.double())
    H = torch.maximum(h[:, None], h[None, :]).double()
    rows, cols = [], []
    for u in order.tolist():
        js = torch.nonzero(d[u] < H[u]).flatten()
        rows.append(torch.full((js.numel(),), u, dtype=torch.int64))
        cols.append(js.to(torch.int64))
    i, j = torch.cat(rows), torch.cat(cols)
    nn = torch.bincount(i, minlength=n)
    off = torch.zeros(n, dtype=torch.int64)
    starts = torch.cumsum(nn[order], 0) - nn[order]
    off[order] = starts
    return i, j, nn, off


def _adjacency(system, seed=1):
    st = system.state
    g = torch.Generator().manual_seed(seed)
    order = torch.randperm(st.numParticles, generator=g)
    i, j, nn, off = _export_form(st.positions, st.supports, order)
    return AdjacencyList.fromExport(i, j, nn, off, st.positions, st.supports), order


def _stages(system, adj, seed=2):
    g = torch.Generator().manual_seed(seed)
    n, D = system.state.positions.shape
    E = adj.n_edges
    out = []
    for k in range(2):
        s = system.state.initializeNewState()
        s.pairs = StoredPairs(torch.randn(E, D, generator=g), torch.randn(E, D, generator=g),
                              torch.rand(E, generator=g))
        upd = CompressibleSystemUpdate(dxdt=torch.randn(n, D, generator=g), dvdt=torch.randn(n, D, generator=g),
                                       dudt=torch.randn(n, generator=g), dEdt=torch.randn(n, generator=g),
                                       drhodt=torch.randn(n, generator=g))
        out.append(StageResult(aux=[adj, s], update=upd))
    return out


def _extras(system):
    st = system.state
    ke = 0.5 * (torch.linalg.norm(st.velocities, dim=-1) ** 2 * st.masses).sum()
    te = (st.internalEnergies * st.masses).sum()
    return dict({'kineticEnergy': ke, 'thermalEnergy': te, 'totalEnergy': ke + te, 'frame_num': 7},
                **{'nx': 8, 'note': 'shear', 'meta': {'a': torch.arange(4), 'b': 'x'},
                   'hist': [torch.ones(3), {'c': torch.zeros(2), 'd': 1.5}, 4]})


def _same_state(a, b):
    for f in dataclasses.fields(CompSPHState):
        if f.name == "pairs":
            continue
        va, vb = getattr(a, f.name), getattr(b, f.name)
        if isinstance(va, torch.Tensor):
            assert isinstance(vb, torch.Tensor) and torch.equal(va, vb), f.name
        else:
            assert va == vb, (f.name, va, vb)


def test_fromExport_restores_the_rows(tmp_path):
    system, _, _ = _case()
    adj, order = _adjacency(system)
    st = system.state
    # the rows' order in the edge arrays is the sorted row order
    assert torch.equal(adj.perm.long(), order)
    assert torch.equal(adj.rank.long()[order], torch.arange(adj.n))
    off = adj.row_offsets.long()
    for k in range(adj.n):
        u = int(adj.perm[k])
        js = adj.perm.long()[adj.nbr.long()[off[k]:off[k + 1]]]
        d = torch.linalg.norm(st.positions[js].double() - st.positions[u].double(), dim=-1)
        H = torch.maximum(st.supports[js], st.supports[u]).double()
        assert bool((d < H).all()) and int(off[k + 1] - off[k]) == int((torch.cdist(
            st.positions[u:u + 1].double(), st.positions.double())[0] < torch.maximum(st.supports, st.supports[u])).sum())
    assert torch.equal(adj.xb, st.positions[order]) and torch.equal(adj.hb, st.supports[order])
    # the export arrays come back unchanged
    i, j, nn, off_u = _export_form(st.positions, st.supports, order)
    assert torch.equal(adj.i, i) and torch.equal(adj.j, j) and torch.equal(adj.numNeighbors, nn)
    assert torch.equal(adj.edgeOffsets, off_u)


def test_fromExport_rejects_inconsistent_arrays():
    system, _, _ = _case()
    st = system.state
    order = torch.arange(st.numParticles)
    i, j, nn, off = _export_form(st.positions, st.supports, order)
    bad_i = i.clone()
    bad_i[3] = (bad_i[3] + 1) % st.numParticles
    with pytest.raises(ValueError):
        AdjacencyList.fromExport(bad_i, j, nn, off, st.positions, st.supports)
    bad_off = off.clone()
    bad_off[1] += 1
    with pytest.raises(ValueError):
        AdjacencyList.fromExport(i, j, nn, bad_off, st.positions, st.supports)
    bad_j = j.clone()
    bad_j[0] = st.numParticles
    with pytest.raises(ValueError):
        AdjacencyList.fromExport(i, bad_j, nn, off, st.positions, st.supports)
    with pytest.raises(ValueError):
        AdjacencyList.fromExport(i[:-1], j, nn, off, st.positions, st.supports)


def test_fromExport_places_empty_rows():
    """A row without edges shares its offset with the next row's block."""
    x = torch.tensor([[0.0, 0.0], [0.1, 0.0], [5.0, 5.0], [0.2, 0.0]])
    h = torch.tensor([0.15, 0.15, 0.15, 0.15])
    # rows laid out 3, 2 (empty), 0, 1
    i = torch.tensor([3, 3, 0, 0, 1, 1, 1])
    j = torch.tensor([3, 1, 0, 1, 0, 1, 3])
    nn = torch.tensor([2, 3, 0, 2])
    off = torch.tensor([2, 4, 2, 0])
    adj = AdjacencyList.fromExport(i, j, nn, off, x, h)
    assert adj.perm.tolist() == [3, 2, 0, 1]
    assert adj.row_offsets.tolist() == [0, 2, 2, 4, 7]
    assert adj.perm[adj.nbr.long()].tolist() == j.tolist()


def test_export_import_round_trip(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    system, config, schemeConfig = _case()
    adj, _ = _adjacency(system)
    system.adjacency = adj
    stages = _stages(system, adj)
    SimulationSystem, SimulationState, SimulationConfig, SimulationUpdate, fn, export_fn, import_fn = \
        buildScheme(SCHEME)
    exportPath = wio.prepExport('shear', config, schemeConfig, SCHEME, export_fn)
    assert exportPath == 'export/shear'
    extras = _extras(system)
    # as the example stores the initial state (keywords) and later frames
    p0 = wio.exportSimulationSystem(exportPath, 'initialState', SCHEME, system, exportAdjacency=False, stages=None,
                                    exportStagesAdjacency=False, extraData=extras)
    p1 = wio.exportSimulationSystem(exportPath, f'state_{7:04d}', SCHEME, system, False, stages, True, extras)
    assert os.path.dirname(p1) == os.path.join(exportPath, 'trajectory')

    back, st_back, schemeEnum, extra = wio.importSimulationSystem(p1, 'cpu', torch.float32)
    assert schemeEnum is SCHEME and back.t == 0.125 and isinstance(back, SimulationSystem)
    _same_state(system.state, back.state)
    # the adjacency (written whenever the system has one) with its CSR
    a2 = back.adjacency
    for k in ('i', 'j', 'numNeighbors', 'edgeOffsets'):
        assert torch.equal(getattr(a2, k), getattr(adj, k)), k
    for k in ('perm', 'rank', 'row_offsets', 'nbr', 'xb', 'hb'):
        assert torch.equal(getattr(a2, k), getattr(adj, k)), k
    # both stages: adjacency, state with its pair data, update
    assert len(st_back) == 2
    for s0, s1 in zip(stages, st_back):
        assert isinstance(s1, StageResult) and isinstance(s1.aux[0], AdjacencyList)
        assert torch.equal(s1.aux[0].j, adj.j)
        assert isinstance(s1.aux[1], SimulationState)
        _same_state(s0.aux[1], s1.aux[1])
        for x0, x1 in zip(s0.aux[1].pairs.full(), s1.aux[1].pairs.full()):
            assert torch.equal(x0, x1)
        assert isinstance(s1.update, SimulationUpdate)
        for f in ('dxdt', 'dvdt', 'dudt', 'dEdt', 'drhodt'):
            assert torch.equal(getattr(s0.update, f), getattr(s1.update, f)), f
    # extras: tensors back as tensors, dictionaries and lists rebuilt
    for k in ('kineticEnergy', 'thermalEnergy', 'totalEnergy'):
        assert torch.equal(extra[k], extras[k]), k
    assert extra['frame_num'] == 7 and extra['nx'] == 8 and extra['note'] == 'shear'
    assert torch.equal(extra['meta']['a'], torch.arange(4)) and extra['meta']['b'] == 'x'
    h = extra['hist']
    assert torch.equal(h[0], torch.ones(3)) and torch.equal(h[1]['c'], torch.zeros(2)) and h[1]['d'] == 1.5
    assert h[2] == 4

    # the initial state: no stages; the adjacency is still written (the
    # reference writes it whenever the system has one)
    b0, s0b, _, e0 = wio.importSimulationSystem(p0, 'cpu', torch.float32)
    assert s0b == [] and b0.adjacency is not None and e0['frame_num'] == 7

    # stages without their adjacency
    p2 = wio.exportSimulationSystem(exportPath, 'noadj', SCHEME, system, False, stages, False, None)
    _, st2, _, e2 = wio.importSimulationSystem(p2, 'cpu', torch.float32)
    assert e2 == {} and [type(a) for a in st2[0].aux] == [SimulationState]

    # a re-export of what was imported is the same file
    p3 = wio.exportSimulationSystem(exportPath, 'again', SCHEME, back, False, st_back, True, extra)
    z1, z3 = np.load(p1), np.load(p3)
    assert sorted(z1.files) == sorted(z3.files)
    for k in z1.files:
        if k != '__attrs__':
            assert np.array_equal(z1[k], z3[k]), k
    m1, m3 = (json.loads(bytes(z['__attrs__']).decode()) for z in (z1, z3))
    assert m1 == m3


def test_prepExport_and_importConfigs(tmp_path, monkeypatch):
    from warpsph_amd.configurations import configurationToDict, crkSPHConfigToDict
    monkeypatch.chdir(tmp_path)
    system, config, schemeConfig = _case()
    schemeConfig.gamma = 1.4
    _, _, _, _, _, export_fn, import_fn = buildScheme(SCHEME)
    path = wio.prepExport('case_a', config, schemeConfig, SCHEME, export_fn)
    with open(f'{path}/config.json') as fh:
        d = json.load(fh)
    assert d['scheme'] == 'CRKSPH' and set(d) == {'scheme', 'config', 'schemeConfig', 'timestamp'}
    assert d['config']['kernel'] == config.kernel.name and d['config']['domain']['dim'] == config.domain.dim
    cfg2, sch2 = wio.importConfigs(f'{path}/config.json', import_fn)
    assert configurationToDict(cfg2) == configurationToDict(config)
    assert crkSPHConfigToDict(sch2) == crkSPHConfigToDict(schemeConfig)


def test_import_needs_a_known_scheme():
    with pytest.raises(ValueError):
        wio.schemeNameToSimulationScheme('NoSuchScheme')
    assert wio.schemeNameToSimulationScheme('crksph') is SCHEME
