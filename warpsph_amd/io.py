"""Trajectory export and import with the reference's layout (io.py:8-284).

File layout (one file per call, ``{exportPath}/trajectory/{tag}.h5``):

  attrs: scheme (name), time, plus scalar extras
  adjacency/   i, j, numNeighbors, edgeOffsets, queryPositions, querySupports,
               referencePositions, referenceSupports; attrs numRows, numCols
  state/       one dataset per tensor field of the state; attrs for scalars
               (and the string 'None' for absent fields)
  stages/stage_k/  attrs index; aux/{a}_adjacency, aux/{a}_state (as above);
               one dataset per update field
  dict_{key}/  extra dictionaries (datasets and attrs)

With h5py installed the file is HDF5, loadable by the reference's
importSimulationSystem.  Without it (h5py is not part of this image) the same
hierarchy is written as one uncompressed .npz archive (``{tag}.h5.npz``):
datasets under their slash-separated paths, attributes as JSON under
``__attrs__``.  The API is the reference's either way.
"""
from __future__ import annotations

import json
import os
from typing import Any, Dict, Optional

import numpy as np
import torch

try:  # optional: the image may not have it
    import h5py  # type: ignore
except ImportError:  # pragma: no cover - depends on the environment
    h5py = None


# ------------------------------------------------------------ a minimal tree
class _Tree:
    """Datasets, groups and attributes of one file, keyed by path."""

    def __init__(self):
        self.data: Dict[str, np.ndarray] = {}
        self.groups = set()
        self.attr_table: Dict[str, Dict[str, Any]] = {"": {}}


class _Group:
    """The part of the h5py Group interface the layout uses."""

    def __init__(self, tree: _Tree, path: str = ""):
        self._t, self._path = tree, path
        self.attrs = tree.attr_table.setdefault(path, {})

    def _p(self, name: str) -> str:
        return f"{self._path}/{name}" if self._path else name

    def create_group(self, name: str) -> "_Group":
        self._t.groups.add(self._p(name))
        return _Group(self._t, self._p(name))

    def create_dataset(self, name: str, data) -> None:
        self._t.data[self._p(name)] = np.asarray(data)

    def __getitem__(self, name: str):
        p = self._p(name)
        if p in self._t.data:
            return self._t.data[p]
        if p in self._t.groups:
            return _Group(self._t, p)
        raise KeyError(name)

    def keys(self):
        pre = f"{self._path}/" if self._path else ""
        names = [p[len(pre):] for p in list(self._t.data) + sorted(self._t.groups) if p.startswith(pre)]
        return [q for q in names if "/" not in q]


def _json(v):
    if isinstance(v, np.integer):
        return int(v)
    if isinstance(v, np.floating):
        return float(v)
    if isinstance(v, np.ndarray):
        return v.tolist()
    return v


def _save_tree(root: _Group, path: str):
    t = root._t
    arrays = dict(t.data)
    meta = {"attrs": {k: {a: _json(b) for a, b in v.items()} for k, v in t.attr_table.items()},
            "groups": sorted(t.groups)}
    arrays["__attrs__"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez(path, **arrays)


def _load_tree(path: str) -> _Group:
    t = _Tree()
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(bytes(z["__attrs__"]).decode())
        for k in z.files:
            if k != "__attrs__":
                t.data[k] = z[k]
    t.groups = set(meta["groups"])
    t.attr_table = meta["attrs"]
    t.attr_table.setdefault("", {})
    return _Group(t, "")


# ------------------------------------------------------------ the layout
def _np(t):
    return t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


def dumpAdjacency(adjacency, adjacencyGroup):
    """io.py:8-20."""
    adjacencyGroup.create_dataset('i', data=_np(adjacency.i))
    adjacencyGroup.create_dataset('j', data=_np(adjacency.j))
    adjacencyGroup.create_dataset('numNeighbors', data=_np(adjacency.numNeighbors))
    adjacencyGroup.create_dataset('edgeOffsets', data=_np(adjacency.edgeOffsets))
    # positions and supports the list was built at, in user order
    perm = adjacency.perm.long()
    D = adjacency.xb.shape[1]
    xu = torch.empty_like(adjacency.xb)
    hu = torch.empty_like(adjacency.hb)
    xu[perm] = adjacency.xb
    hu[perm] = adjacency.hb
    for k, v in (('queryPositions', xu), ('querySupports', hu), ('referencePositions', xu),
                 ('referenceSupports', hu)):
        adjacencyGroup.create_dataset(k, data=_np(v))
    adjacencyGroup.attrs['numRows'] = adjacency.numRows
    adjacencyGroup.attrs['numCols'] = adjacency.numCols


def dumpState(state, stateGroup):
    """io.py:22-33: tensors as datasets, scalars as attrs, absent fields as 'None'
    (the pair data ap_ij, av_ij, f_ij as datasets over the adjacency edges)."""
    import dataclasses
    for f in dataclasses.fields(state):
        name = f.name
        if name == "pairs":
            if state.pairs is not None:
                for k, v in (("ap_ij", state.ap_ij), ("av_ij", state.av_ij), ("f_ij", state.f_ij)):
                    stateGroup.create_dataset(k, data=_np(v))
            continue
        v = getattr(state, name)
        if isinstance(v, torch.Tensor):
            stateGroup.create_dataset(name, data=_np(v))
        elif v is None:
            stateGroup.attrs[name] = 'None'
        else:
            stateGroup.attrs[name] = v


def dumpStage(stage, stageGroups, index, exportStagesAdjacency=True):
    """io.py:35-57."""
    g = stageGroups[index]
    g.attrs['index'] = index
    aux = g.create_group('aux')
    for ai, a in enumerate(stage.aux):
        if hasattr(a, 'perm'):
            if exportStagesAdjacency:
                dumpAdjacency(a, aux.create_group(f'{ai}_adjacency'))
        else:
            dumpState(a, aux.create_group(f'{ai}_state'))
    import dataclasses
    for f in dataclasses.fields(stage.update):
        v = getattr(stage.update, f.name)
        if isinstance(v, torch.Tensor):
            g.create_dataset(f.name, data=_np(v))


def exportSimulationSystem(system, scheme, exportPath: str, tag: str, stages=None, exportAdjacency: bool = True,
                           exportStagesAdjacency: bool = True, extras: Optional[Dict[str, Any]] = None) -> str:
    """io.py:63-127.  Returns the path written."""
    os.makedirs(f'{exportPath}/trajectory', exist_ok=True)
    path = f'{exportPath}/trajectory/{tag}.h5'
    root = h5py.File(path, 'w') if h5py is not None else _Group(_Tree())
    root.attrs['scheme'] = scheme.name if hasattr(scheme, 'name') else str(scheme)
    t = system.t
    root.attrs['time'] = float(t) if np.ndim(t) == 0 else np.asarray(t, dtype=np.float64)
    if exportAdjacency and system.adjacency is not None:
        dumpAdjacency(system.adjacency, root.create_group('adjacency'))
    dumpState(system.state, root.create_group('state'))
    if stages:
        sg = root.create_group('stages')
        groups = [sg.create_group(f'stage_{k}') for k in range(len(stages))]
        for k, st in enumerate(stages):
            dumpStage(st, groups, k, exportStagesAdjacency)
    for key, value in (extras or {}).items():
        if isinstance(value, dict):
            eg = root.create_group(f'dict_{key}')
            for sk, sv in value.items():
                if isinstance(sv, torch.Tensor):
                    eg.create_dataset(sk, data=_np(sv))
                else:
                    eg.attrs[sk] = sv
            root.attrs[key] = 'dict'
        elif isinstance(value, torch.Tensor):
            root.create_dataset(key, data=_np(value))
        else:
            root.attrs[key] = value
    if h5py is not None:
        root.close()
        return path
    _save_tree(root, path + '.npz')
    return path + '.npz'


def _open(path: str):
    if h5py is not None and not path.endswith('.npz'):
        return h5py.File(path, 'r')
    return _load_tree(path if path.endswith('.npz') else path + '.npz')


def loadState(stateGroup, device):
    """io.py:163-172: the state's tensor fields (attrs: scalars)."""
    from .systems import CompSPHState
    import dataclasses
    kw = {}
    for f in dataclasses.fields(CompSPHState):
        if f.name == "pairs":
            continue
        if f.name in stateGroup.keys():
            kw[f.name] = torch.as_tensor(np.asarray(stateGroup[f.name])).to(device)
        elif f.name in stateGroup.attrs and stateGroup.attrs[f.name] != 'None':
            kw[f.name] = stateGroup.attrs[f.name]
    return CompSPHState(**kw)


def importSimulationSystem(importPath: str, device='cuda'):
    """io.py:208-284 (state, time, scheme name; the adjacency is rebuilt by
    the next step).  Returns (system, schemeName, file handle)."""
    from .systems import CompSPHSystem
    f = _open(importPath)
    t = f.attrs['time']
    scheme = f.attrs['scheme']
    state = loadState(f['state'], device)
    system = CompSPHSystem(state=state, adjacency=None, t=t if np.ndim(t) else float(t))
    return system, scheme, f
