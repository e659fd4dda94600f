"""Trajectory export and import with the reference's layout and API
(src/warpSPH/io.py:8-307).

File layout (one file per call, ``{exportPath}/trajectory/{tag}.h5``):

  attrs: scheme (name), time, scalar extras; 'dict' / 'list' for the keys of
         dictionary and list extras
  adjacency/   i, j, numNeighbors, edgeOffsets, queryPositions, querySupports,
               referencePositions, referenceSupports; attrs numRows, numCols
  state/       one dataset per tensor field of the state (ap_ij, av_ij, f_ij
               over the adjacency edges); attrs for scalars, 'None' for
               absent fields
  stages/stage_k/  attrs index; aux/{a}_adjacency, aux/{a}_state (as above);
               one dataset per tensor field of the update
  dict_{key}/  dictionary and list extras (datasets, attrs, subgroups)
  {key}        tensor extras

With h5py installed the file is HDF5, loadable by the reference's
importSimulationSystem.  Without it (h5py is not part of this image) the same
hierarchy is written as one uncompressed .npz archive (``{tag}.h5.npz``):
datasets under their slash-separated paths, attributes as JSON under
``__attrs__``, read back without unpickling.  importSimulationSystem takes
either path.
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
    """The part of the h5py Group interface the layout uses (children listed
    in name order, as h5py lists them)."""

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

    def __contains__(self, name: str) -> bool:
        p = self._p(name)
        return p in self._t.data or p in self._t.groups

    def keys(self):
        pre = f"{self._path}/" if self._path else ""
        names = [p[len(pre):] for p in list(self._t.data) + list(self._t.groups) if p.startswith(pre)]
        return sorted(q for q in names if "/" not in q)

    def items(self):
        return [(k, self[k]) for k in self.keys()]

    def close(self):
        pass


def _json(v):
    if isinstance(v, np.generic):
        return v.item()
    if isinstance(v, np.ndarray):
        return v.tolist()
    if isinstance(v, torch.Tensor):
        return v.detach().cpu().tolist()
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


def _open_write(path: str):
    return h5py.File(path, 'w') if h5py is not None else _Group(_Tree())


def _open_read(path: str):
    if os.path.exists(path) and not path.endswith('.npz'):
        if h5py is None:
            raise RuntimeError(f"{path} is an HDF5 file and h5py is not installed")
        return h5py.File(path, 'r')
    return _load_tree(path if path.endswith('.npz') else path + '.npz')


def _is_dataset(v) -> bool:
    return isinstance(v, np.ndarray) or (h5py is not None and isinstance(v, h5py.Dataset))


def _np(t):
    return t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


def _tensor(v, device):
    """A dataset as a tensor of its stored dtype (0-d datasets as 0-d tensors),
    as hdfDtypeToTorchDtype keeps it (io.py:128-142)."""
    a = np.asarray(v[()])
    return torch.from_numpy(np.array(a)).to(device)


def _attr(v):
    return v.item() if isinstance(v, np.generic) else v


# ------------------------------------------------------------ export
def dumpAdjacency(adjacency, adjacencyGroup):
    """io.py:8-20.  The list is symmetric: query and reference positions
    (supports) are the ones it was built at, in user order."""
    adjacencyGroup.create_dataset('i', data=_np(adjacency.i))
    adjacencyGroup.create_dataset('j', data=_np(adjacency.j))
    adjacencyGroup.create_dataset('numNeighbors', data=_np(adjacency.numNeighbors))
    adjacencyGroup.create_dataset('edgeOffsets', data=_np(adjacency.edgeOffsets))
    perm = adjacency.perm.long()
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
    """io.py:22-33: tensors as datasets, scalars as attrs, absent fields as
    'None'; the pair data as the datasets ap_ij, av_ij, f_ij over the
    adjacency edges."""
    import dataclasses
    for f in dataclasses.fields(state):
        name = f.name
        if name.startswith('_'):
            continue
        if name == "pairs":
            for k in ("ap_ij", "av_ij", "f_ij"):
                v = getattr(state, k)
                if v is None:
                    stateGroup.attrs[k] = 'None'
                else:
                    stateGroup.create_dataset(k, data=_np(v))
            continue
        v = getattr(state, name)
        if isinstance(v, torch.Tensor):
            stateGroup.create_dataset(name, data=_np(v))
        elif v is None:
            stateGroup.attrs[name] = 'None'
        else:
            stateGroup.attrs[name] = v


def dumpStage(stage, stageGroups, index, SimulationState, exportStagesAdjacency):
    """io.py:35-57: aux entries (adjacency lists, states) and the update's
    tensor fields of stage ``index``."""
    from .systems import AdjacencyList
    import dataclasses
    g = stageGroups[index]
    g.attrs['index'] = index
    aux = g.create_group('aux')
    for ai, a in enumerate(stage.aux):
        if isinstance(a, AdjacencyList):
            if exportStagesAdjacency:
                dumpAdjacency(a, aux.create_group(f'{ai}_adjacency'))
        elif isinstance(a, SimulationState):
            dumpState(a, aux.create_group(f'{ai}_state'))
        else:
            print(f'Unsupported aux type: {type(a)}')
    for f in dataclasses.fields(stage.update):
        v = getattr(stage.update, f.name)
        if isinstance(v, torch.Tensor):
            g.create_dataset(f.name, data=_np(v))


def _scheme_name(scheme):
    return scheme.name if hasattr(scheme, 'name') else scheme


def _dump_extra(root, key, value):
    """io.py:92-125: dictionaries and lists as groups dict_{key} (marked by a
    'dict' / 'list' attribute of the key), tensors as datasets, the rest as
    attributes."""
    if isinstance(value, dict):
        eg = root.create_group(f'dict_{key}')
        for sk, sv in value.items():
            if isinstance(sv, torch.Tensor):
                eg.create_dataset(sk, data=_np(sv))
            else:
                eg.attrs[sk] = sv
        root.attrs[key] = 'dict'
    elif isinstance(value, list):
        eg = root.create_group(f'dict_{key}')
        for si, sv in enumerate(value):
            if isinstance(sv, torch.Tensor):
                eg.create_dataset(f'{si}', data=_np(sv))
            elif isinstance(sv, dict):
                sub = eg.create_group(f'{si}')
                for ssk, ssv in sv.items():
                    if isinstance(ssv, torch.Tensor):
                        sub.create_dataset(ssk, data=_np(ssv))
                    else:
                        sub.attrs[ssk] = ssv
            else:
                eg.attrs[f'{si}'] = sv
        root.attrs[key] = 'list'
    elif isinstance(value, torch.Tensor):
        root.create_dataset(key, data=_np(value))
    else:
        root.attrs[key] = value


def exportSimulationSystem(exportPath, tag, scheme, system: Any, exportAdjacency=False, stages=None,
                           exportStagesAdjacency=True, extraData=None):
    """io.py:63-127: write ``{exportPath}/trajectory/{tag}.h5``.

    As in the reference, the system's adjacency is written whenever it has
    one (``exportAdjacency`` is accepted and does not change that), each
    stage's adjacency when ``exportStagesAdjacency``.  Returns the path
    written (``{tag}.h5.npz`` without h5py)."""
    os.makedirs(f'{exportPath}/trajectory/', exist_ok=True)
    path = f'{exportPath}/trajectory/{tag}.h5'
    root = _open_write(path)
    root.attrs['scheme'] = _scheme_name(scheme)
    t = system.t
    if isinstance(t, torch.Tensor):
        t = t.detach().cpu()
        t = t.item() if t.ndim == 0 else t.numpy()
    root.attrs['time'] = float(t) if np.ndim(t) == 0 else np.asarray(t, dtype=np.float64)
    if system.adjacency is not None:
        dumpAdjacency(system.adjacency, root.create_group('adjacency'))
    dumpState(system.state, root.create_group('state'))
    if stages is not None:
        sg = root.create_group('stages')
        groups = [sg.create_group(f'stage_{k}') for k in range(len(stages))]
        for k in range(len(stages)):
            dumpStage(stages[k], groups, k, type(system.state), exportStagesAdjacency)
    for key, value in (extraData.items() if extraData is not None else ()):
        _dump_extra(root, key, value)
    if h5py is not None:
        root.close()
        return path
    _save_tree(root, path + '.npz')
    return path + '.npz'


# ------------------------------------------------------------ import
def loadAdjacency(adjacencyGroup, device):
    """io.py:145-160: the adjacency of the file, with the CSR the passes use
    (AdjacencyList.fromExport)."""
    from .systems import AdjacencyList
    g = {k: _tensor(adjacencyGroup[k], device) for k in
         ('i', 'j', 'numNeighbors', 'edgeOffsets', 'referencePositions', 'referenceSupports')}
    return AdjacencyList.fromExport(g['i'], g['j'], g['numNeighbors'], g['edgeOffsets'], g['referencePositions'],
                                    g['referenceSupports'])


def loadState(stateGroup, device, SimulationState):
    """io.py:162-170: the state's datasets (and its scalar attributes, where
    the state has such a field); ap_ij, av_ij and f_ij become the state's
    pair data."""
    import dataclasses
    from .systems import StoredPairs
    names = {f.name for f in dataclasses.fields(SimulationState)}
    kw, pair = {}, {}
    for k, v in stateGroup.items():
        if not _is_dataset(v):
            continue
        if k in ('ap_ij', 'av_ij', 'f_ij') and 'pairs' in names:
            pair[k] = _tensor(v, device)
        elif k in names:
            kw[k] = _tensor(v, device)
    for k, v in stateGroup.attrs.items():
        if k in names and k not in kw and not (isinstance(v, str) and v == 'None'):
            kw[k] = _attr(v)
    if pair:
        kw['pairs'] = StoredPairs(**pair)
    return SimulationState(**kw)


def loadStage(stageGroup, device, SimulationState, SimulationUpdate):
    """io.py:173-191: StageResult(aux=[adjacency, state], update)."""
    from .integrator import StageResult
    aux = []
    for k, v in stageGroup['aux'].items():
        if '_adjacency' in k:
            aux.append(loadAdjacency(v, device))
        elif '_state' in k:
            aux.append(loadState(v, device, SimulationState))
        else:
            print(f'Unsupported aux type for key {k}')
    upd = {k: _tensor(v, device) for k, v in stageGroup.items() if k != 'aux' and _is_dataset(v)}
    return StageResult(aux=aux, update=SimulationUpdate(**upd))


def schemeNameToSimulationScheme(name: str):
    """io.py:194-203 (the compressible schemes)."""
    from .enumTypes import CompressibleSPHScheme
    for s in CompressibleSPHScheme:
        if s.name.lower() == name.lower():
            return s
    raise ValueError(f'Unsupported scheme name: {name}')


def _load_extra(group, value, device):
    if _is_dataset(value):
        return _tensor(value, device)
    if isinstance(value, _Group) or (h5py is not None and isinstance(value, h5py.Group)):
        return {k: _load_extra(group, v, device) for k, v in value.items()}
    return _attr(value)


def importSimulationSystem(importPath, device, dtype, SimulationSystem: Optional[Any] = None,
                           SimulationState: Optional[Any] = None, SimulationUpdate: Optional[Any] = None):
    """io.py:208-270.  Returns (system, stages, schemeEnum, extraData): the
    system with its state and adjacency (None when the file has none), the
    stages as StageResult(aux, update), the scheme, and the extras (tensor
    extras as tensors, dictionaries and lists rebuilt).  Tensors keep their
    stored dtypes, as in the reference (``dtype`` is accepted for its
    signature)."""
    from .scheme import buildScheme
    f = _open_read(importPath)
    try:
        t = _attr(f.attrs['time'])
        schemeEnum = schemeNameToSimulationScheme(str(f.attrs['scheme']))
        if SimulationSystem is None or SimulationState is None or SimulationUpdate is None:
            SimulationSystem, SimulationState, _, SimulationUpdate, _, _, _ = buildScheme(schemeEnum)
        adjacency = loadAdjacency(f['adjacency'], device) if 'adjacency' in f else None
        state = loadState(f['state'], device, SimulationState)
        stages = []
        if 'stages' in f:
            sg = f['stages']
            for k in sorted(sg.keys(), key=lambda s: int(s.split('_')[-1])):
                stages.append(loadStage(sg[k], device, SimulationState, SimulationUpdate))
        extraData = {}
        for k, v in f.items():
            if k in ('adjacency', 'state', 'stages') or k.startswith('dict_'):
                continue
            if _is_dataset(v):
                extraData[k] = _tensor(v, device)
        for k, v in f.attrs.items():
            if k in ('scheme', 'time'):
                continue
            v = _attr(v)
            if isinstance(v, str) and v in ('dict', 'list'):
                grp = f[f'dict_{k}']
                if v == 'dict':
                    d = {sk: _load_extra(grp, sv, device) for sk, sv in grp.items()}
                    d.update({sk: _attr(sv) for sk, sv in grp.attrs.items()})
                    extraData[k] = d
                else:
                    items = {sk: _load_extra(grp, sv, device) for sk, sv in grp.items()}
                    items.update({sk: _attr(sv) for sk, sv in grp.attrs.items()})
                    for sk, sv in list(items.items()):
                        if isinstance(sv, dict):
                            sub = grp[sk]
                            sv.update({a: _attr(b) for a, b in sub.attrs.items()})
                    extraData[k] = [items[sk] for sk in sorted(items, key=int)]
            else:
                extraData[k] = v
    finally:
        f.close()
    if isinstance(t, (list, np.ndarray)):
        t = np.asarray(t, dtype=np.float64)
    return SimulationSystem(state=state, adjacency=adjacency, t=t), stages, schemeEnum, extraData


# ------------------------------------------------------------ case configuration
def prepExport(caseName, config, schemeConfig, scheme, export_fn):
    """io.py:280-303: write export/{caseName}/config.json (scheme name, the
    simulation configuration, the scheme's parameters through ``export_fn``,
    a timestamp) and return the case's export path."""
    from .configurations import configurationToDict
    from .utils import getCurrentTimestamp
    exportDict = {
        'scheme': _scheme_name(scheme),
        'config': configurationToDict(config),
        'schemeConfig': export_fn(schemeConfig),
        'timestamp': getCurrentTimestamp(),
    }
    exportPath = f'export/{caseName}'
    os.makedirs(exportPath, exist_ok=True)
    with open(f'{exportPath}/config.json', 'w') as fh:
        json.dump(exportDict, fh, indent=4)
    return exportPath


def importConfigs(configPath, import_fn):
    """io.py:307-312: (SimulationConfig, scheme parameters) of a config.json."""
    from .configurations import dictToConfig
    with open(configPath, 'r') as fh:
        d = json.load(fh)
    return dictToConfig(d['config']), import_fn(d['schemeConfig'])
