"""State, update and system objects of the CRKSPH step.

Field names and meaning follow
  CompSPHState             src/warpSPH/systems/compSPH.py:8-32
  CompressibleSystemUpdate src/warpSPH/systems/compressibleMonaghan.py:31-37
  CompSPHSystem            src/warpSPH/systems/compSPH.py:51-129
The reference builds them on warpSPHIntegrators' BaseState descriptors; here
they are plain dataclasses whose tensors live on the GPU (contiguous FP32,
user particle order).
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass
from typing import Any, List, Optional

import torch

from . import abi


def _clone(v):
    return v.clone() if isinstance(v, torch.Tensor) else v


@dataclass
class CompSPHState:
    positions: torch.Tensor = None
    velocities: torch.Tensor = None
    supports: torch.Tensor = None
    masses: torch.Tensor = None
    densities: torch.Tensor = None
    kinds: torch.Tensor = None
    materials: torch.Tensor = None
    UIDs: torch.Tensor = None
    UIDcounter: int = 0
    internalEnergies: torch.Tensor = None
    totalEnergies: Optional[torch.Tensor] = None
    entropies: Optional[torch.Tensor] = None
    pressures: Optional[torch.Tensor] = None
    soundspeeds: Optional[torch.Tensor] = None
    divergence: Optional[torch.Tensor] = None
    alpha0s: Optional[torch.Tensor] = None
    alphas: Optional[torch.Tensor] = None
    ap_ij: Optional[torch.Tensor] = None
    av_ij: Optional[torch.Tensor] = None
    f_ij: Optional[torch.Tensor] = None

    # fields integrated by the RK driver (integrated(...) in the reference)
    INTEGRATED = ("positions", "velocities", "internalEnergies")

    EDGE_FIELDS = ("ap_ij", "av_ij", "f_ij")

    def initializeNewState(self) -> "CompSPHState":
        """Copy of the per-particle fields; the edge arrays are rebuilt by the
        next derivative evaluation and are not copied."""
        return CompSPHState(**{f.name: (None if f.name in self.EDGE_FIELDS else _clone(getattr(self, f.name)))
                               for f in dataclasses.fields(self)})

    @property
    def numParticles(self) -> int:
        return int(self.positions.shape[0])

    def cabi(self) -> abi.CrkState:
        """Raw pointers for the C-ABI (validates layout)."""
        n = self.numParticles
        s = abi.CrkState()
        s.n = n
        for name in ("positions", "velocities", "supports", "masses", "densities", "internalEnergies",
                     "entropies", "pressures", "soundspeeds", "divergence", "alpha0s", "alphas"):
            t = getattr(self, name)
            if t is not None:
                if t.dtype != torch.float32 or not t.is_contiguous() or t.device.type != 'cuda':
                    raise ValueError(f"CompSPHState.{name} must be a contiguous float32 CUDA tensor "
                                     f"(got {t.dtype}, contiguous={t.is_contiguous()}, device={t.device})")
                if t.shape[0] != n:
                    raise ValueError(f"CompSPHState.{name} has {t.shape[0]} rows, expected {n}")
            setattr(s, name, abi.ptr(t).value if t is not None else None)
        return s


@dataclass
class CompressibleSystemUpdate:
    dxdt: torch.Tensor = None
    dvdt: torch.Tensor = None
    dudt: torch.Tensor = None
    dEdt: torch.Tensor = None
    drhodt: torch.Tensor = None
    passive: Optional[torch.Tensor] = None

    def cabi(self) -> abi.CrkUpdate:
        u = abi.CrkUpdate()
        for k in ("dvdt", "dudt", "drhodt", "dEdt"):
            setattr(u, k, abi.ptr(getattr(self, k)).value)
        return u


class AdjacencyList:
    """CSR adjacency built on the GPU.

    Internally rows are stored in Morton-sorted order (``perm`` maps sorted
    row -> user particle).  The reference's AdjacencyList fields (i, j,
    numNeighbors, neighborOffsets, neighborList -- int64 user indices, see
    src/warpSPH/io.py:8-20 and modules/crk/accel.py:305-306) are produced on
    demand; neighborOffsets[i] indexes the edge arrays (ap_ij, av_ij, f_ij).
    """

    def __init__(self, n, grid, keys, perm, rank, cell_start, cell_end, row_offsets, nbr, xb, hb):
        self.n = int(n)
        self.grid = grid
        self.keys = keys
        self.perm = perm
        self.rank = rank
        self.cell_start = cell_start
        self.cell_end = cell_end
        self.row_offsets = row_offsets
        self.nbr = nbr
        self.xb = xb
        self.hb = hb
        self.n_edges = int(nbr.shape[0])
        self._export = None

    @property
    def numRows(self):
        return self.n

    @property
    def numCols(self):
        return self.n

    def cabi(self) -> abi.CrkAdjacency:
        a = abi.CrkAdjacency()
        a.n = self.n
        a.n_edges = self.n_edges
        a.perm = self.perm.data_ptr()
        a.rank = self.rank.data_ptr()
        a.row_offsets = self.row_offsets.data_ptr()
        a.nbr = self.nbr.data_ptr() if self.n_edges > 0 else None
        return a

    def _expand(self):
        if self._export is None:
            from .runtime import runtime
            dev = self.perm.device
            E = self.n_edges
            i = torch.empty(E, dtype=torch.int64, device=dev)
            j = torch.empty(E, dtype=torch.int64, device=dev)
            nn = torch.empty(self.n, dtype=torch.int64, device=dev)
            no = torch.empty(self.n, dtype=torch.int64, device=dev)
            rt = runtime(dev)
            abi.call("crk_adjacency_export", rt.ctx, self.n, abi.ptr(self.perm), abi.ptr(self.row_offsets),
                     abi.ptr(self.nbr if E > 0 else None), E, abi.ptr(i), abi.ptr(j), abi.ptr(nn), abi.ptr(no),
                     rt.stream())
            self._export = (i, j, nn, no)
        return self._export

    @property
    def i(self):
        return self._expand()[0]

    @property
    def j(self):
        return self._expand()[1]

    @property
    def numNeighbors(self):
        return self._expand()[2]

    @property
    def neighborOffsets(self):
        return self._expand()[3]

    @property
    def neighborList(self):
        return self._expand()[1]

    edgeOffsets = neighborOffsets


@dataclass
class CompSPHSystem:
    state: CompSPHState = None
    adjacency: Optional[AdjacencyList] = None
    domain: Any = None
    t: float = 0.0

    def initializeNewState(self, *args, verbose=False, **kwargs) -> "CompSPHSystem":
        return CompSPHSystem(state=self.state.initializeNewState(), adjacency=self.adjacency, t=self.t,
                             domain=self.domain)

    def finalize(self, initialState, dt, returnValues, updateValues, weights, config, compParams, *args, **kwargs):
        """systems/compSPH.py:87-129 with the compatible-energy update of
        compSPH/multistep.py:52-94 (deterministic row sums on the GPU)."""
        from .configurations import toCrkConfig
        from .runtime import runtime
        import ctypes as C
        self.adjacency = returnValues[-1][0]
        lastState = returnValues[-1][1]
        st0 = initialState.state
        n = st0.numParticles
        if len(returnValues) != 2 or len(weights) != 2:
            raise ValueError("the compatible-energy finalize is built for two-stage (RK2) integration")
        if compParams.compatibleEnergy:
            rt = runtime(st0.positions.device)
            cfg = toCrkConfig(config, compParams)
            (a0, s0), (a1, s1) = returnValues
            e0 = abi.CrkEdges(abi.ptr(s0.ap_ij).value, abi.ptr(s0.av_ij).value, abi.ptr(s0.f_ij).value)
            e1 = abi.CrkEdges(abi.ptr(s1.ap_ij).value, abi.ptr(s1.av_ij).value, abi.ptr(s1.f_ij).value)
            u_new = torch.empty_like(st0.internalEnergies)
            A0, A1 = a0.cabi(), a1.cabi()
            from . import timing
            with timing.scope("compatible_energy"):
                    abi.call("crk_compatible_energy", rt.ctx, C.byref(cfg), n, float(dt), abi.ptr(st0.velocities),
                         abi.ptr(st0.internalEnergies), abi.ptr(updateValues[0].dvdt), abi.ptr(updateValues[1].dvdt),
                         float(weights[0]), float(weights[1]), C.byref(A0), C.byref(e0), C.byref(A1), C.byref(e1),
                         abi.ptr(u_new), abi.ptr(updateValues[-1].dudt), rt.stream())
            self.state.internalEnergies = u_new
        if self.state is not lastState:
            for name in ("supports", "densities", "totalEnergies", "entropies", "pressures", "soundspeeds",
                         "divergence", "alpha0s", "alphas"):
                src = getattr(lastState, name)
                dst = getattr(self.state, name)
                if src is not None and dst is not None:
                    dst.copy_(src)
        return self
