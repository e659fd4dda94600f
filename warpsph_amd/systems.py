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

import ctypes as C
import dataclasses
from dataclasses import dataclass
from typing import Any, List, Optional

import torch

from . import abi


def _clone(v):
    return v.clone() if isinstance(v, torch.Tensor) else v


def capacity(count: int) -> int:
    """Allocation size for a per-edge array: rounded up to a multiple of 1/16
    of the largest power of two below it (at least 2^20 entries), so that the
    edge arrays of successive steps reuse the same cached device blocks while
    the adjacency size drifts by a few percent."""
    count = int(count)
    if count <= (1 << 20):
        return max(1, count)
    g = max(1 << 20, (1 << (count.bit_length() - 1)) // 16)
    return ((count + g - 1) // g) * g


class StagePairs:
    """The interacting pairs of one derivative evaluation and their data.

    ``count[k]`` pairs of sorted row k (|x_ij| < max(h_i, h_j)) fill the
    slots ``row_offsets[k] ..`` of ``nbr``; ``ap``, ``av`` and ``f`` use the
    same slots.  The reference keeps ap_ij, av_ij and f_ij over every edge of
    its Verlet list (src/warpSPH/systems/compSPH.py:30-32), with zeros on the
    pairs out of support; ``full()`` produces that layout on demand.
    """

    def __init__(self, adjacency, dim: int, masses: torch.Tensor):
        dev = adjacency.perm.device
        E = capacity(adjacency.n_edges)
        self.adjacency = adjacency
        self.dim = int(dim)
        self.masses = masses
        self.dt = 0.0  # set by the derivative evaluation (energy partition)
        # an ensemble's run table and per-particle dt of the stage (or None)
        self.runs = None
        self.dtp = None
        self.count = torch.empty(adjacency.n, dtype=torch.int32, device=dev)
        self.nbr = torch.empty(E, dtype=torch.int32, device=dev)
        # pair accelerations per slot: a_ij = ap_ij + av_ij, and the viscous
        # fraction th (av_ij = th a_ij)
        self.pf = torch.empty((E, dim), dtype=torch.float32, device=dev)
        self.th = torch.empty(E, dtype=torch.float32, device=dev)
        self.rec = torch.empty((adjacency.n, 4), dtype=torch.float32, device=dev)
        # tile-local neighbour index of each pair (adjacencies with LDS tiles)
        self.loc = torch.empty(E, dtype=torch.int16, device=dev) if adjacency.tiles is not None else None
        self._full = None

    def cabi(self) -> abi.CrkEdges:
        """Pointers for the C-ABI; f_ij stays implicit (evaluated from rec)."""
        return abi.CrkEdges(*(abi.ptr(t).value for t in (self.count, self.nbr, self.pf, self.rec)), None,
                            abi.ptr(self.loc).value, abi.ptr(self.th).value)

    @property
    def n_pairs(self) -> int:
        return int(self.count.sum().item())

    def full(self):
        if self._full is None:
            from .runtime import runtime
            a = self.adjacency
            E = a.n_edges
            dev = self.count.device
            ap = torch.empty((E, self.dim), dtype=torch.float32, device=dev)
            av = torch.empty((E, self.dim), dtype=torch.float32, device=dev)
            f = torch.empty(E, dtype=torch.float32, device=dev)
            if E > 0:
                rt = runtime(dev)
                rt.bind_runs(self.runs, self.dtp)
                A = a.cabi()
                ed = self.cabi()
                abi.call("crk_edges_export", rt.ctx, self.dim, C.byref(A), C.byref(ed), abi.ptr(self.masses),
                         float(self.dt), abi.ptr(ap), abi.ptr(av), abi.ptr(f), rt.stream())
            self._full = (ap, av, f)
        return self._full


class StoredPairs:
    """Pair data read back from a trajectory file: ap_ij, av_ij and f_ij over
    the edges of the stage's adjacency, in the reference's layout
    (src/warpSPH/systems/compSPH.py:30-32); ``full()`` returns them as a
    StagePairs would."""

    def __init__(self, ap_ij=None, av_ij=None, f_ij=None):
        self._full = (ap_ij, av_ij, f_ij)

    def full(self):
        return self._full


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
    # pair data of the last derivative evaluation (ap_ij, av_ij, f_ij below)
    pairs: Optional[StagePairs] = None

    # fields integrated by the RK driver (integrated(...) in the reference)
    INTEGRATED = ("positions", "velocities", "internalEnergies")

    EDGE_FIELDS = ("pairs",)

    def initializeNewState(self) -> "CompSPHState":
        """Copy of the per-particle fields; the edge arrays are rebuilt by the
        next derivative evaluation and are not copied."""
        return CompSPHState(**{f.name: (None if f.name in self.EDGE_FIELDS else _clone(getattr(self, f.name)))
                               for f in dataclasses.fields(self)})

    # How one derivative evaluation uses the per-particle fields: read only;
    # overwritten before any read; read and then updated in place.
    READ_ONLY = ("masses", "kinds", "materials", "UIDs", "totalEnergies")
    OUTPUTS = ("densities", "entropies", "pressures", "soundspeeds")
    CARRIED = ("supports", "divergence", "alpha0s", "alphas")

    def stageState(self, integrated: str) -> "CompSPHState":
        """State of one stage inside the integrator: read-only fields are
        shared, carried fields copied, outputs freshly allocated.  The
        integrated fields (positions, velocities, internalEnergies) are shared
        with this state when ``integrated == "share"`` (the stage only reads
        them) or freshly allocated when ``integrated == "alloc"`` (the
        integrator writes the stage values)."""
        if integrated not in ("share", "alloc"):
            raise ValueError(f"integrated must be 'share' or 'alloc', got {integrated!r}")
        kw = {}
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            if f.name in self.EDGE_FIELDS:
                v = None
            elif isinstance(v, torch.Tensor):
                if f.name in self.INTEGRATED:
                    v = v if integrated == "share" else torch.empty_like(v)
                elif f.name in self.OUTPUTS:
                    v = torch.empty_like(v)
                elif f.name not in self.READ_ONLY:
                    v = v.clone()
            kw[f.name] = v
        return CompSPHState(**kw)

    @property
    def ap_ij(self) -> Optional[torch.Tensor]:
        """[E, dim] pairwise pressure acceleration over the adjacency edges."""
        return None if self.pairs is None else self.pairs.full()[0]

    @property
    def av_ij(self) -> Optional[torch.Tensor]:
        """[E, dim] pairwise viscous acceleration over the adjacency edges."""
        return None if self.pairs is None else self.pairs.full()[1]

    @property
    def f_ij(self) -> Optional[torch.Tensor]:
        """[E] energy partition over the adjacency edges."""
        return None if self.pairs is None else self.pairs.full()[2]

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
        self.runs = None  # the ensemble run table it was built for (per-run grids)
        # LDS tiles (runtime.build_tiles): per-tile run tables and the
        # tile-local index of every entry's neighbour, or None
        self.tiles = None
        self.nbr_loc = None
        self._export = None

    @classmethod
    def fromExport(cls, i, j, numNeighbors, edgeOffsets, referencePositions, referenceSupports):
        """The adjacency of the reference's export form (src/warpSPH/io.py:8-20,
        145-160): edges grouped by row, row u's numNeighbors[u] edges starting
        at edgeOffsets[u], i and j user indices.  The rows' order in the edge
        arrays becomes the sorted row order (``perm``), so that the edges keep
        their slots and a re-export reproduces the arrays; the positions and
        supports it was built at are those of the file.  No cell table or LDS
        tiles are attached: the pair passes read it through its neighbour ids.
        Raises ValueError when the arrays do not describe such a grouping."""
        dev = j.device
        nn = numNeighbors.to(torch.int64)
        off = edgeOffsets.to(torch.int64)
        n = int(nn.shape[0])
        E = int(j.shape[0])
        if off.shape[0] != n or i.shape[0] != E or int(nn.sum()) != E:
            raise ValueError(f"adjacency arrays disagree: {n} rows, {int(nn.sum())} counted edges, "
                             f"{int(i.shape[0])} i, {E} j, {int(off.shape[0])} offsets")
        # an empty row sorts before the row that starts at the same offset
        order = torch.argsort(off * 2 + (nn > 0).to(torch.int64), stable=True)
        start = torch.cumsum(nn[order], 0) - nn[order]
        if not torch.equal(off[order], start):
            raise ValueError("edgeOffsets do not tile the edge arrays row by row")
        rows = torch.repeat_interleave(order, nn[order])
        if not torch.equal(i.to(torch.int64), rows):
            raise ValueError("edge i does not match the row its offset places it in")
        jj = j.to(torch.int64)
        if E and (int(jj.min()) < 0 or int(jj.max()) >= n):
            raise ValueError("edge j out of range")
        perm = order.to(torch.int32)
        rank = torch.empty(n, dtype=torch.int32, device=dev)
        rank[order] = torch.arange(n, dtype=torch.int32, device=dev)
        row_offsets = torch.cat([start, torch.tensor([E], dtype=torch.int64, device=dev)]).to(torch.int32)
        nbr = rank[jj].contiguous()
        xb = referencePositions[order].to(torch.float32).contiguous()
        hb = referenceSupports[order].to(torch.float32).contiguous()
        adj = cls(n, None, None, perm, rank, None, None, row_offsets, nbr, xb, hb)
        adj._export = (i.to(torch.int64), jj, nn, off)
        return adj

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
        if self.tiles is not None:
            a.tiles = self.tiles.data_ptr()
            a.nbr_loc = self.nbr_loc.data_ptr() if self.n_edges > 0 else None
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
    t: Any = 0.0  # per-run times (numpy array) for an ensemble
    # run table of an ensemble (warpsph_amd.ensemble.Runs); None: one simulation
    runs: Any = None

    def stageSystem(self, integrated: str) -> "CompSPHSystem":
        """System of one integrator stage (see CompSPHState.stageState)."""
        return CompSPHSystem(state=self.state.stageState(integrated), adjacency=self.adjacency, t=self.t,
                             domain=self.domain, runs=self.runs)

    def initializeNewState(self, *args, verbose=False, **kwargs) -> "CompSPHSystem":
        return CompSPHSystem(state=self.state.initializeNewState(), adjacency=self.adjacency, t=self.t,
                             domain=self.domain, runs=self.runs)

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
            # an ensemble's per-particle dt (the dt argument is then ignored)
            rt.bind_runs(s0.pairs.runs, s0.pairs.dtp)
            e0 = s0.pairs.cabi()
            e1 = s1.pairs.cabi()
            u_new = torch.empty_like(st0.internalEnergies)
            A0, A1 = a0.cabi(), a1.cabi()
            from . import timing
            with timing.scope("compatible_energy"):
                    abi.call("crk_compatible_energy", rt.ctx, C.byref(cfg), n, float(dt), abi.ptr(st0.masses),
                         abi.ptr(st0.velocities),
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
