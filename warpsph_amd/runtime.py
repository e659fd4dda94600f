"""Per-device runtime: the library context, the Owen table and the
adjacency build (the MI355X replacement of warpSPHCore.buildVerletList).
"""
from __future__ import annotations

import ctypes as C
import threading
from typing import Dict, Optional, Tuple

import numpy as np
import torch

from . import abi, timing
from .utils import nH_to_n_h

_runtimes: Dict[int, "Runtime"] = {}
_lock = threading.Lock()


class Runtime:
    def __init__(self, device: torch.device):
        if device.type != 'cuda':
            raise ValueError(f"the CRKSPH step runs on a GPU; got device {device}")
        self.device = device
        self.index = device.index if device.index is not None else torch.cuda.current_device()
        with torch.cuda.device(self.index):
            p = C.c_void_p()
            abi.call("crk_ctx_create", C.byref(p))
        self.ctx = p
        self._lut_key: Optional[Tuple] = None
        self.scratch = torch.empty(16, dtype=torch.float32, device=device)  # [0,1] h min/max, [2] Owen ratio
        self.nsing = torch.zeros(1, dtype=torch.int32, device=device)
        # evaluate the Cullen-Dehnen switch inside the acceleration sweep when
        # the viscosity does not read alpha (False: separate crk_switch pass)
        self.fuse_switch = True
        # compute V and v_sig in the final Owen sweep and compact the pairs in
        # the moments pass (True), or run the separate crk_support pass
        # (False; measured faster on MI355X: the compaction costs more inside
        # the moments pass than in its own sweep)
        self.fuse_support = False
        # deferred Verlet tests (graph replay): the tests set the context's
        # device flag instead of returning, and nothing is rebuilt
        self.deferred = False
        # LDS tiles for the pair passes of 1D and 2D adjacencies (None: off);
        # the value is the most particles one tile may stage
        self.tile_cap = abi.CRK_TILE_CAP
        # ensembles: after a failed test only the runs that failed it are
        # rebuilt (crk_adjacency_count_keep), so that every run evolves
        # exactly as alone, its signal speed and switch included.  Off by
        # default: the runs' lists then age apart and some run fails almost
        # every step, while a partial rebuild still re-sorts every run and
        # rebuilds every tile (256-run workload: 30.6 against 21.6 ms per
        # step).  False: every run is rebuilt when one fails.
        self.keep_runs = False
        # speculative stages (begin_speculative): the adaptive-support pass
        # and the support sweep are queued before the host reads the Verlet
        # tests that precede them, and do nothing if a list failed; the GPU
        # then never waits for the host's reads (False: blocking tests)
        self.speculate = True

    def stream(self) -> C.c_void_p:
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def bind_runs(self, runs, dtp=None):
        """Bind an ensemble's run table and per-particle dt (None: one simulation)."""
        if runs is None:
            abi.call("crk_bind_runs", self.ctx, 0, None, None, self.stream())
        else:
            runs.bind(self, dtp)

    def ensure_lut(self, kernel: int, dim: int, target_neighbors: float, nlut: int = 4096,
                   nmin: float = 2.0, nmax: float = 6.0):
        """Owen table (optimalSupportOwen.py:52-54: nMin 2, nMax 6, 2^12 entries)."""
        key = (kernel, dim, float(target_neighbors), nlut, nmin, nmax)
        if self._lut_key == key:
            return
        # computed on the GPU (crk_owen_lut_device; within an ulp of crk_owen_lut)
        target = float(nH_to_n_h(float(np.float32(target_neighbors)), dim))
        abi.call("crk_owen_lut_device", self.ctx, kernel, dim, nlut, nmin, nmax, target, self.stream())
        self._lut_key = key


def runtime(device) -> Runtime:
    device = torch.device(device)
    idx = device.index if device.index is not None else torch.cuda.current_device()
    with _lock:
        rt = _runtimes.get(idx)
        if rt is None:
            rt = Runtime(torch.device('cuda', idx))
            _runtimes[idx] = rt
    return rt


def _run_grids(cfg: abi.CrkConfig, rt, runs, positions: torch.Tensor, supports: torch.Tensor):
    """One grid per run of the bound ensemble (the run's h_max and extent)."""
    R = runs.n_runs
    ext = np.empty(7 * R, np.float32)
    abi.call("crk_run_extents", rt.ctx, C.byref(cfg), runs.n, abi.ptr(positions), abi.ptr(supports),
             C.c_void_p(ext.ctypes.data), rt.stream())
    grids = (abi.CrkGrid * R)()
    total = C.c_int64()
    abi.call("crk_grid_plan_runs", C.byref(cfg), R, C.c_void_p(ext.ctypes.data), grids, C.byref(total))
    return grids, int(total.value)


def runs_failed(cfg: abi.CrkConfig, rt, runs, which: int) -> np.ndarray:
    """Which runs of the bound ensemble failed the last Verlet test (blocking):
    which = 0 the stage's, 1 the updated supports'."""
    out = np.zeros(runs.n_runs, dtype=np.uint8)
    abi.call("crk_runs_failed", rt.ctx, C.byref(cfg), int(which), C.c_void_p(out.ctypes.data), rt.stream())
    return out.astype(bool)


def build_adjacency(cfg: abi.CrkConfig, positions: torch.Tensor, supports: torch.Tensor, runs=None, prior=None,
                    failed=None):
    """Morton keys -> stable LSD radix sort -> cell table -> CSR (all on the GPU).
    With ``runs`` (an ensemble, bound to the runtime) every run has its own grid
    and its rows hold only its own particles.  With the ensemble's ``prior``
    list and ``failed`` (a flag per run), only the failed runs are rebuilt;
    the others keep their lists, sort order and build positions
    (crk_adjacency_count_keep), so that each run rebuilds when its own test
    fails, as it would alone."""
    from .systems import AdjacencyList, capacity
    rt = runtime(positions.device)
    n = int(positions.shape[0])
    D = int(cfg.domain.dim)
    dev = positions.device
    keep = None
    if runs is not None and prior is not None and failed is not None and getattr(prior, "runs", None) is runs \
            and prior.n == n and bool(np.any(failed)) and not bool(np.all(failed)):
        keep = ~np.asarray(failed, dtype=bool)
    if runs is not None:
        grid, T = _run_grids(cfg, rt, runs, positions, supports)
        if keep is not None:
            # a kept run's grid (and so its keys and cells) is the prior build's
            R = runs.n_runs
            sz = C.sizeof(abi.CrkGrid)
            new_b = np.frombuffer(grid, dtype=np.uint8).reshape(R, sz)
            old_b = np.frombuffer(prior.grid, dtype=np.uint8).reshape(R, sz)
            new_b[keep] = old_b[keep]
            t_old = np.array([g.table_size for g in prior.grid], dtype=np.int64)
            t_new = np.array([g.table_size for g in grid], dtype=np.int64)
            T = int(t_new.sum())
            delta = (np.cumsum(t_new) - t_new) - (np.cumsum(t_old) - t_old)
            keep_dev = torch.from_numpy(keep.astype(np.uint8)).to(dev)
            delta_dev = torch.from_numpy(delta).to(dev)
        gref = grid
    else:
        hmax = float(supports.max().item())
        lo = np.zeros(3, np.float32)
        hi = np.zeros(3, np.float32)
        if not all(cfg.domain.periodic[d] for d in range(D)):
            lo[:D] = positions.min(dim=0).values.cpu().numpy()
            hi[:D] = positions.max(dim=0).values.cpu().numpy()
        grid = abi.CrkGrid()
        abi.call("crk_grid_plan", C.byref(cfg), hmax, C.c_void_p(lo.ctypes.data), C.c_void_p(hi.ctypes.data),
                 C.byref(grid))
        T = int(grid.table_size)
        gref = C.byref(grid)
    keys = torch.empty(n, dtype=torch.int32, device=dev)  # uint32 bit patterns
    perm = torch.empty(n, dtype=torch.int32, device=dev)
    rank = torch.empty(n, dtype=torch.int32, device=dev)
    cs = torch.empty(T, dtype=torch.int32, device=dev)
    ce = torch.empty(T, dtype=torch.int32, device=dev)
    off = torch.empty(n + 1, dtype=torch.int32, device=dev)
    xb = torch.empty(n * D, dtype=torch.float32, device=dev)
    hb = torch.empty(n, dtype=torch.float32, device=dev)
    E = C.c_int64()
    if keep is not None:
        P = prior.cabi()
        abi.call("crk_adjacency_count_keep", rt.ctx, C.byref(cfg), gref, n, abi.ptr(positions), abi.ptr(supports),
                 abi.ptr(keep_dev), abi.ptr(delta_dev), C.byref(P), abi.ptr(prior.keys), abi.ptr(prior.xb),
                 abi.ptr(prior.hb), abi.ptr(keys), abi.ptr(perm), abi.ptr(rank), abi.ptr(cs), abi.ptr(ce),
                 abi.ptr(off), abi.ptr(xb), abi.ptr(hb), C.byref(E), rt.stream())
    else:
        abi.call("crk_adjacency_count", rt.ctx, C.byref(cfg), gref, n, abi.ptr(positions), abi.ptr(supports),
                 abi.ptr(keys), abi.ptr(perm), abi.ptr(rank), abi.ptr(cs), abi.ptr(ce), abi.ptr(off), abi.ptr(xb),
                 abi.ptr(hb), C.byref(E), rt.stream())
    nbr = torch.empty(capacity(E.value), dtype=torch.int32, device=dev)[:E.value]
    abi.call("crk_adjacency_fill", rt.ctx, C.byref(cfg), gref, n, abi.ptr(cs), abi.ptr(ce), abi.ptr(off),
             abi.ptr(xb), abi.ptr(hb), abi.ptr(nbr), E.value, rt.stream())
    adj = AdjacencyList(n, grid, keys, perm, rank, cs, ce, off, nbr, xb.view(n, D), hb)
    adj.runs = runs
    if rt.tile_cap and D <= 2:
        build_tiles(adj, rt.tile_cap)
    return adj


def build_tiles(adj, cap: int = abi.CRK_TILE_CAP):
    """LDS tiles of an adjacency (crk_adjacency_tiles): the pair passes then
    stage each tile's neighbour records in LDS."""
    rt = runtime(adj.perm.device)
    nt = (adj.n + abi.CRK_TILE_ROWS - 1) // abi.CRK_TILE_ROWS
    adj.tiles = torch.empty(nt * abi.CRK_TILE_WORDS, dtype=torch.int32, device=adj.perm.device)
    adj.nbr_loc = torch.empty(max(adj.n_edges, 1), dtype=torch.int16, device=adj.perm.device)
    A = adj.cabi()
    abi.call("crk_adjacency_tiles", rt.ctx, C.byref(A), abi.ptr(adj.keys), abi.ptr(adj.cell_start),
             abi.ptr(adj.cell_end), int(cap), abi.ptr(adj.tiles), abi.ptr(adj.nbr_loc), rt.stream())
    return adj


def tile_stats(adj):
    """(tiles, wide tiles, mean staged particles of the others) of a tiled adjacency."""
    t = adj.tiles.view(-1, abi.CRK_TILE_WORDS).cpu()
    wide = t[:, 0] < 0
    U = t[~wide, 1].double()
    return int(t.shape[0]), int(wide.sum()), float(U.mean()) if U.numel() else 0.0


def adjacency_valid(cfg: abi.CrkConfig, adj, positions: torch.Tensor, supports: torch.Tensor, runs=None) -> bool:
    if adj is None or adj.n != positions.shape[0] or getattr(adj, "runs", None) is not runs:
        return False
    rt = runtime(positions.device)
    v = C.c_int32()
    abi.call("crk_adjacency_valid", rt.ctx, C.byref(cfg), adj.n, abi.ptr(adj.perm), abi.ptr(adj.xb), abi.ptr(adj.hb),
             abi.ptr(positions), abi.ptr(supports), C.byref(v), rt.stream())
    return bool(v.value)


def rebuild_adjacency(cfg: abi.CrkConfig, positions: torch.Tensor, supports: torch.Tensor, runs=None, prior=None,
                      which=None):
    """A new list after a failed test; for an ensemble with its ``prior``
    list, only the runs that failed test ``which`` (0: the stage's, 1: the
    updated supports') are rebuilt."""
    rt = runtime(positions.device)
    failed = None
    if runs is not None and prior is not None and which is not None and rt.keep_runs:
        failed = runs_failed(cfg, rt, runs, which)
    with timing.scope("adjacency_build"):
        return build_adjacency(cfg, positions, supports, runs, prior, failed)


def flag_wait(rt, slot: int = 0) -> bool:
    """The device flag as the last crk_flag_read_async of ``slot`` copied it
    (True: some list failed)."""
    f = C.c_int32()
    abi.call("crk_flag_wait", rt.ctx, int(slot), C.byref(f))
    return bool(f.value)


def begin_speculative(cfg: abi.CrkConfig, prior, st, cst: abi.CrkState, hminmax: torch.Tensor, runs, adaptive,
                      make_pairs, nsing: torch.Tensor):
    """The stage's Verlet test (with the gather, crk_stage_begin_checked),
    the adaptive-support pass ``adaptive()``, the test of the updated supports
    (crk_owen_valid) and the support sweep over ``make_pairs(adjacency)``, all
    queued before the host reads either test (``Runtime.speculate``).  Both
    tests go to the device flag, which the two passes read (they do nothing
    once it is set), and the flag after each test is copied to the host: the
    GPU never waits for the host's reads.  Returns None without a prior
    list to test (nothing was queued), (None, None, None) when the stage's
    list failed (nothing took effect: the caller rebuilds and takes the
    blocking path), else (adjacency, pairs, True when the updated supports
    failed the test and the list must be rebuilt before the support sweep
    runs again)."""
    rt = runtime(st.positions.device)
    if prior is None or prior.n != st.numParticles or getattr(prior, "runs", None) is not runs:
        return None
    s = rt.stream()
    A = prior.cabi()
    with timing.scope("stage_begin"):
        abi.call("crk_flag_clear", rt.ctx, s)
        abi.call("crk_stage_begin_checked", rt.ctx, C.byref(cfg), C.byref(A), C.byref(cst), abi.ptr(prior.xb),
                 abi.ptr(prior.hb), abi.ptr(hminmax), None, s)
        abi.call("crk_flag_read_async", rt.ctx, 0, s)
    adaptive()
    abi.call("crk_owen_valid", rt.ctx, C.byref(cfg), None, s)  # (ORed into the flag)
    abi.call("crk_flag_read_async", rt.ctx, 1, s)
    pairs = make_pairs(prior)
    nsing.zero_()
    with timing.scope("support"):
        abi.call("crk_support", rt.ctx, C.byref(cfg), C.byref(pairs.cabi()), s)
    if flag_wait(rt, 0):
        return None, None, None
    return prior, pairs, flag_wait(rt, 1)


def begin_checked(cfg: abi.CrkConfig, prior, st, cst: abi.CrkState, hminmax: torch.Tensor, runs=None, then=None):
    """crk_stage_begin fused with the Verlet test of the prior adjacency and the
    support range; rebuilds (and binds the new list) when the prior list no
    longer holds every pair within the Verlet radius, then calls ``then``
    (the pass that follows: it runs on the list returned).  Returns the
    adjacency."""
    rt = runtime(st.positions.device)
    s = rt.stream()
    adj = prior if prior is not None and prior.n == st.numParticles and getattr(prior, "runs", None) is runs \
        else None
    if rt.deferred:
        if adj is None:
            raise RuntimeError("a deferred step needs the adjacency of the state")
        A = adj.cabi()
        abi.call("crk_stage_begin_checked", rt.ctx, C.byref(cfg), C.byref(A), C.byref(cst), abi.ptr(adj.xb),
                 abi.ptr(adj.hb), abi.ptr(hminmax), None, s)
        if then is not None:
            then()
        return adj
    failed_list = None  # (a list that failed its test: an ensemble rebuilds only its failed runs)
    for _ in range(2):
        if adj is None:
            adj = rebuild_adjacency(cfg, st.positions, st.supports, runs, failed_list,
                                    0 if failed_list is not None else None)
        A = adj.cabi()
        ok = C.c_int32()
        with timing.scope("stage_begin"):
            abi.call("crk_stage_begin_checked", rt.ctx, C.byref(cfg), C.byref(A), C.byref(cst), abi.ptr(adj.xb),
                     abi.ptr(adj.hb), abi.ptr(hminmax), C.byref(ok), s)
        if ok.value:
            if then is not None:
                then()
            return adj
        failed_list, adj = (adj if failed_list is None else None), None
    raise RuntimeError("an adjacency built for the current state failed its own validity test")


def ensure_adjacency(cfg: abi.CrkConfig, prior, positions: torch.Tensor, supports: torch.Tensor, runs=None):
    """buildVerletList(..., priorNeighborhood=prior): reuse the prior list
    while every pair inside the current supports is still in it."""
    rt = runtime(positions.device)
    if rt.deferred:
        if prior is None or prior.n != positions.shape[0] or getattr(prior, "runs", None) is not runs:
            raise RuntimeError("a deferred step needs the adjacency of the state")
        abi.call("crk_adjacency_valid", rt.ctx, C.byref(cfg), prior.n, abi.ptr(prior.perm), abi.ptr(prior.xb),
                 abi.ptr(prior.hb), abi.ptr(positions), abi.ptr(supports), None, rt.stream())
        return prior, False
    with timing.scope("adjacency_check"):
        ok = prior is not None and adjacency_valid(cfg, prior, positions, supports, runs)
    if ok:
        return prior, False
    with timing.scope("adjacency_build"):
        adj = build_adjacency(cfg, positions, supports, runs)
    return adj, True
