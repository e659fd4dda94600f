"""LDS tiles of an adjacency (crk_adjacency_tiles) and the tiled passes.

The tile table and the local indices are checked against the adjacency they
index (every entry's local index maps back to its neighbour through the
tile's runs); and a tiled step reproduces the untiled step bit for bit
whether every tile is staged in LDS, some are wide (global gathers) or all
are, since each lane adds the same pairs in the same order either way."""
import numpy as np
import pytest
import torch

from tests.test_gpu_parity import announce, make_case

pytestmark = pytest.mark.gpu


def _decode(adj):
    from warpsph_amd import abi
    W = abi.CRK_TILE_WORDS
    tab = adj.tiles.view(-1, W).cpu().numpy()
    off = adj.row_offsets.cpu().numpy().astype(np.int64)
    nbr = adj.nbr.cpu().numpy()
    loc = adj.nbr_loc.cpu().numpy().view(np.uint16).astype(np.int64)
    return tab, off, nbr, loc


@pytest.mark.parametrize("name", ["shear2d", "sod1d", "box2d_open"])
def test_tile_tables_map_back_to_the_neighbours(name):
    from warpsph_amd import abi
    from warpsph_amd.configurations import toCrkConfig
    from warpsph_amd.runtime import build_adjacency
    announce(f"tile tables {name}")
    system, config, scheme = make_case(name)
    cfg = toCrkConfig(config, scheme)
    adj = build_adjacency(cfg, system.state.positions, system.state.supports)
    tab, off, nbr, loc = _decode(adj)
    R = abi.CRK_TILE_ROWS
    staged = 0
    for t in range(tab.shape[0]):
        nr, U = int(tab[t, 0]), int(tab[t, 1])
        e0, e1 = off[t * R], off[min(adj.n, (t + 1) * R)]
        if nr < 0:
            continue
        staged += 1
        starts = tab[t, 4:4 + 2 * nr:2]
        offs = tab[t, 5:5 + 2 * (nr + 1):2]
        assert offs[0] == 0 and offs[-1] == U and U <= abi.CRK_TILE_CAP and np.all(np.diff(offs) > 0)
        assert e1 - (e0 & ~7) <= abi.CRK_TILE_IDS  # the entries' local indices fit the LDS copy
        lens = np.diff(offs)
        # runs are disjoint, ascending and not adjacent (merged otherwise)
        assert np.all(starts[1:] > starts[:-1] + lens[:-1])
        q = loc[e0:e1]
        assert np.all(q < U)
        r = np.searchsorted(offs, q, side="right") - 1
        assert np.array_equal(starts[r] + (q - offs[r]), nbr[e0:e1])
    assert staged > 0


def _stage(name, cap):
    """One derivative evaluation with the given tile cap (None: no tiles):
    the support-to-velocity-gradient intermediates, the pairs, the update."""
    import ctypes as C
    from tests.test_gpu_parity import EXPORT_FIELDS
    from warpsph_amd import abi, sample
    from warpsph_amd.runtime import runtime
    from warpsph_amd.scheme import crkSPH_step
    system, config, scheme = make_case(name)
    rt = runtime("cuda:0")
    keep = rt.tile_cap
    rt.tile_cap = cap
    try:
        dt = sample.cfl_dt(system.state, config, scheme) * (2.0 / 3.0)
        update, adj, st = crkSPH_step(system, dt, config, scheme)
        n, D = st.positions.shape
        out = {}
        for fid, key in EXPORT_FIELDS:
            comps = {"B": D, "gA": D, "gB": D * D, "J": D * D}.get(key, 1)
            t = torch.empty(n * comps, dtype=torch.float32, device="cuda")
            try:
                abi.call("crk_export_field", rt.ctx, fid, abi.ptr(t), rt.stream())
            except abi.CrkError:
                continue  # not computed in this configuration
            out[key] = t.cpu().numpy()
        p = st.pairs
        A, E = adj.cabi(), p.cabi()
        ap = torch.empty((adj.n_edges, D), dtype=torch.float32, device="cuda")
        abi.call("crk_edges_export", rt.ctx, D, C.byref(A), C.byref(E), abi.ptr(st.masses), float(dt),
                 abi.ptr(ap), None, None, rt.stream())
        torch.cuda.synchronize()
        cnt = p.count.cpu().numpy()
        off = adj.row_offsets.cpu().numpy()
        nb = p.nbr.cpu().numpy()
        out["pairs"] = np.concatenate([nb[off[k]:off[k] + cnt[k]] for k in range(n)])
        out["h"] = st.supports.cpu().numpy()
        upd = {k: getattr(update, k).cpu().numpy() for k in ("dvdt", "dudt", "drhodt", "dEdt")}
        upd["ap"] = ap.cpu().numpy()
    finally:
        rt.tile_cap = keep
    return out, upd, adj


def _one_step(name, cap):
    from warpsph_amd import sample
    from warpsph_amd.integrator import getIntegrator
    from warpsph_amd.runtime import runtime
    from warpsph_amd.scheme import crkSPH_step
    system, config, scheme = make_case(name)
    rt = runtime("cuda:0")
    keep = rt.tile_cap
    rt.tile_cap = cap
    try:
        dt = sample.cfl_dt(system.state, config, scheme) * (2.0 / 3.0)
        res = getIntegrator().function(system, crkSPH_step, dt, config, scheme)
        torch.cuda.synchronize()
    finally:
        rt.tile_cap = keep
    st = res.state.state
    return {k: getattr(st, k).cpu().numpy() for k in ("positions", "velocities", "internalEnergies", "densities",
                                                      "supports", "pressures", "alphas", "divergence")}


@pytest.mark.parametrize("name", ["shear2d", "sod1d", "shear2d_c4_alphaq", "box2d_open"])
def test_tiled_passes_equal_untiled(name):
    """Staged, partly wide and all-wide tiles against no tiles: the support
    sweep, the CRK factors and the velocity gradient (and the pair lists)
    bit for bit; the tiled acceleration pass adds the same pairs in the same
    order but is compiled for a larger block, where the compiler may fuse a
    different few multiply-adds, so its outputs and one whole RK2 step agree
    to within 1e-6 of each field's scale."""
    from tests.util import relerr
    from warpsph_amd import abi
    from warpsph_amd.runtime import tile_stats
    announce(f"tiled passes {name}")
    ref, ref_u, _ = _stage(name, None)
    ref_step = _one_step(name, None)
    full, full_u, adj = _stage(name, abi.CRK_TILE_CAP)
    nt, nwide, mean_u = tile_stats(adj)
    print(f"{name}: {nt} tiles, {nwide} wide, mean staged {mean_u:.0f}", flush=True)
    assert adj.tiles is not None and nwide < nt
    Us = adj.tiles.view(-1, abi.CRK_TILE_WORDS)[:, 1].cpu().numpy()
    mixed_cap = max(1, int(np.median(Us[Us > 0]))) if np.any(Us > 0) else 1
    for label, cap in (("staged", abi.CRK_TILE_CAP), ("mixed", mixed_cap), ("wide", 1)):
        got, got_u, _ = (full, full_u, adj) if label == "staged" else _stage(name, cap)
        for k in ref:
            assert np.array_equal(got[k], ref[k]), (label, k)
        for k in ref_u:
            assert relerr(got_u[k], ref_u[k]) <= 1e-6, (label, k, relerr(got_u[k], ref_u[k]))
        step = _one_step(name, cap)
        for k in ref_step:
            assert relerr(step[k], ref_step[k]) <= 1e-6, (label, k, relerr(step[k], ref_step[k]))


def test_tiles_past_the_entry_budget_are_wide():
    """Rows with many candidates (a two-state run at 64 target neighbours,
    about 90 entries per row): tiles whose rows hold more than CRK_TILE_IDS
    entries are wide in the table, so that every tiled pass gathers them
    from global memory alike (a pass that staged such a tile's particles
    while another gathered it would read local indices nobody wrote).  Three
    RK2 steps with tiles against three without, within 1e-6 of each field's
    scale."""
    from tests.util import relerr
    from warpsph_amd import abi, sample
    from warpsph_amd.configurations import toCrkConfig
    from warpsph_amd.ensemble import two_state_run
    from warpsph_amd.integrator import getIntegrator
    from warpsph_amd.runtime import build_adjacency, runtime
    from warpsph_amd.scheme import crkSPH_step
    announce("tiles past the entry budget")
    rt = runtime("cuda:0")
    keep = rt.tile_cap
    system, config, scheme = two_state_run(48, 7, device="cuda", targetNeighbors=64.0)
    cfg = toCrkConfig(config, scheme)
    adj = build_adjacency(cfg, system.state.positions, system.state.supports)
    tab, off, _, _ = _decode(adj)
    R = abi.CRK_TILE_ROWS
    over = [t for t in range(tab.shape[0])
            if off[min(adj.n, (t + 1) * R)] - (off[t * R] & ~7) > abi.CRK_TILE_IDS]
    print(f"{tab.shape[0]} tiles, {len(over)} past the entry budget, {off[-1] / adj.n:.1f} entries per row",
          flush=True)
    assert over and all(tab[t, 0] < 0 for t in over)
    runs = {}
    try:
        for label, cap in (("tiles", abi.CRK_TILE_CAP), ("none", None)):
            rt.tile_cap = cap
            s = two_state_run(48, 7, device="cuda", targetNeighbors=64.0)[0]
            for _ in range(3):
                dt = sample.cfl_dt(s.state, config, scheme)
                s = getIntegrator().function(s, crkSPH_step, dt, config, scheme).state
            torch.cuda.synchronize()
            st = s.state
            runs[label] = {k: getattr(st, k).cpu().numpy() for k in ("velocities", "internalEnergies", "densities",
                                                                     "supports", "alphas")}
    finally:
        rt.tile_cap = keep
    for k, ref in runs["none"].items():
        assert np.all(np.isfinite(runs["tiles"][k])), k
        assert relerr(runs["tiles"][k], ref) <= 1e-6, (k, relerr(runs["tiles"][k], ref))


@pytest.mark.parametrize("name", ["shear2d", "sod1d"])
def test_tile_block_sizes_give_identical_steps(name):
    """Every tiled pass run with 256, 512 and 1024 threads per tile block,
    including shapes whose rounds are wider than the tile (G = 1 or 2 with
    512 or 1024 threads: the lanes past the tile's last row idle): a row is
    computed by the same code whichever round holds it, so one RK2 step is
    bit-identical for every shape."""
    from warpsph_amd import abi
    from warpsph_amd.autotune import TILED, set_groups
    from warpsph_amd.runtime import runtime
    announce(f"tile block sizes {name}")
    rt = runtime("cuda:0")
    keep = dict(getattr(rt, "groups", {}))
    try:
        for g in (1, 4, 16):  # the lane-group size changes the summation order: compare within one G
            base = None
            for tb in (256, 512, 1024):
                shape = {}
                for p in TILED:
                    shape[p] = g
                    shape[p + ".threads"] = tb
                set_groups("cuda:0", shape)
                got = _one_step(name, abi.CRK_TILE_CAP)
                if base is None:
                    base = got
                    continue
                for k in base:
                    assert np.array_equal(got[k], base[k]), (g, tb, k)
    finally:
        restore = {p: keep.get(p, 4) for p in TILED}
        restore.update({p + ".threads": keep.get(p + ".threads", 1024 if p == "accel" else 256) for p in TILED})
        set_groups("cuda:0", restore)
