"""Pins of the oracle's neighbour search: Morton keys (hand-computed known
answers), stable sort (numpy stable argsort), cell table (partition of the
sorted order) and neighbour sets (O(N^2) minimum-image brute force)."""
import numpy as np
import pytest

from oracle import oracle as O
from warpsph_amd import abi


def test_morton_known_answers():
    assert O.morton(1, 37) == 37
    # 2D: x bits at even positions, y bits at odd positions
    assert [O.morton(2, x, y) for x, y in [(1, 0), (0, 1), (1, 1), (2, 0), (3, 3)]] == [1, 2, 3, 4, 15]
    assert O.morton(2, 5, 9) == 0b10010011  # x=101, y=1001
    assert O.morton(2, 4095, 0) == int("01" * 12, 2) and O.morton(2, 0, 4095) == int("10" * 12, 2)
    # 3D: x, y, z bits at positions 3k, 3k+1, 3k+2
    assert [O.morton(3, *c) for c in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1), (2, 0, 0)]] == [1, 2, 4, 7, 8]
    assert O.morton(3, 255, 255, 255) == (1 << 24) - 1


def _cfg(dim, periodic, lo, hi, verlet):
    c = abi.CrkConfig()
    c.domain.dim = dim
    for d in range(dim):
        c.domain.periodic[d] = int(periodic)
        c.domain.min[d] = lo
        c.domain.max[d] = hi
    c.kernel = 0
    c.verletScale = verlet
    c.targetNeighbors = 16.0
    c.gamma = 1.4
    c.energyScheme = 5
    c.adaptiveSupportScheme = 0
    c.switchParams.scheme = 6
    return c


def brute_force(x, h, s, L, periodic):
    n, D = x.shape
    dx = x[:, None, :] - x[None, :, :]
    if periodic:
        half = np.float32(0.5) * L
        dx = np.where(dx > half, dx - L, np.where(dx < -half, dx + L, dx)).astype(np.float32)
    r2 = np.zeros((n, n), np.float32)
    for d in range(D):
        r2 = (r2 + dx[..., d] * dx[..., d]).astype(np.float32)
    R = (np.float32(s) * np.maximum(h[:, None], h[None, :])).astype(np.float32)
    return r2 < R * R


@pytest.mark.parametrize("dim,periodic", [(1, True), (2, True), (3, True), (2, False), (3, False)])
def test_sort_cells_and_neighbour_sets(dim, periodic):
    rng = np.random.default_rng(7 + dim)
    n = {1: 300, 2: 900, 3: 1500}[dim]
    lo, hi = 0.0, 1.0
    x = rng.random((n, dim)).astype(np.float32)
    spacing = (1.0 / n) ** (1.0 / dim)
    h = (spacing * (2.0 + rng.random(n))).astype(np.float32)
    s = float(np.float32(2.0 ** (1 / dim)))
    cfg = _cfg(dim, periodic, lo, hi, s)
    sim = O.OracleSim(cfg, n, 32)
    sim.set("x", x.astype(np.float64))
    sim.set("h", h.astype(np.float64))
    sim.set("m", np.ones(n))
    sim.build_adjacency()
    adj = sim.adjacency(-1)
    keys, perm = adj["keys"], adj["perm"]
    # keys recomputed from the grid in numpy
    g = adj["grid"]
    cells = np.zeros((n, 3), np.int64)
    for d in range(dim):
        t = ((x[:, d] - np.float32(g.origin[d])) * np.float32(g.inv_cell[d])).astype(np.float32)
        c = np.floor(t).astype(np.int64)
        c = np.mod(c, g.ncell[d]) if periodic else np.clip(c, 0, g.ncell[d] - 1)
        cells[:, d] = c
    ukeys = np.array([O.morton(dim, *cells[i, :3]) for i in range(n)], dtype=np.uint32)
    assert np.array_equal(perm, np.argsort(ukeys, kind="stable"))
    assert np.array_equal(keys, ukeys[perm])
    # cell table: start/end partition the sorted order
    cs, ce = adj["cell_start"], adj["cell_end"]
    occ = np.unique(keys)
    assert int(np.sum(ce[occ] - cs[occ])) == n
    for k in occ[:50]:
        assert np.all(keys[cs[k]:ce[k]] == k)
        assert (cs[k] == 0 or keys[cs[k] - 1] != k) and (ce[k] == n or keys[ce[k]] != k)
    # neighbour sets
    L = np.float32(hi - lo)
    want = brute_force(x, h, s, L, periodic)
    off, nbr = adj["offsets"], adj["nbr"]
    for i in range(n):
        row = nbr[off[i]:off[i + 1]]
        assert len(set(row.tolist())) == len(row), "duplicate neighbour"
        assert set(row.tolist()) == set(np.nonzero(want[i])[0].tolist()), i
