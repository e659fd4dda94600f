"""The C-ABI library loads without a GPU and exports every declared symbol;
its host-side utilities agree with the oracle's restatement."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "crksph.h")).read()
    return sorted(set(re.findall(r"\b(crk_[a-z0-9_]+)\s*\(", text)))


def test_header_and_bindings_agree():
    from warpsph_amd import abi
    assert declared_symbols() == sorted(abi.EXPORTED)


def test_header_constants_match_bindings():
    """Every integer #define of the header that the bindings mirror has the
    header's value (tile shapes, kernel and pass ids)."""
    from warpsph_amd import abi
    text = open(os.path.join(ROOT, "include", "crksph.h")).read()
    consts = dict((k, int(v)) for k, v in re.findall(r"#define\s+(CRK_[A-Z0-9_]+)\s+(-?\d+)\s*$", text, re.M))
    mirrored = {k: v for k, v in consts.items() if hasattr(abi, k)}
    assert {"CRK_TILE_ROWS", "CRK_TILE_CAP", "CRK_TILE_MAXRUNS", "CRK_TILE_IDS"} <= set(mirrored)
    for k, v in mirrored.items():
        assert getattr(abi, k) == v, k
    assert abi.CRK_TILE_WORDS == 4 + 2 * (consts["CRK_TILE_MAXRUNS"] + 1)


def test_library_exports_every_symbol():
    from warpsph_amd import abi
    L = abi.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name
    assert L.crk_version().decode().startswith("crksph-hip")


def test_struct_layout_matches_library():
    from warpsph_amd import abi
    out = (C.c_int64 * 11)()
    abi.call("crk_struct_sizes", out, 11)
    mirrors = [abi.CrkDomain, abi.CrkDiffusionParams, abi.CrkViscosityParams, abi.CrkSwitchParams, abi.CrkConfig,
               abi.CrkGrid, abi.CrkAdjacency, abi.CrkState, abi.CrkUpdate, abi.CrkEdges, abi.CrkMlp]
    assert [C.sizeof(m) for m in mirrors] == list(out)
    # pointer blocks use natural alignment: the leading int32 is padded
    assert abi.CrkState.positions.offset == 8 and abi.CrkAdjacency.n_edges.offset == 8


@pytest.mark.parametrize("kernel", [0, 1, 2, 3])
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_kernel_constants_match_oracle(kernel, dim):
    from oracle import oracle as O
    from warpsph_amd import abi
    c, s = C.c_double(), C.c_double()
    abi.call("crk_kernel_constants", kernel, dim, C.byref(c), C.byref(s))
    oc, os_ = O.kernel_constants(kernel, dim)
    assert c.value == oc and s.value == os_


@pytest.mark.parametrize("kernel,dim", [(0, 2), (1, 1), (1, 3), (3, 2)])
def test_owen_lut_matches_oracle(kernel, dim):
    from oracle import oracle as O
    from warpsph_amd import abi
    nh = np.empty(4096)
    ps = np.empty(4096)
    abi.call("crk_owen_lut", kernel, dim, 4096, 2.0, 6.0, C.c_void_p(nh.ctypes.data), C.c_void_p(ps.ctypes.data))
    onh, ops = O.owen_lut(kernel, dim)
    assert np.array_equal(nh, onh)
    assert np.array_equal(ps, ops)


def test_grid_plan_matches_oracle():
    from oracle import oracle as O
    from warpsph_amd import abi, sample
    from warpsph_amd.configurations import toCrkConfig
    for nx in (32, 64, 1024):
        system, config, scheme = sample.shear_layer_2d(nx=nx, device="cpu")
        cfg = toCrkConfig(config, scheme)
        h = float(system.state.supports.max())
        g = abi.CrkGrid()
        z = np.zeros(3, np.float32)
        abi.call("crk_grid_plan", C.byref(cfg), h, C.c_void_p(z.ctypes.data), C.c_void_p(z.ctypes.data), C.byref(g))
        og = O.grid_plan(cfg, h)
        assert bytes(g) == bytes(og)


def test_grid_plan_rejects_oversized_support():
    from warpsph_amd import abi, sample
    from warpsph_amd.configurations import toCrkConfig
    system, config, scheme = sample.shear_layer_2d(nx=8, device="cpu")
    cfg = toCrkConfig(config, scheme)
    g = abi.CrkGrid()
    z = np.zeros(3, np.float32)
    with pytest.raises(ValueError):
        abi.call("crk_grid_plan", C.byref(cfg), 0.5, C.c_void_p(z.ctypes.data), C.c_void_p(z.ctypes.data), C.byref(g))


def test_missing_library_fails_loudly(monkeypatch):
    from warpsph_amd import abi
    monkeypatch.setattr(abi, "_lib", None)
    monkeypatch.setattr(abi, "LIB_PATH", "/nonexistent/libcrksph_hip.so")
    with pytest.raises(abi.CrkError):
        abi.lib()
