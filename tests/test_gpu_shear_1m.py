"""The 2D shear layer at its benchmark size, 1024 x 1024 = 1,048,576
particles: bit-exact adjacency and one RK2 step against the oracle, the
Verlet reuse tests at that size (their grid-stride reductions then run more
than one round per thread), and 100 steps at 128^2 for the density L1.

Per-element agreement is reported next to the normwise bar: for each field,
percentiles of |gpu - oracle| / max(|oracle_i|, 1e-3 max |oracle|) (elements
within 1e-3 of the field's scale of zero are measured against that floor)."""
import ctypes as C

import numpy as np
import pytest
import torch

from tests.util import RTOL, cpu_copy, relerr, user_rows

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _elementwise(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    floor = 1e-3 * np.max(np.abs(b))
    e = np.abs(a - b) / np.maximum(np.abs(b), floor)
    return {p: float(np.percentile(e, p)) for p in (50, 99, 99.9, 100)}


@pytest.fixture(scope="module")
def shear1m():
    from warpsph_amd import sample
    return sample.shear_layer_2d(nx=1024, device="cuda")


def test_adjacency_bitexact_1m(shear1m):
    from oracle import oracle as O
    from warpsph_amd.configurations import toCrkConfig
    from warpsph_amd.runtime import build_adjacency
    announce("adjacency 1024^2")
    system, config, scheme = shear1m
    cfg = toCrkConfig(config, scheme)
    st = system.state
    adj = build_adjacency(cfg, st.positions, st.supports)
    torch.cuda.synchronize()
    sim = O.from_system(cpu_copy(system, config, scheme), config, scheme, 32)
    sim.build_adjacency()
    ref = sim.adjacency(-1)
    assert adj.grid.bits == ref["grid"].bits and list(adj.grid.ncell) == list(ref["grid"].ncell)
    assert np.array_equal(adj.keys.cpu().numpy().view(np.uint32), ref["keys"])
    assert np.array_equal(adj.perm.cpu().numpy(), ref["perm"])
    assert np.array_equal(adj.cell_start.cpu().numpy(), ref["cell_start"])
    assert np.array_equal(adj.cell_end.cpu().numpy(), ref["cell_end"])
    off, j, _ = user_rows(adj)
    assert adj.n_edges == ref["n_edges"] and adj.n_edges > 60 * st.numParticles
    assert np.array_equal(off, ref["offsets"])
    assert np.array_equal(j, ref["nbr"])


def test_one_step_1m(shear1m):
    from oracle import oracle as O
    from warpsph_amd import sample
    from warpsph_amd.integrator import getIntegrator
    from warpsph_amd.scheme import crkSPH_step
    announce("one RK2 step 1024^2")
    system, config, scheme = shear1m
    cpu = cpu_copy(system, config, scheme)
    dt = sample.cfl_dt(system.state, config, scheme) * (2.0 / 3.0)
    res = getIntegrator().function(system, crkSPH_step, dt, config, scheme)
    torch.cuda.synchronize()
    st = res.state.state
    sim = O.from_system(cpu, config, scheme, 32)
    sim.step(dt, 1)
    x0 = cpu.state.positions.numpy().astype(np.float64)
    xo = sim.get("x")
    mi = lambda d: d - np.round(d)  # unit periodic box
    got = {"v": st.velocities, "u": st.internalEnergies, "rho": st.densities, "h": st.supports,
           "P": st.pressures, "c": st.soundspeeds, "alpha": st.alphas, "div": st.divergence}
    # positions: against the step's displacement, with a floor of 4 ulp of
    # the unit box (one FP32 rounding of x ~ 0.5 is 6e-8, the displacement
    # of one step at this resolution ~1e-4)
    dxmax = float(np.max(np.abs(mi(st.positions.cpu().numpy() - xo))))
    errs = {"x": dxmax / max(float(np.max(np.abs(mi(xo - x0)))), 2.0 ** -22 / RTOL)}
    for k, t in got.items():
        errs[k] = relerr(t.cpu().numpy(), sim.get(k))
        print(k, "elementwise", _elementwise(t.cpu().numpy(), sim.get(k)), flush=True)
    print("normwise", errs, flush=True)
    # the FP32 fields of the state after one step; alpha and div are
    # Cullen-Dehnen intermediates that amplify rounding where div ~ 0
    bad = [k for k in ("x", "v", "u", "rho", "h", "P", "c") if not errs[k] <= RTOL]
    if bad:
        # a field beyond RTOL passes only where FP32 arithmetic itself is
        # that far from FP64 (a small vector component, e.g. v_y of the
        # shear, measured against its own scale): the GPU within 2x of the
        # FP32 oracle's own error against the FP64 oracle
        sim64 = O.from_system(cpu, config, scheme, 64)
        sim64.step(dt, 1)
        for k in list(bad):
            if k == "x":
                continue
            ref64 = sim64.get(k)
            e_gpu, e_o32 = relerr(got[k].cpu().numpy(), ref64), relerr(sim.get(k), ref64)
            print(f"{k}: gpu vs fp64 {e_gpu:.2e}, fp32 oracle vs fp64 {e_o32:.2e}", flush=True)
            if e_gpu <= 2.0 * e_o32:
                bad.remove(k)
    assert not bad, {k: errs[k] for k in bad}
    for k in ("alpha", "div"):
        assert errs[k] <= 10 * RTOL, (k, errs[k])


def test_verlet_reuse_tests_at_1m(shear1m):
    """The list is kept while 2 max|x - x_b| <= (s_v - max h/h_b) min h_b and
    max h/h_b < s_v.  One particle moved just past / just inside the bound,
    or one support grown past s_v h_b, through the host test, the checked
    stage gather (host and deferred device flag) and the post-Owen test."""
    from warpsph_amd import abi
    from warpsph_amd.configurations import toCrkConfig
    from warpsph_amd.runtime import build_adjacency, runtime
    announce("Verlet tests 1024^2")
    system, config, scheme = shear1m
    cfg = toCrkConfig(config, scheme)
    st = system.state.initializeNewState()
    st.positions = st.positions.clone()
    st.supports = st.supports.clone()
    adj = build_adjacency(cfg, st.positions, st.supports)
    rt = runtime("cuda:0")
    s = rt.stream()
    sv = float(np.float32(cfg.verletScale))
    hb = adj.hb.double()
    p = st.numParticles // 2 + 12345
    hmm = torch.zeros(2, dtype=torch.float32, device="cuda")

    def host_valid():
        v = C.c_int32()
        abi.call("crk_adjacency_valid", rt.ctx, C.byref(cfg), adj.n, abi.ptr(adj.perm), abi.ptr(adj.xb),
                 abi.ptr(adj.hb), abi.ptr(st.positions), abi.ptr(st.supports), C.byref(v), s)
        return bool(v.value)

    def checked_valid(owen):
        """(checked gather, post-Owen test) through the host reads, and the
        same through the deferred device flag."""
        h_keep = st.supports.clone()
        cst = st.cabi()
        A = adj.cabi()
        out = []
        for deferred in (False, True):
            st.supports.copy_(h_keep)
            abi.call("crk_flag_clear", rt.ctx, s)
            v = C.c_int32()
            abi.call("crk_stage_begin_checked", rt.ctx, C.byref(cfg), C.byref(A), C.byref(cst), abi.ptr(adj.xb),
                     abi.ptr(adj.hb), abi.ptr(hmm), None if deferred else C.byref(v), s)
            ov = C.c_int32(1)
            if owen:
                abi.call("crk_owen_iteration", rt.ctx, C.byref(cfg), C.byref(cst), abi.ptr(hmm), None, s)
                abi.call("crk_owen_valid", rt.ctx, C.byref(cfg), None if deferred else C.byref(ov), s)
            if deferred:
                f = C.c_int32()
                abi.call("crk_flag_read", rt.ctx, C.byref(f), s)
                out.append(not bool(f.value))
            else:
                out += [bool(v.value), bool(ov.value)]
        st.supports.copy_(h_keep)
        return tuple(out)

    rt.ensure_lut(cfg.kernel, 2, cfg.targetNeighbors)
    assert host_valid() and checked_valid(True) == (True, True, True)
    x0 = st.positions[p].clone()
    bound = 0.5 * (sv - 1.0) * float(hb.min())
    for frac, ok in ((0.98, True), (1.02, False)):
        st.positions[p, 0] = float(x0[0]) + frac * bound
        assert host_valid() == ok, frac
        assert checked_valid(False) == (ok, True, ok), frac
    # half the bound: the post-Owen test (supports moved by a few per cent)
    # still holds; past the bound it fails with the gather's displacement
    st.positions[p, 0] = float(x0[0]) + 0.5 * bound
    assert checked_valid(True) == (True, True, True)
    st.positions[p, 0] = float(x0[0]) + 1.02 * bound
    assert checked_valid(True) == (False, False, False)
    st.positions[p] = x0
    # grown support: g = max h / h_b past s_v fails even without motion
    h0 = float(st.supports[p])
    st.supports[p] = float(hb[int(adj.rank[p])]) * sv * 1.001
    assert not host_valid()
    assert checked_valid(False) == (False, True, False)
    st.supports[p] = h0 * 1.001
    assert host_valid() and checked_valid(False) == (True, True, True)


def test_density_l1_100_steps_128():
    from oracle import oracle as O
    from warpsph_amd import sample
    from warpsph_amd.integrator import getIntegrator
    from warpsph_amd.scheme import crkSPH_step
    announce("100 steps 128^2")
    system, config, scheme = sample.shear_layer_2d(nx=128, device="cuda")
    cpu = cpu_copy(system, config, scheme)
    dt = sample.cfl_dt(system.state, config, scheme) * (2.0 / 3.0)
    integ = getIntegrator()
    s = system
    for _ in range(100):
        s = integ.function(s, crkSPH_step, dt, config, scheme).state
    torch.cuda.synchronize()
    sim = O.from_system(cpu, config, scheme, 32)
    sim.step(dt, 100)
    rg = s.state.densities.cpu().double().numpy()
    ro = sim.get("rho")
    l1 = float(np.mean(np.abs(rg - ro)) / np.mean(np.abs(ro)))
    print(f"density L1 after 100 steps at 128^2: {l1:.3e}", flush=True)
    assert l1 <= 1e-4
