"""Pins of the oracle's pair-level physics by values the mathematics fixes
(no GPU): the CRK slope limiter, the limited viscosity and the heating it
leaves, the energy partition's branches, the Cullen-Dehnen switch under
homologous compression and expansion, Owen's psi table and support update,
and the CFL step.

Each case is chosen so that a plausible slip fails it:
  limiter        hand values of Frontiere, Raskin & Owen (2017):
                 phi = 4 r / (1 + r)^2 at r = min(r_i, r_j), zero for r <= 0,
                 1 for 0/0, times the taper exp(-((eta - eta_c) / eta_f)^2)
                 below eta_c (limiter.py:11-127)
  viscosity      a linear velocity field on a lattice: grad v = M exactly,
                 r_i = r_j = 1, so the reconstructed velocities of accel.py:
                 137-141 differ by (1 - phi) M x_ij and those of dudt.py:123 by
                 M x_ij; a sign slip in either reconstruction breaks the
                 closed-form Q of every pair
  partition      each branch of balance.py:126-154 by hand
  Cullen-Dehnen  v = -k x: div = -d k, R = -1, Xi = 1, A = d k / dt, v_sig =
                 c + k max r over the Verlet row (CullenDehnen2010.py:121-331,
                 wp_vsig.py:55-86)
  Owen           the psi_H table as the radial shell sum of wp_psi.py:65-81,
                 the support update of optimalSupportOwen.py:27-33 by hand,
                 psi0 of wp_psi0.py:72-103 by brute force over the lattice
  CFL            dt = cfl h_min / (c_max xi) with the clamps and growth limit
                 of timestep/compressible.py:12-66
"""
import math

import numpy as np
import pytest
import torch

from test_oracle_pins import _kernel, _lattice_system, _oracle, _pairs


def _cfg(dim=2, **visc):
    from warpsph_amd.configurations import toCrkConfig
    from warpsph_amd import sample
    from warpsph_amd.configurations import CRKSPHConfig, buildConfig
    dom = sample._domain([0.0] * dim, [1.0] * dim, [True] * dim, "cpu")
    config, _ = buildConfig(device=torch.device("cpu"), domain=dom, dim=dim, targetNeighbors=32.0, cflFactor=0.3)
    scheme = CRKSPHConfig(gamma=5.0 / 3.0)
    for k, v in visc.items():
        setattr(scheme.crkViscosityParams, k, v)
    return toCrkConfig(config, scheme)


# ------------------------------------------------------------------ limiter
def _vl(r):
    """Frontiere, Raskin & Owen (2017) limiter: 4 r / (1 + r)^2 for r > 0."""
    return 0.0 if r <= 0 else 4.0 * r / (1.0 + r) ** 2


def test_limiter_hand_values():
    O = _oracle()
    cfg = _cfg()
    M = np.array([[0.7, 0.2], [0.2, -0.4]])
    dx = np.array([0.6, 0.8])  # |x| = 1
    h = 2.0                    # eta = 0.5 >= eta_crit: no taper
    # equal gradients: r_i = r_j = 1 -> phi = 1
    assert O.limiter_phi(cfg, dx, h, h, M, M) == pytest.approx(1.0, abs=1e-15)
    # opposite projected slopes: r < 0 -> phi = 0
    assert O.limiter_phi(cfg, dx, h, h, M, -M) == 0.0
    # slopes in ratio 2: r_i = 1/2, r_j = 2 -> phi = 4 (1/2) / (3/2)^2 = 8/9
    assert O.limiter_phi(cfg, dx, h, h, M, 2.0 * M) == pytest.approx(8.0 / 9.0, rel=1e-14)
    assert O.limiter_phi(cfg, dx, h, h, 2.0 * M, M) == pytest.approx(8.0 / 9.0, rel=1e-14)
    # ratio 5 -> 4 (1/5) / (6/5)^2 = 5/9
    assert O.limiter_phi(cfg, dx, h, h, M, 5.0 * M) == pytest.approx(5.0 / 9.0, rel=1e-14)
    # both projected slopes zero (0/0 -> 1): an antisymmetric gradient
    W = np.array([[0.0, 1.3], [-1.3, 0.0]])
    assert O.limiter_phi(cfg, dx, h, h, W, W) == pytest.approx(1.0, abs=1e-15)
    # taper below eta_crit: eta = min(|x|/h_i, |x|/h_j) = 0.2
    f = math.exp(-(((0.2 - cfg.crkViscosity.eta_crit) / cfg.crkViscosity.eta_fold) ** 2))
    assert O.limiter_phi(cfg, dx, 5.0, 5.0, M, M) == pytest.approx(f, rel=1e-12)
    # the larger support sets eta: |x|/h_i = 0.25, |x|/h_j = 1.0 -> eta = 0.25
    f = math.exp(-(((0.25 - cfg.crkViscosity.eta_crit) / cfg.crkViscosity.eta_fold) ** 2))
    assert O.limiter_phi(cfg, dx, 4.0, 1.0, M, 2.0 * M) == pytest.approx(f * 8.0 / 9.0, rel=1e-12)
    # switches
    assert O.limiter_phi(_cfg(forceVanLeerOff=True), dx, h, h, M, M) == 0.0
    assert O.limiter_phi(_cfg(forceVanLeerOn=True), dx, h, h, M, -M) == 1.0
    assert O.limiter_phi(_cfg(enableCRKLimiter=False), dx, 5.0, 5.0, M, M) == pytest.approx(1.0, abs=1e-15)
    assert O.limiter_phi(_cfg(enableVanLeerLimiter=False), dx, h, h, M, M) == 0.0


# --------------------------------------------------- limited viscosity, heating
def _linear_flow(M, ns=(24, 24)):
    """FP64 oracle evaluation of v = M x on an unjittered open lattice."""
    from warpsph_amd.configurations import toCrkConfig
    system, config, scheme = _lattice_system(ns, (0.0, 0.0), (1.0, 1.0), (False, False),
                                             lambda x: torch.zeros_like(x), jitter=0.0)
    O = _oracle()
    sim = O.from_system(system, config, scheme, 64)
    x = sim.get("x").reshape(-1, 2)
    sim.set("v", x @ M.T)
    sim.evaluate(sim.cfl_dt())
    return sim, toCrkConfig(config, scheme)


def _q(rho, c, mu, Cl, Cq):
    return rho * (-Cl * c * mu + Cq * mu * mu)


def _mu(w, dx, s):
    return np.minimum(0.0, np.einsum("ea,ea->e", w, dx) / s / (np.sum(dx * dx, 1) / s ** 2 + 1e-7 * s ** 2))


@pytest.mark.parametrize("M", [np.array([[-0.8, 0.3], [0.3, -0.5]]), np.array([[0.4, -0.9], [0.2, -1.1]])])
def test_limited_viscosity_of_a_linear_flow(M):
    """grad v = M everywhere, so r_i = r_j = 1 and phi is the taper alone.
    accel (accel.py:137-141): vdot_i - vdot_j = (1 - phi) M x_ij, zero for
    eta >= eta_crit; the pair's av_ij = (Q_i + Q_j) / (P_i + P_j) ap_ij.
    dudt (dudt.py:123): vdot_i - vdot_j^U = M x_ij for any phi, and
    du_i/dt = sum_j [P_j + (Q^U_i + Q^U_j) / 2] (v_i - v_j).G_ij with
    G_ij = -ap_ij / (P_i + P_j)."""
    O = _oracle()
    sim, cfg = _linear_flow(M)
    n = sim.n
    x, h = sim.get("x").reshape(n, 2), sim.get("h")
    assert np.max(np.abs(sim.get("J").reshape(n, 2, 2) - M[None])) < 1e-10
    P, rho, c = sim.get("P"), sim.get("rho"), sim.get("c")
    _, kscale = O.kernel_constants(cfg.kernel, 2)
    i, j, dx = _pairs(sim, x, [1.0, 1.0], [False, False])
    r = np.linalg.norm(dx, axis=1)
    ap, av = sim.edge_field(0, "ap"), sim.edge_field(0, "av")
    act = r < np.maximum(h[i], h[j])
    eta = np.minimum(r / h[i], r / h[j])
    ec, ef = cfg.crkViscosity.eta_crit, cfg.crkViscosity.eta_fold
    phi = np.where(eta < ec, np.exp(-(((eta - ec) / ef) ** 2)), 1.0)
    Cl, Cq = cfg.diffusion.C_l, cfg.diffusion.C_q
    si, sj = h[i] / kscale, h[j] / kscale
    # accel viscosity
    w = (1.0 - phi)[:, None] * (dx @ M.T)
    Qi = _q(rho[i], c[i], _mu(w, dx, si), Cl, Cq)
    Qj = _q(rho[j], c[j], _mu(w, dx, sj), Cl, Cq)
    want = ((Qi + Qj) / (P[i] + P[j]))[:, None] * ap
    scale = np.max(np.abs(ap[act]))
    assert np.max(np.abs(av[act] - want[act])) < 1e-9 * scale
    far = act & (eta >= ec)
    assert far.sum() > 1000 and np.max(np.abs(av[far])) < 1e-9 * scale
    near = act & (eta < ec) & (np.abs(want).sum(1) > 1e-7 * scale)
    assert near.sum() > 100  # the taper's pairs carry viscosity
    # heating: raw velocity difference, half the sum of the U-side Q
    wU = dx @ M.T
    QiU = _q(rho[i], c[i], _mu(wU, dx, si), Cl, Cq)
    QjU = _q(rho[j], c[j], _mu(wU, dx, sj), Cl, Cq)
    G = -ap / (P[i] + P[j])[:, None]
    term = np.where(act, (P[j] + 0.5 * (QiU + QjU)) * np.einsum("ea,ea->e", wU, G), 0.0)
    du = np.bincount(i, weights=term, minlength=n)
    assert np.max(np.abs(sim.get("dudt") - du)) < 1e-9 * np.max(np.abs(du))
    # and the acceleration is the sum of both pair terms
    acc = np.stack([np.bincount(i, weights=(ap + av)[:, a] * act, minlength=n) for a in range(2)], 1)
    assert np.allclose(sim.get("dvdt").reshape(n, 2), acc, rtol=0, atol=1e-10 * np.max(np.abs(acc)))


# -------------------------------------------------------------- partition
@pytest.mark.parametrize("si,sj,dE,want", [
    (2.0, 1.0, +1.0, 1.0 / 3.0),   # heating, i has the larger s: the smaller share
    (2.0, 1.0, -1.0, 2.0 / 3.0),   # cooling, i has the larger s: the larger share
    (1.0, 2.0, +1.0, 2.0 / 3.0),
    (1.0, 2.0, -1.0, 1.0 / 3.0),
    (1.0, 2.0, 0.0, 2.0 / 3.0),    # dE = 0 counts as heating
    (3.0, 3.0, +5.0, 0.5),         # equal s
    (3.0, 3.0, -5.0, 0.5),
    (-1.0, 3.0, +1.0, 3.0 / 4.0),  # |s| enters the shares, the signed s the branch
    (0.0, 2.0, -1.0, 0.0),
])
def test_energy_partition_branches(si, sj, dE, want):
    O = _oracle()
    assert O.energy_fraction(dE, si, sj) == pytest.approx(want, rel=1e-12, abs=1e-13)
    # the pair's two shares add up to one (up to the 1e-14 guard)
    if si != sj:
        assert O.energy_fraction(dE, si, sj) + O.energy_fraction(dE, sj, si) == pytest.approx(1.0, abs=1e-12)


# ------------------------------------------------------------ Cullen-Dehnen
@pytest.mark.parametrize("k,alpha0", [(0.8, 0.05), (0.8, 1.9), (-0.5, 1.0)])
def test_cullen_dehnen_homologous_flow(k, alpha0):
    """v = -k x on an open lattice, div_prev = 0.  Interior rows: div = -2 k
    (the kernel sum of r W'(r) equals -d to the lattice's quadrature error),
    R = -sgn(k) (every neighbour's divergence has one sign), Xi = 1, A =
    max(0, 2 k / dt), v_sig = c + max(0, k) max r over the Verlet row, and
    alpha from the local value and the relaxation of CullenDehnen2010.py:281-314."""
    from warpsph_amd.configurations import toCrkConfig
    system, config, scheme = _lattice_system((40, 40), (0.0, 0.0), (1.0, 1.0), (False, False),
                                             lambda x: torch.zeros_like(x), jitter=0.0)
    O = _oracle()
    sim = O.from_system(system, config, scheme, 64)
    n = sim.n
    x = sim.get("x").reshape(n, 2)
    sim.set("v", -k * x)
    sim.set("a0", np.full(n, alpha0))
    sim.set("div", np.zeros(n))
    dt = 1e-3
    sim.evaluate(dt)
    cfg = toCrkConfig(config, scheme)
    h, c = sim.get("h"), sim.get("c")
    adj = sim.adjacency(0)
    i = np.repeat(np.arange(n), np.diff(adj["offsets"]))
    j = adj["nbr"].astype(np.int64)
    rmax = np.zeros(n)
    np.maximum.at(rmax, i, np.linalg.norm(x[i] - x[j], axis=1))
    R = np.sqrt(2.0) * h.max()
    interior = np.all((x > 2 * R) & (x < 1 - 2 * R), axis=1)
    assert interior.sum() > 300
    div = sim.get("divcd")
    assert np.allclose(div[interior], -2.0 * k, rtol=0.02)
    assert np.allclose(sim.get("Rcd")[interior], -np.sign(k), rtol=0.02)
    vsig = sim.get("vsig")
    assert np.allclose(vsig, c + max(k, 0.0) * rmax, rtol=1e-12)
    sp = scheme.viscositySwitchParams
    _, kscale = O.kernel_constants(cfg.kernel, 2)
    A = np.maximum(0.0, -div / dt)  # Xi = 1
    loc = np.clip(sp.alpha_max * h * h * A / (vsig ** 2 + h * h * A + 1e-14 * h), sp.alpha_min, sp.alpha_max)
    tau = h * (1.0 / kscale) / (0.1 * vsig)
    want = np.where(loc > alpha0, loc, alpha0 + (loc - alpha0) / tau * dt)
    want = np.clip(want, sp.alpha_min, sp.alpha_max)
    assert np.allclose(sim.get("alpha")[interior], want[interior], rtol=1e-10)
    if k > 0 and alpha0 < 0.5:
        assert np.all(sim.get("alpha")[interior] > 0.5)  # the switch opens in compression


# ------------------------------------------------------------------- Owen
@pytest.mark.parametrize("kernel", [0, 1, 2, 3])
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_owen_table_is_the_radial_shell_sum(kernel, dim):
    """psi_H(n_h) = (sum over shells eta = k/n_h < 1 of |C k'(eta)| times the
    shell count 2, 2 pi eta n_h, 4 pi eta^2 n_h^2)^(1/d), wp_psi.py:65-81;
    and psi_H / n_h approaches (int_0^1 S_d q^(d-1) |C k'(q)| dq)^(1/d)."""
    from scipy.integrate import quad
    O = _oracle()
    nh, ps = O.owen_lut(kernel, dim)
    C, _ = O.kernel_constants(kernel, dim)
    for idx in (0, 1, 777, 2048, 4095):
        n_h = nh[idx]
        assert n_h == pytest.approx(2.0 + idx * 4.0 / 4095.0, rel=1e-15)
        deta = 1.0 / n_h
        total, eta = 0.0, deta
        while eta < 1.0:
            _, dk = O.kernel_eval(kernel, dim, np.array([eta]))
            shells = 2.0 if dim == 1 else (2 * math.pi * eta / deta if dim == 2 else 4 * math.pi * eta ** 2 / deta ** 2)
            total += abs(C * dk[0]) * shells
            eta += deta
        assert ps[idx] == pytest.approx(total ** (1.0 / dim), rel=1e-13)
    S = {1: lambda q: 2.0, 2: lambda q: 2 * math.pi * q, 3: lambda q: 4 * math.pi * q * q}[dim]
    I = quad(lambda q: S(q) * abs(C * O.kernel_eval(kernel, dim, np.array([q]))[1][0]), 0.0, 1.0)[0]
    assert ps[-1] / nh[-1] == pytest.approx(I ** (1.0 / dim), rel=0.08)


@pytest.mark.parametrize("s,want", [(1.0, 1.0), (2.0, 1.45), (0.5, 0.75), (4.0, 1.0 - 0.4 * (1 + 1 / 64) * (1 - 4)),
                                    (0.25, 1.0 - 0.4 * (1 + 1 / 16) * 0.75)])
def test_owen_support_update_by_hand(s, want):
    """h' / h = 1 - a + a s with a = 0.4 (1 + s^-3) for s >= 1 and 0.4 (1 + s^2)
    below (s = n_target / n_h)."""
    O = _oracle()
    assert O.owen_new_support(4.0 * s, 4.0, 1.0) == pytest.approx(want, rel=1e-14)
    assert O.owen_new_support(4.0 * s, 4.0, 0.37) == pytest.approx(0.37 * want, rel=1e-14)


def test_owen_psi0_and_update_on_a_lattice():
    """One Owen iteration on a periodic lattice: psi0_i by brute force over
    every particle (r < h_i, minimum image), n_h by the table, h' by the
    update; and on the uniform lattice the estimated n_h is the lattice's
    H / dx within the table's shell-versus-lattice difference."""
    from warpsph_amd.configurations import toCrkConfig
    system, config, scheme = _lattice_system((24, 24), (0.0, 0.0), (1.0, 1.0), (True, True),
                                             lambda x: torch.zeros_like(x), jitter=0.3, owen=True)
    O = _oracle()
    cfg = toCrkConfig(config, scheme)
    sim = O.from_system(system, config, scheme, 64)
    n = sim.n
    x0, h0 = sim.get("x").reshape(n, 2).copy(), sim.get("h").copy()
    sim.evaluate(sim.cfl_dt())
    C, _ = O.kernel_constants(cfg.kernel, 2)
    d = x0[:, None, :] - x0[None, :, :]
    d -= np.round(d)  # unit periodic box
    r = np.sqrt((d * d).sum(-1))
    q = r / h0[:, None]
    _, dk = O.kernel_eval(cfg.kernel, 2, np.minimum(q, 1.0))
    psi0 = np.sqrt(np.where(q < 1.0, np.abs(C * dk), 0.0).sum(1))
    assert np.allclose(sim.get("psiH"), psi0, rtol=1e-12)
    nh_tab, ps_tab = O.owen_lut(cfg.kernel, 2)
    nh = np.interp(psi0, ps_tab, nh_tab)
    target = math.sqrt(cfg.targetNeighbors / math.pi)  # nH_to_n_h in 2D
    want = np.array([O.owen_new_support(target, a, b) for a, b in zip(nh, h0)])
    want = np.clip(want, 0.25 * h0.min(), 4.0 * h0.max())
    assert np.allclose(sim.get("h"), want, rtol=1e-12)
    # uniform lattice with H = 4 dx exactly
    system, config, scheme = _lattice_system((32, 32), (0.0, 0.0), (1.0, 1.0), (True, True),
                                             lambda x: torch.zeros_like(x), jitter=0.0, owen=True,
                                             targetNeighbors=16.0 * math.pi)
    sim = O.from_system(system, config, scheme, 64)
    sim.set("h", np.full(sim.n, 4.0 / 32.0))
    sim.evaluate(sim.cfl_dt())
    nh = np.interp(sim.get("psiH"), ps_tab, nh_tab)
    assert np.allclose(nh, 4.0, rtol=0.05)


# -------------------------------------------------------------------- CFL
def test_cfl_closed_form_and_limits():
    """dt = cfl h_min / (c_max xi), c = sqrt(|u| gamma (gamma - 1)), xi the
    kernel scale; clamped to [minDt, maxDt]; at most dtGrowthFactor times a
    smaller previous step."""
    from warpsph_amd.configurations import toCrkConfig
    system, config, scheme = _lattice_system((16, 16), (0.0, 0.0), (1.0, 1.0), (True, True),
                                             lambda x: torch.zeros_like(x), jitter=0.1)
    O = _oracle()
    cfg = toCrkConfig(config, scheme)
    sim = O.from_system(system, config, scheme, 64)
    n = sim.n
    g = torch.Generator().manual_seed(3)
    u = 0.5 + torch.rand(n, generator=g, dtype=torch.float64).numpy()
    h = 0.05 + 0.02 * torch.rand(n, generator=g, dtype=torch.float64).numpy()
    sim.set("u", u)
    sim.set("h", h)
    gam = float(np.float32(cfg.gamma))
    _, xi = O.kernel_constants(cfg.kernel, 2)
    cmax = math.sqrt(u.max() * gam * (gam - 1.0))
    want = float(np.float32(cfg.cflFactor)) * h.min() / (cmax * xi)
    assert want < cfg.maxDt and want > cfg.minDt
    assert sim.cfl_dt() == pytest.approx(want, rel=1e-12)
    # growth limit from a smaller previous step; none from a larger one
    prev = 0.5 * want
    assert sim.cfl_dt(prev) == pytest.approx(float(np.float32(cfg.dtGrowthFactor)) * prev, rel=1e-12)
    assert sim.cfl_dt(2.0 * want) == pytest.approx(want, rel=1e-12)
    # the sound speed of the hottest particle sets the step, not the smallest h's own
    hot = int(np.argmax(u))
    sim.set("u", np.where(np.arange(n) == hot, 4.0 * u.max(), u))
    assert sim.cfl_dt() == pytest.approx(want / 2.0, rel=1e-12)
