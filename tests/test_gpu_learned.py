"""The learned viscosity switch on the GPU: the BF16 MFMA forward and
backward kernels against the numpy oracle (same BF16 rounding points, FP64
sums), Adam, the switch inside a stage (fallback and network), the features'
density and pressure gradients against an FP64 assembly over the oracle's
pairs, and training reducing the loss."""
import ctypes as C
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _model(seed=3):
    from oracle import mlp_oracle as M
    from warpsph_amd.learned import LearnedSwitch
    p = M.init_params(seed)
    rng = np.random.default_rng(seed)
    p[M.NH * M.NIN:M.NH * M.NIN + M.NH] = rng.normal(0, 0.2, M.NH).astype(np.float32)
    return LearnedSwitch(p, device="cuda"), p


@pytest.mark.parametrize("n", [1, 16, 1000, 70001])
def test_mlp_forward_matches_oracle(n):
    from oracle import mlp_oracle as M
    announce(f"mlp forward n={n}")
    model, p = _model()
    x = np.random.default_rng(n).normal(size=(n, 8)).astype(np.float32)
    a = model.forward(torch.tensor(x, device="cuda")).cpu().numpy()
    ref, _ = M.forward(p, x)
    assert np.max(np.abs(a - ref)) < 2e-3, np.max(np.abs(a - ref))


@pytest.mark.parametrize("n", [16, 1000, 50001])
def test_mlp_gradient_matches_oracle(n):
    from oracle import mlp_oracle as M
    announce(f"mlp gradient n={n}")
    model, p = _model(5)
    rng = np.random.default_rng(n)
    x = rng.normal(size=(n, 8)).astype(np.float32)
    t = rng.uniform(size=n).astype(np.float32)
    g, loss = model.grad(torch.tensor(x, device="cuda"), torch.tensor(t, device="cuda"))
    g = g.cpu().numpy().astype(np.float64)
    lref, gref = M.loss_grad(p, x, t)
    assert abs(float(loss) - lref) < 1e-5 * max(1.0, lref)
    # blockwise normwise agreement: W1, b1, W2, b2, w3, b3
    H, I = M.NH, M.NIN
    cuts = np.cumsum([0, H * I, H, H * H, H, H, 1])
    for a, b in zip(cuts[:-1], cuts[1:]):
        err = np.linalg.norm(g[a:b] - gref[a:b]) / max(np.linalg.norm(gref[a:b]), 1e-12)
        assert err < 1e-2, (a, b, err)
    # determinism: the same batch gives the same bits
    g2, _ = model.grad(torch.tensor(x, device="cuda"), torch.tensor(t, device="cuda"))
    assert np.array_equal(g2.cpu().numpy(), g.astype(np.float32))


def test_adam_step_matches_oracle():
    from oracle import mlp_oracle as M
    from warpsph_amd.learned import Adam
    announce("adam")
    model, p = _model(7)
    rng = np.random.default_rng(0)
    opt = Adam(model, lr=1e-2)
    m = np.zeros(M.NPARAM)
    v = np.zeros(M.NPARAM)
    pr = p.astype(np.float64)
    for step in range(1, 4):
        g = rng.normal(size=M.NPARAM).astype(np.float32)
        opt.step(torch.tensor(g, device="cuda"))
        pr, m, v = M.adam(pr, m, v, g.astype(np.float64), 1e-2, 0.9, 0.999, 1e-8, step)
    got = model.params.cpu().numpy()
    assert np.max(np.abs(got - pr)) < 1e-5


def test_training_reduces_loss():
    from warpsph_amd.learned import fit
    announce("fit")
    rng = np.random.default_rng(1)
    X = rng.normal(size=(40000, 8)).astype(np.float32)
    t = 1.0 / (1.0 + np.exp(-(1.5 * X[:, 0] - X[:, 3] + 0.5 * X[:, 5] * X[:, 6])))
    model, hist = fit(torch.tensor(X, device="cuda"), torch.tensor(t.astype(np.float32), device="cuda"), epochs=15,
                      batch=4096, lr=3e-3)
    print(hist)
    assert hist[-1] < 0.25 * hist[0]


def _stage(switch_model=None, fallback_all=False):
    from warpsph_amd import sample
    from warpsph_amd.enumTypes import ViscositySwitch
    from warpsph_amd.scheme import crkSPH_step
    system, config, scheme = sample.shear_layer_2d(nx=48, device="cuda")
    scheme.alphaWeightedViscosity = True
    dt = sample.cfl_dt(system.state, config, scheme) * (2.0 / 3.0)
    if switch_model is not None:
        scheme.viscositySwitchParams.scheme = ViscositySwitch.Learned
        scheme.viscositySwitchParams.model = switch_model
    update, adj, st = crkSPH_step(system, dt, config, scheme)
    torch.cuda.synchronize()
    return update, st, system, config, scheme, dt


def test_learned_switch_falls_back_to_cullen_dehnen():
    """Features outside the model's range everywhere: the stage equals the
    Cullen-Dehnen stage bit for bit."""
    announce("learned switch fallback")
    model, _ = _model()
    model.lo[:] = np.inf
    model.hi[:] = np.inf
    u_cd, st_cd, *_ = _stage()
    u_l, st_l, *_ = _stage(model)
    assert torch.equal(st_cd.alphas, st_l.alphas)
    assert torch.equal(u_cd.dvdt, u_l.dvdt) and torch.equal(u_cd.dudt, u_l.dudt)


@pytest.mark.parametrize("multiply", [False, True])
def test_learned_switch_uses_the_network(multiply):
    """Every feature in range: alpha = scale MLP((features - mu) / sigma),
    times the Cullen-Dehnen alpha of the same stage when the model
    multiplies, checked against the oracle's network on the exported
    features."""
    from oracle import mlp_oracle as M
    from warpsph_amd import abi
    from warpsph_amd.configurations import toCrkConfig
    from warpsph_amd.runtime import runtime
    announce(f"learned switch network multiply={multiply}")
    model, p = _model(11)
    model.mu[:] = 0.0
    model.sigma[:] = 2.0
    model.scale, model.multiply = (2.0, True) if multiply else (1.0, False)
    u_l, st_l, system, config, scheme, dt = _stage(model)
    rt = runtime("cuda:0")
    n = st_l.numParticles
    feats = torch.empty((n, 8), dtype=torch.float32, device="cuda")
    abi.call("crk_learned_features", rt.ctx, C.byref(toCrkConfig(config, scheme)), float(dt), abi.ptr(feats),
             rt.stream())
    f = feats.cpu().numpy()
    assert np.all(np.isfinite(f))
    ref, _ = M.forward(p, (f - model.mu) / model.sigma)
    u_cd, st_cd, *_ = _stage()
    if multiply:
        a_cd = st_cd.alphas.cpu().numpy().astype(np.float64)
        ref = 2.0 * ref * a_cd
        tol = 2e-3 * 2.0 * max(1.0, float(np.max(a_cd)))
    else:
        tol = 2e-3
    got = st_l.alphas.cpu().numpy()
    assert np.max(np.abs(got - ref)) < tol
    # and it changes the dynamics against the Cullen-Dehnen stage
    assert not torch.equal(u_cd.dvdt, u_l.dvdt)


def test_constant_network_scales_the_cullen_dehnen_switch():
    """LearnedSwitch.constant(m): alpha = m alpha_CD on every particle."""
    from warpsph_amd.learned import LearnedSwitch
    announce("constant multiplier network")
    u_cd, st_cd, *_ = _stage()
    for m in (0.5, 1.0, 1.5):
        _, st_m, *_ = _stage(LearnedSwitch.constant(m, device="cuda"))
        a_cd = st_cd.alphas.cpu().numpy().astype(np.float64)
        got = st_m.alphas.cpu().numpy().astype(np.float64)
        # zero weights: the output is the FP32 bias alone, sigmoid(logit(m / 2))
        assert np.max(np.abs(got - m * a_cd)) <= 1e-5 * m * np.max(a_cd), m


def test_feature_gradients_against_fp64_assembly():
    """h |grad rho| / rho and h |grad P| / P of the features against the CRK
    gradient assembled in numpy over the FP64 oracle's pairs and factors."""
    from oracle import oracle as O
    from tests.test_oracle_pins import _crk_gradients
    from tests.util import cpu_copy
    from warpsph_amd import abi, sample
    from warpsph_amd.configurations import toCrkConfig
    from warpsph_amd.runtime import runtime
    from warpsph_amd.scheme import crkSPH_step
    announce("feature gradients")
    system, config, scheme = sample.shear_layer_2d(nx=40, device="cuda")
    cpu = cpu_copy(system, config, scheme)
    dt = sample.cfl_dt(system.state, config, scheme) * (2.0 / 3.0)
    crkSPH_step(system, dt, config, scheme)
    rt = runtime("cuda:0")
    n = system.state.numParticles
    feats = torch.empty((n, 8), dtype=torch.float32, device="cuda")
    cfg = toCrkConfig(config, scheme)
    abi.call("crk_learned_features", rt.ctx, C.byref(cfg), float(dt), abi.ptr(feats), rt.stream())
    f = feats.cpu().numpy().astype(np.float64)
    sim = O.from_system(cpu, config, scheme, 64)
    sim.evaluate(dt)
    L = (config.domain.max - config.domain.min).cpu().numpy()
    i, j, dx, WR, gWR = _crk_gradients(sim, cfg.kernel, 2, L, (True, True))
    V, rho, P, h = sim.get("V"), sim.get("rho"), sim.get("P"), sim.get("h")
    inside = np.linalg.norm(dx, axis=1) < h[j]
    gr = np.zeros((n, 2))
    gp = np.zeros((n, 2))
    np.add.at(gr, i[inside], (V[j] * (rho[j] - rho[i]))[inside, None] * gWR[inside])
    np.add.at(gp, i[inside], (V[j] * (P[j] - P[i]))[inside, None] * gWR[inside])
    f5 = h * np.linalg.norm(gr, axis=1) / rho
    f6 = h * np.linalg.norm(gp, axis=1) / P
    assert np.max(np.abs(f[:, 5] - f5)) <= 1e-4 * max(1.0, np.max(np.abs(f5)))
    assert np.max(np.abs(f[:, 6] - f6)) <= 1e-4 * max(1.0, np.max(np.abs(f6)))
    assert np.max(f5) > 0.1  # the shear layer's density jumps are seen
    # the neighbour-count feature: in-support pairs / n_H
    assert np.all(f[:, 7] > 0.3) and np.all(f[:, 7] < 3.0)


def test_learned_switch_gain_on_held_out_pairs():
    """What the trained network gains: train on solver-produced pairs of the
    two-state generator (64^2 / 128^2, three seeds), then run four held-out
    seeds' coarse states with the learned and with the Cullen-Dehnen switch
    and compare the density field (SPH densities averaged per cell of a 16^2
    grid) against each seed's 4x-resolution run at t = 0.09.  The network
    rescales the Cullen-Dehnen alpha; it must lower the error on every
    held-out seed and by more than 3 % on average."""
    from warpsph_amd import learned as Lm
    from warpsph_amd.ensemble import two_state_run
    from warpsph_amd.enumTypes import ViscositySwitch
    announce("learned gain")
    feats, targs = [], []
    for k in range(3):
        f, t = Lm.pair_samples(2000 + k, nx=64, snapshots=(0.03, 0.06, 0.09), grid=16)
        feats.append(f)
        targs.append(t)
    model, hist = Lm.fit(torch.cat(feats), torch.cat(targs), epochs=40, batch=4096, lr=3e-3)
    assert model.multiply and model.scale == 2.0

    def l1(a, b):
        return float((a - b).abs().mean() / b.abs().mean())

    gains, used = [], []
    for seed in (2999, 2998, 2997, 2996):
        coarse, config, scheme = two_state_run(64, seed, device="cuda")
        fine, _, _ = two_state_run(128, seed, device="cuda")
        res = {}
        for name, sw, mdl, sysm in (("fine", ViscositySwitch.CullenDehnen2010, None, fine),
                                    ("cd", ViscositySwitch.CullenDehnen2010, None, coarse),
                                    ("learned", ViscositySwitch.Learned, model, coarse)):
            sc = Lm._with_switch(scheme, sw)
            sc.viscositySwitchParams.model = mdl
            s = Lm.advance(sysm.initializeNewState(), config, sc, 0.09)
            res[name] = Lm.density_field(s.state, 16)[0]
            if mdl is not None:
                f = Lm.stage_features(s, config, sc)
                lo = torch.as_tensor(mdl.lo, device=f.device)
                hi = torch.as_tensor(mdl.hi, device=f.device)
                used.append(float(((f >= lo) & (f <= hi)).all(dim=1).float().mean()))
        l1_cd, l1_ml = l1(res["cd"], res["fine"]), l1(res["learned"], res["fine"])
        gains.append((l1_cd - l1_ml) / l1_cd)
        print(f"seed {seed}: density L1 vs 4x: Cullen-Dehnen {l1_cd:.5f}, learned {l1_ml:.5f}, "
              f"gain {gains[-1]:+.1%}", flush=True)
    mean_gain = sum(gains) / len(gains)
    print(f"loss {hist[0]:.4f} -> {hist[-1]:.4f}; mean gain {mean_gain:+.1%}; network used on "
          f"{min(used):.1%} .. {max(used):.1%} of the particles", flush=True)
    assert hist[-1] < hist[0]
    assert min(used) > 0.9
    assert all(g > 0.0 for g in gains), gains
    assert mean_gain > 0.03, mean_gain
