"""The learned viscosity switch: model object, training on the GPU, and the
solver-produced training data.

No counterpart in the reference: a per-particle MLP (8 -> 64 -> 64 -> 1,
SiLU, BF16 on the matrix cores, ``csrc/mlp.hip``) gives the viscosity alpha in
[0, 1] in place of the Cullen-Dehnen switch (CullenDehnen2010.py:218-331);
particles whose features leave the training range keep the Cullen-Dehnen
alpha.  alpha reaches the dynamics through alpha-weighted Q
(``CRKSPHConfig.alphaWeightedViscosity``).

The network rescales the Cullen-Dehnen switch: alpha = alpha_CD * scale *
sigmoid(MLP(x)) with scale = 2, so an output of 1/2 is the plain switch and
the network can lower the viscosity where the plain switch over-dissipates
and raise it where it under-dissipates (``multiply``; the absolute form
alpha = scale * sigmoid(MLP(x)) is kept for models trained that way).

Training targets come from the solver itself.  For a coarse run and the same
initial state at 4x the particle count (the config-4 generator's two-state
runs): the coarse run is repeated with the Cullen-Dehnen alpha scaled by
constant factors (the learned switch with constant networks); at snapshot
times the density of each run is binned on a uniform grid, and a coarse
particle's target is the soft arg-min over the factors of the grid cell's
density error against the fine run.  The features are those of the plain
(Cullen-Dehnen) coarse run at the same snapshot.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import abi
from .runtime import runtime

FEATURES = ("h div/c", "h^2 ddiv/dt / c^2", "h |curl v|/c", "h (rho/m)^(1/d)", "c/v_sig", "h |grad rho|/rho",
            "h |grad P|/P", "neighbours/n_H")


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initial weights, zero biases (FP32)."""
    rng = np.random.default_rng(seed)
    H, I = abi.MLP_HIDDEN, abi.MLP_IN
    W1 = rng.normal(0.0, math.sqrt(2.0 / I), (H, I))
    W2 = rng.normal(0.0, math.sqrt(2.0 / H), (H, H))
    w3 = rng.normal(0.0, math.sqrt(1.0 / H), H)
    return np.concatenate([W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), w3, [0.0]]).astype(np.float32)


class LearnedSwitch:
    """Parameters (FP32 master and BF16 operand copies, on the GPU) and the
    feature normalisation / range of the training data."""

    def __init__(self, params=None, mu=None, sigma=None, lo=None, hi=None, device="cuda", scale=2.0,
                 multiply=True):
        self.device = torch.device(device)
        self.scale = float(scale)
        self.multiply = bool(multiply)
        p = init_params() if params is None else np.asarray(params, dtype=np.float32)
        if p.shape != (abi.MLP_NPARAM,):
            raise ValueError(f"expected {abi.MLP_NPARAM} parameters, got {p.shape}")
        self.params = torch.tensor(p, device=self.device)
        self.wbf16 = torch.empty(abi.MLP_NBF16, dtype=torch.int16, device=self.device)
        self.mu = np.zeros(8, np.float32) if mu is None else np.asarray(mu, np.float32)
        self.sigma = np.ones(8, np.float32) if sigma is None else np.asarray(sigma, np.float32)
        self.lo = np.full(8, -np.inf, np.float32) if lo is None else np.asarray(lo, np.float32)
        self.hi = np.full(8, np.inf, np.float32) if hi is None else np.asarray(hi, np.float32)
        self.pack()

    def pack(self):
        rt = runtime(self.device)
        abi.call("crk_mlp_pack", abi.ptr(self.params), abi.ptr(self.wbf16), rt.stream())

    def cabi(self) -> abi.CrkMlp:
        m = abi.CrkMlp()
        m.params = self.params.data_ptr()
        m.wbf16 = self.wbf16.data_ptr()
        for k in ("mu", "sigma", "lo", "hi"):
            getattr(m, k)[:] = [float(v) for v in getattr(self, k)]
        m.scale = self.scale
        m.multiply = int(self.multiply)
        self._c = m  # kept alive while the caller uses the pointer
        return m

    def normalise(self, feats: torch.Tensor) -> torch.Tensor:
        mu = torch.as_tensor(self.mu, device=feats.device)
        sg = torch.as_tensor(self.sigma, device=feats.device)
        return ((feats - mu) / sg).contiguous()

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """alpha for normalised features x [n][8] (GPU)."""
        x = x.to(torch.float32).contiguous()
        out = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
        rt = runtime(self.device)
        abi.call("crk_mlp_forward", C.byref(self.cabi()), int(x.shape[0]), abi.ptr(x), abi.ptr(out), rt.stream())
        return out

    def grad(self, x: torch.Tensor, t: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
        """(gradient [NPARAM], loss) of mean (alpha - t)^2 on the GPU."""
        x = x.to(torch.float32).contiguous()
        t = t.to(torch.float32).contiguous()
        g = torch.empty(abi.MLP_NPARAM + 1, dtype=torch.float32, device=x.device)
        rt = runtime(self.device)
        abi.call("crk_mlp_grad", rt.ctx, C.byref(self.cabi()), int(x.shape[0]), abi.ptr(x), abi.ptr(t), abi.ptr(g),
                 rt.stream())
        return g[:abi.MLP_NPARAM], g[abi.MLP_NPARAM]

    def save(self, path: str):
        np.savez(path, params=self.params.cpu().numpy(), mu=self.mu, sigma=self.sigma, lo=self.lo, hi=self.hi,
                 scale=np.float32(self.scale), multiply=np.int32(self.multiply))

    @classmethod
    def load(cls, path: str, device="cuda") -> "LearnedSwitch":
        with np.load(path, allow_pickle=False) as z:
            scale = float(z["scale"]) if "scale" in z.files else 1.0
            multiply = bool(z["multiply"]) if "multiply" in z.files else False
            return cls(z["params"], z["mu"], z["sigma"], z["lo"], z["hi"], device=device, scale=scale,
                       multiply=multiply)

    @classmethod
    def constant(cls, value: float, scale: float = 2.0, multiply: bool = True, device="cuda") -> "LearnedSwitch":
        """A network whose output is the constant value / scale for every
        input (zero weights, output bias logit(value / scale)) and that
        accepts every finite feature: alpha = value * alpha_CD (multiply) or
        alpha = value."""
        s = float(value) / float(scale)
        if not 0.0 < s < 1.0:
            raise ValueError(f"constant {value} outside (0, {scale})")
        p = np.zeros(abi.MLP_NPARAM, np.float32)
        p[-1] = math.log(s / (1.0 - s))
        inf = np.full(8, np.inf, np.float32)
        return cls(p, None, None, -inf, inf, device=device, scale=scale, multiply=multiply)


class Adam:
    """Adam on the model's FP32 parameters (crk_mlp_adam), BF16 copies repacked."""

    def __init__(self, model: LearnedSwitch, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.model = model
        self.m1 = torch.zeros_like(model.params)
        self.m2 = torch.zeros_like(model.params)
        self.lr, self.b1, self.b2, self.eps = lr, beta1, beta2, eps
        self.t = 0

    def step(self, grad: torch.Tensor):
        self.t += 1
        rt = runtime(self.model.device)
        abi.call("crk_mlp_adam", abi.MLP_NPARAM, abi.ptr(self.model.params), abi.ptr(self.m1), abi.ptr(self.m2),
                 abi.ptr(grad), self.lr, self.b1, self.b2, self.eps, self.t, rt.stream())
        self.model.pack()


def fit(features: torch.Tensor, targets: torch.Tensor, epochs: int = 20, batch: int = 8192, lr: float = 2e-3,
        seed: int = 0, model: Optional[LearnedSwitch] = None, verbose: bool = False, scale: float = 2.0,
        multiply: bool = True):
    """Train on raw features [N][8] and targets [N] in [0, 1] (the network's
    sigmoid output; GPU tensors).  The normalisation is the data's mean and
    standard deviation; the range [lo, hi] is the 0.1 % .. 99.9 % quantile
    span widened by 5 % on both sides.  Returns (model, per-epoch mean loss)."""
    X = features.to(torch.float32)
    T = targets.to(torch.float32)
    ok = torch.isfinite(X).all(dim=1) & torch.isfinite(T)
    X, T = X[ok], T[ok]
    Xc = X.cpu().numpy().astype(np.float64)
    mu = Xc.mean(axis=0)
    sg = np.maximum(Xc.std(axis=0), 1e-6)
    qlo, qhi = np.quantile(Xc, 0.001, axis=0), np.quantile(Xc, 0.999, axis=0)
    span = qhi - qlo
    lo, hi = qlo - 0.05 * span, qhi + 0.05 * span
    if model is None:
        model = LearnedSwitch(init_params(seed), mu, sg, lo, hi, device=X.device, scale=scale, multiply=multiply)
    else:
        model.mu, model.sigma, model.lo, model.hi = (np.asarray(a, np.float32) for a in (mu, sg, lo, hi))
    # train on the in-range samples (the others fall back to Cullen-Dehnen at run time)
    inr = torch.ones(X.shape[0], dtype=torch.bool, device=X.device)
    for q in range(8):
        inr &= (X[:, q] >= float(model.lo[q])) & (X[:, q] <= float(model.hi[q]))
    Xn = model.normalise(X[inr])
    Tn = T[inr].contiguous()
    opt = Adam(model, lr=lr)
    g = torch.Generator(device="cpu").manual_seed(seed)
    hist = []
    n = Xn.shape[0]
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g).to(X.device)
        tot, cnt = 0.0, 0
        for b0 in range(0, n, batch):
            idx = perm[b0:b0 + batch]
            grad, loss = model.grad(Xn[idx], Tn[idx])
            opt.step(grad)
            tot += float(loss) * idx.numel()
            cnt += idx.numel()
        hist.append(tot / max(cnt, 1))
        if verbose:
            print(f"epoch {ep}: loss {hist[-1]:.5f}", flush=True)
    return model, hist


# ------------------------------------------------------------ training data
def grid_density(positions: torch.Tensor, masses: torch.Tensor, G: int, lo=(0.0, 0.0), hi=(1.0, 1.0)):
    """Mass per cell / cell area on a G x G grid (2D)."""
    L = torch.tensor([hi[0] - lo[0], hi[1] - lo[1]], dtype=torch.float64, device=positions.device)
    lo_t = torch.tensor(lo, dtype=torch.float64, device=positions.device)
    c = torch.floor((positions.double() - lo_t) / L * G).long().clamp_(0, G - 1)
    idx = c[:, 1] * G + c[:, 0]
    rho = torch.zeros(G * G, dtype=torch.float64, device=positions.device)
    rho.index_add_(0, idx, masses.double())
    return (rho / (float(L[0] * L[1]) / (G * G))).view(G, G), idx


def density_field(state, G: int, lo=(0.0, 0.0), hi=(1.0, 1.0)):
    """The particles' SPH densities averaged per cell of a G x G grid, mass
    weighted (2D): unlike the binned mass per cell it changes continuously
    with the trajectory, so small changes of the dynamics are seen.  Returns
    (field [G][G], cell index of each particle)."""
    msum, idx = grid_density(state.positions, state.masses, G, lo, hi)
    w = torch.zeros(G * G, dtype=torch.float64, device=state.positions.device)
    w.index_add_(0, idx, state.masses.double() * state.densities.double())
    area = (hi[0] - lo[0]) * (hi[1] - lo[1]) / (G * G)
    m = (msum * area).view(-1)
    return (w / torch.where(m > 0, m, torch.ones_like(m))).view(G, G), idx


def advance(system, config, scheme, t_target: float, integ=None):
    """Step with the CFL dt until t_target (the last step shortened to land on it)."""
    from . import sample
    from .integrator import getIntegrator
    from .scheme import crkSPH_step
    integ = integ or getIntegrator()
    s = system
    prev = None
    while s.t < t_target - 1e-12:
        dt = sample.cfl_dt(s.state, config, scheme, prev)
        prev = dt
        dt = min(dt, t_target - s.t)
        s = integ.function(s, crkSPH_step, dt, config, scheme).state
    return s


def stage_features(system, config, scheme) -> torch.Tensor:
    """Raw features [n][8] (user order) of one derivative evaluation of the
    current state (the Cullen-Dehnen or learned switch must be configured)."""
    from . import sample
    from .configurations import toCrkConfig
    from .scheme import crkSPH_step
    s = system.initializeNewState()
    s.adjacency = system.adjacency
    dt = sample.cfl_dt(s.state, config, scheme)
    crkSPH_step(s, dt, config, scheme)
    n = s.state.numParticles
    out = torch.empty((n, 8), dtype=torch.float32, device=s.state.positions.device)
    rt = runtime(s.state.positions.device)
    cfg = toCrkConfig(config, scheme)
    abi.call("crk_learned_features", rt.ctx, C.byref(cfg), float(dt), abi.ptr(out), rt.stream())
    return out


def _with_switch(scheme, switch, alpha_weighted=True):
    import copy
    from .enumTypes import ViscositySwitch
    sc = copy.deepcopy(scheme)
    sc.viscositySwitchParams.scheme = switch
    sc.alphaWeightedViscosity = alpha_weighted
    return sc


def pair_samples(seed: int, nx: int = 128, snapshots: Sequence[float] = (0.02, 0.04, 0.06),
                 alphas: Sequence[float] = (0.25, 0.5, 0.75, 1.0, 1.5, 1.9), grid: int = 32, tau: float = 0.002,
                 device="cuda", scale: float = 2.0, multiply: bool = True):
    """Training samples from one coarse / 4x-resolution pair of the two-state
    generator: (features [M][8], targets [M] in [0, 1]).  ``alphas`` are the
    constant factors on the Cullen-Dehnen alpha (multiply) or the constant
    alphas; targets are the soft arg-min divided by ``scale``.  A candidate
    run that fails (its supports outgrow the periodic box when too little
    viscosity lets a strong shock scatter particles) leaves the candidate
    set from then on."""
    from .ensemble import two_state_run
    from .enumTypes import ViscositySwitch
    coarse, config, scheme = two_state_run(nx, seed, device=device)
    fine, _, _ = two_state_run(2 * nx, seed, device=device)
    cd = _with_switch(scheme, ViscositySwitch.CullenDehnen2010)
    runs_a, schemes_a, values = [], [], []
    for a in alphas:
        s = coarse.initializeNewState()
        if multiply:
            sc = _with_switch(scheme, ViscositySwitch.Learned)
            sc.viscositySwitchParams.model = LearnedSwitch.constant(a, scale, True, device=device)
        else:
            sc = _with_switch(scheme, ViscositySwitch.NoneSwitch)
            s.state.alphas.fill_(float(a))
        runs_a.append(s)
        schemes_a.append(sc)
        values.append(float(a))
    s_cd, s_f = coarse.initializeNewState(), fine.initializeNewState()
    feats, targets = [], []
    for ts in snapshots:
        s_cd = advance(s_cd, config, cd, ts)
        s_f = advance(s_f, config, cd, ts)
        alive = []
        for s, sc, a in zip(runs_a, schemes_a, values):
            try:
                alive.append((advance(s, config, sc, ts), sc, a))
            except ValueError:
                pass
        if not alive:
            break
        runs_a, schemes_a, values = (list(c) for c in zip(*alive))
        rho_f, _ = density_field(s_f.state, grid)
        errs = []
        for s in runs_a:
            rho_a, _ = density_field(s.state, grid)
            errs.append((rho_a - rho_f).abs() / rho_f.mean())
        E = torch.stack(errs)  # [n_alpha, G, G]
        w = torch.softmax(-E / tau, dim=0)  # soft arg-min over the constant alphas (tau: relative error)
        a_cell = (w * torch.tensor(values, dtype=torch.float64, device=E.device)[:, None, None]).sum(dim=0) / scale
        f = stage_features(s_cd, config, cd)
        _, idx = density_field(s_cd.state, grid)
        feats.append(f)
        targets.append(a_cell.view(-1)[idx].to(torch.float32))
    if not feats:
        raise RuntimeError(f"pair_samples: every candidate run of seed {seed} failed")
    return torch.cat(feats), torch.cat(targets)
