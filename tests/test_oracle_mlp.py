"""Pins of the MLP oracle (no GPU): its back-propagation against central
finite differences of the FP64 loss, and its BF16 rounding against the
definition of round-to-nearest-even."""
import numpy as np

from oracle import mlp_oracle as M


def test_bf16_rounding_is_nearest_even():
    # 1 + 2^-8 lies halfway between 1 and 1 + 2^-7: ties to the even mantissa (1.0)
    assert M.bf16(1.0 + 2.0 ** -8) == 1.0
    # 1 + 3 * 2^-8 lies halfway between 1 + 2^-7 and 1 + 2^-6: ties to 1 + 2^-6
    assert M.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert M.bf16(1.0 + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7
    x = np.random.default_rng(0).normal(size=1000)
    r = M.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8 + 1e-300)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = M.init_params(3).astype(np.float64)
    p[M.NH * M.NIN:M.NH * M.NIN + M.NH] = rng.normal(0, 0.1, M.NH)  # nonzero biases
    x = rng.normal(size=(37, 8))
    t = rng.uniform(size=37)
    loss, g = M.loss_grad(p, x, t, rounding=False)
    idx = rng.choice(M.NPARAM, 60, replace=False)
    idx = np.concatenate([idx, [M.NPARAM - 1, 0, M.NH * M.NIN]])
    for i in idx:
        e = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += e
        pm[i] -= e
        lp, _ = M.loss_grad(pp, x, t, rounding=False)
        lm, _ = M.loss_grad(pm, x, t, rounding=False)
        fd = (lp - lm) / (2 * e)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_rounded_forward_close_to_exact():
    rng = np.random.default_rng(2)
    p = M.init_params(5)
    x = rng.normal(size=(100, 8))
    a, _ = M.forward(p, x, rounding=True)
    b, _ = M.forward(p, x, rounding=False)
    assert np.all((a > 0) & (a < 1))
    assert np.max(np.abs(a - b)) < 2e-2
