"""Pins of the network oracle (oracle/mlp_ref.py): its gradient is the exact derivative
of its FP64 forward (central differences), its bfloat16 rounding is round-to-nearest-
even, and the flat parameter order matches the C ABI's parameter count."""
import numpy as np

from oracle import mlp_ref as R


def test_bf16_round_ties_to_even_and_specials():
    one = 1.0
    assert R.bf16_round(np.array([one]))[0] == 1.0
    assert R.bf16_round(np.array([1.0 + 2.0 ** -8]))[0] == 1.0            # tie -> even (1.0)
    assert R.bf16_round(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6   # tie -> even (up)
    assert R.bf16_round(np.array([1.0 + 2.0 ** -8 + 2.0 ** -12]))[0] == 1.0 + 2.0 ** -7
    assert np.isnan(R.bf16_round(np.array([np.nan]))[0])
    assert R.bf16_round(np.array([-3.0]))[0] == -3.0
    x = np.random.default_rng(0).normal(size=1000)
    r = R.bf16_round(x)
    assert np.all(np.abs(r - x) <= 2.0 ** -8 * np.abs(x))


def test_parameter_count_matches_design():
    assert R.layout(24, 22)[1] == 39062     # drm19: T, log p, 21 log10 Y, log dt -> dT, 21 dY
    assert R.layout(56, 54)[1] == 56 * 128 + 128 + 2 * (128 * 128 + 128) + 128 * 54 + 54


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    n_in, n_out, B = 7, 5, 6
    _, n = R.layout(n_in, n_out)
    p = rng.normal(size=n) * 0.2
    x = rng.normal(size=(B, n_in))
    t = rng.normal(size=(B, n_out))
    _, g = R.loss_and_grad(p, x, t)
    idx = rng.choice(n, size=60, replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (R.loss_and_grad(pp, x, t)[0] - R.loss_and_grad(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(g[i])), (i, fd, g[i])


def test_bf16_gradient_is_close_to_exact():
    rng = np.random.default_rng(2)
    _, n = R.layout(24, 22)
    p = rng.normal(size=n) * 0.05
    x = rng.normal(size=(64, 24))
    t = rng.normal(size=(64, 22))
    _, g = R.loss_and_grad(p, x, t)
    _, gb = R.loss_and_grad(p, x, t, bf16=True)
    cos = g @ gb / np.linalg.norm(g) / np.linalg.norm(gb)
    assert cos > 0.999
