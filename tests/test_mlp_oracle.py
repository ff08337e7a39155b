"""Pins the MLP oracle: its FP64 backward pass against central finite
differences of its own loss (the mathematics fixes the gradient)."""
import numpy as np

from oracle import mlp_oracle as mo


def test_fp64_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, mo.PARAM_COUNT)
    u = mo.unpack(p)
    u["W1"][13:] = 0.0  # padding rows, as the engine keeps them
    u["W4"][:, 7:] = 0.0
    u["b4"][7:] = 0.0
    p = mo.pack(u)
    x = rng.normal(0, 1, (37, 13))
    t = rng.normal(0, 1, (37, 7))
    loss, g = mo.loss_and_grad(p, x, t, bf16_mode=False)
    idx = rng.choice(mo.PARAM_COUNT, 300, replace=False)
    live = [i for i in idx if not (i < 1024 and i // 64 >= 13)]
    for i in live[:150]:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss_and_grad(pp, x, t, False)[0] - mo.loss_and_grad(pm, x, t, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.140625], np.float32)
    r = mo.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0            # tie -> even
    assert r[2] == 1.0 + 2 ** -6  # tie -> even (up)
    assert r[3] == -2.5 and r[4] == 3.140625
