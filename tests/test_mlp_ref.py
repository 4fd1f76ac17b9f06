"""Pins the CPU reference of the learned-bracket network by central differences (CPU)."""
import numpy as np

import _mlp_ref as R


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    w = R.init(1)
    x = rng.normal(size=(40, 4))
    mu = rng.uniform(0.01, 0.99, 40)
    mean, std = (0.5, 0.0, -0.3, 0.0), (1.5, 1.0, 2.0, 1.0)
    L, g = R.loss_and_grad(w, x, mu, 0.3, mean, std, bf=False)
    # the o2 target depends on o1 but is held fixed: freeze it for the differences
    o, _ = R.forward(w, x, bf=False)
    z0 = R.prior_z0(x, mean, std)
    t1, t2 = R.targets(o[:, 0], mu, z0)

    def loss(wv):
        oo, _ = R.forward(wv, x, bf=False)
        return np.mean((oo[:, 0] - t1) ** 2 + 0.3 * (oo[:, 1] - t2) ** 2)

    assert abs(loss(w) - L) < 1e-12
    idx = np.concatenate([rng.choice(R.NPARAM, 300, replace=False),
                          np.arange(R.NPARAM - 130, R.NPARAM)])
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (loss(wp) - loss(wm)) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * (1 + abs(fd)), (i, fd, g[i])


def test_prior_tracks_hW():
    # the energy equation (tau + D)/D = hW - P/(rho W) + b^2 (1 + v^2)/2 - (b.v)^2/2 in a
    # flat frame makes z0 = 1 + q - b^2/2 = hW - P/(rho W) + (b^2 v^2 - (b.v)^2)/2:
    # never below hW/Gamma, never below 1, equal to hW - P/(rho W) for b = 0
    rng = np.random.default_rng(9)
    n, g = 20000, 4 / 3
    rho = 10 ** rng.uniform(-6, 2, n)
    P = rho * 10 ** rng.uniform(-6, 2, n)
    W = rng.uniform(1, 10, n)
    v = np.sqrt(1 - 1 / W ** 2)[:, None] * _unit(rng, n)
    B = np.sqrt(rho * 10 ** rng.uniform(-4, 2, n))[:, None] * _unit(rng, n)
    h = 1 + g / (g - 1) * P / rho
    D = rho * W
    vB = np.sum(v * B, axis=1)
    B2 = np.sum(B * B, axis=1)
    v2 = np.sum(v * v, axis=1)
    tau = rho * h * W ** 2 - P + 0.5 * B2 * (1 + v2) - 0.5 * vB ** 2 - D
    q, bsq = tau / D, B2 / D
    z0 = 1 + q - 0.5 * bsq
    expect = h * W - P / (rho * W) + 0.5 * (B2 * v2 - vB ** 2) / D
    np.testing.assert_allclose(z0, expect, rtol=1e-9)
    assert np.all(z0 >= h * W / g * (1 - 1e-9))
    assert np.all(z0 >= 1 - 1e-9)


def _unit(rng, n):
    u = rng.normal(size=(n, 3))
    return u / np.linalg.norm(u, axis=1, keepdims=True)


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.00390625, 1.005859375, 3.14159265, -2.5e-3, 65504.0], np.float32)
    r = R.bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to the even mantissa (1.0)
    assert r[1] == 1.0
    assert r[2] == 1.0078125
    assert abs(r[3] - 3.140625) < 1e-12
    assert np.all(np.abs(r - vals) <= np.abs(vals) * 2 ** -8)
