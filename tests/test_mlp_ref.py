"""Pins the CPU reference of the learned-bracket network by central differences (CPU)."""
import numpy as np

import _mlp_ref as R


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    w = R.init(1)
    x = rng.normal(size=(40, 4))
    mu = rng.uniform(0.01, 0.99, 40)
    L, g = R.loss_and_grad(w, x, mu, 0.3, bf=False)
    # the o2 target depends on o1 but is held fixed: freeze it for the differences
    o, _ = R.forward(w, x, bf=False)
    c = 1 / (1 + np.exp(-o[:, 0]))
    t2 = np.log(np.abs(c - mu) + 1e-12)
    t1 = np.log(mu / (1 - mu))

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


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.00390625, 1.005859375, 3.14159265, -2.5e-3, 65504.0], np.float32)
    r = R.bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to the even mantissa (1.0)
    assert r[1] == 1.0
    assert r[2] == 1.0078125
    assert abs(r[3] - 3.140625) < 1e-12
    assert np.all(np.abs(r - vals) <= np.abs(vals) * 2 ** -8)
