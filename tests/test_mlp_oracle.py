"""CPU: the surrogate-MLP oracle's backward pass equals finite differences of its loss (FP64,
no rounding), and its BF16 rounding helper rounds to nearest even."""
import numpy as np

from _mlp_oracle import MLPOracle, bf16


def test_bf16_rounding_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38], np.float32)
    r = bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                       # halfway -> even (mantissa ...0)
    assert r[2] == np.float32(1.0 + 2 ** -7)  # representable
    assert r[3] == np.float32(1.0 + 2 ** -6)  # halfway -> even (up)
    assert r[4] == -2.5
    assert np.isfinite(r[5])


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    din, w, dout, B = 7, 6, 5, 4
    layers = [(rng.normal(size=(w, din)) * 0.5, rng.normal(size=w) * 0.1),
              (rng.normal(size=(w, w)) * 0.5, rng.normal(size=w) * 0.1),
              (rng.normal(size=(dout, w)) * 0.5, rng.normal(size=dout) * 0.1)]
    mean, scale, s = rng.normal(size=din), rng.uniform(0.5, 2, din), 3.0
    x, t, wt = rng.normal(size=(B, din)), rng.normal(size=(B, dout)), rng.uniform(0.5, 1.5, dout)
    o = MLPOracle(layers, mean, scale, s, exact=True, out_mean=0.3)
    loss, g = o.grads(x, t, wt)
    eps = 1e-6
    for i, (W, b) in enumerate(layers):
        for (r, c) in [(0, 0), (W.shape[0] - 1, W.shape[1] - 1), (1, 2)]:
            Wp, Wm = W.copy(), W.copy()
            Wp[r, c] += eps
            Wm[r, c] -= eps
            lp = MLPOracle([(Wp, b) if k == i else layers[k] for k in range(3)], mean, scale, s,
                           exact=True, out_mean=0.3).grads(x, t, wt)[0]
            lm = MLPOracle([(Wm, b) if k == i else layers[k] for k in range(3)], mean, scale, s,
                           exact=True, out_mean=0.3).grads(x, t, wt)[0]
            assert abs((lp - lm) / (2 * eps) - g[i][0][r, c]) < 1e-8 * max(1.0, abs(g[i][0][r, c]))
        bp, bm = b.copy(), b.copy()
        bp[0] += eps
        bm[0] -= eps
        lp = MLPOracle([(W, bp) if k == i else layers[k] for k in range(3)], mean, scale, s,
                       exact=True, out_mean=0.3).grads(x, t, wt)[0]
        lm = MLPOracle([(W, bm) if k == i else layers[k] for k in range(3)], mean, scale, s,
                       exact=True, out_mean=0.3).grads(x, t, wt)[0]
        assert abs((lp - lm) / (2 * eps) - g[i][1][0]) < 1e-8 * max(1.0, abs(g[i][1][0]))
