"""CPU oracle of the surrogate MLP (test infrastructure only).

A numpy restatement of what libmemb's kernels compute: inputs normalised and rounded to BF16,
weights rounded to BF16, every GEMM accumulated in FP32 from BF16 operands, hidden
activations H = bf16(silu(Z)) with Z = acc + b kept in FP32, loss
mean((Y - t/s)^2 w), and the backward pass dY = bf16(2 (Y - t/s) w / (B D)),
dW = dZ^T H, db = column sums of dZ, dZ_prev = bf16((dZ W) * silu'(Z_prev)).
`exact=True` evaluates the same network in FP64 without any rounding (used to pin the
backward formulas by finite differences).
"""
from __future__ import annotations

import numpy as np


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


class MLPOracle:
    def __init__(self, layers, mean, scale, out_scale, exact=False, out_mean=0.0):
        """layers: list of (W [out x in], b [out]) in float32; out_scale / out_mean: scalars or
        per-output arrays."""
        self.layers = layers
        self.mean, self.scale, self.out_scale = mean, scale, out_scale
        self.out_mean = out_mean
        self.exact = exact

    def _r(self, a):
        return np.asarray(a, np.float64) if self.exact else bf16(a)

    def _mm(self, a, b):
        if self.exact:
            return np.asarray(a, np.float64) @ np.asarray(b, np.float64)
        return (np.asarray(a, np.float32) @ np.asarray(b, np.float32)).astype(np.float32)

    def forward(self, x):
        dt = np.float64 if self.exact else np.float32
        h = self._r(((np.asarray(x, dt) - self.mean) * self.scale).astype(dt))
        Zs, Hs = [], [h]
        n = len(self.layers)
        for i, (W, b) in enumerate(self.layers):
            z = (self._mm(h, self._r(W).T) + np.asarray(b, dt)).astype(dt)
            Zs.append(z)
            if i < n - 1:
                h = self._r(silu(z).astype(dt))
                Hs.append(h)
        return Zs, Hs

    def predict(self, x):
        Zs, _ = self.forward(x)
        return self.out_mean + self.out_scale * Zs[-1]

    def grads(self, x, t, wt=None):
        dt = np.float64 if self.exact else np.float32
        Zs, Hs = self.forward(x)
        Y = Zs[-1]
        B, D = Y.shape
        w = np.ones(D, dt) if wt is None else np.asarray(wt, dt)
        d = Y - ((np.asarray(t, dt) - np.asarray(self.out_mean, dt)) / np.asarray(self.out_scale, dt)).astype(dt)
        loss = float(np.sum(d * d * w) / (B * D))
        g = self._r((2.0 * d * w / (B * D)).astype(dt))
        out = [None] * len(self.layers)
        for i in range(len(self.layers) - 1, -1, -1):
            W, _ = self.layers[i]
            dW = self._mm(g.T, Hs[i])
            db = g.astype(np.float64).sum(axis=0).astype(dt)
            out[i] = (dW, db)
            if i > 0:
                g = self._r((self._mm(g, self._r(W)) * dsilu(Zs[i - 1])).astype(dt))
        return loss, out
