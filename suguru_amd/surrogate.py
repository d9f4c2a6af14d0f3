"""Learned membrane traction: training data from the GPU FP64 law, training with the library's
BF16 MFMA kernels, inference with a per-capsule residual indicator and FP64 fallback.

Workload (degree 16, 17 x 34 = 578 points per capsule): shapes are the unit sphere (the
reference's stress-free state, sph.f90:56-70) with a seeded random radial perturbation of
spherical-harmonic degree 1..8 and relative amplitude up to 0.3; inputs are the 3 x 289 real
shape coefficients (memb_coeffs_f64_dev), labels the FP64 traction of the reference law at its
configured moduli (gscell = 1/We = 6250, ebend = 0.3 gscell; param.f90:57,88,91).

Residual indicator (cheap, per capsule, no law evaluation): a membrane load in equilibrium has
zero net force and zero net torque (the oracle and the law satisfy both to ~1e-12).  The
indicator is |sum f A| / sum |f| A + |sum (x - c) x f A| / (R sum |f| A) of the predicted
traction; capsules above the tolerance are recomputed with the FP64 law.
"""
from __future__ import annotations

import math
import time

import numpy as np

from .membrane import MembranePlan, SurrogateMLP

GSCELL = 6250.0         # 1 / We, We = Rep * Ca = 0.0016 * 0.1 (param.f90:57,88)
EBEND = 0.3 * GSCELL    # cbend * gscell (param.f90:91)


def grid_angles(L: int):
    """Gauss-Legendre colatitudes (increasing) and uniform longitudes of the degree-L grid."""
    x, _ = np.polynomial.legendre.leggauss(L + 1)
    th = np.arccos(x[::-1])
    ph = 2 * np.pi * np.arange(2 * (L + 1)) / (2 * (L + 1))
    return th, ph


def _real_sh_basis(th, ph, nmax):
    """[npts, nmodes] real spherical harmonics of degrees 1..nmax on the grid (any fixed real
    orthogonal basis will do: it only parametrises the random perturbations)."""
    from scipy.special import sph_harm_y

    TH, PH = np.meshgrid(th, ph, indexing="ij")
    cols = []
    for n in range(1, nmax + 1):
        for m in range(-n, n + 1):
            Y = sph_harm_y(n, abs(m), TH, PH)
            cols.append((np.sqrt(2) * (Y.imag if m < 0 else Y.real) if m != 0 else Y.real).ravel())
    return np.stack(cols, axis=1)


def random_shapes(L: int, n: int, seed: int, amp_max: float = 0.3, nmax: int = 8):
    """n shapes x = r(t,p) (sin t sin p, sin t cos p, cos t), r = 1 + a dr / max|dr|,
    a ~ U(0, amp_max), dr a random combination of real harmonics of degrees 1..nmax whose
    coefficients decay as 1/n.  Returns x, y, z as [n, npts] float64."""
    th, ph = grid_angles(L)
    B = _real_sh_basis(th, ph, nmax)
    deg = np.concatenate([[k] * (2 * k + 1) for k in range(1, nmax + 1)])
    rng = np.random.default_rng(seed)
    c = rng.normal(size=(n, B.shape[1])) / deg[None, :]
    dr = c @ B.T
    a = rng.uniform(0.0, amp_max, size=(n, 1))
    r = 1.0 + a * dr / np.abs(dr).max(axis=1, keepdims=True)
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    ex = (np.sin(TH) * np.sin(PH)).ravel()[None, :]
    ey = (np.sin(TH) * np.cos(PH)).ravel()[None, :]
    ez = np.cos(TH).ravel()[None, :]
    return r * ex, r * ey, r * ez


class Surrogate:
    """Plan + MLP + data on one GPU.  Tensors are torch (device memory only)."""

    def __init__(self, L=16, width=512, nhidden=4, max_batch=1024, device=0, gs=GSCELL, eb=EBEND):
        import torch

        self.torch = torch
        self.dev = torch.device(f"cuda:{device}")
        self.plan = MembranePlan(L, device)
        self.L, self.npts, self.ncoef = L, self.plan.npts, self.plan.ncoef
        self.din, self.dout = 3 * self.ncoef, 3 * self.npts
        self.mlp = SurrogateMLP(self.din, width, nhidden, self.dout, max_batch, device)
        self.max_batch = max_batch
        self.gs, self.eb = gs, eb
        self.timings = {}
        # one stream for the library kernels and torch's own work (indexing, copies)
        s = torch.cuda.current_stream(self.dev).cuda_stream
        self.plan.set_stream(s)
        self.mlp.set_stream(s)

    # ---- data --------------------------------------------------------------------------------
    def upload(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64).reshape(-1)).to(self.dev)

    def law(self, x, y, z):
        """FP64 law on device point arrays [ncap * npts]: (f [3, n], normal [3, n], area [n])."""
        t = self.torch
        n = x.numel()
        out = [t.empty(n, dtype=t.float64, device=self.dev) for _ in range(7)]
        self.plan.traction_dev(x, y, z, self.gs, self.eb, *out)
        return t.stack(out[:3]), t.stack(out[3:6]), out[6]

    def coeffs(self, x, y, z):
        t = self.torch
        ncap = x.numel() // self.npts
        c = t.empty(ncap * self.din, dtype=t.float64, device=self.dev)
        self.plan.coeffs_dev(x, y, z, c)
        return c.view(ncap, self.din)

    def dataset(self, n, seed, chunk=8192, amp_max=0.3):
        """(inputs fp32 [n, din], targets fp32 [n, dout], positions fp64 [3, n, npts])."""
        t = self.torch
        X = t.empty((n, self.din), dtype=t.float32, device=self.dev)
        Y = t.empty((n, self.dout), dtype=t.float32, device=self.dev)
        P = t.empty((3, n, self.npts), dtype=t.float64, device=self.dev)
        t_law = 0.0
        for a in range(0, n, chunk):
            b = min(n, a + chunk)
            xs, ys, zs = random_shapes(self.L, b - a, seed * 1000003 + a, amp_max=amp_max)
            x, y, z = self.upload(xs), self.upload(ys), self.upload(zs)
            X[a:b] = self.coeffs(x, y, z).float()
            t.cuda.synchronize()
            t0 = time.perf_counter()
            f, _, _ = self.law(x, y, z)
            t.cuda.synchronize()
            t_law += time.perf_counter() - t0
            Y[a:b] = f.view(3, b - a, self.npts).permute(1, 0, 2).reshape(b - a, self.dout).float()
            P[0, a:b], P[1, a:b], P[2, a:b] = x.view(-1, self.npts), y.view(-1, self.npts), z.view(-1, self.npts)
        self.timings["law_s_per_capsule"] = t_law / n
        return X, Y, P

    # ---- training ----------------------------------------------------------------------------
    def fit_norm(self, X, Y):
        """Per-feature input standardisation and per-output target standardisation."""
        mean = X.mean(0).cpu().numpy().astype(np.float32)
        std = X.std(0).cpu().numpy()
        scale = np.where(std > 1e-12, 1.0 / np.maximum(std, 1e-12), 0.0).astype(np.float32)
        omean = Y.mean(0).cpu().numpy().astype(np.float32)
        ostd = Y.std(0).cpu().numpy().astype(np.float32)
        ostd = np.maximum(ostd, 1e-6 * float(ostd.max())).astype(np.float32)
        self._norm = (mean, scale, omean, ostd)
        self.mlp.set_norm(mean, scale, omean, ostd)

    def train(self, X, Y, steps, batch=512, lr=1e-3, lr_min=1e-5, seed=0, log_every=0):
        t = self.torch
        g = t.Generator(device=self.dev)
        g.manual_seed(seed)
        n = X.shape[0]
        losses = []
        t0 = time.perf_counter()
        for s in range(steps):
            idx = t.randint(0, n, (batch,), device=self.dev, generator=g)
            xb, yb = X[idx].contiguous(), Y[idx].contiguous()
            lr_s = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * s / max(1, steps)))
            losses.append(self.mlp.train_step(xb, yb, lr_s))
            if log_every and (s % log_every == 0 or s == steps - 1):
                print(f"step {s} loss {losses[-1]:.4e}", flush=True)
        t.cuda.synchronize()
        self.timings["train_s"] = time.perf_counter() - t0
        return losses

    def predict(self, X):
        t = self.torch
        n = X.shape[0]
        out = t.empty((n, self.dout), dtype=t.float32, device=self.dev)
        for a in range(0, n, self.max_batch):
            b = min(n, a + self.max_batch)
            self.mlp.forward(X[a:b].contiguous(), out[a:b], batch=b - a)
        return out

    # ---- evaluation --------------------------------------------------------------------------
    @staticmethod
    def rel_error(pred, ref):
        """per-capsule ||pred - ref|| / ||ref|| over all 3 x npts values."""
        d = (pred.double() - ref.double()).norm(dim=1)
        return d / ref.double().norm(dim=1)

    def indicator(self, pred, P, area):
        """Equilibrium residual of predicted tractions, per capsule: |net force| / sum |f| A
        + |net torque| / (R sum |f| A).  pred [ncap, 3 npts], P [ncap, 3, npts] positions,
        area [ncap * npts]."""
        t = self.torch
        n = pred.shape[0]
        f = pred.double().view(n, 3, self.npts)
        A = area.view(n, self.npts)
        fa = f * A[:, None, :]
        mag = (f.norm(dim=1) * A).sum(1)
        F = fa.sum(2).norm(dim=1)
        r = P - P.mean(2, keepdim=True)
        R = r.norm(dim=1).mean(1)
        tq = t.stack([r[:, 1] * fa[:, 2] - r[:, 2] * fa[:, 1], r[:, 2] * fa[:, 0] - r[:, 0] * fa[:, 2],
                      r[:, 0] * fa[:, 1] - r[:, 1] * fa[:, 0]], dim=1).sum(2).norm(dim=1)
        return F / mag + tq / (R * mag)

    # ---- the surrogate step ----------------------------------------------------------------
    def geometry(self, x, y, z):
        t = self.torch
        n = x.numel()
        out = [t.empty(n, dtype=t.float64, device=self.dev) for _ in range(4)]
        self.plan.geometry_dev(x, y, z, *out)
        return t.stack(out[:3]), out[3]

    def step(self, x, y, z, tol):
        """Traction of every capsule (device point arrays [ncap * npts]): the surrogate's
        prediction where the equilibrium residual is within `tol`, the FP64 law elsewhere.
        Returns (f [3, ncap*npts] fp64, area [ncap*npts], fallback mask [ncap])."""
        t = self.torch
        ncap = x.numel() // self.npts
        pred = self.predict(self.coeffs(x, y, z).float())
        _, area = self.geometry(x, y, z)
        P = t.stack([x.view(ncap, -1), y.view(ncap, -1), z.view(ncap, -1)]).permute(1, 0, 2)
        ind = self.indicator(pred, P, area)
        # fail closed: a NaN indicator (NaN / Inf prediction, or an all-zero one: 0 / 0) and
        # any non-finite predicted value send the capsule to the law
        bad = ~(ind <= tol) | ~t.isfinite(pred).all(dim=1)
        f = pred.double().view(ncap, 3, self.npts).permute(1, 0, 2).reshape(3, -1).contiguous()
        nb = int(bad.sum().item())
        if 2 * nb >= ncap:
            # most capsules fall back: the law on the whole batch and a per-capsule select.
            # The law runs one workgroup per capsule, so a capsule's traction does not depend
            # on the batch it is computed in, and this equals the gathered path bit for bit.
            fl, _, _ = self.law(x, y, z)
            sel = bad.view(1, ncap, 1)
            f = t.where(sel, fl.view(3, ncap, self.npts), f.view(3, ncap, self.npts))
            f = f.reshape(3, -1).contiguous()
        elif nb:
            idx = t.nonzero(bad).view(-1)
            sub = [a.view(ncap, -1).index_select(0, idx).reshape(-1) for a in (x, y, z)]
            fl, _, _ = self.law(*sub)
            f.view(3, ncap, self.npts).index_copy_(1, idx, fl.view(3, nb, self.npts))
        return f, area, bad

    # ---- persistence ---------------------------------------------------------------------------
    def save(self, path):
        """Weights, biases and normalisation as an .npz (plain arrays, no pickles)."""
        arrs = {}
        for i in range(self.mlp.nl):
            W, b = self.mlp.get_layer(i)
            arrs[f"W{i}"], arrs[f"b{i}"] = W, b
        arrs["norm_mean"], arrs["norm_scale"], arrs["out_mean"], arrs["out_scale"] = self._norm
        arrs["meta"] = np.array([self.L, self.mlp.width, self.mlp.nhidden], np.int64)
        np.savez(path, **arrs)

    def load(self, path):
        z = np.load(path, allow_pickle=False)
        for i in range(self.mlp.nl):
            self.mlp.set_layer(i, z[f"W{i}"], z[f"b{i}"])
        self._norm = (z["norm_mean"], z["norm_scale"], z["out_mean"], z["out_scale"])
        self.mlp.set_norm(*self._norm)


def run_experiment(n_train=200_000, n_test=2048, steps=20_000, batch=512, lr=1e-3, seed=1,
                   tol=1e-3, width=512, nhidden=4, log_every=0, save=None, amp_max=0.3):
    """Train the surrogate on GPU-law labels and measure what it gains: per-capsule relative
    traction error (trained vs the untrained network), the equilibrium indicator, and the
    step time of surrogate + fallback against the FP64 law on the held-out capsules."""
    import torch

    s = Surrogate(L=16, width=width, nhidden=nhidden, max_batch=max(batch, 1024))
    t0 = time.perf_counter()
    X, Y, _ = s.dataset(n_train, seed, amp_max=amp_max)
    Xt, Yt, Pt = s.dataset(n_test, seed + 7919, amp_max=amp_max)
    t_data = time.perf_counter() - t0
    s.fit_norm(X, Y)
    s.mlp.init(seed)
    err0 = s.rel_error(s.predict(Xt), Yt)
    losses = s.train(X, Y, steps, batch=batch, lr=lr, seed=seed, log_every=log_every)
    pred = s.predict(Xt)
    err = s.rel_error(pred, Yt)
    q = lambda v, p: float(torch.quantile(v.double().cpu(), p))
    x, y, z = (Pt[k].reshape(-1).contiguous() for k in range(3))
    _, area = s.geometry(x, y, z)
    ind = s.indicator(pred, Pt.permute(1, 0, 2), area)
    # timing: FP64 law vs surrogate step (coefficients + MLP + geometry + indicator + fallback)
    reps = 5
    for _ in range(2):
        s.law(x, y, z)
        s.step(x, y, z, tol)
    torch.cuda.synchronize()
    a = time.perf_counter()
    for _ in range(reps):
        s.law(x, y, z)
    torch.cuda.synchronize()
    t_law = (time.perf_counter() - a) / reps
    a = time.perf_counter()
    for _ in range(reps):
        f, _, bad = s.step(x, y, z, tol)
    torch.cuda.synchronize()
    t_sur = (time.perf_counter() - a) / reps
    # the surrogate path without any fallback (coefficients, MLP, geometry, indicator)
    a = time.perf_counter()
    for _ in range(reps):
        s.step(x, y, z, float("inf"))
    torch.cuda.synchronize()
    t_inf = (time.perf_counter() - a) / reps
    # other tolerances: fallback share, step time, and the largest error that is accepted
    sweep = []
    for tl in (3e-3, 1e-2, 3e-2, 1e-1):
        s.step(x, y, z, tl)  # untimed, like the warm-up of the timings above
        torch.cuda.synchronize()
        a = time.perf_counter()
        for _ in range(reps):
            _, _, bd = s.step(x, y, z, tl)
        torch.cuda.synchronize()
        acc = ~bd
        sweep.append({"tol": tl, "fallback_fraction": float(bd.double().mean()),
                      "step_ms": 1e3 * (time.perf_counter() - a) / reps,
                      "accepted_rel_err_max": float(err[acc].max()) if bool(acc.any()) else 0.0,
                      "accepted_rel_err_median": q(err[acc], 0.5) if bool(acc.any()) else 0.0})
    fl, _, _ = s.law(x, y, z)
    final = s.rel_error(f.view(3, n_test, -1).permute(1, 0, 2).reshape(n_test, -1),
                        fl.view(3, n_test, -1).permute(1, 0, 2).reshape(n_test, -1))
    if save:
        s.save(save)
    return {
        "n_train": n_train, "n_test": n_test, "steps": steps, "batch": batch, "lr": lr,
        "amp_max": amp_max,
        "data_s": t_data, "train_s": s.timings.get("train_s"),
        "law_s_per_capsule": s.timings.get("law_s_per_capsule"),
        "loss_first": losses[0], "loss_last": float(np.mean(losses[-50:])),
        "rel_err_untrained_median": q(err0, 0.5),
        "rel_err_median": q(err, 0.5), "rel_err_p90": q(err, 0.9), "rel_err_max": float(err.max()),
        "indicator_median": q(ind, 0.5),
        "indicator_corr_with_error": float(np.corrcoef(ind.cpu().numpy(), err.cpu().numpy())[0, 1]),
        "tol": tol, "fallback_fraction": float(bad.double().mean()),
        "final_rel_err_max": float(final.max()),
        "t_law_ms": 1e3 * t_law, "t_surrogate_step_ms": 1e3 * t_sur,
        "speedup": t_law / t_sur, "t_inference_ms": 1e3 * t_inf, "tol_sweep": sweep,
    }
