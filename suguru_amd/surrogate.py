"""Learned membrane traction: a co-rotated linear response of the reference law plus a BF16-MFMA
MLP for its nonlinear remainder, a per-capsule error gate and the FP64 law as fallback.

The reference law (complagrforces, forcing.f90:11-114; sph.f90:1523-1755) is translation
invariant and rotation covariant, f(R x + c) = R f(x), with the unit sphere as stress-free
state (sph.f90:56-70).  Per capsule (degree 16: 578 points):

  1. co-rotated frame (memb_corot_f64_dev): centroid c, rotation R = polar factor of
     sum (x_p - c) X_p^T, deformation delta = R^T (x - c) - X, RMS amplitude s;
  2. linear response: J delta, J = d f / d x at the unit sphere, built once by central
     differences of the GPU FP64 law (3 x 578 columns); an FP64 GEMM over all capsules;
  3. the learned remainder: the MLP (4 hidden layers of 512, SiLU; BF16 MFMA, FP32
     accumulation; trained here on GPU-law labels by the library's own kernels) maps the real
     spherical-harmonic coefficients of the direction delta / s (3 x 289) to
     (R^T f - J delta) / s^2, the law's second-order term;
  4. f = R (J delta + s^2 MLP) (memb_corot_apply_f64_dev);
  5. gate: the remainder of a second-order model is O(s^2) relative to J delta, so the
     relative error grows with s: a capsule is kept when alpha s <= tol, alpha calibrated on
     held-out shapes (a high quantile of error / s), and when every value is finite; the others
     are recomputed by the FP64 law in place (memb_traction_select_f64_dev).
Every stage is a device launch on one stream: the step needs no host round trip and can be
captured in a hipGraph.

Shapes for training and evaluation: the unit sphere displaced by a random vector field of
spherical-harmonic degree 1..8 in all three components (coefficients ~ N(0,1)/n), scaled to an
RMS amplitude drawn log-uniformly, then rotated and translated at random.  The law's moduli:
gscell = 1/We (6249.999999999999 in FP64), ebend = 0.3 gscell (param.f90:57,88,91).
"""
from __future__ import annotations

import contextlib
import math
import time

import numpy as np

from .membrane import MembranePlan, SurrogateMLP
from .params import EBEND, GSCELL  # 1 / We (folded: 6249.999999999999) and 0.3 gscell


def grid_angles(L: int):
    """Gauss-Legendre colatitudes (increasing) and uniform longitudes of the degree-L grid."""
    x, _ = np.polynomial.legendre.leggauss(L + 1)
    th = np.arccos(x[::-1])
    ph = 2 * np.pi * np.arange(2 * (L + 1)) / (2 * (L + 1))
    return th, ph


def sphere_points(L: int):
    """Unit-sphere points of the degree-L grid as [3, npts] (x = sin t sin p, y = sin t cos p,
    z = cos t, latitude-major; sph.f90:84-86)."""
    th, ph = grid_angles(L)
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    return np.stack([(np.sin(TH) * np.sin(PH)).ravel(), (np.sin(TH) * np.cos(PH)).ravel(),
                     np.cos(TH).ravel()])


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
    """n radially perturbed spheres x = r(t,p) (sin t sin p, sin t cos p, cos t),
    r = 1 + a dr / max|dr|, a ~ U(0, amp_max), dr a random combination of real harmonics of
    degrees 1..nmax with coefficients decaying as 1/n.  Returns x, y, z as [n, npts]."""
    th, ph = grid_angles(L)
    B = _real_sh_basis(th, ph, nmax)
    deg = np.concatenate([[k] * (2 * k + 1) for k in range(1, nmax + 1)])
    rng = np.random.default_rng(seed)
    c = rng.normal(size=(n, B.shape[1])) / deg[None, :]
    dr = c @ B.T
    a = rng.uniform(0.0, amp_max, size=(n, 1))
    r = 1.0 + a * dr / np.abs(dr).max(axis=1, keepdims=True)
    X = sphere_points(L)
    return r * X[0][None, :], r * X[1][None, :], r * X[2][None, :]


def random_rotations(n, rng, max_angle=math.pi):
    """n rotation matrices about random axes by angles U(0, max_angle)."""
    ax = rng.normal(size=(n, 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    ang = rng.uniform(0.0, max_angle, size=n)
    K = np.zeros((n, 3, 3))
    K[:, 0, 1], K[:, 0, 2], K[:, 1, 2] = -ax[:, 2], ax[:, 1], -ax[:, 0]
    K = K - K.transpose(0, 2, 1)
    s, c = np.sin(ang)[:, None, None], np.cos(ang)[:, None, None]
    return np.eye(3)[None] + s * K + (1 - c) * (K @ K)


def random_deformed(L: int, n: int, seed: int, amp_lo: float = 1e-4, amp_hi: float = 0.05,
                    nmax: int = 8, kind: str = "field"):
    """n deformed, rotated and translated unit spheres X + delta, scaled to an RMS amplitude
    ~ log-uniform(amp_lo, amp_hi), then rotated and shifted at random.
    kind "field": delta a random vector field of degrees 1..nmax in all three components
    (coefficients N(0,1)/degree).  kind "flow": what a capsule in a viscous flow accumulates,
    an affine displacement A X (A with N(0,1) entries: strain, spin and dilation) plus a vector
    field of degrees 2..nmax with coefficients N(0,1)/degree^2 at 30 % of its RMS.
    Returns x, y, z as [n, npts] and the amplitudes."""
    th, ph = grid_angles(L)
    B = _real_sh_basis(th, ph, nmax)
    deg = np.concatenate([[k] * (2 * k + 1) for k in range(1, nmax + 1)])
    rng = np.random.default_rng(seed)
    X = sphere_points(L)
    if kind == "field":
        c = rng.normal(size=(n, 3, B.shape[1])) / deg[None, None, :]
        d = c @ B.T                                    # [n, 3, npts]
    elif kind == "flow":
        A = rng.normal(size=(n, 3, 3))
        d = A @ X[None]
        c = rng.normal(size=(n, 3, B.shape[1])) / (deg ** 2)[None, None, :]
        c[:, :, deg < 2] = 0.0
        h = c @ B.T
        rd = np.sqrt((d ** 2).sum(1).mean(1))
        rh = np.sqrt((h ** 2).sum(1).mean(1))
        d = d + h * (0.3 * rd / rh)[:, None, None]
    else:
        raise ValueError(kind)
    rms = np.sqrt((d ** 2).sum(1).mean(1))
    a = np.exp(rng.uniform(np.log(amp_lo), np.log(amp_hi), size=n))
    d *= (a / rms)[:, None, None]
    P = X[None] + d
    R = random_rotations(n, rng)
    P = R @ P + rng.uniform(-5.0, 5.0, size=(n, 3, 1))
    return P[:, 0], P[:, 1], P[:, 2], a


class Surrogate:
    """The co-rotated surrogate of one degree on one GPU (module docstring).  Device arrays are
    torch tensors (torch provides memory and streams only)."""

    def __init__(self, L=16, width=512, nhidden=4, max_batch=8192, device=0, gs=GSCELL, eb=EBEND):
        import torch

        self.torch = torch
        self.dev = torch.device(f"cuda:{device}")
        self.plan = MembranePlan(L, device)
        self.L, self.npts, self.ncoef = L, self.plan.npts, self.plan.ncoef
        self.din, self.dout = 3 * self.ncoef, 3 * self.npts
        self.mlp = SurrogateMLP(self.din, width, nhidden, self.dout, max_batch, device)
        self.max_batch = max_batch
        self.gs, self.eb = gs, eb
        self.X = torch.from_numpy(sphere_points(L).reshape(-1).copy()).to(self.dev)
        self.JT = None
        self.alpha = None   # gate: error ~ alpha * amplitude
        self.timings = {}
        # a stream of its own: a null handle would send each library to its own stream
        self._own = torch.cuda.Stream(self.dev)
        self.set_stream(self._own.cuda_stream)

    def set_stream(self, stream):
        """All library work and the torch work of the surrogate on this (non-null) HIP stream."""
        if not stream:
            raise ValueError("a non-null stream handle is needed")
        self.stream_handle = stream
        self.plan.set_stream(stream)
        self.mlp.set_stream(stream)
        self.stream = self.torch.cuda.ExternalStream(stream, device=self.dev)

    @contextlib.contextmanager
    def _streamed(self):
        """Run the body on the surrogate's stream, ordered after the caller's stream and
        before its later work (device-side waits only)."""
        t = self.torch
        cur = t.cuda.current_stream(self.dev)
        same = cur.cuda_stream == self.stream_handle
        if not same:
            self.stream.wait_stream(cur)
        with t.cuda.stream(self.stream):
            yield
        if not same:
            cur.wait_stream(self.stream)

    # ---- the law and the frame -----------------------------------------------------------
    def upload(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64).reshape(-1)).to(self.dev)

    def law(self, x, y, z):
        """FP64 law on device point arrays [ncap * npts]: (f [3, n], normal [3, n], area [n])."""
        t = self.torch
        n = x.numel()
        with self._streamed():
            out = [t.empty(n, dtype=t.float64, device=self.dev) for _ in range(7)]
            self.plan.traction_dev(x, y, z, self.gs, self.eb, *out)
            res = t.stack(out[:3]), t.stack(out[3:6]), out[6]
        return res

    def corot(self, x, y, z):
        """(delta [ncap, 3 npts], direction (dx, dy, dz) [ncap * npts] each, rot [ncap, 9],
        amp [ncap])."""
        t = self.torch
        ncap = x.numel() // self.npts
        e = lambda *s: t.empty(*s, dtype=t.float64, device=self.dev)
        with self._streamed():
            delta, rot, amp = e(ncap, self.dout), e(ncap, 9), e(ncap)
            dh = [e(ncap * self.npts) for _ in range(3)]
            self.plan.corot_dev(self.X, x, y, z, delta, *dh, rot, amp)
        return delta, dh, rot, amp

    def build_jacobian(self, h=1e-6):
        """J = d f / d x at the unit sphere (FP64, [3 npts, 3 npts], x block, y block, z block),
        by central differences of the GPU law: one batch of 2 x 3 npts perturbed spheres."""
        t = self.torch
        n, m = self.npts, self.dout
        with self._streamed():
            base = self.X.view(1, m).repeat(m, 1)
            eye = t.eye(m, dtype=t.float64, device=self.dev) * h
            cols = []
            for sgn in (1.0, -1.0):
                P = base + sgn * eye                   # capsule i: coordinate i moved by +-h
                x, y, z = (P[:, k * n:(k + 1) * n].contiguous().view(-1) for k in range(3))
                f, _, _ = self.law(x, y, z)
                cols.append(f.view(3, m, n).permute(1, 0, 2).reshape(m, m))  # [capsule][3 n]
            J = ((cols[0] - cols[1]) / (2 * h)).T      # column i = d f / d x_i
            self.JT = J.T.contiguous()
        return J

    def features(self, dh):
        """Real SH coefficients of the deformation direction, FP32 [ncap, 3 ncoef]."""
        t = self.torch
        ncap = dh[0].numel() // self.npts
        with self._streamed():
            c = t.empty(ncap * self.din, dtype=t.float64, device=self.dev)
            self.plan.coeffs_dev(dh[0], dh[1], dh[2], c)
            out = c.view(ncap, self.din).float()
        return out

    # ---- data ---------------------------------------------------------------------------------
    def dataset(self, n, seed, chunk=4096, amp_lo=1e-4, amp_hi=0.05, kind="flow"):
        """(features fp32 [n, din], targets fp32 [n, dout] = (R^T f - J delta) / s^2,
        positions fp64 [3, n, npts], amplitudes fp64 [n], law traction fp64 [n, dout])."""
        t = self.torch
        X = t.empty((n, self.din), dtype=t.float32, device=self.dev)
        Y = t.empty((n, self.dout), dtype=t.float32, device=self.dev)
        P = t.empty((3, n, self.npts), dtype=t.float64, device=self.dev)
        A = t.empty(n, dtype=t.float64, device=self.dev)
        F = t.empty((n, self.dout), dtype=t.float64, device=self.dev)
        t.cuda.synchronize()
        for a in range(0, n, chunk):
            b = min(n, a + chunk)
            xs, ys, zs, _ = random_deformed(self.L, b - a, seed * 1000003 + a, amp_lo, amp_hi,
                                            kind=kind)
            x, y, z = self.upload(xs), self.upload(ys), self.upload(zs)
            f, _, _ = self.law(x, y, z)
            fc = f.view(3, b - a, self.npts).permute(1, 0, 2).reshape(b - a, self.dout)
            delta, dh, rot, amp = self.corot(x, y, z)
            R = rot.view(-1, 3, 3)
            fr = (R.transpose(1, 2) @ fc.view(-1, 3, self.npts)).reshape(b - a, self.dout)
            r = (fr - delta @ self.JT) / (amp * amp)[:, None]
            X[a:b] = self.features(dh)
            Y[a:b] = r.float()
            P[0, a:b], P[1, a:b], P[2, a:b] = x.view(-1, self.npts), y.view(-1, self.npts), z.view(-1, self.npts)
            A[a:b] = amp
            F[a:b] = fc
        t.cuda.synchronize()
        return X, Y, P, A, F

    # ---- training ------------------------------------------------------------------------------
    def fit_norm(self, X, Y):
        """Per-feature input standardisation and per-output target standardisation."""
        mean = X.mean(0).cpu().numpy().astype(np.float32)
        std = X.std(0).cpu().numpy()
        # features the co-rotated frame pins (the centroid and rotation modes of the direction
        # field) vary only by round-off: scaling them up would feed round-off to the network
        floor = 1e-3 * float(std.max())
        scale = np.where(std > floor, 1.0 / np.maximum(std, floor), 0.0).astype(np.float32)
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
            with self._streamed():
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
        with self._streamed():
            out = t.empty((n, self.dout), dtype=t.float32, device=self.dev)
            for a in range(0, n, self.max_batch):
                b = min(n, a + self.max_batch)
                self.mlp.forward(X[a:b].contiguous(), out[a:b], batch=b - a)
        return out

    # ---- the surrogate traction ----------------------------------------------------------------
    def predict_traction(self, x, y, z, learned=True):
        """f [ncap, 3 npts] (x block, y block, z block per capsule), amp [ncap]: the co-rotated
        linear response plus (learned = True) the MLP remainder."""
        t = self.torch
        ncap = x.numel() // self.npts
        with self._streamed():
            delta, dh, rot, amp = self.corot(x, y, z)
            flin = delta @ self.JT
            resid = self.predict(self.features(dh)) if learned else None
            f = [t.empty(ncap * self.npts, dtype=t.float64, device=self.dev) for _ in range(3)]
            self.plan.corot_apply_dev(flin, resid, rot, amp, *f)
            out = t.stack([a.view(ncap, self.npts) for a in f], 1).reshape(ncap, self.dout)
        return out, amp

    @staticmethod
    def rel_error(pred, ref):
        """per-capsule ||pred - ref|| / ||ref|| over all 3 x npts values."""
        d = (pred.double() - ref.double()).norm(dim=1)
        return d / ref.double().norm(dim=1)

    def calibrate(self, err, amp, q=0.99):
        """Gate slope: the q-quantile of relative error / amplitude on held-out shapes."""
        ratio = (err / amp).double()
        self.alpha = float(self.torch.quantile(ratio[self.torch.isfinite(ratio)].cpu(), q))
        return self.alpha

    def enqueue(self, x, y, z, fx, fy, fz, nx, ny, nz, area, tol=1e-3):
        """The surrogate step's traction on the current stream, no host round trip: prediction
        and geometry for every capsule, then the FP64 law in place for the capsules the gate
        rejects (alpha * amplitude > tol, or any non-finite value).  Returns the int32 mask
        of recomputed capsules (device)."""
        t = self.torch
        ncap = x.numel() // self.npts
        with self._streamed():
            f, amp = self.predict_traction(x, y, z)
            fv = f.view(ncap, 3, self.npts)
            fx.view(ncap, self.npts).copy_(fv[:, 0])
            fy.view(ncap, self.npts).copy_(fv[:, 1])
            fz.view(ncap, self.npts).copy_(fv[:, 2])
            self.plan.geometry_dev(x, y, z, nx, ny, nz, area)
            alpha = self.alpha if self.alpha is not None else float("inf")
            ok = (alpha * amp <= tol) & t.isfinite(f).all(dim=1)
            sel = (~ok).to(t.int32)
            self.plan.traction_select_dev(sel, x, y, z, self.gs, self.eb, fx, fy, fz, nx, ny, nz,
                                          area)
        return sel

    # ---- persistence ---------------------------------------------------------------------------
    def save(self, path):
        """Weights, biases, normalisation, gate and Jacobian as an .npz (plain arrays)."""
        arrs = {}
        for i in range(self.mlp.nl):
            W, b = self.mlp.get_layer(i)
            arrs[f"W{i}"], arrs[f"b{i}"] = W, b
        arrs["norm_mean"], arrs["norm_scale"], arrs["out_mean"], arrs["out_scale"] = self._norm
        arrs["meta"] = np.array([self.L, self.mlp.width, self.mlp.nhidden], np.int64)
        arrs["alpha"] = np.array([self.alpha if self.alpha is not None else np.inf])
        arrs["JT"] = self.JT.cpu().numpy()
        np.savez(path, **arrs)

    def load(self, path):
        z = np.load(path, allow_pickle=False)
        for i in range(self.mlp.nl):
            self.mlp.set_layer(i, z[f"W{i}"], z[f"b{i}"])
        self._norm = (z["norm_mean"], z["norm_scale"], z["out_mean"], z["out_scale"])
        self.mlp.set_norm(*self._norm)
        self.alpha = float(z["alpha"][0])
        self.JT = self.torch.from_numpy(z["JT"]).to(self.dev)


AMP_BINS = (1e-4, 3e-4, 1e-3, 3e-3, 1e-2, 3e-2, 5e-2)


def run_experiment(n_train=200_000, n_cal=4096, n_test=4096, steps=20_000, batch=512, lr=1e-3,
                   seed=1, tol=1e-3, width=512, nhidden=4, log_every=0, save=None, kind="flow",
                   amp_lo=1e-4, amp_hi=0.05):
    """Build J, train the remainder network on GPU-law labels, calibrate the gate, and measure
    on held-out shapes: relative traction error against amplitude with and without the learned
    part (the gain), the gate's rank correlation with the error, the accepted error and the
    fallback share at tol, and the traction stage's time against the FP64 law."""
    import torch

    s = Surrogate(L=16, width=width, nhidden=nhidden, max_batch=8192)
    t0 = time.perf_counter()
    s.build_jacobian()
    torch.cuda.synchronize()
    t_jac = time.perf_counter() - t0
    t0 = time.perf_counter()
    X, Y, _, _, _ = s.dataset(n_train, seed, kind=kind, amp_lo=amp_lo, amp_hi=amp_hi)
    Xc, _, Pc, Ac, Fc = s.dataset(n_cal, seed + 101, kind=kind, amp_lo=amp_lo, amp_hi=amp_hi)
    Xt, _, Pt, At, Ft = s.dataset(n_test, seed + 7919, kind=kind, amp_lo=amp_lo, amp_hi=amp_hi)
    t_data = time.perf_counter() - t0
    s.fit_norm(X, Y)
    s.mlp.init(seed)
    losses = s.train(X, Y, steps, batch=batch, lr=lr, seed=seed, log_every=log_every)
    del X, Y

    def errors(P, F, learned):
        x, y, z = (P[k].reshape(-1).contiguous() for k in range(3))
        pred, amp = s.predict_traction(x, y, z, learned=learned)
        return s.rel_error(pred, F), amp

    err_c, amp_c = errors(Pc, Fc, True)
    alpha = s.calibrate(err_c, amp_c)
    err_lin, amp_t = errors(Pt, Ft, False)
    err, _ = errors(Pt, Ft, True)
    q = lambda v, p: float(torch.quantile(v.double().cpu(), p))
    by_amp = []
    for lo, hi in zip(AMP_BINS[:-1], AMP_BINS[1:]):
        m = (amp_t >= lo) & (amp_t < hi)
        if bool(m.any()):
            by_amp.append({"amp": [lo, hi], "n": int(m.sum()),
                           "linear_median": q(err_lin[m], 0.5), "learned_median": q(err[m], 0.5),
                           "learned_max": float(err[m].max())})
    acc = alpha * amp_t <= tol
    from scipy.stats import spearmanr

    rho = float(spearmanr((alpha * amp_t).cpu().numpy(), err.cpu().numpy()).statistic)
    # timing of the traction stage on the held-out batch: FP64 law vs the surrogate (gate on)
    x, y, z = (Pt[k].reshape(-1).contiguous() for k in range(3))
    n = x.numel()
    e = lambda: torch.empty(n, dtype=torch.float64, device=s.dev)
    outs = [e() for _ in range(7)]
    reps = 10
    for _ in range(2):
        s.law(x, y, z)
        s.enqueue(x, y, z, *outs, tol=tol)
    torch.cuda.synchronize()
    a = time.perf_counter()
    for _ in range(reps):
        s.law(x, y, z)
    torch.cuda.synchronize()
    t_law = (time.perf_counter() - a) / reps
    a = time.perf_counter()
    for _ in range(reps):
        sel = s.enqueue(x, y, z, *outs, tol=tol)
    torch.cuda.synchronize()
    t_sur = (time.perf_counter() - a) / reps
    fl, _, _ = s.law(x, y, z)
    fs = torch.stack([o.view(n_test, -1) for o in outs[:3]], 1).reshape(n_test, -1)
    final = s.rel_error(fs, fl.view(3, n_test, -1).permute(1, 0, 2).reshape(n_test, -1))
    if save:
        s.save(save)
    out = {
        "n_train": n_train, "n_test": n_test, "steps": steps, "batch": batch, "lr": lr,
        "shapes": kind, "amp_range": [amp_lo, amp_hi],
        "jacobian_s": t_jac, "data_s": t_data, "train_s": s.timings.get("train_s"),
        "loss_first": losses[0], "loss_last": float(np.mean(losses[-50:])),
        "rel_err_linear_median": q(err_lin, 0.5), "rel_err_median": q(err, 0.5),
        "rel_err_p90": q(err, 0.9), "rel_err_max": float(err.max()),
        "by_amplitude": by_amp, "gate_alpha": alpha, "gate_spearman": rho, "tol": tol,
        "fallback_fraction": float((~acc).double().mean()),
        "accepted_rel_err_median": q(err[acc], 0.5) if bool(acc.any()) else 0.0,
        "accepted_rel_err_max": float(err[acc].max()) if bool(acc.any()) else 0.0,
        "step_fallback_fraction": float(sel.double().mean()),
        "step_rel_err_max": float(final.max()),
        "t_law_ms": 1e3 * t_law, "t_surrogate_ms": 1e3 * t_sur, "speedup": t_law / t_sur,
    }
    return out



class AnchoredTraction:
    """The surrogate as a time-stepping traction stage: each capsule's prediction is anchored
    at the last shape where the FP64 law was evaluated for it (all in the co-rotated frame,
    f = R f'):

        f' = J delta + c^2 m0 + b,     m0 = s0^2 MLP(delta0 / s0),   c = <delta, delta0> / |delta0|^2
        b  = R0^T f_law(anchor) - J delta0 - m0

    At the anchor the prediction is the law's value; away from it the error is what the
    model misses of the law's change, O(|delta - delta0|) relative instead of O(|delta|).  The
    learned second-order term Q is evaluated at the anchor and carried by the projection c of
    the current deformation on the anchor's: with delta = c delta0 + d, Q(delta) =
    c^2 Q(delta0) + 2 c Q(delta0, d) + Q(d), so growth along the anchor's own direction (how a
    capsule deforms in a steady flow) is followed exactly up to the network's accuracy, and an
    orthogonal change leaves the anchored linear model as it is.  (A BF16 network re-evaluated
    every step would add its rounding steps to the small differences the anchoring works
    with.)
    The gate: a capsule keeps its prediction while beta * rms(delta - delta0) <= tol (beta
    calibrated on held-out shape pairs) and every value is finite; otherwise the law
    recomputes it in place and its anchor moves to the current shape.  Anchors start at the
    unit sphere (delta0 = 0, b = m0 = 0: the law's load there is zero).  Every operation is a
    device launch on the surrogate's stream: capturable in a hipGraph.

    learned = False drops the MLP term (m0 = 0): the same scheme without the learned part.
    The network runs only for the capsules whose anchor moves (their SH coefficients
    gathered into max_refresh rows).  The drift test needs no linear response, so it runs
    first (memb_anchored_gate_f64_dev): the law then recomputes the capsules it rejects on a
    second stream, and the geometry pass (normals, areas) the others on a third, while J
    delta, the fused prediction (which skips the law's capsules and flags any non-finite
    prediction for a second law call after the streams join) and the network run on the
    surrogate's stream.  Every capsule gets the same values as with the law run after the
    prediction.
    """

    def __init__(self, sur: Surrogate, ncap: int, beta: float, tol: float = 1e-3,
                 learned: bool = True, max_refresh: int = None):
        t = sur.torch
        self.s, self.ncap, self.beta, self.tol, self.learned = sur, ncap, float(beta), tol, learned
        # anchors moved per step with the learned term: the network runs on these rows only
        # (a capsule beyond them gets the law's traction and moves its anchor a step later)
        if max_refresh is None:
            max_refresh = ncap if ncap <= 2048 else max(2048, ncap // 4)
        self.kmax = int(min(max(1, max_refresh), ncap))
        z = lambda *shape: t.zeros(*shape, dtype=t.float64, device=sur.dev)
        self.delta0, self.b, self.m0 = z(ncap, sur.dout), z(ncap, sur.dout), z(ncap, sur.dout)
        self.s0 = z(ncap)
        self.refreshed = t.zeros(ncap, dtype=t.int64, device=sur.dev)   # law calls per capsule
        self.steps = 0
        self._side = None   # stream of the geometry pass
        self._lawst = None  # stream of the law on the drift-rejected capsules

    def reset(self):
        with self.s._streamed():
            for a in (self.delta0, self.b, self.m0, self.s0, self.refreshed):
                a.zero_()
        self.steps = 0

    def _linear(self, delta):
        """J delta over all capsules: one FP64 GEMM.  (In FP32 the traction, a small difference
        of large stiffness terms, loses up to ~4e-5 of itself; the calibrated gate then becomes
        five times stricter and the law runs five times as often.)"""
        return delta @ self.s.JT

    def _geometry(self, x, y, z, nx, ny, nz, area, skip=None):
        """Normals and areas on a side stream (the law's geometry pass), concurrent with the
        prediction; capsules with skip[c] != 0 are left to the law.  Joined before the
        second law call, which writes normals and areas too."""
        s, t = self.s, self.s.torch
        if self._side is None:
            self._side = t.cuda.Stream(s.dev)
        self._side.wait_stream(t.cuda.current_stream(s.dev))
        s.plan.set_stream(self._side.cuda_stream)
        try:
            s.plan.geometry_dev(x, y, z, nx, ny, nz, area, skip)
        finally:
            s.plan.set_stream(s.stream_handle)

    def _law_beside(self, sel, x, y, z, fx, fy, fz, nx, ny, nz, area):
        """The FP64 law for the capsules with sel != 0 on a stream of its own (it writes only
        their rows of the outputs and allocates nothing)."""
        s, t = self.s, self.s.torch
        if self._lawst is None:
            self._lawst = t.cuda.Stream(s.dev)
        self._lawst.wait_stream(t.cuda.current_stream(s.dev))
        s.plan.set_stream(self._lawst.cuda_stream)
        try:
            s.plan.traction_select_dev(sel, x, y, z, s.gs, s.eb, fx, fy, fz, nx, ny, nz, area)
        finally:
            s.plan.set_stream(s.stream_handle)

    def _predict_into(self, x, y, z, fx, fy, fz, sel, drift=None):
        """corot, J delta, and the fused prediction + gate (memb_anchored_predict_f64_dev)."""
        s = self.s
        delta, dh, rot, amp = s.corot(x, y, z)
        lin = self._linear(delta)
        s.plan.anchored_predict_dev(lin, self.m0 if self.learned else None, self.b, delta,
                                    self.delta0, rot, self.beta, self.tol, fx, fy, fz, sel,
                                    drift)
        return delta, dh, rot, amp, lin

    def predict(self, x, y, z):
        """(f [ncap, 3, npts] in the global frame, rms drift [ncap]) from the current anchors."""
        s, t = self.s, self.s.torch
        n, ncap = s.npts, self.ncap
        with s._streamed():
            f = t.empty(3, ncap * n, dtype=t.float64, device=s.dev)
            sel = t.empty(ncap, dtype=t.int32, device=s.dev)
            drift = t.empty(ncap, dtype=t.float64, device=s.dev)
            self._predict_into(x, y, z, f[0], f[1], f[2], sel, drift)
            f = f.view(3, ncap, n).permute(1, 0, 2)
        return f, drift

    def enqueue(self, x, y, z, fx, fy, fz, nx, ny, nz, area):
        """Traction, normals and areas of every capsule on the surrogate's stream.  Returns
        the int32 mask of the capsules the law recomputed this step (device)."""
        s, t = self.s, self.s.torch
        i32 = lambda n: t.empty(n, dtype=t.int32, device=s.dev)
        with s._streamed():
            main = t.cuda.current_stream(s.dev)
            delta, dh, rot, amp = s.corot(x, y, z)
            pre = i32(self.ncap)
            s.plan.anchored_gate_dev(delta, self.delta0, self.beta, self.tol, pre)
            self._geometry(x, y, z, nx, ny, nz, area, skip=pre)
            self._law_beside(pre, x, y, z, fx, fy, fz, nx, ny, nz, area)
            lin = self._linear(delta)
            sel, late = i32(self.ncap), i32(self.ncap)
            s.plan.anchored_predict_dev(lin, self.m0 if self.learned else None, self.b, delta,
                                        self.delta0, rot, self.beta, self.tol, fx, fy, fz,
                                        sel, None, pre, late)
            slot = m_k = None
            if self.learned:
                # the network's term for the (first kmax) capsules whose anchor moves
                idx, slot = i32(self.kmax), i32(self.ncap)
                s.plan.compact_sel_dev(sel, self.kmax, idx, slot)
                il = idx.long()
                dh_k = [d.view(self.ncap, -1)[il].reshape(-1) for d in dh]
                a_k = amp[il]
                m_k = (a_k * a_k)[:, None] * s.predict(s.features(dh_k)).double()
            main.wait_stream(self._lawst)
            main.wait_stream(self._side)
            # non-finite predictions: the law for them too
            s.plan.traction_select_dev(late, x, y, z, s.gs, s.eb, fx, fy, fz, nx, ny, nz, area)
            s.plan.anchored_update_dev(sel, slot, rot, fx, fy, fz, lin, m_k, delta, amp,
                                       self.b, self.delta0, self.m0, self.s0, self.refreshed)
        self.steps += 1
        return sel


def anchored_pairs(L: int, n: int, seed: int, amp0=(1e-4, 2e-2), damp=(1e-6, 3e-3),
                   along=0.0):
    """Held-out (anchor, current) shape pairs: anchors are flow-like deformations of amplitude
    log-uniform in amp0 (a quarter of them the sphere itself); the current shape adds a
    flow-like increment of amplitude log-uniform in damp and turns by up to 0.05 rad; both
    then move by one random rigid motion.  along in [0, 1]: that share of the increment is the
    anchor's own deformation scaled (growth along its direction, as in a steady flow), the
    rest an independent direction.  Returns two [n, 3, npts] arrays."""
    rng = np.random.default_rng(seed)
    X = sphere_points(L)

    def deformation(lo, hi, sd):
        x, y, z, _ = random_deformed(L, n, sd, lo, hi, kind="flow")
        D = np.stack([x, y, z], 1)
        D = D - D.mean(2, keepdims=True)
        U, _, Vt = np.linalg.svd(D @ X.T)     # undo the generator's rigid motion
        return (U @ Vt).transpose(0, 2, 1) @ D - X[None]

    d0 = deformation(amp0[0], amp0[1], seed)
    d0[rng.uniform(size=n) < 0.25] = 0.0
    inc = deformation(damp[0], damp[1], seed + 1)
    if along > 0.0:
        r0 = np.sqrt((d0 ** 2).sum(1).mean(1))
        ri = np.sqrt((inc ** 2).sum(1).mean(1))
        grow = np.where(r0 > 0, ri / np.maximum(r0, 1e-300), 0.0)[:, None, None] * d0
        inc = along * grow + (1.0 - along) * inc
    P0 = X[None] + d0
    P1 = random_rotations(n, rng, max_angle=0.05) @ (P0 + inc)
    R1 = random_rotations(n, rng)
    c = rng.uniform(-5.0, 5.0, size=(n, 3, 1))
    return R1 @ P0 + c, R1 @ P1 + c


def calibrate_anchored(sur: Surrogate, learned=True, n=4096, seed=77, q=0.99, **kw):
    """beta of the anchored gate: the q-quantile of (relative traction error) / rms drift over
    held-out (anchor, current) pairs.  Returns (beta, error, drift) with error and drift on
    the device."""
    t = sur.torch
    P0, P1 = anchored_pairs(sur.L, n, seed, **kw)
    up = lambda A, k: sur.upload(np.ascontiguousarray(A[:, k]))
    at = AnchoredTraction(sur, n, beta=float("inf"), learned=learned,
                          max_refresh=n)   # inf: every capsule anchors
    e = lambda: t.empty(n * sur.npts, dtype=t.float64, device=sur.dev)
    outs = [e() for _ in range(7)]
    at.enqueue(up(P0, 0), up(P0, 1), up(P0, 2), *outs)
    x1, y1, z1 = up(P1, 0), up(P1, 1), up(P1, 2)
    f, drift = at.predict(x1, y1, z1)
    fl, _, _ = sur.law(x1, y1, z1)
    ref = fl.view(3, n, -1).permute(1, 0, 2).reshape(n, -1)
    err = Surrogate.rel_error(f.reshape(n, -1), ref)
    ratio = (err / drift).double()
    beta = float(t.quantile(ratio[t.isfinite(ratio)].cpu(), q))
    return beta, err, drift
