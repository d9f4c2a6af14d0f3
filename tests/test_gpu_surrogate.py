"""GPU: the learned membrane traction (suguru_amd.surrogate) -- what the trained network gains
over the untrained one, the exactness of the FP64 fallback, the equilibrium indicator on the
law's own traction, and inference against the law it stands in for.

Sizes are small (20 000 training shapes, 3 000 Adam steps, 256 held-out capsules) so the
whole file runs in seconds; bench.py's surrogate leg is the full-size measurement.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    assert torch.cuda.is_available(), "no HIP device visible to torch"
    return torch


@pytest.fixture(scope="module")
def trained():
    torch = _torch()
    from suguru_amd.surrogate import Surrogate

    print("case: surrogate, 20000 shapes, 3000 steps", flush=True)
    s = Surrogate(L=16, width=512, nhidden=4, max_batch=1024)
    X, Y, _ = s.dataset(20_000, seed=5)
    Xt, Yt, Pt = s.dataset(256, seed=77)
    s.fit_norm(X, Y)
    s.mlp.init(5)
    err0 = s.rel_error(s.predict(Xt), Yt)
    losses = s.train(X, Y, 3000, batch=512, lr=1e-3, seed=5)
    err = s.rel_error(s.predict(Xt), Yt)
    torch.cuda.synchronize()
    return s, Xt, Yt, Pt, err0, err, losses


def test_trained_network_beats_untrained(trained):
    """Training on the law's labels cuts the per-capsule relative traction error from ~1 (the
    untrained network predicts about the mean traction) to well below half of it."""
    _, _, _, _, err0, err, losses = trained
    m0, m1 = float(err0.median()), float(err.median())
    print(f"median relative error: untrained {m0:.3f}, trained {m1:.3f}", flush=True)
    assert m0 > 0.8
    assert m1 < 0.5 * m0
    assert np.mean(losses[-100:]) < 0.5 * np.mean(losses[:10])


def test_step_fallback_is_exact_and_surrogate_only_path(trained):
    """tol = 0 sends every capsule to the FP64 law: the step's traction is the law's, bit for
    bit.  tol = inf keeps every prediction: the step's traction is the network's output."""
    torch = _torch()
    s, Xt, _, Pt, _, _, _ = trained
    x, y, z = (Pt[k].reshape(-1).contiguous() for k in range(3))
    f_law, _, _ = s.law(x, y, z)
    f0, _, bad0 = s.step(x, y, z, 0.0)
    assert bool(bad0.all())
    assert torch.equal(f0, f_law)
    finf, _, badinf = s.step(x, y, z, float("inf"))
    assert not bool(badinf.any())
    n = Xt.shape[0]
    pred = s.predict(s.coeffs(x, y, z).float()).double().view(n, 3, -1).permute(1, 0, 2)
    assert torch.equal(finf, pred.reshape(3, -1))


def test_indicator_separates_law_and_network(trained):
    """The equilibrium indicator (net force + net torque of the traction, relative) of the
    FP64 law's own traction is small for typical shapes (its discrete equilibrium holds to
    ~1e-5; for the most deformed shapes it reaches ~1e-2 -- a limit of the indicator as a
    gate), and the trained network's predictions sit well above it."""
    s, _, Yt, Pt, _, _, _ = trained
    x, y, z = (Pt[k].reshape(-1).contiguous() for k in range(3))
    f_law, _, area = s.law(x, y, z)
    n = Pt.shape[1]
    fl = f_law.view(3, n, -1).permute(1, 0, 2).reshape(n, -1)
    ind_law = s.indicator(fl, Pt.permute(1, 0, 2), area)
    ind_net = s.indicator(s.predict(s.coeffs(x, y, z).float()), Pt.permute(1, 0, 2), area)
    print(f"indicator: law median {float(ind_law.median()):.2e} max {float(ind_law.max()):.2e}, "
          f"network median {float(ind_net.median()):.2e}", flush=True)
    assert float(ind_law.median()) < 1e-4
    assert float(ind_net.median()) > 10 * float(ind_law.median())


def test_inference_is_faster_than_the_law(trained):
    """The surrogate path without fallback (coefficients, MLP, geometry, indicator) against
    the FP64 law on the same 256 capsules (a generous bar; the bench measures 2.6x at 2048)."""
    torch = _torch()
    s, _, _, Pt, _, _, _ = trained
    x, y, z = (Pt[k].reshape(-1).contiguous() for k in range(3))
    for _ in range(2):
        s.law(x, y, z)
        s.step(x, y, z, float("inf"))
    torch.cuda.synchronize()
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        s.law(x, y, z)
    torch.cuda.synchronize()
    t_law = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        s.step(x, y, z, float("inf"))
    torch.cuda.synchronize()
    t_sur = (time.perf_counter() - t0) / reps
    print(f"law {1e3 * t_law:.3f} ms, surrogate path {1e3 * t_sur:.3f} ms", flush=True)
    assert t_sur < t_law


@pytest.mark.parametrize("q", [0.25, 0.75])
def test_partial_fallback_selects_per_capsule(trained, q):
    """A tolerance between the indicator's quantiles sends part of the batch to the FP64 law
    (q = 0.25: three quarters fall back, the whole-batch-law branch; q = 0.75: a quarter, the
    gathered branch).  Each capsule then carries the law's traction bit for bit if it fell
    back, and the network's prediction otherwise."""
    torch = _torch()
    s, Xt, _, Pt, _, _, _ = trained
    x, y, z = (Pt[k].reshape(-1).contiguous() for k in range(3))
    n = Xt.shape[0]
    pred = s.predict(s.coeffs(x, y, z).float())
    _, area = s.geometry(x, y, z)
    ind = s.indicator(pred, Pt.permute(1, 0, 2), area)
    tol = float(torch.quantile(ind.double(), q))
    f, _, bad = s.step(x, y, z, tol)
    nb = int(bad.sum())
    print(f"q={q}: {nb} of {n} capsules fall back", flush=True)
    assert 0 < nb < n
    f_law, _, _ = s.law(x, y, z)
    fp = pred.double().view(n, 3, -1).permute(1, 0, 2)
    fv, lv = f.view(3, n, -1), f_law.view(3, n, -1)
    assert torch.equal(fv[:, bad], lv[:, bad])
    assert torch.equal(fv[:, ~bad], fp[:, ~bad])


def test_gate_fails_closed_on_nan_inf_and_zero_predictions(trained, monkeypatch):
    """A prediction row that is NaN, holds an Inf, or is all zero (indicator 0/0 = NaN) must
    fall back to the FP64 law at any tolerance, even tol = inf: the capsule then carries the
    law's traction bit for bit."""
    torch = _torch()
    s, Xt, _, Pt, _, _, _ = trained
    x, y, z = (Pt[k].reshape(-1).contiguous() for k in range(3))
    n = Xt.shape[0]
    real_predict = s.predict
    broken = {3: "nan", 17: "inf", 40: "zero"}

    def predict(X):
        p = real_predict(X).clone()
        p[3] = float("nan")
        p[17, 5] = float("inf")
        p[40] = 0.0
        return p

    monkeypatch.setattr(s, "predict", predict)
    f_law, _, _ = s.law(x, y, z)
    for tol in (1e-3, float("inf")):
        f, _, bad = s.step(x, y, z, tol)
        for i in broken:
            assert bool(bad[i]), (tol, broken[i])
            assert torch.equal(f.view(3, n, -1)[:, i], f_law.view(3, n, -1)[:, i]), (tol, broken[i])
        if tol == float("inf"):
            assert int(bad.sum()) == len(broken)
