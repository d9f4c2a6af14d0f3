"""GPU: the learned membrane traction (suguru_amd.surrogate) -- the co-rotated linear response,
the learned second-order term and what it gains over the method without it, the anchored
time-stepping stage with its gate, the selective FP64 fallback (bit for bit the law), the
fail-closed gate and graph capture.

Sizes are small (40 000 training shapes, 3 000 Adam steps, a few hundred held-out capsules) so
the file runs in seconds; bench.py's coupled-step leg is the full-size measurement.
"""
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

    print("case: surrogate, Jacobian + 40000 shapes, 3000 steps", flush=True)
    s = Surrogate(L=16, width=512, nhidden=4, max_batch=2048)
    s.build_jacobian()
    X, Y, _, _, _ = s.dataset(40_000, seed=5, kind="flow")
    s.fit_norm(X, Y)
    s.mlp.init(5)
    losses = s.train(X, Y, 3000, batch=512, lr=1e-3, seed=5)
    Xt, _, Pt, At, Ft = s.dataset(512, seed=77, kind="flow")
    torch.cuda.synchronize()
    return s, Pt, At, Ft, losses


def _xyz(P):
    return tuple(P[k].reshape(-1).contiguous() for k in range(3))


def test_jacobian_annihilates_rigid_motions(trained):
    """J = d f / d x at the unit sphere: the law is translation invariant and rotation
    covariant with zero load at the sphere, so J maps translations and infinitesimal
    rotations to zero (to the central differences' accuracy)."""
    torch = _torch()
    s = trained[0]
    n = s.npts
    X = s.X.view(3, n)
    J = s.JT.T
    scale = float(J.abs().max()) * n ** 0.5
    for k in range(3):
        tr = torch.zeros(3 * n, dtype=torch.float64, device=s.dev)
        tr[k * n:(k + 1) * n] = 1.0
        assert float((J @ tr).norm()) <= 1e-8 * scale
    for a, b in ((0, 1), (1, 2), (2, 0)):  # generators of rotations about the three axes
        g = torch.zeros(3, n, dtype=torch.float64, device=s.dev)
        g[a], g[b] = -X[b], X[a]
        assert float((J @ g.reshape(-1)).norm()) <= 1e-8 * scale


def test_learned_term_beats_the_linear_response(trained):
    """What the trained network gains over the method without it: on held-out flow-like
    shapes the co-rotated prediction with the learned second-order term is closer to the FP64
    law than the co-rotated linear response alone, and the training loss fell."""
    s, Pt, At, Ft, losses = trained
    x, y, z = _xyz(Pt)
    f_lin, _ = s.predict_traction(x, y, z, learned=False)
    f_mlp, _ = s.predict_traction(x, y, z, learned=True)
    e_lin = s.rel_error(f_lin, Ft)
    e_mlp = s.rel_error(f_mlp, Ft)
    r = float((e_mlp / e_lin).median())
    print(f"median error: linear {float(e_lin.median()):.2e}, learned {float(e_mlp.median()):.2e} "
          f"(ratio {r:.2f})", flush=True)
    assert r < 0.8
    assert np.mean(losses[-100:]) < 0.6 * np.mean(losses[:10])
    # the error of a second-order model grows with the deformation
    small = At < 1e-3
    assert float(e_mlp[small].max()) < float(e_mlp[~small].max())


def test_selective_law_is_the_law_bit_for_bit(trained):
    """memb_traction_select_f64_dev recomputes the selected capsules exactly as the full law
    and leaves the others untouched."""
    torch = _torch()
    s, Pt, _, _, _ = trained
    x, y, z = _xyz(Pt)
    n = x.numel()
    ncap = n // s.npts
    f_law, nrm, area = s.law(x, y, z)
    outs = [torch.full((n,), 7.0, dtype=torch.float64, device=s.dev) for _ in range(7)]
    sel = (torch.arange(ncap, device=s.dev) % 3 == 1).to(torch.int32)
    with s._streamed():
        s.plan.traction_select_dev(sel, x, y, z, s.gs, s.eb, *outs)
    torch.cuda.synchronize()
    m = sel.bool().repeat_interleave(s.npts)
    for k in range(3):
        assert torch.equal(outs[k][m], f_law[k][m])
        assert bool((outs[k][~m] == 7.0).all())
    assert torch.equal(outs[6][m], area[m])


def test_anchored_stage_is_the_law_at_its_anchors_and_gates(trained):
    """After an anchoring step the prediction at the same shapes is the law's value (to
    rounding); the gate recomputes a capsule once beta * drift exceeds tol, and a capsule it
    keeps is within tol of the law on these shapes."""
    torch = _torch()
    from suguru_amd.surrogate import AnchoredTraction, anchored_pairs

    s = trained[0]
    n = 256
    P0, P1 = anchored_pairs(s.L, n, seed=3, damp=(1e-6, 1e-4), along=0.8)
    up = lambda A, k: s.upload(np.ascontiguousarray(A[:, k]))
    x0, y0, z0 = up(P0, 0), up(P0, 1), up(P0, 2)
    x1, y1, z1 = up(P1, 0), up(P1, 1), up(P1, 2)
    e = lambda: torch.empty(n * s.npts, dtype=torch.float64, device=s.dev)
    outs = [e() for _ in range(7)]
    at = AnchoredTraction(s, n, beta=float("inf"), tol=1e-3)
    sel = at.enqueue(x0, y0, z0, *outs)          # beta = inf: every capsule anchors here
    assert bool(sel.all())
    f, drift = at.predict(x0, y0, z0)
    fl, _, _ = s.law(x0, y0, z0)
    ref = fl.view(3, n, -1).permute(1, 0, 2)
    err = s.rel_error(f.reshape(n, -1), ref.reshape(n, -1))
    assert float(err.max()) < 1e-9 and float(drift.max()) < 1e-12
    # at the moved shapes: the gate keeps what it may keep, and that is within tol
    at.beta, at.tol = 7.6, 1e-3
    sel = at.enqueue(x1, y1, z1, *outs)
    kept = ~sel.bool()
    fl, _, _ = s.law(x1, y1, z1)
    ref = fl.view(3, n, -1).permute(1, 0, 2).reshape(n, -1)
    got = torch.stack([o.view(n, -1) for o in outs[:3]], 1).reshape(n, -1)
    err = s.rel_error(got, ref)
    print(f"kept {int(kept.sum())} of {n}; max error kept {float(err[kept].max()):.2e}", flush=True)
    assert int(kept.sum()) > n // 2
    assert float(err[kept].max()) <= 1e-3
    assert torch.equal(got[~kept], ref[~kept])   # recomputed capsules: the law's own bits


def test_gate_fails_closed_on_nonfinite_predictions(trained):
    """A capsule whose prediction holds NaN or Inf (planted in its anchor offset b) is
    recomputed by the law at any tolerance, even beta = 0 (which would keep every finite
    prediction), and its anchor becomes finite again."""
    torch = _torch()
    from suguru_amd.surrogate import AnchoredTraction

    s, Pt, _, _, _ = trained
    x, y, z = _xyz(Pt)
    ncap = x.numel() // s.npts
    at = AnchoredTraction(s, ncap, beta=0.0, tol=1e-3)
    at.b[3] = float("nan")
    at.b[17, 5] = float("inf")
    e = lambda: torch.empty(x.numel(), dtype=torch.float64, device=s.dev)
    outs = [e() for _ in range(7)]
    sel = at.enqueue(x, y, z, *outs)
    fl, _, _ = s.law(x, y, z)
    assert int(sel.sum()) == 2 and bool(sel[3]) and bool(sel[17])
    for c in (3, 17):
        sl = slice(c * s.npts, (c + 1) * s.npts)
        for k in range(3):
            assert torch.equal(outs[k][sl], fl[k][sl])
    assert bool(torch.isfinite(at.b).all())


def test_anchored_stage_replays_in_a_graph(trained):
    """The anchored stage (prediction, gate, selective law, anchor update) captured in a
    hipGraph: every replay equals the same eager step, with eager work between replays."""
    torch = _torch()
    from suguru_amd.surrogate import AnchoredTraction, anchored_pairs

    s = trained[0]
    n = 128
    P0, P1 = anchored_pairs(s.L, n, seed=9, damp=(1e-5, 3e-4), along=0.5)
    up = lambda A, k: s.upload(np.ascontiguousarray(A[:, k]))
    shapes = [(up(P, 0), up(P, 1), up(P, 2)) for P in (P0, P1)]
    x, y, z = (torch.empty_like(a) for a in shapes[0])
    e = lambda: torch.empty(n * s.npts, dtype=torch.float64, device=s.dev)
    outs = [e() for _ in range(7)]
    stream = torch.cuda.Stream()
    s.set_stream(stream.cuda_stream)
    try:
        at = AnchoredTraction(s, n, beta=7.6, tol=1e-3)

        def set_shape(k):
            with torch.cuda.stream(stream):
                for d, src in zip((x, y, z), shapes[k]):
                    d.copy_(src)
            stream.synchronize()

        eager = []
        for k in (0, 1, 1):
            set_shape(k)
            with torch.cuda.stream(stream):
                at.enqueue(x, y, z, *outs)
            stream.synchronize()
            eager.append([o.clone() for o in outs[:3]] + [at.b.clone()])
        at.reset()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            at.enqueue(x, y, z, *outs)
        at.reset()
        for i, k in enumerate((0, 1, 1)):
            set_shape(k)
            g.replay()
            stream.synchronize()
            junk = torch.randn(1024, device="cuda:0").sum()
            torch.cuda.synchronize()
            assert torch.isfinite(junk)
            for a, b in zip([o for o in outs[:3]] + [at.b], eager[i]):
                assert torch.equal(a, b), i
    finally:
        s.set_stream(s._own.cuda_stream)
