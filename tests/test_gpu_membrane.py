"""GPU: libmemb -- the FP64 membrane law against the CPU oracle, and the BF16-MFMA surrogate's
forward pass and gradients against the numpy oracle (tests/_mlp_oracle.py).

The law is FP64 on both sides with different summation orders (and FMA contraction on the
GPU), so the bar is a tolerance: 1e-10 of the largest traction component, 1e-12 for normals,
areas and shape coefficients.  The MLP kernels round operands to BF16 exactly where the
oracle does and accumulate in FP32 in another order: gradients must agree to 2e-3 in
relative Frobenius norm.
"""
import ctypes

import numpy as np
import pytest
from scipy.special import sph_harm_y

import _oracle as O
from _mlp_oracle import MLPOracle

pytestmark = pytest.mark.gpu

D = ctypes.POINTER(ctypes.c_double)


def _torch():
    import torch

    assert torch.cuda.is_available(), "no HIP device visible to torch"
    return torch


def oracle_lib():
    o = O.lib()
    o.mem_grid.argtypes = [ctypes.c_int, D, D]
    o.mem_traction.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double] + [D] * 10
    o.mem_analysis.argtypes = [ctypes.c_int, D, D, D]
    return o


def capsules(L, ncap, seed, amp=0.15):
    o = oracle_lib()
    nt, np_ = L + 1, 2 * (L + 1)
    th, w = np.zeros(nt), np.zeros(nt)
    o.mem_grid(L, th.ctypes.data_as(D), w.ctypes.data_as(D))
    ph = 2 * np.pi * np.arange(np_) / np_
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    rng = np.random.default_rng(seed)
    xs, ys, zs = [], [], []
    for c in range(ncap):
        r = np.ones_like(TH)
        for n in range(1, 6):
            for m in range(0, n + 1):
                r = r + rng.normal() * amp / n * np.real(sph_harm_y(n, m, TH, PH))
        xs.append((r * np.sin(TH) * np.sin(PH)).ravel() + rng.normal())
        ys.append((r * np.sin(TH) * np.cos(PH)).ravel() + rng.normal())
        zs.append((r * np.cos(TH)).ravel() + rng.normal())
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(zs)


@pytest.mark.parametrize("L,eb", [(16, 0.3), (16, 0.0), (8, 0.3), (4, 0.3), (23, 0.3)])
def test_membrane_law_matches_oracle(L, eb):
    torch = _torch()
    from suguru_amd.membrane import MembranePlan

    print(f"case: membrane law L={L} eb={eb}", flush=True)
    ncap = 5
    x, y, z = capsules(L, ncap, seed=L)
    plan = MembranePlan(L)
    plan.set_stream(torch.cuda.current_stream().cuda_stream)  # ordered after torch's fills
    n = plan.npts
    assert n == (L + 1) * 2 * (L + 1)
    dev = "cuda:0"
    dx, dy, dz = (torch.from_numpy(a).to(dev) for a in (x, y, z))
    outs = [torch.full((ncap * n,), np.nan, dtype=torch.float64, device=dev) for _ in range(7)]
    gs = 6250.0
    plan.traction_dev(dx, dy, dz, gs, eb * gs, *outs)
    plan.sync()
    got = [o.cpu().numpy() for o in outs]
    o = oracle_lib()
    for c in range(ncap):
        sl = slice(c * n, (c + 1) * n)
        want = [np.zeros(n) for _ in range(7)]
        xc, yc, zc = (np.ascontiguousarray(a[sl]) for a in (x, y, z))
        o.mem_traction(L, gs, eb * gs, xc.ctypes.data_as(D), yc.ctypes.data_as(D),
                       zc.ctypes.data_as(D), *[w.ctypes.data_as(D) for w in want])
        fmax = max(np.abs(want[k]).max() for k in range(3))
        for k in range(3):
            assert np.abs(got[k][sl] - want[k]).max() < 1e-10 * fmax, (c, k)
        for k in range(3, 7):
            assert np.abs(got[k][sl] - want[k]).max() < 1e-12 * max(1.0, np.abs(want[k]).max()), (c, k)
    plan.close()


@pytest.mark.parametrize("L", [8, 16, 23])
def test_geometry_matches_the_laws_areas_and_oracle(L):
    """memb_geometry_f64_dev (normals and areas only): the areas are the law's own (same
    formula from the coarse shape; 1e-12 of the largest against the oracle), the normals are
    unit vectors of the force grid, within the dealiasing's difference of the law's filtered
    normals, and the same for every capsule of a larger batch (persistent workgroups)."""
    torch = _torch()
    from suguru_amd.membrane import MembranePlan

    print(f"case: geometry L={L}", flush=True)
    ncap = 7
    x, y, z = capsules(L, ncap, seed=40 + L)
    plan = MembranePlan(L)
    plan.set_stream(torch.cuda.current_stream().cuda_stream)
    n = plan.npts
    dev = "cuda:0"
    dx, dy, dz = (torch.from_numpy(a).to(dev) for a in (x, y, z))
    geo = [torch.full((ncap * n,), np.nan, dtype=torch.float64, device=dev) for _ in range(4)]
    plan.geometry_dev(dx, dy, dz, *geo)
    plan.sync()
    gn = np.stack([g.cpu().numpy() for g in geo[:3]])
    ga = geo[3].cpu().numpy()
    o = oracle_lib()
    for c in range(ncap):
        sl = slice(c * n, (c + 1) * n)
        want = [np.zeros(n) for _ in range(7)]
        xc, yc, zc = (np.ascontiguousarray(a[sl]) for a in (x, y, z))
        o.mem_traction(L, 6250.0, 1875.0, xc.ctypes.data_as(D), yc.ctypes.data_as(D),
                       zc.ctypes.data_as(D), *[w.ctypes.data_as(D) for w in want])
        assert np.abs(ga[sl] - want[6]).max() < 1e-12 * np.abs(want[6]).max(), c
        assert np.abs(np.linalg.norm(gn[:, sl], axis=0) - 1.0).max() < 1e-13
        assert np.min((gn[:, sl] * np.stack(want[3:6])).sum(0)) > 0.99
    # a batch larger than one round of workgroups gives each capsule the same values
    reps = 3000
    big = [t.repeat(reps) for t in (dx, dy, dz)]
    out = [torch.empty(reps * ncap * n, dtype=torch.float64, device=dev) for _ in range(4)]
    plan.geometry_dev(*big, *out)
    plan.sync()
    for g, b in zip(geo, out):
        assert torch.equal(b.view(reps, -1), g.view(1, -1).expand(reps, -1))
    plan.close()


def test_shape_coefficients_match_oracle():
    torch = _torch()
    from suguru_amd.membrane import MembranePlan

    L = 16
    ncap = 3
    x, y, z = capsules(L, ncap, seed=3)
    plan = MembranePlan(L)
    plan.set_stream(torch.cuda.current_stream().cuda_stream)
    n, N = plan.npts, L + 1
    coef = torch.full((ncap * 3 * N * N,), np.nan, dtype=torch.float64, device="cuda:0")
    plan.coeffs_dev(*(torch.from_numpy(a).to("cuda:0") for a in (x, y, z)), coef)
    plan.sync()
    got = coef.cpu().numpy().reshape(ncap, 3, N * N)
    o = oracle_lib()
    for c in range(ncap):
        for k, arr in enumerate((x, y, z)):
            f = np.ascontiguousarray(arr[c * n:(c + 1) * n])
            a, b = np.zeros(N * N), np.zeros(N * N)
            o.mem_analysis(L, f.ctypes.data_as(D), a.ctypes.data_as(D), b.ctypes.data_as(D))
            packed = [a[m * N + nn] for m in range(N) for nn in range(m, N)] + \
                     [b[m * N + nn] for m in range(1, N) for nn in range(m, N)]
            assert np.abs(got[c, k] - np.array(packed)).max() < 1e-13


def _mlp_case(seed, din=867, width=512, nh=4, dout=1734, batch=128):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(batch, din)).astype(np.float32)
    t = rng.normal(size=(batch, dout)).astype(np.float32) * 3.0
    wt = rng.uniform(0.5, 1.5, dout).astype(np.float32)
    mean = rng.normal(size=din).astype(np.float32) * 0.1
    scale = rng.uniform(0.5, 2.0, din).astype(np.float32)
    return x, t, wt, mean, scale


def test_mlp_forward_and_gradients_match_oracle():
    torch = _torch()
    from suguru_amd.membrane import SurrogateMLP

    print("case: mlp forward/grad 867-512x4-1734, batch 128", flush=True)
    x, t, wt, mean, scale = _mlp_case(1)
    batch, din = x.shape
    dout = t.shape[1]
    m = SurrogateMLP(din, 512, 4, dout, max_batch=256)
    m.set_stream(torch.cuda.current_stream().cuda_stream)
    m.init(7)
    m.set_norm(mean, scale, 0.5, 3.0)
    layers = [m.get_layer(i) for i in range(m.nl)]
    ref = MLPOracle(layers, mean, scale, 3.0, out_mean=0.5)
    dev = "cuda:0"
    dx, dt_, dw = (torch.from_numpy(a).to(dev) for a in (x, t, wt))
    y = torch.empty((batch, dout), dtype=torch.float32, device=dev)
    m.forward(dx, y)
    torch.cuda.synchronize()
    yr = ref.predict(x)
    err = np.linalg.norm(y.cpu().numpy() - yr) / np.linalg.norm(yr)
    assert err < 2e-3, err
    loss = m.grad_only(dx, dt_, dw)
    lr_, g = ref.grads(x, t, wt)
    assert abs(loss - lr_) < 2e-3 * lr_
    for i in range(m.nl):
        gW, gb = m.get_grad(i)
        eW = np.linalg.norm(gW - g[i][0]) / np.linalg.norm(g[i][0])
        eb = np.linalg.norm(gb - g[i][1]) / np.linalg.norm(g[i][1])
        assert eW < 2e-3 and eb < 2e-3, (i, eW, eb)
    m.close()


def test_mlp_padded_shapes_and_training_reduces_loss():
    """Odd sizes exercise the zero padding; a few Adam steps on a fixed batch must lower the
    loss, and padded weights stay zero (the padded network equals the unpadded one)."""
    torch = _torch()
    from suguru_amd.membrane import SurrogateMLP

    x, t, wt, mean, scale = _mlp_case(2, din=37, width=70, nh=2, dout=45, batch=50)
    m = SurrogateMLP(37, 70, 2, 45, max_batch=64)
    m.set_stream(torch.cuda.current_stream().cuda_stream)
    m.init(3)
    m.set_norm(mean, scale, 0.5, 3.0)
    dev = "cuda:0"
    dx, dt_ = torch.from_numpy(x).to(dev), torch.from_numpy(t).to(dev)
    l0 = m.grad_only(dx, dt_)
    layers = [m.get_layer(i) for i in range(m.nl)]
    lr_, g = MLPOracle(layers, mean, scale, 3.0, out_mean=0.5).grads(x, t)
    assert abs(l0 - lr_) < 2e-3 * lr_
    for i in range(m.nl):
        gW, _ = m.get_grad(i)
        assert np.linalg.norm(gW - g[i][0]) < 2e-3 * np.linalg.norm(g[i][0])
    losses = [m.train_step(dx, dt_, 1e-3) for _ in range(30)]
    assert losses[-1] < 0.8 * losses[0], losses[::10]
    assert all(b < a for a, b in zip(losses[::10], losses[10::10])), losses[::10]
    y = torch.empty((50, 45), dtype=torch.float32, device=dev)
    m.forward(dx, y)
    torch.cuda.synchronize()
    ref = MLPOracle([m.get_layer(i) for i in range(m.nl)], mean, scale, 3.0, out_mean=0.5).predict(x)
    assert np.linalg.norm(y.cpu().numpy() - ref) < 3e-3 * np.linalg.norm(ref)
    m.close()
