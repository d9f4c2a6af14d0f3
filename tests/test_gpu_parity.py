"""GPU parity: libibx on gfx950 against the CPU oracle, bit for bit.

The library restates the reference's single-rank loop order (interp_spread.f90:315-450,
:817-912, :1274-1378) without fused multiply-adds, so the bar is bitwise equality of every
FP64 output -- stricter than the 1e-12 relative tolerance the task allows.  Sizes are those
the oracle finishes in seconds; the full 2000-capsule grid is checked through a bitwise
subset and through size-independent identities (conservation, adjointness, determinism).
"""
import numpy as np
import pytest

import _oracle as O
from suguru_amd import workload as W

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def assert_bitwise(got, want, what):
    g, w = bits(got), bits(want)
    if not np.array_equal(g, w):
        bad = np.nonzero(g != w)[0]
        gf = np.ascontiguousarray(got).reshape(-1)
        wf = np.ascontiguousarray(want).reshape(-1)
        rel = np.max(np.abs(gf[bad] - wf[bad])) / max(np.max(np.abs(wf)), 1e-300)
        raise AssertionError(f"{what}: {bad.size} of {g.size} values differ (max rel {rel:.3e}); "
                             f"first at {bad[:5]}")


def run_all(ctx, g, caps, u, v, w, f):
    fx, fy, fz = f
    off = caps.cap_off
    gi = ctx.eulr2lagr(u, v, w, off, caps.x, caps.y, caps.z)
    oi = O.eulr2lagr(g, u, v, w, off, caps.x, caps.y, caps.z)
    for a, b, nm in zip(gi, oi, ("ul", "vl", "wl")):
        assert_bitwise(a, b, nm)
    gs = ctx.lagr2eulr(off, caps.x, caps.y, caps.z, fx, fy, fz, caps.area)
    os_ = O.lagr2eulr(g, off, caps.x, caps.y, caps.z, fx, fy, fz, caps.area)
    for a, b, nm in zip(gs, os_, ("forcex", "forcey", "forcez")):
        assert_bitwise(a, b, nm)
    gc = ctx.lagr2eulrind(off, caps.x, caps.y, caps.z, fx, fy, fz, caps.area)
    oc = O.lagr2eulrind(g, off, caps.x, caps.y, caps.z, fx, fy, fz, caps.area)
    for a, b, nm in zip(gc, oc, ("gx", "gy", "gz")):
        assert_bitwise(a, b, nm)
    return gi, gs, gc


def test_plain_small_config_bitexact(gpu_ctx_factory):
    wl = W.config_small()
    print("case: 64^3, 1 capsule deg 12", flush=True)
    ctx = gpu_ctx_factory(wl.grid)
    run_all(ctx, wl.grid, wl.caps, wl.u, wl.v, wl.w, (wl.fx, wl.fy, wl.fz))


def test_plain_moderate_config_bitexact(gpu_ctx_factory):
    wl = W.config_moderate()
    print("case: 256x128x256, 64 capsules deg 16", flush=True)
    ctx = gpu_ctx_factory(wl.grid)
    run_all(ctx, wl.grid, wl.caps, wl.u, wl.v, wl.w, (wl.fx, wl.fy, wl.fz))


def test_dense_share_config_bitexact(gpu_ctx_factory):
    """One GPU's share of the dense 1024 x 512 x 256 channel: 128 x 512 x 256 cells, 1000
    capsules of radius 10.6h at 30 % volume fraction (578 000 points), all three transfers."""
    wl = W.config_dense_share()
    print("case: dense share 128x512x256, 1000 capsules deg 16", flush=True)
    ctx = gpu_ctx_factory(wl.grid)
    run_all(ctx, wl.grid, wl.caps, wl.u, wl.v, wl.w, (wl.fx, wl.fy, wl.fz))


def _edge_caps(g, seed):
    """Capsules straddling the periodic edges and corners, touching both walls, one empty,
    uneven degrees; plus points on exact lattice and half-lattice coordinates."""
    centers = [
        (0.0, 0.0, 0.5 * g.lz),                 # corner, straddles both periodic seams
        (g.lx - 0.5 * g.dx, 0.3 * g.ly, 2.0 * g.dz),  # right seam, near the bottom wall
        (0.4 * g.lx, g.ly, g.lz - 2.5 * g.dz),   # back seam, near the top wall
        (0.5 * g.lx, 0.5 * g.ly, 0.5 * g.lz),
    ]
    radii = [3.0 * g.dx, 2.5 * g.dx, 3.5 * g.dx, 4.0 * g.dx]
    degs = [8, 12, 6, 16]
    caps = W.build_capsules(centers, radii, degs)
    # insert an empty capsule and lattice-exact points
    x, y, z, a = caps.x.copy(), caps.y.copy(), caps.z.copy(), caps.area.copy()
    n0 = int(caps.cap_off[1])
    x[:4] = [0.0, g.lx, 2.0 * g.dx, 2.5 * g.dx]
    y[:4] = [g.ly, 0.0, 3.5 * g.dy, 4.0 * g.dy]
    z[4:8] = [0.0, g.lz, 0.5 * g.dz, g.dz]
    off = np.concatenate([[0, n0, n0], caps.cap_off[2:]]).astype(np.int64)
    caps.cap_off, caps.x, caps.y, caps.z, caps.area = off, x, y, z, a
    return caps


@pytest.mark.parametrize("shape", [(32, 24, 16), (20, 12, 10), (37, 19, 23)])
def test_edges_uneven_odd_grids_bitexact(gpu_ctx_factory, shape):
    g = W.make_grid(*shape, h=1.0 / 32)
    print("case: edge capsules on grid", shape, flush=True)
    caps = _edge_caps(g, 3)
    u, v, w = W.velocity_field(g, 5)
    f = W.forces(caps.npts, 6)
    ctx = gpu_ctx_factory(g)
    run_all(ctx, g, caps, u, v, w, f)


def test_dense_tile_overflow_path_bitexact(gpu_ctx_factory):
    """More than 2048 candidates in one tile: the in-HBM sort path and many 64-wide chunks."""
    g = W.make_grid(32, 32, 32, 1.0 / 32)
    rng = np.random.default_rng(9)
    n = 6000
    c = np.array([10.2, 11.7, 9.4]) * g.dx
    x = c[0] + rng.uniform(-1.0, 1.0, n) * g.dx
    y = c[1] + rng.uniform(-1.0, 1.0, n) * g.dy
    z = c[2] + rng.uniform(-1.0, 1.0, n) * g.dz
    caps = W.Capsules(np.array([0, 2500, 6000], dtype=np.int64), x, y, z,
                      rng.uniform(0.5, 1.5, n) * g.dx * g.dy, np.zeros((2, 3)), np.ones(2),
                      np.ones(2, dtype=np.int64))
    print("case: 6000 points in one tile", flush=True)
    u, v, w = W.velocity_field(g, 1)
    f = W.forces(n, 2)
    ctx = gpu_ctx_factory(g)
    run_all(ctx, g, caps, u, v, w, f)


def test_empty_and_nonfinite(gpu_ctx_factory):
    from suguru_amd import IbxError

    g = W.make_grid(16, 16, 16, 1.0 / 16)
    ctx = gpu_ctx_factory(g)
    u, v, w = W.velocity_field(g, 1)
    e = np.zeros(0)
    ul, vl, wl = ctx.eulr2lagr(u, v, w, [0], e, e, e)
    assert ul.size == 0
    fx, fy, fz = ctx.lagr2eulr([0], e, e, e, e, e, e, e)
    assert not fx.any() and not fy.any() and not fz.any()
    x = np.array([0.3, np.nan, 0.2])
    y = np.array([0.3, 0.3, 0.2])
    z = np.array([0.3, 0.3, np.inf])
    with pytest.raises(IbxError) as ei:
        ctx.eulr2lagr(u, v, w, [0, 3], x, y, z)
    assert ei.value.code == -3
    with pytest.raises(IbxError) as ei:
        ctx.lagr2eulr([0, 3], x, y, z, x, y, z, y)
    assert ei.value.code == -3
    # the context stays usable
    x = np.array([0.3])
    ctx.eulr2lagr(u, v, w, [0, 1], x, x, x)


def test_update_pos_bitexact(gpu_ctx_factory):
    wl = W.config_moderate(with_fields=False)
    g = wl.grid
    caps = wl.caps
    rng = np.random.default_rng(4)
    n = caps.npts
    ul, vl, wl_ = rng.standard_normal((3, n)) * 50.0
    # push a few capsules across the periodic seams and some points into the walls
    x0 = caps.x.copy()
    x0[: caps.cap_off[1]] += g.lx - caps.centers[0, 0] + 0.001 * g.dx
    y0 = caps.y.copy()
    z0 = caps.z.copy()
    z0[5] = 1e-9
    wl_[5] = -1.0
    xo = x0 - rng.uniform(-1, 1, n) * 0.6 * g.lx
    yo = y0 + rng.uniform(-1, 1, n) * 0.6 * g.ly
    ctx = gpu_ctx_factory(g)
    a = [x0.copy(), y0.copy(), z0.copy(), xo.copy(), yo.copy()]
    b = [x0.copy(), y0.copy(), z0.copy(), xo.copy(), yo.copy()]
    cg, ug, ncg = ctx.update_pos(caps.cap_off, a[0], a[1], a[2], ul, vl, wl_, 1e-4, a[3], a[4])
    co, uo, nco = O.update_pos(g, caps.cap_off, 1e-4, b[0], b[1], b[2], ul, vl, wl_, b[3], b[4])
    assert ncg == nco >= 1
    for p, q, nm in zip(a + [cg, ug], b + [co, uo], ("x", "y", "z", "xold", "yold", "cen", "umean")):
        assert_bitwise(p, q, nm)


def test_full_size_uneven_batch_subset_bitexact_and_identities(gpu_ctx_factory):
    """512 x 256 x 512 channel with its 2000 uneven capsules: a bitwise subset against the
    oracle, and size-independent identities on the full batch."""
    wl = W.config_uneven(with_fields=False)
    g, caps = wl.grid, wl.caps
    print("case: 512x256x512 uneven 2000 capsules", flush=True)
    u, v, w = W.velocity_field(g, 3, nmodes=2)
    ctx = gpu_ctx_factory(g)
    # subset of 40 capsules: bitwise interp + spread on the full-size grid
    sel = np.arange(0, 2000, 50)
    xs, ys, zs, As, off = [], [], [], [], [0]
    for p in sel:
        s = slice(caps.cap_off[p], caps.cap_off[p + 1])
        xs.append(caps.x[s]), ys.append(caps.y[s]), zs.append(caps.z[s]), As.append(caps.area[s])
        off.append(off[-1] + xs[-1].size)
    sub = W.Capsules(np.array(off, dtype=np.int64), np.concatenate(xs), np.concatenate(ys),
                     np.concatenate(zs), np.concatenate(As), caps.centers[sel], caps.radii[sel],
                     caps.degrees[sel])
    fs = W.forces(sub.npts, 8)
    gi = ctx.eulr2lagr(u, v, w, sub.cap_off, sub.x, sub.y, sub.z)
    oi = O.eulr2lagr(g, u, v, w, sub.cap_off, sub.x, sub.y, sub.z)
    for a, b, nm in zip(gi, oi, ("ul", "vl", "wl")):
        assert_bitwise(a, b, nm)
    gs = ctx.lagr2eulr(sub.cap_off, sub.x, sub.y, sub.z, *fs, sub.area)
    os_ = O.lagr2eulr(g, sub.cap_off, sub.x, sub.y, sub.z, *fs, sub.area)
    for a, b, nm in zip(gs, os_, ("forcex", "forcey", "forcez")):
        assert_bitwise(a, b, nm)
    del gs, os_
    # full batch: conservation and adjointness, and run-to-run bitwise determinism
    f = W.forces(caps.npts, 9)
    F1 = ctx.lagr2eulr(caps.cap_off, caps.x, caps.y, caps.z, *f, caps.area)
    dV = g.dx * g.dy * g.dz
    inner = (slice(1, -1), slice(1, -1), slice(1, -1))
    ul = ctx.eulr2lagr(u, v, w, caps.cap_off, caps.x, caps.y, caps.z)
    for F, fl, a, l in zip(F1, f, (u, v, w), ul):
        tot = (fl * caps.area).sum()
        assert abs(F[inner].sum() * dV - tot) < 1e-11 * np.abs(fl * caps.area).sum()
        lhs = np.sum(a[inner] * F[inner]) * dV
        rhs = np.sum(l * fl * caps.area)
        assert abs(lhs - rhs) < 1e-11 * np.abs(a).max() * np.abs(fl * caps.area).sum()
    F2 = ctx.lagr2eulr(caps.cap_off, caps.x, caps.y, caps.z, *f, caps.area)
    for a, b, nm in zip(F1, F2, "xyz"):
        assert_bitwise(a, b, "repeat force" + nm)


def test_dense_overlapping_capsules_capacity_and_fallback(gpu_ctx_factory):
    """40 capsules packed into one spot: every work item of the object spread sees 40 source
    objects, more than its 32.  Object configurations then report IBX_E_CAPACITY, the
    host entry points retry with the tile spread, and the batch autotuner skips them; every
    result that is returned is the oracle's, bit for bit."""
    import torch
    from suguru_amd import Batch, IbxError

    g = W.make_grid(32, 32, 32, 1.0 / 32)
    rng = np.random.default_rng(11)
    ncap, per = 40, 60
    n = ncap * per
    c = np.array([15.3, 16.8, 14.1]) * g.dx
    x = c[0] + rng.uniform(-2.0, 2.0, n) * g.dx
    y = c[1] + rng.uniform(-2.0, 2.0, n) * g.dy
    z = c[2] + rng.uniform(-2.0, 2.0, n) * g.dz
    caps = W.Capsules(np.arange(0, n + 1, per, dtype=np.int64), x, y, z,
                      rng.uniform(0.5, 1.5, n) * g.dx * g.dy, np.zeros((ncap, 3)), np.ones(ncap),
                      np.ones(ncap, dtype=np.int64))
    print("case: 40 overlapping capsules, host entry points", flush=True)
    u, v, w = W.velocity_field(g, 1)
    f = W.forces(n, 2)
    ctx = gpu_ctx_factory(g)
    run_all(ctx, g, caps, u, v, w, f)
    # the batch: object configurations overflow, the autotuner picks a tile one
    dev = "cuda:0"
    up = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).reshape(-1, order="F"))).to(dev)
    t = {"u": up(u), "v": up(v), "w": up(w), "fx": up(f[0]), "fy": up(f[1]), "fz": up(f[2])}
    for k in ("x", "y", "z", "area"):
        t[k] = up(getattr(caps, k))
    for k in ("ul", "vl", "wl"):
        t[k] = torch.empty(n, dtype=torch.float64, device=dev)
    for k in ("forcex", "forcey", "forcez"):
        t[k] = torch.empty(g.ncells_halo, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    b = Batch(ctx, caps.cap_off)
    try:
        b.bind(**t)
        names = Batch.config_names()
        print("case: 40 overlapping capsules, object configuration overflows", flush=True)
        b.set_config(names.index("objects-w4"))
        with pytest.raises(IbxError) as ei:  # at the capture's sizing run, or at the sync
            b.capture(use_graph=False)
            b.run()
            ctx.sync()
        assert ei.value.code == -4
        print("case: 40 overlapping capsules, autotune", flush=True)
        choice = b.autotune(None)
        assert not names[choice].startswith("objects"), names[choice]
        b.capture(use_graph=True)
        b.run()
        ctx.sync()
        want = O.lagr2eulr(g, caps.cap_off, caps.x, caps.y, caps.z, *f, caps.area)
        for k, r in zip(("forcex", "forcey", "forcez"), want):
            assert_bitwise(t[k].cpu().numpy(), r.reshape(-1, order="F"), k)
    finally:
        b.close()


def test_device_kernel_sweep_and_kerneltest(gpu_ctx_factory):
    """The device kernel (the one every weight uses) against the reference's own compiled
    kernel.f90 values (tests/golden/kernel_ref.npz) and the oracle on a dense sweep that holds
    the branch points |r| = 0.5 and 1.5 exactly and their neighbouring doubles; and kerneltest
    (kernel.f90:43-63) summed on the device equals out.log:1 bit for bit."""
    import ctypes
    import os

    import torch
    from suguru_amd import _lib

    lib = _lib.load()
    ctx = gpu_ctx_factory(W.config_small(with_fields=False).grid)
    gold = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kernel_ref.npz"))
    edges = np.array([0.5, 1.5, -0.5, -1.5, 0.0, -0.0, 1.0, -1.0])
    near = np.concatenate([np.nextafter(edges, np.inf), np.nextafter(edges, -np.inf)])
    sweep = np.concatenate([gold["r"], edges, near, np.linspace(-2.0, 2.0, 100001),
                            np.random.default_rng(9).uniform(-1.6, 1.6, 50000)])
    r = torch.from_numpy(sweep).to("cuda:0")
    out = torch.empty_like(r)
    print(f"case: device kernel sweep, {sweep.size} arguments", flush=True)
    p = lambda t: ctypes.c_void_p(int(t.data_ptr()))
    assert lib.ibx_kernel_dev(ctx.h, int(r.numel()), p(r), p(out)) == 0
    ctx.sync()
    got = out.cpu().numpy()
    ng = gold["r"].size
    assert np.array_equal(bits(got[:ng]), bits(gold["kernel"]))
    want = np.array([O.kernel(float(x)) for x in sweep[ng:]])
    assert np.array_equal(bits(got[ng:]), bits(want))
    print("case: device kerneltest", flush=True)
    s = ctypes.c_double()
    assert lib.ibx_kerneltest_dev(ctx.h, ctypes.byref(s)) == 0
    assert s.value == float("0.99999999999996703") == float(gold["kerneltest"][0])


def test_full_size_dense_channel_subset_bitexact_and_identities(gpu_ctx_factory):
    """The dense coupled channel at full size: 1024 x 512 x 256 cells, 8000 capsules of
    degree 16 (4.62 M points) at 30 %.  A 40-capsule subset is bitwise against the oracle on
    the full grid; the whole batch runs through the batched path (capsule-object spread,
    captured hipGraph), where conservation, adjointness, replay-to-replay determinism and
    graph = eager are checked on the device."""
    import torch
    from suguru_amd import Batch

    wl = W.config_dense_full(with_fields=False)
    g, caps = wl.grid, wl.caps
    print("case: dense 1024x512x256, 8000 capsules", flush=True)
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(17)
    nf = g.ncells_halo
    dev = "cuda:0"
    t = {k: torch.randn(nf, dtype=torch.float64, device=dev, generator=gen) for k in ("u", "v", "w")}
    host = {k: t[k].cpu().numpy().reshape(g.shape_f, order="F") for k in ("u", "v", "w")}
    ctx = gpu_ctx_factory(g)
    sel = np.arange(0, caps.ncap, 200)
    xs, ys, zs, As, off = [], [], [], [], [0]
    for p in sel:
        s = slice(caps.cap_off[p], caps.cap_off[p + 1])
        xs.append(caps.x[s]), ys.append(caps.y[s]), zs.append(caps.z[s]), As.append(caps.area[s])
        off.append(off[-1] + xs[-1].size)
    sub = W.Capsules(np.array(off, dtype=np.int64), np.concatenate(xs), np.concatenate(ys),
                     np.concatenate(zs), np.concatenate(As), caps.centers[sel], caps.radii[sel],
                     caps.degrees[sel])
    fs = W.forces(sub.npts, 8)
    print("case: dense full, 40-capsule subset vs oracle", flush=True)
    gi = ctx.eulr2lagr(host["u"], host["v"], host["w"], sub.cap_off, sub.x, sub.y, sub.z)
    oi = O.eulr2lagr(g, host["u"], host["v"], host["w"], sub.cap_off, sub.x, sub.y, sub.z)
    for a, b, nm in zip(gi, oi, ("ul", "vl", "wl")):
        assert_bitwise(a, b, nm)
    gs = ctx.lagr2eulr(sub.cap_off, sub.x, sub.y, sub.z, *fs, sub.area)
    os_ = O.lagr2eulr(g, sub.cap_off, sub.x, sub.y, sub.z, *fs, sub.area)
    for a, b, nm in zip(gs, os_, ("forcex", "forcey", "forcez")):
        assert_bitwise(a, b, nm)
    del gs, os_, host
    # the whole batch on the device
    print("case: dense full, whole batch, objects-w8 graph", flush=True)
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    f = W.forces(caps.npts, 9)
    # capsules of the lattice come within 1.1 cells of the walls, where the clamped windows
    # drop taps: conservation is checked with the forces of those points set to zero
    zf = caps.z * g.dzi
    near = (zf < 3.0) | (zf > g.ktot - 3.0)
    assert near.any() and not near.all()
    for a in f:
        a[near] = 0.0
    for k, a in zip(("fx", "fy", "fz"), f):
        t[k] = up(a)
    for k in ("x", "y", "z", "area"):
        t[k] = up(getattr(caps, k))
    n = caps.npts
    for k in ("ul", "vl", "wl"):
        t[k] = torch.empty(n, dtype=torch.float64, device=dev)
    for k in ("forcex", "forcey", "forcez"):
        t[k] = torch.full((nf,), np.nan, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    b = Batch(ctx, caps.cap_off)
    try:
        b.bind(**t)
        b.set_config(Batch.config_names().index("objects-w8"))
        b.capture(use_graph=True)
        b.run()
        ctx.sync()
        first = [t[k].clone() for k in ("ul", "vl", "wl", "forcex", "forcey", "forcez")]
        dV = g.dx * g.dy * g.dz
        shp = (g.ktot + 2, g.jtot + 2, g.itot + 2)
        inner = lambda a: a.view(shp)[1:-1, 1:-1, 1:-1]
        for comp, (vel, fl, l, F) in enumerate(zip(("u", "v", "w"), ("fx", "fy", "fz"),
                                                   ("ul", "vl", "wl"),
                                                   ("forcex", "forcey", "forcez"))):
            fa = t[fl] * t["area"]
            tot = float(fa.sum())
            assert abs(float(inner(t[F]).sum()) * dV - tot) < 1e-11 * float(fa.abs().sum()), comp
            lhs = float((inner(t[vel]) * inner(t[F])).sum()) * dV
            rhs = float((t[l] * fa).sum())
            assert abs(lhs - rhs) < 1e-11 * float(t[vel].abs().max()) * float(fa.abs().sum()), comp
        for rep in range(2):
            for k in ("ul", "vl", "wl", "forcex", "forcey", "forcez"):
                t[k].fill_(np.nan)
            torch.cuda.synchronize()
            b.run()
            ctx.sync()
            for k, a in zip(("ul", "vl", "wl", "forcex", "forcey", "forcez"), first):
                assert torch.equal(t[k].view(torch.int64), a.view(torch.int64)), (rep, k)
        print("case: dense full, eager batch", flush=True)
        b.capture(use_graph=False)
        b.run()
        ctx.sync()
        for k, a in zip(("ul", "vl", "wl", "forcex", "forcey", "forcez"), first):
            assert torch.equal(t[k].view(torch.int64), a.view(torch.int64)), ("eager", k)
    finally:
        b.close()
