"""GPU: the HBM-resident entry points and the batched, graph-captured path.

Every tile configuration of the autotuner, eager or replayed from a hipGraph, must give the
same bits as the CPU oracle: the tile shape changes only which workgroup computes a cell,
never the order in which the cell receives its terms.
"""
import numpy as np
import pytest

import _oracle as O
from suguru_amd import workload as W

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    assert torch.cuda.is_available(), "no HIP device visible to torch"
    return torch


def bits(a):
    return np.ascontiguousarray(a).reshape(-1).view(np.uint64)


def to_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).reshape(-1, order="F"))).to("cuda:0")


@pytest.fixture(scope="module")
def moderate():
    wl = W.config_moderate()
    g, caps = wl.grid, wl.caps
    ref_i = O.eulr2lagr(g, wl.u, wl.v, wl.w, caps.cap_off, caps.x, caps.y, caps.z)
    ref_s = O.lagr2eulr(g, caps.cap_off, caps.x, caps.y, caps.z, wl.fx, wl.fy, wl.fz, caps.area)
    return wl, ref_i, ref_s


def test_device_entry_points_bitexact(gpu_ctx_factory, moderate):
    torch = _torch()
    wl, ref_i, ref_s = moderate
    g, caps = wl.grid, wl.caps
    print("case: device entry points, 256x128x256", flush=True)
    ctx = gpu_ctx_factory(g)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    d = {k: to_dev(torch, getattr(wl, k)) for k in ("u", "v", "w", "fx", "fy", "fz")}
    for k in ("x", "y", "z", "area"):
        d[k] = to_dev(torch, getattr(caps, k))
    off = torch.from_numpy(caps.cap_off).to("cuda:0")
    n = caps.npts
    ul, vl, wl_ = (torch.empty(n, dtype=torch.float64, device="cuda:0") for _ in range(3))
    F = [torch.full((g.ncells_halo,), np.nan, dtype=torch.float64, device="cuda:0") for _ in range(3)]
    ctx.interp_dev(d["u"], d["v"], d["w"], off, d["x"], d["y"], d["z"], ul, vl, wl_)
    ctx.spread_dev(off, d["x"], d["y"], d["z"], d["fx"], d["fy"], d["fz"], d["area"], *F)
    ctx.sync()
    torch.cuda.synchronize()
    for a, b in zip((ul, vl, wl_), ref_i):
        assert np.array_equal(bits(a.cpu().numpy()), bits(b))
    for a, b in zip(F, ref_s):
        assert np.array_equal(bits(a.cpu().numpy()), bits(b.reshape(-1, order="F")))
    # the cell-centred spread writes every element too
    G = [torch.full((g.ncells_halo,), np.nan, dtype=torch.float64, device="cuda:0") for _ in range(3)]
    ctx.spread_dev(off, d["x"], d["y"], d["z"], d["fx"], d["fy"], d["fz"], d["area"], *G, cc=True)
    ctx.sync()
    ref_c = O.lagr2eulrind(g, caps.cap_off, caps.x, caps.y, caps.z, wl.fx, wl.fy, wl.fz, caps.area)
    for a, b in zip(G, ref_c):
        assert np.array_equal(bits(a.cpu().numpy()), bits(b.reshape(-1, order="F")))
    ctx.set_stream(0)


def test_batch_every_config_graph_and_eager_bitexact(gpu_ctx_factory, moderate):
    torch = _torch()
    from suguru_amd import Batch

    wl, ref_i, ref_s = moderate
    g, caps = wl.grid, wl.caps
    ctx = gpu_ctx_factory(g)
    t = {k: to_dev(torch, getattr(wl, k)) for k in ("u", "v", "w", "fx", "fy", "fz")}
    for k in ("x", "y", "z", "area"):
        t[k] = to_dev(torch, getattr(caps, k))
    n = caps.npts
    for k in ("ul", "vl", "wl"):
        t[k] = torch.empty(n, dtype=torch.float64, device="cuda:0")
    for k in ("forcex", "forcey", "forcez"):
        t[k] = torch.empty(g.ncells_halo, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    b = Batch(ctx, caps.cap_off)
    try:
        b.bind(**t)
        names = Batch.config_names()
        want_i = [bits(a) for a in ref_i]
        want_s = [bits(a.reshape(-1, order="F")) for a in ref_s]
        for cfg, name in enumerate(names):
            for graph in (True, False):
                print("case: batch config", name, "graph" if graph else "eager", flush=True)
                b.set_config(cfg)
                b.capture(use_graph=graph)
                for rep in range(3):
                    for k in ("ul", "vl", "wl", "forcex", "forcey", "forcez"):
                        t[k].fill_(np.nan)  # torch's eager work between replays
                    torch.cuda.synchronize()
                    b.run()
                    ctx.sync()
                    for k, w in zip(("ul", "vl", "wl"), want_i):
                        assert np.array_equal(bits(t[k].cpu().numpy()), w), (name, graph, rep, k)
                    for k, w in zip(("forcex", "forcey", "forcez"), want_s):
                        assert np.array_equal(bits(t[k].cpu().numpy()), w), (name, graph, rep, k)
        prof = b.profile(reps=3)
        assert prof["total_us"] > 0
        # interpolation on the side stream, concurrent with the spread: same bits
        b.set_config(0)
        b.set_overlap(True)
        for graph in (True, False):
            print("case: overlap", "graph" if graph else "eager", flush=True)
            b.capture(use_graph=graph)
            for rep in range(2):
                for k in ("ul", "vl", "wl", "forcex", "forcey", "forcez"):
                    t[k].fill_(np.nan)
                torch.cuda.synchronize()
                b.run()
                ctx.sync()
                torch.cuda.synchronize()
                for k, w in zip(("ul", "vl", "wl"), want_i):
                    assert np.array_equal(bits(t[k].cpu().numpy()), w), ("overlap", graph, rep, k)
                for k, w in zip(("forcex", "forcey", "forcez"), want_s):
                    assert np.array_equal(bits(t[k].cpu().numpy()), w), ("overlap", graph, rep, k)
        b.set_overlap(False)
        # timed graph instances (the benchmark's timed region): same bits, one timing record each
        print("case: timed slots", flush=True)
        b.set_config(0)
        b.capture_timed(3)
        b.reset_slots()
        for rep in range(3):
            for k in ("forcex", "forcey", "forcez"):
                t[k].fill_(np.nan)
            torch.cuda.synchronize()
            b.run()
            ctx.sync()
            for k, w in zip(("forcex", "forcey", "forcez"), want_s):
                assert np.array_equal(bits(t[k].cpu().numpy()), w), ("timed", rep, k)
        times = b.slot_times()
        assert len(times) == 3
        for s in times:
            assert 0 < s["tile_spread_us"] < s["step_us"] and s["interp_us"] > 0
    finally:
        b.close()


def test_batch_replays_follow_moving_points(gpu_ctx_factory):
    """The graph reads the bound buffers at replay time: after Update_Pos moves the points in
    place, the next replay equals the oracle at the new positions."""
    torch = _torch()
    from suguru_amd import Batch

    wl = W.config_small()
    g, caps = wl.grid, wl.caps
    ctx = gpu_ctx_factory(g)
    t = {k: to_dev(torch, getattr(wl, k)) for k in ("u", "v", "w", "fx", "fy", "fz")}
    for k in ("x", "y", "z", "area"):
        t[k] = to_dev(torch, getattr(caps, k))
    for k in ("ul", "vl", "wl"):
        t[k] = torch.empty(caps.npts, dtype=torch.float64, device="cuda:0")
    for k in ("forcex", "forcey", "forcez"):
        t[k] = torch.empty(g.ncells_halo, dtype=torch.float64, device="cuda:0")
    off = torch.from_numpy(caps.cap_off).to("cuda:0")
    cen = torch.empty((caps.ncap, 3), dtype=torch.float64, device="cuda:0")
    um = torch.empty((caps.ncap, 3), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    b = Batch(ctx, caps.cap_off)
    try:
        b.bind(**t)
        b.capture(use_graph=True)
        x, y, z = caps.x.copy(), caps.y.copy(), caps.z.copy()
        for step in range(4):
            b.run()
            ctx.update_pos_dev(off, t["x"], t["y"], t["z"], t["ul"], t["vl"], t["wl"], 2e-3, cen, um)
            ctx.sync()
            ri = O.eulr2lagr(g, wl.u, wl.v, wl.w, caps.cap_off, x, y, z)
            O.update_pos(g, caps.cap_off, 2e-3, x, y, z, *[np.array(a) for a in ri])
            for k, a in zip("xyz", (x, y, z)):
                assert np.array_equal(bits(t[k].cpu().numpy()), bits(a)), (step, k)
    finally:
        b.close()


def test_nonfinite_forces_rejected_every_config(gpu_ctx_factory):
    """A non-finite force or area is an input error (IBX_E_DOMAIN), as a non-finite position
    is: the reference would spread NaN over its whole window [nint(c-2), nint(c+2)], whose
    outer taps carry weight 0 and which this build does not visit.  After the error the
    batch keeps working and gives the oracle's bits again on finite input."""
    torch = _torch()
    from suguru_amd import Batch, IbxError

    wl = W.config_small()
    g, caps = wl.grid, wl.caps
    ref = [a.reshape(-1, order="F") for a in
           O.lagr2eulr(g, caps.cap_off, caps.x, caps.y, caps.z, wl.fx, wl.fy, wl.fz, caps.area)]
    ctx = gpu_ctx_factory(g)
    t = {k: to_dev(torch, getattr(wl, k)) for k in ("u", "v", "w", "fx", "fy", "fz")}
    for k in ("x", "y", "z", "area"):
        t[k] = to_dev(torch, getattr(caps, k))
    for k in ("ul", "vl", "wl"):
        t[k] = torch.empty(caps.npts, dtype=torch.float64, device="cuda:0")
    for k in ("forcex", "forcey", "forcez"):
        t[k] = torch.empty(g.ncells_halo, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    b = Batch(ctx, caps.cap_off)
    b.bind(**t)
    try:
        for cfg, name in enumerate(Batch.config_names()):
            b.set_config(cfg)
            b.capture(use_graph=False)
            for key, idx, val in (("fx", 3, np.nan), ("fy", 50, np.inf), ("area", 200, np.nan)):
                print("case: non-finite", key, "config", name, flush=True)
                keep = float(t[key][idx])
                t[key][idx] = val
                torch.cuda.synchronize()
                b.run()
                with pytest.raises(IbxError) as ei:
                    ctx.sync()
                assert ei.value.code == -3, (name, key)
                t[key][idx] = keep
                torch.cuda.synchronize()
            b.run()
            ctx.sync()
            for k, r in zip(("forcex", "forcey", "forcez"), ref):
                assert np.array_equal(bits(t[k].cpu().numpy()), bits(r)), (name, k)
    finally:
        b.close()


def test_context_closed_before_its_batch():
    """ibx_destroy also destroys the context's live batches; the Python Batch then becomes
    a no-op to close (no double free, no use of the freed context)."""
    from suguru_amd import Batch, IBTransfer

    wl = W.config_small(with_fields=False)
    ctx = IBTransfer(wl.grid, 0)
    b = Batch(ctx, wl.caps.cap_off)
    ctx.close()
    assert b.h is None
    b.close()
    del b


def test_dense_share_every_config_bitexact(gpu_ctx_factory):
    """One GPU's share of the dense channel (128 x 512 x 256 cells, 1000 close-packed capsules
    at 30 % volume fraction, 578 000 points): every configuration's batched interpolation and
    spread, eager and replayed from a graph, against the oracle bit for bit.  Close packing
    gives items several source capsules, the case the per-item source lists exist for."""
    torch = _torch()
    from suguru_amd import Batch

    wl = W.config_dense_share()
    g, caps = wl.grid, wl.caps
    print("case: dense share 128x512x256, 1000 capsules, oracle", flush=True)
    ref_i = O.eulr2lagr(g, wl.u, wl.v, wl.w, caps.cap_off, caps.x, caps.y, caps.z)
    ref_s = O.lagr2eulr(g, caps.cap_off, caps.x, caps.y, caps.z, wl.fx, wl.fy, wl.fz, caps.area)
    ctx = gpu_ctx_factory(g)
    t = {k: to_dev(torch, getattr(wl, k)) for k in ("u", "v", "w", "fx", "fy", "fz")}
    for k in ("x", "y", "z", "area"):
        t[k] = to_dev(torch, getattr(caps, k))
    for k in ("ul", "vl", "wl"):
        t[k] = torch.empty(caps.npts, dtype=torch.float64, device="cuda:0")
    for k in ("forcex", "forcey", "forcez"):
        t[k] = torch.empty(g.ncells_halo, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    b = Batch(ctx, caps.cap_off)
    b.bind(**t)
    sy, sz = g.itot + 2, (g.itot + 2) * (g.jtot + 2)
    failures = []
    try:
        for cfg, name in enumerate(Batch.config_names()):
            for graph in (False, True):
                print("case: dense share, config", name, "graph", graph, flush=True)
                for k in ("ul", "vl", "wl", "forcex", "forcey", "forcez"):
                    t[k].fill_(np.nan)
                torch.cuda.synchronize()
                b.set_config(cfg)
                b.capture(use_graph=graph)
                b.run()
                ctx.sync()
                for k, r in list(zip(("ul", "vl", "wl"), ref_i)) + list(
                        zip(("forcex", "forcey", "forcez"), ref_s)):
                    got = t[k].cpu().numpy()
                    want = np.asarray(r).reshape(-1, order="F")  # fields: i fastest
                    bad = np.nonzero(bits(got) != bits(want))[0]
                    if bad.size:
                        m = int(bad[0])
                        where = (m % sy, (m // sy) % (g.jtot + 2), m // sz) if k.startswith("f") else m
                        msg = (f"{name} graph={graph} {k}: {bad.size} differ, {int(np.isnan(got[bad]).sum())} NaN; "
                               f"first {where}: got {got[m]!r} want {want[m]!r}")
                        print(msg, flush=True)
                        failures.append(msg)
    finally:
        b.close()
    assert not failures, failures[:6]


def test_device_wrappers_reject_wrong_sizes(gpu_ctx_factory):
    """The device entry points take raw pointers; the Python wrappers check every tensor's
    size, dtype and placement before the call (a short field would otherwise be read out of
    bounds), and nothing reaches the library."""
    torch = _torch()
    from suguru_amd import Batch

    wl = W.config_small(with_fields=False)
    g, caps = wl.grid, wl.caps
    ctx = gpu_ctx_factory(g)
    e = lambda n, dt=torch.float64: torch.zeros(n, dtype=dt, device="cuda:0")
    n, nf = caps.npts, g.ncells_halo
    off = torch.from_numpy(caps.cap_off).to("cuda:0")
    ok = dict(u=e(nf), v=e(nf), w=e(nf), x=e(n), y=e(n), z=e(n), ul=e(n), vl=e(n), wl=e(n))
    with pytest.raises(ValueError, match="u:"):
        ctx.interp_dev(e(nf - 1), ok["v"], ok["w"], off, ok["x"], ok["y"], ok["z"], ok["ul"],
                       ok["vl"], ok["wl"])
    with pytest.raises(ValueError, match="forcez:"):
        ctx.spread_dev(off, ok["x"], ok["y"], ok["z"], e(n), e(n), e(n), e(n), e(nf), e(nf),
                       e(nf // 2))
    with pytest.raises(ValueError, match="area:"):
        ctx.spread_dev(off, ok["x"], ok["y"], ok["z"], e(n), e(n), e(n), e(n, torch.float32),
                       e(nf), e(nf), e(nf))
    with pytest.raises(ValueError, match="centroid:"):
        ctx.update_pos_dev(off, ok["x"], ok["y"], ok["z"], ok["ul"], ok["vl"], ok["wl"], 1e-3,
                           e(caps.ncap), e(3 * caps.ncap))
    with pytest.raises(ValueError, match="x:"):
        ctx.interp_dev(ok["u"], ok["v"], ok["w"], off, ok["x"].cpu(), ok["y"], ok["z"],
                       ok["ul"], ok["vl"], ok["wl"])
    b = Batch(ctx, caps.cap_off)
    try:
        bufs = dict(ok, fx=e(n), fy=e(n), fz=e(n), area=e(n), forcex=e(nf), forcey=e(nf),
                    forcez=e(nf - 16))
        with pytest.raises(ValueError, match="forcez:"):
            b.bind(**bufs)
    finally:
        b.close()


def test_set_config_of_the_one_call_entry_points(gpu_ctx_factory):
    """ibx_set_config: an unknown configuration is rejected; every configuration, chosen by
    name for the one-call entry points, gives the oracle's bits."""
    _torch()
    from suguru_amd import Batch, IbxError

    wl = W.config_small()
    g, caps = wl.grid, wl.caps
    ctx = gpu_ctx_factory(g)
    with pytest.raises(IbxError):
        ctx.set_config(len(Batch.config_names()))
    ref = O.lagr2eulr(g, caps.cap_off, caps.x, caps.y, caps.z, wl.fx, wl.fy, wl.fz, caps.area)
    for name in Batch.config_names():
        print("case: one-call spread, config", name, flush=True)
        ctx.set_config(name)
        got = ctx.lagr2eulr(caps.cap_off, caps.x, caps.y, caps.z, wl.fx, wl.fy, wl.fz, caps.area)
        for a, b, nm in zip(got, ref, ("forcex", "forcey", "forcez")):
            assert np.array_equal(bits(a), bits(b)), (name, nm)


def test_two_batches_one_context_replay_after_growth(gpu_ctx_factory, moderate):
    """Batches share the context's work buffers.  A second, larger batch grows them (frees and
    reallocates); the first batch's graph, captured before, holds the freed addresses, so its
    next run captures again: every replay still gives the oracle's bits."""
    torch = _torch()
    from suguru_amd import Batch

    wl, ref_i, ref_s = moderate
    g, caps = wl.grid, wl.caps
    nsmall = 4
    off_s = caps.cap_off[:nsmall + 1].copy()
    ns = int(off_s[-1])
    sub = lambda a: np.ascontiguousarray(np.asarray(a)[:ns])
    ri_s = O.eulr2lagr(g, wl.u, wl.v, wl.w, off_s, sub(caps.x), sub(caps.y), sub(caps.z))
    rs_s = O.lagr2eulr(g, off_s, sub(caps.x), sub(caps.y), sub(caps.z), sub(wl.fx), sub(wl.fy),
                       sub(wl.fz), sub(caps.area))
    ctx = gpu_ctx_factory(g)

    def bufs(npts, src_caps, src_wl):
        t = {k: to_dev(torch, getattr(wl, k)) for k in ("u", "v", "w")}
        for k in ("fx", "fy", "fz"):
            t[k] = to_dev(torch, np.asarray(getattr(src_wl, k))[:npts])
        for k in ("x", "y", "z", "area"):
            t[k] = to_dev(torch, np.asarray(getattr(src_caps, k))[:npts])
        for k in ("ul", "vl", "wl"):
            t[k] = torch.empty(npts, dtype=torch.float64, device="cuda:0")
        for k in ("forcex", "forcey", "forcez"):
            t[k] = torch.empty(g.ncells_halo, dtype=torch.float64, device="cuda:0")
        return t

    ta, tb = bufs(ns, caps, wl), bufs(caps.npts, caps, wl)
    torch.cuda.synchronize()
    a, b = Batch(ctx, off_s), Batch(ctx, caps.cap_off)
    try:
        a.bind(**ta)
        b.bind(**tb)

        def check(t, ri, rs, what):
            for k, r in zip(("ul", "vl", "wl"), ri):
                assert np.array_equal(bits(t[k].cpu().numpy()), bits(r)), (what, k)
            for k, r in zip(("forcex", "forcey", "forcez"), rs):
                assert np.array_equal(bits(t[k].cpu().numpy()), bits(r.reshape(-1, order="F"))), (what, k)

        def run(batch, t, what):
            for k in ("ul", "vl", "wl", "forcex", "forcey", "forcez"):
                t[k].fill_(np.nan)
            torch.cuda.synchronize()
            print("case: two batches,", what, flush=True)
            batch.run()
            ctx.sync()

        a.set_config(0), b.set_config(0)
        a.capture(use_graph=True)        # small: sizes the work buffers for 4 capsules
        run(a, ta, "small first replay")
        check(ta, ri_s, rs_s, "small first")
        b.capture(use_graph=True)        # large: frees and reallocates them
        for rep in range(2):             # the small graph captures again, then replays
            run(a, ta, f"small after growth {rep}")
            check(ta, ri_s, rs_s, f"small after growth {rep}")
        run(b, tb, "large")
        check(tb, ref_i, ref_s, "large")
        # the same with the tile spread, whose candidate list also grows
        a.set_config(4), b.set_config(4)
        a.capture(use_graph=True)
        b.capture(use_graph=True)
        run(a, ta, "tile small after tile large")
        check(ta, ri_s, rs_s, "tile small")
        run(b, tb, "tile large")
        check(tb, ref_i, ref_s, "tile large")
    finally:
        a.close()
        b.close()


def test_device_offsets_checked_on_device(gpu_ctx_factory):
    """ibx_spread_f64_dev / ibx_update_pos_dev read cap_off from HBM: a table whose last entry
    is not npts, or that decreases, is caught on the device before any kernel indexes
    through it (IBX_E_ARG at the next sync); a valid call afterwards gives the oracle's bits."""
    torch = _torch()
    from suguru_amd import IbxError

    wl = W.config_moderate()
    g, caps = wl.grid, wl.caps
    ctx = gpu_ctx_factory(g)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    d = {k: to_dev(torch, getattr(wl, k)) for k in ("fx", "fy", "fz")}
    for k in ("x", "y", "z", "area"):
        d[k] = to_dev(torch, getattr(caps, k))
    F = [torch.empty(g.ncells_halo, dtype=torch.float64, device="cuda:0") for _ in range(3)]
    n = caps.npts
    long_end = caps.cap_off.copy()
    long_end[-1] = n + 4096
    decreasing = caps.cap_off.copy()
    decreasing[5] = decreasing[7] + 1
    nonzero = caps.cap_off.copy()
    nonzero[0] = -3
    try:
        for nm, bad in (("last != npts", long_end), ("decreasing", decreasing), ("first != 0", nonzero)):
            print("case: device offsets,", nm, flush=True)
            off = torch.from_numpy(bad).to("cuda:0")
            ctx.spread_dev(off, d["x"], d["y"], d["z"], d["fx"], d["fy"], d["fz"], d["area"], *F)
            with pytest.raises(IbxError) as ei:
                ctx.sync()
            assert ei.value.code == -1 and "cap_off" in str(ei.value)
            ul = torch.zeros(n, dtype=torch.float64, device="cuda:0")
            cen = torch.empty(3 * caps.ncap, dtype=torch.float64, device="cuda:0")
            um = torch.empty(3 * caps.ncap, dtype=torch.float64, device="cuda:0")
            xs, ys, zs = (d[k].clone() for k in "xyz")
            ctx.update_pos_dev(off, xs, ys, zs, ul, ul, ul, 1e-4, cen, um)
            with pytest.raises(IbxError) as ei:
                ctx.sync()
            assert ei.value.code == -1
        off = torch.from_numpy(caps.cap_off).to("cuda:0")
        ctx.spread_dev(off, d["x"], d["y"], d["z"], d["fx"], d["fy"], d["fz"], d["area"], *F)
        ctx.sync()
        ref = O.lagr2eulr(g, caps.cap_off, caps.x, caps.y, caps.z, wl.fx, wl.fy, wl.fz, caps.area)
        for a, r in zip(F, ref):
            assert np.array_equal(bits(a.cpu().numpy()), bits(r.reshape(-1, order="F")))
    finally:
        ctx.set_stream(0)
