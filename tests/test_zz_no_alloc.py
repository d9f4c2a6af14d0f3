"""GPU: ibx_spread_f64_dev_noalloc, the spread entry point that allocates nothing.

Its work buffers live in one fixed device array per GPU (64 MiB), so a caller that cannot
allocate (graph capture, a tight serving loop) can still spread.  Through the C ABI it must give
exactly the bits of ibx_spread_f64_dev on the same inputs, on two calls in a row, the second
with a smaller input (the array keeps the first call's contents), and it must refuse a call
whose work buffers do not fit, before any device work.
"""
import ctypes

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
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).reshape(-1))).to("cuda:0")


def _call(lib, fn, ctx, off, d, F):
    p = lambda t: ctypes.c_void_p(int(t.data_ptr()))
    return getattr(lib, fn)(ctx.h, int(off.numel()) - 1, p(off), int(d["x"].numel()),
                            p(d["x"]), p(d["y"]), p(d["z"]), p(d["fx"]), p(d["fy"]),
                            p(d["fz"]), p(d["area"]), *[p(f) for f in F])


@pytest.mark.parametrize("make", [W.config_moderate, W.config_dense_share], ids=["moderate", "dense-share"])
def test_noalloc_spread_matches_allocating_spread(gpu_ctx_factory, make):
    torch = _torch()
    from suguru_amd import _lib

    lib = _lib.load()
    wl = make()
    g, caps = wl.grid, wl.caps
    ctx = gpu_ctx_factory(g)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    try:
        # call 1: every capsule; call 2: the first fifth of them (a smaller input)
        for ncap in (caps.ncap, max(1, caps.ncap // 5)):
            off_h = caps.cap_off[:ncap + 1].copy()
            n = int(off_h[-1])
            print(f"case: noalloc spread {wl.name}, {ncap} capsules, {n} points", flush=True)
            d = {k: to_dev(torch, np.asarray(getattr(wl, k))[:n]) for k in ("fx", "fy", "fz")}
            for k in ("x", "y", "z", "area"):
                d[k] = to_dev(torch, np.asarray(getattr(caps, k))[:n])
            off = torch.from_numpy(off_h).to("cuda:0")
            want = [torch.full((g.ncells_halo,), np.nan, dtype=torch.float64, device="cuda:0")
                    for _ in range(3)]
            got = [torch.full((g.ncells_halo,), np.nan, dtype=torch.float64, device="cuda:0")
                   for _ in range(3)]
            torch.cuda.synchronize()
            assert _call(lib, "ibx_spread_f64_dev", ctx, off, d, want) == 0
            assert _call(lib, "ibx_spread_f64_dev_noalloc", ctx, off, d, got) == 0, \
                lib.ibx_last_error(ctx.h).decode()
            ctx.sync()
            for a, b, nm in zip(got, want, ("forcex", "forcey", "forcez")):
                assert np.array_equal(bits(a.cpu().numpy()), bits(b.cpu().numpy())), (ncap, nm)
            if make is W.config_moderate:
                ref = O.lagr2eulr(g, off_h, *(np.asarray(getattr(caps, k))[:n] for k in "xyz"),
                                  *(np.asarray(getattr(wl, k))[:n] for k in ("fx", "fy", "fz")),
                                  np.asarray(caps.area)[:n])
                for a, r in zip(got, ref):
                    assert np.array_equal(bits(a.cpu().numpy()), bits(r.reshape(-1, order="F")))
    finally:
        ctx.set_stream(0)


def test_noalloc_spread_refuses_what_does_not_fit(gpu_ctx_factory):
    """24 000 one-point capsules need 72 MB of capsule neighbour masks alone: refused with
    IBX_E_ARG and a message naming the 64 MiB limit, before any device work."""
    torch = _torch()
    from suguru_amd import _lib

    lib = _lib.load()
    wl = W.config_moderate(with_fields=False)
    g = wl.grid
    ctx = gpu_ctx_factory(g)
    ncap = 24000
    rng = np.random.default_rng(5)
    d = {"x": rng.uniform(0.3, 0.7, ncap) * g.lx, "y": rng.uniform(0.3, 0.7, ncap) * g.ly,
         "z": rng.uniform(0.3, 0.7, ncap) * g.lz, "fx": rng.normal(size=ncap),
         "fy": rng.normal(size=ncap), "fz": rng.normal(size=ncap), "area": np.full(ncap, 1e-3)}
    d = {k: to_dev(torch, v) for k, v in d.items()}
    off = torch.arange(ncap + 1, dtype=torch.int64, device="cuda:0")
    F = [torch.zeros(g.ncells_halo, dtype=torch.float64, device="cuda:0") for _ in range(3)]
    torch.cuda.synchronize()
    print("case: noalloc spread over its limit", flush=True)
    rc = _call(lib, "ibx_spread_f64_dev_noalloc", ctx, off, d, F)
    assert rc == -1
    assert "64 MiB" in lib.ibx_last_error(ctx.h).decode()
    ctx.sync()
    # nothing was written
    assert all(float(f.abs().max()) == 0.0 for f in F)
