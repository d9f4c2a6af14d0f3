"""GPU: the capsule step without the fluid update (suguru_amd.capsule_step), stage by stage
against the oracle chain over several steps.

Each step the CPU side runs the same chain on its own copy of the positions: eulr2lagr,
the membrane law (mem_traction), lagr2eulr of the GPU's traction, Update_Pos.  The
transfer stages and the position update are bitwise (as in their own tests); the law
agrees to 1e-10 of the largest traction, as in tests/test_gpu_membrane.py, so the spread
is checked bitwise on the GPU's own traction.  The positions therefore stay bitwise in
lockstep: with the velocity field frozen they depend on the interpolation alone.
"""
import ctypes

import numpy as np
import pytest

import _oracle as O
from suguru_amd import workload as W

pytestmark = pytest.mark.gpu
D = ctypes.POINTER(ctypes.c_double)


def bits(a):
    return np.ascontiguousarray(a).reshape(-1).view(np.uint64)


def test_capsule_step_chain_matches_oracle():
    import torch

    from suguru_amd.capsule_step import CapsuleStep

    assert torch.cuda.is_available(), "no HIP device visible to torch"
    # 64^3 cells of h = 1/10.6, so that the capsules' rest radius is 1 (the law's unit
    # sphere); 8 ellipsoidal capsules of degree 16 on a jittered 2 x 2 x 2 lattice
    g = W.make_grid(64, 64, 64, h=1.0 / 10.6)
    r = 10.6 * g.dx
    centers = W.lattice_capsules(g, 1.1 * r, (2, 2, 2), seed=5)
    caps = W.build_capsules(centers, np.full(8, r), [16] * 8)
    x, y, z = caps.x.copy(), caps.y.copy(), caps.z.copy()
    for p, c in enumerate(caps.centers):  # semi-axes 1.1, 0.95, 1.0: a loaded membrane
        s = slice(caps.cap_off[p], caps.cap_off[p + 1])
        x[s] = c[0] + 1.1 * (x[s] - c[0])
        y[s] = c[1] + 0.95 * (y[s] - c[1])
    u, v, w = W.velocity_field(g, 3)
    gs, eb, dt = 6250.0, 0.3 * 6250.0, 0.05  # dt moves points by about half a cell
    print("case: capsule step, 64^3, 8 capsules deg 16, 3 steps", flush=True)
    st = CapsuleStep(g, caps.cap_off, 16, gs, eb, dt)
    dev = st.dev
    up = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).reshape(-1, order="F"))).to(dev)
    du, dv, dw = up(u), up(v), up(w)
    dx, dy, dz = up(x), up(y), up(z)
    F = [torch.full((g.ncells_halo,), np.nan, dtype=torch.float64, device=dev) for _ in range(3)]
    torch.cuda.synchronize()
    st.bind(du, dv, dw, dx, dy, dz, *F)
    o = O.lib()
    o.mem_traction.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double] + [D] * 10
    off, n = caps.cap_off, st.plan.npts
    try:
        for k in range(3):
            st.step()
            st.sync()
            # interpolation on the pre-step positions
            ul = O.eulr2lagr(g, u, v, w, off, x, y, z)
            for a, b, nm in zip((st.ul, st.vl, st.wl), ul, ("ul", "vl", "wl")):
                assert np.array_equal(bits(a.cpu().numpy()), bits(b)), (k, nm)
            # membrane law, capsule by capsule
            got = [t.cpu().numpy() for t in (st.fx, st.fy, st.fz, st.nx, st.ny, st.nz, st.area)]
            for p in range(8):
                sl = slice(p * n, (p + 1) * n)
                want = [np.zeros(n) for _ in range(7)]
                xc, yc, zc = (np.ascontiguousarray(a[sl]) for a in (x, y, z))
                o.mem_traction(16, gs, eb, xc.ctypes.data_as(D), yc.ctypes.data_as(D),
                               zc.ctypes.data_as(D), *[q.ctypes.data_as(D) for q in want])
                fmax = max(np.abs(want[c]).max() for c in range(3))
                assert fmax > 1.0  # the ellipsoids carry a real load
                for c in range(3):
                    assert np.abs(got[c][sl] - want[c]).max() < 1e-10 * fmax, (k, p, c)
                for c in range(3, 7):
                    assert np.abs(got[c][sl] - want[c]).max() < 1e-12 * max(1.0, np.abs(want[c]).max())
            # spread of the GPU's traction, bitwise
            Fo = O.lagr2eulr(g, off, x, y, z, got[0], got[1], got[2], got[6])
            for a, b, nm in zip(F, Fo, ("forcex", "forcey", "forcez")):
                assert np.array_equal(bits(a.cpu().numpy()), bits(b.reshape(-1, order="F"))), (k, nm)
            # position update (in place on the CPU copies), bitwise
            cen, um, _ = O.update_pos(g, off, dt, x, y, z, *ul)
            for a, b, nm in zip((dx, dy, dz), (x, y, z), "xyz"):
                assert np.array_equal(bits(a.cpu().numpy()), bits(b)), (k, nm)
            assert np.array_equal(bits(st.centroid.cpu().numpy()), bits(cen)), k
            assert np.array_equal(bits(st.umean.cpu().numpy()), bits(um)), k
    finally:
        st.close()


def test_capsule_step_rejects_bad_bindings():
    """Uneven capsules (one membrane plan serves one degree), short fields and a null stream
    are refused before anything reaches the libraries."""
    import torch

    from suguru_amd.capsule_step import CapsuleStep

    g = W.make_grid(64, 64, 64, h=1.0 / 10.6)
    r = 10.6 * g.dx
    caps = W.build_capsules([(3.0, 3.0, 3.0)], [r], [16])
    with pytest.raises(ValueError, match="points of degree"):
        CapsuleStep(g, np.array([0, caps.npts - 1], dtype=np.int64), 16, 6250.0, 1875.0, 1e-4)
    st = CapsuleStep(g, caps.cap_off, 16, 6250.0, 1875.0, 1e-4)
    try:
        e = lambda n: torch.zeros(n, dtype=torch.float64, device=st.dev)
        nf, n = g.ncells_halo, caps.npts
        with pytest.raises(ValueError, match="forcez"):
            st.bind(e(nf), e(nf), e(nf), e(n), e(n), e(n), e(nf), e(nf), e(nf - 1))
        with pytest.raises(ValueError, match="non-null"):
            st.bind(e(nf), e(nf), e(nf), e(n), e(n), e(n), e(nf), e(nf), e(nf), stream=0)
        with pytest.raises(RuntimeError):
            st.step()  # nothing bound
    finally:
        st.close()
