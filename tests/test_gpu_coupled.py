"""GPU: the coupled step (suguru_amd.coupled_step) -- interpolation, FP64 membrane law, spread,
Update_Pos and the flow update, in the reference's order (main.f90:188-410) -- against the CPU
oracle chain (oracle/ibx_oracle.c, membrane_oracle.c, fluid_oracle.c), and its hipGraph
replays against its eager steps.

Tolerances.  The transfer kernels are bitwise to the oracle and the flow update agrees to
~1e-14; the membrane law runs with fused multiply-adds and agrees to ~1e-10 of its largest
traction, which then feeds the force field.  After three steps the velocities and pressure
agree to 1e-8 of their largest values and the positions to 1e-12 of the domain; the graph
replays repeat the eager launches and are compared bit for bit, replay by replay, with eager
torch work and synchronisations between the replays.
"""
import ctypes

import numpy as np
import pytest

import _oracle as O
from suguru_amd import params as PR
from suguru_amd import workload as W

pytestmark = pytest.mark.gpu
REP, DT = 0.0016, 1e-4
GS, EB = 6250.0, 0.3 * 6250.0


def _torch():
    import torch

    assert torch.cuda.is_available(), "no HIP device visible to torch"
    return torch


def bits(a):
    return np.ascontiguousarray(a).reshape(-1).view(np.uint64)


def small_case(degree=8):
    """32 x 16 x 16 channel, h = 1/4, two unit-radius capsules of degree 8 (one of them across
    the periodic x boundary), Poiseuille + perturbation velocity."""
    g = W.make_grid(32, 16, 16, 0.25)
    centers = [(0.5, 2.1, 1.9), (4.6, 1.8, 2.2)]
    caps = W.build_capsules(centers, [1.0, 1.0], [degree, degree])
    u, v, w = W.velocity_field(g, 21, nmodes=4)
    return g, caps, (u, v, w)


def oracle_law(L, x, y, z, n):
    o = O.lib()
    D = ctypes.POINTER(ctypes.c_double)
    o.mem_traction.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double] + [D] * 10
    out = [np.zeros_like(x) for _ in range(7)]
    for c in range(x.size // n):
        s = slice(c * n, (c + 1) * n)
        xc, yc, zc = (np.ascontiguousarray(a[s]) for a in (x, y, z))
        part = [np.zeros(n) for _ in range(7)]
        o.mem_traction(L, GS, EB, xc.ctypes.data_as(D), yc.ctypes.data_as(D), zc.ctypes.data_as(D),
                       *[p.ctypes.data_as(D) for p in part])
        for k in range(7):
            out[k][s] = part[k]
    return out


def oracle_steps(g, caps, vel, L, nsteps, full=False):
    """The reference's step order on the CPU oracle; returns per step (u, v, w, pnew, x, y, z).
    full: with the wall collisions and the indicator function (main.f90:231,237)."""
    n = caps.npts // caps.ncap
    cur = [np.asfortranarray(a.copy()) for a in vel]
    O.bounduvw(g, *cur)
    zero = lambda: np.zeros(g.shape_f, order="F")
    old = [zero(), zero(), zero()]
    pn = zero()
    x, y, z = caps.x.copy(), caps.y.copy(), caps.z.copy()
    out = []
    for _ in range(nsteps):
        older, old = old, [a.copy(order="F") for a in cur]
        ul, vl, wl = O.eulr2lagr(g, old[0], old[1], old[2], caps.cap_off, x, y, z)
        fx, fy, fz, nx, ny, nz, area = oracle_law(L, x, y, z, n)
        if full:
            O.collisions(g, z, area, fz, PR.collision_factor(REP), PR.rad0(g.dz), PR.BETACOL)
        F = O.lagr2eulr(g, caps.cap_off, x, y, z, fx, fy, fz, area)
        if full:
            st, ct, tc = PR.centrifugal()
            O.indicator(g, REP, DT, caps.cap_off, x, y, z, nx, ny, nz, area, F[1], F[2],
                        PR.LAMBDA, PR.VISC0, st, ct, tc)
        O.update_pos(g, caps.cap_off, DT, x, y, z, ul, vl, wl)
        O.fluid_step(g, REP, DT, *cur, pn, *old, *older, *F)
        out.append(([a.copy(order="F") for a in cur], pn.copy(order="F"), x.copy(), y.copy(), z.copy()))
    return out


def state(st, g):
    u, v, w = (a.cpu().numpy().reshape(g.shape_f, order="F") for a in st.current())
    return (u, v, w), st.pnew.cpu().numpy().reshape(g.shape_f, order="F"), st.x.cpu().numpy(), \
        st.y.cpu().numpy(), st.z.cpu().numpy()


def test_coupled_steps_match_oracle_chain():
    _torch()
    from suguru_amd.coupled_step import CoupledStep

    g, caps, vel = small_case()
    print("case: coupled 32x16x16, 2 capsules deg 8, 3 steps vs oracle", flush=True)
    want = oracle_steps(g, caps, vel, 8, 3)
    st = CoupledStep(g, caps.cap_off, degree=8, gs=GS, eb=EB, rep=REP, dt=DT)
    try:
        st.load(*vel, caps.x, caps.y, caps.z)
        for k in range(3):
            st.step()
            st.sync()
            (u, v, w), pn, x, y, z = state(st, g)
            (wu, wv, ww), wpn, wx, wy, wz = want[k]
            for a, b, nm in ((u, wu, "u"), (v, wv, "v"), (w, ww, "w"), (pn, wpn, "pnew")):
                err = np.abs(a - b).max()
                assert err <= 1e-8 * np.abs(b).max(), (k, nm, err)
            for a, b, nm in ((x, wx, "x"), (y, wy, "y"), (z, wz, "z")):
                assert np.abs(a - b).max() <= 1e-12 * g.lx, (k, nm)
    finally:
        st.close()


def test_graph_replays_match_eager_steps():
    torch = _torch()
    from suguru_amd.coupled_step import CoupledStep

    g, caps, vel = small_case()
    st = CoupledStep(g, caps.cap_off, degree=8, gs=GS, eb=EB, rep=REP, dt=DT)
    try:
        print("case: coupled eager reference run, 5 steps", flush=True)
        st.load(*vel, caps.x, caps.y, caps.z)
        eager = []
        for _ in range(5):
            st.step()
            st.sync()
            eager.append(state(st, g))
        print("case: coupled graph capture and 5 replays", flush=True)
        st.load(*vel, caps.x, caps.y, caps.z)
        st.capture()
        for k in range(5):
            st.run()
            st.sync()
            # eager torch work between the replays
            junk = torch.randn(4096, device="cuda:0", dtype=torch.float64).sum()
            torch.cuda.synchronize()
            assert torch.isfinite(junk)
            got = state(st, g)
            for a, b, nm in zip(got[0] + (got[1],) + got[2:], eager[k][0] + (eager[k][1],) + eager[k][2:],
                                ("u", "v", "w", "pnew", "x", "y", "z")):
                assert np.array_equal(bits(a), bits(b)), (k, nm)
    finally:
        st.close()


def test_full_physics_steps_match_oracle_chain():
    """physics="full": the wall collisions and the indicator function in the reference's
    places (collisions after the traction, the indicator after the spread), one capsule close
    enough to the bottom wall for the Morse force (rad0 = 2.5 dz), three steps against the
    oracle chain.  Also every graph replay against the eager steps."""
    torch = _torch()
    from suguru_amd.coupled_step import CoupledStep

    g = W.make_grid(32, 16, 16, 0.25)
    caps = W.build_capsules([(0.5, 2.1, 1.45), (4.6, 1.8, 2.2)], [1.0, 1.0], [8, 8])
    vel = W.velocity_field(g, 23, nmodes=4)
    assert caps.z.min() < PR.rad0(g.dz)
    print("case: coupled full physics 32x16x16, 2 capsules deg 8, 3 steps vs oracle", flush=True)
    want = oracle_steps(g, caps, vel, 8, 3, full=True)
    st = CoupledStep(g, caps.cap_off, degree=8, gs=GS, eb=EB, rep=REP, dt=DT, physics="full")
    try:
        st.load(*vel, caps.x, caps.y, caps.z)
        eager = []
        for k in range(3):
            st.step()
            st.sync()
            got = state(st, g)
            eager.append(got)
            (u, v, w), pn, x, y, z = got
            (wu, wv, ww), wpn, wx, wy, wz = want[k]
            for a, b, nm in ((u, wu, "u"), (v, wv, "v"), (w, ww, "w"), (pn, wpn, "pnew")):
                err = np.abs(a - b).max()
                assert err <= 1e-8 * np.abs(b).max(), (k, nm, err)
            for a, b, nm in ((x, wx, "x"), (y, wy, "y"), (z, wz, "z")):
                assert np.abs(a - b).max() <= 1e-12 * g.lx, (k, nm)
        print("case: coupled full physics, graph capture and 3 replays", flush=True)
        st.load(*vel, caps.x, caps.y, caps.z)
        st.capture()
        for k in range(3):
            st.run()
            st.sync()
            junk = torch.randn(4096, device="cuda:0", dtype=torch.float64).sum()
            torch.cuda.synchronize()
            assert torch.isfinite(junk)
            got = state(st, g)
            for a, b, nm in zip(got[0] + (got[1],) + got[2:], eager[k][0] + (eager[k][1],) + eager[k][2:],
                                ("u", "v", "w", "pnew", "x", "y", "z")):
                assert np.array_equal(bits(a), bits(b)), (k, nm)
    finally:
        st.close()


def test_restart_continues_bit_for_bit(tmp_path):
    """save_restart after two steps (fldbg / allpartdata, loadd.f90), a new CoupledStep
    load_restart, two more steps: the same bits as four uninterrupted steps.  write_output's
    text files hold the points (7 digits) and one center_velocity line per capsule."""
    _torch()
    from suguru_amd import formats as F
    from suguru_amd.coupled_step import CoupledStep

    g, caps, vel = small_case()
    print("case: coupled restart, 4 steps vs 2 + restart + 2", flush=True)
    st = CoupledStep(g, caps.cap_off, degree=8, gs=GS, eb=EB, rep=REP, dt=DT)
    try:
        st.load(*vel, caps.x, caps.y, caps.z)
        for _ in range(4):
            st.step()
        st.sync()
        want = state(st, g)
        st.load(*vel, caps.x, caps.y, caps.z)
        for _ in range(2):
            st.step()
        st.save_restart(str(tmp_path), 2, 2 * DT)
        st.write_output(str(tmp_path), 2, 2 * DT)
        x2 = st.x.cpu().numpy()
    finally:
        st.close()
    n = caps.npts // caps.ncap
    for p in range(caps.ncap):
        rx, ry, rz = F.read_capsule_txt(str(tmp_path / F.capsule_txt_name(p + 1, 2)))
        assert np.allclose(rx, x2[p * n:(p + 1) * n], rtol=1e-6, atol=0)
    assert len((tmp_path / "center_velocity.txt").read_text().splitlines()) == caps.ncap
    st2 = CoupledStep(g, caps.cap_off, degree=8, gs=GS, eb=EB, rep=REP, dt=DT)
    try:
        time, nr, dt = st2.load_restart(str(tmp_path), 2)
        assert (time, nr, dt) == (2 * DT, 2, DT)
        for _ in range(2):
            st2.step()
        st2.sync()
        got = state(st2, g)
        for a, b, nm in zip(got[0] + (got[1],) + got[2:], want[0] + (want[1],) + want[2:],
                            ("u", "v", "w", "pnew", "x", "y", "z")):
            assert np.array_equal(bits(a), bits(b)), nm
    finally:
        st2.close()
