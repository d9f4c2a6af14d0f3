"""GPU: the flow update (libfluid) against the CPU oracle (oracle/fluid_oracle.c).

The stencil kernels (bounduvw, boundp, fillps, correc) evaluate the reference's expressions in
its order without fused multiply-adds: bitwise equal to the oracle.  The solves use this
library's radix-2 FFTs where the oracle uses exact DFT sums (and the reference VFFTPACK), so
solver1d, fftuvw and the whole step are compared normwise: max |GPU - oracle| <= 1e-12 of
the largest |oracle| value (fields are O(1); the FFT rounding is ~1e-15 per pass).  At sizes
the oracle cannot reach, size-independent properties: the solution satisfies the discrete
equations, and the projected velocity is divergence-free.
"""
import numpy as np
import pytest

import _oracle as O
from suguru_amd import workload as W

pytestmark = pytest.mark.gpu
REP, DT = 0.0016, 1e-4
TOL = 1e-12


def _torch():
    import torch

    assert torch.cuda.is_available(), "no HIP device visible to torch"
    return torch


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).reshape(-1, order="F"))).to("cuda:0")


def host(t, g):
    return t.cpu().numpy().reshape(g.shape_f, order="F")


def bits(a):
    return np.ascontiguousarray(a).reshape(-1).view(np.uint64)


def close(got, want, what, tol=TOL):
    err = np.abs(got - want).max()
    scale = max(np.abs(want).max(), 1e-300)
    assert err <= tol * scale, f"{what}: max |diff| {err:.3e} vs scale {scale:.3e}"


@pytest.fixture(scope="module")
def solvers():
    from suguru_amd.fluid import FlowSolver

    made = {}

    def get(shape, h=0.1):
        key = (shape, h)
        if key not in made:
            made[key] = (W.make_grid(*shape, h), None)
            made[key] = (made[key][0], FlowSolver(made[key][0], REP, DT, 0))
        return made[key]

    yield get
    for _, s in made.values():
        s.close()


# (8, 16, 300): more than 257 planes, the two-pass elimination; the others the segmented
# one with 1, 2 and 3 segments
@pytest.mark.parametrize("shape", [(16, 8, 12), (32, 16, 24), (8, 64, 9), (128, 64, 48),
                                   (8, 16, 300)])
@pytest.mark.parametrize("ne", [1, 2, 3])
def test_solver1d_matches_oracle(solvers, shape, ne):
    torch = _torch()
    g, s = solvers(shape)
    print(f"case: solver1d ne={ne} grid {shape}", flush=True)
    rng = np.random.default_rng(100 * ne + shape[0])
    p = np.asfortranarray(rng.standard_normal(g.shape_f))
    if ne == 1:
        p[1:-1, 1:-1, 1:-1] -= p[1:-1, 1:-1, 1:-1].mean()
    if ne == 3:
        p[1:-1, 1:-1, -2] = 0.0
    want = p.copy(order="F")
    O.solver1d(g, REP, DT, ne, want)
    t = dev(torch, p)
    s.solver1d(t, ne)
    s.sync()
    got = host(t, g)
    close(got[1:-1, 1:-1, 1:-1], want[1:-1, 1:-1, 1:-1], f"solver1d ne={ne}")
    # halos are not touched
    mask = np.ones(g.shape_f, bool)
    mask[1:-1, 1:-1, 1:-1] = False
    assert np.array_equal(bits(got[mask]), bits(p[mask]))


def test_stencils_bitwise(solvers):
    torch = _torch()
    g, s = solvers((32, 16, 24))
    rng = np.random.default_rng(5)
    u, v, w, p = (np.asfortranarray(rng.standard_normal(g.shape_f)) for _ in range(4))
    print("case: bounduvw / boundp / fillps / correc", flush=True)
    tu, tv, tw, tp = (dev(torch, a) for a in (u, v, w, p))
    s.bounduvw(tu, tv, tw)
    s.boundp(tp)
    O.bounduvw(g, u, v, w)
    O.boundp(g, p)
    for a, b, nm in ((tu, u, "u"), (tv, v, "v"), (tw, w, "w"), (tp, p, "p")):
        assert np.array_equal(bits(host(a, g)), bits(b)), nm
    tq = torch.zeros_like(tp)
    s.fillps(tu, tv, tw, tq)
    q = O.fillps(g, DT, u, v, w)
    assert np.array_equal(bits(host(tq, g)[1:-1, 1:-1, 1:-1]), bits(q[1:-1, 1:-1, 1:-1]))
    s.correc(tu, tv, tw, tp)
    O.correc(g, DT, u, v, w, p)
    for a, b, nm in ((tu, u, "u"), (tv, v, "v"), (tw, w, "w")):
        assert np.array_equal(bits(host(a, g)), bits(b)), "correc " + nm


def _state(g, seed):
    rng = np.random.default_rng(seed)
    un, vn, wn = W.velocity_field(g, seed, nmodes=4)
    for a in (un, vn, wn):
        a += 0.01 * rng.standard_normal(g.shape_f)
    pn = np.asfortranarray(0.1 * rng.standard_normal(g.shape_f))
    F = [np.asfortranarray(50.0 * rng.standard_normal(g.shape_f)) for _ in range(3)]
    O.bounduvw(g, un, vn, wn)
    O.boundp(g, pn)
    return [np.asfortranarray(a) for a in (un, vn, wn)], pn, F


def test_fftuvw_matches_oracle(solvers):
    torch = _torch()
    g, s = solvers((32, 16, 24), 0.25)
    (un, vn, wn), pn, F = _state(g, 3)
    uo = [a.copy(order="F") for a in (un, vn, wn)]
    uoo = [np.asfortranarray(a + 0.001) for a in (un, vn, wn)]
    print("case: fftuvw 32x16x24", flush=True)
    want = O.fftuvw(g, REP, DT, *uo, *uoo, pn, *F)
    out = [torch.zeros(g.ncells_halo, dtype=torch.float64, device="cuda:0") for _ in range(3)]
    s.fftuvw(*out, *(dev(torch, a) for a in uo), *(dev(torch, a) for a in uoo), dev(torch, pn),
             *(dev(torch, a) for a in F))
    s.sync()
    for o, w_, nm in zip(out, want, "uvw"):
        close(host(o, g)[1:-1, 1:-1, 1:-1], w_[1:-1, 1:-1, 1:-1], "fftuvw " + nm)


def test_three_steps_match_oracle_and_are_divergence_free(solvers):
    """main.f90:188-205 rotates (uoo <- uo <- unew) before each step; three steps of the whole
    flow update against the oracle's, then chkdiv on both."""
    torch = _torch()
    g, s = solvers((32, 16, 24), 0.25)
    (un, vn, wn), pn, F = _state(g, 9)
    cur = [a.copy(order="F") for a in (un, vn, wn)]
    old = [a.copy(order="F") for a in cur]
    pno = pn.copy(order="F")
    # device: three velocity sets rotated by reference, like the reference's arrays
    d_cur = [dev(torch, a) for a in cur]
    d_old = [dev(torch, a) for a in cur]
    d_older = [torch.empty_like(a) for a in d_cur]
    d_pn, d_p = dev(torch, pn), torch.zeros(g.ncells_halo, dtype=torch.float64, device="cuda:0")
    dF = [dev(torch, a) for a in F]
    for step in range(3):
        print(f"case: fluid step {step}", flush=True)
        older = old
        old = [a.copy(order="F") for a in cur]
        O.fluid_step(g, REP, DT, *cur, pno, *old, *older, *F)
        # uoo <- uo <- unew; the new unew is the freed buffer (fully rewritten by the step)
        d_older, d_old, d_cur = d_old, d_cur, d_older
        for a, b in zip(d_cur, d_old):
            a.copy_(b)  # the step updates unew in place from the current values' layout
        s.step(*d_cur, d_pn, d_p, *d_old, *d_older, *dF)
        s.sync()
        for a, b, nm in zip(d_cur, cur, "uvw"):
            close(host(a, g), b, f"step {step} {nm}")
        close(host(d_pn, g), pno, f"step {step} pnew")
    t_gpu, m_gpu = s.chkdiv(*d_cur)
    t_ora, m_ora = O.chkdiv(g, *cur)
    scale = g.dxi * max(np.abs(a).max() for a in cur)
    assert m_gpu <= 1e-12 * scale and m_ora <= 1e-12 * scale
    assert abs(t_gpu) <= 1e-9 * scale


def test_large_grid_solve_satisfies_discrete_equation(solvers):
    """256 x 128 x 64 (beyond the oracle's reach): the GPU solution of every solve satisfies
    the discrete equations to rounding, evaluated with torch on the device."""
    torch = _torch()
    g, s = solvers((256, 128, 64), 0.05)
    for ne in (1, 2, 3):
        print(f"case: large solve ne={ne}", flush=True)
        gen = torch.Generator(device="cuda:0")
        gen.manual_seed(ne)
        p = torch.randn(g.ktot + 2, g.jtot + 2, g.itot + 2, dtype=torch.float64, device="cuda:0",
                        generator=gen)
        inner = p[1:-1, 1:-1, 1:-1]
        if ne == 1:
            inner -= inner.mean()
        if ne == 3:
            inner[-1] = 0.0
        rhs = inner.clone()
        s.solver1d(p.view(-1), ne)
        s.sync()
        x = p[1:-1, 1:-1, 1:-1]
        if ne == 2:
            ext = torch.cat([-x[:1], x, -x[-1:]], 0)
        elif ne == 1:
            ext = torch.cat([x[:1], x, x[-1:]], 0)
        else:
            ext = torch.cat([torch.zeros_like(x[:1]), x], 0)
        lz = (ext[2:] - 2 * ext[1:-1] + ext[:-2]) * g.dzi ** 2
        xc = x if ne != 3 else x[:-1]
        lx = (torch.roll(xc, -1, 2) - 2 * xc + torch.roll(xc, 1, 2)) * g.dxi ** 2
        ly = (torch.roll(xc, -1, 1) - 2 * xc + torch.roll(xc, 1, 1)) * g.dyi ** 2
        L = lx + ly + lz - (0.0 if ne == 1 else REP / DT) * xc
        r = rhs if ne != 3 else rhs[:-1]
        scale = 4 * (g.dxi ** 2 + g.dyi ** 2 + g.dzi ** 2 + REP / DT) * float(x.abs().max())
        assert float((L - r).abs().max()) <= 1e-13 * scale, ne
