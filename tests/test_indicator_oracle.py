"""CPU: the wall-collision and indicator-function oracle (oracle/indicator_oracle.c), pinned by
what the mathematics fixes.

The reference's collisions.f90 and indicator.f90 need MPI and the 2DECOMP modules, so they
cannot be compiled here.  Pins:

* collisions (collisions.f90:10-32): no force at or above the cutoff rad0; repulsive (+z) and
  decreasing with height below it; each lattice term is the Morse force, minus the
  derivative of D (e^{-2 beta (s - r0)} - 2 e^{-beta (s - r0)}) (checked by central
  differences of that potential, not by the force formula); the reference's truncated
  lattice [-jm, jm]^2 holds every lattice point inside the cutoff (brute force over a wide
  lattice agrees); the area factor min/max is symmetric in area <-> dx^4 / area.
* the indicator (indicator.f90:19-182): for a spherical capsule (inward normals and
  Gauss-quadrature areas of the membrane grid) it is 1 in the cells well inside, 0 in the cells
  well outside, and its integral is the sphere's volume; the viscosity and the centrifugal
  increments follow from it (exact values where it is 0 or 1); the divergence stencil is exact
  on linear fields; the solve satisfies the discrete Neumann Poisson equation; and the points
  come back as (z + s) - s.
"""
import math

import numpy as np
import pytest

import _oracle as O
from suguru_amd import params as PR
from suguru_amd import workload as W


def test_collisions_zero_at_and_above_cutoff():
    g = W.make_grid(16, 16, 16, 1.0 / 16)
    r0 = PR.rad0(g.dz)
    z = np.array([r0, r0 * 1.0000001, 2 * r0, 0.5 * g.lz, g.lz])
    area = np.full(z.size, 0.7 * g.dx * g.dx)
    fz = np.array([-0.0, 1.5, -2.25, 3.0, 0.0])
    want = fz.copy() + 0.0  # the reference adds fc = 0: -0 becomes +0
    O.collisions(g, z, area, fz, PR.collision_factor(), r0, PR.BETACOL)
    assert np.array_equal(fz.view(np.uint64), want.view(np.uint64))


def test_collisions_repulsive_and_decreasing():
    g = W.make_grid(16, 16, 16, 1.0 / 16)
    r0 = PR.rad0(g.dz)
    z = np.linspace(0.05, 0.999, 200) * r0
    area = np.full(z.size, g.dx * g.dx)
    fz = np.zeros(z.size)
    O.collisions(g, z, area, fz, PR.collision_factor(), r0, PR.BETACOL)
    assert np.all(fz > 0)
    assert np.all(np.diff(fz) < 0)


def _morse_force_fd(s, r0, beta, D):
    """-dU/ds of the Morse potential U = D (e^{-2 beta (s - r0)} - 2 e^{-beta (s - r0)}), by
    central differences (a pin independent of the force expression)."""
    U = lambda t: D * (np.exp(-2 * beta * (t - r0)) - 2 * np.exp(-beta * (t - r0)))
    h = 1e-6 * r0
    return -(U(s + h) - U(s - h)) / (2 * h)


def test_collisions_single_term_is_the_morse_force():
    """A point so close to the cutoff that the lattice holds only (0, 0): the force is the
    Morse force of depth D = col / (beta Rep^2) at s = z (col is beta De, param.f90:64), times
    |z / s| = 1 and the area factor."""
    g = W.make_grid(16, 16, 16, 1.0 / 16)
    r0 = PR.rad0(g.dz)
    beta = PR.BETACOL
    # jm = nint(sqrt(r0^2 - z^2) dyi) = 0 needs sqrt(r0^2 - z^2) < dy / 2
    z = np.array([math.sqrt(r0 * r0 - (0.45 * g.dy) ** 2), math.sqrt(r0 * r0 - (0.2 * g.dy) ** 2)])
    area = np.array([0.5, 2.0]) * g.dx * g.dx
    fz = np.zeros(2)
    O.collisions(g, z, area, fz, PR.collision_factor(), r0, beta)
    # the Morse length scale in these units: beta multiplies (rad0 - s) directly
    D = PR.COL / (PR.REP ** 2) / beta
    want = _morse_force_fd(z, r0, beta, D) * np.array([0.5, 0.5])
    assert np.allclose(fz, want, rtol=1e-7, atol=0)


def test_collisions_lattice_is_complete():
    """The reference truncates the plane sum at jm = nint(sqrt(rad0^2 - z^2) dyi); on a uniform
    grid that holds every lattice point with s < rad0.  Brute force over [-12, 12]^2."""
    g = W.make_grid(16, 16, 16, 1.0 / 16)
    r0 = PR.rad0(g.dz)
    rng = np.random.default_rng(4)
    z = rng.uniform(0.01, 1.0, 300) * r0
    area = rng.uniform(0.3, 3.0, 300) * g.dx * g.dx
    fz = np.zeros(300)
    a2 = PR.collision_factor()
    O.collisions(g, z, area, fz, a2, r0, PR.BETACOL)
    J, K = np.meshgrid(np.arange(-12, 13), np.arange(-12, 13), indexing="ij")
    ratio = np.minimum(area, g.dx ** 2) / np.maximum(area, g.dx ** 2)
    for q in range(300):
        s = np.sqrt(z[q] ** 2 + (J * g.dy) ** 2 + (K * g.dx) ** 2)
        m = s < r0
        x = PR.BETACOL * (r0 - s[m])
        want = np.sum(a2 * (np.exp(2 * x) - np.exp(x)) * np.abs(z[q] / s[m])) * ratio[q]
        assert abs(fz[q] - want) <= 1e-13 * abs(want), q


def test_collisions_area_factor_symmetric():
    g = W.make_grid(16, 16, 16, 1.0 / 16)
    r0 = PR.rad0(g.dz)
    z = np.linspace(0.1, 0.9, 9) * r0
    a = np.linspace(0.2, 0.9, 9) * g.dx ** 2
    f1, f2 = np.zeros(9), np.zeros(9)
    O.collisions(g, z, a, f1, PR.collision_factor(), r0, PR.BETACOL)
    O.collisions(g, z, g.dx ** 4 / a, f2, PR.collision_factor(), r0, PR.BETACOL)
    assert np.allclose(f1, f2, rtol=1e-14, atol=0)


def test_indicator_divergence_exact_on_linear_fields():
    g = W.make_grid(16, 8, 12, 0.125)
    I, J, K = np.meshgrid(np.arange(g.itot + 2), np.arange(g.jtot + 2), np.arange(g.ktot + 2),
                          indexing="ij")
    gx = np.asfortranarray(3.0 * I * g.dx)
    gy = np.asfortranarray(-2.0 * J * g.dy)
    gz = np.asfortranarray(0.5 * K * g.dz)
    indi = O.indicator_div(g, gx, gy, gz)
    assert np.array_equal(indi[1:-1, 1:-1, 1:-1], np.full((16, 8, 12), 3.0 - 2.0 + 0.5))


def _sphere_case(R_cells=6.0, n=48, degree=16, zc_cells=10.3):
    g = W.make_grid(n, n, n, 1.0 / n)
    c = np.array([0.43, 0.57, zc_cells * g.dz])
    R = R_cells * g.dx
    caps = W.build_capsules([c], [R], [degree])
    # inward unit normals (the membrane law's convention) and Gauss areas
    nx, ny, nz = (c[k] - p for k, p in enumerate((caps.x, caps.y, caps.z)))
    nrm = np.sqrt(nx * nx + ny * ny + nz * nz)
    return g, caps, c, R, nx / nrm, ny / nrm, nz / nrm


def test_indicator_of_a_sphere():
    g, caps, c, R, nx, ny, nz = _sphere_case()
    z = caps.z.copy()
    s = (g.ktot // 4) * g.dz
    fy = np.asfortranarray(np.zeros(g.shape_f))
    fz = np.asfortranarray(np.zeros(g.shape_f))
    lam = 3.0
    st, ct, tc = PR.centrifugal()
    gx, gy, gz, indi, indic, visc = O.indicator(g, PR.REP, 1e-4, caps.cap_off, caps.x, caps.y, z,
                                                nx, ny, nz, caps.area, fy, fz, lam, PR.VISC0,
                                                st, ct, tc)
    assert np.array_equal(z, (caps.z + s) - s)
    ii = (np.arange(g.itot) + 0.5) * g.dx
    jj = (np.arange(g.jtot) + 0.5) * g.dy
    kk = (np.arange(g.ktot) + 0.5) * g.dz
    X, Y, Z = np.meshgrid(ii, jj, kk, indexing="ij")
    d = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)
    I = indic[1:-1, 1:-1, 1:-1]
    inside, outside = d < R - 2.5 * g.dx, d > R + 2.5 * g.dx
    assert inside.sum() > 100
    assert np.all(I[inside] == 1.0)
    assert np.all(I[outside] == 0.0)
    vol = I.sum() * g.dx * g.dy * g.dz
    assert abs(vol / (4.0 / 3.0 * math.pi * R ** 3) - 1.0) < 0.03
    # viscosity: visc0 (1 + (lambda - 1) indic), exact where indic is 0 or 1
    V = visc[1:-1, 1:-1, 1:-1]
    assert np.all(V[inside] == PR.VISC0 * (1.0 + (lam - 1.0)))
    assert np.all(V[outside] == PR.VISC0)
    # centrifugal increments: sin_t tcoef and -cos_t tcoef deep inside, 0 far outside
    Fy, Fz = fy[1:-1, 1:-1, 1:-1], fz[1:-1, 1:-1, 1:-1]
    deep = d < R - 3.5 * g.dx
    assert np.all(Fy[deep] == st * tc) and np.all(Fz[deep] == -ct * tc)
    far = d > R + 3.5 * g.dx
    # j = jmax and k = kmax read the unthresholded halos (small values, not 0)
    far_y, far_z = far.copy(), far.copy()
    far_y[:, -1, :] = False
    far_z[:, :, -1] = False
    assert np.all(Fy[far_y] == 0.0) and np.all(Fz[far_z] == 0.0)
    halo = indic[1:-1, g.jtot + 1, 1:-1]
    assert np.array_equal(Fy[:, -1, :][far[:, -1, :]],
                          (0.5 * (I[:, -1, :] + halo) * st * tc)[far[:, -1, :]])
    assert np.any(Fy[:, -1, :][far[:, -1, :]] != 0.0)


def test_indicator_solve_is_the_neumann_poisson_solution():
    """indi after the solve satisfies the discrete Laplacian = the divergence of G (the
    Neumann pressure operator, periodic in x and y) to rounding, up to the null-space
    constant the solver fixes."""
    g, caps, c, R, nx, ny, nz = _sphere_case(R_cells=4.0, n=32, degree=12, zc_cells=6.7)
    z = caps.z.copy()
    fy = np.asfortranarray(np.zeros(g.shape_f))
    fz = np.asfortranarray(np.zeros(g.shape_f))
    st, ct, tc = PR.centrifugal()
    gx, gy, gz, indi, indic, visc = O.indicator(g, PR.REP, 1e-4, caps.cap_off, caps.x, caps.y, z,
                                                nx, ny, nz, caps.area, fy, fz, 1.0, PR.VISC0,
                                                st, ct, tc)
    rhs = O.indicator_div(g, gx, gy, gz)[1:-1, 1:-1, 1:-1]
    x = indi[1:-1, 1:-1, 1:-1]
    ext = np.concatenate([x[:, :, :1], x, x[:, :, -1:]], axis=2)  # Neumann ghosts
    L = ((np.roll(x, -1, 0) - 2 * x + np.roll(x, 1, 0)) * g.dxi ** 2 +
         (np.roll(x, -1, 1) - 2 * x + np.roll(x, 1, 1)) * g.dyi ** 2 +
         (ext[:, :, 2:] - 2 * ext[:, :, 1:-1] + ext[:, :, :-2]) * g.dzi ** 2)
    scale = 4 * (g.dxi ** 2 + g.dyi ** 2 + g.dzi ** 2) * np.abs(x).max()
    # the right-hand side of a closed surface's G integrates to ~0; the solver drops the mean
    assert abs(rhs.mean()) < 1e-9 * np.abs(rhs).max()
    assert np.abs(L - (rhs - rhs.mean())).max() <= 1e-10 * scale


def test_indicator_finish_halo_and_threshold_order():
    """indicator.f90:104-165: boundind(indic) runs before the thresholds, so the j = j1 and
    k = k1 halos keep unthresholded copies, and the forces at j = jmax and k = kmax read them."""
    g = W.make_grid(8, 8, 8, 0.125)
    rng = np.random.default_rng(3)
    indi = np.asfortranarray(rng.choice([0.005, 0.5, 0.95], size=g.shape_f))
    fy = np.asfortranarray(np.zeros(g.shape_f))
    fz = np.asfortranarray(np.zeros(g.shape_f))
    indic, visc = O.indicator_finish(g, indi.copy(order="F"), fy, fz, 1.0, 1.0, 1.0, 1.0, 1.0)
    kq = g.ktot // 4
    raw = indi.copy(order="F")
    O.boundp(g, raw)
    # interior thresholded, halo j1 = raw copy of j = 1 (shifted down by kq)
    src = raw[1:-1, 1, 1 + kq]
    assert np.array_equal(indic[1:-1, g.jtot + 1, 1], src)
    thr = np.where(src < 0.01, 0.0, np.where(src >= 0.9, 1.0, src))
    assert np.array_equal(indic[1:-1, 1, 1], thr)
    top = indic[1:-1, g.jtot, 1]
    assert np.allclose(fy[1:-1, g.jtot, 1], 0.5 * (top + src), rtol=0, atol=0)
