"""CPU: the FP64 membrane-law oracle (oracle/membrane_oracle.c) pinned without the reference.

The reference's law (forcing.f90:11-114, sph.f90:1523-1755) needs GSL, which is absent, so it
cannot be compiled here; its neo-Hookean coefficients (membmodel.c) can, and are compared
directly.  Everything else is pinned by what the mathematics fixes:

* spherical-harmonic machinery: Gauss-Legendre nodes/weights against numpy and the colatitudes
  against SPHEREPACK's own test log (gaqd at nlat = 15 and 33, as its tdiv and tgrad tests
  print them), Legendre
  normalisation against scipy (without the Condon-Shortley phase), theta-derivatives against
  finite differences, and exact analysis/synthesis round trips;
* the law: zero traction on the reference sphere; Laplace-law pressure 2 tau / lambda with the
  neo-Hookean tension tau = Gs (1 - lambda^-6) on a dilated sphere, no tangential part, and no
  bending contribution (uniform moment); zero net force and torque on deformed shapes
  (equilibrium of a closed membrane); virtual work: the traction equals minus the gradient of
  the neo-Hookean energy; rotation covariance on the grid's own symmetry (one longitude step);
  translation invariance; linearity in the moduli;
* the bending branch: the linear response of the unit sphere to r = 1 + eps Y_lm is radial,
  mode-diagonal, independent of m and equal to a (l-1) l (l+1) (l+2) eps Y_lm with one a > 0
  for l = 2..8 (the isotropic fourth-order operator with the dilation and translation null
  modes).
"""
import ctypes
import os

import numpy as np
import pytest
from scipy.special import lpmv, sph_harm_y
from math import factorial

import _oracle as O

D = ctypes.POINTER(ctypes.c_double)
L = 16
NT, NP = L + 1, 2 * (L + 1)


def ptr(a):
    return a.ctypes.data_as(D)


@pytest.fixture(scope="module")
def lib():
    o = O.lib()
    o.mem_gauss.argtypes = [ctypes.c_int, D, D]
    o.mem_legendre.argtypes = [ctypes.c_int, ctypes.c_double, D]
    o.mem_legendre_derivs.argtypes = [ctypes.c_int, ctypes.c_double, D]
    o.mem_grid.argtypes = [ctypes.c_int, D, D]
    o.mem_analysis.argtypes = [ctypes.c_int, D, D, D]
    o.mem_synthesis.argtypes = [ctypes.c_int, D, D, D]
    o.mem_traction.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double] + [D] * 10
    o.mem_shear_energy.argtypes = [ctypes.c_int, ctypes.c_double, D, D, D]
    o.mem_shear_energy.restype = ctypes.c_double
    o.mem_nh_coeffs.argtypes = [ctypes.c_double, ctypes.c_double, D, D, D, D, D]
    return o


@pytest.fixture(scope="module")
def grid(lib):
    th, w = np.zeros(NT), np.zeros(NT)
    assert lib.mem_grid(L, ptr(th), ptr(w)) == 0
    ph = 2 * np.pi * np.arange(NP) / NP
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    return th, w, TH, PH


def shape(TH, PH, r):
    """Points of the surface r(theta, phi) in the reference's parametrisation (sph.f90:84-86)."""
    return [np.ascontiguousarray(v.ravel()) for v in
            (r * np.sin(TH) * np.sin(PH), r * np.sin(TH) * np.cos(PH), r * np.cos(TH))]


def law(lib, x, y, z, gs=1.0, eb=0.3):
    out = [np.zeros(x.size) for _ in range(7)]
    assert lib.mem_traction(L, gs, eb, ptr(x), ptr(y), ptr(z), *[ptr(o) for o in out]) == 0
    return out


def random_radius(TH, PH, seed, amp, nmax=4):
    rng = np.random.default_rng(seed)
    r = np.ones_like(TH)
    for n in range(1, nmax + 1):
        for m in range(0, n + 1):
            r = r + rng.normal() * amp / n * np.real(sph_harm_y(n, m, TH, PH))
    return r


# ---- spherical-harmonic machinery ---------------------------------------------------------

def test_gauss_nodes_match_numpy(lib):
    for n in (5, 17, 34, 48):
        th, w = np.zeros(n), np.zeros(n)
        lib.mem_gauss(n, ptr(th), ptr(w))
        x, wx = np.polynomial.legendre.leggauss(n)  # increasing x = decreasing theta
        assert np.allclose(np.cos(th), x[::-1], rtol=0, atol=1e-14)
        assert np.allclose(w, wx[::-1], rtol=2e-12, atol=0)
        assert np.all(np.diff(th) > 0)
        # exact for every polynomial of degree <= 2n-1 (numpy's own weights reach ~6e-15)
        xc = np.cos(th)
        for k in range(0, 2 * n):
            exact = 2.0 / (k + 1) if k % 2 == 0 else 0.0
            assert abs((w * xc ** k).sum() - exact) < 1e-15, (n, k)


# SPHEREPACK's own test log: the Gauss colatitudes gaqd returns, as tdiv (nlat = 15) and
# tgrad (nlat = 33) print them in E11.4 (spherepack3.2_impi-15/output/darwin.dp:250-252 and
# :448-453)
SPHEREPACK_THTG = {
    15: "0.1551E+00 0.3561E+00 0.5582E+00 0.7606E+00 0.9631E+00 0.1166E+01 0.1368E+01 "
        "0.1571E+01 0.1773E+01 0.1976E+01 0.2178E+01 0.2381E+01 0.2583E+01 0.2786E+01 "
        "0.2986E+01",
    33: "0.7178E-01 0.1648E+00 0.2583E+00 0.3520E+00 0.4457E+00 0.5394E+00 0.6332E+00 "
        "0.7269E+00 0.8207E+00 0.9144E+00 0.1008E+01 0.1102E+01 0.1196E+01 0.1289E+01 "
        "0.1383E+01 0.1477E+01 0.1571E+01 0.1665E+01 0.1758E+01 0.1852E+01 0.1946E+01 "
        "0.2040E+01 0.2133E+01 0.2227E+01 0.2321E+01 0.2415E+01 0.2508E+01 0.2602E+01 "
        "0.2696E+01 0.2790E+01 0.2883E+01 0.2977E+01 0.3070E+01",
}


@pytest.mark.parametrize("n", sorted(SPHEREPACK_THTG))
def test_gauss_colatitudes_match_spherepack_log(lib, n):
    from suguru_amd.formats import fortran_e

    th, w = np.zeros(n), np.zeros(n)
    lib.mem_gauss(n, ptr(th), ptr(w))
    got = [fortran_e(t, 11, 4).strip() for t in th]
    assert got == SPHEREPACK_THTG[n].split()


def test_legendre_normalisation_matches_scipy(lib):
    nmax = 12
    for theta in (0.3, 1.1, 2.7):
        P = np.zeros(nmax * nmax)
        lib.mem_legendre(nmax, theta, ptr(P))
        for m in range(nmax):
            for n in range(m, nmax):
                norm = np.sqrt((2 * n + 1) / 2 * factorial(n - m) / factorial(n + m))
                want = norm * (-1) ** m * lpmv(m, n, np.cos(theta))  # remove the CS phase
                assert P[m * nmax + n] == pytest.approx(want, rel=1e-12, abs=1e-14)


def test_legendre_theta_derivatives_match_finite_differences(lib):
    nmax = 10
    h = 1e-3

    def P_at(t):
        P = np.zeros(nmax * nmax)
        lib.mem_legendre(nmax, t, ptr(P))
        return P

    for theta in (0.4, 1.3, 2.2):
        Dk = np.zeros(5 * nmax * nmax)
        lib.mem_legendre_derivs(nmax, theta, ptr(Dk))
        Dk = Dk.reshape(5, nmax * nmax)
        s = [P_at(theta + k * h) for k in range(-3, 4)]  # s[3] is the centre
        fd = {
            0: s[3],
            1: (-s[5] + 8 * s[4] - 8 * s[2] + s[1]) / (12 * h),
            2: (-s[5] + 16 * s[4] - 30 * s[3] + 16 * s[2] - s[1]) / (12 * h * h),
            3: (-s[6] + 8 * s[5] - 13 * s[4] + 13 * s[2] - 8 * s[1] + s[0]) / (8 * h ** 3),
            4: (-s[6] + 12 * s[5] - 39 * s[4] + 56 * s[3] - 39 * s[2] + 12 * s[1] - s[0])
            / (6 * h ** 4),
        }
        tol = {0: 1e-14, 1: 1e-9, 2: 1e-8, 3: 1e-5, 4: 1e-4}
        for k in range(5):
            scale = max(1.0, np.abs(Dk[k]).max())
            assert np.abs(Dk[k] - fd[k]).max() / scale < tol[k], k


def test_analysis_synthesis_round_trip(lib, grid):
    rng = np.random.default_rng(3)
    a = np.zeros(NT * NT)
    b = np.zeros(NT * NT)
    for m in range(NT):
        for n in range(m, NT):
            a[m * NT + n] = rng.normal()
            if m > 0:
                b[m * NT + n] = rng.normal()
    f = np.zeros(NT * NP)
    lib.mem_synthesis(L, ptr(a), ptr(b), ptr(f))
    a2, b2 = np.zeros_like(a), np.zeros_like(b)
    lib.mem_analysis(L, ptr(f), ptr(a2), ptr(b2))
    assert np.abs(a2 - a).max() < 1e-12 and np.abs(b2 - b).max() < 1e-12
    # the unit sphere's z = cos(theta) is Pbar_1^0 sqrt(2/3) with the m = 0 term halved
    th, w, TH, PH = grid
    lib.mem_analysis(L, ptr(np.ascontiguousarray(np.cos(TH).ravel())), ptr(a2), ptr(b2))
    want = np.zeros_like(a2)
    want[0 * NT + 1] = 2 * np.sqrt(2.0 / 3.0)
    assert np.abs(a2 - want).max() < 1e-14 and np.abs(b2).max() < 1e-14


def test_neo_hookean_coefficients_match_reference_membmodel(lib):
    """membmodel.c:60-79 compiled from the reference tree (oracle/_ref) where present."""
    ref = O.ref_lib()
    if ref is None or not hasattr(ref, "calpwspi_deri_"):
        pytest.skip("reference membmodel.c not built here (oracle/_ref)")
    gs = 6250.0
    ref.ref_set_hf_gscell.argtypes = [ctypes.c_double]
    ref.ref_set_hf_gscell(0.5 * gs)  # capsule_init_cal_coeff, sph.f90:35
    rng = np.random.default_rng(1)
    for _ in range(50):
        I1, I2 = rng.uniform(-0.5, 2.0), rng.uniform(-0.6, 3.0)
        I1d, I2d = rng.normal(size=2), rng.normal(size=2)
        r = [np.zeros(1), np.zeros(1), np.zeros(2), np.zeros(2)]
        args = [np.array([gs]), np.array([50.0]), np.array([I1]), np.array([I2]), I1d, I2d]
        ref.calpwspi_deri_(ptr(r[0]), ptr(r[1]), ptr(r[2]), ptr(r[3]), *[ptr(a) for a in args])
        m = [np.zeros(1), np.zeros(1), np.zeros(2), np.zeros(2)]
        lib.mem_nh_coeffs(gs, I2, ptr(np.ascontiguousarray(I2d)), ptr(m[0]), ptr(m[1]), ptr(m[2]),
                          ptr(m[3]))
        for u, v in zip(m, r):
            assert np.allclose(u, v, rtol=4e-16, atol=0)


# ---- the law ------------------------------------------------------------------------------

@pytest.mark.parametrize("eb", [0.0, 0.3])
def test_reference_sphere_has_zero_traction(lib, grid, eb):
    th, w, TH, PH = grid
    x, y, z = shape(TH, PH, 1.0)
    fx, fy, fz, nx, ny, nz, area = law(lib, x, y, z, eb=eb)
    assert max(np.abs(fx).max(), np.abs(fy).max(), np.abs(fz).max()) < 1e-13
    assert area.sum() == pytest.approx(4 * np.pi, rel=1e-14)
    # n = a_theta x a_phi / |.|: with x = sin t sin p, y = sin t cos p this is the inward normal
    assert np.allclose(nx * x + ny * y + nz * z, -1.0, atol=1e-14)


@pytest.mark.parametrize("lam", [0.9, 1.1, 1.3])
def test_dilated_sphere_laplace_law(lib, grid, lam):
    th, w, TH, PH = grid
    x, y, z = shape(TH, PH, lam)
    gs = 2.5
    for eb in (0.0, 0.7):
        fx, fy, fz, nx, ny, nz, area = law(lib, x, y, z, gs=gs, eb=eb)
        fn = fx * nx + fy * ny + fz * nz
        tau = gs * (1.0 - lam ** -6)  # neo-Hookean isotropic tension
        assert np.allclose(fn, 2 * tau / lam, rtol=1e-10, atol=1e-12 * gs)
        tang = np.stack([fx, fy, fz]) - fn * np.stack([nx, ny, nz])
        assert np.abs(tang).max() < 1e-10 * gs
        assert area.sum() == pytest.approx(4 * np.pi * lam ** 2, rel=1e-13)


@pytest.mark.parametrize("seed", [1, 2])
def test_deformed_shape_net_force_and_torque_vanish(lib, grid, seed):
    th, w, TH, PH = grid
    x, y, z = shape(TH, PH, random_radius(TH, PH, seed, 0.08))
    for eb in (0.0, 0.3):
        fx, fy, fz, nx, ny, nz, A = law(lib, x, y, z, eb=eb)
        scale = (np.sqrt(fx ** 2 + fy ** 2 + fz ** 2) * A).sum()
        F = [(fx * A).sum(), (fy * A).sum(), (fz * A).sum()]
        T = [((y * fz - z * fy) * A).sum(), ((z * fx - x * fz) * A).sum(),
             ((x * fy - y * fx) * A).sum()]
        assert np.abs(F).max() < 1e-10 * scale and np.abs(T).max() < 1e-10 * scale


def test_virtual_work_traction_is_energy_gradient(lib, grid):
    """sum_l f_l . dx_l area_l = -dE/deps for the neo-Hookean energy (no bending)."""
    th, w, TH, PH = grid
    x, y, z = shape(TH, PH, random_radius(TH, PH, 7, 0.1))
    fx, fy, fz, *_, A = law(lib, x, y, z, eb=0.0)
    rng = np.random.default_rng(11)
    for _ in range(3):
        dr = np.zeros_like(TH)
        for n in range(0, 6):
            for m in range(0, n + 1):
                dr = dr + rng.normal() * np.real(sph_harm_y(n, m, TH, PH))
        d = [np.ascontiguousarray((dr * v).ravel()) for v in
             (np.sin(TH) * np.sin(PH), np.sin(TH) * np.cos(PH), np.cos(TH))]
        eps = 1e-6
        Ep = lib.mem_shear_energy(L, 1.0, ptr(x + eps * d[0]), ptr(y + eps * d[1]),
                                  ptr(z + eps * d[2]))
        Em = lib.mem_shear_energy(L, 1.0, ptr(x - eps * d[0]), ptr(y - eps * d[1]),
                                  ptr(z - eps * d[2]))
        dE = (Ep - Em) / (2 * eps)
        work = ((fx * d[0] + fy * d[1] + fz * d[2]) * A).sum()
        assert abs(dE + work) < 1e-8 * abs(dE)


def test_rotation_covariance_and_translation_invariance(lib, grid):
    th, w, TH, PH = grid
    x, y, z = shape(TH, PH, random_radius(TH, PH, 4, 0.1))
    fx, fy, fz, *_ = law(lib, x, y, z)

    def roll(a, k):
        return np.ascontiguousarray(np.roll(a.reshape(NT, NP), -k, axis=1).ravel())

    for k in (1, 5):
        gx, gy, gz, *_ = law(lib, roll(x, k), roll(y, k), roll(z, k))
        err = max(np.abs(gx - roll(fx, k)).max(), np.abs(gy - roll(fy, k)).max(),
                  np.abs(gz - roll(fz, k)).max())
        assert err < 1e-10 * np.abs(fx).max()
    gx, gy, gz, *_ = law(lib, x + 0.37, y - 1.2, z + 3.0)
    assert max(np.abs(gx - fx).max(), np.abs(gy - fy).max(), np.abs(gz - fz).max()) < 1e-9


def test_linear_in_moduli(lib, grid):
    th, w, TH, PH = grid
    x, y, z = shape(TH, PH, random_radius(TH, PH, 9, 0.1))
    f1 = law(lib, x, y, z, gs=1.0, eb=0.0)[:3]
    f2 = law(lib, x, y, z, gs=0.0, eb=1.0)[:3]
    f3 = law(lib, x, y, z, gs=2.0, eb=3.0)[:3]
    for a, b, c in zip(f1, f2, f3):
        assert np.allclose(c, 2 * a + 3 * b, rtol=1e-12, atol=1e-12 * np.abs(c).max())


def _real_ylm(l, m, TH, PH):
    Y = sph_harm_y(l, abs(m), TH, PH)
    return (np.sqrt(2) * (Y.imag if m < 0 else Y.real)) if m != 0 else Y.real


def test_bending_linear_response_of_the_sphere(lib, grid):
    """The bending branch (eb != 0, gs = 0) about the unit sphere, pinned by symmetry: for the
    radial perturbation r = 1 + eps Y_lm the radial part of the load is -kappa_l eps Y_lm
    (mode-diagonal; kappa independent of m by isotropy), and for a fourth-order isotropic
    surface operator kappa is quadratic in lambda = l (l + 1) with roots at l = 0 (a dilated
    sphere carries no bending load: the Laplace test) and l = 1 (translation), so
    kappa_l = a (l - 1) l (l + 1) (l + 2) with a > 0 (the load restores the sphere).  A
    sign flip of the moment, a transposed index or a dropped curvature-derivative term breaks
    one of these."""
    th, w, TH, PH = grid
    wq = (w[:, None] * np.full((1, NP), 2 * np.pi / NP)).ravel()
    X = np.stack([(np.sin(TH) * np.sin(PH)).ravel(), (np.sin(TH) * np.cos(PH)).ravel(),
                  np.cos(TH).ravel()])
    eps, eb = 1e-5, 1.0
    a_l = []
    for l in range(2, 9):
        kap = []
        for m in range(-l, l + 1):
            Y = _real_ylm(l, m, TH, PH)
            fp = np.stack(law(lib, *shape(TH, PH, 1.0 + eps * Y), gs=0.0, eb=eb)[:3])
            fm = np.stack(law(lib, *shape(TH, PH, 1.0 - eps * Y), gs=0.0, eb=eb)[:3])
            q = ((fp - fm) / (2 * eps) * X).sum(0)           # radial part of d f / d eps
            y = Y.ravel()
            k = -(q * y * wq).sum() / (y * y * wq).sum()
            assert np.abs(q + k * y).max() <= 1e-6 * np.abs(k * y).max(), (l, m)
            kap.append(k)
        kap = np.array(kap)
        assert np.abs(kap - kap.mean()).max() <= 1e-7 * abs(kap.mean()), (l, kap)
        a_l.append(kap.mean() / ((l - 1) * l * (l + 1) * (l + 2)))
    a_l = np.array(a_l)
    assert np.all(a_l > 0)
    assert np.abs(a_l - a_l.mean()).max() <= 1e-6 * a_l.mean(), a_l
