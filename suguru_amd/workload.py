"""Seeded synthetic inputs for the immersed-boundary transfer.

These generators make the fields and capsule surfaces that the parity tests and the
benchmark run on.  Everything is FP64 and reproducible from the seed alone.

Axis convention.  The reference hard-codes its wall-normal direction on the third index k
(bound.f90:42-48, interp_spread.f90:344-351): i = x spanwise (periodic), j = y streamwise
(periodic), k = z wall-normal.  A channel quoted as NX x NY x NZ with "x, z periodic and y
the walls" is therefore stored as itot = NX, jtot = NZ, ktot = NY.

Capsule surfaces follow the reference's Lagrangian layout: nlat Gauss-Legendre
colatitudes (north to south) times nlon = 2 nlat uniform longitudes, point index
l = i*nlon + j (initparticles.f90:227-233), positions
x = xc + R sin(theta) sin(phi), y = yc + R sin(theta) cos(phi), z = zc + R cos(theta)
(sph.f90:84-86) and areas R^2 w_i 2 pi / nlon, the sphere's surfmet*wgtdsin
(sph.f90:129-130, forcing.f90:91).
"""
from __future__ import annotations

import dataclasses
import functools
import math

import numpy as np

H_DEFAULT = 1.0 / 64.0


@dataclasses.dataclass(frozen=True)
class Grid:
    """One rank's grid; mirrors param.f90:19-27 with imax = itot etc."""

    itot: int
    jtot: int
    ktot: int
    lx: float
    ly: float
    lz: float
    dxi: float
    dyi: float
    dzi: float
    dx: float
    dy: float
    dz: float
    boundleft: float = 0.0
    boundfront: float = 0.0

    @property
    def shape_f(self):
        """Fortran field shape (0:i1, 0:j1, 0:k1), i fastest."""
        return (self.itot + 2, self.jtot + 2, self.ktot + 2)

    @property
    def ncells_halo(self) -> int:
        return (self.itot + 2) * (self.jtot + 2) * (self.ktot + 2)

    def empty_field(self) -> np.ndarray:
        return np.zeros(self.shape_f, dtype=np.float64, order="F")


def make_grid(itot: int, jtot: int, ktot: int, h: float = H_DEFAULT) -> Grid:
    """Grid with uniform spacing h, computed in the reference's order:
    lx = itot*h, dxi = itot/lx, dx = 1/dxi (param.f90:24-26)."""
    lx, ly, lz = itot * h, jtot * h, ktot * h
    dxi, dyi, dzi = itot / lx, jtot / ly, ktot / lz
    return Grid(itot, jtot, ktot, lx, ly, lz, dxi, dyi, dzi, 1.0 / dxi, 1.0 / dyi, 1.0 / dzi)


def grid_from_channel(nx: int, ny: int, nz: int, h: float = H_DEFAULT) -> Grid:
    """NX x NY x NZ channel with x,z periodic and y walls -> itot=NX, jtot=NZ, ktot=NY."""
    return make_grid(nx, nz, ny, h)


# ---------------------------------------------------------------------------------------
# Capsule surfaces
# ---------------------------------------------------------------------------------------


@functools.lru_cache(maxsize=None)
def sphere_layout(degree: int):
    """Unit-sphere Gauss grid for spherical-harmonic degree L: nlat = L+1, nlon = 2 nlat.

    Returns (theta[nlat] north->south, phi[nlon], gauss weights w[nlat]), read-only (cached
    per degree)."""
    nlat = degree + 1
    nlon = 2 * nlat
    xg, wg = np.polynomial.legendre.leggauss(nlat)  # ascending cos(theta)
    theta = np.arccos(xg[::-1])  # north (theta ~ 0) to south
    w = wg[::-1].copy()
    phi = np.arange(nlon, dtype=np.float64) * (2.0 * math.pi / nlon)
    for a in (theta, phi, w):
        a.setflags(write=False)
    return theta, phi, w


def capsule_points(center, radius: float, degree: int):
    """Points (x, y, z) and areas of one spherical capsule, latitude-major order."""
    theta, phi, w = sphere_layout(degree)
    nlat, nlon = theta.size, phi.size
    st = np.sin(theta)[:, None]
    ct = np.cos(theta)[:, None]
    sp = np.sin(phi)[None, :]
    cp = np.cos(phi)[None, :]
    x = center[0] + radius * st * sp
    y = center[1] + radius * st * cp
    z = center[2] + radius * ct * np.ones_like(sp)
    area = (radius * radius) * w[:, None] * (2.0 * math.pi / nlon) * np.ones_like(sp)
    return (x.reshape(-1), y.reshape(-1), z.reshape(-1), area.reshape(-1))


@dataclasses.dataclass
class Capsules:
    """Structure-of-arrays capsule batch: cap_off[p] .. cap_off[p+1]-1 are capsule p."""

    cap_off: np.ndarray  # int64 [ncap+1]
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    area: np.ndarray
    centers: np.ndarray  # [ncap, 3]
    radii: np.ndarray  # [ncap]
    degrees: np.ndarray  # [ncap]

    @property
    def ncap(self) -> int:
        return int(self.cap_off.size - 1)

    @property
    def npts(self) -> int:
        return int(self.cap_off[-1])


def build_capsules(centers, radii, degrees) -> Capsules:
    xs, ys, zs, As, off = [], [], [], [], [0]
    for c, r, d in zip(centers, radii, degrees):
        x, y, z, a = capsule_points(c, float(r), int(d))
        xs.append(x), ys.append(y), zs.append(z), As.append(a)
        off.append(off[-1] + x.size)
    cat = lambda v: np.ascontiguousarray(np.concatenate(v)) if v else np.zeros(0)
    return Capsules(
        np.asarray(off, dtype=np.int64),
        cat(xs),
        cat(ys),
        cat(zs),
        cat(As),
        np.asarray(centers, dtype=np.float64).reshape(-1, 3),
        np.asarray(radii, dtype=np.float64),
        np.asarray(degrees, dtype=np.int64),
    )


def place_capsules(grid: Grid, radii, seed: int, gap_h: float = 2.0, max_trials: int = 20000):
    """Seeded random sequential placement without overlap: centre distance >= R1+R2+gap
    under the periodic minimum image in x and y, wall clearance >= R+gap in z."""
    rng = np.random.default_rng(seed)
    gap = gap_h * grid.dx
    radii = np.asarray(radii, dtype=np.float64)
    rmax = float(radii.max()) if radii.size else 0.0
    cell = 2.0 * rmax + gap
    nbx = max(1, int(grid.lx // cell))
    nby = max(1, int(grid.ly // cell))
    nbz = max(1, int(grid.lz // cell))
    buckets: dict = {}
    centers = np.zeros((radii.size, 3))
    for p, r in enumerate(radii):
        for _ in range(max_trials):
            c = np.array(
                [
                    rng.uniform(0.0, grid.lx),
                    rng.uniform(0.0, grid.ly),
                    rng.uniform(r + gap, grid.lz - r - gap),
                ]
            )
            bx = int(c[0] / grid.lx * nbx) % nbx
            by = int(c[1] / grid.ly * nby) % nby
            bz = min(int(c[2] / grid.lz * nbz), nbz - 1)
            ok = True
            for dx_ in (-1, 0, 1):
                for dy_ in (-1, 0, 1):
                    for dz_ in (-1, 0, 1):
                        kz = bz + dz_
                        if kz < 0 or kz >= nbz:
                            continue
                        for q in buckets.get(((bx + dx_) % nbx, (by + dy_) % nby, kz), ()):
                            d = c - centers[q]
                            d[0] -= grid.lx * round(d[0] / grid.lx)
                            d[1] -= grid.ly * round(d[1] / grid.ly)
                            if float(np.dot(d, d)) < (r + radii[q] + gap) ** 2:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                centers[p] = c
                buckets.setdefault((bx, by, bz), []).append(p)
                break
        else:
            raise RuntimeError(f"could not place capsule {p} (radius {r}) without overlap")
    return centers


def degree_for_radius(r_over_h: float) -> int:
    """Spherical-harmonic degree scaled to radius for the uneven batch."""
    if r_over_h < 5.5:
        return 8
    if r_over_h < 7.0:
        return 12
    if r_over_h < 8.5:
        return 16
    return 24


# ---------------------------------------------------------------------------------------
# Eulerian velocity and Lagrangian forces
# ---------------------------------------------------------------------------------------


def velocity_field(grid: Grid, seed: int, ub: float = 1.0, amp: float = 0.05, nmodes: int = 6):
    """Laminar Poiseuille profile in the streamwise component v (reference j) plus a
    divergence-free perturbation: the discrete MAC curl of a random vector potential
    A = sum of separable modes cos(2 pi n1 x/lx + a) cos(2 pi n2 y/ly + b) sin(m pi z/lz),
    each A-component sampled on its cell edge (A_x at ((i-1/2)h, jh, kh), A_y at
    (ih, (j-1/2)h, kh), A_z at (ih, jh, (k-1/2)h)).  u = dAz/dy - dAy/dz etc. by backward
    differences, so the discrete divergence of the perturbation vanishes identically.
    Returns Fortran-ordered (u, v, w) on (0:i1, 0:j1, 0:k1)."""
    rng = np.random.default_rng(seed)
    g = grid
    ii = np.arange(g.itot + 2, dtype=np.float64)
    jj = np.arange(g.jtot + 2, dtype=np.float64)
    kk = np.arange(g.ktot + 2, dtype=np.float64)
    xc, xf = (ii - 0.5) * g.dx, ii * g.dx
    yc, yf = (jj - 0.5) * g.dy, jj * g.dy
    zc, zf = (kk - 0.5) * g.dz, kk * g.dz

    def modes(xs, ys, zs):
        out = []
        for _ in range(nmodes):
            n1 = int(rng.integers(-3, 4))
            n2 = int(rng.integers(-3, 4))
            m3 = int(rng.integers(1, 4))
            c = float(rng.normal())
            ph1, ph2 = rng.uniform(0, 2 * math.pi, size=2)
            out.append((c * np.cos(2 * math.pi * n1 * xs / g.lx + ph1),
                        np.cos(2 * math.pi * n2 * ys / g.ly + ph2),
                        np.sin(m3 * math.pi * zs / g.lz)))
        return out

    ax = modes(xc, yf, zf)
    ay = modes(xf, yc, zf)
    az = modes(xf, yf, zc)

    def bd(a, inv):  # backward difference; index 0 wraps (a halo value)
        return (a - np.roll(a, 1)) * inv

    # each component is a sum of separable terms (fx, fy, fz, sign)
    terms_u = [(X, bd(Y, g.dyi), Z, 1.0) for X, Y, Z in az] + \
              [(X, Y, bd(Z, g.dzi), -1.0) for X, Y, Z in ay]
    terms_v = [(X, Y, bd(Z, g.dzi), 1.0) for X, Y, Z in ax] + \
              [(bd(X, g.dxi), Y, Z, -1.0) for X, Y, Z in az]
    terms_w = [(bd(X, g.dxi), Y, Z, 1.0) for X, Y, Z in ay] + \
              [(X, bd(Y, g.dyi), Z, -1.0) for X, Y, Z in ax]
    u = _separable_sum(terms_u, g)
    v = _separable_sum(terms_v, g)
    w = _separable_sum(terms_w, g)
    scale = max(np.abs(u).max(), np.abs(v).max(), np.abs(w).max(), 1e-300)
    s = amp * ub / scale
    u *= s
    v *= s
    w *= s
    zeta = (kk - 0.5) / g.ktot
    v += (6.0 * ub * zeta * (1.0 - zeta))[None, None, :]
    return u, v, w


def _separable_sum(terms, g):
    """sum_t sign_t X_t(i) Y_t(j) Z_t(k) as one GEMM, returned Fortran-ordered (I, J, K)."""
    I, J, K = g.itot + 2, g.jtot + 2, g.ktot + 2
    A = np.stack([s * (Y[:, None] * X[None, :]).reshape(-1) for X, Y, Z, s in terms])  # (M, J*I)
    C = np.stack([Z for X, Y, Z, s in terms], axis=1)  # (K, M)
    out = C @ A  # (K, J*I): memory order i fastest, then j, then k
    return out.reshape(K, J, I).transpose(2, 1, 0)


def forces(npts: int, seed: int):
    """i.i.d. N(0,1) FP64 Lagrangian force components."""
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((3, npts))
    return f[0].copy(), f[1].copy(), f[2].copy()


# ---------------------------------------------------------------------------------------
# The benchmark configurations
# ---------------------------------------------------------------------------------------


@dataclasses.dataclass
class Workload:
    name: str
    grid: Grid
    caps: Capsules
    u: object
    v: object
    w: object
    fx: np.ndarray
    fy: np.ndarray
    fz: np.ndarray


def config_small(seed: int = 1234, with_fields: bool = True) -> Workload:
    """64^3 channel, one capsule of radius 8h at the channel centre, degree 12 (338 points)."""
    g = grid_from_channel(64, 64, 64)
    caps = build_capsules([(g.lx / 2, g.ly / 2, g.lz / 2)], [8 * g.dx], [12])
    u, v, w = velocity_field(g, seed) if with_fields else (None, None, None)
    fx, fy, fz = forces(caps.npts, seed + 1)
    return Workload("plain-small-64x64x64-1cap-deg12", g, caps, u, v, w, fx, fy, fz)


def config_moderate(seed: int = 1234, with_fields: bool = True) -> Workload:
    """256 x 128 x 256 channel, 64 capsules of radius 8h, degree 16 (578 points each)."""
    g = grid_from_channel(256, 128, 256)
    radii = np.full(64, 8 * g.dx)
    centers = place_capsules(g, radii, seed)
    caps = build_capsules(centers, radii, [16] * 64)
    u, v, w = velocity_field(g, seed) if with_fields else (None, None, None)
    fx, fy, fz = forces(caps.npts, seed + 1)
    return Workload("plain-256x128x256-64cap-deg16", g, caps, u, v, w, fx, fy, fz)


def config_uneven(seed: int = 1234, with_fields: bool = True, ncap: int = 2000) -> Workload:
    """512 x 256 x 512 channel, 2000 capsules, radii U(4h,10h), degree from the radius."""
    g = grid_from_channel(512, 256, 512)
    rng = np.random.default_rng(seed + 7)
    r_h = rng.uniform(4.0, 10.0, size=ncap)
    radii = r_h * g.dx
    degs = [degree_for_radius(r) for r in r_h]
    centers = place_capsules(g, radii, seed)
    caps = build_capsules(centers, radii, degs)
    u, v, w = velocity_field(g, seed) if with_fields else (None, None, None)
    fx, fy, fz = forces(caps.npts, seed + 1)
    return Workload("batched-512x256x512-2000cap-uneven", g, caps, u, v, w, fx, fy, fz)


def lattice_capsules(grid: Grid, radius: float, counts, seed: int, gap_h: float = 1.0):
    """Centres on a counts[0] x counts[1] x counts[2] lattice (periodic in x and y, wall
    clearance radius + gap in z), each jittered uniformly by up to half the spare spacing, so
    every centre distance stays >= 2 radius + gap.  Dense suspensions, where sequential random
    placement jams."""
    nx, ny, nz = counts
    gap = gap_h * grid.dx
    sx, sy = grid.lx / nx, grid.ly / ny
    z0, z1 = radius + gap, grid.lz - radius - gap
    sz = (z1 - z0) / max(nz - 1, 1)
    jit = 0.5 * (min(sx, sy, sz if nz > 1 else sx) - 2.0 * radius - gap)
    if jit < 0:
        raise ValueError("lattice too dense for the radius and gap")
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                d = rng.uniform(-jit, jit, size=3)
                out.append(((i + 0.5) * sx + d[0], (j + 0.5) * sy + d[1],
                            min(max(z0 + k * sz + d[2], z0), z1)))
    return np.asarray(out)


def config_dense_share(seed: int = 1234, with_fields: bool = True,
                       h: float = H_DEFAULT) -> Workload:
    """One GPU's share of the dense 1024 x 512 x 256 (itot, jtot, ktot) channel with 8000
    capsules at 30 % volume fraction: 128 x 512 x 256 cells, 1000 capsules of radius 10.6h,
    degree 16 (578 points each, 578 000 in all), on a jittered 5 x 20 x 10 lattice.  With
    h = 1/10.6 the capsules have unit radius, the membrane law's stress-free size."""
    g = make_grid(128, 512, 256, h)
    r = 10.6 * g.dx
    centers = lattice_capsules(g, r, (5, 20, 10), seed + 11)
    caps = build_capsules(centers, np.full(len(centers), r), [16] * len(centers))
    u, v, w = velocity_field(g, seed) if with_fields else (None, None, None)
    fx, fy, fz = forces(caps.npts, seed + 1)
    return Workload("share-128x512x256-1000cap-30pct", g, caps, u, v, w, fx, fy, fz)


def config_dense_full(seed: int = 1234, with_fields: bool = True, h: float = 1.0 / 10.6,
                      counts=(40, 20, 10), grid_cells=(1024, 512, 256)) -> Workload:
    """The dense coupled channel: 1024 x 512 x 256 (itot, jtot, ktot) cells with 8000 capsules
    of radius 10.6h and degree 16 (578 points each, 4 624 000 in all) at 30 % volume fraction,
    on a jittered 40 x 20 x 10 lattice.  With h = 1/10.6 the capsules have unit radius, the
    membrane law's stress-free size (param.f90:135 radius = 1).  counts / grid_cells scale the
    same arrangement down (one lattice cell per 25.6 x 25.6 x 24.5 h)."""
    g = make_grid(*grid_cells, h)
    r = 10.6 * g.dx
    centers = lattice_capsules(g, r, counts, seed + 11)
    caps = build_capsules(centers, np.full(len(centers), r), [16] * len(centers))
    u, v, w = velocity_field(g, seed) if with_fields else (None, None, None)
    fx, fy, fz = forces(caps.npts, seed + 1)
    name = f"dense-{grid_cells[0]}x{grid_cells[1]}x{grid_cells[2]}-{len(centers)}cap-30pct"
    return Workload(name, g, caps, u, v, w, fx, fy, fz)


def volume_fraction(grid: Grid, caps: Capsules) -> float:
    return float(np.sum(4.0 / 3.0 * math.pi * caps.radii**3) / (grid.lx * grid.ly * grid.lz))
