"""ctypes access to the CPU oracle (oracle/liboracle.so) and, where it was built, to the
reference's own compiled kernel (oracle/_ref/libref.so).  Test infrastructure only."""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libref.so")

import sys

sys.path.insert(0, ROOT)
from suguru_amd._lib import IbxGrid, grid_struct  # noqa: E402  (plain struct, no GPU)

P = ctypes.c_void_p
_o = None


def lib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE):
            import subprocess

            subprocess.run(["make", "-s", "liboracle.so"], cwd=os.path.join(ROOT, "oracle"),
                           check=True)
        o = ctypes.CDLL(ORACLE)
        o.ora_kernel.restype = ctypes.c_double
        o.ora_kernel.argtypes = [ctypes.c_double]
        o.ora_kerneltest.restype = ctypes.c_double
        o.ora_kerneltest.argtypes = []
        G = ctypes.POINTER(IbxGrid)
        o.ora_updthalos.argtypes = [G, P, ctypes.c_int]
        o.ora_eulr2lagr.argtypes = [G, P, P, P, ctypes.c_int64, P, P, P, P, P, P, P]
        o.ora_lagr2eulr.argtypes = [G, ctypes.c_int64, P, P, P, P, P, P, P, P, P, P, P]
        o.ora_lagr2eulrind.argtypes = [G, ctypes.c_int64, P, P, P, P, P, P, P, P, P, P, P]
        o.ora_update_pos.restype = ctypes.c_int64
        o.ora_update_pos.argtypes = [G, ctypes.c_int64, P, ctypes.c_double, P, P, P, P, P, P, P,
                                     P, P, P]
        _o = o
    return _o


def ref_lib():
    """The reference's compiled kernel.f90 + membmodel.c, or None where not built."""
    if not os.path.exists(REF):
        return None
    r = ctypes.CDLL(REF)
    r.ref_kernel.restype = ctypes.c_double
    r.ref_kernel.argtypes = [ctypes.c_double]
    r.ref_kerneltest.restype = ctypes.c_double
    return r


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


def kernel(r):
    return lib().ora_kernel(float(r))


def kerneltest():
    return lib().ora_kerneltest()


def eulr2lagr(grid, u, v, w, off, x, y, z):
    gs = grid_struct(grid)
    off = np.ascontiguousarray(off, dtype=np.int64)
    n = int(off[-1])
    ul, vl, wl = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().ora_eulr2lagr(ctypes.byref(gs), _p(u), _p(v), _p(w), off.size - 1, _p(off),
                        _p(np.ascontiguousarray(x)), _p(np.ascontiguousarray(y)),
                        _p(np.ascontiguousarray(z)), _p(ul), _p(vl), _p(wl))
    return ul, vl, wl


def _scatter(fn, grid, off, x, y, z, f1, f2, f3, area):
    gs = grid_struct(grid)
    off = np.ascontiguousarray(off, dtype=np.int64)
    o = [grid.empty_field() for _ in range(3)]
    c = np.ascontiguousarray
    fn(ctypes.byref(gs), off.size - 1, _p(off), _p(c(x)), _p(c(y)), _p(c(z)), _p(c(f1)),
       _p(c(f2)), _p(c(f3)), _p(c(area)), _p(o[0]), _p(o[1]), _p(o[2]))
    return tuple(o)


def lagr2eulr(grid, off, x, y, z, fx, fy, fz, area):
    return _scatter(lib().ora_lagr2eulr, grid, off, x, y, z, fx, fy, fz, area)


def lagr2eulrind(grid, off, x, y, z, xn, yn, zn, area):
    return _scatter(lib().ora_lagr2eulrind, grid, off, x, y, z, xn, yn, zn, area)


def updthalos(grid, var, d):
    gs = grid_struct(grid)
    lib().ora_updthalos(ctypes.byref(gs), _p(var), int(d))


def update_pos(grid, off, dt, x, y, z, ul, vl, wl, xold=None, yold=None):
    """In place; returns (centroids, umean, nclamp)."""
    gs = grid_struct(grid)
    off = np.ascontiguousarray(off, dtype=np.int64)
    ncap = off.size - 1
    cen = np.zeros((ncap, 3))
    um = np.zeros((ncap, 3))
    nc = lib().ora_update_pos(ctypes.byref(gs), ncap, _p(off), float(dt), _p(x), _p(y), _p(z),
                              _p(xold), _p(yold), _p(ul), _p(vl), _p(wl), _p(cen), _p(um))
    return cen, um, int(nc)


# ---- the flow update (oracle/fluid_oracle.c) ---------------------------------------------
def _flib():
    o = lib()
    if not getattr(o, "_fluid_sig", False):
        G = ctypes.POINTER(IbxGrid)
        D, I = ctypes.c_double, ctypes.c_int
        o.ora_rfftf.argtypes = [I, P]
        o.ora_rfftb.argtypes = [I, P]
        o.ora_initsolver.argtypes = [G, D, D, I, P, P, P, P]
        o.ora_solver1d.argtypes = [G, D, D, I, P]
        o.ora_mom_rhs.argtypes = [G, D, D] + [P] * 13
        o.ora_fftuvw.argtypes = [G, D, D] + [P] * 13
        o.ora_bounduvw.argtypes = [G, P, P, P]
        o.ora_boundp.argtypes = [G, P]
        o.ora_fillps.argtypes = [G, D, P, P, P, P]
        o.ora_correc.argtypes = [G, D, P, P, P, P]
        o.ora_fluid_step.argtypes = [G, D, D] + [P] * 14
        o.ora_chkdiv.argtypes = [G, P, P, P, ctypes.POINTER(D), ctypes.POINTER(D)]
        o._fluid_sig = True
    return o


def rfftf(x):
    x = np.array(x, dtype=np.float64)
    _flib().ora_rfftf(x.size, _p(x))
    return x


def rfftb(x):
    x = np.array(x, dtype=np.float64)
    _flib().ora_rfftb(x.size, _p(x))
    return x


def ref_vrfft(x, inverse=False):
    """The reference's own compiled VFFTPACK (one sequence), or None where not built."""
    r = ref_lib()
    if r is None:
        return None
    r.ref_vrfft.argtypes = [ctypes.c_int, P, ctypes.c_int]
    x = np.array(x, dtype=np.float64)
    r.ref_vrfft(x.size, _p(x), 1 if inverse else 0)
    return x


def initsolver(grid, rep, dt, e):
    k, n = grid.ktot, grid.itot * grid.jtot
    a, b, c, xy = np.zeros(k), np.zeros(k), np.zeros(k), np.zeros(n)
    gs = grid_struct(grid)
    _flib().ora_initsolver(ctypes.byref(gs), rep, dt, e, _p(a), _p(b), _p(c), _p(xy))
    return a, b, c, xy.reshape(grid.jtot, grid.itot).T


def solver1d(grid, rep, dt, ne, p):
    """In place on the interior of the Fortran-ordered field p."""
    gs = grid_struct(grid)
    _flib().ora_solver1d(ctypes.byref(gs), rep, dt, ne, _p(p))


def mom_rhs(grid, rep, dt, uo, vo, wo, uoo, voo, woo, p, fx, fy, fz):
    out = [grid.empty_field() for _ in range(3)]
    gs = grid_struct(grid)
    _flib().ora_mom_rhs(ctypes.byref(gs), rep, dt, *[_p(a) for a in (uo, vo, wo, uoo, voo, woo, p, fx, fy, fz)],
                        *[_p(a) for a in out])
    return out


def fftuvw(grid, rep, dt, uo, vo, wo, uoo, voo, woo, pnew, fx, fy, fz):
    out = [grid.empty_field() for _ in range(3)]
    gs = grid_struct(grid)
    _flib().ora_fftuvw(ctypes.byref(gs), rep, dt, *[_p(a) for a in out],
                       *[_p(a) for a in (uo, vo, wo, uoo, voo, woo, pnew, fx, fy, fz)])
    return out


def bounduvw(grid, u, v, w):
    gs = grid_struct(grid)
    _flib().ora_bounduvw(ctypes.byref(gs), _p(u), _p(v), _p(w))


def boundp(grid, p):
    gs = grid_struct(grid)
    _flib().ora_boundp(ctypes.byref(gs), _p(p))


def fillps(grid, dt, u, v, w):
    p = grid.empty_field()
    gs = grid_struct(grid)
    _flib().ora_fillps(ctypes.byref(gs), dt, _p(u), _p(v), _p(w), _p(p))
    return p


def correc(grid, dt, u, v, w, p):
    gs = grid_struct(grid)
    _flib().ora_correc(ctypes.byref(gs), dt, _p(u), _p(v), _p(w), _p(p))


def fluid_step(grid, rep, dt, unew, vnew, wnew, pnew, uo, vo, wo, uoo, voo, woo, fx, fy, fz):
    """main.f90:388-410 in place on unew, vnew, wnew, pnew; returns the correction pressure."""
    p = grid.empty_field()
    gs = grid_struct(grid)
    _flib().ora_fluid_step(ctypes.byref(gs), rep, dt, *[_p(a) for a in (unew, vnew, wnew, pnew, p, uo, vo, wo,
                                                                         uoo, voo, woo, fx, fy, fz)])
    return p


def chkdiv(grid, u, v, w):
    gs = grid_struct(grid)
    t, m = ctypes.c_double(), ctypes.c_double()
    _flib().ora_chkdiv(ctypes.byref(gs), _p(u), _p(v), _p(w), ctypes.byref(t), ctypes.byref(m))
    return t.value, m.value


# ---- wall collisions and the indicator function (oracle/indicator_oracle.c) ----------------
def _ilib():
    o = _flib()
    if not getattr(o, "_ind_sig", False):
        G = ctypes.POINTER(IbxGrid)
        D = ctypes.c_double
        o.ora_collisions.argtypes = [G, ctypes.c_int64, P, P, P, D, D, D]
        o.ora_indicator_div.argtypes = [G, P, P, P, P]
        o.ora_indicator_finish.argtypes = [G] + [P] * 5 + [D] * 5
        o.ora_indicator.argtypes = [G, D, D, ctypes.c_int64] + [P] * 16 + [D] * 5
        o._ind_sig = True
    return o


def collisions(grid, z, area, fz, a2, rad0, betacol):
    """collisions.f90:10-32: adds the wall force to fz in place."""
    gs = grid_struct(grid)
    _ilib().ora_collisions(ctypes.byref(gs), z.size, _p(z), _p(area), _p(fz), a2, rad0, betacol)


def indicator_div(grid, gx, gy, gz):
    indi = grid.empty_field()
    gs = grid_struct(grid)
    _ilib().ora_indicator_div(ctypes.byref(gs), _p(gx), _p(gy), _p(gz), _p(indi))
    return indi


def indicator_finish(grid, indi, forcey, forcez, lambda_, visc0, sin_t, cos_t, tcoef):
    """In place on indi (halos), forcey, forcez; returns (indic, visc)."""
    indic, visc = grid.empty_field(), grid.empty_field()
    gs = grid_struct(grid)
    _ilib().ora_indicator_finish(ctypes.byref(gs), _p(indi), _p(indic), _p(visc), _p(forcey),
                                 _p(forcez), lambda_, visc0, sin_t, cos_t, tcoef)
    return indic, visc


def indicator(grid, rep, dt, off, x, y, z, xn, yn, zn, area, forcey, forcez, lambda_, visc0,
              sin_t, cos_t, tcoef):
    """indicator.f90:19-182; z (shifted and back), forcey and forcez in place.  Returns
    (gx, gy, gz, indi, indic, visc)."""
    off = np.ascontiguousarray(off, dtype=np.int64)
    out = [grid.empty_field() for _ in range(6)]
    gs = grid_struct(grid)
    _ilib().ora_indicator(ctypes.byref(gs), rep, dt, off.size - 1, _p(off), _p(x), _p(y), _p(z),
                          _p(xn), _p(yn), _p(zn), _p(area), *[_p(a) for a in out], _p(forcey),
                          _p(forcez), lambda_, visc0, sin_t, cos_t, tcoef)
    return tuple(out)


# ---- the Fortran port of the transfer loops (oracle/ibx_port.f90, CPU baseline) ------------
PORT_VARIANTS = ("O0", "O2", "O0_omp", "O2_omp")
_ports = {}


def port_lib(variant):
    """oracle/_port/libibxport_<variant>.so (built by oracle/Makefile), or None if absent."""
    if variant not in _ports:
        path = os.path.join(ROOT, "oracle", "_port", f"libibxport_{variant}.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        D, I, I64 = ctypes.c_double, ctypes.c_int, ctypes.c_int64
        L.port_eulr2lagr.argtypes = [I, I, I, D, D, D, D, D, P, P, P, I64, P, P, P, P, P, P, P]
        L.port_lagr2eulr.argtypes = [I, I, I, D, D, D, D, D, D, I64, P, P, P, P, P, P, P, P, P,
                                     P, P]
        L.port_max_threads.restype = I
        L.port_set_threads.argtypes = [I]
        _ports[variant] = L
    return _ports[variant]


def port_eulr2lagr(variant, grid, u, v, w, off, x, y, z, threads=1):
    L = port_lib(variant)
    L.port_set_threads(int(threads))
    off = np.ascontiguousarray(off, dtype=np.int64)
    n = int(off[-1])
    out = [np.zeros(n) for _ in range(3)]
    c = np.ascontiguousarray
    L.port_eulr2lagr(grid.itot, grid.jtot, grid.ktot, grid.lx, grid.ly, grid.dxi, grid.dyi,
                     grid.dzi, _p(u), _p(v), _p(w), off.size - 1, _p(off), _p(c(x)), _p(c(y)),
                     _p(c(z)), *[_p(a) for a in out])
    return tuple(out)


def port_lagr2eulr(variant, grid, off, x, y, z, fx, fy, fz, area, threads=1, out=None):
    L = port_lib(variant)
    L.port_set_threads(int(threads))
    off = np.ascontiguousarray(off, dtype=np.int64)
    o = out if out is not None else [grid.empty_field() for _ in range(3)]
    c = np.ascontiguousarray
    L.port_lagr2eulr(grid.itot, grid.jtot, grid.ktot, grid.lx, grid.ly, grid.dxi, grid.dyi,
                     grid.dzi, grid.dx * grid.dy * grid.dz, off.size - 1, _p(off), _p(c(x)),
                     _p(c(y)), _p(c(z)), _p(c(fx)), _p(c(fy)), _p(c(fz)), _p(c(area)),
                     *[_p(a) for a in o])
    return tuple(o)
