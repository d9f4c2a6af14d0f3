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
