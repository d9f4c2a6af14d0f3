"""Python host mirror of the reference's `mod_interp_spread` (interp_spread.f90:1-9).

The reference exports three argument-less subroutines that read and write module state
(particles `ap(1:pmax)` of `type particle`, common.f90:67-88; fields unew/vnew/wnew,
forcex/y/z, gx/gy/gz, common.f90:6-9).  Here the same routines take that state explicitly:

    IBTransfer(grid).eulr2lagr(u, v, w, caps)        -> (ul, vl, wl)     interp_spread.f90:12
    IBTransfer(grid).lagr2eulr(caps, fx, fy, fz)     -> (forcex, ...)    interp_spread.f90:510
    IBTransfer(grid).lagr2eulrind(caps, xn, yn, zn)  -> (gx, gy, gz)     interp_spread.f90:978
    IBTransfer(grid).update_pos(caps, ul, vl, wl, dt)                     Update_Pos.f90:77-186
    kernel(r)                                                             kernel.f90:7

Every call runs on the GPU through libibx (include/ibx.h); results are bit-identical to the
reference loop order.  Arrays are numpy (host) or, for the *_dev methods, device tensors
that expose `data_ptr()` (torch), in which case nothing leaves HBM.
"""
from __future__ import annotations

import ctypes
import weakref

import numpy as np

from . import _lib
from ._lib import IBX_OK


class IbxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libibx error {code}: {msg}")
        self.code = code


def kernel(r: float) -> float:
    """Regularized delta times h, kernel.f90:7-25 (host FP64)."""
    return _lib.load().ibx_kernel(float(r))


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(ctypes.c_void_p)
    return ctypes.c_void_p(int(a.data_ptr()))


def _f64(a, name):
    a = np.asarray(a)
    if a.dtype != np.float64:
        raise TypeError(f"{name} must be float64")
    return a


def _field(a, grid, name):
    a = _f64(a, name)
    if a.size != grid.ncells_halo:
        raise ValueError(f"{name} must hold (itot+2)(jtot+2)(ktot+2) = {grid.ncells_halo} values")
    if a.ndim == 3:
        if not a.flags.f_contiguous:
            raise ValueError(f"{name} must be Fortran-ordered (i fastest)")
    elif not a.flags.c_contiguous:
        raise ValueError(f"{name} must be contiguous")
    return a


def _pts(a, n, name):
    a = np.ascontiguousarray(_f64(a, name))
    if a.size != n:
        raise ValueError(f"{name} must hold {n} values")
    return a


def _offsets(cap_off):
    off = np.ascontiguousarray(np.asarray(cap_off, dtype=np.int64))
    if off.ndim != 1 or off.size < 1:
        raise ValueError("cap_off must be a 1-D array of ncap+1 offsets")
    return off


def _dev_check(name, t, n, integer=False):
    """A tensor that the library reads or writes through a raw pointer: its size, dtype,
    contiguity and placement are checked here, because the library cannot see them."""
    import torch

    want = torch.int64 if integer else torch.float64
    if not isinstance(t, torch.Tensor) or t.dtype != want or not t.is_cuda:
        raise ValueError(f"{name}: a {want} tensor on the GPU expected")
    if not t.is_contiguous():
        raise ValueError(f"{name}: contiguous tensor expected")
    if t.numel() != n:
        raise ValueError(f"{name}: {n} values expected, got {t.numel()}")


class IBTransfer:
    """One libibx context: a grid on one device."""

    def __init__(self, grid, device: int = 0):
        self.grid = grid
        self.lib = _lib.load()
        self._gs = _lib.grid_struct(grid)
        h = ctypes.c_void_p()
        rc = self.lib.ibx_create(ctypes.byref(h), ctypes.byref(self._gs), int(device))
        if rc != IBX_OK:
            raise IbxError(rc, "ibx_create failed (invalid grid or no HIP device)")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            # ibx_destroy also destroys the context's live batches: mark them closed
            for b in list(getattr(self, "_batches", ())):
                b.h = None
            self.lib.ibx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != IBX_OK:
            raise IbxError(rc, self.lib.ibx_last_error(self.h).decode())

    def set_stream(self, stream_handle):
        self._check(self.lib.ibx_set_stream(self.h, ctypes.c_void_p(stream_handle or 0)))

    def sync(self):
        self._check(self.lib.ibx_sync(self.h))

    def set_config(self, choice):
        """Spread configuration of the one-call entry points: an index or a name of
        Batch.config_names() (default objects-w8-o4)."""
        if isinstance(choice, str):
            choice = Batch.config_names().index(choice)
        self._check(self.lib.ibx_set_config(self.h, int(choice)))

    # ---- host arrays -------------------------------------------------------------------
    def eulr2lagr(self, u, v, w, cap_off, x, y, z):
        g = self.grid
        off = _offsets(cap_off)
        n = int(off[-1])
        u, v, w = (_field(a, g, nm) for a, nm in ((u, "u"), (v, "v"), (w, "w")))
        x, y, z = (_pts(a, n, nm) for a, nm in ((x, "x"), (y, "y"), (z, "z")))
        ul, vl, wl = (np.empty(n) for _ in range(3))
        self._check(self.lib.ibx_eulr2lagr(self.h, _ptr(u), _ptr(v), _ptr(w), off.size - 1,
                                           _ptr(off), _ptr(x), _ptr(y), _ptr(z), _ptr(ul),
                                           _ptr(vl), _ptr(wl)))
        return ul, vl, wl

    def _spread(self, fn, cap_off, x, y, z, f1, f2, f3, area):
        g = self.grid
        off = _offsets(cap_off)
        n = int(off[-1])
        x, y, z, f1, f2, f3, area = (
            _pts(a, n, nm) for a, nm in ((x, "x"), (y, "y"), (z, "z"), (f1, "f1"), (f2, "f2"),
                                         (f3, "f3"), (area, "area")))
        o = [g.empty_field() for _ in range(3)]
        self._check(fn(self.h, off.size - 1, _ptr(off), _ptr(x), _ptr(y), _ptr(z), _ptr(f1),
                       _ptr(f2), _ptr(f3), _ptr(area), _ptr(o[0]), _ptr(o[1]), _ptr(o[2])))
        return tuple(o)

    def lagr2eulr(self, cap_off, x, y, z, fx, fy, fz, area):
        return self._spread(self.lib.ibx_lagr2eulr, cap_off, x, y, z, fx, fy, fz, area)

    def lagr2eulrind(self, cap_off, x, y, z, xn, yn, zn, area):
        return self._spread(self.lib.ibx_lagr2eulrind, cap_off, x, y, z, xn, yn, zn, area)

    def update_pos(self, cap_off, x, y, z, ul, vl, wl, dt, xold=None, yold=None):
        """In-place Update_Pos.f90:77-186 on x,y,z (and xold,yold); returns
        (centroids[ncap,3], mean velocities[ncap,3], number of wall-clamped points)."""
        off = _offsets(cap_off)
        n = int(off[-1])
        for a in (x, y, z, ul, vl, wl) + ((xold, yold) if xold is not None else ()):
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous
                    and a.size == n):
                raise ValueError("update_pos needs contiguous float64 arrays of npts values")
        ncap = off.size - 1
        cen = np.empty((ncap, 3))
        um = np.empty((ncap, 3))
        nclamp = ctypes.c_int64(0)
        self._check(self.lib.ibx_update_pos(self.h, ncap, _ptr(off), float(dt), _ptr(x), _ptr(y),
                                            _ptr(z), _ptr(xold), _ptr(yold), _ptr(ul), _ptr(vl),
                                            _ptr(wl), _ptr(cen), _ptr(um), ctypes.byref(nclamp)))
        return cen, um, int(nclamp.value)

    # ---- HBM-resident tensors (torch) ----------------------------------------------------
    def interp_dev(self, u, v, w, cap_off, x, y, z, ul, vl, wl):
        n = int(x.numel())
        nf = self.grid.ncells_halo
        for nm, t in (("u", u), ("v", v), ("w", w)):
            _dev_check(nm, t, nf)
        for nm, t in (("x", x), ("y", y), ("z", z), ("ul", ul), ("vl", vl), ("wl", wl)):
            _dev_check(nm, t, n)
        _dev_check("cap_off", cap_off, int(cap_off.numel()), integer=True)
        self._check(self.lib.ibx_interp_f64_dev(self.h, _ptr(u), _ptr(v), _ptr(w),
                                                int(cap_off.numel()) - 1, _ptr(cap_off), n,
                                                _ptr(x), _ptr(y), _ptr(z), _ptr(ul), _ptr(vl),
                                                _ptr(wl)))

    def spread_dev(self, cap_off, x, y, z, fx, fy, fz, area, forcex, forcey, forcez, cc=False,
                   noalloc=False):
        """lagr2eulr (cc=False) or lagr2eulrind (cc=True) on device tensors.  noalloc=True:
        the plain spread through ibx_spread_f64_dev_noalloc, whose work buffers live in a
        fixed device array (no allocation in the call; calls on one GPU must not overlap)."""
        if cc and noalloc:
            raise ValueError("the no-allocation entry point is the plain spread (cc=False)")
        n, nf = int(x.numel()), self.grid.ncells_halo
        for nm, t in (("x", x), ("y", y), ("z", z), ("fx", fx), ("fy", fy), ("fz", fz),
                      ("area", area)):
            _dev_check(nm, t, n)
        for nm, t in (("forcex", forcex), ("forcey", forcey), ("forcez", forcez)):
            _dev_check(nm, t, nf)
        _dev_check("cap_off", cap_off, int(cap_off.numel()), integer=True)
        fn = (self.lib.ibx_spread_cc_f64_dev if cc else
              self.lib.ibx_spread_f64_dev_noalloc if noalloc else self.lib.ibx_spread_f64_dev)
        self._check(fn(self.h, int(cap_off.numel()) - 1, _ptr(cap_off), int(x.numel()), _ptr(x),
                       _ptr(y), _ptr(z), _ptr(fx), _ptr(fy), _ptr(fz), _ptr(area),
                       _ptr(forcex), _ptr(forcey), _ptr(forcez)))

    def collisions_dev(self, z, area, fz, a2, rad0, betacol):
        """collisions.f90:10-32 on device tensors: fz += the bottom wall's Morse force.
        a2 = 2 col / Rep^2, rad0 = 2.5 dz, betacol (param.f90:64; suguru_amd.params)."""
        n = int(z.numel())
        for nm, t in (("z", z), ("area", area), ("fz", fz)):
            _dev_check(nm, t, n)
        self._check(self.lib.ibx_collisions_dev(self.h, n, _ptr(z), _ptr(area), _ptr(fz),
                                                float(a2), float(rad0), float(betacol)))

    def shift_z_dev(self, z, s):
        """z += s in place (the indicator's shift of the points, indicator.f90:25-43)."""
        _dev_check("z", z, int(z.numel()))
        self._check(self.lib.ibx_shift_z_dev(self.h, int(z.numel()), _ptr(z), float(s)))

    def update_pos_dev(self, cap_off, x, y, z, ul, vl, wl, dt, centroid, umean, xold=None,
                       yold=None):
        n, ncap = int(x.numel()), int(cap_off.numel()) - 1
        for nm, t in (("x", x), ("y", y), ("z", z), ("ul", ul), ("vl", vl), ("wl", wl)) + (
                (("xold", xold), ("yold", yold)) if xold is not None else ()):
            _dev_check(nm, t, n)
        for nm, t in (("centroid", centroid), ("umean", umean)):
            _dev_check(nm, t, 3 * ncap)
        _dev_check("cap_off", cap_off, ncap + 1, integer=True)
        self._check(self.lib.ibx_update_pos_dev(self.h, int(cap_off.numel()) - 1, _ptr(cap_off),
                                                int(x.numel()), float(dt), _ptr(x), _ptr(y),
                                                _ptr(z), _ptr(xold), _ptr(yold), _ptr(ul),
                                                _ptr(vl), _ptr(wl), _ptr(centroid),
                                                _ptr(umean)))


class Batch:
    """Batched, graph-captured interp + spread over fixed HBM buffers (ibx_batch_*)."""

    def __init__(self, ctx: IBTransfer, cap_off_host):
        self.ctx = ctx
        self.lib = ctx.lib
        off = _offsets(cap_off_host)
        h = ctypes.c_void_p()
        ctx._check(self.lib.ibx_batch_create(ctx.h, ctypes.byref(h), off.size - 1, _ptr(off)))
        self.h = h
        self.npts = int(off[-1])
        self._keep = None
        if not hasattr(ctx, "_batches"):
            ctx._batches = weakref.WeakSet()
        ctx._batches.add(self)

    def bind(self, **bufs):
        nf, n = self.ctx.grid.ncells_halo, self.npts
        for k, v in bufs.items():
            _dev_check(k, v, nf if k in ("u", "v", "w", "forcex", "forcey", "forcez") else n)
        b = _lib.IbxBatchBufs(**{k: int(v.data_ptr()) for k, v in bufs.items()})
        self._keep = bufs
        self.ctx._check(self.lib.ibx_batch_bind(self.h, ctypes.byref(b)))

    def autotune(self, cache_path=None) -> int:
        ch = ctypes.c_int(-1)
        self.ctx._check(self.lib.ibx_batch_autotune(
            self.h, cache_path.encode() if cache_path else None, ctypes.byref(ch)))
        return ch.value

    def set_config(self, i: int):
        self.ctx._check(self.lib.ibx_batch_set_config(self.h, int(i)))

    def set_overlap(self, on: bool):
        """Interpolation on a second stream, concurrent with the spread."""
        self.ctx._check(self.lib.ibx_batch_set_overlap(self.h, 1 if on else 0))

    @staticmethod
    def config_names():
        lib = _lib.load()
        return [lib.ibx_batch_config_name(i).decode() for i in range(lib.ibx_batch_num_configs())]

    def capture(self, use_graph: bool = True):
        self.ctx._check(self.lib.ibx_batch_capture(self.h, 1 if use_graph else 0))

    def capture_timed(self, nslots: int):
        """Timed runs: run() enqueues the step eagerly with HIP events around each stage, run n
        recording into slot n % nslots (ibx_batch_capture_timed)."""
        self.ctx._check(self.lib.ibx_batch_capture_timed(self.h, int(nslots)))
        self._nslots = int(nslots)

    def run(self):
        self.ctx._check(self.lib.ibx_batch_run(self.h))

    def reset_slots(self):
        self.ctx._check(self.lib.ibx_batch_reset_slots(self.h))

    SLOT_KEYS = ("interp_us", "bin_us", "tile_spread_us", "step_us")

    def slot_times(self):
        """Per timed slot (its last replay): stage durations in microseconds."""
        n = getattr(self, "_nslots", 0)
        vals = (ctypes.c_double * (4 * n))()
        self.ctx._check(self.lib.ibx_batch_slot_times(self.h, vals))
        v = list(vals)
        return [dict(zip(self.SLOT_KEYS, v[4 * i:4 * i + 4])) for i in range(n)]

    PROFILE_KEYS = ("interp_us", "bin_us", "spread_us", "total_us", "spread_tiles_only_us",
                    "spread_zero_only_us", "nonempty_tiles", "tiles", "entries")

    def profile(self, reps: int = 10):
        """Average stage durations (HIP events, eager launches) and work counts."""
        vals = (ctypes.c_double * len(self.PROFILE_KEYS))()
        self.ctx._check(self.lib.ibx_batch_profile(self.h, int(reps), vals))
        return dict(zip(self.PROFILE_KEYS, list(vals)))

    def close(self):
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            self.lib.ibx_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
