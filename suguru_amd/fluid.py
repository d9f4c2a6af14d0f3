"""The flow update on the GPU (libfluid.so, include/fluid.h): the reference's FFT solver
(fsolver = "f") on one rank -- fftuvw (fftuvw.f90:13-127), solver1d (solver.f90:43-157),
bounduvw / boundp (bound.f90:7-107), fillps (fillps.f90:8-47), correc (correc.f90:11-38) and
their sequence in the time step (main.f90:388-410).

Fields are torch float64 tensors of (itot+2)(jtot+2)(ktot+2) values on the GPU (Fortran order,
i fastest), passed to the library as raw device pointers after a size / dtype / placement
check.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes

from . import _lib
from .interp_spread import _dev_check, _ptr

REP = 0.0016  # particle Reynolds number (param.f90:57; suguru_amd.params.REP)
DT = 1e-4     # time step (main.f90:119, 491)


def _ind_consts(lambda_, visc0, sin_t, cos_t, tcoef):
    """The indicator's constants, each defaulting to param.f90's value."""
    from . import params

    st, ct, tc = params.centrifugal()
    pick = lambda v, d: float(d if v is None else v)
    return (pick(lambda_, params.LAMBDA), pick(visc0, params.VISC0), pick(sin_t, st),
            pick(cos_t, ct), pick(tcoef, tc))


class FluidError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libfluid error {code}: {msg}")
        self.code = code


class FlowSolver:
    """One libfluid context: the flow update of a grid on one device."""

    def __init__(self, grid, rep: float = REP, dt: float = DT, device: int = 0):
        self.grid = grid
        self.rep, self.dt = rep, dt
        self.lib = _lib.load_fluid()
        self._gs = _lib.grid_struct(grid)
        h = ctypes.c_void_p()
        rc = self.lib.fld_create(ctypes.byref(h), ctypes.byref(self._gs), float(rep), float(dt),
                                 int(device))
        if rc != 0:
            raise FluidError(rc, "fld_create failed (itot and jtot must be powers of two, "
                                 "ktot >= 2; or no HIP device)")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.fld_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise FluidError(rc, self.lib.fld_last_error(self.h).decode())

    def _fields(self, **kw):
        n = self.grid.ncells_halo
        for name, t in kw.items():
            _dev_check(name, t, n)
        return [_ptr(t) for t in kw.values()]

    def set_stream(self, stream_handle):
        self._check(self.lib.fld_set_stream(self.h, ctypes.c_void_p(stream_handle or 0)))

    def sync(self):
        self._check(self.lib.fld_sync(self.h))

    def fftuvw(self, unew, vnew, wnew, uo, vo, wo, uoo, voo, woo, pnew, fx, fy, fz):
        p = self._fields(unew=unew, vnew=vnew, wnew=wnew, uo=uo, vo=vo, wo=wo, uoo=uoo, voo=voo,
                         woo=woo, pnew=pnew, forcex=fx, forcey=fy, forcez=fz)
        self._check(self.lib.fld_fftuvw_dev(self.h, *p))

    def solver1d(self, p, ne: int):
        self._check(self.lib.fld_solver1d_dev(self.h, *self._fields(p=p), int(ne)))

    def bounduvw(self, u, v, w):
        self._check(self.lib.fld_bounduvw_dev(self.h, *self._fields(u=u, v=v, w=w)))

    def boundp(self, p):
        self._check(self.lib.fld_boundp_dev(self.h, *self._fields(p=p)))

    def fillps(self, u, v, w, p):
        self._check(self.lib.fld_fillps_dev(self.h, *self._fields(u=u, v=v, w=w, p=p)))

    def correc(self, u, v, w, p):
        self._check(self.lib.fld_correc_dev(self.h, *self._fields(u=u, v=v, w=w, p=p)))

    def step(self, unew, vnew, wnew, pnew, p, uo, vo, wo, uoo, voo, woo, fx, fy, fz):
        """main.f90:388-410 in place on unew, vnew, wnew, pnew; p receives the correction."""
        ptrs = self._fields(unew=unew, vnew=vnew, wnew=wnew, pnew=pnew, p=p, uo=uo, vo=vo, wo=wo,
                            uoo=uoo, voo=voo, woo=woo, forcex=fx, forcey=fy, forcez=fz)
        self._check(self.lib.fld_step_dev(self.h, *ptrs))

    def indicator_div(self, gx, gy, gz, indi):
        """indicator.f90:47-67: the interior of indi = divergence of G."""
        self._check(self.lib.fld_indicator_div_dev(self.h, *self._fields(gx=gx, gy=gy, gz=gz,
                                                                         indi=indi)))

    def indicator_finish(self, indi, indic, visc, forcey, forcez, lambda_=None, visc0=None,
                         sin_t=None, cos_t=None, tcoef=None):
        """indicator.f90:91-168 after the solve (constants default to param.f90's)."""
        k = _ind_consts(lambda_, visc0, sin_t, cos_t, tcoef)
        self._check(self.lib.fld_indicator_finish_dev(
            self.h, *self._fields(indi=indi, indic=indic, visc=visc, forcey=forcey,
                                  forcez=forcez), *k))

    def indicator(self, gx, gy, gz, indi, indic, visc, forcey, forcez, lambda_=None,
                  visc0=None, sin_t=None, cos_t=None, tcoef=None):
        """indicator.f90:47-168: divergence, Poisson solve, shift, thresholds, viscosity and
        centrifugal buoyancy; G is the cell-centred spread of the normals at z + (ktot/4) dz."""
        k = _ind_consts(lambda_, visc0, sin_t, cos_t, tcoef)
        self._check(self.lib.fld_indicator_dev(
            self.h, *self._fields(gx=gx, gy=gy, gz=gz, indi=indi, indic=indic, visc=visc,
                                  forcey=forcey, forcez=forcez), *k))

    def chkdiv(self, u, v, w):
        """(total, largest absolute) discrete divergence of the interior; waits."""
        t, m = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.fld_chkdiv_dev(self.h, *self._fields(u=u, v=v, w=w),
                                            ctypes.byref(t), ctypes.byref(m)))
        return t.value, m.value
