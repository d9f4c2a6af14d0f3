"""One capsule time step on the GPU, in the reference's order, without the fluid update.

Per step the reference runs (main.f90:225-241):

    eulr2lagr -> complagrforces -> (collisions) -> lagr2eulr -> (indicator) -> Update_Pos
    -> (surf_vol) -> flow solver

The dense configuration leaves out collisions and the indicator.  The flow solver is not part
of these libraries: the bound velocity field stays as it is (frozen), and the spread force
field is what the step hands to it.  All four stages are enqueued on one HIP stream, so the
step is stream-ordered with no host round trip:

    1. ibx_interp_f64_dev      velocity at every point             (eulr2lagr)
    2. memb_traction_f64_dev   membrane traction, normals, areas   (complagrforces)
    3. ibx_spread_f64_dev      traction x area onto the grid       (lagr2eulr)
    4. ibx_update_pos_dev      x += u dt, centroid, periodic shift (Update_Pos.f90:77-186)

Every capsule has the same degree L (one membrane plan) and its points in the law's
latitude-major Gauss-grid order; the law's stress-free state is the unit sphere
(sph.f90:56-70), so the capsules' rest radius is 1 in grid units of length.
"""
from __future__ import annotations

import numpy as np

from .interp_spread import IBTransfer, _offsets
from .membrane import MembranePlan


class CapsuleStep:
    """Device buffers and the four stage launches of one step (see the module docstring).

    gs, eb: shear and bending moduli of the law (gscell, ebend; param.f90:88,91).
    dt: the time step of the position update (main.f90:119).
    spread_config: the spread's configuration (index or name; default: the library's).
    """

    def __init__(self, grid, cap_off, degree: int, gs: float, eb: float, dt: float,
                 device: int = 0, spread_config=None):
        import torch

        off = _offsets(cap_off)
        self.plan = MembranePlan(degree, device)
        n = self.plan.npts
        ncap = off.size - 1
        if not np.array_equal(off, np.arange(ncap + 1, dtype=np.int64) * n):
            self.plan.close()
            raise ValueError(f"every capsule must have the {n} points of degree {degree}")
        self.ctx = IBTransfer(grid, device)
        if spread_config is not None:
            self.ctx.set_config(spread_config)
        self.grid, self.ncap, self.npts = grid, ncap, ncap * n
        self.gs, self.eb, self.dt = float(gs), float(eb), float(dt)
        self.dev = torch.device("cuda", device)
        self.off = torch.from_numpy(off).to(self.dev)
        e = lambda k: torch.empty(k, dtype=torch.float64, device=self.dev)
        # per point: velocity, traction, normal, area; per capsule: centroid, mean velocity
        self.ul, self.vl, self.wl = e(self.npts), e(self.npts), e(self.npts)
        self.fx, self.fy, self.fz = e(self.npts), e(self.npts), e(self.npts)
        self.nx, self.ny, self.nz = e(self.npts), e(self.npts), e(self.npts)
        self.area = e(self.npts)
        self.centroid, self.umean = e(3 * ncap), e(3 * ncap)
        self.bound = None

    def bind(self, u, v, w, x, y, z, forcex, forcey, forcez, stream=None):
        """HBM-resident inputs and outputs (float64 torch tensors on the step's device):
        the velocity fields u, v, w ((itot+2)(jtot+2)(ktot+2), i fastest), the positions
        x, y, z (updated in place every step) and the force fields (written every step).
        stream: the HIP stream handle both libraries run the stages on (default: a stream of
        the step's own).  A null handle is refused: it selects each library's own stream,
        and the two would then not be ordered with each other.  Work that produced the
        inputs on another stream must be complete (or waited for) before step()."""
        import torch

        if stream is None:
            self._stream = torch.cuda.Stream(self.dev)
            stream = self._stream.cuda_stream
        if not stream:
            raise ValueError("a non-null stream handle is needed (0 selects each library's own "
                             "stream, which are not ordered with each other)")
        nf = self.grid.ncells_halo
        for name, t, k in (("u", u, nf), ("v", v, nf), ("w", w, nf), ("x", x, self.npts),
                           ("y", y, self.npts), ("z", z, self.npts), ("forcex", forcex, nf),
                           ("forcey", forcey, nf), ("forcez", forcez, nf)):
            if t.dtype != torch.float64 or t.device != self.dev or not t.is_contiguous():
                raise ValueError(f"{name}: contiguous float64 on {self.dev} expected")
            if t.numel() != k:  # the libraries see raw pointers: sizes are checked here
                raise ValueError(f"{name}: {k} values expected, got {t.numel()}")
        self.ctx.set_stream(stream)
        self.plan.set_stream(stream)
        self.stream = torch.cuda.ExternalStream(stream, device=self.dev)
        self.bound = (u, v, w, x, y, z, forcex, forcey, forcez)

    def step(self, events=None):
        """Enqueues one step on the bound stream (asynchronous).  events: optionally five
        torch.cuda.Event, recorded on that stream before, between and after the four stages."""
        if self.bound is None:
            raise RuntimeError("bind() first")
        u, v, w, x, y, z, F1, F2, F3 = self.bound
        mark = (lambda i: events[i].record(self.stream)) if events is not None else (lambda i: None)
        mark(0)
        self.ctx.interp_dev(u, v, w, self.off, x, y, z, self.ul, self.vl, self.wl)
        mark(1)
        self.plan.traction_dev(x, y, z, self.gs, self.eb, self.fx, self.fy, self.fz, self.nx,
                               self.ny, self.nz, self.area)
        mark(2)
        self.ctx.spread_dev(self.off, x, y, z, self.fx, self.fy, self.fz, self.area, F1, F2, F3)
        mark(3)
        self.ctx.update_pos_dev(self.off, x, y, z, self.ul, self.vl, self.wl, self.dt,
                                self.centroid, self.umean)
        mark(4)

    def sync(self):
        """Waits for the enqueued steps and raises the first error they flagged."""
        self.ctx.sync()
        self.plan.sync()

    def close(self):
        self.plan.close()
        self.ctx.close()
