"""The coupled immersed-boundary time step on one GPU, in the reference's order.

Per step the reference runs (main.f90:188-410, fsolver = "f", no overlap mesh):

    uoo <- uo <- unew                                     (main.f90:197-205)
    eulr2lagr -> complagrforces -> (collisions) -> lagr2eulr -> (indicator) -> Update_Pos
    -> (surf_vol) -> fftuvw -> bounduvw -> fillps -> solver1d(p) -> boundp -> correc
    -> bounduvw -> pnew += p -> boundp(pnew)

surf_vol is a diagnostic and is left out.  The dense configuration (physics="dense") also
leaves out the wall collisions and the indicator function; physics="full" runs them in their
places.  Every stage is enqueued on one HIP stream, with no host round trip:

    1. ibx_interp_f64_dev          velocity at every point             (eulr2lagr)
    2. traction                    the FP64 law (memb_traction_f64_dev) or a surrogate
       [ibx_collisions_dev         wall Morse force added to fz        (collisions.f90:10-32)]
    3. ibx_spread_f64_dev_noalloc  traction x area onto the grid       (lagr2eulr)
       [indicator                  ibx_shift_z_dev(+s), ibx_spread_cc_f64_dev of the normals,
                                   ibx_shift_z_dev(-s), fld_indicator_dev: indicator function,
                                   viscosity, centrifugal buoyancy added to forcey / forcez
                                   (indicator.f90:19-182, s = (ktot/4) dz)]
    4. ibx_update_pos_dev          x += u dt, centroid, periodic shift (Update_Pos.f90:77-186)
    5. fld_step_dev                the flow update                     (main.f90:388-410)

The three velocity sets rotate by reference (no copies): the buffer that becomes unew is the
old uoo, whose interior fftuvw rewrites and whose halos bounduvw rewrites.  The rotation has
period three, so `capture()` records three hipGraphs, one per phase, and `run()` replays them
in turn; the spread's no-allocation entry point keeps hipMalloc out of the captured work.

The reference starts with uo = 0 (module arrays), so the first step's uoo is zero
(main.f90:190-205); the same holds here.
"""
from __future__ import annotations

import numpy as np

from .fluid import DT, REP, FlowSolver
from .interp_spread import IBTransfer, _offsets
from .membrane import MembranePlan
from .params import EBEND, GSCELL  # 1 / We (folded: 6249.999999999999) and 0.3 gscell


class CoupledStep:
    """Device state and the stage launches of the coupled step (module docstring).
    overlap: run the interpolation on a second stream beside the traction stage in the graphs
    and in untimed eager steps (steps timed with events stay sequential).  None (default):
    only with a traction object -- beside the FP64 law, whose persistent workgroups then
    share the compute units, the step was slower (37.3 against 36.5 ms at 8000 capsules); beside
    the surrogate stage it saves 0.17 ms (25.15 against 25.33 ms).

    grid: ibx-style grid (itot, jtot powers of two); cap_off: every capsule has the points of
    degree `degree`, latitude-major; gs, eb: the law's moduli; rep, dt: param.f90:57 and
    main.f90:119.  traction: None for the FP64 law, or an object with
    ``enqueue(x, y, z, fx, fy, fz, nx, ny, nz, area)`` that writes the same outputs on the
    step's stream without waiting (a surrogate).
    """

    def __init__(self, grid, cap_off, degree: int = 16, gs: float = GSCELL, eb: float = EBEND,
                 rep: float = REP, dt: float = DT, device: int = 0, spread_config="objects-w8",
                 traction=None, physics: str = "dense", overlap=None):
        import torch

        if physics not in ("dense", "full"):
            raise ValueError("physics must be 'dense' or 'full'")

        self.torch = torch
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
        self.fluid = FlowSolver(grid, rep, dt, device)
        self.grid, self.ncap, self.npts = grid, ncap, ncap * n
        self.gs, self.eb, self.dt = float(gs), float(eb), float(dt)
        self.traction = traction
        self.dev = torch.device("cuda", device)
        self.off = torch.from_numpy(off).to(self.dev)
        e = lambda k: torch.zeros(k, dtype=torch.float64, device=self.dev)
        nf = grid.ncells_halo
        # velocity sets [unew, uo, uoo] x (u, v, w); pressure, correction, spread force
        self.vel = [[e(nf) for _ in range(3)] for _ in range(3)]
        self.pnew, self.p = e(nf), e(nf)
        self.force = [e(nf) for _ in range(3)]
        self.x, self.y, self.z = e(self.npts), e(self.npts), e(self.npts)
        self.ul, self.vl, self.wl = e(self.npts), e(self.npts), e(self.npts)
        self.fx, self.fy, self.fz = e(self.npts), e(self.npts), e(self.npts)
        self.nx, self.ny, self.nz = e(self.npts), e(self.npts), e(self.npts)
        self.area = e(self.npts)
        self.centroid, self.umean = e(3 * ncap), e(3 * ncap)
        self.stream = torch.cuda.Stream(self.dev)
        self.side = torch.cuda.Stream(self.dev)   # the interpolation beside the traction
        self.overlap = None if overlap is None else bool(overlap)
        s = self.stream.cuda_stream
        self.ctx.set_stream(s)
        self.plan.set_stream(s)
        self.fluid.set_stream(s)
        self.physics = physics
        self.stage_names = ("interp", "traction", "spread", "update_pos", "fluid")
        if physics == "full":
            from . import params

            self.stage_names = ("interp", "traction", "collisions", "spread", "indicator",
                                "update_pos", "fluid")
            self.G = [e(nf) for _ in range(3)]
            self.indi, self.indic, self.visc = e(nf), e(nf), e(nf)
            self.a2 = params.collision_factor(rep)
            self.betacol = params.BETACOL
            self.rad0 = params.rad0(grid.dz)
            self.zshift = (grid.ktot // 4) * grid.dz
            # the normals' spread sizes its work buffers by the capsule count alone: one call
            # now, so that no allocation falls inside a later graph capture
            self.ctx.spread_dev(self.off, self.x, self.y, self.z, self.nx, self.ny, self.nz,
                                self.area, *self.G, cc=True)
            self.stream.synchronize()
            self.ctx.sync()
        self.phase = 0          # which velocity set is unew
        self.graphs = None
        self.steps_done = 0

    # ---- state ------------------------------------------------------------------------------
    def load(self, u, v, w, x, y, z, pnew=None):
        """Initial state (numpy Fortran-ordered fields, point arrays): unew = (u, v, w) with
        halos made consistent by bounduvw, uo = uoo = 0, pnew (default 0)."""
        t = self.torch

        def up(a):
            return t.from_numpy(np.ascontiguousarray(np.asarray(a).reshape(-1, order="F"))).to(self.dev)

        with t.cuda.stream(self.stream):
            for k, a in enumerate((u, v, w)):
                self.vel[0][k].copy_(up(a))
            for st in (1, 2):
                for k in range(3):
                    self.vel[st][k].zero_()
            self.pnew.copy_(up(pnew) if pnew is not None else t.zeros_like(self.pnew))
            self.p.zero_()
            for dst, a in zip((self.x, self.y, self.z), (x, y, z)):
                dst.copy_(up(a))
        self.phase = 0
        self._x0 = tuple(np.asarray(a, dtype=np.float64).copy() for a in (x, y, z))
        self.fluid.bounduvw(*self.vel[0])
        self.fluid.boundp(self.pnew)
        self.sync()

    def current(self):
        """(u, v, w) of the current unew set."""
        return self.vel[self.phase]

    # ---- one step -----------------------------------------------------------------------------
    def _sets(self, phase):
        """unew, uo, uoo sets after the rotation that starts the step at `phase` (the index of
        the set holding the current unew)."""
        uo = self.vel[phase]            # uo <- unew
        uoo = self.vel[(phase + 1) % 3]  # uoo <- uo (the previous step's uo)
        unew = self.vel[(phase + 2) % 3]  # the freed old uoo
        return unew, uo, uoo

    def _enqueue(self, phase, events=None):
        unew, uo, uoo = self._sets(phase)
        n = [0]

        def mark():
            if events is not None:
                events[n[0]].record(self.stream)
            n[0] += 1

        full = self.physics == "full"
        u, v, w = uo  # the velocity of this step: unew before the rotation (main.f90:225)
        # Without timing events the interpolation runs on a side stream beside the traction
        # (which reads only the positions) and joins before the first stage that writes the
        # positions (the indicator's shift, or Update_Pos) or reads the velocities (Update_Pos):
        # the same operations on the same data, so the same bits.
        ov = self.overlap if self.overlap is not None else self.traction is not None
        side = ov and events is None
        mark()
        if side:
            self.side.wait_stream(self.stream)
            self.ctx.set_stream(self.side.cuda_stream)
            try:
                self.ctx.interp_dev(u, v, w, self.off, self.x, self.y, self.z, self.ul, self.vl,
                                    self.wl)
            finally:
                self.ctx.set_stream(self.stream.cuda_stream)
        else:
            self.ctx.interp_dev(u, v, w, self.off, self.x, self.y, self.z, self.ul, self.vl,
                                self.wl)
        mark()
        if self.traction is None:
            self.plan.traction_dev(self.x, self.y, self.z, self.gs, self.eb, self.fx, self.fy,
                                   self.fz, self.nx, self.ny, self.nz, self.area)
        else:
            self.traction.enqueue(self.x, self.y, self.z, self.fx, self.fy, self.fz, self.nx,
                                  self.ny, self.nz, self.area)
        mark()
        if full:
            self.ctx.collisions_dev(self.z, self.area, self.fz, self.a2, self.rad0, self.betacol)
            mark()
        self.ctx.spread_dev(self.off, self.x, self.y, self.z, self.fx, self.fy, self.fz,
                            self.area, *self.force, noalloc=True)
        mark()
        if side:
            self.stream.wait_stream(self.side)
        if full:
            self.ctx.shift_z_dev(self.z, self.zshift)
            self.ctx.spread_dev(self.off, self.x, self.y, self.z, self.nx, self.ny, self.nz,
                                self.area, *self.G, cc=True)
            self.ctx.shift_z_dev(self.z, -self.zshift)
            self.fluid.indicator(*self.G, self.indi, self.indic, self.visc, self.force[1],
                                 self.force[2])
            mark()
        self.ctx.update_pos_dev(self.off, self.x, self.y, self.z, self.ul, self.vl, self.wl,
                                self.dt, self.centroid, self.umean)
        mark()
        self.fluid.step(*unew, self.pnew, self.p, *uo, *uoo, *self.force)
        mark()

    def step(self, events=None):
        """One eager step (asynchronous on the step's stream).  events: len(stage_names) + 1
        torch.cuda.Event recorded before, between and after the stages."""
        self._enqueue(self.phase, events)
        self.phase = (self.phase + 2) % 3
        self.steps_done += 1

    def capture(self):
        """Record the three phases of the step as hipGraphs (torch.cuda.CUDAGraph on the step's
        stream).  The state is not advanced: capture launches nothing."""
        t = self.torch
        self.graphs = []
        for ph in range(3):
            g = t.cuda.CUDAGraph()
            with t.cuda.graph(g, stream=self.stream):
                self._enqueue(ph)
            self.graphs.append(g)

    def run(self):
        """One step by replaying the captured graph of the current phase."""
        if self.graphs is None:
            raise RuntimeError("capture() first")
        with self.torch.cuda.stream(self.stream):
            self.graphs[self.phase].replay()
        self.phase = (self.phase + 2) % 3
        self.steps_done += 1

    # ---- output and restart files (suguru_amd.formats) --------------------------------------
    def _interior(self, t):
        g = self.grid
        a = t.view(g.ktot + 2, g.jtot + 2, g.itot + 2)[1:-1, 1:-1, 1:-1]
        return a.cpu().numpy().transpose(2, 1, 0)  # (itot, jtot, ktot), Fortran order

    def write_output(self, datadir: str, istep: int, time: float, dt=None):
        """The text output of main.f90:455-489 after step istep: Capsule###step#######.txt of
        every capsule, the center_velocity.txt lines (one per capsule: the first capsule's
        centroid and the last capsule's mean point velocity, as Update_Pos leaves them) and the
        timestep.txt line.  Waits for the stream."""
        import os

        from . import formats as F

        self.sync()
        n = self.npts // self.ncap
        x, y, z = (a.cpu().numpy().reshape(self.ncap, n) for a in (self.x, self.y, self.z))
        for p in range(self.ncap):
            F.write_capsule_txt(os.path.join(datadir, F.capsule_txt_name(p + 1, istep)), x[p],
                                y[p], z[p])
        cen = self.centroid.view(-1, 3).cpu().numpy()
        um = self.umean.view(-1, 3).cpu().numpy()
        F.append_center_velocity(os.path.join(datadir, "center_velocity.txt"), time, cen[0],
                                 um[-1], lines=self.ncap)
        F.append_timestep(os.path.join(datadir, "timestep.txt"), istep, time,
                          self.dt if dt is None else dt)

    def save_restart(self, datadir: str, istep: int, time: float):
        """fldbg####### (unew, vnew, wnew, uo, vo, wo, pnew and time, istep, dt) and
        allpartdata####### (per capsule: centroid, points, velocities, forces, areas; the
        initial points of the last load as xfp0; integralv, volume, volume0, dxdt.. and the
        older point sets 0: this step keeps none of them), loadd.f90:84-103, 588-666.  Waits."""
        import os

        from . import formats as F

        self.sync()
        unew, uo = self.vel[self.phase], self.vel[(self.phase + 1) % 3]
        fields = {"unew": unew[0], "vnew": unew[1], "wnew": unew[2], "uo": uo[0], "vo": uo[1],
                  "wo": uo[2], "pnew": self.pnew}
        F.write_fldbg(os.path.join(datadir, F.fldbg_name(istep)),
                      {k: self._interior(t) for k, t in fields.items()}, time, istep, self.dt)
        n = self.npts // self.ncap
        rec = np.zeros(self.ncap, dtype=F.particle_record_dtype(n))
        cen = self.centroid.view(-1, 3).cpu().numpy()
        rec["x"], rec["y"], rec["z"] = cen[:, 0], cen[:, 1], cen[:, 2]
        pts = {"xfp": self.x, "yfp": self.y, "zfp": self.z, "ul": self.ul, "vl": self.vl,
               "wl": self.wl, "fxl": self.fx, "fyl": self.fy, "fzl": self.fz, "area": self.area}
        for k, t in pts.items():
            rec[k] = t.cpu().numpy().reshape(self.ncap, n)
        if getattr(self, "_x0", None) is not None:
            for k, a in zip(("xfp0", "yfp0", "zfp0"), self._x0):
                rec[k] = a.reshape(self.ncap, n)
        F.write_allpartdata(os.path.join(datadir, F.allpartdata_name(istep)), rec)

    def load_restart(self, datadir: str, istep: int):
        """Reads fldbg####### and allpartdata####### (loadflds(0) / loadpart(0)): unew and uo
        interiors with their halos rebuilt by bounduvw, pnew with boundp, the capsule points;
        uoo is not stored (the next step makes it uo).  Returns (time, istep, dt)."""
        import os

        from . import formats as F

        g, t = self.grid, self.torch
        fields, time, nr, dt = F.read_fldbg(os.path.join(datadir, F.fldbg_name(istep)), g.itot,
                                            g.jtot, g.ktot)
        n = self.npts // self.ncap
        rec = F.read_allpartdata(os.path.join(datadir, F.allpartdata_name(istep)), self.ncap, n)

        def put(dst, a):  # interior (itot, jtot, ktot) into a halo field
            v = dst.view(g.ktot + 2, g.jtot + 2, g.itot + 2)
            v[1:-1, 1:-1, 1:-1] = t.from_numpy(np.ascontiguousarray(a.transpose(2, 1, 0))).to(self.dev)

        with t.cuda.stream(self.stream):
            for st in range(3):
                for k in range(3):
                    self.vel[st][k].zero_()
            for k, nm in enumerate(("unew", "vnew", "wnew")):
                put(self.vel[0][k], fields[nm])
            for k, nm in enumerate(("uo", "vo", "wo")):
                put(self.vel[1][k], fields[nm])
            self.pnew.zero_()
            put(self.pnew, fields["pnew"])
            self.p.zero_()
            for dst, nm in zip((self.x, self.y, self.z), ("xfp", "yfp", "zfp")):
                dst.copy_(t.from_numpy(np.ascontiguousarray(rec[nm].reshape(-1))).to(self.dev))
        self.phase = 0
        self.fluid.bounduvw(*self.vel[0])
        self.fluid.bounduvw(*self.vel[1])
        self.fluid.boundp(self.pnew)
        self.sync()
        return time, nr, dt

    def sync(self):
        """Waits for the stream and raises the first error any library flagged."""
        self.stream.synchronize()
        self.ctx.sync()
        self.plan.sync()
        self.fluid.sync()

    def close(self):
        self.graphs = None
        self.plan.close()
        self.ctx.close()
        self.fluid.close()
