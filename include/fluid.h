/*
 * fluid.h -- C ABI of the MI355X (gfx950) flow update of the capsule solver (libfluid.so).
 *
 * The library replaces the reference's FFT flow solver (fsolver = "f", param.f90:114) on one
 * rank: the momentum step with explicit Adams-Bashforth advection and implicit viscous terms,
 * then the pressure projection.
 *
 *   reference                                  | this ABI
 *   -------------------------------------------+--------------------------------------
 *   fftuvw(durk,dvrk,dwrk)  fftuvw.f90:13-127  | fld_fftuvw_dev
 *     momxp/momxad/...      mom.f90:9-394      |   (inside fld_fftuvw_dev)
 *   initsolver(e)           initsolver.f90     | fld_create (tables for e = 1, 2, 3)
 *   solver1d(p,ne)          solver.f90:43-157  | fld_solver1d_dev
 *   bounduvw(u,v,w)         bound.f90:7-80     | fld_bounduvw_dev
 *   boundp(p)               bound.f90:82-107   | fld_boundp_dev
 *   fillps(p)               fillps.f90:8-47    | fld_fillps_dev
 *   correc(p)               correc.f90:11-38   | fld_correc_dev
 *   the sequence            main.f90:388-410   | fld_step_dev
 *   chkdiv                  chkdiv.f90:7-61    | fld_chkdiv_dev
 *   indicator (ifoverlap=0) indicator.f90:47-168 | fld_indicator_dev (after the spread of
 *                                              |   the normals, ibx_spread_cc_f64_dev)
 *     divergence of G       indicator.f90:47-67  | fld_indicator_div_dev
 *     boundind .. forces    indicator.f90:91-168 | fld_indicator_finish_dev
 *
 * Fields are the reference's real(8) (0:i1,0:j1,0:k1) arrays (common.f90:6-9), i fastest,
 * in device memory.  The grid is the transfer library's ibx_grid (include/ibx.h); itot and
 * jtot must be powers of two (the x and y transforms are radix-2), ktot >= 2.  Boundary
 * conditions are those of iniu = 'poi' / 'zer' without free slip (bound.f90:38-56): no-slip
 * walls at k = 1/2 and kmax + 1/2, periodic in i and j.
 *
 * Arithmetic: IEEE FP64 without fused multiply-adds.  The stencil parts (momentum right-hand
 * side, fillps, correc, the boundary copies, the Gaussian elimination in k) evaluate the
 * reference's expressions in its order; the x and y transforms are this library's own radix-2
 * FFTs (the reference's VFFTPACK uses mixed radices and its own pi), so results agree with the
 * reference to rounding (~1e-14 relative), not bit for bit.
 *
 * All calls are asynchronous on the context stream unless stated; errors found on the
 * device are reported by fld_sync.  Return codes: 0 or a negative FLD_E_* code, with
 * fld_last_error() giving the message.
 */
#ifndef FLUID_H
#define FLUID_H

#include <stdint.h>

#include "ibx.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FLD_OK 0
#define FLD_E_ARG (-1)  /* invalid argument (grid not a power of two in i or j, null pointer) */
#define FLD_E_HIP (-2)  /* HIP runtime error */

typedef struct fld_ctx fld_ctx;

/* Context for one grid on HIP device `device`: Rep and dt as in param.f90:57 / main.f90:119.
 * Builds the eigenvalues and the elimination tables of initsolver(1..3) (initsolver.f90) on
 * the device and owns the spectral work arrays (3 x itot*jtot*ktot doubles).  No reference
 * counterpart for the context itself (module state). */
int fld_create(fld_ctx** out, const ibx_grid* grid, double rep, double dt, int device);
void fld_destroy(fld_ctx* ctx);
const char* fld_last_error(const fld_ctx* ctx);
/* Run subsequent work on `hip_stream` (a hipStream_t; NULL = the context's own stream). */
int fld_set_stream(fld_ctx* ctx, void* hip_stream);
/* Wait for the context stream. */
int fld_sync(fld_ctx* ctx);

/* fftuvw, fftuvw.f90:13-127: writes the interior of unew, vnew, wnew from the two previous
 * velocity fields (uo.., uoo..), the pressure pnew and the spread force field. */
int fld_fftuvw_dev(fld_ctx* ctx, double* unew, double* vnew, double* wnew, const double* uo,
                   const double* vo, const double* wo, const double* uoo, const double* voo,
                   const double* woo, const double* pnew, const double* forcex,
                   const double* forcey, const double* forcez);
/* solver1d(p, ne), solver.f90:43-157, in place on the interior of p: ne = 1 pressure
 * (Neumann in k), 2 u or v, 3 w (initsolver.f90:38-99). */
int fld_solver1d_dev(fld_ctx* ctx, double* p, int ne);
/* bounduvw, bound.f90:7-80 (wall ghosts, w = 0 on the wall faces, periodic halos). */
int fld_bounduvw_dev(fld_ctx* ctx, double* u, double* v, double* w);
/* boundp, bound.f90:82-107. */
int fld_boundp_dev(fld_ctx* ctx, double* p);
/* fillps, fillps.f90:8-47: the interior of p = divergence(u, v, w) / dt. */
int fld_fillps_dev(fld_ctx* ctx, const double* u, const double* v, const double* w, double* p);
/* correc, correc.f90:11-38. */
int fld_correc_dev(fld_ctx* ctx, double* u, double* v, double* w, const double* p);
/* The flow update of one time step, main.f90:388-410: fftuvw, bounduvw, fillps,
 * solver1d(p, 1), boundp(p), correc, bounduvw, pnew += p, boundp(pnew).  unew.. and pnew are
 * updated in place; p is the correction pressure (out). */
int fld_step_dev(fld_ctx* ctx, double* unew, double* vnew, double* wnew, double* pnew,
                 double* p, const double* uo, const double* vo, const double* wo,
                 const double* uoo, const double* voo, const double* woo,
                 const double* forcex, const double* forcey, const double* forcez);
/* The indicator function's divergence, indicator.f90:47-67: the interior of indi =
 * (gx(i+1)-gx(i-1)) .5 dxi + (gy(j+1)-gy(j-1)) .5 dyi + (gz(k+1)-gz(k-1)) .5 dzi, reading the
 * halos of G as lagr2eulrind leaves them (x and y periodic copies, k halos with their own
 * contributions).  Bit for bit the reference's expression. */
int fld_indicator_div_dev(fld_ctx* ctx, const double* gx, const double* gy, const double* gz,
                          double* indi);
/* The indicator after its Poisson solve, indicator.f90:91-168 (ifoverlap = 0):
 * boundind(indi) (= boundp, bound.f90:113-139); indic(k) = indi(k + ktot/4) for
 * k <= kmax - ktot/4, else 0; boundind(indic); on the interior indic < .01 -> 0,
 * indic >= .9 -> 1; visc = visc0 (1 + (lambda - 1) indic), boundp(visc);
 * forcey += .5 (indic(j) + indic(j+1)) sin_t tcoef, forcez += .5 (indic(k) + indic(k+1))
 * (-cos_t) tcoef, with sin_t = sin(thetainc), cos_t = cos(thetainc), tcoef =
 * (ta / Rep^2)^.25 (param.f90:52,57,59 folded on the host).  The halos of indic keep their
 * values from before the thresholds, as in the reference.  Bit for bit the reference's
 * expressions. */
int fld_indicator_finish_dev(fld_ctx* ctx, double* indi, double* indic, double* visc,
                             double* forcey, double* forcez, double lambda, double visc0,
                             double sin_t, double cos_t, double tcoef);
/* The indicator function, indicator.f90:47-168 (ifoverlap = 0): fld_indicator_div_dev,
 * the Poisson solve of initsolver(0) / solver1d(indi, 0) (the pressure solve, Neumann walls),
 * fld_indicator_finish_dev.  G = (gx, gy, gz) is the cell-centred spread of the normals with
 * the points shifted up by (ktot/4) dz (indicator.f90:25-43; ibx_shift_z_dev and
 * ibx_spread_cc_f64_dev).  The solve uses this library's FFTs: indi agrees with the reference
 * to rounding, and a cell whose value lies within rounding of .01 or .9 may threshold the
 * other way. */
int fld_indicator_dev(fld_ctx* ctx, const double* gx, const double* gy, const double* gz,
                      double* indi, double* indic, double* visc, double* forcey, double* forcez,
                      double lambda, double visc0, double sin_t, double cos_t, double tcoef);
/* chkdiv, chkdiv.f90:7-61: total and largest absolute discrete divergence of the interior
 * (host outputs; waits for the stream).  The total is summed in a different order than the
 * reference's sequential loop. */
int fld_chkdiv_dev(fld_ctx* ctx, const double* u, const double* v, const double* w,
                   double* divtot, double* divmax);

#ifdef __cplusplus
}
#endif

#endif /* FLUID_H */
