/* CPU oracle (test infrastructure only; never linked by the product): the wall collision
 * force and the indicator function of the reference, restated in plain C, FP64, without
 * fused multiply-adds (built with -ffp-contract=off), in the reference's operation order.
 *
 *   ora_collisions        collisions.f90:10-32   Morse wall force added to fzl
 *   ora_indicator_div     indicator.f90:47-67    indi = central divergence of G
 *   ora_indicator_finish  indicator.f90:89-168   (after the solve) boundind, the shift back
 *                                                by ktot/4 planes, thresholds, viscosity,
 *                                                centrifugal buoyancy added to forcey/forcez
 *   ora_indicator         indicator.f90:19-182   the whole routine: z shifted up by
 *                                                (ktot/4) dz, lagr2eulrind, z shifted back,
 *                                                divergence, solver1d(indi, 0), finish
 *
 * ifoverlap = 0 (param.f90:32): the fine-mesh branches are not restated.  boundind
 * (bound.f90:113-139) is boundp (bound.f90:88-111) under another name: Neumann copies of the
 * k halos, then the periodic x and y halos.  initsolver(0) builds the same main-mesh matrix
 * as initsolver(1) (initsolver.f90:37-60), and solver1d treats ne = 0 as ne = 1
 * (solver.f90:110-115), so the solve is ora_solver1d(..., ne = 1, ...).
 *
 * Constants: the reference folds them at compile time (param.f90:48-64; -autodouble
 * -fpconstant, Makefile:7); the caller passes their FP64 values (sin thetainc, cos thetainc,
 * (ta / Rep^2)^0.25, col / Rep^2 and so on are computed once on the host).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/ibx.h"

#define IDX(g, i, j, k) \
  ((size_t)(i) + (size_t)((g)->itot + 2) * ((size_t)(j) + (size_t)((g)->jtot + 2) * (size_t)(k)))

void ora_updthalos(const ibx_grid* g, double* var, int dir);                 /* ibx_oracle.c */
void ora_boundp(const ibx_grid* g, double* p);                                /* fluid_oracle.c */
void ora_solver1d(const ibx_grid* g, double rep, double dt, int ne, double* p); /* fluid_oracle.c */
void ora_lagr2eulrind(const ibx_grid* g, int64_t ncap, const int64_t* off, const double* x,
                      const double* y, const double* z, const double* xn, const double* yn,
                      const double* zn, const double* area, double* gx, double* gy,
                      double* gz); /* ibx_oracle.c */

/* collisions.f90:10-32 for every point (one rank: every capsule is a master).
 * a2 = 2 col / Rep^2 (the factor 2.*(col/(Rep**2.)), :24), rad0 = the cutoff and zero-force
 * distance (:14, param.f90:64), betacol (param.f90:64).  The plane sum runs over
 * j, k in [-jmaxcol, jmaxcol], jmaxcol = nint(sqrt(max(rad0^2 - z^2, 0)) dyi) (:19); j steps
 * dy and k steps dx (:22). */
void ora_collisions(const ibx_grid* g, int64_t npts, const double* z, const double* area,
                    double* fz, double a2, double rad0, double betacol) {
  const double dx2 = g->dx * g->dx;
  for (int64_t q = 0; q < npts; ++q) {
    const double zq = z[q];
    double fc = 0.0;
    const double r2 = rad0 * rad0 - zq * zq;
    const long jm = lround(sqrt(r2 > 0.0 ? r2 : 0.0) * g->dyi);
    for (long j = -jm; j <= jm; ++j)
      for (long k = -jm; k <= jm; ++k) {
        const double jy = (double)j * g->dy, kx = (double)k * g->dx;
        const double ss = sqrt(zq * zq + jy * jy + kx * kx);
        if (ss < rad0) {
          const double e2 = exp(2.0 * betacol * (rad0 - ss));
          const double e1 = exp(betacol * (rad0 - ss));
          const double a = area[q];
          const double lo = a < dx2 ? a : dx2, hi = a > dx2 ? a : dx2;
          fc = fc + a2 * (e2 - e1) * fabs(zq / ss) * (lo / hi);
        }
      }
    fz[q] = fz[q] + fc;
  }
}

/* indicator.f90:51-67: indi(i,j,k) = dgx + dgy + dgz with dgx = (gx(i+1)-gx(i-1)) .5 dxi
 * etc., over the interior. */
void ora_indicator_div(const ibx_grid* g, const double* gx, const double* gy, const double* gz,
                       double* indi) {
  for (int k = 1; k <= g->ktot; ++k)
    for (int j = 1; j <= g->jtot; ++j)
      for (int i = 1; i <= g->itot; ++i) {
        const double dgx = (gx[IDX(g, i + 1, j, k)] - gx[IDX(g, i - 1, j, k)]) * 0.5 * g->dxi;
        const double dgy = (gy[IDX(g, i, j + 1, k)] - gy[IDX(g, i, j - 1, k)]) * 0.5 * g->dyi;
        const double dgz = (gz[IDX(g, i, j, k + 1)] - gz[IDX(g, i, j, k - 1)]) * 0.5 * g->dzi;
        indi[IDX(g, i, j, k)] = dgx + dgy + dgz;
      }
}

/* indicator.f90:91-168 after the solve (ifoverlap = 0):
 *   boundind(indi); indic(k) = indi(k + ktot/4) for k <= kmax - ktot/4, else 0;
 *   boundind(indic); thresholds on the interior (< .01 -> 0, >= .9 -> 1);
 *   visc = visc0 (1 + (lambda - 1) indic) on the interior, boundp(visc);
 *   forcey += .5 (indic(j) + indic(j+1)) sin(thetainc) tcoef,
 *   forcez += .5 (indic(k) + indic(k+1)) (-cos(thetainc)) tcoef, tcoef = (ta / Rep^2)^.25.
 * The halos of indic keep the values boundind gave them before the thresholds (the
 * thresholds touch the interior only), and the forces at j = jmax, k = kmax read them. */
void ora_indicator_finish(const ibx_grid* g, double* indi, double* indic, double* visc,
                          double* forcey, double* forcez, double lambda, double visc0,
                          double sin_t, double cos_t, double tcoef) {
  const int kq = g->ktot / 4;
  ora_boundp(g, indi);
  for (int k = 1; k <= g->ktot; ++k)
    for (int j = 1; j <= g->jtot; ++j)
      for (int i = 1; i <= g->itot; ++i)
        indic[IDX(g, i, j, k)] = k <= g->ktot - kq ? indi[IDX(g, i, j, k + kq)] : 0.0;
  ora_boundp(g, indic);
  for (int k = 1; k <= g->ktot; ++k)
    for (int j = 1; j <= g->jtot; ++j)
      for (int i = 1; i <= g->itot; ++i) {
        double* v = &indic[IDX(g, i, j, k)];
        if (*v < 0.01) *v = 0.0;
        if (*v >= 0.9) *v = 1.0;
      }
  for (int k = 1; k <= g->ktot; ++k)
    for (int j = 1; j <= g->jtot; ++j)
      for (int i = 1; i <= g->itot; ++i)
        visc[IDX(g, i, j, k)] = visc0 * (1.0 + (lambda - 1.0) * indic[IDX(g, i, j, k)]);
  ora_boundp(g, visc);
  const double mcos = -cos_t;
  for (int k = 1; k <= g->ktot; ++k)
    for (int j = 1; j <= g->jtot; ++j)
      for (int i = 1; i <= g->itot; ++i) {
        const size_t c = IDX(g, i, j, k);
        forcey[c] = forcey[c] + 0.5 * (indic[c] + indic[IDX(g, i, j + 1, k)]) * sin_t * tcoef;
        forcez[c] = forcez[c] + 0.5 * (indic[c] + indic[IDX(g, i, j, k + 1)]) * mcos * tcoef;
      }
}

/* indicator.f90:19-182 (ifoverlap = 0).  z is shifted up by (ktot/4) dz and back in place,
 * as the reference does with ap(p)%zfp (:25-43), so it returns as (z + s) - s.  gx, gy, gz,
 * indi, indic, visc: full fields with halos (outputs); forcey, forcez: added to. */
void ora_indicator(const ibx_grid* g, double rep, double dt, int64_t ncap, const int64_t* off,
                   const double* x, const double* y, double* z, const double* xn,
                   const double* yn, const double* zn, const double* area, double* gx,
                   double* gy, double* gz, double* indi, double* indic, double* visc,
                   double* forcey, double* forcez, double lambda, double visc0, double sin_t,
                   double cos_t, double tcoef) {
  const double s = (double)(g->ktot / 4) * g->dz;
  const int64_t n = off[ncap];
  for (int64_t q = 0; q < n; ++q) z[q] = z[q] + s;
  ora_lagr2eulrind(g, ncap, off, x, y, z, xn, yn, zn, area, gx, gy, gz);
  for (int64_t q = 0; q < n; ++q) z[q] = z[q] - s;
  ora_indicator_div(g, gx, gy, gz, indi);
  ora_solver1d(g, rep, dt, 1, indi);
  ora_indicator_finish(g, indi, indic, visc, forcey, forcez, lambda, visc0, sin_t, cos_t, tcoef);
}
