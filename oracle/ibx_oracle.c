/*
 * ibx_oracle.c -- CPU restatement of the reference's immersed-boundary transfer.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke test and the
 * CPU-baseline leg of bench.py load it.  The product (suguru_amd/, libibx.so) never links,
 * calls or falls back to it.
 *
 * It restates, loop for loop and in the same floating-point evaluation order, the
 * reference's single-rank arithmetic:
 *   ora_kernel          kernel.f90:7-25
 *   ora_kerneltest      kernel.f90:43-63
 *   ora_updthalos       bound.f90:141-190 (one rank: MPI_SENDRECV to self = periodic copy;
 *                       halo datatypes initmpi.f90:144-151)
 *   ora_eulr2lagr       interp_spread.f90:12-475 (image copies :60-146, partial integration
 *                       :315-383, owner sum over slots :426-450)
 *   ora_lagr2eulr       interp_spread.f90:510-922 (zeroing :811-815, scatter :817-884,
 *                       halos :904-912)
 *   ora_lagr2eulrind    interp_spread.f90:978-1387 (clamps :1311-1316, weights :1334-1338)
 *   ora_update_pos      Update_Pos.f90:77-186
 *
 * One rank (dims = 1x1): every capsule is a master whose eight neighbours are itself, so
 * for nb = 1..8 the copy is stored in slot nbrecv = nb+4 (mod 8) and shifted by the table
 * at interp_spread.f90:82-141.  All eight images are formed for every capsule.  The
 * reference forms image nb only when some point of the capsule lies within
 * `offset` = 2.61 h of the matching boundary (initparticles.f90:241-295, param.f90:135);
 * an image that is not formed there contributes exactly nothing here, because a shifted
 * point reaches a tap inside [1,imax] only if it lies less than 1 h outside the domain.
 *
 * Compile with -ffp-contract=off (no fused multiply-add) so that every product and sum is
 * rounded as the reference's scalar Fortran does at -O0 (Makefile:7).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/ibx.h"

/* kernel.f90:7-25.  (1./6.) and (1./3.) are double constants (-autodouble, Makefile:7);
 * x**2 is x*x; a-b-c is (a-b)-c. */
double ora_kernel(double r) {
  double ar = fabs(r);
  if (ar > 1.5) return 0.0;
  if (ar > 0.5) {
    double t = 1.0 - ar;
    return (1.0 / 6.0) * ((5.0 - 3.0 * ar) - sqrt(-(3.0 * (t * t)) + 1.0));
  }
  return (1.0 / 3.0) * (1.0 + sqrt(-(3.0 * (r * r)) + 1.0));
}

/* kernel.f90:43-63: Riemann sum of the tensor-product kernel on a 1/12 lattice. */
double ora_kerneltest(void) {
  double sum = 0.0;
  for (int k = -25; k <= 25; ++k) {
    double coorz = (double)k / 12.0;
    for (int j = -25; j <= 25; ++j) {
      double coory = (double)j / 12.0;
      for (int i = -25; i <= 25; ++i) {
        double coorx = (double)i / 12.0;
        sum = sum + (1.0 / 12.0) * (1.0 / 12.0) * (1.0 / 12.0) * ora_kernel(coorx) *
                        ora_kernel(coory) * ora_kernel(coorz);
      }
    }
  }
  return sum;
}

#define IDX(g, i, j, k) \
  ((size_t)(i) + (size_t)((g)->itot + 2) * ((size_t)(j) + (size_t)((g)->jtot + 2) * (size_t)(k)))

/* bound.f90:141-190 on one rank.  dir 1: var(i1,:,:) = var(1,:,:) then var(0,:,:) =
 * var(imax,:,:) over all j,k (xhalo = (j1+1)(k1+1) strided singles).  dir 2: var(:,j1,:) =
 * var(:,1,:) then var(:,0,:) = var(:,jmax,:) over all i,k (yhalo). */
void ora_updthalos(const ibx_grid* g, double* var, int dir) {
  int i1 = g->itot + 1, j1 = g->jtot + 1, k1 = g->ktot + 1;
  if (dir == 1) {
    for (int k = 0; k <= k1; ++k)
      for (int j = 0; j <= j1; ++j) var[IDX(g, i1, j, k)] = var[IDX(g, 1, j, k)];
    for (int k = 0; k <= k1; ++k)
      for (int j = 0; j <= j1; ++j) var[IDX(g, 0, j, k)] = var[IDX(g, g->itot, j, k)];
  } else {
    for (int k = 0; k <= k1; ++k)
      for (int i = 0; i <= i1; ++i) var[IDX(g, i, j1, k)] = var[IDX(g, i, 1, k)];
    for (int k = 0; k <= k1; ++k)
      for (int i = 0; i <= i1; ++i) var[IDX(g, i, 0, k)] = var[IDX(g, i, g->jtot, k)];
  }
}

/* Shift of the copy held in slot s (interp_spread.f90:82-141), in units of (lx, ly). */
static const int SLOT_SX[9] = {0, +1, +1, 0, -1, -1, -1, 0, +1};
static const int SLOT_SY[9] = {0, 0, -1, -1, -1, 0, +1, +1, +1};

/* Loop index nb -> data slot nbrecv for a master that is its own neighbour
 * (interp_spread.f90:333-336). */
static int nb_to_slot(int nb) {
  if (nb == 0) return 0;
  int s = nb + 4;
  if (s > 8) s -= 8;
  return s;
}

/* Copy of a point as placed in slot s: anb(s)%xfp = anb(0)%xfp, then +-lx / +-ly. */
static void shifted(const ibx_grid* g, int s, double x, double y, double* xs, double* ys) {
  *xs = x;
  *ys = y;
  if (SLOT_SX[s] > 0) *xs = *xs + g->lx;
  if (SLOT_SX[s] < 0) *xs = *xs - g->lx;
  if (SLOT_SY[s] > 0) *ys = *ys + g->ly;
  if (SLOT_SY[s] < 0) *ys = *ys - g->ly;
}

/* Fortran nint: round half away from zero. */
static long fnint(double a) { return lround(a); }

/* eulr2lagr, interp_spread.f90:12-475, one rank. */
void ora_eulr2lagr(const ibx_grid* g, const double* u, const double* v, const double* w,
                   int64_t ncap, const int64_t* off, const double* x, const double* y,
                   const double* z, double* ul, double* vl, double* wl) {
  const int imax = g->itot, jmax = g->jtot, kmax = g->ktot;
  for (int64_t p = 0; p < ncap; ++p) {
    int64_t n = off[p + 1] - off[p];
    if (n <= 0) continue;
    const double* xp = x + off[p];
    const double* yp = y + off[p];
    const double* zp = z + off[p];
    /* anb(0:8)%ul/vl/wl, zeroed (:300-312). */
    double* an = (double*)calloc((size_t)(9 * 3 * n), sizeof(double));
    for (int nb = 0; nb <= 8; ++nb) {
      int s = nb_to_slot(nb);
      double* aul = an + (size_t)(s * 3 + 0) * n;
      double* avl = an + (size_t)(s * 3 + 1) * n;
      double* awl = an + (size_t)(s * 3 + 2) * n;
      for (int64_t l = 0; l < n; ++l) {
        double xs, ys;
        shifted(g, s, xp[l], yp[l], &xs, &ys);
        double coorzfp = zp[l] * g->dzi;
        double coorxfp = (xs - g->boundleft) * g->dxi;
        double cooryfp = (ys - g->boundfront) * g->dyi;
        long ilow = fnint(coorxfp - 2.0), ihigh = fnint(coorxfp + 2.0);
        long jlow = fnint(cooryfp - 2.0), jhigh = fnint(cooryfp + 2.0);
        long klow = fnint(coorzfp - 2.0), khigh = fnint(coorzfp + 2.0);
        if (ilow < 1) ilow = 1;
        if (jlow < 1) jlow = 1;
        if (klow < 1) klow = 1;
        if (ihigh > imax) ihigh = imax;
        if (jhigh > jmax) jhigh = jmax;
        if (khigh > kmax) khigh = kmax;
        for (long k = klow; k <= khigh; ++k) {
          double coorzs = (1.0 * (double)k) - coorzfp;
          double coorz = coorzs - 0.5;
          double kernelzs = ora_kernel(coorzs);
          double kernelz = ora_kernel(coorz);
          for (long j = jlow; j <= jhigh; ++j) {
            double coorys = (1.0 * (double)j) - cooryfp;
            double coory = coorys - 0.5;
            double kernelys = ora_kernel(coorys);
            double kernely = ora_kernel(coory);
            for (long i = ilow; i <= ihigh; ++i) {
              double coorxs = (1.0 * (double)i) - coorxfp;
              double coorx = coorxs - 0.5;
              double kernelxs = ora_kernel(coorxs);
              double kernelx = ora_kernel(coorx);
              size_t c = IDX(g, i, j, k);
              aul[l] = aul[l] + u[c] * kernelxs * kernely * kernelz;
              avl[l] = avl[l] + v[c] * kernelx * kernelys * kernelz;
              awl[l] = awl[l] + w[c] * kernelx * kernely * kernelzs;
            }
          }
        }
      }
    }
    /* Owner sum over slots 0..8 (:426-440). */
    for (int64_t l = 0; l < n; ++l) {
      double su = 0.0, sv = 0.0, sw = 0.0;
      for (int s = 0; s <= 8; ++s) {
        su = su + an[(size_t)(s * 3 + 0) * n + l];
        sv = sv + an[(size_t)(s * 3 + 1) * n + l];
        sw = sw + an[(size_t)(s * 3 + 2) * n + l];
      }
      ul[off[p] + l] = su;
      vl[off[p] + l] = sv;
      wl[off[p] + l] = sw;
    }
    free(an);
  }
}

/* Shared scatter of lagr2eulr (cc = 0) and lagr2eulrind (cc = 1). */
static void scatter(const ibx_grid* g, int cc, int64_t ncap, const int64_t* off,
                    const double* x, const double* y, const double* z, const double* f1,
                    const double* f2, const double* f3, const double* area, double* o1,
                    double* o2, double* o3) {
  const int imax = g->itot, jmax = g->jtot, kmax = g->ktot;
  size_t nf = (size_t)(g->itot + 2) * (size_t)(g->jtot + 2) * (size_t)(g->ktot + 2);
  /* dVeul = dx*dy*dz (:534). */
  double dVeul = g->dx * g->dy * g->dz;
  /* force(:,:,:) = 0 (:811-815). */
  memset(o1, 0, nf * sizeof(double));
  memset(o2, 0, nf * sizeof(double));
  memset(o3, 0, nf * sizeof(double));
  for (int64_t p = 0; p < ncap; ++p) {
    int64_t n = off[p + 1] - off[p];
    for (int nb = 0; nb <= 8; ++nb) {
      int s = nb_to_slot(nb);
      for (int64_t l = 0; l < n; ++l) {
        int64_t q = off[p] + l;
        double xs, ys;
        shifted(g, s, x[q], y[q], &xs, &ys);
        double coorxfp = (xs - g->boundleft) * g->dxi;
        double cooryfp = (ys - g->boundfront) * g->dyi;
        double coorzfp = z[q] * g->dzi;
        long ilow = fnint(coorxfp - 2.0), ihigh = fnint(coorxfp + 2.0);
        long jlow = fnint(cooryfp - 2.0), jhigh = fnint(cooryfp + 2.0);
        long klow = fnint(coorzfp - 2.0), khigh = fnint(coorzfp + 2.0);
        if (!cc) {
          /* :836-851 */
          if (ilow < 1) ilow = 1;
          if (jlow < 1) jlow = 1;
          if (klow < 1) klow = 1;
          if (ihigh > imax) ihigh = imax;
          if (jhigh > jmax) jhigh = jmax;
          if (khigh > kmax) khigh = kmax;
        } else {
          /* :1311-1316 */
          if (ilow < 1) ilow = 0;
          if (jlow < 1) jlow = 0;
          if (klow < 1) klow = 0;
          if (ihigh > imax) ihigh = imax + 1;
          if (jhigh > jmax) jhigh = jmax + 1;
          if (khigh > kmax) khigh = kmax + 1;
        }
        for (long k = klow; k <= khigh; ++k) {
          double coorzs = (1.0 * (double)k) - coorzfp;
          double coorz = coorzs - 0.5;
          double kernelzs = ora_kernel(coorzs);
          double kernelz = ora_kernel(coorz);
          for (long j = jlow; j <= jhigh; ++j) {
            double coorys = (1.0 * (double)j) - cooryfp;
            double coory = coorys - 0.5;
            double kernelys = ora_kernel(coorys);
            double kernely = ora_kernel(coory);
            for (long i = ilow; i <= ihigh; ++i) {
              double coorxs = (1.0 * (double)i) - coorxfp;
              double coorx = coorxs - 0.5;
              double kernelxs = ora_kernel(coorxs);
              double kernelx = ora_kernel(coorx);
              size_t c = IDX(g, i, j, k);
              double adv = area[q] / dVeul;
              if (!cc) {
                /* :869-874 */
                o1[c] = o1[c] + f1[q] * kernelxs * kernely * kernelz * adv;
                o2[c] = o2[c] + f2[q] * kernelx * kernelys * kernelz * adv;
                o3[c] = o3[c] + f3[q] * kernelx * kernely * kernelzs * adv;
              } else {
                /* :1334-1338 */
                o1[c] = o1[c] + f1[q] * kernelx * kernely * kernelz * adv;
                o2[c] = o2[c] + f2[q] * kernelx * kernely * kernelz * adv;
                o3[c] = o3[c] + f3[q] * kernelx * kernely * kernelz * adv;
              }
            }
          }
        }
      }
    }
  }
  /* :904-912 -- x halos of all three fields, then y halos. */
  ora_updthalos(g, o1, 1);
  ora_updthalos(g, o2, 1);
  ora_updthalos(g, o3, 1);
  ora_updthalos(g, o1, 2);
  ora_updthalos(g, o2, 2);
  ora_updthalos(g, o3, 2);
}

void ora_lagr2eulr(const ibx_grid* g, int64_t ncap, const int64_t* off, const double* x,
                   const double* y, const double* z, const double* fx, const double* fy,
                   const double* fz, const double* area, double* forcex, double* forcey,
                   double* forcez) {
  scatter(g, 0, ncap, off, x, y, z, fx, fy, fz, area, forcex, forcey, forcez);
}

void ora_lagr2eulrind(const ibx_grid* g, int64_t ncap, const int64_t* off, const double* x,
                      const double* y, const double* z, const double* xn, const double* yn,
                      const double* zn, const double* area, double* gx, double* gy,
                      double* gz) {
  scatter(g, 1, ncap, off, x, y, z, xn, yn, zn, area, gx, gy, gz);
}

/* Update_Pos.f90:77-186 for one rank (every capsule a master, mslv > 0).  The point count
 * NL of the reference is the capsule's own count here. */
int64_t ora_update_pos(const ibx_grid* g, int64_t ncap, const int64_t* off, double dt,
                       double* x, double* y, double* z, double* xold, double* yold,
                       const double* ul, const double* vl, const double* wl,
                       double* centroid, double* umean) {
  int64_t nclamp = 0;
  double lx = g->lx, ly = g->ly, lz = g->lz;
  for (int64_t p = 0; p < ncap; ++p) {
    int64_t a = off[p], b = off[p + 1];
    /* :80-92 periodicity of the old positions */
    if (xold && yold) {
      for (int64_t q = a; q < b; ++q) {
        if ((xold[q] - x[q]) > (lx * .5))
          xold[q] = xold[q] - lx;
        else if ((xold[q] - x[q]) < -(lx * .5))
          xold[q] = xold[q] + lx;
        if ((yold[q] - y[q]) > (ly * .5))
          yold[q] = yold[q] - ly;
        else if ((yold[q] - y[q]) < -(ly * .5))
          yold[q] = yold[q] + ly;
      }
    }
    /* :95-99 */
    for (int64_t q = a; q < b; ++q) {
      x[q] = x[q] + (ul[q] * dt);
      y[q] = y[q] + (vl[q] * dt);
      z[q] = z[q] + (wl[q] * dt);
    }
    /* :118-126 */
    for (int64_t q = a; q < b; ++q) {
      if (z[q] < 0.0) {
        z[q] = 0.0;
        ++nclamp;
      } else if (z[q] > lz) {
        z[q] = lz;
        ++nclamp;
      }
    }
  }
  /* :132-183 */
  for (int64_t p = 0; p < ncap; ++p) {
    int64_t a = off[p], b = off[p + 1];
    double nl = (double)(b - a);
    double cx = 0.0, cy = 0.0, cz = 0.0, mu = 0.0, mv = 0.0, mw = 0.0;
    for (int64_t q = a; q < b; ++q) {
      cx = cx + x[q];
      cy = cy + y[q];
      cz = cz + z[q];
      mu = mu + ul[q];
      mv = mv + vl[q];
      mw = mw + wl[q];
    }
    cx = cx / nl;
    cy = cy / nl;
    cz = cz / nl;
    mu = mu / nl;
    mv = mv / nl;
    mw = mw / nl;
    if (cx > lx) {
      cx = cx - lx;
      for (int64_t q = a; q < b; ++q) x[q] = x[q] - lx;
    }
    if (cx < 0.0) {
      cx = cx + lx;
      for (int64_t q = a; q < b; ++q) x[q] = x[q] + lx;
    }
    if (cy > ly) {
      cy = cy - ly;
      for (int64_t q = a; q < b; ++q) y[q] = y[q] - ly;
    }
    if (cy < 0.0) {
      cy = cy + ly;
      for (int64_t q = a; q < b; ++q) y[q] = y[q] + ly;
    }
    centroid[3 * p + 0] = cx;
    centroid[3 * p + 1] = cy;
    centroid[3 * p + 2] = cz;
    umean[3 * p + 0] = mu;
    umean[3 * p + 1] = mv;
    umean[3 * p + 2] = mw;
  }
  return nclamp;
}
