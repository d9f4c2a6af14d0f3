/* fluid_oracle.c -- CPU restatement of the reference's flow update (fsolver = "f"), one rank.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of libfluid, never linked by the product.
 *
 *   ora_rfftf / ora_rfftb   VFFTPACK vrfftf / vrfftb (libs/fft/fft.f:693-839, 1-155): one
 *                           real periodic sequence, packed layout r0, re1, im1, ..., r_{N/2}
 *                           and sqrt(1/N) on both directions.  Evaluated here as the plain
 *                           DFT sums of the documented formulas (fft.f:750-770, 60-80) in long
 *                           double, so the transform itself is more accurate than VFFTPACK.
 *   ora_initsolver          initsolver.f90:11-166 (main mesh only; e = 1 pressure, 2 u and v,
 *                           3 w): tridiagonal a, b, c and the eigenvalues xyrt.
 *   ora_solver1d            solver.f90:43-157: x then y real transforms of every k plane,
 *                           the Gaussian elimination in k per mode (:74-129, with the ne == 3
 *                           top row and the z == 0 guard), y then x inverse transforms.
 *   ora_mom_rhs             fftuvw.f90:30-93: momxp/momyp/momzp (mom.f90:137-156,
 *                           264-284, 376-393), momxad/momyad/momzad (mom.f90:9-50, 158-199,
 *                           286-327), minus uo/dt, times Rep, dwdt(:,:,kmax) = 0.
 *   ora_fftuvw              fftuvw.f90:13-127: ora_mom_rhs, solver1d(w, 3), solver1d(u, 2),
 *                           solver1d(v, 2) into unew, vnew, wnew (interior).
 *   ora_bounduvw            bound.f90:7-80 (iniu = 'poi' / 'zer', no free slip): wall
 *                           ghosts in k, then the periodic halos in x and y.
 *   ora_boundp              bound.f90:82-107.
 *   ora_fillps              fillps.f90:8-47.
 *   ora_correc              correc.f90:11-38.
 *   ora_fluid_step          main.f90:388-410: fftuvw, bounduvw, fillps, solver1d(p, 1),
 *                           boundp(p), correc, bounduvw, pnew += p, boundp(pnew).
 *   ora_chkdiv              chkdiv.f90:7-61 (total and maximum |divergence|).
 *
 * Fields are Fortran (0:i1, 0:j1, 0:k1) arrays, i fastest.  Products and sums follow the
 * reference's expressions left to right (compile with -ffp-contract=off); Fortran `x**2.`
 * with a real exponent is pow(x, 2.0) = x*x exactly.  Grid: one rank, imax = itot,
 * jmax = jtot, kmax = ktot (param.f90:22).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/ibx.h"

#define IDX(g, i, j, k) \
  ((size_t)(i) + (size_t)((g)->itot + 2) * ((size_t)(j) + (size_t)((g)->jtot + 2) * (size_t)(k)))

/* ---- VFFTPACK transforms, one sequence ----------------------------------------------- */

void ora_rfftf(int n, double* x) {
  long double* y = (long double*)malloc(sizeof(long double) * (size_t)n);
  const long double pi = 3.141592653589793238462643383279502884L;
  const long double s = sqrtl(1.0L / (long double)n);
  for (int m = 0; m < n; ++m) {
    int k = (m + 1) / 2; /* r0 | re1 im1 | re2 im2 | ... | r_{n/2} */
    long double acc = 0.0L;
    for (int j = 0; j < n; ++j) {
      long double ang = 2.0L * pi * (long double)(((long long)j * k) % n) / (long double)n;
      if (m == 0 || (m % 2 == 1)) acc += (long double)x[j] * cosl(ang);
      else acc -= (long double)x[j] * sinl(ang);
    }
    /* for even n the last slot (m = n-1, odd) is the real part at k = n/2 */
    y[m] = s * acc;
  }
  for (int m = 0; m < n; ++m) x[m] = (double)y[m];
  free(y);
}

void ora_rfftb(int n, double* x) {
  long double* y = (long double*)malloc(sizeof(long double) * (size_t)n);
  const long double pi = 3.141592653589793238462643383279502884L;
  const long double s = sqrtl(1.0L / (long double)n);
  for (int j = 0; j < n; ++j) {
    long double acc = (long double)x[0];
    for (int k = 1; 2 * k < n; ++k) {
      long double ang = 2.0L * pi * (long double)(((long long)j * k) % n) / (long double)n;
      acc += 2.0L * ((long double)x[2 * k - 1] * cosl(ang) - (long double)x[2 * k] * sinl(ang));
    }
    if (n % 2 == 0) acc += ((j % 2) ? -1.0L : 1.0L) * (long double)x[n - 1];
    y[j] = s * acc;
  }
  for (int j = 0; j < n; ++j) x[j] = (double)y[j];
  free(y);
}

/* ---- initsolver.f90:11-166 (main mesh) -------------------------------------------------- */

/* a, b, c: [ktot] (k = 1..kmax at index k-1); xyrt: [itot * jtot], i fastest. */
void ora_initsolver(const ibx_grid* g, double rep, double dt, int e, double* a, double* b,
                    double* c, double* xyrt) {
  const int kmax = g->ktot, itot = g->itot, jtot = g->jtot;
  const double dxi = g->dxi, dyi = g->dyi, dzi = g->dzi;
  for (int k = 0; k < kmax; ++k) {
    a[k] = dzi * dzi;
    c[k] = dzi * dzi;
    b[k] = -(a[k] + c[k]);
  }
  if (e == 0 || e == 1) {
    b[0] = b[0] + a[0];
    b[kmax - 1] = b[kmax - 1] + c[kmax - 1];
    a[0] = 0.0;
    c[kmax - 1] = 0.0;
  } else if (e == 2 || e == 3) {
    if (e == 2) {
      b[0] = b[0] - a[0];
      b[kmax - 1] = b[kmax - 1] - c[kmax - 1];
      a[0] = 0.0;
      c[kmax - 1] = 0.0;
    } else {
      b[kmax - 1] = 1.0;
      a[0] = 0.0;
      a[kmax - 1] = 0.0;
      c[kmax - 1] = 0.0;
    }
    for (int k = 0; k < kmax; ++k) b[k] = b[k] - (rep / dt);
    if (e == 3) b[kmax - 1] = 1.0;
  }
  const double pi = acos(-1.0);
  double* xrt = (double*)malloc(sizeof(double) * (size_t)itot);
  double* yrt = (double*)malloc(sizeof(double) * (size_t)jtot);
  for (int i = 3; i <= itot; i += 2) {
    double s = sin((double)(i - 1) * pi / (2.0 * itot));
    xrt[i - 2] = -(4.0 * dxi * dxi * (s * s));
    xrt[i - 1] = xrt[i - 2];
  }
  xrt[0] = 0.0;
  xrt[itot - 1] = -(4.0 * dxi * dxi);
  for (int j = 3; j <= jtot; j += 2) {
    double s = sin((double)(j - 1) * pi / (2.0 * jtot));
    yrt[j - 2] = -(4.0 * dyi * dyi * (s * s));
    yrt[j - 1] = yrt[j - 2];
  }
  yrt[0] = 0.0;
  yrt[jtot - 1] = -(4.0 * dyi * dyi);
  for (int j = 0; j < jtot; ++j)
    for (int i = 0; i < itot; ++i) xyrt[(size_t)i + (size_t)itot * j] = xrt[i] + yrt[j];
  free(xrt);
  free(yrt);
}

/* ---- solver.f90:43-157 ------------------------------------------------------------------ */

void ora_solver1d(const ibx_grid* g, double rep, double dt, int ne, double* p) {
  const int itot = g->itot, jtot = g->jtot, kmax = g->ktot;
  double* a = (double*)malloc(sizeof(double) * (size_t)kmax);
  double* b = (double*)malloc(sizeof(double) * (size_t)kmax);
  double* c = (double*)malloc(sizeof(double) * (size_t)kmax);
  double* xyrt = (double*)malloc(sizeof(double) * (size_t)itot * jtot);
  ora_initsolver(g, rep, dt, ne, a, b, c, xyrt);
  double* line = (double*)malloc(sizeof(double) * (size_t)(itot > jtot ? itot : jtot));
  double* d = (double*)malloc(sizeof(double) * (size_t)itot * jtot * kmax);
  /* forward: x then y, every k plane (:58-67) */
  for (int k = 1; k <= kmax; ++k) {
    for (int j = 1; j <= jtot; ++j) {
      for (int i = 0; i < itot; ++i) line[i] = p[IDX(g, i + 1, j, k)];
      ora_rfftf(itot, line);
      for (int i = 0; i < itot; ++i) p[IDX(g, i + 1, j, k)] = line[i];
    }
    for (int i = 1; i <= itot; ++i) {
      for (int j = 0; j < jtot; ++j) line[j] = p[IDX(g, i, j + 1, k)];
      ora_rfftf(jtot, line);
      for (int j = 0; j < jtot; ++j) p[IDX(g, i, j + 1, k)] = line[j];
    }
  }
  /* Gaussian elimination per mode (:74-129) */
#define D(i, j, k) d[(size_t)((i) - 1) + (size_t)itot * ((size_t)((j) - 1) + (size_t)jtot * ((k) - 1))]
#define XY(i, j) xyrt[(size_t)((i) - 1) + (size_t)itot * ((j) - 1)]
  for (int j = 1; j <= jtot; ++j)
    for (int i = 1; i <= itot; ++i) {
      double z = 1.0 / (b[0] + XY(i, j));
      D(i, j, 1) = c[0] * z;
      p[IDX(g, i, j, 1)] = p[IDX(g, i, j, 1)] * z;
    }
  for (int k = 2; k <= kmax - 1; ++k)
    for (int j = 1; j <= jtot; ++j)
      for (int i = 1; i <= itot; ++i) {
        double bb = b[k - 1] + XY(i, j);
        double z = 1.0 / (bb - a[k - 1] * D(i, j, k - 1));
        D(i, j, k) = c[k - 1] * z;
        p[IDX(g, i, j, k)] = (p[IDX(g, i, j, k)] - a[k - 1] * p[IDX(g, i, j, k - 1)]) * z;
      }
  for (int j = 1; j <= jtot; ++j)
    for (int i = 1; i <= itot; ++i) {
      double bb = (ne == 3) ? b[kmax - 1] : b[kmax - 1] + XY(i, j);
      double z = bb - a[kmax - 1] * D(i, j, kmax - 1);
      if (z != 0.0)
        p[IDX(g, i, j, kmax)] = (p[IDX(g, i, j, kmax)] - a[kmax - 1] * p[IDX(g, i, j, kmax - 1)]) / z;
      else
        p[IDX(g, i, j, kmax)] = 0.0;
    }
  for (int k = kmax - 1; k >= 1; --k)
    for (int j = 1; j <= jtot; ++j)
      for (int i = 1; i <= itot; ++i)
        p[IDX(g, i, j, k)] = p[IDX(g, i, j, k)] - D(i, j, k) * p[IDX(g, i, j, k + 1)];
#undef D
#undef XY
  /* backward: y then x (:137-146) */
  for (int k = 1; k <= kmax; ++k) {
    for (int i = 1; i <= itot; ++i) {
      for (int j = 0; j < jtot; ++j) line[j] = p[IDX(g, i, j + 1, k)];
      ora_rfftb(jtot, line);
      for (int j = 0; j < jtot; ++j) p[IDX(g, i, j + 1, k)] = line[j];
    }
    for (int j = 1; j <= jtot; ++j) {
      for (int i = 0; i < itot; ++i) line[i] = p[IDX(g, i + 1, j, k)];
      ora_rfftb(itot, line);
      for (int i = 0; i < itot; ++i) p[IDX(g, i + 1, j, k)] = line[i];
    }
  }
  free(a), free(b), free(c), free(xyrt), free(line), free(d);
}

/* ---- the momentum right-hand side, fftuvw.f90:30-93 ------------------------------------- */

/* mom.f90:9-50 (momxad) at one cell */
static double momxad(const ibx_grid* g, const double* Uo, const double* Vo, const double* Wo,
                     const double* Uoo, const double* Voo, const double* Woo, int i, int j, int k) {
  const int ip = i + 1, jp = j + 1, kp = k + 1, im = i - 1, jm = j - 1, km = k - 1;
#define F(A, a, b, c) A[IDX(g, a, b, c)]
  double uuip = 0.25 * (F(Uo, ip, j, k) + F(Uo, i, j, k)) * (F(Uo, ip, j, k) + F(Uo, i, j, k));
  double uuim = 0.25 * (F(Uo, im, j, k) + F(Uo, i, j, k)) * (F(Uo, im, j, k) + F(Uo, i, j, k));
  double uvjp = 0.25 * (F(Uo, i, jp, k) + F(Uo, i, j, k)) * (F(Vo, ip, j, k) + F(Vo, i, j, k));
  double uvjm = 0.25 * (F(Uo, i, jm, k) + F(Uo, i, j, k)) * (F(Vo, ip, jm, k) + F(Vo, i, jm, k));
  double uwkp = 0.25 * (F(Uo, i, j, kp) + F(Uo, i, j, k)) * (F(Wo, ip, j, k) + F(Wo, i, j, k));
  double uwkm = 0.25 * (F(Uo, i, j, km) + F(Uo, i, j, k)) * (F(Wo, ip, j, km) + F(Wo, i, j, km));
  double uuipo = 0.25 * (F(Uoo, ip, j, k) + F(Uoo, i, j, k)) * (F(Uoo, ip, j, k) + F(Uoo, i, j, k));
  double uuimo = 0.25 * (F(Uoo, im, j, k) + F(Uoo, i, j, k)) * (F(Uoo, im, j, k) + F(Uoo, i, j, k));
  double uvjpo = 0.25 * (F(Uoo, i, jp, k) + F(Uoo, i, j, k)) * (F(Voo, ip, j, k) + F(Voo, i, j, k));
  double uvjmo = 0.25 * (F(Uoo, i, jm, k) + F(Uoo, i, j, k)) * (F(Voo, ip, jm, k) + F(Voo, i, jm, k));
  double uwkpo = 0.25 * (F(Uoo, i, j, kp) + F(Uoo, i, j, k)) * (F(Woo, ip, j, k) + F(Woo, i, j, k));
  double uwkmo = 0.25 * (F(Uoo, i, j, km) + F(Uoo, i, j, k)) * (F(Woo, ip, j, km) + F(Woo, i, j, km));
  return g->dxi * ((1.5 * (-uuip + uuim)) - (.5 * (-uuipo + uuimo))) +
         g->dyi * ((1.5 * (-uvjp + uvjm)) - (.5 * (-uvjpo + uvjmo))) +
         g->dzi * ((1.5 * (-uwkp + uwkm)) - (.5 * (-uwkpo + uwkmo)));
}

/* mom.f90:158-199 (momyad) */
static double momyad(const ibx_grid* g, const double* Uo, const double* Vo, const double* Wo,
                     const double* Uoo, const double* Voo, const double* Woo, int i, int j, int k) {
  const int ip = i + 1, jp = j + 1, kp = k + 1, im = i - 1, jm = j - 1, km = k - 1;
  double uvip = 0.25 * (F(Uo, i, j, k) + F(Uo, i, jp, k)) * (F(Vo, i, j, k) + F(Vo, ip, j, k));
  double uvim = 0.25 * (F(Uo, im, j, k) + F(Uo, im, jp, k)) * (F(Vo, i, j, k) + F(Vo, im, j, k));
  double vvjp = 0.25 * (F(Vo, i, j, k) + F(Vo, i, jp, k)) * (F(Vo, i, j, k) + F(Vo, i, jp, k));
  double vvjm = 0.25 * (F(Vo, i, j, k) + F(Vo, i, jm, k)) * (F(Vo, i, j, k) + F(Vo, i, jm, k));
  double wvkp = 0.25 * (F(Wo, i, j, k) + F(Wo, i, jp, k)) * (F(Vo, i, j, kp) + F(Vo, i, j, k));
  double wvkm = 0.25 * (F(Wo, i, j, km) + F(Wo, i, jp, km)) * (F(Vo, i, j, km) + F(Vo, i, j, k));
  double uvipo = 0.25 * (F(Uoo, i, j, k) + F(Uoo, i, jp, k)) * (F(Voo, i, j, k) + F(Voo, ip, j, k));
  double uvimo = 0.25 * (F(Uoo, im, j, k) + F(Uoo, im, jp, k)) * (F(Voo, i, j, k) + F(Voo, im, j, k));
  double vvjpo = 0.25 * (F(Voo, i, j, k) + F(Voo, i, jp, k)) * (F(Voo, i, j, k) + F(Voo, i, jp, k));
  double vvjmo = 0.25 * (F(Voo, i, j, k) + F(Voo, i, jm, k)) * (F(Voo, i, j, k) + F(Voo, i, jm, k));
  double wvkpo = 0.25 * (F(Woo, i, j, k) + F(Woo, i, jp, k)) * (F(Voo, i, j, kp) + F(Voo, i, j, k));
  double wvkmo = 0.25 * (F(Woo, i, j, km) + F(Woo, i, jp, km)) * (F(Voo, i, j, km) + F(Voo, i, j, k));
  return g->dxi * ((1.5 * (-uvip + uvim)) - (.5 * (-uvipo + uvimo))) +
         g->dyi * ((1.5 * (-vvjp + vvjm)) - (.5 * (-vvjpo + vvjmo))) +
         g->dzi * ((1.5 * (-wvkp + wvkm)) - (.5 * (-wvkpo + wvkmo)));
}

/* mom.f90:286-327 (momzad) */
static double momzad(const ibx_grid* g, const double* Uo, const double* Vo, const double* Wo,
                     const double* Uoo, const double* Voo, const double* Woo, int i, int j, int k) {
  const int ip = i + 1, jp = j + 1, kp = k + 1, im = i - 1, jm = j - 1, km = k - 1;
  double uwip = 0.25 * (F(Wo, i, j, k) + F(Wo, ip, j, k)) * (F(Uo, i, j, k) + F(Uo, i, j, kp));
  double uwim = 0.25 * (F(Wo, i, j, k) + F(Wo, im, j, k)) * (F(Uo, im, j, k) + F(Uo, im, j, kp));
  double vwjp = 0.25 * (F(Wo, i, j, k) + F(Wo, i, jp, k)) * (F(Vo, i, j, k) + F(Vo, i, j, kp));
  double vwjm = 0.25 * (F(Wo, i, j, k) + F(Wo, i, jm, k)) * (F(Vo, i, jm, k) + F(Vo, i, jm, kp));
  double wwkp = 0.25 * (F(Wo, i, j, k) + F(Wo, i, j, kp)) * (F(Wo, i, j, k) + F(Wo, i, j, kp));
  double wwkm = 0.25 * (F(Wo, i, j, k) + F(Wo, i, j, km)) * (F(Wo, i, j, k) + F(Wo, i, j, km));
  double uwipo = 0.25 * (F(Woo, i, j, k) + F(Woo, ip, j, k)) * (F(Uoo, i, j, k) + F(Uoo, i, j, kp));
  double uwimo = 0.25 * (F(Woo, i, j, k) + F(Woo, im, j, k)) * (F(Uoo, im, j, k) + F(Uoo, im, j, kp));
  double vwjpo = 0.25 * (F(Woo, i, j, k) + F(Woo, i, jp, k)) * (F(Voo, i, j, k) + F(Voo, i, j, kp));
  double vwjmo = 0.25 * (F(Woo, i, j, k) + F(Woo, i, jm, k)) * (F(Voo, i, jm, k) + F(Voo, i, jm, kp));
  double wwkpo = 0.25 * (F(Woo, i, j, k) + F(Woo, i, j, kp)) * (F(Woo, i, j, k) + F(Woo, i, j, kp));
  double wwkmo = 0.25 * (F(Woo, i, j, k) + F(Woo, i, j, km)) * (F(Woo, i, j, k) + F(Woo, i, j, km));
  return g->dxi * ((1.5 * (-uwip + uwim)) - (.5 * (-uwipo + uwimo))) +
         g->dyi * ((1.5 * (-vwjp + vwjm)) - (.5 * (-vwjpo + vwjmo))) +
         g->dzi * ((1.5 * (-wwkp + wwkm)) - (.5 * (-wwkpo + wwkmo)));
}

/* dudt, dvdt, dwdt (interior written; halos untouched) of fftuvw.f90:30-93:
 *   d*dt = ( -(momp) - (momad) - (o/dt) ) * Rep, with dwdt(:,:,kmax) = 0 (:97-105). */
void ora_mom_rhs(const ibx_grid* g, double rep, double dt, const double* uo, const double* vo,
                 const double* wo, const double* uoo, const double* voo, const double* woo,
                 const double* p, const double* fx, const double* fy, const double* fz,
                 double* dudt, double* dvdt, double* dwdt) {
  for (int k = 1; k <= g->ktot; ++k)
    for (int j = 1; j <= g->jtot; ++j)
      for (int i = 1; i <= g->itot; ++i) {
        const size_t q = IDX(g, i, j, k);
        /* momxp, momyp, momzp (mom.f90:150,279,388) */
        double durk = (-g->dxi * ((F(p, i + 1, j, k) - F(p, i, j, k)))) + fx[q];
        double dvrk = (-g->dyi * (F(p, i, j + 1, k) - F(p, i, j, k))) + fy[q];
        double dwrk = (-g->dzi * (F(p, i, j, k + 1) - F(p, i, j, k))) + fz[q];
        double du = -durk, dv = -dvrk, dw = -dwrk;
        du = du - momxad(g, uo, vo, wo, uoo, voo, woo, i, j, k);
        dv = dv - momyad(g, uo, vo, wo, uoo, voo, woo, i, j, k);
        dw = dw - momzad(g, uo, vo, wo, uoo, voo, woo, i, j, k);
        du = du - (uo[q] / dt);
        dv = dv - (vo[q] / dt);
        dw = dw - (wo[q] / dt);
        dudt[q] = du * rep;
        dvdt[q] = dv * rep;
        dwdt[q] = (k == g->ktot) ? 0.0 : dw * rep;
      }
}
#undef F

static double* field_alloc(const ibx_grid* g) {
  return (double*)calloc((size_t)(g->itot + 2) * (g->jtot + 2) * (g->ktot + 2), sizeof(double));
}

/* fftuvw.f90:13-127: unew, vnew, wnew interior from the old fields, pressure and force */
void ora_fftuvw(const ibx_grid* g, double rep, double dt, double* unew, double* vnew,
                double* wnew, const double* uo, const double* vo, const double* wo,
                const double* uoo, const double* voo, const double* woo, const double* pnew,
                const double* fx, const double* fy, const double* fz) {
  double* du = field_alloc(g);
  double* dv = field_alloc(g);
  double* dw = field_alloc(g);
  ora_mom_rhs(g, rep, dt, uo, vo, wo, uoo, voo, woo, pnew, fx, fy, fz, du, dv, dw);
  ora_solver1d(g, rep, dt, 3, dw);
  ora_solver1d(g, rep, dt, 2, du);
  ora_solver1d(g, rep, dt, 2, dv);
  for (int k = 1; k <= g->ktot; ++k)
    for (int j = 1; j <= g->jtot; ++j)
      for (int i = 1; i <= g->itot; ++i) {
        const size_t q = IDX(g, i, j, k);
        unew[q] = du[q], vnew[q] = dv[q], wnew[q] = dw[q];
      }
  free(du), free(dv), free(dw);
}

void ora_updthalos(const ibx_grid* g, double* var, int dir); /* ibx_oracle.c */

/* bound.f90:7-80, iniu = 'poi' or 'zer' without free slip, then the periodic halos */
void ora_bounduvw(const ibx_grid* g, double* u, double* v, double* w) {
  const int kmax = g->ktot, k1 = g->ktot + 1;
  for (int j = 0; j <= g->jtot + 1; ++j)
    for (int i = 0; i <= g->itot + 1; ++i) {
      u[IDX(g, i, j, 0)] = -u[IDX(g, i, j, 1)];
      v[IDX(g, i, j, 0)] = -v[IDX(g, i, j, 1)];
      w[IDX(g, i, j, 0)] = 0.0;
      u[IDX(g, i, j, k1)] = -u[IDX(g, i, j, kmax)];
      v[IDX(g, i, j, k1)] = -v[IDX(g, i, j, kmax)];
      w[IDX(g, i, j, kmax)] = 0.0;
      w[IDX(g, i, j, k1)] = w[IDX(g, i, j, kmax - 1)];
    }
  ora_updthalos(g, u, 1);
  ora_updthalos(g, v, 1);
  ora_updthalos(g, w, 1);
  ora_updthalos(g, u, 2);
  ora_updthalos(g, v, 2);
  ora_updthalos(g, w, 2);
}

/* bound.f90:82-107 */
void ora_boundp(const ibx_grid* g, double* p) {
  for (int j = 0; j <= g->jtot + 1; ++j)
    for (int i = 0; i <= g->itot + 1; ++i) {
      p[IDX(g, i, j, 0)] = p[IDX(g, i, j, 1)];
      p[IDX(g, i, j, g->ktot + 1)] = p[IDX(g, i, j, g->ktot)];
    }
  ora_updthalos(g, p, 1);
  ora_updthalos(g, p, 2);
}

/* fillps.f90:8-47 */
void ora_fillps(const ibx_grid* g, double dt, const double* u, const double* v, const double* w,
                double* p) {
  const double dti = 1.0 / dt;
  const double dtidxi = dti * g->dxi, dtidyi = dti * g->dyi, dtidzi = dti * g->dzi;
  for (int k = 1; k <= g->ktot; ++k)
    for (int j = 1; j <= g->jtot; ++j)
      for (int i = 1; i <= g->itot; ++i)
        p[IDX(g, i, j, k)] = ((w[IDX(g, i, j, k)] - w[IDX(g, i, j, k - 1)]) * dtidzi +
                              (v[IDX(g, i, j, k)] - v[IDX(g, i, j - 1, k)]) * dtidyi +
                              (u[IDX(g, i, j, k)] - u[IDX(g, i - 1, j, k)]) * dtidxi);
}

/* correc.f90:11-38 */
void ora_correc(const ibx_grid* g, double dt, double* u, double* v, double* w, const double* p) {
  const double factori = dt * g->dxi, factorj = dt * g->dyi, factork = dt * g->dzi;
  for (int k = 1; k <= g->ktot; ++k)
    for (int j = 1; j <= g->jtot; ++j)
      for (int i = 1; i <= g->itot; ++i) {
        const size_t q = IDX(g, i, j, k);
        u[q] = u[q] - factori * (p[IDX(g, i + 1, j, k)] - p[q]);
        v[q] = v[q] - factorj * (p[IDX(g, i, j + 1, k)] - p[q]);
        w[q] = w[q] - factork * (p[IDX(g, i, j, k + 1)] - p[q]);
      }
}

/* main.f90:388-410 (fsolver = "f", no overlap mesh).  unew, vnew, wnew and pnew are updated
 * in place; uo.. and uoo.. are the previous two velocity fields (main.f90:197-205 rotates
 * them before the step); p is the correction-pressure work field (out). */
void ora_fluid_step(const ibx_grid* g, double rep, double dt, double* unew, double* vnew,
                    double* wnew, double* pnew, double* p, const double* uo, const double* vo,
                    const double* wo, const double* uoo, const double* voo, const double* woo,
                    const double* fx, const double* fy, const double* fz) {
  ora_fftuvw(g, rep, dt, unew, vnew, wnew, uo, vo, wo, uoo, voo, woo, pnew, fx, fy, fz);
  ora_bounduvw(g, unew, vnew, wnew);
  ora_fillps(g, dt, unew, vnew, wnew, p);
  ora_solver1d(g, rep, dt, 1, p);
  ora_boundp(g, p);
  ora_correc(g, dt, unew, vnew, wnew, p);
  ora_bounduvw(g, unew, vnew, wnew);
  const size_t n = (size_t)(g->itot + 2) * (g->jtot + 2) * (g->ktot + 2);
  for (size_t q = 0; q < n; ++q) pnew[q] = pnew[q] + p[q];
  ora_boundp(g, pnew);
}

/* chkdiv.f90:7-61: divtot = sum of the divergence, divmax = max |divergence| */
void ora_chkdiv(const ibx_grid* g, const double* u, const double* v, const double* w,
                double* divtot, double* divmax) {
  double tot = 0.0, mx = 0.0;
  for (int k = 1; k <= g->ktot; ++k)
    for (int j = 1; j <= g->jtot; ++j)
      for (int i = 1; i <= g->itot; ++i) {
        double div = (w[IDX(g, i, j, k)] - w[IDX(g, i, j, k - 1)]) * g->dzi +
                     (v[IDX(g, i, j, k)] - v[IDX(g, i, j - 1, k)]) * g->dyi +
                     (u[IDX(g, i, j, k)] - u[IDX(g, i - 1, j, k)]) * g->dxi;
        tot = tot + div;
        if (fabs(div) > mx) mx = fabs(div);
      }
  *divtot = tot;
  *divmax = mx;
}
