/*
 * membrane_oracle.c -- CPU restatement of the reference's FP64 capsule membrane law.
 *
 * TEST INFRASTRUCTURE ONLY.  Like ibx_oracle.c this is the checker: tests/, the smoke test
 * and the CPU-baseline leg of bench.py may load it; the product never does.
 *
 * What it restates (one capsule, neo-Hookean shear + bending, dealiased):
 *   complagrforces        forcing.f90:11-114 (with the force grid equal to the spherical-
 *                         harmonic grid, TRSSPH is the identity)
 *   cal_spectral_force    sph.f90:1523-1755
 *   shape coefficients    sph1.f90:29-79 (SPHEREPACK shagc on a Gauss grid)
 *   x and its derivatives sph3.f90:9-182 (up to 4th order, evaluated on the 2x grid)
 *   normal, bases         sph1.f90:157-296, 581-655
 *   curvature             sph.f90:442-733
 *   Christoffel symbols   sph.f90:735-895
 *   metrics               sph.f90:896-1113
 *   invariants, tension   sph.f90:1115-1235; membmodel.c:60-79 (-DNONHOOK, Makefile:7-8)
 *   bending moment, Q     sph.f90:1236-1455
 *   dealias filter        sph1.f90:723-951 (analysis on the 2x grid, synthesis of the
 *                         degrees < nlat on the force grid), cal_unitO_filter
 *                         sph1.f90:1378-1406, surface metric sph3.f90:186-260
 *   reference state       capsule_init_cal_coeff sph.f90:16-191 (unit sphere)
 *
 * The reference evaluates the same continuum quantities (metric, curvature and their first
 * and second derivatives, Christoffel symbols and their derivatives) from analytic
 * derivatives of the spherical-harmonic expansion of the surface; so does this file, with
 * its own algebra (inverse-metric derivatives through the matrix identity
 * d(A^-1) = -A^-1 dA A^-1 instead of derivatives of cross products).  The two agree to
 * rounding, not bit for bit.  The reference itself cannot be compiled here (it needs GSL,
 * absent from the image), so this restatement is pinned by closed forms: the undeformed
 * sphere gives zero traction, a uniformly dilated sphere gives the Laplace-law pressure of
 * the neo-Hookean tension, and the net force and torque of any shape vanish.
 *
 * Spherical-harmonic conventions (SPHEREPACK, shagc/shsgc documentation):
 *   f(theta,phi) = sum_{m>=0} sum_{n>=m} Pbar_n^m(cos theta) (a_mn cos m phi - b_mn sin m phi),
 * with the m = 0 terms halved, Pbar_n^m = sqrt((2n+1)/2 (n-m)!/(n+m)!) P_n^m without the
 * Condon-Shortley phase (sphharm.c:38-53 multiplies GSL's sphPlm by sqrt(2 pi) (-1)^m).
 * Colatitudes are Gauss-Legendre points in increasing order (gaqd), longitudes
 * phi_j = 2 pi (j-1)/nlon, and point l = i*nlon + j (latitude-major, forcing.f90:37-43).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MEM_LMAX 64

typedef struct {
  int L, nt2, np2, nt, np;
  double th2[MEM_LMAX], w2[MEM_LMAX];          /* coarse (force / harmonic) grid */
  double th[2 * MEM_LMAX], w[2 * MEM_LMAX];    /* fine (dealias) grid */
  double* Pc;   /* [nt2][m][n] Pbar at the coarse colatitudes, degrees < nt2 */
  double* Pcd;  /* [nt2][m][n] d Pbar / d theta at the coarse colatitudes */
  double* Pfd;  /* [nt][k=0..4][m][n] theta-derivatives at the fine colatitudes, n < nt2 */
  double* Pfa;  /* [nt][m][n] Pbar at the fine colatitudes, degrees < nt (analysis) */
  /* reference state on the fine grid: contravariant metric and curvature with their first
   * and second derivatives, per node */
  double* AR;   /* [node][2][2] */
  double* ARd;  /* [node][2 (d)][2][2] */
  double* ARdd; /* [node][2][2 (d)][2][2] */
  double* BR;
  double* BRd;
  double* BRdd;
} mem_plan;

static mem_plan g_plan;
static int g_plan_L = -1;

/* ---------------------------------------------------------------------------------------
 * Gauss-Legendre colatitudes (increasing) and weights for the integral over theta of
 * f sin(theta): Newton iteration on P_n(x), x = cos theta.
 * ------------------------------------------------------------------------------------- */
void mem_gauss(int n, double* theta, double* w) {
  for (int i = 0; i < n; ++i) {
    double x = cos(M_PI * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = x;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1, p1 = p2;
      }
      double pn = n == 0 ? 1.0 : p1, pnm1 = n == 0 ? 0.0 : p0;
      if (n == 1) pn = x, pnm1 = 1.0;
      dp = n * (x * pn - pnm1) / (x * x - 1.0);
      double dx = pn / dp;
      x -= dx;
      if (fabs(dx) < 1e-16) {
        /* one more evaluation of the derivative at the converged node */
        p0 = 1.0, p1 = x;
        for (int k = 2; k <= n; ++k) {
          double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
          p0 = p1, p1 = p2;
        }
        pn = n == 1 ? x : p1, pnm1 = n == 1 ? 1.0 : p0;
        dp = n * (x * pn - pnm1) / (x * x - 1.0);
        break;
      }
    }
    theta[i] = acos(x);
    w[i] = 2.0 / ((1.0 - x * x) * dp * dp);
  }
}

/* Pbar_n^m(cos theta) for 0 <= m <= n < nmax, P[m * nmax + n] (zero for n < m). */
void mem_legendre(int nmax, double theta, double* P) {
  const double c = cos(theta), s = sin(theta);
  memset(P, 0, sizeof(double) * nmax * nmax);
  double pmm = sqrt(0.5);
  for (int m = 0; m < nmax; ++m) {
    if (m > 0) pmm = pmm * sqrt((2.0 * m + 1.0) / (2.0 * m)) * s;
    P[m * nmax + m] = pmm;
    if (m + 1 < nmax) P[m * nmax + m + 1] = sqrt(2.0 * m + 3.0) * c * pmm;
    for (int n = m + 2; n < nmax; ++n) {
      const double anm = sqrt((4.0 * n * n - 1.0) / ((double)n * n - (double)m * m));
      const double anm1 =
          sqrt((4.0 * (n - 1.0) * (n - 1.0) - 1.0) / ((n - 1.0) * (n - 1.0) - (double)m * m));
      P[m * nmax + n] = anm * (c * P[m * nmax + n - 1] - P[m * nmax + n - 2] / anm1);
    }
  }
}

/* theta-derivatives of order 0..4 of Pbar_n^m, n < nmax, D[(k * nmax + m) * nmax + n], by the
 * ladder identity d/dtheta Pbar^m = (c-_m Pbar^{m-1} - c+_m Pbar^{m+1}) / 2 with
 * c-_m = sqrt((n+m)(n-m+1)), c+_m = sqrt((n+m+1)(n-m)) and Pbar^{-m} = (-1)^m Pbar^m. */
void mem_legendre_derivs(int nmax, double theta, double* D) {
  double* P = (double*)malloc(sizeof(double) * nmax * nmax);
  mem_legendre(nmax, theta, P);
  memset(D, 0, sizeof(double) * 5 * nmax * nmax);
  double v[5][2 * MEM_LMAX + 3];
  for (int n = 0; n < nmax; ++n) {
    const int off = n + 1;  /* v[k][m + off], m in [-n-1, n+1] */
    for (int k = 0; k < 5; ++k)
      for (int q = 0; q < 2 * n + 3; ++q) v[k][q] = 0.0;
    for (int m = -n; m <= n; ++m) {
      const int am = m < 0 ? -m : m;
      v[0][m + off] = ((m < 0 && (am & 1)) ? -1.0 : 1.0) * P[am * nmax + n];
    }
    for (int k = 1; k < 5; ++k)
      for (int m = -n; m <= n; ++m) {
        const double cm = sqrt((double)(n + m) * (double)(n - m + 1));
        const double cp = sqrt((double)(n + m + 1) * (double)(n - m));
        v[k][m + off] = 0.5 * (cm * v[k - 1][m - 1 + off] - cp * v[k - 1][m + 1 + off]);
      }
    for (int k = 0; k < 5; ++k)
      for (int m = 0; m <= n; ++m) D[(k * nmax + m) * nmax + n] = v[k][m + off];
  }
  free(P);
}

/* cos / sin(m phi_j), phi_j = 2 pi j / n, with the angle reduced first to 2 pi ((m j) mod n) / n:
 * SPHEREPACK's longitude transforms are real FFTs whose twiddle factors are cos / sin of
 * angles below 2 pi (hrfft.f:214-267, hrfti1), never of m phi_j itself, whose rounding
 * (1e-14 at m phi_j ~ 2 pi L) the fourth phi derivatives multiply by m^4. */
static inline double cos_mj(int m, int j, int n) { return cos(2.0 * M_PI * ((m * j) % n) / n); }
static inline double sin_mj(int m, int j, int n) { return sin(2.0 * M_PI * ((m * j) % n) / n); }

/* Analysis on an nt x np Gauss grid (f[i * np + j]) with the Pbar table P[i][m][n]
 * (degrees < nmax <= nt): a[m * nmax + n], b[m * nmax + n]. */
static void analysis(int nt, int np, const double* w, const double* P, int nmax, const double* f,
                     double* a, double* b) {
  memset(a, 0, sizeof(double) * nmax * nmax);
  memset(b, 0, sizeof(double) * nmax * nmax);
  for (int i = 0; i < nt; ++i)
    for (int m = 0; m < nmax; ++m) {
      double cs = 0.0, sn = 0.0;
      for (int j = 0; j < np; ++j) {
        cs += f[i * np + j] * cos_mj(m, j, np);
        sn += f[i * np + j] * sin_mj(m, j, np);
      }
      for (int n = m; n < nmax; ++n) {
        const double p = w[i] * P[((size_t)i * nmax + m) * nmax + n];
        a[m * nmax + n] += 2.0 / np * p * cs;
        b[m * nmax + n] -= 2.0 / np * p * sn;
      }
    }
}

/* Synthesis of degrees < nmax on an nt x np grid with the table P[i][m][n]. */
static void synthesis(int nt, int np, const double* P, int nmax, const double* a, const double* b,
                      double* f) {
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < np; ++j) {
      double s = 0.0;
      for (int m = 0; m < nmax; ++m) {
        const double cm = cos_mj(m, j, np), sm = sin_mj(m, j, np);
        for (int n = m; n < nmax; ++n) {
          const double p = P[((size_t)i * nmax + m) * nmax + n] * (m == 0 ? 0.5 : 1.0);
          s += p * (a[m * nmax + n] * cm - b[m * nmax + n] * sm);
        }
      }
      f[i * np + j] = s;
    }
}

/* x and all its partial derivatives of total order <= 4 at the fine nodes, from the coarse
 * coefficients (sph3.f90:9-182).  out[node][c][q] with q = index of (k_theta, k_phi):
 * order s = k_theta + k_phi, q = s(s+1)/2 + k_phi. */
#define NQ 15
static inline int qidx(int kt, int kp) {
  const int s = kt + kp;
  return s * (s + 1) / 2 + kp;
}

static void derivs_fine(const mem_plan* p, const double* a3, const double* b3, double* out) {
  const int N = p->nt2;
  double A[5][MEM_LMAX], B[5][MEM_LMAX];
  for (int i = 0; i < p->nt; ++i)
    for (int c = 0; c < 3; ++c) {
      const double* a = a3 + c * N * N;
      const double* b = b3 + c * N * N;
      for (int k = 0; k < 5; ++k)
        for (int m = 0; m < N; ++m) {
          double sa = 0.0, sb = 0.0;
          for (int n = m; n < N; ++n) {
            const double d = p->Pfd[(((size_t)i * 5 + k) * N + m) * N + n] * (m == 0 ? 0.5 : 1.0);
            sa += d * a[m * N + n];
            sb += d * b[m * N + n];
          }
          A[k][m] = sa, B[k][m] = sb;
        }
      for (int j = 0; j < p->np; ++j) {
        double* o = out + ((size_t)(i * p->np + j) * 3 + c) * NQ;
        for (int q = 0; q < NQ; ++q) o[q] = 0.0;
        for (int m = 0; m < N; ++m) {
          const double cm = cos_mj(m, j, p->np), sm = sin_mj(m, j, p->np);
          /* k_phi-th phi derivative of (A cos m phi - B sin m phi) */
          for (int kt = 0; kt <= 4; ++kt)
            for (int kp = 0; kt + kp <= 4; ++kp) {
              const double mk = pow((double)m, kp);
              double h;
              switch (kp & 3) {
                case 0: h = A[kt][m] * cm - B[kt][m] * sm; break;
                case 1: h = -A[kt][m] * sm - B[kt][m] * cm; break;
                case 2: h = -(A[kt][m] * cm - B[kt][m] * sm); break;
                default: h = A[kt][m] * sm + B[kt][m] * cm; break;
              }
              o[qidx(kt, kp)] += mk * h;
            }
        }
      }
    }
}

/* ---- small vector helpers ------------------------------------------------------------- */
static inline double dot3(const double* u, const double* v) {
  return u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
}
static inline void cross3(const double* u, const double* v, double* r) {
  r[0] = u[1] * v[2] - u[2] * v[1];
  r[1] = u[2] * v[0] - u[0] * v[2];
  r[2] = u[0] * v[1] - u[1] * v[0];
}
static inline void inv2(const double g[2][2], double G[2][2]) {
  const double det = g[0][0] * g[1][1] - g[0][1] * g[1][0];
  G[0][0] = g[1][1] / det, G[1][1] = g[0][0] / det;
  G[0][1] = -g[0][1] / det, G[1][0] = -g[1][0] / det;
}
static inline void mul2(const double X[2][2], const double Y[2][2], double R[2][2]) {
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) R[i][j] = X[i][0] * Y[0][j] + X[i][1] * Y[1][j];
}

/* Geometry of one node from the derivatives X[c][q]. */
typedef struct {
  double a[2][3];          /* covariant base vectors x_theta, x_phi */
  double n[3], J;          /* unit normal (a1 x a2)/|a1 x a2| (sph1.f90:611-655), |a1 x a2| */
  double g[2][2];          /* covariant metric a_ab */
  double gd[2][2][2];      /* [c][a][b] d_c a_ab */
  double G[2][2];          /* contravariant metric a^ab */
  double Gd[2][2][2];      /* [c][a][b] d_c a^ab */
  double Gdd[2][2][2][2];  /* [c][d][a][b] d_c d_d a^ab */
  double b[2][2];          /* curvature b_ab = x_ab . n */
  double bd[2][2][2];      /* [c][a][b] */
  double bdd[2][2][2][2];  /* [c][d][a][b] */
  double Gam[2][2][2];     /* [b][a][c] Gamma^b_ac = x_ac . a^b */
  double Gamd[2][2][2][2]; /* [l][b][a][c] d_l Gamma^b_ac */
} node_geom;

static void node_geometry(const double* X, node_geom* o) {
  /* X[c * NQ + q]; derivative with s phi-indices among `ord` indices */
#define XV(v, kt, kp) \
  do {                \
    v[0] = X[0 * NQ + qidx(kt, kp)]; v[1] = X[1 * NQ + qidx(kt, kp)]; v[2] = X[2 * NQ + qidx(kt, kp)]; \
  } while (0)
  double x1[2][3], x2[2][2][3], x3[2][2][2][3], x4[2][2][2][2][3];
  for (int a = 0; a < 2; ++a) {
    XV(x1[a], 1 - a, a);
    for (int b = 0; b < 2; ++b) {
      XV(x2[a][b], 2 - a - b, a + b);
      for (int c = 0; c < 2; ++c) {
        XV(x3[a][b][c], 3 - a - b - c, a + b + c);
        for (int d = 0; d < 2; ++d) XV(x4[a][b][c][d], 4 - a - b - c - d, a + b + c + d);
      }
    }
  }
#undef XV
  memcpy(o->a, x1, sizeof x1);
  double cvec[3], cd[2][3], nd[2][3], Jd[2];
  cross3(x1[0], x1[1], cvec);
  o->J = sqrt(dot3(cvec, cvec));
  for (int k = 0; k < 3; ++k) o->n[k] = cvec[k] / o->J;
  for (int c = 0; c < 2; ++c) {
    double t1[3], t2[3];
    cross3(x2[0][c], x1[1], t1);
    cross3(x1[0], x2[1][c], t2);
    for (int k = 0; k < 3; ++k) cd[c][k] = t1[k] + t2[k];
    Jd[c] = dot3(o->n, cd[c]);
    for (int k = 0; k < 3; ++k) nd[c][k] = (cd[c][k] - Jd[c] * o->n[k]) / o->J;
  }
  /* covariant metric and derivatives */
  double gdd[2][2][2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      o->g[a][b] = dot3(x1[a], x1[b]);
      for (int c = 0; c < 2; ++c) {
        o->gd[c][a][b] = dot3(x2[a][c], x1[b]) + dot3(x1[a], x2[b][c]);
        for (int d = 0; d < 2; ++d)
          gdd[c][d][a][b] = dot3(x3[a][c][d], x1[b]) + dot3(x2[a][c], x2[b][d]) +
                            dot3(x2[a][d], x2[b][c]) + dot3(x1[a], x3[b][c][d]);
      }
    }
  /* contravariant metric: G = g^-1, dG = -G dg G, ddG = G dg_d G dg_c G + G dg_c G dg_d G - G ddg G */
  inv2(o->g, o->G);
  double GdgG[2][2][2];
  for (int c = 0; c < 2; ++c) {
    double t[2][2], u[2][2];
    mul2(o->G, o->gd[c], t);
    mul2(t, o->G, u);
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) GdgG[c][a][b] = u[a][b], o->Gd[c][a][b] = -u[a][b];
  }
  for (int c = 0; c < 2; ++c)
    for (int d = 0; d < 2; ++d) {
      double t1[2][2], t2[2][2], t3[2][2], t4[2][2], t5[2][2];
      mul2(GdgG[d], o->gd[c], t1);
      mul2(t1, o->G, t2); /* G dg_d G dg_c G */
      mul2(GdgG[c], o->gd[d], t3);
      mul2(t3, o->G, t4); /* G dg_c G dg_d G */
      double gcd[2][2];
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) gcd[a][b] = gdd[c][d][a][b];
      mul2(o->G, gcd, t1);
      mul2(t1, o->G, t5);
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) o->Gdd[c][d][a][b] = t2[a][b] + t4[a][b] - t5[a][b];
    }
  /* contravariant base vectors a^a = G^ab a_b and their derivatives */
  double ac[2][3], acd[2][2][3];
  for (int a = 0; a < 2; ++a)
    for (int k = 0; k < 3; ++k) {
      ac[a][k] = o->G[a][0] * x1[0][k] + o->G[a][1] * x1[1][k];
      for (int c = 0; c < 2; ++c)
        acd[c][a][k] = o->Gd[c][a][0] * x1[0][k] + o->Gd[c][a][1] * x1[1][k] +
                       o->G[a][0] * x2[0][c][k] + o->G[a][1] * x2[1][c][k];
    }
  /* curvature and derivatives; second derivative of n by the Weingarten relation
   * n_cd = -(d_d b_cl a^l + b_cl d_d a^l) (sph.f90:553-567) */
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      o->b[a][b] = dot3(x2[a][b], o->n);
      for (int c = 0; c < 2; ++c) o->bd[c][a][b] = dot3(x3[a][b][c], o->n) + dot3(x2[a][b], nd[c]);
    }
  double ndd[2][2][3];
  for (int c = 0; c < 2; ++c)
    for (int d = 0; d < 2; ++d)
      for (int k = 0; k < 3; ++k) {
        double s = 0.0;
        for (int l = 0; l < 2; ++l) s += o->bd[d][c][l] * ac[l][k] + o->b[c][l] * acd[d][l][k];
        ndd[c][d][k] = -s;
      }
  for (int c = 0; c < 2; ++c)
    for (int d = 0; d < 2; ++d)
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
          o->bdd[c][d][a][b] = dot3(x4[a][b][c][d], o->n) + dot3(x3[a][b][c], nd[d]) +
                               dot3(x3[a][b][d], nd[c]) + dot3(x2[a][b], ndd[c][d]);
  /* Christoffel symbols Gamma^b_ac = x_ac . a^b (sph.f90:735-772) and derivatives */
  for (int b = 0; b < 2; ++b)
    for (int a = 0; a < 2; ++a)
      for (int c = 0; c < 2; ++c) {
        o->Gam[b][a][c] = dot3(x2[a][c], ac[b]);
        for (int l = 0; l < 2; ++l)
          o->Gamd[l][b][a][c] = dot3(x3[a][c][l], ac[b]) + dot3(x2[a][c], acd[l][b]);
      }
}

/* ---- plan ------------------------------------------------------------------------------- */
static void sphere_point(double th, double ph, double R, double* x) {
  /* sph.f90:84-86 */
  x[0] = R * sin(th) * sin(ph);
  x[1] = R * sin(th) * cos(ph);
  x[2] = R * cos(th);
}

int mem_plan_init(int L) {
  if (L < 2 || L + 1 > MEM_LMAX) return -1;
  if (g_plan_L == L) return 0;
  mem_plan* p = &g_plan;
  if (g_plan_L >= 0) {
    free(p->Pc), free(p->Pcd), free(p->Pfd), free(p->Pfa);
    free(p->AR), free(p->ARd), free(p->ARdd), free(p->BR), free(p->BRd), free(p->BRdd);
  }
  p->L = L, p->nt2 = L + 1, p->np2 = 2 * (L + 1), p->nt = 2 * p->nt2, p->np = 2 * p->np2;
  const int N = p->nt2, NF = p->nt;
  mem_gauss(N, p->th2, p->w2);
  mem_gauss(NF, p->th, p->w);
  p->Pc = (double*)malloc(sizeof(double) * N * N * N);
  p->Pcd = (double*)malloc(sizeof(double) * N * N * N);
  double* D = (double*)malloc(sizeof(double) * 5 * N * N);
  for (int i = 0; i < N; ++i) {
    mem_legendre_derivs(N, p->th2[i], D);
    memcpy(p->Pc + (size_t)i * N * N, D, sizeof(double) * N * N);
    memcpy(p->Pcd + (size_t)i * N * N, D + N * N, sizeof(double) * N * N);
  }
  p->Pfd = (double*)malloc(sizeof(double) * NF * 5 * N * N);
  for (int i = 0; i < NF; ++i) {
    mem_legendre_derivs(N, p->th[i], D);
    memcpy(p->Pfd + (size_t)i * 5 * N * N, D, sizeof(double) * 5 * N * N);
  }
  free(D);
  p->Pfa = (double*)malloc(sizeof(double) * NF * NF * NF);
  for (int i = 0; i < NF; ++i) mem_legendre(NF, p->th[i], p->Pfa + (size_t)i * NF * NF);
  /* reference state: the unit sphere (sph.f90:56-70, 152-161) */
  const int nf = NF * p->np;
  p->AR = (double*)malloc(sizeof(double) * nf * 4);
  p->ARd = (double*)malloc(sizeof(double) * nf * 8);
  p->ARdd = (double*)malloc(sizeof(double) * nf * 16);
  p->BR = (double*)malloc(sizeof(double) * nf * 4);
  p->BRd = (double*)malloc(sizeof(double) * nf * 8);
  p->BRdd = (double*)malloc(sizeof(double) * nf * 16);
  const int nc = N * p->np2;
  double* xs = (double*)malloc(sizeof(double) * nc * 3);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < p->np2; ++j) {
      double v[3];
      sphere_point(p->th2[i], 2.0 * M_PI * j / p->np2, 1.0, v);
      for (int c = 0; c < 3; ++c) xs[c * nc + i * p->np2 + j] = v[c];
    }
  double* a3 = (double*)malloc(sizeof(double) * 3 * N * N);
  double* b3 = (double*)malloc(sizeof(double) * 3 * N * N);
  for (int c = 0; c < 3; ++c)
    analysis(N, p->np2, p->w2, p->Pc, N, xs + c * nc, a3 + c * N * N, b3 + c * N * N);
  double* X = (double*)malloc(sizeof(double) * nf * 3 * NQ);
  derivs_fine(p, a3, b3, X);
  g_plan_L = L;
  for (int q = 0; q < nf; ++q) {
    node_geom ng;
    node_geometry(X + (size_t)q * 3 * NQ, &ng);
    memcpy(p->AR + q * 4, ng.G, sizeof(double) * 4);
    memcpy(p->ARd + q * 8, ng.Gd, sizeof(double) * 8);
    memcpy(p->ARdd + q * 16, ng.Gdd, sizeof(double) * 16);
    memcpy(p->BR + q * 4, ng.b, sizeof(double) * 4);
    memcpy(p->BRd + q * 8, ng.bd, sizeof(double) * 8);
    memcpy(p->BRdd + q * 16, ng.bdd, sizeof(double) * 16);
  }
  free(xs), free(a3), free(b3), free(X);
  return 0;
}

int mem_grid(int L, double* theta, double* wts) {
  if (mem_plan_init(L)) return -1;
  for (int i = 0; i < g_plan.nt2; ++i) theta[i] = g_plan.th2[i], wts[i] = g_plan.w2[i];
  return 0;
}

/* coefficients of one scalar field on the force grid: a[m * (L+1) + n], b[...] */
int mem_analysis(int L, const double* f, double* a, double* b) {
  if (mem_plan_init(L)) return -1;
  analysis(g_plan.nt2, g_plan.np2, g_plan.w2, g_plan.Pc, g_plan.nt2, f, a, b);
  return 0;
}

int mem_synthesis(int L, const double* a, const double* b, double* f) {
  if (mem_plan_init(L)) return -1;
  synthesis(g_plan.nt2, g_plan.np2, g_plan.Pc, g_plan.nt2, a, b, f);
  return 0;
}

/* Neo-Hookean strain-energy derivatives (membmodel.c:60-79): dW/dI1 = gs/2,
 * dW/dI2 = -gs/2 (1+I2)^-2, and their derivatives along the two surface coordinates. */
void mem_nh_coeffs(double gs, double I2, const double I2d[2], double* w1, double* w2,
                   double w1d[2], double w2d[2]) {
  *w1 = 0.5 * gs;
  *w2 = -0.5 * gs / ((I2 + 1.0) * (I2 + 1.0));
  for (int c = 0; c < 2; ++c) {
    w1d[c] = 0.0;
    w2d[c] = gs / ((I2 + 1.0) * (I2 + 1.0) * (I2 + 1.0)) * I2d[c];
  }
}

/* ---------------------------------------------------------------------------------------
 * The law at one fine node (sph.f90:1615-1700): tension of the neo-Hookean membrane
 * (membmodel.c:60-79 with hf_gscell = gs/2, sph.f90:35), bending moment about the
 * reference curvature (sph.f90:1236-1372), transverse shear Q (:1375-1425), and the load
 * f = q1 a_1 + q2 a_2 + q3 n.
 * ------------------------------------------------------------------------------------- */
static void node_law(const node_geom* o, const double* AR4, const double* ARd8,
                     const double* ARdd16, const double* BR4, const double* BRd8,
                     const double* BRdd16, double gs, double eb, double f[3]) {
  double A[2][2], Ad[2][2][2], Add[2][2][2][2], B[2][2], Bd[2][2][2], Bdd[2][2][2][2];
  memcpy(A, AR4, sizeof A), memcpy(Ad, ARd8, sizeof Ad), memcpy(Add, ARdd16, sizeof Add);
  memcpy(B, BR4, sizeof B), memcpy(Bd, BRd8, sizeof Bd), memcpy(Bdd, BRdd16, sizeof Bdd);
  /* invariants (sph.f90:1115-1157) */
  const double detA = A[0][0] * A[1][1] - A[0][1] * A[0][1];
  const double deta = o->g[0][0] * o->g[1][1] - o->g[0][1] * o->g[0][1];
  const double I2 = detA * deta - 1.0;
  const double Js = sqrt(1.0 + I2);
  double I2d[2], Jsd[2];
  for (int c = 0; c < 2; ++c) {
    const double dA = Ad[c][0][0] * A[1][1] + A[0][0] * Ad[c][1][1] - 2.0 * A[0][1] * Ad[c][0][1];
    const double da = o->gd[c][0][0] * o->g[1][1] + o->g[0][0] * o->gd[c][1][1] -
                      2.0 * o->g[0][1] * o->gd[c][0][1];
    I2d[c] = dA * deta + da * detA;
    Jsd[c] = 0.5 / Js * I2d[c];
  }
  double w1, w2, w1d[2], w2d[2];
  mem_nh_coeffs(gs, I2, I2d, &w1, &w2, w1d, w2d);
  /* Cauchy tension T^ab = 2/Js w1 A^ab + 2 Js w2 a^ab and d_c T^ab (sph.f90:1179-1234) */
  double T[2][2], Td[2][2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      T[a][b] = 2.0 / Js * w1 * A[a][b] + 2.0 * Js * w2 * o->G[a][b];
      for (int c = 0; c < 2; ++c)
        Td[c][a][b] = -2.0 / (Js * Js) * Jsd[c] * w1 * A[a][b] + 2.0 / Js * w1 * Ad[c][a][b] +
                      2.0 * Jsd[c] * w2 * o->G[a][b] + 2.0 * Js * w2d[c] * o->G[a][b] +
                      2.0 * Js * w2 * o->Gd[c][a][b];
    }
  /* membrane equilibrium: q^b = T^ab_|a, q3 = T^ab b_ab */
  double q[2];
  for (int b = 0; b < 2; ++b) {
    double s = Td[0][0][b] + Td[1][1][b];
    for (int a = 0; a < 2; ++a)
      for (int l = 0; l < 2; ++l) s += o->Gam[a][a][l] * T[l][b] + o->Gam[b][a][l] * T[a][l];
    q[b] = s;
  }
  double q3 = 0.0;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) q3 += T[a][b] * o->b[a][b];
  if (eb != 0.0) {
    /* K[c][a] = A^{c'a} B_{cc'} - a^{c'a} b_{cc'}; Mt^{ab} = a^{bc} K[c][a] (unsymmetrised),
     * M^{ab}: entries (a,b) = (0,0), (0,1), (1,1) of Mt with M^{10} = M^{01} */
    double K[2][2], Kd[2][2][2], Kdd[2][2][2][2];
    for (int c = 0; c < 2; ++c)
      for (int a = 0; a < 2; ++a) {
        double s = 0.0;
        for (int cp = 0; cp < 2; ++cp) s += A[cp][a] * B[c][cp] - o->G[cp][a] * o->b[c][cp];
        K[c][a] = s;
        for (int l = 0; l < 2; ++l) {
          double t = 0.0;
          for (int cp = 0; cp < 2; ++cp)
            t += Ad[l][cp][a] * B[c][cp] + A[cp][a] * Bd[l][c][cp] - o->Gd[l][cp][a] * o->b[c][cp] -
                 o->G[cp][a] * o->bd[l][c][cp];
          Kd[l][c][a] = t;
          for (int k = 0; k < 2; ++k) {
            double u = 0.0;
            for (int cp = 0; cp < 2; ++cp)
              u += Add[l][k][cp][a] * B[c][cp] + Ad[k][cp][a] * Bd[l][c][cp] +
                   Ad[l][cp][a] * Bd[k][c][cp] + A[cp][a] * Bdd[l][k][c][cp] -
                   o->Gdd[l][k][cp][a] * o->b[c][cp] - o->Gd[k][cp][a] * o->bd[l][c][cp] -
                   o->Gd[l][cp][a] * o->bd[k][c][cp] - o->G[cp][a] * o->bdd[l][k][c][cp];
            Kdd[l][k][c][a] = u;
          }
        }
      }
    double M[2][2], Md[2][2][2], dds = 0.0;  /* Md[l][a][b] = d_l M^ab */
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) {
        double m = 0.0;
        double md[2] = {0.0, 0.0};
        for (int c = 0; c < 2; ++c) {
          m += o->G[b][c] * K[c][a];
          for (int l = 0; l < 2; ++l) md[l] += o->Gd[l][b][c] * K[c][a] + o->G[b][c] * Kd[l][c][a];
          /* d_b d_a Mt^{ab}, summed over all (a,b) (sph.f90:1301-1327) */
          dds += o->Gdd[b][a][b][c] * K[c][a] + o->Gd[a][b][c] * Kd[b][c][a] +
                 o->Gd[b][b][c] * Kd[a][c][a] + o->G[b][c] * Kdd[b][a][c][a];
        }
        M[a][b] = m;
        Md[0][a][b] = md[0], Md[1][a][b] = md[1];
      }
    M[1][0] = M[0][1];
    Md[0][1][0] = Md[0][0][1], Md[1][1][0] = Md[1][0][1];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) {
        M[a][b] *= eb;
        for (int l = 0; l < 2; ++l) Md[l][a][b] *= eb;
      }
    dds *= eb;
    /* Q^b = M^ab_|a and Q^b_|b (sph.f90:1375-1425) */
    double Q[2], Qc = dds;
    for (int b = 0; b < 2; ++b) {
      double s = 0.0;
      for (int a = 0; a < 2; ++a) {
        s += Md[a][a][b];
        for (int l = 0; l < 2; ++l) {
          s += o->Gam[a][a][l] * M[l][b] + o->Gam[b][a][l] * M[a][l];
          Qc += o->Gamd[b][a][a][l] * M[l][b] + o->Gam[a][a][l] * Md[b][l][b] +
                o->Gamd[b][b][a][l] * M[a][l] + o->Gam[b][a][l] * Md[b][a][l];
        }
      }
      Q[b] = s;
    }
    for (int b = 0; b < 2; ++b)
      for (int l = 0; l < 2; ++l) Qc += Q[l] * o->Gam[b][b][l];
    /* mixed curvature b^a_b = a^{la} b_{bl} (sph.f90:1427-1454) */
    double bm[2][2]; /* bm[b][a] */
    for (int b = 0; b < 2; ++b)
      for (int a = 0; a < 2; ++a) bm[b][a] = o->G[0][a] * o->b[b][0] + o->G[1][a] * o->b[b][1];
    q[0] -= bm[0][0] * Q[0] + bm[1][0] * Q[1];
    q[1] -= bm[0][1] * Q[0] + bm[1][1] * Q[1];
    q3 += Qc;
  }
  for (int k = 0; k < 3; ++k) f[k] = q[0] * o->a[0][k] + q[1] * o->a[1][k] + q3 * o->n[k];
}

/* ---------------------------------------------------------------------------------------
 * complagrforces for one capsule whose (L+1) x 2(L+1) points x,y,z (latitude-major) lie on
 * the Gauss grid: writes the traction f (per unit area), the filtered unit normal and the
 * point areas surfmet * w_i / sin(theta_i) * dphi (forcing.f90:91, sph.f90:131).
 * gs: shear modulus (param.f90:88); eb: bending modulus (param.f90:91), 0 = no bending.
 * ------------------------------------------------------------------------------------- */
int mem_traction(int L, double gs, double eb, const double* x, const double* y, const double* z,
                 double* fx, double* fy, double* fz, double* nx, double* ny, double* nz,
                 double* area) {
  if (mem_plan_init(L)) return -1;
  const mem_plan* p = &g_plan;
  const int N = p->nt2, NF = p->nt, nc = N * p->np2, nf = NF * p->np;
  double* a3 = (double*)malloc(sizeof(double) * 3 * N * N);
  double* b3 = (double*)malloc(sizeof(double) * 3 * N * N);
  const double* xyz[3] = {x, y, z};
  for (int c = 0; c < 3; ++c)
    analysis(N, p->np2, p->w2, p->Pc, N, xyz[c], a3 + c * N * N, b3 + c * N * N);
  double* X = (double*)malloc(sizeof(double) * nf * 3 * NQ);
  derivs_fine(p, a3, b3, X);
  double* F = (double*)malloc(sizeof(double) * nf * 6);  /* fine-grid load and normal */
  for (int q = 0; q < nf; ++q) {
    node_geom ng;
    node_geometry(X + (size_t)q * 3 * NQ, &ng);
    double f[3];
    node_law(&ng, p->AR + q * 4, p->ARd + q * 8, p->ARdd + q * 16, p->BR + q * 4,
             p->BRd + q * 8, p->BRdd + q * 16, gs, eb, f);
    for (int k = 0; k < 3; ++k) F[k * nf + q] = f[k], F[(3 + k) * nf + q] = ng.n[k];
  }
  /* dealias: analysis on the fine grid, synthesis of degrees < L+1 on the force grid */
  double* af = (double*)malloc(sizeof(double) * NF * NF);
  double* bf = (double*)malloc(sizeof(double) * NF * NF);
  double* at = (double*)malloc(sizeof(double) * N * N);
  double* bt = (double*)malloc(sizeof(double) * N * N);
  double* outs[6] = {fx, fy, fz, nx, ny, nz};
  for (int c = 0; c < 6; ++c) {
    analysis(NF, p->np, p->w, p->Pfa, NF, F + (size_t)c * nf, af, bf);
    for (int m = 0; m < N; ++m)
      for (int n = 0; n < N; ++n) at[m * N + n] = af[m * NF + n], bt[m * N + n] = bf[m * NF + n];
    synthesis(N, p->np2, p->Pc, N, at, bt, outs[c]);
  }
  for (int q = 0; q < nc; ++q) {
    const double r = sqrt(nx[q] * nx[q] + ny[q] * ny[q] + nz[q] * nz[q]);
    nx[q] /= r, ny[q] /= r, nz[q] /= r;
  }
  /* surface metric on the force grid from the coarse coefficients (sph3.f90:186-260) */
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < p->np2; ++j) {
      double t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
      for (int c = 0; c < 3; ++c)
        for (int m = 0; m < N; ++m) {
          const double cm = cos_mj(m, j, p->np2), sm = sin_mj(m, j, p->np2), h = m == 0 ? 0.5 : 1.0;
          for (int n = m; n < N; ++n) {
            const double an = a3[c * N * N + m * N + n], bn = b3[c * N * N + m * N + n];
            t1[c] += h * p->Pcd[((size_t)i * N + m) * N + n] * (an * cm - bn * sm);
            t2[c] += h * p->Pc[((size_t)i * N + m) * N + n] * m * (-an * sm - bn * cm);
          }
        }
      double cr[3];
      cross3(t1, t2, cr);
      area[i * p->np2 + j] =
          sqrt(dot3(cr, cr)) * p->w2[i] / sin(p->th2[i]) * (2.0 * M_PI / p->np2);
    }
  free(a3), free(b3), free(X), free(F), free(af), free(bf), free(at), free(bt);
  return 0;
}

/* Neo-Hookean shear energy of one capsule (cal_energy_nh, membmodel.c:44-48:
 * W = gs/2 (I1 - 1 + 1/(I2+1)) per unit reference area), integrated on the fine grid with
 * the reference (unit-sphere) area element.  Used to check the tension by virtual work. */
double mem_shear_energy(int L, double gs, const double* x, const double* y, const double* z) {
  if (mem_plan_init(L)) return NAN;
  const mem_plan* p = &g_plan;
  const int N = p->nt2, nf = p->nt * p->np;
  double* a3 = (double*)malloc(sizeof(double) * 3 * N * N);
  double* b3 = (double*)malloc(sizeof(double) * 3 * N * N);
  const double* xyz[3] = {x, y, z};
  for (int c = 0; c < 3; ++c)
    analysis(N, p->np2, p->w2, p->Pc, N, xyz[c], a3 + c * N * N, b3 + c * N * N);
  double* X = (double*)malloc(sizeof(double) * nf * 3 * NQ);
  derivs_fine(p, a3, b3, X);
  double e = 0.0;
  for (int q = 0; q < nf; ++q) {
    node_geom ng;
    node_geometry(X + (size_t)q * 3 * NQ, &ng);
    const double* A = p->AR + q * 4;  /* A^ab, row-major 2x2 */
    const double I1 = A[0] * ng.g[0][0] + A[3] * ng.g[1][1] + 2.0 * A[1] * ng.g[0][1] - 2.0;
    const double detA = A[0] * A[3] - A[1] * A[1];
    const double I2 = detA * (ng.g[0][0] * ng.g[1][1] - ng.g[0][1] * ng.g[0][1]) - 1.0;
    const int i = q / p->np;
    /* reference area element sin(theta) dtheta dphi -> Gauss weight w_i * dphi */
    e += 0.5 * gs * (I1 - 1.0 + 1.0 / (I2 + 1.0)) * p->w[i] * (2.0 * M_PI / p->np);
  }
  free(a3), free(b3), free(X);
  return e;
}
