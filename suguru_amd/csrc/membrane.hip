// membrane.hip -- the reference's FP64 capsule membrane law on gfx950, one workgroup per
// capsule (libmemb.so, C ABI in include/memb.h).
//
// Per capsule (forcing.f90:11-114 with the force grid equal to the harmonic grid):
//   1. shape coefficients: longitude DFT per latitude, then Gauss-Legendre quadrature
//      (the SPHEREPACK analysis of sph1.f90:29-79);
//   2. on the 2x (dealias) grid, latitude ring by ring: x and all its derivatives up to 4th
//      order (sph3.f90:9-182), then per node the geometry (normal, metrics, curvature,
//      Christoffel symbols and their derivatives) and the load q1 a1 + q2 a2 + q3 n of the
//      neo-Hookean tension and the bending moment (sph.f90:1523-1700);
//   3. the dealias filter: analysis of the 2x-grid load and normal, synthesis of degrees
//      <= L on the force grid (sph1.f90:723-951, 1378-1406), and the point area from the
//      surface metric (sph3.f90:186-260; forcing.f90:91).
// Everything is FP64.  Tables (Gauss nodes, Legendre functions and their theta-derivatives,
// the unit-sphere reference state) are built once per degree by memb_plan_create.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/memb.h"

namespace memb {

constexpr int NQ = 15;        // derivatives of total order <= 4
constexpr int RING = 3;       // fine latitude rings processed together
constexpr int THREADS = 128;     // per capsule workgroup (two wavefronts)
constexpr int NREF = 56;      // reference-state doubles per fine node
constexpr int LMAX = 23;

__host__ __device__ inline int qidx(int kt, int kp) {
  const int s = kt + kp;
  return s * (s + 1) / 2 + kp;
}

struct Tables {
  int L, N, np2, nt, np;       // N = L+1 coarse latitudes; nt = 2N, np = 2 np2
  const double* w2;            // [N] coarse Gauss weights
  const double* w;             // [nt] fine Gauss weights
  const double* s2;            // [N] sin(theta) coarse
  const double* Pc;            // [N][N][N] (i, m, n) coarse Pbar
  const double* Pcd;           // [N][N][N] coarse dPbar/dtheta
  const double* Pfd;           // [nt][5][N][N] fine theta-derivatives 0..4
  const double* Pfa;           // [nt][N][N] fine Pbar, degrees <= L
  const double* cs2;           // [np2][N] cos(m phi_j) coarse
  const double* sn2;
  const double* csf;           // [np][N] fine
  const double* snf;
  const double* ref;           // [nt*np][NREF] unit-sphere state (mode 0 reads it)
};

// ---- per-node geometry and law (device versions of the oracle's formulas) ------------------
struct Geom {
  double a[2][3], n[3], J;
  double g[2][2], gd[2][2][2];
  double G[2][2], Gd[2][2][2], Gdd[2][2][2][2];
  double b[2][2], bd[2][2][2], bdd[2][2][2][2];
  double Gam[2][2][2], Gamd[2][2][2][2];
};

__device__ inline double dot3(const double* u, const double* v) {
  return u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
}
__device__ inline void cross3(const double* u, const double* v, double* r) {
  r[0] = u[1] * v[2] - u[2] * v[1];
  r[1] = u[2] * v[0] - u[0] * v[2];
  r[2] = u[0] * v[1] - u[1] * v[0];
}
__device__ inline void mul2(const double X[2][2], const double Y[2][2], double R[2][2]) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) R[i][j] = X[i][0] * Y[0][j] + X[i][1] * Y[1][j];
}

// v of the adjacent lane (lanes 2k and 2k + 1 swap): DPP quad_perm [1, 0, 3, 2] on the two
// halves, a VALU move instead of an LDS-crossbar permute; both lanes of a pair must be active
__device__ __forceinline__ double swap_pair(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, 0xB1, 0xF, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), 0xB1, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// X[c][q]: derivatives of coordinate c; returns the node geometry
__device__ void geometry(const double (*X)[NQ], Geom& o) {
  double x1[2][3], x2[2][2][3], x3[2][2][2][3], x4[2][2][2][2][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      x1[a][k] = X[k][qidx(1 - a, a)];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        x2[a][b][k] = X[k][qidx(2 - a - b, a + b)];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          x3[a][b][c][k] = X[k][qidx(3 - a - b - c, a + b + c)];
#pragma unroll
          for (int d = 0; d < 2; ++d) x4[a][b][c][d][k] = X[k][qidx(4 - a - b - c - d, a + b + c + d)];
        }
      }
    }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int k = 0; k < 3; ++k) o.a[a][k] = x1[a][k];
  double cv[3], nd[2][3];
  cross3(x1[0], x1[1], cv);
  o.J = sqrt(dot3(cv, cv));
  const double rJ = 1.0 / o.J;  // one division per node; the others multiply by it
#pragma unroll
  for (int k = 0; k < 3; ++k) o.n[k] = cv[k] * rJ;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    double t1[3], t2[3], cd[3];
    cross3(x2[0][c], x1[1], t1);
    cross3(x1[0], x2[1][c], t2);
#pragma unroll
    for (int k = 0; k < 3; ++k) cd[k] = t1[k] + t2[k];
    const double Jd = dot3(o.n, cd);
#pragma unroll
    for (int k = 0; k < 3; ++k) nd[c][k] = (cd[k] - Jd * o.n[k]) * rJ;
  }
  double gdd[2][2][2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      o.g[a][b] = dot3(x1[a], x1[b]);
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        o.gd[c][a][b] = dot3(x2[a][c], x1[b]) + dot3(x1[a], x2[b][c]);
#pragma unroll
        for (int d = 0; d < 2; ++d)
          gdd[c][d][a][b] = dot3(x3[a][c][d], x1[b]) + dot3(x2[a][c], x2[b][d]) +
                            dot3(x2[a][d], x2[b][c]) + dot3(x1[a], x3[b][c][d]);
      }
    }
  const double det = o.g[0][0] * o.g[1][1] - o.g[0][1] * o.g[1][0];
  const double rdet = 1.0 / det;
  o.G[0][0] = o.g[1][1] * rdet, o.G[1][1] = o.g[0][0] * rdet;
  o.G[0][1] = -o.g[0][1] * rdet, o.G[1][0] = -o.g[1][0] * rdet;
  double GdgG[2][2][2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    double t[2][2], u[2][2];
    mul2(o.G, o.gd[c], t);
    mul2(t, o.G, u);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) GdgG[c][a][b] = u[a][b], o.Gd[c][a][b] = -u[a][b];
  }
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      double t1[2][2], t2[2][2], t3[2][2], t4[2][2], t5[2][2], gcd[2][2];
      mul2(GdgG[d], o.gd[c], t1);
      mul2(t1, o.G, t2);
      mul2(GdgG[c], o.gd[d], t3);
      mul2(t3, o.G, t4);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) gcd[a][b] = gdd[c][d][a][b];
      mul2(o.G, gcd, t1);
      mul2(t1, o.G, t5);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) o.Gdd[c][d][a][b] = t2[a][b] + t4[a][b] - t5[a][b];
    }
  double ac[2][3], acd[2][2][3];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      ac[a][k] = o.G[a][0] * x1[0][k] + o.G[a][1] * x1[1][k];
#pragma unroll
      for (int c = 0; c < 2; ++c)
        acd[c][a][k] = o.Gd[c][a][0] * x1[0][k] + o.Gd[c][a][1] * x1[1][k] +
                       o.G[a][0] * x2[0][c][k] + o.G[a][1] * x2[1][c][k];
    }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      o.b[a][b] = dot3(x2[a][b], o.n);
#pragma unroll
      for (int c = 0; c < 2; ++c) o.bd[c][a][b] = dot3(x3[a][b][c], o.n) + dot3(x2[a][b], nd[c]);
    }
  double ndd[2][2][3];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int d = 0; d < 2; ++d)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double s = 0.0;
#pragma unroll
        for (int l = 0; l < 2; ++l) s += o.bd[d][c][l] * ac[l][k] + o.b[c][l] * acd[d][l][k];
        ndd[c][d][k] = -s;
      }
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int d = 0; d < 2; ++d)
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
          o.bdd[c][d][a][b] = dot3(x4[a][b][c][d], o.n) + dot3(x3[a][b][c], nd[d]) +
                              dot3(x3[a][b][d], nd[c]) + dot3(x2[a][b], ndd[c][d]);
#pragma unroll
  for (int b = 0; b < 2; ++b)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        o.Gam[b][a][c] = dot3(x2[a][c], ac[b]);
#pragma unroll
        for (int l = 0; l < 2; ++l)
          o.Gamd[l][b][a][c] = dot3(x3[a][c][l], ac[b]) + dot3(x2[a][c], acd[l][b]);
      }
}

// load q1 a1 + q2 a2 + q3 n at one node (sph.f90:1615-1700; membmodel.c:60-79)
__device__ void law(const Geom& o, const double* __restrict__ R, double gs, double eb,
                    double f[3]) {
  double A[2][2], Ad[2][2][2], Add[2][2][2][2], B[2][2], Bd[2][2][2], Bdd[2][2][2][2];
#pragma unroll
  for (int q = 0; q < 4; ++q) (&A[0][0])[q] = R[q], (&B[0][0])[q] = R[28 + q];
#pragma unroll
  for (int q = 0; q < 8; ++q) (&Ad[0][0][0])[q] = R[4 + q], (&Bd[0][0][0])[q] = R[32 + q];
#pragma unroll
  for (int q = 0; q < 16; ++q) (&Add[0][0][0][0])[q] = R[12 + q], (&Bdd[0][0][0][0])[q] = R[40 + q];
  const double detA = A[0][0] * A[1][1] - A[0][1] * A[0][1];
  const double deta = o.g[0][0] * o.g[1][1] - o.g[0][1] * o.g[0][1];
  const double I2 = detA * deta - 1.0;
  const double Js = sqrt(1.0 + I2);
  const double rJs = 1.0 / Js, rI = 1.0 / (I2 + 1.0);  // one division each
  double I2d[2], Jsd[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const double dA = Ad[c][0][0] * A[1][1] + A[0][0] * Ad[c][1][1] - 2.0 * A[0][1] * Ad[c][0][1];
    const double da = o.gd[c][0][0] * o.g[1][1] + o.g[0][0] * o.gd[c][1][1] -
                      2.0 * o.g[0][1] * o.gd[c][0][1];
    I2d[c] = dA * deta + da * detA;
    Jsd[c] = 0.5 * rJs * I2d[c];
  }
  const double w1 = 0.5 * gs;
  const double w2 = -0.5 * gs * (rI * rI);
  double w2d[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) w2d[c] = gs * (rI * rI * rI) * I2d[c];
  double T[2][2], Td[2][2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      T[a][b] = 2.0 * rJs * w1 * A[a][b] + 2.0 * Js * w2 * o.G[a][b];
#pragma unroll
      for (int c = 0; c < 2; ++c)
        Td[c][a][b] = -2.0 * (rJs * rJs) * Jsd[c] * w1 * A[a][b] + 2.0 * rJs * w1 * Ad[c][a][b] +
                      2.0 * Jsd[c] * w2 * o.G[a][b] + 2.0 * Js * w2d[c] * o.G[a][b] +
                      2.0 * Js * w2 * o.Gd[c][a][b];
    }
  double q[2];
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    double s = Td[0][0][b] + Td[1][1][b];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int l = 0; l < 2; ++l) s += o.Gam[a][a][l] * T[l][b] + o.Gam[b][a][l] * T[a][l];
    q[b] = s;
  }
  double q3 = 0.0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) q3 += T[a][b] * o.b[a][b];
  if (eb != 0.0) {
    double K[2][2], Kd[2][2][2], Kdd[2][2][2][2];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        double s = 0.0;
#pragma unroll
        for (int cp = 0; cp < 2; ++cp) s += A[cp][a] * B[c][cp] - o.G[cp][a] * o.b[c][cp];
        K[c][a] = s;
#pragma unroll
        for (int l = 0; l < 2; ++l) {
          double t = 0.0;
#pragma unroll
          for (int cp = 0; cp < 2; ++cp)
            t += Ad[l][cp][a] * B[c][cp] + A[cp][a] * Bd[l][c][cp] - o.Gd[l][cp][a] * o.b[c][cp] -
                 o.G[cp][a] * o.bd[l][c][cp];
          Kd[l][c][a] = t;
#pragma unroll
          for (int k = 0; k < 2; ++k) {
            double u = 0.0;
#pragma unroll
            for (int cp = 0; cp < 2; ++cp)
              u += Add[l][k][cp][a] * B[c][cp] + Ad[k][cp][a] * Bd[l][c][cp] +
                   Ad[l][cp][a] * Bd[k][c][cp] + A[cp][a] * Bdd[l][k][c][cp] -
                   o.Gdd[l][k][cp][a] * o.b[c][cp] - o.Gd[k][cp][a] * o.bd[l][c][cp] -
                   o.Gd[l][cp][a] * o.bd[k][c][cp] - o.G[cp][a] * o.bdd[l][k][c][cp];
            Kdd[l][k][c][a] = u;
          }
        }
      }
    double M[2][2], Md[2][2][2], dds = 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        double m = 0.0, md0 = 0.0, md1 = 0.0;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          m += o.G[b][c] * K[c][a];
          md0 += o.Gd[0][b][c] * K[c][a] + o.G[b][c] * Kd[0][c][a];
          md1 += o.Gd[1][b][c] * K[c][a] + o.G[b][c] * Kd[1][c][a];
          dds += o.Gdd[b][a][b][c] * K[c][a] + o.Gd[a][b][c] * Kd[b][c][a] +
                 o.Gd[b][b][c] * Kd[a][c][a] + o.G[b][c] * Kdd[b][a][c][a];
        }
        M[a][b] = m, Md[0][a][b] = md0, Md[1][a][b] = md1;
      }
    M[1][0] = M[0][1];
    Md[0][1][0] = Md[0][0][1], Md[1][1][0] = Md[1][0][1];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        M[a][b] *= eb;
        Md[0][a][b] *= eb, Md[1][a][b] *= eb;
      }
    dds *= eb;
    double Q[2], Qc = dds;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        s += Md[a][a][b];
#pragma unroll
        for (int l = 0; l < 2; ++l) {
          s += o.Gam[a][a][l] * M[l][b] + o.Gam[b][a][l] * M[a][l];
          Qc += o.Gamd[b][a][a][l] * M[l][b] + o.Gam[a][a][l] * Md[b][l][b] +
                o.Gamd[b][b][a][l] * M[a][l] + o.Gam[b][a][l] * Md[b][a][l];
        }
      }
      Q[b] = s;
    }
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int l = 0; l < 2; ++l) Qc += Q[l] * o.Gam[b][b][l];
    double bm[2][2];
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int a = 0; a < 2; ++a) bm[b][a] = o.G[0][a] * o.b[b][0] + o.G[1][a] * o.b[b][1];
    q[0] -= bm[0][0] * Q[0] + bm[1][0] * Q[1];
    q[1] -= bm[0][1] * Q[0] + bm[1][1] * Q[1];
    q3 += Qc;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) f[k] = q[0] * o.a[0][k] + q[1] * o.a[1][k] + q3 * o.n[k];
}

// ---- the per-capsule kernel ---------------------------------------------------------------
// MODE 0: the law; 1: the reference state of one capsule (the unit sphere) into `ref`;
// 2: shape coefficients only; 3: geometry only (unit normal and point area on the force
// grid, no load).  Persistent: workgroup b handles capsules b, b + gridDim.x, ...
// LDS: coarse coefficients [3][N][N] a and b, the longitude factors of both grids (2 (np + np2)
// values), then a work region reused by the phases.
// scratch: per workgroup, longitude DFT of the 2x-grid outputs [nt][6][N][2].
// Registers: capped for two wavefronts per SIMD (256 VGPRs, about 42 of them spilled).  With
// no cap the law takes 256 VGPRs + 40 AGPRs, one wavefront per SIMD, and runs 1.6x slower.
// 128 threads per capsule and four capsules per CU (LDS: 39 KB each at L = 16); the per-node
// phase takes chunks of 128 consecutive nodes, one per thread (step 2 below).
// The transform loops carry `#pragma unroll 4`.  Their trip counts are run-time values, so
// without the hint each iteration waits for its own table loads.  Each sum keeps its order,
// so the output is bitwise unchanged; 1000 capsules at L = 16 take 1.32 instead of 1.47 ms
// (unrolled by 8: 1.35 ms).
template <int MODE>
__global__ void __launch_bounds__(THREADS, 2)
    k_mem(Tables T, long long ncap, double gs, double eb, const double* __restrict__ x,
          const double* __restrict__ y, const double* __restrict__ z, double* __restrict__ o0,
          double* __restrict__ o1, double* __restrict__ o2, double* __restrict__ o3,
          double* __restrict__ o4, double* __restrict__ o5, double* __restrict__ o6,
          double* __restrict__ scratch, const int* __restrict__ sel) {
  extern __shared__ double lds[];
  const int tid = threadIdx.x;
  const int N = T.N, NN = N * N, np2 = T.np2, nt = T.nt, np = T.np;
  const int nc = N * np2, NT = N * (N + 1) / 2;  // NT: pairs (m, n >= m)
  double* ca = lds;            // [3][N][N]
  double* cb = ca + 3 * NN;    // [3][N][N]
  // cos / sin(2 pi q / np) for q < np, then the same for q < np2: every longitude factor
  // cos / sin(m phi_j) of the fine (force) grid is entry (m j) mod np (np2) of these, so the
  // transforms read them from LDS instead of from the [j][m] tables in HBM
  double* tcf = cb + 3 * NN;
  double* tsf = tcf + np;
  double* tc2 = tsf + np;
  double* ts2 = tc2 + np2;
  double* U = ts2 + np2;       // work region
  double* scr = scratch + (size_t)blockIdx.x * nt * 6 * N * 2;  // >= 6 nc doubles too
  for (int q = tid; q < np; q += THREADS) tcf[q] = T.csf[q * N + 1], tsf[q] = T.snf[q * N + 1];
  for (int q = tid; q < np2; q += THREADS) tc2[q] = T.cs2[q * N + 1], ts2[q] = T.sn2[q * N + 1];
  __syncthreads();
  for (long long cap = blockIdx.x; cap < ncap; cap += gridDim.x) {
    if (sel && sel[cap] == 0) continue;  // (uniform in the workgroup)
    const double* xyz[3] = {x + cap * nc, y + cap * nc, z + cap * nc};
    // 1. analysis: longitude DFT per (c, i, m) into U, then quadrature per (c, m, n)
    // one thread per (i, m) for the three coordinates, every point in order: the raw
    // coordinates carry the capsule's position, and pairing points j and j + np2/2 (the
    // phi -> phi + pi symmetry used for the outputs below) rounds the near-zero high-m
    // coefficients differently enough for the fourth derivatives' m^4 to show (4x the
    // difference to the oracle on an ellipsoid)
    for (int t = tid; t < NN; t += THREADS) {
      const int i = t / N, m = t % N;
      double cs[3] = {0.0, 0.0, 0.0}, sn[3] = {0.0, 0.0, 0.0};
      int q = 0;  // (m j) mod np2
      #pragma unroll 4
      for (int j = 0; j < np2; ++j) {
        const double c = tc2[q], s = ts2[q];
        q += m;
        if (q >= np2) q -= np2;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double v = xyz[k][i * np2 + j];
          cs[k] += v * c;
          sn[k] += v * s;
        }
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) U[2 * (k * NN + t)] = cs[k], U[2 * (k * NN + t) + 1] = sn[k];
    }
    __syncthreads();
    // one thread per (m, n >= m) for the three coordinates (entries n < m are never read)
    for (int t = tid; t < NT; t += THREADS) {
      int m = 0, n = t;
      while (n >= N - m) n -= N - m, ++m;
      n += m;
      double sa[3] = {0.0, 0.0, 0.0}, sb[3] = {0.0, 0.0, 0.0};
      #pragma unroll 4
      for (int i = 0; i < N; ++i) {
        const double p = T.w2[i] * T.Pc[(i * N + m) * N + n];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const int u = (c * N + i) * N + m;
          sa[c] += p * U[2 * u];
          sb[c] += p * U[2 * u + 1];
        }
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        ca[(c * N + m) * N + n] = 2.0 / np2 * sa[c];
        cb[(c * N + m) * N + n] = -2.0 / np2 * sb[c];
      }
    }
    __syncthreads();
    if (MODE == 2) {
      // packed: a(m, n>=m) then b(m>=1, n>=m), per coordinate
      for (int t = tid; t < 3 * NN; t += THREADS) {
        const int c = t / NN, m = (t / N) % N, n = t % N;
        if (n < m) continue;
        const int pa = m * N - m * (m - 1) / 2 + (n - m);
        o0[cap * 3 * NN + c * NN + pa] = ca[(c * N + m) * N + n];
        if (m >= 1) {
          const int pb = (N * (N + 1)) / 2 + (m - 1) * N - (m - 1) * (m - 2) / 2 - (m - 1) + (n - m);
          o0[cap * 3 * NN + c * NN + pb] = cb[(c * N + m) * N + n];
        }
      }
      __syncthreads();
      continue;
    }
    // 2. the 2x grid in chunks of W consecutive nodes, one node per thread (not for MODE 3:
    //    geometry only).  A chunk of W <= (RING - 1) np nodes touches at most RING rings.
    //    Ring r's coefficients and outputs live in slot r % RING of AB and F: its coefficients
    //    are summed before the first chunk that touches it, its longitude DFT runs after the
    //    chunk that completes it, and only ring r + RING takes the slot again, whose first node
    //    lies (RING - 1) np + 1 >= W nodes after ring r's last (so in a later chunk).  With
    //    W = THREADS every round of the per-node phase is full (one partial round per capsule
    //    instead of one per group of RING rings).
    double* AB = U;                          // [RING][3][5][N][2]
    double* F = AB + RING * 3 * 5 * N * 2;   // [RING][np][6]
    const int nnode = MODE == 3 ? 0 : nt * np;
    const int W = min(THREADS, (RING - 1) * np);  // even: node pairs stay in one chunk
    int rnext = 0, rdone = 0;  // first ring without coefficients / without its DFT
    for (int n0 = 0; n0 < nnode; n0 += W) {
      const int n1 = min(n0 + W, nnode);
      const int rhi = (n1 - 1) / np;  // last ring of the chunk
      // a thread sums the three coordinates of one (ring, k, m): each table load feeds six FMAs
      for (int t = tid; t < (rhi + 1 - rnext) * 5 * N; t += THREADS) {
        const int ring = rnext + t / (5 * N), k = (t / N) % 5, m = t % N;
        const int r = ring % RING;
        const double* Pd = T.Pfd + (((size_t)ring * 5 + k) * N + m) * N;
        const double h = m == 0 ? 0.5 : 1.0;
        double sa0 = 0.0, sb0 = 0.0, sa1 = 0.0, sb1 = 0.0, sa2 = 0.0, sb2 = 0.0;
        #pragma unroll 4
        for (int n = m; n < N; ++n) {
          const double p = Pd[n];
          sa0 += p * ca[m * N + n], sb0 += p * cb[m * N + n];
          sa1 += p * ca[(N + m) * N + n], sb1 += p * cb[(N + m) * N + n];
          sa2 += p * ca[(2 * N + m) * N + n], sb2 += p * cb[(2 * N + m) * N + n];
        }
        const int t0 = ((r * 3) * 5 + k) * N + m, cs = 5 * N;  // AB slot of coordinate 0; stride
        AB[2 * t0] = h * sa0, AB[2 * t0 + 1] = h * sb0;
        AB[2 * (t0 + cs)] = h * sa1, AB[2 * (t0 + cs) + 1] = h * sb1;
        AB[2 * (t0 + 2 * cs)] = h * sa2, AB[2 * (t0 + 2 * cs) + 1] = h * sb2;
      }
      rnext = rhi + 1;
      __syncthreads();
      // Lanes 2k and 2k + 1 take the node pair (jj, jj + np/2) of one ring, whose factors
      // agree up to (-1)^m (phi -> phi + pi): lane 2k sums the even-m terms E and lane 2k + 1
      // the odd-m terms O, both at phi_jj; after one exchange node jj holds E + O and node
      // jj + np/2 holds E - O, and each lane goes on with its own node.
      for (int pn = n0 + tid; pn < n1; pn += THREADS) {
        const int par = pn & 1, pr = pn >> 1, hp = np / 2;
        const int ring = pr / hp, jj = pr - ring * hp, j = jj + par * hp, r = ring % RING;
        const int node = ring * np + j;
        double X[3][NQ];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int q = 0; q < NQ; ++q) X[c][q] = 0.0;
        int q = par * jj;  // (m jj) mod np
        const int dq = (2 * jj) % np;
        for (int m = par; m < N; m += 2) {
          const double cm = tcf[q], sm = tsf[q];
          q += dq;
          if (q >= np) q -= np;
          const double m1 = (double)m, m2 = m1 * m1, m3 = m2 * m1, m4 = m2 * m2;
#pragma unroll
          for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int kt = 0; kt <= 4; ++kt) {
              const int ab = ((r * 3 + c) * 5 + kt) * N + m;
              const double A = AB[2 * ab], B = AB[2 * ab + 1];
              const double h0 = A * cm - B * sm, h1 = -A * sm - B * cm;
              X[c][qidx(kt, 0)] += h0;
              if (kt <= 3) X[c][qidx(kt, 1)] += m1 * h1;
              if (kt <= 2) X[c][qidx(kt, 2)] -= m2 * h0;
              if (kt <= 1) X[c][qidx(kt, 3)] -= m3 * h1;
              if (kt == 0) X[c][qidx(0, 4)] += m4 * h0;
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int qq = 0; qq < NQ; ++qq) {
            const double other = swap_pair(X[c][qq]);
            X[c][qq] = par ? other - X[c][qq] : X[c][qq] + other;
          }
        Geom o;
        geometry(X, o);
        if (MODE == 1) {
          double* R = o0 + node * NREF;
#pragma unroll
          for (int q = 0; q < 4; ++q) R[q] = (&o.G[0][0])[q], R[28 + q] = (&o.b[0][0])[q];
#pragma unroll
          for (int q = 0; q < 8; ++q) R[4 + q] = (&o.Gd[0][0][0])[q], R[32 + q] = (&o.bd[0][0][0])[q];
#pragma unroll
          for (int q = 0; q < 16; ++q)
            R[12 + q] = (&o.Gdd[0][0][0][0])[q], R[40 + q] = (&o.bdd[0][0][0][0])[q];
        } else {
          double f[3];
          law(o, T.ref + node * NREF, gs, eb, f);
          double* Fo = F + ((size_t)r * np + j) * 6;
          Fo[0] = f[0], Fo[1] = f[1], Fo[2] = f[2];
          Fo[3] = o.n[0], Fo[4] = o.n[1], Fo[5] = o.n[2];
        }
      }
      __syncthreads();
      if (MODE == 0 && n1 / np > rdone) {
        // rings rdone .. n1 / np - 1 are complete: longitude DFT of their six outputs, degrees
        // <= L only.  A thread sums three outputs for one (ring, m), so that each pair of
        // table loads feeds six FMAs.
        // Nodes j and j + np/2 share their factors up to (-1)^m (phi -> phi + pi), so the
        // sum runs over the first half of the ring on v_j + (-1)^m v_{j + np/2}.
        for (int t = tid; t < (n1 / np - rdone) * 2 * N; t += THREADS) {
          const int ring = rdone + t / (2 * N), f0 = 3 * ((t / N) % 2), m = t % N;
          const int r = ring % RING;
          const double sg = (m & 1) ? -1.0 : 1.0;
          const int hp = np / 2;
          double cs0 = 0.0, sn0 = 0.0, cs1 = 0.0, sn1 = 0.0, cs2 = 0.0, sn2 = 0.0;
          int q = 0;  // (m j) mod np
          #pragma unroll 4
          for (int j = 0; j < hp; ++j) {
            const double* v = F + ((size_t)r * np + j) * 6 + f0;
            const double* w = v + (size_t)hp * 6;
            const double c = tcf[q], s = tsf[q];
            q += m;
            if (q >= np) q -= np;
            const double e0 = v[0] + sg * w[0], e1 = v[1] + sg * w[1], e2 = v[2] + sg * w[2];
            cs0 += e0 * c, sn0 += e0 * s;
            cs1 += e1 * c, sn1 += e1 * s;
            cs2 += e2 * c, sn2 += e2 * s;
          }
          double* s = scr + (((size_t)ring * 6 + f0) * N + m) * 2;
          s[0] = cs0, s[1] = sn0;
          s[2 * N] = cs1, s[2 * N + 1] = sn1;
          s[4 * N] = cs2, s[4 * N + 1] = sn2;
        }
        __syncthreads();
        rdone = n1 / np;
      }
    }
    if (MODE == 1) continue;
    // 3. filter: quadrature of the 2x-grid outputs (degrees <= L), synthesis on the force
    //    grid, two fields at a time (each table load feeds both); the area from the coarse
    //    shape (sph3.f90:186-260)
    // fc: the six fields' coefficients, packed by (m, n >= m) at pk(m) + n - m with
    // pk(m) = m N - m (m - 1) / 2, [6][N (N + 1) / 2][2]; tmp: per (i, m) partial sums of the
    // current pair of fields
    double* fc = U;
    double* tmp = U + 12 * NT;   // [2][N][N][2]
    double* xt = scr;            // [nc][3] x_theta on the force grid (scratch reused)
    double* xp = scr + (size_t)nc * 3;  // [nc][3] x_phi
    if (MODE == 0) {
      // quadrature of all six fields in one pass, one item per (m, n >= m) pair (each table
      // load feeds twelve sums), 2 THREADS items per round.  The scratch rows of nq rings at
      // a time are first copied into the tmp region (all of a thread's loads in flight, one
      // wait per chunk): read in place, each iteration waited on its own scratch loads.
      const int nq = N / 3;  // 12 N nq <= 4 NN doubles
      double* stg = tmp;
      for (int t0 = 0; t0 < NT; t0 += 2 * THREADS) {
        int mm[2], nn[2];
        double sa[2][6], sb[2][6];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          int m = 0, n = min(t0 + k * THREADS + tid, NT - 1);
          while (n >= N - m) n -= N - m, ++m;
          mm[k] = m, nn[k] = n + m;
#pragma unroll
          for (int f = 0; f < 6; ++f) sa[k][f] = 0.0, sb[k][f] = 0.0;
        }
        for (int c0 = 0; c0 < nt; c0 += nq) {
          const int cr = min(nq, nt - c0);
          __syncthreads();  // the previous chunk is consumed
          const double2* src = reinterpret_cast<const double2*>(scr + (size_t)c0 * 12 * N);
          double2* dst = reinterpret_cast<double2*>(stg);
          for (int e = tid; e < cr * 6 * N; e += THREADS) dst[e] = src[e];
          __syncthreads();
          for (int r = 0; r < cr; ++r) {
            const int i = c0 + r;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
              const double p = T.w[i] * T.Pfa[((size_t)i * N + mm[k]) * N + nn[k]];
              const double* sr = stg + ((size_t)r * 6 * N + mm[k]) * 2;
#pragma unroll
              for (int f = 0; f < 6; ++f) {
                sa[k][f] += p * sr[2 * f * N];
                sb[k][f] += p * sr[2 * f * N + 1];
              }
            }
          }
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          const int t = t0 + k * THREADS + tid;
          if (t < NT)
#pragma unroll
            for (int f = 0; f < 6; ++f)
              fc[2 * (f * NT + t)] = 2.0 / np * sa[k][f], fc[2 * (f * NT + t) + 1] = -2.0 / np * sb[k][f];
        }
      }
      __syncthreads();
    }
    for (int f2 = MODE == 3 ? 6 : 0; f2 < 12; f2 += 2) {
      if (f2 < 6) {
        const double* fa = fc + 2 * f2 * NT;
        const double* fb = fa + 2 * NT;
        for (int t = tid; t < NN; t += THREADS) {
          const int i = t / N, m = t % N;
          const double h = m == 0 ? 0.5 : 1.0;
          const int pk = m * N - m * (m - 1) / 2 - m;  // + n
          double sa0 = 0.0, sb0 = 0.0, sa1 = 0.0, sb1 = 0.0;
          #pragma unroll 4
          for (int n = m; n < N; ++n) {
            const double p = T.Pc[(i * N + m) * N + n];
            const int u = pk + n;
            sa0 += p * fa[2 * u];
            sb0 += p * fa[2 * u + 1];
            sa1 += p * fb[2 * u];
            sb1 += p * fb[2 * u + 1];
          }
          tmp[2 * t] = h * sa0, tmp[2 * t + 1] = h * sb0;
          tmp[2 * (NN + t)] = h * sa1, tmp[2 * (NN + t) + 1] = h * sb1;
        }
      } else {
        // fi 6..8: x_theta, fi 9..11: x_phi of coordinate c, from the coarse shape
        const int c0 = f2 < 9 ? f2 - 6 : f2 - 9, c1 = f2 + 1 < 9 ? f2 - 5 : f2 - 8;
        const double* P0 = f2 < 9 ? T.Pcd : T.Pc;
        const double* P1 = f2 + 1 < 9 ? T.Pcd : T.Pc;
        for (int t = tid; t < NN; t += THREADS) {
          const int i = t / N, m = t % N;
          const double h = m == 0 ? 0.5 : 1.0;
          const double* Q0 = P0 + (i * N + m) * N;
          const double* Q1 = P1 + (i * N + m) * N;
          double sa0 = 0.0, sb0 = 0.0, sa1 = 0.0, sb1 = 0.0;
          #pragma unroll 4
          for (int n = m; n < N; ++n) {
            sa0 += Q0[n] * ca[(c0 * N + m) * N + n];
            sb0 += Q0[n] * cb[(c0 * N + m) * N + n];
            sa1 += Q1[n] * ca[(c1 * N + m) * N + n];
            sb1 += Q1[n] * cb[(c1 * N + m) * N + n];
          }
          tmp[2 * t] = h * sa0, tmp[2 * t + 1] = h * sb0;
          tmp[2 * (NN + t)] = h * sa1, tmp[2 * (NN + t) + 1] = h * sb1;
        }
      }
      __syncthreads();
      // the 2x-grid longitude DFT partials are dead once f2 >= 6: x_theta / x_phi reuse them.
      // One thread per node (i, j < np2/2) and its partner j + np2/2, whose factors are the
      // same up to (-1)^m (phi -> phi + pi).
      const int hp2 = np2 / 2;
      for (int t = tid; t < N * hp2; t += THREADS) {
        const int i = t / hp2, j = t % hp2;
        double v0 = 0.0, v1 = 0.0, w0 = 0.0, w1 = 0.0;  // v: node j, w: node j + np2/2
        int q = 0;  // (m j) mod np2
        #pragma unroll 4
        for (int m = 0; m < N; ++m) {
          const double A0 = tmp[2 * (i * N + m)], B0 = tmp[2 * (i * N + m) + 1];
          const double A1 = tmp[2 * (NN + i * N + m)], B1 = tmp[2 * (NN + i * N + m) + 1];
          const double cm = tc2[q], sm = ts2[q], sg = (m & 1) ? -1.0 : 1.0;
          q += j;
          if (q >= np2) q -= np2;
          const double t0 = f2 < 9 ? A0 * cm - B0 * sm : (double)m * (-A0 * sm - B0 * cm);
          const double t1 = f2 + 1 < 9 ? A1 * cm - B1 * sm : (double)m * (-A1 * sm - B1 * cm);
          v0 += t0, w0 += sg * t0;
          v1 += t1, w1 += sg * t1;
        }
        const double v[4] = {v0, v1, w0, w1};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int fi = f2 + (e & 1);
          const int node = i * np2 + j + (e >> 1) * hp2;
          if (fi < 6) {
            double* dst = fi == 0 ? o0 : fi == 1 ? o1 : fi == 2 ? o2 : fi == 3 ? o3 : fi == 4 ? o4 : o5;
            dst[(size_t)cap * nc + node] = v[e];
          } else if (fi < 9) {
            xt[(size_t)node * 3 + (fi - 6)] = v[e];
          } else {
            xp[(size_t)node * 3 + (fi - 9)] = v[e];
          }
        }
      }
      __syncthreads();
    }
    // unit normal (cal_unitO_filter, sph1.f90:1378-1406) and point area
    // surfmet * w_i / sin(theta_i) * dphi (forcing.f90:91; sph.f90:131)
    for (int node = tid; node < nc; node += THREADS) {
      const size_t o = (size_t)cap * nc + node;
      double cr[3];
      cross3(xt + (size_t)node * 3, xp + (size_t)node * 3, cr);
      const double J = sqrt(dot3(cr, cr));
      if (MODE == 0) {
        const double a = o3[o], b = o4[o], c = o5[o];
        const double r = sqrt(a * a + b * b + c * c);
        o3[o] = a / r, o4[o] = b / r, o5[o] = c / r;
      } else {
        // MODE 3: the unit normal of the force grid itself, (x_t x x_p) / |x_t x x_p|
        o3[o] = cr[0] / J, o4[o] = cr[1] / J, o5[o] = cr[2] / J;
      }
      const int i = node / np2;
      o6[o] = J * T.w2[i] / T.s2[i] * (2.0 * M_PI / np2);
    }
    __syncthreads();
  }
}

// ---- host tables ------------------------------------------------------------------------
void gauss(int n, double* theta, double* w) {
  for (int i = 0; i < n; ++i) {
    double x = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = x;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1, p1 = p2;
      }
      const double pn = n == 1 ? x : p1, pm = n == 1 ? 1.0 : p0;
      dp = n * (x * pn - pm) / (x * x - 1.0);
      const double dx = pn / dp;
      x -= dx;
      if (std::fabs(dx) < 1e-16) {
        p0 = 1.0, p1 = x;
        for (int k = 2; k <= n; ++k) {
          const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
          p0 = p1, p1 = p2;
        }
        const double qn = n == 1 ? x : p1, qm = n == 1 ? 1.0 : p0;
        dp = n * (x * qn - qm) / (x * x - 1.0);
        break;
      }
    }
    theta[i] = std::acos(x);
    w[i] = 2.0 / ((1.0 - x * x) * dp * dp);
  }
}

// Pbar_n^m(cos t), m <= n < nmax: P[m * nmax + n]
void legendre(int nmax, double t, double* P) {
  const double c = std::cos(t), s = std::sin(t);
  std::fill(P, P + nmax * nmax, 0.0);
  double pmm = std::sqrt(0.5);
  for (int m = 0; m < nmax; ++m) {
    if (m > 0) pmm *= std::sqrt((2.0 * m + 1.0) / (2.0 * m)) * s;
    P[m * nmax + m] = pmm;
    if (m + 1 < nmax) P[m * nmax + m + 1] = std::sqrt(2.0 * m + 3.0) * c * pmm;
    for (int n = m + 2; n < nmax; ++n) {
      const double a = std::sqrt((4.0 * n * n - 1.0) / ((double)n * n - (double)m * m));
      const double a1 =
          std::sqrt((4.0 * (n - 1.0) * (n - 1.0) - 1.0) / ((n - 1.0) * (n - 1.0) - (double)m * m));
      P[m * nmax + n] = a * (c * P[m * nmax + n - 1] - P[m * nmax + n - 2] / a1);
    }
  }
}

// theta-derivatives 0..4 by the ladder d/dt Pbar^m = (c-_m Pbar^{m-1} - c+_m Pbar^{m+1}) / 2
void legendre_derivs(int nmax, double t, double* D /* [5][nmax][nmax] */) {
  std::vector<double> P(nmax * nmax);
  legendre(nmax, t, P.data());
  std::fill(D, D + 5 * nmax * nmax, 0.0);
  std::vector<double> v(5 * (2 * nmax + 3));
  for (int n = 0; n < nmax; ++n) {
    const int off = n + 1, W = 2 * nmax + 3;
    std::fill(v.begin(), v.end(), 0.0);
    for (int m = -n; m <= n; ++m) {
      const int am = m < 0 ? -m : m;
      v[m + off] = ((m < 0 && (am & 1)) ? -1.0 : 1.0) * P[am * nmax + n];
    }
    for (int k = 1; k < 5; ++k)
      for (int m = -n; m <= n; ++m) {
        const double cm = std::sqrt((double)(n + m) * (double)(n - m + 1));
        const double cp = std::sqrt((double)(n + m + 1) * (double)(n - m));
        v[k * W + m + off] = 0.5 * (cm * v[(k - 1) * W + m - 1 + off] - cp * v[(k - 1) * W + m + 1 + off]);
      }
    for (int k = 0; k < 5; ++k)
      for (int m = 0; m <= n; ++m) D[(k * nmax + m) * nmax + n] = v[k * W + m + off];
  }
}

}  // namespace memb

using namespace memb;

struct memb_plan {
  int device = 0, L = 0;
  hipStream_t own = nullptr, stream = nullptr;
  std::string err;
  Tables T{};
  std::vector<void*> bufs;  // device allocations of the tables
  double* ref = nullptr;
  double* scratch = nullptr;
  size_t scratch_wg = 0;
  int num_wg = 0;
  size_t lds = 0;
};

namespace {
int fail(memb_plan* p, int code, const std::string& m) {
  if (p) p->err = m;
  return code;
}
#define MCHK(p, call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail((p), MEMB_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
  } while (0)

int upload(memb_plan* p, const std::vector<double>& h, const double** dst) {
  void* d = nullptr;
  MCHK(p, hipMalloc(&d, h.size() * sizeof(double)));
  p->bufs.push_back(d);
  MCHK(p, hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  *dst = (const double*)d;
  return MEMB_OK;
}

// ---- co-rotated frame of the surrogate ---------------------------------------------------------
// The law is invariant under translation and covariant under rotation of the current shape
// (f(R x + c) = R f(x)): its load depends on the capsule's deformation only.  One workgroup
// per capsule: centroid c, the cross-covariance H = sum (x_p - c) X_p^T against the
// stress-free points X, the rotation R = the orthogonal polar factor of H (the rotation that
// brings X closest to the shape, Kabsch), the displacement delta_p = R^T (x_p - c) - X_p, its
// RMS amplitude s and the direction delta / s.
constexpr int CR_THREADS = 256;

__device__ inline void inv3t(const double A[3][3], double B[3][3], double& det) {
  // B = inverse(A)^T
  const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
  const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
  const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
  det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
  const double id = 1.0 / det;
  B[0][0] = c00 * id, B[0][1] = c01 * id, B[0][2] = c02 * id;
  B[1][0] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
  B[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
  B[1][2] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
  B[2][0] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
  B[2][1] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
  B[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
}

__global__ void __launch_bounds__(CR_THREADS)
    k_corot(int n, const double* __restrict__ X, const double* __restrict__ x,
            const double* __restrict__ y, const double* __restrict__ z,
            double* __restrict__ delta, double* __restrict__ dhx, double* __restrict__ dhy,
            double* __restrict__ dhz, double* __restrict__ rot, double* __restrict__ amp) {
  __shared__ double red[12][CR_THREADS / 64];
  __shared__ double sh[12];
  const long long cap = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const double* px = x + cap * n;
  const double* py = y + cap * n;
  const double* pz = z + cap * n;
  auto block_sum = [&](double* v, int k) {  // sums v[0..k) over the workgroup into sh
    for (int a = 0; a < k; ++a) {
      double t = v[a];
      for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o);
      if (lane == 0) red[a][wv] = t;
    }
    __syncthreads();
    if (tid < k) {
      double t = 0.0;
      for (int w = 0; w < CR_THREADS / 64; ++w) t += red[tid][w];
      sh[tid] = t;
    }
    __syncthreads();
  };
  double v[12];
  v[0] = v[1] = v[2] = 0.0;
  for (int p = tid; p < n; p += CR_THREADS) v[0] += px[p], v[1] += py[p], v[2] += pz[p];
  block_sum(v, 3);
  const double c[3] = {sh[0] / n, sh[1] / n, sh[2] / n};
  __syncthreads();
  for (int a = 0; a < 9; ++a) v[a] = 0.0;
  for (int p = tid; p < n; p += CR_THREADS) {
    const double d[3] = {px[p] - c[0], py[p] - c[1], pz[p] - c[2]};
    const double r[3] = {X[p], X[n + p], X[2 * n + p]};
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) v[3 * a + b] += d[a] * r[b];
  }
  block_sum(v, 9);
  __shared__ double R[3][3];
  __shared__ int ok;
  if (tid == 0) {
    // polar factor by Newton's iteration Y <- (Y + Y^-T) / 2 from the scaled H
    double Y[3][3], Z[3][3], det = 0.0, nrm = 0.0;
    for (int a = 0; a < 9; ++a) nrm += sh[a] * sh[a];
    nrm = sqrt(nrm / 3.0);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) Y[a][b] = nrm > 0.0 ? sh[3 * a + b] / nrm : (a == b ? 1.0 : 0.0);
    for (int it = 0; it < 24; ++it) {
      inv3t(Y, Z, det);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) Y[a][b] = 0.5 * (Y[a][b] + Z[a][b]);
    }
    inv3t(Y, Z, det);
    ok = isfinite(det) && det > 0.5;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) R[a][b] = Y[a][b], rot[cap * 9 + 3 * a + b] = Y[a][b];
  }
  __syncthreads();
  // delta = R^T (x - c) - X, layout [cap][3][n]
  double ss = 0.0;
  for (int p = tid; p < n; p += CR_THREADS) {
    const double d[3] = {px[p] - c[0], py[p] - c[1], pz[p] - c[2]};
    for (int a = 0; a < 3; ++a) {
      const double q = R[0][a] * d[0] + R[1][a] * d[1] + R[2][a] * d[2] - X[a * n + p];
      delta[(cap * 3 + a) * n + p] = q;
      ss += q * q;
    }
  }
  v[0] = ss;
  block_sum(v, 1);
  const double s = sqrt(sh[0] / n);
  const double inv = s > 0.0 ? 1.0 / s : 0.0;
  for (int p = tid; p < n; p += CR_THREADS) {
    dhx[cap * n + p] = delta[(cap * 3 + 0) * n + p] * inv;
    dhy[cap * n + p] = delta[(cap * 3 + 1) * n + p] * inv;
    dhz[cap * n + p] = delta[(cap * 3 + 2) * n + p] * inv;
  }
  if (tid == 0) amp[cap] = ok ? s : __builtin_inf();
}

// f = R (flin + s^2 resid) per point, into the capsule-major traction arrays
__global__ void __launch_bounds__(256)
    k_corot_apply(long long ncap, int n, const double* __restrict__ flin,
                  const float* __restrict__ resid, const double* __restrict__ rot,
                  const double* __restrict__ amp, double* __restrict__ fx,
                  double* __restrict__ fy, double* __restrict__ fz) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncap * n) return;
  const long long cap = q / n;
  const int p = (int)(q - cap * n);
  const double s2 = amp[cap] * amp[cap];
  double g[3];
  for (int a = 0; a < 3; ++a) {
    const long long o = (cap * 3 + a) * n + p;
    g[a] = flin[o] + (resid ? s2 * (double)resid[o] : 0.0);
  }
  const double* R = rot + cap * 9;
  fx[q] = R[0] * g[0] + R[1] * g[1] + R[2] * g[2];
  fy[q] = R[3] * g[0] + R[4] * g[1] + R[5] * g[2];
  fz[q] = R[6] * g[0] + R[7] * g[1] + R[8] * g[2];
}

// The anchored surrogate's prediction and gate, one workgroup per capsule (rows of 3n values,
// component-major: [c][3][n]):
//   cc = <delta, delta0> / |delta0|^2 (0 when delta0 = 0), drift = sqrt(sum (delta - delta0)^2 / n),
//   f' = (lin + cc^2 m0) + b (m0 = NULL: f' = lin + b), f = R f' into fx, fy, fz,
//   sel = 1 unless beta drift <= tol and every f is finite; drift_out (optional).
// pre (optional): capsules with pre[c] != 0 (already rejected by the drift test, k_anch_gate)
// are left to the law: sel[c] = 1, nothing else written.  late (optional): 1 for the other
// capsules this kernel rejects (their prediction is not finite), else 0.
__global__ void __launch_bounds__(CR_THREADS)
    k_anch_predict(int n, const double* __restrict__ lin, const double* __restrict__ m0,
                   const double* __restrict__ b, const double* __restrict__ delta,
                   const double* __restrict__ delta0, const double* __restrict__ rot,
                   double beta, double tol, double* __restrict__ fx, double* __restrict__ fy,
                   double* __restrict__ fz, int* __restrict__ sel,
                   double* __restrict__ drift_out, const int* __restrict__ pre,
                   int* __restrict__ late) {
  __shared__ double red[3][CR_THREADS / 64];
  __shared__ int bad;
  const long long cap = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const long long r0 = cap * 3 * n;
  if (pre && pre[cap]) {
    if (tid == 0) {
      sel[cap] = 1;
      if (late) late[cap] = 0;
    }
    return;
  }
  if (tid == 0) bad = 0;
  double v[3] = {0.0, 0.0, 0.0};
  for (int q = tid; q < 3 * n; q += CR_THREADS) {
    const double d = delta[r0 + q], d0 = delta0[r0 + q], e = d - d0;
    v[0] += d * d0, v[1] += d0 * d0, v[2] += e * e;
  }
  for (int a = 0; a < 3; ++a) {
    double t = v[a];
    for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o);
    if (lane == 0) red[a][wv] = t;
  }
  __syncthreads();
  double sdd0 = 0.0, s00 = 0.0, sdr = 0.0;
  for (int w = 0; w < CR_THREADS / 64; ++w) sdd0 += red[0][w], s00 += red[1][w], sdr += red[2][w];
  const double c = s00 > 0.0 ? sdd0 / s00 : 0.0, c2 = c * c;
  const double drift = sqrt(sdr / (double)n);
  const double* R = rot + cap * 9;
  int nonfinite = 0;
  for (int p = tid; p < n; p += CR_THREADS) {
    double g[3];
    for (int a = 0; a < 3; ++a) {
      const long long o = r0 + (long long)a * n + p;
      g[a] = (m0 ? lin[o] + c2 * m0[o] : lin[o]) + b[o];
    }
    const long long q = cap * n + p;
    const double ox = R[0] * g[0] + R[1] * g[1] + R[2] * g[2];
    const double oy = R[3] * g[0] + R[4] * g[1] + R[5] * g[2];
    const double oz = R[6] * g[0] + R[7] * g[1] + R[8] * g[2];
    fx[q] = ox, fy[q] = oy, fz[q] = oz;
    nonfinite |= !(isfinite(ox) && isfinite(oy) && isfinite(oz));
  }
  if (nonfinite) atomicOr(&bad, 1);
  __syncthreads();
  if (tid == 0) {
    sel[cap] = (beta * drift <= tol && !bad) ? 0 : 1;
    if (late) late[cap] = sel[cap];
    if (drift_out) drift_out[cap] = drift;
  }
}

// The drift test of k_anch_predict alone (the same sums in the same order, so the same
// drift): pre[c] = 1 unless beta rms(delta - delta0) <= tol.  The law can then start on these
// capsules while the linear response is computed for the others.
__global__ void __launch_bounds__(CR_THREADS)
    k_anch_gate(int n, const double* __restrict__ delta, const double* __restrict__ delta0,
                double beta, double tol, int* __restrict__ pre) {
  __shared__ double red[CR_THREADS / 64];
  const long long cap = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const long long r0 = cap * 3 * n;
  double v = 0.0;
  for (int q = tid; q < 3 * n; q += CR_THREADS) {
    const double e = delta[r0 + q] - delta0[r0 + q];
    v += e * e;
  }
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  if (lane == 0) red[wv] = v;
  __syncthreads();
  if (tid == 0) {
    double sdr = 0.0;
    for (int w = 0; w < CR_THREADS / 64; ++w) sdr += red[w];
    const double drift = sqrt(sdr / (double)n);
    pre[cap] = beta * drift <= tol ? 0 : 1;
  }
}

// New anchors for the capsules with sel != 0 (the law has just written their traction):
//   b = R^T f - lin (- m_now), delta0 = delta, m0 = m_now, s0 = amp, refreshed += 1.
// slot (optional): m_now holds rows only for the capsules with slot[c] >= 0 (row slot[c]); a
// selected capsule without a row keeps its old anchor (its traction this step is the law's
// all the same, and it is selected again at the next step).
__global__ void __launch_bounds__(CR_THREADS)
    k_anch_update(int n, const int* __restrict__ sel, const int* __restrict__ slot,
                  const double* __restrict__ rot, const double* __restrict__ fx,
                  const double* __restrict__ fy, const double* __restrict__ fz,
                  const double* __restrict__ lin, const double* __restrict__ m_now,
                  const double* __restrict__ delta, const double* __restrict__ amp,
                  double* __restrict__ b, double* __restrict__ delta0, double* __restrict__ m0,
                  double* __restrict__ s0, long long* __restrict__ refreshed) {
  const long long cap = blockIdx.x;
  if (sel[cap] == 0) return;
  const int tid = threadIdx.x;
  if (tid == 0) refreshed[cap] += 1;
  const long long row = slot ? (long long)slot[cap] : cap;
  if (row < 0) return;
  const long long r0 = cap * 3 * n, m_r0 = row * 3 * n;
  const double* R = rot + cap * 9;
  for (int p = tid; p < n; p += CR_THREADS) {
    const long long q = cap * n + p;
    const double f0 = fx[q], f1 = fy[q], f2 = fz[q];
    const double g[3] = {R[0] * f0 + R[3] * f1 + R[6] * f2, R[1] * f0 + R[4] * f1 + R[7] * f2,
                         R[2] * f0 + R[5] * f1 + R[8] * f2};
    for (int a = 0; a < 3; ++a) {
      const long long o = r0 + (long long)a * n + p;
      double bl = g[a] - lin[o];
      if (m_now) {
        const double mn = m_now[m_r0 + (long long)a * n + p];
        bl = bl - mn;
        m0[o] = mn;
      }
      b[o] = bl;
      delta0[o] = delta[o];
    }
  }
  if (tid == 0 && m_now && s0) s0[cap] = amp[cap];
}

// The first k capsules with sel != 0, in capsule order: idx[r] = c (padded with 0) and
// slot[c] = r, or -1 for the others and for selected capsules beyond k.  One workgroup.
constexpr int CS_THREADS = 1024;
__global__ void __launch_bounds__(CS_THREADS)
    k_compact_sel(long long ncap, const int* __restrict__ sel, int k, int* __restrict__ idx,
                  int* __restrict__ slot) {
  __shared__ int wsum[CS_THREADS / 64];
  __shared__ int base;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (tid == 0) base = 0;
  for (int r = tid; r < k; r += CS_THREADS) idx[r] = 0;
  __syncthreads();
  for (long long c0 = 0; c0 < ncap; c0 += CS_THREADS) {
    const long long c = c0 + tid;
    const bool on = c < ncap && sel[c] != 0;
    const unsigned long long bal = __ballot(on);
    if (lane == 0) wsum[wv] = __popcll(bal);
    __syncthreads();
    int before = 0, total = 0;
    for (int w = 0; w < CS_THREADS / 64; ++w) {
      before += w < wv ? wsum[w] : 0;
      total += wsum[w];
    }
    const int r = base + before + __popcll(bal & __lanemask_lt());
    if (c < ncap) {
      const bool kept = on && r < k;
      slot[c] = kept ? r : -1;
      if (kept) idx[r] = (int)c;
    }
    __syncthreads();
    if (tid == 0) base += total;
    __syncthreads();
  }
}

// Geometry only (unit normal and point area on the force grid) with the arithmetic of
// k_mem<3> -- analysis of the three coordinates, x_theta and x_phi synthesised from the coarse
// coefficients, n = (x_t x x_p) / |x_t x x_p|, area = |x_t x x_p| w_i / sin(theta_i) dphi
// (forcing.f90:91; sph.f90:131) -- laid out for throughput rather than for the law's
// registers: 256 threads per capsule, the longitude tables staged in LDS once per workgroup,
// the six (i, m) partial sums of x_theta and x_phi in one pass, and one thread per node for
// all six node sums, the cross product, the normal and the area.  Persistent over capsules.
constexpr int GM_THREADS = 256;
__global__ void __launch_bounds__(GM_THREADS)
    k_geom(Tables T, long long ncap, const double* __restrict__ x, const double* __restrict__ y,
           const double* __restrict__ z, double* __restrict__ nx, double* __restrict__ ny,
           double* __restrict__ nz, double* __restrict__ area, const int* __restrict__ skip) {
  extern __shared__ double lds[];
  const int tid = threadIdx.x;
  const int N = T.N, NN = N * N, np2 = T.np2, nc = N * np2, NT = N * (N + 1) / 2, hp = np2 / 2;
  // cos / sin(2 pi q / np2), q < np2: the factor cos / sin(m phi_j) is entry (m j) mod np2, and
  // nodes j and j + np2/2 share their factors up to (-1)^m (phi -> phi + pi)
  double* tc = lds;
  double* ts = tc + np2;
  double* ca = ts + np2;            // [3][N][N]
  double* cb = ca + 3 * NN;         // [3][N][N]
  double* tmp = cb + 3 * NN;        // [6][N][N][2]: x_theta (0..2), x_phi (3..5) partial sums
  double* pts = tmp;                // [3][nc] the capsule's points and
  double* U = tmp + 3 * nc;         // [3][N][N][2] their longitude DFT, both dead before tmp
  for (int q = tid; q < np2; q += GM_THREADS) tc[q] = T.cs2[q * N + 1], ts[q] = T.sn2[q * N + 1];
  for (long long cap = blockIdx.x; cap < ncap; cap += gridDim.x) {
    if (skip && skip[cap]) continue;  // uniform over the workgroup
    // the capsule's points into LDS with coalesced loads
    for (int t = tid; t < 3 * nc; t += GM_THREADS) {
      const int c = t / nc, q = t - c * nc;
      pts[t] = (c == 0 ? x : (c == 1 ? y : z))[cap * nc + q];
    }
    __syncthreads();
    // longitude DFT, one thread per (i, m) for the three coordinates, every point in order
    // (the raw coordinates are not paired by the phi -> phi + pi symmetry, as in the law)
    for (int t = tid; t < NN; t += GM_THREADS) {
      const int i = t / N, m = t % N;
      double cs[3] = {0.0, 0.0, 0.0}, sn[3] = {0.0, 0.0, 0.0};
      int q = 0;  // (m j) mod np2
      #pragma unroll 4
      for (int j = 0; j < np2; ++j) {
        const double c = tc[q], s = ts[q];
        q += m;
        if (q >= np2) q -= np2;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double v = pts[k * nc + i * np2 + j];
          cs[k] += v * c;
          sn[k] += v * s;
        }
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) U[2 * (k * NN + t)] = cs[k], U[2 * (k * NN + t) + 1] = sn[k];
    }
    __syncthreads();
    // quadrature, one thread per (m, n >= m) for the three coordinates (n < m is never read)
    for (int t = tid; t < NT; t += GM_THREADS) {
      int m = 0, n = t;
      while (n >= N - m) n -= N - m, ++m;
      n += m;
      double sa[3] = {0.0, 0.0, 0.0}, sb[3] = {0.0, 0.0, 0.0};
      #pragma unroll 4
      for (int i = 0; i < N; ++i) {
        const double p = T.w2[i] * T.Pc[(i * N + m) * N + n];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int u = (k * N + i) * N + m;
          sa[k] += p * U[2 * u];
          sb[k] += p * U[2 * u + 1];
        }
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        ca[(k * N + m) * N + n] = 2.0 / np2 * sa[k];
        cb[(k * N + m) * N + n] = -2.0 / np2 * sb[k];
      }
    }
    __syncthreads();
    // x_theta and x_phi of coordinate c per (c, i, m): both from the same coefficients
    for (int t = tid; t < 3 * NN; t += GM_THREADS) {
      const int c = t / NN, u = t % NN, i = u / N, m = u % N;
      const double h = m == 0 ? 0.5 : 1.0;
      const double* Pt = T.Pcd + (i * N + m) * N;
      const double* Pp = T.Pc + (i * N + m) * N;
      double sa0 = 0.0, sb0 = 0.0, sa1 = 0.0, sb1 = 0.0;
      #pragma unroll 4
      for (int n = m; n < N; ++n) {
        const double a = ca[(c * N + m) * N + n], b = cb[(c * N + m) * N + n];
        sa0 += Pt[n] * a;
        sb0 += Pt[n] * b;
        sa1 += Pp[n] * a;
        sb1 += Pp[n] * b;
      }
      tmp[2 * t] = h * sa0, tmp[2 * t + 1] = h * sb0;
      tmp[2 * (3 * NN + t)] = h * sa1, tmp[2 * (3 * NN + t) + 1] = h * sb1;
    }
    __syncthreads();
    // node sums, one thread per (i, j < np2/2) and its partner j + np2/2; normal and area
    for (int t = tid; t < N * hp; t += GM_THREADS) {
      const int i = t / hp, j = t % hp;
      double d[2][6];
#pragma unroll
      for (int f = 0; f < 6; ++f) d[0][f] = 0.0, d[1][f] = 0.0;
      int q = 0;  // (m j) mod np2
      #pragma unroll 2
      for (int m = 0; m < N; ++m) {
        const double cm = tc[q], sm = ts[q], sg = (m & 1) ? -1.0 : 1.0;
        q += j;
        if (q >= np2) q -= np2;
#pragma unroll
        for (int f = 0; f < 6; ++f) {
          const double A = tmp[2 * (f * NN + i * N + m)], B = tmp[2 * (f * NN + i * N + m) + 1];
          const double e = f < 3 ? A * cm - B * sm : (double)m * (-A * sm - B * cm);
          d[0][f] += e;
          d[1][f] += sg * e;
        }
      }
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        double cr[3];
        cross3(&d[e][0], &d[e][3], cr);
        const double J = sqrt(dot3(cr, cr));
        const size_t o = (size_t)cap * nc + i * np2 + j + e * hp;
        nx[o] = cr[0] / J, ny[o] = cr[1] / J, nz[o] = cr[2] / J;
        area[o] = J * T.w2[i] / T.s2[i] * (2.0 * M_PI / np2);
      }
    }
    __syncthreads();
  }
}

size_t geom_lds_bytes(const Tables& T) {
  const int N = T.N, NN = N * N;
  // the two longitude tables, the coefficients, and the partial sums (which hold the points
  // and their longitude DFT before)
  return (size_t)(2 * T.np2 + 6 * NN + std::max(12 * NN, 3 * N * T.np2 + 6 * NN)) * sizeof(double);
}

size_t lds_bytes(const Tables& T) {
  const int N = T.N, NN = N * N;
  size_t work = std::max<size_t>(3 * NN * 2, (size_t)RING * 3 * 5 * N * 2 + (size_t)RING * T.np * 6);
  // phase 3: the six fields' packed coefficients and the partial sums of two fields
  work = std::max<size_t>(work, (size_t)6 * N * (N + 1) + 4 * NN);
  return (6 * NN + 2 * (T.np + T.np2) + work) * sizeof(double);
}

template <int MODE>
int launch(memb_plan* p, long long ncap, double gs, double eb, const double* x, const double* y,
           const double* z, double* o0, double* o1, double* o2, double* o3, double* o4,
           double* o5, double* o6, const int* sel = nullptr) {
  const int nwg = (int)std::min<long long>(ncap, p->num_wg);
  if (nwg <= 0) return MEMB_OK;
  k_mem<MODE><<<nwg, THREADS, p->lds, p->stream>>>(p->T, ncap, gs, eb, x, y, z, o0, o1, o2, o3,
                                                   o4, o5, o6, p->scratch, sel);
  MCHK(p, hipGetLastError());
  return MEMB_OK;
}
}  // namespace

extern "C" {

int memb_plan_create(memb_plan** out, int L, int device) {
  if (!out || L < 2 || L > LMAX) return MEMB_E_ARG;
  *out = nullptr;
  memb_plan* p = new memb_plan();
  p->device = device, p->L = L;
  int rc;
  auto bail = [&](int code) {
    *out = p;  // keep the message readable through memb_last_error
    return code;
  };
  if (hipSetDevice(device) != hipSuccess) return bail(fail(p, MEMB_E_HIP, "hipSetDevice failed"));
  if (hipStreamCreate(&p->own) != hipSuccess)
    return bail(fail(p, MEMB_E_HIP, "hipStreamCreate failed"));
  p->stream = p->own;
  Tables& T = p->T;
  T.L = L, T.N = L + 1, T.np2 = 2 * (L + 1), T.nt = 2 * T.N, T.np = 2 * T.np2;
  const int N = T.N, nt = T.nt, np = T.np, np2 = T.np2;
  std::vector<double> th2(N), w2(N), th(nt), w(nt), s2(N);
  gauss(N, th2.data(), w2.data());
  gauss(nt, th.data(), w.data());
  for (int i = 0; i < N; ++i) s2[i] = std::sin(th2[i]);
  std::vector<double> Pc((size_t)N * N * N), Pcd((size_t)N * N * N), Pfd((size_t)nt * 5 * N * N),
      Pfa((size_t)nt * N * N), D(5 * N * N);
  for (int i = 0; i < N; ++i) {
    legendre_derivs(N, th2[i], D.data());
    std::copy(D.begin(), D.begin() + N * N, Pc.begin() + (size_t)i * N * N);
    std::copy(D.begin() + N * N, D.begin() + 2 * N * N, Pcd.begin() + (size_t)i * N * N);
  }
  for (int i = 0; i < nt; ++i) {
    legendre_derivs(N, th[i], D.data());
    std::copy(D.begin(), D.end(), Pfd.begin() + (size_t)i * 5 * N * N);
    std::copy(D.begin(), D.begin() + N * N, Pfa.begin() + (size_t)i * N * N);
  }
  std::vector<double> cs2((size_t)np2 * N), sn2((size_t)np2 * N), csf((size_t)np * N),
      snf((size_t)np * N);
  // cos / sin(m phi_j) with the angle reduced first, 2 pi ((m j) mod n) / n, as the longitude
  // FFT's twiddle factors are formed (hrfft.f:214-267): exact to the last bit of cos / sin
  // of an angle below 2 pi, where m phi_j itself (up to 2 pi L) carries 1e-14 of rounding
  // that the fourth phi derivatives multiply by m^4
  for (int j = 0; j < np2; ++j)
    for (int m = 0; m < N; ++m) {
      const double ph = 2.0 * M_PI * ((m * j) % np2) / np2;
      cs2[j * N + m] = std::cos(ph), sn2[j * N + m] = std::sin(ph);
    }
  for (int j = 0; j < np; ++j)
    for (int m = 0; m < N; ++m) {
      const double ph = 2.0 * M_PI * ((m * j) % np) / np;
      csf[j * N + m] = std::cos(ph), snf[j * N + m] = std::sin(ph);
    }
  if ((rc = upload(p, w2, &T.w2)) || (rc = upload(p, w, &T.w)) || (rc = upload(p, s2, &T.s2)) ||
      (rc = upload(p, Pc, &T.Pc)) || (rc = upload(p, Pcd, &T.Pcd)) ||
      (rc = upload(p, Pfd, &T.Pfd)) || (rc = upload(p, Pfa, &T.Pfa)) ||
      (rc = upload(p, cs2, &T.cs2)) || (rc = upload(p, sn2, &T.sn2)) ||
      (rc = upload(p, csf, &T.csf)) || (rc = upload(p, snf, &T.snf)))
    return bail(rc);
  p->lds = lds_bytes(T);
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  p->num_wg = 4 * cus;  // four capsule workgroups per CU (LDS and the register cap allow it)
  // per workgroup: longitude DFT of the 2x-grid outputs, and the force-grid basis vectors
  p->scratch_wg = std::max<size_t>((size_t)nt * 6 * N * 2, (size_t)N * np2 * 6);
  void* s = nullptr;
  if (hipMalloc(&s, p->scratch_wg * p->num_wg * sizeof(double)) != hipSuccess)
    return bail(fail(p, MEMB_E_HIP, "hipMalloc scratch failed"));
  p->bufs.push_back(s);
  p->scratch = (double*)s;
  // the reference state: the unit sphere through the same kernel (sph.f90:56-70, 152-161)
  void* r = nullptr;
  if (hipMalloc(&r, (size_t)nt * np * NREF * sizeof(double)) != hipSuccess)
    return bail(fail(p, MEMB_E_HIP, "hipMalloc reference state failed"));
  p->bufs.push_back(r);
  p->ref = (double*)r;
  std::vector<double> sx((size_t)N * np2 * 3);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < np2; ++j) {
      const double ph = 2.0 * M_PI * j / np2;
      sx[i * np2 + j] = std::sin(th2[i]) * std::sin(ph);
      sx[N * np2 + i * np2 + j] = std::sin(th2[i]) * std::cos(ph);
      sx[2 * N * np2 + i * np2 + j] = std::cos(th2[i]);
    }
  const double* dsx = nullptr;
  if ((rc = upload(p, sx, &dsx))) return bail(rc);
  T.ref = nullptr;
  // scratch stride is per workgroup; one capsule uses workgroup 0's region
  if (hipFuncSetAttribute((const void*)k_mem<0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)p->lds) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_mem<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)p->lds) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_mem<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)p->lds) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_mem<3>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)p->lds) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_geom, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)geom_lds_bytes(p->T)) != hipSuccess)
    return bail(fail(p, MEMB_E_HIP, "hipFuncSetAttribute failed"));
  if ((rc = launch<1>(p, 1, 0.0, 0.0, dsx, dsx + N * np2, dsx + 2 * N * np2, p->ref, nullptr,
                      nullptr, nullptr, nullptr, nullptr, nullptr)))
    return bail(rc);
  if (hipStreamSynchronize(p->stream) != hipSuccess)
    return bail(fail(p, MEMB_E_HIP, "reference state failed"));
  T.ref = p->ref;
  *out = p;
  return MEMB_OK;
}

void memb_plan_destroy(memb_plan* p) {
  if (!p) return;
  (void)hipSetDevice(p->device);
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  for (void* b : p->bufs) (void)hipFree(b);
  if (p->own) (void)hipStreamDestroy(p->own);
  delete p;
}

const char* memb_last_error(const memb_plan* p) { return p ? p->err.c_str() : "null plan"; }

int memb_set_stream(memb_plan* p, void* s) {
  if (!p) return MEMB_E_ARG;
  p->stream = s ? (hipStream_t)s : p->own;
  return MEMB_OK;
}

int memb_sync(memb_plan* p) {
  if (!p) return MEMB_E_ARG;
  MCHK(p, hipSetDevice(p->device));
  MCHK(p, hipStreamSynchronize(p->stream));
  return MEMB_OK;
}

int memb_npts(const memb_plan* p) { return p ? p->T.N * p->T.np2 : 0; }
int memb_ncoef(const memb_plan* p) { return p ? p->T.N * p->T.N : 0; }

int memb_traction_f64_dev(memb_plan* p, int64_t ncap, double gs, double eb, const double* x,
                          const double* y, const double* z, double* fx, double* fy, double* fz,
                          double* nx, double* ny, double* nz, double* area) {
  if (!p) return MEMB_E_ARG;
  if (!p->T.ref) return fail(p, MEMB_E_STATE, "plan not initialised");
  if (ncap < 0) return fail(p, MEMB_E_ARG, "ncap < 0");
  if (ncap > 0 && !(x && y && z && fx && fy && fz && nx && ny && nz && area))
    return fail(p, MEMB_E_ARG, "null pointer");
  MCHK(p, hipSetDevice(p->device));
  return launch<0>(p, ncap, gs, eb, x, y, z, fx, fy, fz, nx, ny, nz, area);
}

int memb_coeffs_f64_dev(memb_plan* p, int64_t ncap, const double* x, const double* y,
                        const double* z, double* coef) {
  if (!p) return MEMB_E_ARG;
  if (ncap < 0) return fail(p, MEMB_E_ARG, "ncap < 0");
  if (ncap > 0 && !(x && y && z && coef)) return fail(p, MEMB_E_ARG, "null pointer");
  MCHK(p, hipSetDevice(p->device));
  return launch<2>(p, ncap, 0.0, 0.0, x, y, z, coef, nullptr, nullptr, nullptr, nullptr, nullptr,
                   nullptr);
}

int memb_geometry_f64_dev(memb_plan* p, int64_t ncap, const double* x, const double* y,
                          const double* z, double* nx, double* ny, double* nz, double* area,
                          const int* skip) {
  if (!p) return MEMB_E_ARG;
  if (ncap < 0) return fail(p, MEMB_E_ARG, "ncap < 0");
  if (ncap > 0 && !(x && y && z && nx && ny && nz && area))
    return fail(p, MEMB_E_ARG, "null pointer");
  MCHK(p, hipSetDevice(p->device));
  if (ncap == 0) return MEMB_OK;
  // three 256-thread workgroups per CU (LDS: 3 x 42 KB at L = 16)
  const unsigned nwg = (unsigned)std::min<long long>(ncap, (long long)p->num_wg * 3 / 4);
  k_geom<<<nwg, GM_THREADS, geom_lds_bytes(p->T), p->stream>>>(p->T, ncap, x, y, z, nx, ny, nz,
                                                               area, skip);
  MCHK(p, hipGetLastError());
  return MEMB_OK;
}

int memb_traction_select_f64_dev(memb_plan* p, int64_t ncap, const int* sel, double gs,
                                 double eb, const double* x, const double* y, const double* z,
                                 double* fx, double* fy, double* fz, double* nx, double* ny,
                                 double* nz, double* area) {
  if (!p) return MEMB_E_ARG;
  if (!p->T.ref) return fail(p, MEMB_E_STATE, "plan not initialised");
  if (ncap < 0) return fail(p, MEMB_E_ARG, "ncap < 0");
  if (ncap > 0 && !(sel && x && y && z && fx && fy && fz && nx && ny && nz && area))
    return fail(p, MEMB_E_ARG, "null pointer");
  MCHK(p, hipSetDevice(p->device));
  return launch<0>(p, ncap, gs, eb, x, y, z, fx, fy, fz, nx, ny, nz, area, sel);
}

int memb_corot_f64_dev(memb_plan* p, int64_t ncap, const double* X, const double* x,
                       const double* y, const double* z, double* delta, double* dhx,
                       double* dhy, double* dhz, double* rot, double* amp) {
  if (!p) return MEMB_E_ARG;
  if (ncap < 0) return fail(p, MEMB_E_ARG, "ncap < 0");
  if (ncap > 0 && !(X && x && y && z && delta && dhx && dhy && dhz && rot && amp))
    return fail(p, MEMB_E_ARG, "null pointer");
  if (ncap == 0) return MEMB_OK;
  MCHK(p, hipSetDevice(p->device));
  k_corot<<<(unsigned)ncap, CR_THREADS, 0, p->stream>>>(p->T.N * p->T.np2, X, x, y, z, delta,
                                                         dhx, dhy, dhz, rot, amp);
  MCHK(p, hipGetLastError());
  return MEMB_OK;
}

int memb_anchored_predict_f64_dev(memb_plan* p, int64_t ncap, const double* lin,
                                  const double* m0, const double* b, const double* delta,
                                  const double* delta0, const double* rot, double beta,
                                  double tol, double* fx, double* fy, double* fz, int* sel,
                                  double* drift, const int* pre, int* late) {
  if (!p) return MEMB_E_ARG;
  if (ncap < 0) return fail(p, MEMB_E_ARG, "ncap < 0");
  if (ncap > 0 && !(lin && b && delta && delta0 && rot && fx && fy && fz && sel))
    return fail(p, MEMB_E_ARG, "null pointer");
  if (ncap == 0) return MEMB_OK;
  MCHK(p, hipSetDevice(p->device));
  const int n = p->T.N * p->T.np2;
  k_anch_predict<<<(unsigned)ncap, CR_THREADS, 0, p->stream>>>(n, lin, m0, b, delta, delta0, rot,
                                                              beta, tol, fx, fy, fz, sel, drift,
                                                              pre, late);
  MCHK(p, hipGetLastError());
  return MEMB_OK;
}

int memb_anchored_gate_f64_dev(memb_plan* p, int64_t ncap, const double* delta,
                               const double* delta0, double beta, double tol, int* pre) {
  if (!p) return MEMB_E_ARG;
  if (ncap < 0) return fail(p, MEMB_E_ARG, "ncap < 0");
  if (ncap > 0 && !(delta && delta0 && pre)) return fail(p, MEMB_E_ARG, "null pointer");
  if (ncap == 0) return MEMB_OK;
  MCHK(p, hipSetDevice(p->device));
  const int n = p->T.N * p->T.np2;
  k_anch_gate<<<(unsigned)ncap, CR_THREADS, 0, p->stream>>>(n, delta, delta0, beta, tol, pre);
  MCHK(p, hipGetLastError());
  return MEMB_OK;
}

int memb_anchored_update_f64_dev(memb_plan* p, int64_t ncap, const int* sel, const int* slot,
                                 const double* rot, const double* fx, const double* fy,
                                 const double* fz, const double* lin, const double* m_now,
                                 const double* delta, const double* amp, double* b,
                                 double* delta0, double* m0, double* s0, int64_t* refreshed) {
  if (!p) return MEMB_E_ARG;
  if (ncap < 0) return fail(p, MEMB_E_ARG, "ncap < 0");
  if (ncap > 0 && !(sel && rot && fx && fy && fz && lin && delta && amp && b && delta0 &&
                    refreshed))
    return fail(p, MEMB_E_ARG, "null pointer");
  if (m_now && !m0) return fail(p, MEMB_E_ARG, "m_now needs m0");
  if (ncap == 0) return MEMB_OK;
  MCHK(p, hipSetDevice(p->device));
  const int n = p->T.N * p->T.np2;
  k_anch_update<<<(unsigned)ncap, CR_THREADS, 0, p->stream>>>(n, sel, slot, rot, fx, fy, fz, lin,
                                                             m_now, delta, amp, b, delta0, m0,
                                                             s0, (long long*)refreshed);
  MCHK(p, hipGetLastError());
  return MEMB_OK;
}

int memb_compact_sel_dev(memb_plan* p, int64_t ncap, const int* sel, int k, int* idx,
                         int* slot) {
  if (!p) return MEMB_E_ARG;
  if (ncap < 0 || k < 1) return fail(p, MEMB_E_ARG, "ncap < 0 or k < 1");
  if (!(sel && idx && slot)) return fail(p, MEMB_E_ARG, "null pointer");
  if (ncap > 0x7fffffffLL) return fail(p, MEMB_E_ARG, "more than 2^31 - 1 capsules");
  MCHK(p, hipSetDevice(p->device));
  k_compact_sel<<<1, CS_THREADS, 0, p->stream>>>(ncap, sel, k, idx, slot);
  MCHK(p, hipGetLastError());
  return MEMB_OK;
}

int memb_corot_apply_f64_dev(memb_plan* p, int64_t ncap, const double* flin, const float* resid,
                             const double* rot, const double* amp, double* fx, double* fy,
                             double* fz) {
  if (!p) return MEMB_E_ARG;
  if (ncap < 0) return fail(p, MEMB_E_ARG, "ncap < 0");
  if (ncap > 0 && !(flin && rot && amp && fx && fy && fz))
    return fail(p, MEMB_E_ARG, "null pointer");
  if (ncap == 0) return MEMB_OK;
  MCHK(p, hipSetDevice(p->device));
  const int n = p->T.N * p->T.np2;
  const long long tot = ncap * n;
  k_corot_apply<<<(unsigned)((tot + 255) / 256), 256, 0, p->stream>>>(ncap, n, flin, resid, rot,
                                                                      amp, fx, fy, fz);
  MCHK(p, hipGetLastError());
  return MEMB_OK;
}

}  // extern "C"
