// ibx_device.hpp -- device code of libibx: the regularized delta, the gather (interpolation),
// the deterministic tile spread and the Lagrangian advance, for gfx950.
//
// Every FP64 expression keeps the reference's evaluation order and is built with
// -ffp-contract=off (and the pragma below), so no multiply-add is fused: the results are
// bit-identical to the reference loop nest (interp_spread.f90:352-376, :852-878,
// :1318-1340) on the same inputs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace ibx {

// Device copy of the grid constants, passed by value.
struct GridD {
  int itot, jtot, ktot;
  long long sy;  // stride of j: itot+2
  long long sz;  // stride of k: (itot+2)*(jtot+2)
  double lx, ly, dxi, dyi, dzi, bl, bf, dveul;
};

// Tile geometry of one spread launch.  Tiles cover i in [1,itot], j in [1,jtot] and
// k in [kbase, kbase+nkp-1] (kbase = 1, nkp = ktot for lagr2eulr; kbase = 0, nkp = ktot+2
// for lagr2eulrind whose window reaches the wall halo planes, interp_spread.f90:1311-1316).
struct TileGeom {
  int ti, tj, tk;     // tile extent in cells
  int nti, ntj, ntk;  // tiles per direction
  int kbase, nkp;     // first k plane and number of k planes covered
  int klo, khi;       // k clamp of the stencil window
  long long ntiles;
};

enum : int { ERR_NONFINITE = 1, ERR_CAPACITY = 2 };

// kernel.f90:7-25, in the reference's operation order.
__device__ __forceinline__ double phi(double r) {
  double ar = fabs(r);
  if (ar > 1.5) return 0.0;
  if (ar > 0.5) {
    double t = 1.0 - ar;
    return (1.0 / 6.0) * ((5.0 - 3.0 * ar) - sqrt(-(3.0 * (t * t)) + 1.0));
  }
  return (1.0 / 3.0) * (1.0 + sqrt(-(3.0 * (r * r)) + 1.0));
}

// Image slots of a master that is its own neighbour (interp_spread.f90:82-141):
// slot -> shift in units of (lx, ly).
__device__ __forceinline__ int slot_sx(int s) {
  // s: 0  1  2  3  4  5  6  7  8
  //    0 +1 +1  0 -1 -1 -1  0 +1
  return (s == 1 || s == 2 || s == 8) ? 1 : ((s == 4 || s == 5 || s == 6) ? -1 : 0);
}
__device__ __forceinline__ int slot_sy(int s) {
  //    0  0 -1 -1 -1  0 +1 +1 +1
  return (s == 2 || s == 3 || s == 4) ? -1 : ((s == 6 || s == 7 || s == 8) ? 1 : 0);
}
// Loop index nb -> data slot (interp_spread.f90:333-336).
__device__ __forceinline__ int nb_slot(int nb) { return nb == 0 ? 0 : (nb + 4 > 8 ? nb - 4 : nb + 4); }

// The copy of a point in slot s: anb(s)%xfp = xfp then +-lx, anb(s)%yfp = yfp then +-ly.
__device__ __forceinline__ void slot_coords(const GridD& g, int s, double x, double y, double& cx,
                                            double& cy) {
  int sx = slot_sx(s), sy = slot_sy(s);
  double xs = x, ys = y;
  if (sx > 0) xs = xs + g.lx;
  if (sx < 0) xs = xs - g.lx;
  if (sy > 0) ys = ys + g.ly;
  if (sy < 0) ys = ys - g.ly;
  cx = (xs - g.bl) * g.dxi;
  cy = (ys - g.bf) * g.dyi;
}

// Support of the stencil along one direction.  With f = floor(c), every tap whose weight can
// be non-zero -- |i-c| < 1.5 (centred) or |i-c-1/2| < 1.5 (staggered) -- lies in [f-1, f+2];
// every tap of the reference's window nint(c-2)..nint(c+2) outside it has weight exactly 0
// (its rounded distance is >= 1.5), and adding +-0 leaves a non-negative-zero sum unchanged.
// Returns false if [f-1,f+2] misses [lo,hi].  `c` must be finite.
__device__ __forceinline__ bool window(double c, int lo, int hi, int& w0, int& a, int& b) {
  if (!(c > (double)lo - 4.0 && c < (double)hi + 4.0)) return false;
  int f = (int)floor(c);
  w0 = f - 1;
  a = w0 < lo ? lo : w0;
  b = (f + 2) > hi ? hi : (f + 2);
  return a <= b;
}

// Weights on the 4-tap support [w0, w0+3]: ns = kernel(i-c) (kernel*s in the reference),
// st = kernel(i-c-1/2) (kernel* without s), zero outside [a,b].
__device__ __forceinline__ void weights4(double c, int w0, int a, int b, double ns[4],
                                         double st[4]) {
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    int i = w0 + t;
    double rs = (1.0 * (double)i) - c;
    double r = rs - 0.5;
    bool in = (i >= a) && (i <= b);
    ns[t] = in ? phi(rs) : 0.0;
    st[t] = in ? phi(r) : 0.0;
  }
}

// ---------------------------------------------------------------------------------------
// Interpolation: eulr2lagr (interp_spread.f90:315-383 and the owner sum :426-440).
// One thread per Lagrangian point.  For every image slot whose clamped window meets the
// grid, the partial sum runs k -> j -> i over the taps of non-zero weight (skipping a zero
// term is exact), then the partials are added in slot order 0..8.
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_interp(GridD g, const double* __restrict__ u,
                                                const double* __restrict__ v,
                                                const double* __restrict__ w, long long npts,
                                                const double* __restrict__ x,
                                                const double* __restrict__ y,
                                                const double* __restrict__ z,
                                                double* __restrict__ ul, double* __restrict__ vl,
                                                double* __restrict__ wl, int* __restrict__ err) {
  long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npts) return;
  double xq = x[q], yq = y[q], zq = z[q];
  double su = 0.0, sv = 0.0, sw = 0.0;
  if (!(isfinite(xq) && isfinite(yq) && isfinite(zq))) {
    atomicOr(err, ERR_NONFINITE);
    ul[q] = 0.0, vl[q] = 0.0, wl[q] = 0.0;
    return;
  }
  double cz = zq * g.dzi;
  int zw0, za, zb;
  bool zin = window(cz, 1, g.ktot, zw0, za, zb);
  if (zin) {
    double zns[4], zst[4];
    weights4(cz, zw0, za, zb, zns, zst);
    for (int s = 0; s < 9; ++s) {
      double cx, cy;
      slot_coords(g, s, xq, yq, cx, cy);
      int xw0, xa, xb, yw0, ya, yb;
      if (!window(cx, 1, g.itot, xw0, xa, xb)) continue;
      if (!window(cy, 1, g.jtot, yw0, ya, yb)) continue;
      double xns[4], xst[4], yns[4], yst[4];
      weights4(cx, xw0, xa, xb, xns, xst);
      weights4(cy, yw0, ya, yb, yns, yst);
      const long long base = (long long)xw0 + (long long)yw0 * g.sy + (long long)zw0 * g.sz;
      double pu = 0.0, pv = 0.0, pw = 0.0;
      // u: kernelxs * kernely * kernelz   (interp_spread.f90:370)
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        if (zst[kk] == 0.0) continue;
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          if (yst[jj] == 0.0) continue;
#pragma unroll
          for (int ii = 0; ii < 4; ++ii) {
            if (xns[ii] == 0.0) continue;
            pu = pu + u[base + ii + jj * g.sy + kk * g.sz] * xns[ii] * yst[jj] * zst[kk];
          }
        }
      }
      // v: kernelx * kernelys * kernelz   (:371)
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        if (zst[kk] == 0.0) continue;
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          if (yns[jj] == 0.0) continue;
#pragma unroll
          for (int ii = 0; ii < 4; ++ii) {
            if (xst[ii] == 0.0) continue;
            pv = pv + v[base + ii + jj * g.sy + kk * g.sz] * xst[ii] * yns[jj] * zst[kk];
          }
        }
      }
      // w: kernelx * kernely * kernelzs   (:372)
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        if (zns[kk] == 0.0) continue;
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          if (yst[jj] == 0.0) continue;
#pragma unroll
          for (int ii = 0; ii < 4; ++ii) {
            if (xst[ii] == 0.0) continue;
            pw = pw + w[base + ii + jj * g.sy + kk * g.sz] * xst[ii] * yst[jj] * zns[kk];
          }
        }
      }
      su = su + pu;
      sv = sv + pv;
      sw = sw + pw;
    }
  }
  ul[q] = su;
  vl[q] = sv;
  wl[q] = sw;
}

// ---------------------------------------------------------------------------------------
// Spread, step 1: binning.  Every (point, image slot) pair whose clamped window meets the
// grid is one candidate; it is listed in every tile its window overlaps (at most 2 per
// direction).  Its 64-bit entry is (K << 32) | (point | nb << 28) where
// K = 9 off_p + nb n_p + l orders candidates exactly as the reference's loop nest visits
// them: capsule p, then image loop index nb, then point l (interp_spread.f90:825-878).
// MODE 0 counts per tile; MODE 1 writes entries at tile_off[t] + cursor[t].
// ---------------------------------------------------------------------------------------
template <int MODE>
__global__ void __launch_bounds__(256) k_bin(GridD g, TileGeom tg, bool cc, long long ncap,
                                             const long long* __restrict__ cap_off,
                                             long long npts, const double* __restrict__ x,
                                             const double* __restrict__ y,
                                             const double* __restrict__ z,
                                             unsigned* __restrict__ count,
                                             const unsigned* __restrict__ tile_off,
                                             unsigned long long* __restrict__ entries,
                                             unsigned long long ecap,
                                             int* __restrict__ err) {
  long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npts) return;
  double xq = x[q], yq = y[q], zq = z[q];
  if (!(isfinite(xq) && isfinite(yq) && isfinite(zq))) {
    if (MODE == 0) atomicOr(err, ERR_NONFINITE);
    return;
  }
  double cz = zq * g.dzi;
  int zw0, za, zb;
  if (!window(cz, tg.klo, tg.khi, zw0, za, zb)) return;
  // capsule of point q: largest p with cap_off[p] <= q
  long long lo = 0, hi = ncap - 1;
  while (lo < hi) {
    long long mid = (lo + hi + 1) >> 1;
    if (cap_off[mid] <= q)
      lo = mid;
    else
      hi = mid - 1;
  }
  const long long p = lo;
  const long long offp = cap_off[p];
  const long long np_ = cap_off[p + 1] - offp;
  const long long l = q - offp;
  const int tk0 = (za - tg.kbase) / tg.tk, tk1 = (zb - tg.kbase) / tg.tk;
  for (int nb = 0; nb < 9; ++nb) {
    int s = nb_slot(nb);
    double cx, cy;
    slot_coords(g, s, xq, yq, cx, cy);
    int xw0, xa, xb, yw0, ya, yb;
    if (!window(cx, 1, g.itot, xw0, xa, xb)) continue;
    if (!window(cy, 1, g.jtot, yw0, ya, yb)) continue;
    const int ti0 = (xa - 1) / tg.ti, ti1 = (xb - 1) / tg.ti;
    const int tj0 = (ya - 1) / tg.tj, tj1 = (yb - 1) / tg.tj;
    unsigned long long key = 0;
    if (MODE == 1) {
      unsigned long long K = 9ull * (unsigned long long)offp +
                             (unsigned long long)nb * (unsigned long long)np_ +
                             (unsigned long long)l;
      key = (K << 32) | (unsigned long long)((unsigned)q | ((unsigned)nb << 28));
    }
    for (int tk = tk0; tk <= tk1; ++tk)
      for (int tj = tj0; tj <= tj1; ++tj)
        for (int ti = ti0; ti <= ti1; ++ti) {
          long long t = (long long)ti + (long long)tg.nti * ((long long)tj + (long long)tg.ntj * tk);
          if (MODE == 0) {
            atomicAdd(&count[t], 1u);
          } else {
            unsigned pos = tile_off[t] + atomicAdd(&count[t], 1u);
            if ((unsigned long long)pos < ecap)
              entries[pos] = key;
            else
              atomicOr(err, ERR_CAPACITY);
          }
        }
  }
}

// ---------------------------------------------------------------------------------------
// Spread, step 2: one workgroup per tile.
//   (a) load the tile's candidate entries and sort them by K (bitonic, in LDS);
//   (b) in chunks of 64 candidates: compute each candidate's window and its 24 weights;
//       mark, for every cell of the tile, which chunk candidates touch it (one 64-bit mask
//       per cell, integer LDS atomics -- the set is order-independent);
//   (c) every thread walks the set bits of its cells in ascending order, i.e. in K order,
//       and adds the terms to register accumulators: cell (i,j,k) receives exactly the
//       reference's sequence force = force + f*wx*wy*wz*(area/dVeul) (:869-874 / :1334-1338);
//   (d) the tile is written once, with the periodic halo copies of bound.f90:141-190.
// No floating-point atomics: results are deterministic and bit-identical to the reference.
// ---------------------------------------------------------------------------------------
constexpr int SP_THREADS = 256;
constexpr int SP_CH = 64;      // candidates per chunk (one bit each in a 64-bit mask)
constexpr int SP_CAPK = 2048;  // candidates sorted in LDS; larger lists are sorted in place in HBM

// Ascending sort of a[0..n) (n <= P, P a power of two) with the "flip" form of the bitonic
// network: every compare-exchange puts the minimum at the lower index, so the virtual +inf
// elements a[n..P) never move and need no storage.  Works in LDS or in HBM (one workgroup;
// __syncthreads orders the workgroup's own global accesses).
__device__ __forceinline__ void cmpx(unsigned long long* a, int lo, int hi, int n) {
  if (hi < n) {
    unsigned long long x = a[lo], y = a[hi];
    if (x > y) {
      a[lo] = y;
      a[hi] = x;
    }
  }
}

__device__ void bitonic_sort(unsigned long long* a, int n, int tid) {
  int P = 1;
  while (P < n) P <<= 1;
  const int half = P >> 1;
  for (int k = 2; k <= P; k <<= 1) {
    const int hk = k >> 1;
    for (int t = tid; t < half; t += SP_THREADS) {
      const int blk = t / hk, o = t % hk;
      cmpx(a, blk * k + o, blk * k + k - 1 - o, n);
    }
    __syncthreads();
    for (int j = k >> 2; j > 0; j >>= 1) {
      for (int t = tid; t < half; t += SP_THREADS) {
        const int lo = (t / j) * 2 * j + (t % j);
        cmpx(a, lo, lo + j, n);
      }
      __syncthreads();
    }
  }
}

template <int TI, int TJ, int TK, bool CC>
struct SpreadTile {
  static constexpr int NC = TI * TJ * TK;
  static constexpr int CPT = NC / SP_THREADS;
  static_assert(NC % SP_THREADS == 0, "tile must hold a multiple of 256 cells");
  static_assert(TI >= 4 && TJ >= 4 && TK >= 4, "tile must be at least the stencil width");

  struct Smem {
    unsigned long long keys[SP_CAPK];
    unsigned long long mask[NC];
    double wt[SP_CH][3][8];  // per candidate, per direction: ns[4], st[4]
    double fv[SP_CH][4];     // f1, f2, f3, area/dVeul
    int w0[SP_CH][4];        // window origin relative to the tile, per direction; [3]: valid
    int lim[SP_CH][6];       // clamped window [a,b] relative to the tile, per direction
  };

  __device__ static void run(Smem& sm, long long t, const GridD& g, const TileGeom& tg,
                             const unsigned* __restrict__ tile_off,
                             unsigned long long* __restrict__ entries,
                             unsigned long long ecap, const double* __restrict__ x,
                             const double* __restrict__ y, const double* __restrict__ z,
                             const double* __restrict__ f1, const double* __restrict__ f2,
                             const double* __restrict__ f3, const double* __restrict__ area,
                             double* __restrict__ o1, double* __restrict__ o2,
                             double* __restrict__ o3) {
    const int tid = threadIdx.x;
    const int ti = (int)(t % tg.nti);
    const int tj = (int)((t / tg.nti) % tg.ntj);
    const int tk = (int)(t / ((long long)tg.nti * tg.ntj));
    const int i0 = 1 + ti * TI, j0 = 1 + tj * TJ, k0 = tg.kbase + tk * TK;

    unsigned base = tile_off[t];
    unsigned long long endl = tile_off[t + 1];
    if (endl > ecap) endl = ecap;  // overflowed lists are flagged by k_bin; never read past
    int n = (unsigned long long)base < endl ? (int)(endl - base) : 0;

    double acc1[CPT], acc2[CPT], acc3[CPT];
#pragma unroll
    for (int r = 0; r < CPT; ++r) acc1[r] = acc2[r] = acc3[r] = 0.0;

    if (n > 0) {
      const unsigned long long* src;
      if (n <= SP_CAPK) {
        for (int i = tid; i < n; i += SP_THREADS) sm.keys[i] = entries[base + i];
        __syncthreads();
        bitonic_sort(sm.keys, n, tid);
        src = sm.keys;
      } else {
        bitonic_sort(entries + base, n, tid);
        src = entries + base;
      }
      for (int c0 = 0; c0 < n; c0 += SP_CH) {
        const int m = (n - c0) < SP_CH ? (n - c0) : SP_CH;
#pragma unroll
        for (int r = 0; r < CPT; ++r) sm.mask[tid + r * SP_THREADS] = 0ull;
        // (b) candidate preparation: 4 threads per candidate
        {
          const int c = tid >> 2, d = tid & 3;
          if (c < m) {
            unsigned long long e = src[c0 + c];
            unsigned lo32 = (unsigned)(e & 0xffffffffull);
            long long q = (long long)(lo32 & 0x0fffffffu);
            int nb = (int)(lo32 >> 28);
            int s = nb_slot(nb);
            if (d < 3) {
              double cc_;
              int lo, hi, base_;
              if (d == 0) {
                double cy_;
                slot_coords(g, s, x[q], y[q], cc_, cy_);
                lo = 1, hi = g.itot, base_ = i0;
              } else if (d == 1) {
                double cx_;
                slot_coords(g, s, x[q], y[q], cx_, cc_);
                lo = 1, hi = g.jtot, base_ = j0;
              } else {
                cc_ = z[q] * g.dzi;
                lo = tg.klo, hi = tg.khi, base_ = k0;
              }
              int w0, a, b;
              bool ok = window(cc_, lo, hi, w0, a, b);
              double ns[4], st[4];
              if (ok) {
                weights4(cc_, w0, a, b, ns, st);
              } else {
#pragma unroll
                for (int tt = 0; tt < 4; ++tt) ns[tt] = st[tt] = 0.0;
                w0 = 0, a = 1, b = 0;
              }
#pragma unroll
              for (int tt = 0; tt < 4; ++tt) {
                sm.wt[c][d][tt] = ns[tt];
                sm.wt[c][d][4 + tt] = st[tt];
              }
              sm.w0[c][d] = w0 - base_;
              sm.lim[c][2 * d] = a - base_;
              sm.lim[c][2 * d + 1] = b - base_;
            } else {
              sm.fv[c][0] = f1[q];
              sm.fv[c][1] = f2[q];
              sm.fv[c][2] = f3[q];
              sm.fv[c][3] = area[q] / g.dveul;
            }
          }
        }
        __syncthreads();
        // mark the cells each candidate touches: candidate c = tid/4 marks 16 of its 64
        {
          const int c = tid >> 2, part = tid & 3;
          if (c < m) {
            const int wx = sm.w0[c][0], wy = sm.w0[c][1], wz = sm.w0[c][2];
            const int ax = max(sm.lim[c][0], 0), bx = min(sm.lim[c][1], TI - 1);
            const int ay = max(sm.lim[c][2], 0), by = min(sm.lim[c][3], TJ - 1);
            const int az = max(sm.lim[c][4], 0), bz = min(sm.lim[c][5], TK - 1);
            const unsigned long long bit = 1ull << c;
#pragma unroll
            for (int h = 0; h < 16; ++h) {
              const int idx = part * 16 + h;
              const int li = wx + (idx & 3), lj = wy + ((idx >> 2) & 3), lk = wz + (idx >> 4);
              if (li >= ax && li <= bx && lj >= ay && lj <= by && lk >= az && lk <= bz)
                atomicOr(&sm.mask[li + TI * (lj + TJ * lk)], bit);
            }
          }
        }
        __syncthreads();
        // (c) ordered accumulation
#pragma unroll
        for (int r = 0; r < CPT; ++r) {
          const int cell = tid + r * SP_THREADS;
          const int li = cell % TI, lj = (cell / TI) % TJ, lk = cell / (TI * TJ);
          unsigned long long mk = sm.mask[cell];
          while (mk) {
            const int c = __builtin_ctzll(mk);
            mk &= mk - 1;
            const int a = li - sm.w0[c][0], b = lj - sm.w0[c][1], d = lk - sm.w0[c][2];
            const double adv = sm.fv[c][3];
            if (!CC) {
              const double wxs = sm.wt[c][0][a], wx = sm.wt[c][0][4 + a];
              const double wys = sm.wt[c][1][b], wy = sm.wt[c][1][4 + b];
              const double wzs = sm.wt[c][2][d], wz = sm.wt[c][2][4 + d];
              acc1[r] = acc1[r] + sm.fv[c][0] * wxs * wy * wz * adv;
              acc2[r] = acc2[r] + sm.fv[c][1] * wx * wys * wz * adv;
              acc3[r] = acc3[r] + sm.fv[c][2] * wx * wy * wzs * adv;
            } else {
              const double wx = sm.wt[c][0][4 + a], wy = sm.wt[c][1][4 + b],
                           wz = sm.wt[c][2][4 + d];
              acc1[r] = acc1[r] + sm.fv[c][0] * wx * wy * wz * adv;
              acc2[r] = acc2[r] + sm.fv[c][1] * wx * wy * wz * adv;
              acc3[r] = acc3[r] + sm.fv[c][2] * wx * wy * wz * adv;
            }
          }
        }
        __syncthreads();
      }
    }
    // (d) write the tile and its periodic halo images
#pragma unroll
    for (int r = 0; r < CPT; ++r) {
      const int cell = tid + r * SP_THREADS;
      const int li = cell % TI, lj = (cell / TI) % TJ, lk = cell / (TI * TJ);
      const int i = i0 + li, j = j0 + lj, k = k0 + lk;
      if (i > g.itot || j > g.jtot || k >= tg.kbase + tg.nkp) continue;
      const long long kk = (long long)k * g.sz;
      const double a1 = acc1[r], a2 = acc2[r], a3 = acc3[r];
      long long idx = (long long)i + (long long)j * g.sy + kk;
      o1[idx] = a1, o2[idx] = a2, o3[idx] = a3;
      const int ih = (i == 1) ? g.itot + 1 : ((i == g.itot) ? 0 : -1);
      const int jh = (j == 1) ? g.jtot + 1 : ((j == g.jtot) ? 0 : -1);
      if (ih >= 0) {
        long long h = (long long)ih + (long long)j * g.sy + kk;
        o1[h] = a1, o2[h] = a2, o3[h] = a3;
      }
      if (jh >= 0) {
        long long h = (long long)i + (long long)jh * g.sy + kk;
        o1[h] = a1, o2[h] = a2, o3[h] = a3;
        if (ih >= 0) {
          long long h2 = (long long)ih + (long long)jh * g.sy + kk;
          o1[h2] = a1, o2[h2] = a2, o3[h2] = a3;
        }
      }
      // itot == 1 or jtot == 1 would need both images; the context rejects such grids.
    }
  }
};

// Workgroup -> tile: consecutive workgroups are dealt round-robin to the 8 XCDs, so give
// each XCD a contiguous run of tiles (neighbouring tiles share candidates in its L2).
__device__ __forceinline__ long long xcd_tile(long long bid, long long nblk) {
  long long per = nblk >> 3;
  return (bid & 7) * per + (bid >> 3);
}

template <int TI, int TJ, int TK, bool CC, bool PERSIST>
__global__ void __launch_bounds__(SP_THREADS)
    k_spread(GridD g, TileGeom tg, const unsigned* __restrict__ tile_off,
             unsigned long long* __restrict__ entries, unsigned long long ecap,
             const double* __restrict__ x, const double* __restrict__ y,
             const double* __restrict__ z, const double* __restrict__ f1,
             const double* __restrict__ f2, const double* __restrict__ f3,
             const double* __restrict__ area, double* __restrict__ o1, double* __restrict__ o2,
             double* __restrict__ o3) {
  using T = SpreadTile<TI, TJ, TK, CC>;
  __shared__ typename T::Smem sm;
  if (PERSIST) {
    // static round-robin over tiles; every workgroup leaves after its last tile
    for (long long t = blockIdx.x; t < tg.ntiles; t += gridDim.x) {
      T::run(sm, t, g, tg, tile_off, entries, ecap, x, y, z, f1, f2, f3, area, o1, o2, o3);
      __syncthreads();
    }
  } else {
    long long t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= tg.ntiles) return;
    T::run(sm, t, g, tg, tile_off, entries, ecap, x, y, z, f1, f2, f3, area, o1, o2, o3);
  }
}

// Wall halo planes k = 0 and k = k1 of lagr2eulr: zero (the scatter never reaches them and
// updthalos copies only x and y halos, bound.f90:141-190).
__global__ void __launch_bounds__(256) k_zero_kplanes(GridD g, double* __restrict__ o1,
                                                      double* __restrict__ o2,
                                                      double* __restrict__ o3) {
  long long n = g.sz;  // one plane
  long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 2 * n) return;
  long long idx = q < n ? q : q - n + (long long)(g.ktot + 1) * g.sz;
  o1[idx] = 0.0, o2[idx] = 0.0, o3[idx] = 0.0;
}

__global__ void __launch_bounds__(256) k_fill_u32(unsigned* __restrict__ a, long long n,
                                                  unsigned v) {
  long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) a[q] = v;
}

__global__ void __launch_bounds__(64) k_fill_i32(int* __restrict__ a, int n, int v) {
  int q = threadIdx.x;
  if (q < n) a[q] = v;
}

// ---------------------------------------------------------------------------------------
// Exclusive scan of the per-tile counts (integers, order-independent).
// ---------------------------------------------------------------------------------------
constexpr int SCAN_B = 1024;  // elements per block (256 threads x 4)

__device__ __forceinline__ unsigned block_excl_scan(unsigned v, unsigned* sh, unsigned& total) {
  // 256 threads
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    unsigned t = tid >= o ? sh[tid - o] : 0u;
    __syncthreads();
    sh[tid] += t;
    __syncthreads();
  }
  total = sh[255];
  unsigned incl = sh[tid];
  __syncthreads();
  return incl - v;
}

__global__ void __launch_bounds__(256) k_scan_reduce(const unsigned* __restrict__ in, long long n,
                                                     unsigned* __restrict__ part) {
  __shared__ unsigned sh[256];
  long long b0 = (long long)blockIdx.x * SCAN_B + threadIdx.x * 4;
  unsigned s = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r)
    if (b0 + r < n) s += in[b0 + r];
  unsigned tot;
  block_excl_scan(s, sh, tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(256) k_scan_parts(unsigned* __restrict__ part, long long nb) {
  __shared__ unsigned sh[256];
  unsigned carry = 0;
  for (long long c = 0; c < nb; c += 1024) {
    long long b0 = c + threadIdx.x * 4;
    unsigned v[4], s = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      v[r] = (b0 + r < nb) ? part[b0 + r] : 0u;
      s += v[r];
    }
    unsigned tot;
    unsigned ex = block_excl_scan(s, sh, tot) + carry;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (b0 + r < nb) part[b0 + r] = ex;
      ex += v[r];
    }
    carry += tot;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256) k_scan_apply(const unsigned* __restrict__ in, long long n,
                                                    const unsigned* __restrict__ part,
                                                    unsigned* __restrict__ out) {
  __shared__ unsigned sh[256];
  long long b0 = (long long)blockIdx.x * SCAN_B + threadIdx.x * 4;
  unsigned v[4], s = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    v[r] = (b0 + r < n) ? in[b0 + r] : 0u;
    s += v[r];
  }
  unsigned tot;
  unsigned ex = block_excl_scan(s, sh, tot) + part[blockIdx.x];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    if (b0 + r < n) out[b0 + r] = ex;
    ex += v[r];
  }
  if (b0 <= n - 1 && n - 1 < b0 + 4) out[n] = part[blockIdx.x] + tot;  // total
}

// ---------------------------------------------------------------------------------------
// Update_Pos.f90:77-186 on the device.
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_advance(GridD g, double lz, double dt, long long npts,
                                                 double* __restrict__ x, double* __restrict__ y,
                                                 double* __restrict__ z, double* __restrict__ xo,
                                                 double* __restrict__ yo,
                                                 const double* __restrict__ ul,
                                                 const double* __restrict__ vl,
                                                 const double* __restrict__ wl) {
  long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npts) return;
  double xq = x[q], yq = y[q], zq = z[q];
  if (xo) {
    double a = xo[q], b = yo[q];
    if ((a - xq) > (g.lx * .5))
      a = a - g.lx;
    else if ((a - xq) < -(g.lx * .5))
      a = a + g.lx;
    if ((b - yq) > (g.ly * .5))
      b = b - g.ly;
    else if ((b - yq) < -(g.ly * .5))
      b = b + g.ly;
    xo[q] = a, yo[q] = b;
  }
  xq = xq + (ul[q] * dt);
  yq = yq + (vl[q] * dt);
  zq = zq + (wl[q] * dt);
  if (zq < 0.0)
    zq = 0.0;
  else if (zq > lz)
    zq = lz;
  x[q] = xq, y[q] = yq, z[q] = zq;
}

// one thread per capsule: sequential sums in point order (:135-145), then the shift decision
__global__ void __launch_bounds__(64) k_centroid(GridD g, long long ncap,
                                                 const long long* __restrict__ off,
                                                 const double* __restrict__ x,
                                                 const double* __restrict__ y,
                                                 const double* __restrict__ z,
                                                 const double* __restrict__ ul,
                                                 const double* __restrict__ vl,
                                                 const double* __restrict__ wl,
                                                 double* __restrict__ cen,
                                                 double* __restrict__ um,
                                                 signed char* __restrict__ shift) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ncap) return;
  long long a = off[p], b = off[p + 1];
  double nl = (double)(b - a);
  double cx = 0.0, cy = 0.0, cz = 0.0, mu = 0.0, mv = 0.0, mw = 0.0;
  for (long long q = a; q < b; ++q) {
    cx = cx + x[q];
    cy = cy + y[q];
    cz = cz + z[q];
    mu = mu + ul[q];
    mv = mv + vl[q];
    mw = mw + wl[q];
  }
  cx = cx / nl, cy = cy / nl, cz = cz / nl;
  mu = mu / nl, mv = mv / nl, mw = mw / nl;
  int sx = 0, sy = 0;
  if (cx > g.lx) {
    cx = cx - g.lx;
    sx = -1;
  }
  if (cx < 0.0) {
    cx = cx + g.lx;
    sx = (sx == -1) ? 2 : 1;  // 2: both branches taken (x - lx then + lx)
  }
  if (cy > g.ly) {
    cy = cy - g.ly;
    sy = -1;
  }
  if (cy < 0.0) {
    cy = cy + g.ly;
    sy = (sy == -1) ? 2 : 1;
  }
  cen[3 * p] = cx, cen[3 * p + 1] = cy, cen[3 * p + 2] = cz;
  um[3 * p] = mu, um[3 * p + 1] = mv, um[3 * p + 2] = mw;
  shift[2 * p] = (signed char)sx;
  shift[2 * p + 1] = (signed char)sy;
}

__global__ void __launch_bounds__(256) k_shift(GridD g, long long ncap,
                                               const long long* __restrict__ off,
                                               long long npts, const signed char* __restrict__ shift,
                                               double* __restrict__ x, double* __restrict__ y) {
  long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npts) return;
  long long lo = 0, hi = ncap - 1;
  while (lo < hi) {
    long long mid = (lo + hi + 1) >> 1;
    if (off[mid] <= q)
      lo = mid;
    else
      hi = mid - 1;
  }
  int sx = shift[2 * lo], sy = shift[2 * lo + 1];
  // same per-point operation sequence as the reference's two successive loops
  if (sx == -1 || sx == 2) x[q] = x[q] - g.lx;
  if (sx == 1 || sx == 2) x[q] = x[q] + g.lx;
  if (sy == -1 || sy == 2) y[q] = y[q] - g.ly;
  if (sy == 1 || sy == 2) y[q] = y[q] + g.ly;
}

}  // namespace ibx
