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
  int ti, tj, tk;     // tile extent: ti = chunk width (16 positions per row), tj x tk rows
  int nti, ntj, ntk;  // tiles per direction (nti: chunk columns per row)
  int kbase, nkp;     // first k plane and number of k planes covered
  int klo, khi;       // k clamp of the stencil window
  long long ntiles;
};

enum : int { ERR_NONFINITE = 1, ERR_CAPACITY = 2, ERR_FORCE = 4, ERR_OFFSETS = 8 };

// A device-held capsule offset table that is not 0 at the start, non-decreasing and npts at
// the end sets ERR_OFFSETS; every kernel that indexes points through the table (or through
// work lists built from it) then returns at once, so no point index leaves [0, npts).
__device__ __forceinline__ bool offsets_bad(const int* err) {
  return (*(volatile const int*)err & ERR_OFFSETS) != 0;
}

__global__ void __launch_bounds__(256) k_check_off(long long ncap,
                                                   const long long* __restrict__ off,
                                                   long long npts, int* __restrict__ err) {
  int bad = 0;
  for (long long p = threadIdx.x; p < ncap; p += 256) bad |= off[p + 1] < off[p];
  if (threadIdx.x == 0) bad |= off[0] != 0 || off[ncap] != npts;
  if (__syncthreads_or(bad) && threadIdx.x == 0) atomicOr(err, ERR_OFFSETS);
}


// kernel.f90:7-25, in the reference's operation order.  Both branches are evaluated and one
// is selected, so a wavefront with mixed |r| runs a single square root; each branch performs
// exactly the reference's operations, hence the same rounded result.
__device__ __forceinline__ double phi(double r) {
  const double ar = fabs(r);
  const bool outer = ar > 0.5;
  const double t = 1.0 - ar;
  const double arg = outer ? (-(3.0 * (t * t)) + 1.0) : (-(3.0 * (r * r)) + 1.0);
  const double sq = sqrt(arg);
  const double v = outer ? (1.0 / 6.0) * ((5.0 - 3.0 * ar) - sq) : (1.0 / 3.0) * (1.0 + sq);
  return ar > 1.5 ? 0.0 : v;
}

// kernel.f90:7-25 over an array: the phi() of every interpolation and spread weight.
__global__ void __launch_bounds__(256) k_phi(long long n, const double* __restrict__ r,
                                             double* __restrict__ out) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) out[q] = phi(r[q]);
}

// kerneltest, kernel.f90:43-63: the 51 kernel values phi(i/12), i = -25..25, are formed by
// the lanes of one wavefront; lane 0 then adds the 51^3 products k -> j -> i in the
// reference's order, each product formed left to right, (1/12)(1/12)(1/12) phi_x phi_y phi_z.
__global__ void __launch_bounds__(64) k_kerneltest(double* __restrict__ sum) {
  __shared__ double ph[51];
  const int t = threadIdx.x;
  if (t < 51) ph[t] = phi((double)(t - 25) / 12.0);
  __syncthreads();
  if (t != 0) return;
  const double c = 1.0 / 12.0;
  double s = 0.0;
  for (int k = 0; k < 51; ++k)
    for (int j = 0; j < 51; ++j)
      for (int i = 0; i < 51; ++i) s = s + c * c * c * ph[i] * ph[j] * ph[k];
  *sum = s;
}

// Image slots of a master that is its own neighbour (interp_spread.f90:82-141):
// slot -> shift in units of (lx, ly), coded 0: none, 1: +, 2: -.
//   s:  0  1  2  3  4  5  6  7  8
//   x:  0  +  +  0  -  -  -  0  +
//   y:  0  0  -  -  -  0  +  +  +
__device__ __forceinline__ int slot_xs(int s) {
  return (s == 1 || s == 2 || s == 8) ? 1 : ((s == 4 || s == 5 || s == 6) ? 2 : 0);
}
__device__ __forceinline__ int slot_ys(int s) {
  return (s == 2 || s == 3 || s == 4) ? 2 : ((s == 6 || s == 7 || s == 8) ? 1 : 0);
}
// Loop index nb -> data slot (interp_spread.f90:333-336).
__device__ __forceinline__ int nb_slot(int nb) { return nb == 0 ? 0 : (nb + 4 > 8 ? nb - 4 : nb + 4); }

// Stencil coordinate of the copy shifted by code sh (0 none, 1 +L, 2 -L):
// anb(s)%xfp = xfp (+-lx), coorxfp = (anb(s)%xfp - boundleftmyid)*dxi (:337-338).
__device__ __forceinline__ double shifted_coord(double x, int sh, double L, double b, double inv) {
  double xs = x;
  if (sh == 1) xs = xs + L;
  if (sh == 2) xs = xs - L;
  return (xs - b) * inv;
}

// Support of the stencil along one direction.  With f = floor(c), every tap whose weight can
// be non-zero -- |i-c| < 1.5 (centred) or |i-c-1/2| < 1.5 (staggered) -- lies in [f-1, f+2];
// every tap of the reference's window nint(c-2)..nint(c+2) outside it has weight exactly 0
// (its rounded distance is >= 1.5), and adding +-0 leaves a sum that is not -0 unchanged.
// Returns false if [f-1,f+2] misses [lo,hi].  `c` must be finite.
__device__ __forceinline__ bool window(double c, int lo, int hi, int& w0, int& a, int& b) {
  if (!(c > (double)lo - 4.0 && c < (double)hi + 4.0)) return false;
  int f = (int)floor(c);
  w0 = f - 1;
  a = w0 < lo ? lo : w0;
  b = (f + 2) > hi ? hi : (f + 2);
  return a <= b;
}

// Weights of one copy of a point along one direction, on the 4-tap support [w0, w0+3] of
// window(): ns = kernel(i-c) (the reference's kernel*s), st = kernel(i-c-1/2) (kernel*),
// zero outside the clamped window [a,b].  The staggered weight of tap w0 is always 0
// (distance <= -1.5), so st is kept for taps 1..3; the centred weights are non-zero on at
// most three consecutive taps, {0,1,2} when tap 0's distance exceeds -1.5, else within
// {1,2,3}, so ns is kept for taps na..na+2.  Only those six are evaluated.
struct Axis {
  int w0, na;
  double ns[3];  // taps na, na+1, na+2
  double st[3];  // taps 1, 2, 3
};

__device__ __forceinline__ bool axis_of(double c, int lo, int hi, Axis& ax) {
  int a, b;
  if (!window(c, lo, hi, ax.w0, a, b)) return false;
  const double rs0 = (1.0 * (double)ax.w0) - c;
  ax.na = (rs0 > -1.5) ? 0 : 1;
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    const int in = ax.w0 + ax.na + t;
    const double rn = (1.0 * (double)in) - c;
    ax.ns[t] = (in >= a && in <= b) ? phi(rn) : 0.0;
    const int is = ax.w0 + 1 + t;
    const double rs = (1.0 * (double)is) - c;
    const double r = rs - 0.5;
    ax.st[t] = (is >= a && is <= b) ? phi(r) : 0.0;
  }
  return true;
}

// ---------------------------------------------------------------------------------------
// Interpolation: eulr2lagr (interp_spread.f90:315-383 and the owner sum :426-440).
// The partial sum of one image slot runs k -> j -> i over the 3x3x3 taps that carry the
// component's non-zero weights (skipping a zero term is exact).  A tap outside the clamp has
// weight 0 and is read at a clamped in-bounds index (adding +-0 is exact).
// ---------------------------------------------------------------------------------------
// Weights of one staggering along one direction: STAG = 0 centred (kernel*s, taps na..na+2),
// STAG = 1 staggered (kernel*, taps 1..3); zero outside the clamped window.  Returns false if
// the window misses [lo,hi].
template <int STAG>
__device__ __forceinline__ bool axis1(double c, int lo, int hi, int& first, double w[3]) {
  int w0, a, b;
  if (!window(c, lo, hi, w0, a, b)) return false;
  int t0 = 1;
  if (STAG == 0) {
    const double rs0 = (1.0 * (double)w0) - c;
    t0 = (rs0 > -1.5) ? 0 : 1;
  }
  first = w0 + t0;
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    const int i = first + t;
    const double rs = (1.0 * (double)i) - c;
    w[t] = (i >= a && i <= b) ? phi(STAG == 0 ? rs : rs - 0.5) : 0.0;
  }
  return true;
}

// One thread per (point, velocity component); the component is blockIdx.y (0: u, 1: v, 2: w).
// For every image slot whose clamped window meets the grid, all 27 taps of the component's
// stencil are loaded first, then summed k -> j -> i in the reference's order; partials are
// added in slot order 0..8.
template <int C>
__device__ __forceinline__ double interp_component(const GridD& g, const double* __restrict__ f,
                                                   double xq, double yq, double zq) {
  constexpr int SX = C == 0 ? 0 : 1, SY = C == 1 ? 0 : 1, SZ = C == 2 ? 0 : 1;
  double wz[3];
  int kz;
  if (!axis1<SZ>(zq * g.dzi, 1, g.ktot, kz, wz)) return 0.0;
  int okx = 0, oky = 0;
#pragma unroll
  for (int sh = 0; sh < 3; ++sh) {
    int w0, a, b;
    if (window(shifted_coord(xq, sh, g.lx, g.bl, g.dxi), 1, g.itot, w0, a, b)) okx |= 1 << sh;
    if (window(shifted_coord(yq, sh, g.ly, g.bf, g.dyi), 1, g.jtot, w0, a, b)) oky |= 1 << sh;
  }
  double sum = 0.0;
  for (int s = 0; s < 9; ++s) {
    const int xs = slot_xs(s), ys = slot_ys(s);
    if (!((okx >> xs) & 1) || !((oky >> ys) & 1)) continue;
    double wx[3], wy[3];
    int ix0, jy0;
    axis1<SX>(shifted_coord(xq, xs, g.lx, g.bl, g.dxi), 1, g.itot, ix0, wx);
    axis1<SY>(shifted_coord(yq, ys, g.ly, g.bf, g.dyi), 1, g.jtot, jy0, wy);
    long long ix[3], iy[3], iz[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      int i = ix0 + t, j = jy0 + t, k = kz + t;  // clamped in-bounds address for zero taps
      i = i < 0 ? 0 : (i > g.itot + 1 ? g.itot + 1 : i);
      j = j < 0 ? 0 : (j > g.jtot + 1 ? g.jtot + 1 : j);
      k = k < 0 ? 0 : (k > g.ktot + 1 ? g.ktot + 1 : k);
      ix[t] = i, iy[t] = (long long)j * g.sy, iz[t] = (long long)k * g.sz;
    }
    // the three x-taps of a row are consecutive cells unless the window is clamped: then one
    // 16-byte and one 8-byte load per row (fewer cache-line lookups than three gathers)
    typedef double dv2 __attribute__((ext_vector_type(2), aligned(8)));
    const bool row = ix0 >= 0 && ix0 + 2 <= g.itot + 1;
    double v[27];
#pragma unroll
    for (int kk = 0; kk < 3; ++kk)
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) {
        double* vr = v + 3 * jj + 9 * kk;
        if (row) {
          const double* pr = f + (ix0 + iy[jj] + iz[kk]);
          const dv2 a = *reinterpret_cast<const dv2*>(pr);
          vr[0] = a.x, vr[1] = a.y, vr[2] = pr[2];
        } else {
#pragma unroll
          for (int ii = 0; ii < 3; ++ii) vr[ii] = f[ix[ii] + iy[jj] + iz[kk]];
        }
      }
    double p = 0.0;
#pragma unroll
    for (int kk = 0; kk < 3; ++kk)
#pragma unroll
      for (int jj = 0; jj < 3; ++jj)
#pragma unroll
        for (int ii = 0; ii < 3; ++ii) p = p + v[ii + 3 * jj + 9 * kk] * wx[ii] * wy[jj] * wz[kk];
    sum = sum + p;
  }
  return sum;
}

__global__ void __launch_bounds__(256) k_interp(GridD g, const double* __restrict__ u,
                                                const double* __restrict__ v,
                                                const double* __restrict__ w, long long npts,
                                                const double* __restrict__ x,
                                                const double* __restrict__ y,
                                                const double* __restrict__ z,
                                                double* __restrict__ ul, double* __restrict__ vl,
                                                double* __restrict__ wl, int* __restrict__ err) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npts) return;
  const int c = blockIdx.y;
  const double xq = x[q], yq = y[q], zq = z[q];
  if (!(isfinite(xq) && isfinite(yq) && isfinite(zq))) {
    if (c == 0) atomicOr(err, ERR_NONFINITE);
    (c == 0 ? ul : (c == 1 ? vl : wl))[q] = 0.0;
    return;
  }
  // u: kernelxs*kernely*kernelz, v: kernelx*kernelys*kernelz, w: kernelx*kernely*kernelzs
  // (interp_spread.f90:370-372)
  if (c == 0)
    ul[q] = interp_component<0>(g, u, xq, yq, zq);
  else if (c == 1)
    vl[q] = interp_component<1>(g, v, xq, yq, zq);
  else
    wl[q] = interp_component<2>(g, w, xq, yq, zq);
}

// ---------------------------------------------------------------------------------------
// Spread tiles are cut along the aligned 128-byte chunks of the field's memory, so that
// every tile writes whole chunks.  In row (j,k) the memory index of cell i is base + i with
// base = j*(itot+2) + k*(itot+2)(jtot+2); the row phase is s = base mod 16 and cell i lies in
// tile column (i + s) / 16 at position (i + s) mod 16.  A tile is one column over TJ x TK
// rows: 16 positions per row, whose cells shift with the row phase.  The periodic halo
// cells i = 0 and i = itot+1 of a row are tiled like the others and computed as duplicates
// of cells itot and 1 (their values after bound.f90:141-190).
// ---------------------------------------------------------------------------------------
constexpr int SP_CW = 16;  // doubles per aligned chunk = tile column width

__device__ __forceinline__ int row_phase(const GridD& g, int j, int k) {
  return (int)(((long long)j * g.sy + (long long)k * g.sz) & (SP_CW - 1));
}

// ---------------------------------------------------------------------------------------
// Spread, step 1: binning.  Every (point, image slot) pair whose clamped window meets the
// grid is one candidate; it is listed in every tile whose cells its window covers: for each
// tile row-block (tj, tk) it meets, the columns that hold its x-taps in those rows, plus the
// columns of the halo duplicates (cell i = 0 when the window covers itot, cell itot+1 when it
// covers 1).  Its 64-bit entry is (K << 32) | (point | nb << 28) where
// K = 9 off_p + nb n_p + l orders candidates exactly as the reference's loop nest visits
// them: capsule p, then image loop index nb, then point l (interp_spread.f90:825-878).
// MODE 0 counts per tile; MODE 1 writes entries at tile_off[t] + cursor[t].  Lanes of a
// wavefront that target the same tile share one atomic (wave-aggregated).
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ unsigned agg_add(unsigned* __restrict__ cnt, int tile, bool valid) {
  unsigned long long active = __ballot(valid);
  unsigned res = 0;
  while (active) {
    const int leader = __ffsll((long long)active) - 1;
    const int lt = __shfl(tile, leader);
    const bool mine = valid && tile == lt;
    const unsigned long long peers = __ballot(mine);
    unsigned base = 0;
    if ((int)(threadIdx.x & 63) == leader) base = atomicAdd(&cnt[lt], (unsigned)__popcll(peers));
    base = __shfl(base, leader);
    if (mine) res = base + (unsigned)__popcll(peers & __lanemask_lt());
    active &= ~peers;
  }
  return res;
}

template <int MODE>
__global__ void __launch_bounds__(256) k_bin(GridD g, TileGeom tg, long long ncap,
                                             const long long* __restrict__ cap_off,
                                             long long npts, const double* __restrict__ x,
                                             const double* __restrict__ y,
                                             const double* __restrict__ z,
                                             unsigned* __restrict__ count,
                                             const unsigned* __restrict__ tile_off,
                                             unsigned long long* __restrict__ entries,
                                             unsigned long long ecap,
                                             int* __restrict__ err) {
  if (offsets_bad(err)) return;
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  bool live = q < npts;
  double xq = 0, yq = 0, zq = 0;
  if (live) {
    xq = x[q], yq = y[q], zq = z[q];
    if (!(isfinite(xq) && isfinite(yq) && isfinite(zq))) {
      if (MODE == 0) atomicOr(err, ERR_NONFINITE);
      live = false;
    }
  }
  int zw0 = 0, za = 1, zb = 0;
  if (live) live = window(zq * g.dzi, tg.klo, tg.khi, zw0, za, zb);
  // clamped windows of the x / y copies (unshifted, +L, -L); a > b marks "misses the grid"
  int xa[3], xb[3], ya[3], yb[3];
#pragma unroll
  for (int sh = 0; sh < 3; ++sh) {
    int w0;
    xa[sh] = ya[sh] = 1, xb[sh] = yb[sh] = 0;
    if (live && !window(shifted_coord(xq, sh, g.lx, g.bl, g.dxi), 1, g.itot, w0, xa[sh], xb[sh]))
      xa[sh] = 1, xb[sh] = 0;
    if (live && !window(shifted_coord(yq, sh, g.ly, g.bf, g.dyi), 1, g.jtot, w0, ya[sh], yb[sh]))
      ya[sh] = 1, yb[sh] = 0;
  }
  const int tk0 = live ? (za - tg.kbase) / tg.tk : 0;
  unsigned long long offp = 0, np_ = 0, l = 0;
  if (MODE == 1 && live) {
    long long lo = 0, hi = ncap - 1;  // capsule of point q: largest p with cap_off[p] <= q
    while (lo < hi) {
      long long mid = (lo + hi + 1) >> 1;
      if (cap_off[mid] <= q)
        lo = mid;
      else
        hi = mid - 1;
    }
    offp = (unsigned long long)cap_off[lo];
    np_ = (unsigned long long)(cap_off[lo + 1] - cap_off[lo]);
    l = (unsigned long long)q - offp;
  }
  for (int nb = 0; nb < 9; ++nb) {
    const int s = nb_slot(nb);
    const int xs = slot_xs(s), ys = slot_ys(s);
    const bool has = live && xa[xs] <= xb[xs] && ya[ys] <= yb[ys];
    if (!__ballot(has)) continue;
    unsigned long long key = 0;
    if (MODE == 1) {
      const unsigned long long K = 9ull * offp + (unsigned long long)nb * np_ + l;
      key = (K << 32) | (unsigned long long)((unsigned)q | ((unsigned)nb << 28));
    }
    const int cxa = has ? xa[xs] : 1, cxb = has ? xb[xs] : 0;
    const int cya = has ? ya[ys] : 1, cyb = has ? yb[ys] : 0;
    const int tj0 = (cya - 1) / tg.tj;
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      const int tj = tj0 + (rb & 1), tk = tk0 + (rb >> 1);
      // rows of this row-block inside the window
      const int j0 = max(cya, 1 + tj * tg.tj), j1 = min(cyb, (tj + 1) * tg.tj);
      const int k0 = max(za, tg.kbase + tk * tg.tk), k1 = min(zb, tg.kbase + (tk + 1) * tg.tk - 1);
      const bool rv = has && j0 <= j1 && k0 <= k1;
      int smin = SP_CW, smax = -1;
      if (rv)
        for (int k = k0; k <= k1; ++k)
          for (int j = j0; j <= j1; ++j) {
            const int ph = row_phase(g, j, k);
            smin = min(smin, ph), smax = max(smax, ph);
          }
      // candidate columns: the window's taps (<= 3 columns), the duplicate of cell itot
      // (column 0) and the duplicate of cell 1 (<= 2 columns at the row end)
      int col[6];
      bool cv[6];
      const int cm0 = (cxa + smin) / SP_CW, cm1 = (cxb + smax) / SP_CW;
      const int ch0 = (g.itot + 1 + smin) / SP_CW, ch1 = (g.itot + 1 + smax) / SP_CW;
#pragma unroll
      for (int u = 0; u < 3; ++u) col[u] = cm0 + u, cv[u] = rv && cm0 + u <= cm1;
      col[3] = 0, cv[3] = rv && cxb == g.itot;
      col[4] = ch0, cv[4] = rv && cxa == 1;
      col[5] = ch0 + 1, cv[5] = rv && cxa == 1 && ch0 + 1 <= ch1;
#pragma unroll
      for (int u = 3; u < 6; ++u)
#pragma unroll
        for (int w = 0; w < u; ++w)
          if (cv[w] && col[w] == col[u]) cv[u] = false;
#pragma unroll
      for (int u = 0; u < 6; ++u) {
        const bool val = cv[u];
        const int t = val ? col[u] + tg.nti * (tj + tg.ntj * tk) : 0;
        const unsigned pos = agg_add(count, t, val);
        if (MODE == 1 && val) {
          const unsigned long long at = (unsigned long long)tile_off[t] + pos;
          if (at < ecap)
            entries[at] = key;
          else
            atomicOr(err, ERR_CAPACITY);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// Exclusive scan of the per-tile counts, fused with the compaction of the non-empty tiles.
// Each tile contributes the packed value count | (count > 0) << 32, so one 64-bit scan gives
// both the list offset (low word) and the tile's rank among non-empty tiles (high word).
// Integer arithmetic: order-independent.
// ---------------------------------------------------------------------------------------
constexpr int SCAN_B = 1024;  // elements per block (256 threads x 4)

__device__ __forceinline__ unsigned long long block_excl_scan(unsigned long long v,
                                                              unsigned long long* sh,
                                                              unsigned long long& total) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    unsigned long long t = tid >= o ? sh[tid - o] : 0ull;
    __syncthreads();
    sh[tid] += t;
    __syncthreads();
  }
  total = sh[255];
  unsigned long long incl = sh[tid];
  __syncthreads();
  return incl - v;
}

__device__ __forceinline__ unsigned long long packc(unsigned c) {
  return (unsigned long long)c | (c ? (1ull << 32) : 0ull);
}

__global__ void __launch_bounds__(256) k_scan_reduce(const unsigned* __restrict__ in, long long n,
                                                     unsigned long long* __restrict__ part) {
  __shared__ unsigned long long sh[256];
  long long b0 = (long long)blockIdx.x * SCAN_B + threadIdx.x * 4;
  unsigned long long s = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r)
    if (b0 + r < n) s += packc(in[b0 + r]);
  unsigned long long tot;
  block_excl_scan(s, sh, tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(256) k_scan_parts(unsigned long long* __restrict__ part,
                                                    long long nb) {
  __shared__ unsigned long long sh[256];
  unsigned long long carry = 0;
  for (long long c = 0; c < nb; c += 1024) {
    long long b0 = c + threadIdx.x * 4;
    unsigned long long v[4], s = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      v[r] = (b0 + r < nb) ? part[b0 + r] : 0ull;
      s += v[r];
    }
    unsigned long long tot;
    unsigned long long ex = block_excl_scan(s, sh, tot) + carry;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (b0 + r < nb) part[b0 + r] = ex;
      ex += v[r];
    }
    carry += tot;
    __syncthreads();
  }
}

// off[0..n]: candidate list offsets; list[0..nne): non-empty tiles in increasing order;
// meta[0] = nne, meta[1..287] = 0 (work-queue counters of the spread kernel, reset here).
__global__ void __launch_bounds__(256) k_scan_apply(const unsigned* __restrict__ in, long long n,
                                                    const unsigned long long* __restrict__ part,
                                                    unsigned* __restrict__ off,
                                                    int* __restrict__ list,
                                                    int* __restrict__ meta) {
  __shared__ unsigned long long sh[256];
  long long b0 = (long long)blockIdx.x * SCAN_B + threadIdx.x * 4;
  unsigned c[4];
  unsigned long long s = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    c[r] = (b0 + r < n) ? in[b0 + r] : 0u;
    s += packc(c[r]);
  }
  unsigned long long tot;
  unsigned long long ex = block_excl_scan(s, sh, tot) + part[blockIdx.x];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    if (b0 + r < n) {
      off[b0 + r] = (unsigned)(ex & 0xffffffffull);
      if (c[r]) list[ex >> 32] = (int)(b0 + r);
    }
    ex += packc(c[r]);
  }
  if (b0 <= n - 1 && n - 1 < b0 + 4) {
    const unsigned long long all = part[blockIdx.x] + tot;
    off[n] = (unsigned)(all & 0xffffffffull);
    meta[0] = (int)(all >> 32);
    for (int q = 1; q < 9 * 32; ++q) meta[q] = 0;  // work-queue counters of the spread kernel
  }
}

// ---------------------------------------------------------------------------------------
// Spread, step 2: one workgroup per tile, for every tile of the grid, taken in tile order
// from one work queue per XCD over a contiguous range of tiles.  A tile without candidates
// is written as zeros.  A tile with candidates:
//   (a) loads its candidate entries and sorts them by K (rank sort in LDS; bitonic in place
//       in HBM beyond SP_CAPK);
//   (b) in chunks of SP_CH candidates: computes each candidate's window and weights, and
//       marks for every cell of the tile which chunk candidates touch it (one bit each,
//       integer LDS atomics -- the set is order-independent);
//   (c) every thread walks the set bits of its cells in ascending order, i.e. in K order, and
//       adds the terms to register accumulators: cell (i,j,k) receives exactly the
//       reference's sequence force = force + f*wx*wy*wz*(area/dVeul) (:869-874 / :1334-1338);
//   (d) writes the tile once: whole aligned 128-byte chunks, 16-byte stores.
// No floating-point atomics: results are deterministic and bit-identical to the reference.
// The j halo rows are copied afterwards by k_jhalo (bound.f90:172-177).
// ---------------------------------------------------------------------------------------
constexpr int SP_THREADS = 256; // threads per spread workgroup
constexpr int SP_CH = 64;       // candidates per chunk (one bit each in a 64-bit mask)
constexpr int SP_CAPK = 512;    // candidates sorted in LDS; larger lists are sorted in HBM
constexpr int SP_TG = 8;        // tiles per work item
constexpr int SP_QSTRIDE = 32;  // ints between the per-XCD queue counters (128 bytes)

__device__ __forceinline__ void cmpx(unsigned long long* a, int lo, int hi, int n) {
  if (hi < n) {
    unsigned long long x = a[lo], y = a[hi];
    if (x > y) {
      a[lo] = y;
      a[hi] = x;
    }
  }
}

// Ascending sort of a[0..n) with the "flip" form of the bitonic network: every
// compare-exchange puts the minimum at the lower index, so the virtual +inf elements
// a[n..P) never move and need no storage.  One workgroup; __syncthreads orders the
// workgroup's own accesses (LDS or HBM).
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

template <int TJ, int TK, bool CC>
struct SpreadTile {
  static constexpr int NR = TJ * TK;           // rows per tile
  static constexpr int NC = SP_CW * NR;        // cells per tile
  static constexpr int NP = NC / 2;            // cell pairs per tile
  static constexpr int PPT = NP / SP_THREADS;  // pairs per thread in the write
  static constexpr int CPT = NC / SP_THREADS;  // cells per thread in the per-cell passes
  static_assert(NP % SP_THREADS == 0, "tile must hold a multiple of 512 cells");
  static_assert(TJ >= 4 && TK >= 4 && TJ <= 64 && TK <= 64, "row block 4..64 each way");

  struct Smem {
    unsigned long long keys[SP_CAPK];
    unsigned long long mask[NC];
    double acc[3][NC];          // running sums of the tile's cells, per force component
    unsigned short order[NC];   // cells with set bits, by decreasing number of bits
    int hist[SP_CH + 1];        // per bit count: cells, then scatter cursors
    int nz;                     // cells with set bits in this chunk
    double4 fw[SP_CH][4];       // per candidate and x tap: f1*wx, f2*wx, f3*wx, area/dVeul
    double2 wyz[SP_CH][2][4];   // per candidate, y / z tap: (centred, staggered) weight
    int org[SP_CH];             // window origin: x (bits 0-15, +2^15), y, z rel. tile (+8)
    int xr[SP_CH];              // clamped x window: a | b << 16
    int box[SP_CH][2];          // clamped y, z window within the tile: a | b << 8
    int item;
  };

  // cell c of the tile (position p = c % 16, row r = c / 16): grid coordinates; returns
  // false for positions outside the row or rows outside the grid
  __device__ static bool cell_of(const GridD& g, const TileGeom& tg, int col, int j0, int k0,
                                 int c, int& i, int& vi, int& lj, int& lk, long long& m) {
    const int p = c % SP_CW, r = c / SP_CW;
    lj = r % TJ, lk = r / TJ;
    const int j = j0 + lj, k = k0 + lk;
    const long long base = (long long)j * g.sy + (long long)k * g.sz;
    const int s = (int)(base & (SP_CW - 1));
    i = SP_CW * col + p - s;
    m = base + i;
    vi = i == 0 ? g.itot : (i == g.itot + 1 ? 1 : i);
    return j <= g.jtot && k < tg.kbase + tg.nkp && i >= 0 && i <= g.itot + 1;
  }

  __device__ static void run(Smem& sm, long long t, int parts, const GridD& g,
                             const TileGeom& tg, const unsigned* __restrict__ tile_off,
                             unsigned long long* __restrict__ entries,
                             unsigned long long ecap, const double* __restrict__ x,
                             const double* __restrict__ y, const double* __restrict__ z,
                             const double* __restrict__ f1, const double* __restrict__ f2,
                             const double* __restrict__ f3, const double* __restrict__ area,
                             double* __restrict__ o1, double* __restrict__ o2,
                             double* __restrict__ o3) {
    const int tid = threadIdx.x;
    const int col = (int)(t % tg.nti);
    const int tj = (int)((t / tg.nti) % tg.ntj);
    const int tk = (int)(t / ((long long)tg.nti * tg.ntj));
    const int j0 = 1 + tj * TJ, k0 = tg.kbase + tk * TK;

    const unsigned base = tile_off[t];
    unsigned long long endl = tile_off[t + 1];
    if (endl > ecap) endl = ecap;  // overflowed lists are flagged by k_bin; never read past
    const int n = (unsigned long long)base < endl ? (int)(endl - base) : 0;
    if (!(parts & (n > 0 ? 1 : 2))) return;

    if (n > 0) {
#pragma unroll
      for (int r = 0; r < CPT; ++r)
        sm.acc[0][tid + r * SP_THREADS] = sm.acc[1][tid + r * SP_THREADS] =
            sm.acc[2][tid + r * SP_THREADS] = 0.0;
      // (a) K-sort of the tile's candidates
      const unsigned long long* src;
      if (n <= SP_THREADS) {
        // rank sort: keys are unique, rank = number of smaller keys
        unsigned long long* in = sm.keys + SP_THREADS;
        const unsigned long long mine = tid < n ? entries[base + tid] : 0ull;
        in[tid] = mine;
        __syncthreads();
        if (tid < n) {
          int rank = 0;
          for (int j = 0; j < n; ++j) rank += in[j] < mine ? 1 : 0;
          sm.keys[rank] = mine;
        }
        __syncthreads();
        src = sm.keys;
      } else if (n <= SP_CAPK) {
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
        if (tid <= SP_CH) sm.hist[tid] = 0;
        // (b) candidate preparation, 4 threads per candidate: x weights times the values,
        // y weights, z weights
        for (int c = tid >> 2; c < m; c += SP_THREADS / 4) {
          const int d = tid & 3;
          if (d == 3) continue;
          const unsigned long long e = src[c0 + c];
          const unsigned lo32 = (unsigned)(e & 0xffffffffull);
          const long long q = (long long)(lo32 & 0x0fffffffu);
          const int s = nb_slot((int)(lo32 >> 28));
          double cc_;
          int lo, hi, base_, ext;
          if (d == 0) {
            cc_ = shifted_coord(x[q], slot_xs(s), g.lx, g.bl, g.dxi);
            lo = 1, hi = g.itot, base_ = 0, ext = 1 << 30;
          } else if (d == 1) {
            cc_ = shifted_coord(y[q], slot_ys(s), g.ly, g.bf, g.dyi);
            lo = 1, hi = g.jtot, base_ = j0, ext = TJ;
          } else {
            cc_ = z[q] * g.dzi;
            lo = tg.klo, hi = tg.khi, base_ = k0, ext = TK;
          }
          Axis ax;
          int w0 = 0, bl = 1, bh = 0;
          double wn[4] = {0.0, 0.0, 0.0, 0.0}, ws[4] = {0.0, 0.0, 0.0, 0.0};
          if (axis_of(cc_, lo, hi, ax)) {
            const bool n0 = ax.na == 0;
            wn[0] = n0 ? ax.ns[0] : 0.0;
            wn[1] = n0 ? ax.ns[1] : ax.ns[0];
            wn[2] = n0 ? ax.ns[2] : ax.ns[1];
            wn[3] = n0 ? 0.0 : ax.ns[2];
            ws[1] = ax.st[0], ws[2] = ax.st[1], ws[3] = ax.st[2];
            int a, b;
            window(cc_, lo, hi, w0, a, b);
            w0 -= base_;
            bl = max(a - base_, 0);
            bh = min(b - base_, ext - 1);
          }
          if (d == 0) {
            const double v1 = f1[q], v2 = f2[q], v3 = f3[q];
            const double adv = area[q] / g.dveul;
#pragma unroll
            for (int u = 0; u < 4; ++u)
              sm.fw[c][u] = CC ? make_double4(v1 * ws[u], v2 * ws[u], v3 * ws[u], adv)
                               : make_double4(v1 * wn[u], v2 * ws[u], v3 * ws[u], adv);
            sm.xr[c] = bl | (bh << 16);
            // the origin's bytes are written separately by the x, y and z threads
            reinterpret_cast<unsigned short*>(&sm.org[c])[0] = (unsigned short)(w0 + 32768);
          } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) sm.wyz[c][d - 1][u] = make_double2(wn[u], ws[u]);
            sm.box[c][d - 1] = bl | (bh << 8);
            reinterpret_cast<unsigned char*>(&sm.org[c])[d + 1] = (unsigned char)(w0 + 8);
          }
        }
        __syncthreads();
        // (c) mark the cells each candidate touches: 4 threads per candidate, one x tap each,
        // over the window's rows within the tile; a tap at i = itot also marks the halo
        // duplicate at i = 0, a tap at i = 1 the one at i = itot+1
        for (int c = tid >> 2; c < m; c += SP_THREADS / 4) {
          const int part = tid & 3;
          const int xr = sm.xr[c];
          const int vi = (xr & 0xffff) + part;
          if (vi <= (xr >> 16)) {
            const unsigned long long bit = 1ull << c;
            const int by = sm.box[c][0], bz = sm.box[c][1];
            for (int lk = bz & 0xff; lk <= (bz >> 8); ++lk)
              for (int lj = by & 0xff; lj <= (by >> 8); ++lj) {
                const int s = row_phase(g, j0 + lj, k0 + lk);
                const int rowc = SP_CW * (lj + TJ * lk);
                const int p = vi + s - SP_CW * col;
                if (p >= 0 && p < SP_CW) atomicOr(&sm.mask[rowc + p], bit);
                if (vi == g.itot) {
                  const int ph = s - SP_CW * col;  // cell i = 0
                  if (ph >= 0 && ph < SP_CW) atomicOr(&sm.mask[rowc + ph], bit);
                }
                if (vi == 1) {
                  const int ph = g.itot + 1 + s - SP_CW * col;  // cell i = itot+1
                  if (ph >= 0 && ph < SP_CW) atomicOr(&sm.mask[rowc + ph], bit);
                }
              }
          }
        }
        __syncthreads();
        // (d) order the marked cells by decreasing bit count (a counting sort), so that the
        // lanes of a wavefront get cells with about equal numbers of terms
        int pc[CPT];
#pragma unroll
        for (int r = 0; r < CPT; ++r) {
          pc[r] = __popcll(sm.mask[tid + r * SP_THREADS]);
          if (pc[r]) atomicAdd(&sm.hist[SP_CH - pc[r]], 1);
        }
        __syncthreads();
        if (tid < 64) {  // exclusive scan of hist[0..SP_CH) by one wavefront
          const int v = tid < SP_CH ? sm.hist[tid] : 0;
          int incl = v;
#pragma unroll
          for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(incl, o);
            if (tid >= o) incl += u;
          }
          if (tid < SP_CH) sm.hist[tid] = incl - v;
          if (tid == 63) sm.nz = incl;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < CPT; ++r)
          if (pc[r])
            sm.order[atomicAdd(&sm.hist[SP_CH - pc[r]], 1)] = (unsigned short)(tid + r * SP_THREADS);
        __syncthreads();
        // (e) ordered accumulation: each cell adds its terms in ascending bit order, i.e. in
        // K order: cell (i,j,k) receives the reference's sequence
        // force = force + f*wx*wy*wz*(area/dVeul) (:869-874 / :1334-1338)
        const int nz = sm.nz;
        for (int rk = tid; rk < nz; rk += SP_THREADS) {
          const int cell = sm.order[rk];
          unsigned long long mk = sm.mask[cell];
          int i, vi, lj, lk;
          long long mm;
          cell_of(g, tg, col, j0, k0, cell, i, vi, lj, lk, mm);
          double a1 = sm.acc[0][cell], a2 = sm.acc[1][cell], a3 = sm.acc[2][cell];
          const int ox = vi + 32768, oy = lj + 8, oz = lk + 8;
          while (mk) {
            const int c = __builtin_ctzll(mk);
            mk &= mk - 1;
            const unsigned o = (unsigned)sm.org[c];
            const int a = ox - (int)(o & 0xffff), b = oy - (int)((o >> 16) & 0xff),
                      d = oz - (int)(o >> 24);
            const double4 fw = sm.fw[c][a];
            const double2 wy = sm.wyz[c][0][b], wz = sm.wyz[c][1][d];
            if (!CC) {
              a1 = a1 + fw.x * wy.y * wz.y * fw.w;
              a2 = a2 + fw.y * wy.x * wz.y * fw.w;
              a3 = a3 + fw.z * wy.y * wz.x * fw.w;
            } else {
              a1 = a1 + fw.x * wy.y * wz.y * fw.w;
              a2 = a2 + fw.y * wy.y * wz.y * fw.w;
              a3 = a3 + fw.z * wy.y * wz.y * fw.w;
            }
          }
          sm.acc[0][cell] = a1, sm.acc[1][cell] = a2, sm.acc[2][cell] = a3;
        }
        __syncthreads();
      }
    }
    // (f) write: each thread's pairs are two adjacent positions of one row, 16 bytes
#pragma unroll
    for (int h = 0; h < PPT; ++h) {
      const int cell = 2 * (tid + h * SP_THREADS);
      int i0, vi0, lj, lk, i1, vi1, lj1, lk1;
      long long m0, m1;
      const bool v0 = cell_of(g, tg, col, j0, k0, cell, i0, vi0, lj, lk, m0);
      const bool v1 = cell_of(g, tg, col, j0, k0, cell + 1, i1, vi1, lj1, lk1, m1);
      double2 r1 = make_double2(0.0, 0.0), r2 = r1, r3 = r1;
      if (n > 0) {
        r1 = *reinterpret_cast<const double2*>(&sm.acc[0][cell]);
        r2 = *reinterpret_cast<const double2*>(&sm.acc[1][cell]);
        r3 = *reinterpret_cast<const double2*>(&sm.acc[2][cell]);
      }
      if (v0 && v1 && (m0 & 1) == 0) {
        *reinterpret_cast<double2*>(o1 + m0) = r1;
        *reinterpret_cast<double2*>(o2 + m0) = r2;
        *reinterpret_cast<double2*>(o3 + m0) = r3;
      } else {
        if (v0) o1[m0] = r1.x, o2[m0] = r2.x, o3[m0] = r3.x;
        if (v1) o1[m1] = r1.y, o2[m1] = r2.y, o3[m1] = r3.y;
      }
    }
    if (n > 0) __syncthreads();  // the next tile reuses the accumulators
  }
};

// One persistent launch visits every tile.  Workgroup b runs on XCD b % 8 and takes work
// items (SP_TG consecutive tiles) from that XCD's queue counter (its own 128-byte line,
// meta[SP_QSTRIDE * (1 + xcd)]) over the XCD's contiguous share of the tiles; the next item
// is requested before the current one is processed.  Every workgroup leaves when its queue
// is drained; results do not depend on which workgroup takes which item.  `parts` selects
// the tiles (1: with candidates, 2: without, 3: all); a split launch does the same work in
// two launches (used to time the two parts).
template <int TJ, int TK, bool CC>
__global__ void __launch_bounds__(SP_THREADS)
    k_spread(GridD g, TileGeom tg, const unsigned* __restrict__ tile_off,
             int* __restrict__ meta, int parts, unsigned long long* __restrict__ entries,
             unsigned long long ecap, const double* __restrict__ x,
             const double* __restrict__ y, const double* __restrict__ z,
             const double* __restrict__ f1, const double* __restrict__ f2,
             const double* __restrict__ f3, const double* __restrict__ area,
             double* __restrict__ o1, double* __restrict__ o2, double* __restrict__ o3) {
  using T = SpreadTile<TJ, TK, CC>;
  __shared__ typename T::Smem sm;
  const int xcd = blockIdx.x & 7;
  const long long lo = (long long)xcd * tg.ntiles / 8, hi = (long long)(xcd + 1) * tg.ntiles / 8;
  const long long nitems = (hi - lo + SP_TG - 1) / SP_TG;
  int* ctr = meta + SP_QSTRIDE * (1 + xcd);
  if (threadIdx.x == 0) sm.item = atomicAdd(ctr, 1);
  __syncthreads();
  long long item = sm.item;
  __syncthreads();
  while (item < nitems) {
    if (threadIdx.x == 0) sm.item = atomicAdd(ctr, 1);  // request the next item
    for (int h = 0; h < SP_TG; ++h) {
      const long long t = lo + item * SP_TG + h;
      if (t >= hi) break;
      T::run(sm, t, parts, g, tg, tile_off, entries, ecap, x, y, z, f1, f2, f3, area, o1, o2,
             o3);
    }
    __syncthreads();
    item = sm.item;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------------
// Spread by capsule objects.  An object is one capsule p in one image slot (loop index nb),
// o = 9 p + nb; its candidates are the capsule's points in point order.  The reference adds
// the terms of a cell in the loop order capsule -> nb -> point (interp_spread.f90:825-878),
// i.e. in ascending (o, point) order, so an object needs neither binning nor a sort.
// Work items are sub-boxes (at most OB_NC cells) of an object's box, the union of its
// points' clamped windows.  An item takes as its sources every object whose box meets the
// item's box, in ascending o, and accumulates their candidates in that order: every cell of
// the item then receives all of its terms in the reference order, a cell lying in the boxes
// of several items gets the same value from each, and each item writes its whole box (zeros
// included).  Cells outside every item box are zeroed by the same persistent launch (zero
// units interleaved with the items), which finds them in a bitmap of the covered cells.
// Steps: k_obj_clear (bitmap, counters); k_obj_boxes (object boxes, items per capsule);
// k_obj_scan (item offsets); k_obj_items (items, bitmap, neighbouring capsules);
// k_obj_sources (source lists); k_obj_spread (items and zero units); k_jhalo.
// ---------------------------------------------------------------------------------------
constexpr int OB_BX = 32;          // item extent limit in x and y (cells)
constexpr int OB_NZ = 32;          // item extent limit in z
constexpr int OB_NC = 2048;        // cells per item
constexpr int OB_THREADS = 256;
constexpr int OB_CPT = OB_NC / OB_THREADS;  // cells per thread
constexpr int OB_PEND = 64 + OB_THREADS;    // pending candidate list
constexpr int OB_NSRC = 32;        // source objects per item
constexpr int OB_ZIT = 8;          // cell pairs per thread in one zero unit
constexpr long long OB_ZCELLS = 2LL * OB_THREADS * OB_ZIT;  // cells per zero unit
constexpr int OB_QBITS = 27;       // pending entry: point index | nb << OB_QBITS
constexpr long long OB_MAXPTS = 1LL << OB_QBITS;

struct ObjItem {
  int p, nb, x0, y0, z0, nx, ny, nz;
};

// counters in `meta` (OB_META ints): [0] items; [OB_QSTRIDE * (1 + x)] the item queue of
// XCD x = 0..7 (its own 128-byte line), which serves items [x n / 8, (x + 1) n / 8)
constexpr int OB_QSTRIDE = 32;
constexpr int OB_META = OB_QSTRIDE * 9;
// boxes: [nobj][6] clamped window bounds xa, xb, ya, yb, za, zb of object o = 9 p + nb
// (xa > xb: no candidate)

// the coverage bitmap (bit = linear cell index), the bin masks and the counters
__global__ void __launch_bounds__(256) k_obj_clear(long long nw, unsigned* __restrict__ cover,
                                                   long long nbw,
                                                   unsigned long long* __restrict__ binm,
                                                   int* __restrict__ meta) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nw) cover[q] = 0u;
  if (q < nbw) binm[q] = 0ull;
  if (q < OB_META) meta[q] = 0;
}

// Item grid of one object box: cx x cy x cz sub-boxes of extent nx x ny x nz (the last ones
// cut short), at most OB_BX x OB_BX x OB_NZ and OB_NC cells each.
struct ObjGrid {
  int cx, cy, cz, nx, ny, nz;
  __device__ explicit ObjGrid(const int* b) {
    const int X = b[1] - b[0] + 1, Y = b[3] - b[2] + 1, Z = b[5] - b[4] + 1;
    cx = (X + OB_BX - 1) / OB_BX, cy = (Y + OB_BX - 1) / OB_BX;
    nx = (X + cx - 1) / cx, ny = (Y + cy - 1) / cy;
    nz = max(1, min(min(Z, OB_NZ), OB_NC / (nx * ny)));
    cz = (Z + nz - 1) / nz;
  }
  __device__ int count() const { return cx * cy * cz; }
};

constexpr int OB_NBL = 256;  // bins of one capsule's boxes held in LDS
constexpr int OB_NQL = 512;  // candidate capsules per mask chunk held in LDS
constexpr int OB_NWC = 64;   // neighbour-mask words per chunk

// Coarse bins of OB_BIN^3 cells over i in [1,itot], j in [1,jtot], k in [0,k1]: bin mask
// binm[bin][0..nwd) has bit p set when a box of capsule p overlaps the bin.
constexpr int OB_BIN = 32;
struct ObjBins {
  int nx, ny, nz;
  __device__ explicit ObjBins(const GridD& g)
      : nx((g.itot + OB_BIN - 1) / OB_BIN), ny((g.jtot + OB_BIN - 1) / OB_BIN),
        nz((g.ktot + 2 + OB_BIN - 1) / OB_BIN) {}
  // bin ranges of a (non-empty) box: bx0..bx1, by0..by1, bz0..bz1
  __device__ void range(const int* b, int (&r)[6]) const {
    r[0] = (b[0] - 1) / OB_BIN, r[1] = (b[1] - 1) / OB_BIN;
    r[2] = (b[2] - 1) / OB_BIN, r[3] = (b[3] - 1) / OB_BIN;
    r[4] = b[4] / OB_BIN, r[5] = b[5] / OB_BIN;
  }
  __device__ long long index(int bx, int by, int bz) const {
    return (long long)bx + (long long)nx * (by + (long long)ny * bz);
  }
};

// One workgroup per capsule: the clamped window boxes of its nine objects (boxes[o], o =
// 9 p + nb), its registration in the coarse bins, and its number of items (cnt[p]).
__global__ void __launch_bounds__(256) k_obj_boxes(GridD g, int klo, int khi,
                                                   const long long* __restrict__ cap_off,
                                                   const double* __restrict__ x,
                                                   const double* __restrict__ y,
                                                   const double* __restrict__ z,
                                                   int* __restrict__ boxes,
                                                   int* __restrict__ cnt, int nwd,
                                                   unsigned long long* __restrict__ binm,
                                                   int* __restrict__ err) {
  if (offsets_bad(err)) return;
  __shared__ int red[4][9][6];
  __shared__ int ncnt[9];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const long long p = blockIdx.x;
  const long long lo = cap_off[p], hi = cap_off[p + 1];
  int bx[9][6];
#pragma unroll
  for (int nb = 0; nb < 9; ++nb)
    bx[nb][0] = bx[nb][2] = bx[nb][4] = 1 << 30, bx[nb][1] = bx[nb][3] = bx[nb][5] = -(1 << 30);
  for (long long q = lo + tid; q < hi; q += 256) {
    const double xq = x[q], yq = y[q], zq = z[q];
    if (!(isfinite(xq) && isfinite(yq) && isfinite(zq))) {
      atomicOr(err, ERR_NONFINITE);
      continue;
    }
    int w0, za, zb;
    if (!window(zq * g.dzi, klo, khi, w0, za, zb)) continue;
    int xa[3], xb[3], ya[3], yb[3];
#pragma unroll
    for (int sh = 0; sh < 3; ++sh) {
      if (!window(shifted_coord(xq, sh, g.lx, g.bl, g.dxi), 1, g.itot, w0, xa[sh], xb[sh]))
        xa[sh] = 1, xb[sh] = 0;
      if (!window(shifted_coord(yq, sh, g.ly, g.bf, g.dyi), 1, g.jtot, w0, ya[sh], yb[sh]))
        ya[sh] = 1, yb[sh] = 0;
    }
#pragma unroll
    for (int nb = 0; nb < 9; ++nb) {
      const int s = nb_slot(nb), xs = slot_xs(s), ys = slot_ys(s);
      if (xa[xs] <= xb[xs] && ya[ys] <= yb[ys]) {
        bx[nb][0] = min(bx[nb][0], xa[xs]), bx[nb][1] = max(bx[nb][1], xb[xs]);
        bx[nb][2] = min(bx[nb][2], ya[ys]), bx[nb][3] = max(bx[nb][3], yb[ys]);
        bx[nb][4] = min(bx[nb][4], za), bx[nb][5] = max(bx[nb][5], zb);
      }
    }
  }
#pragma unroll
  for (int nb = 0; nb < 9; ++nb) {
    if (__ballot(bx[nb][0] <= bx[nb][1]) == 0ull) {  // no lane of the wavefront: skip
      if (lane < 6) red[wv][nb][lane] = (lane & 1) ? -(1 << 30) : (1 << 30);
      continue;
    }
#pragma unroll
    for (int e = 0; e < 6; ++e) {
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const int v = __shfl_xor(bx[nb][e], o);
        bx[nb][e] = (e & 1) ? max(bx[nb][e], v) : min(bx[nb][e], v);
      }
      if (lane == 0) red[wv][nb][e] = bx[nb][e];
    }
  }
  __syncthreads();
  if (tid < 54) {
    const int nb = tid / 6, e = tid % 6;
    int v = red[0][nb][e];
    for (int w = 1; w < 4; ++w) v = (e & 1) ? max(v, red[w][nb][e]) : min(v, red[w][nb][e]);
    red[1][nb][e] = v;  // only this thread reads or writes element (nb, e) of red[1] here
  }
  __syncthreads();
  const int(&B)[9][6] = red[1];
  if (tid < 9) {
    const int nb = tid;
    const bool has = B[nb][0] <= B[nb][1];
    for (int e = 0; e < 6; ++e) boxes[(p * 9 + nb) * 6 + e] = has ? B[nb][e] : (e & 1 ? 0 : 1);
    ncnt[nb] = has ? ObjGrid(B[nb]).count() : 0;
    if (has) {
      const ObjBins bins(g);
      int r[6];
      bins.range(B[nb], r);
      const unsigned long long bit = 1ull << (p & 63);
      for (int bz = r[4]; bz <= r[5]; ++bz)
        for (int by = r[2]; by <= r[3]; ++by)
          for (int bx_ = r[0]; bx_ <= r[1]; ++bx_)
            atomicOr(&binm[bins.index(bx_, by, bz) * nwd + (p >> 6)], bit);
    }
  }
  __syncthreads();
  if (tid == 0) {
    int total = 0;
    for (int nb = 0; nb < 9; ++nb) total += ncnt[nb];
    cnt[p] = total;
  }
}

// One workgroup of 1024 threads: off[p] = exclusive prefix sum of cnt over capsules, meta[0]
// = the total number of items (flagged if above icap).
__global__ void __launch_bounds__(1024) k_obj_scan(long long ncap, const int* __restrict__ cnt,
                                                   int* __restrict__ off, int icap,
                                                   int* __restrict__ meta,
                                                   int* __restrict__ err) {
  if (offsets_bad(err)) return;
  __shared__ int ws[16];
  __shared__ int carry;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (long long b = 0; b < ncap; b += 1024) {
    const long long p = b + tid;
    const int v = p < ncap ? cnt[p] : 0;
    int s = v;  // inclusive scan within the wavefront
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(s, o);
      if (lane >= o) s += t;
    }
    if (lane == 63) ws[wv] = s;
    __syncthreads();
    int before = carry;
    for (int w = 0; w < wv; ++w) before += ws[w];
    if (p < ncap) off[p] = before + s - v;
    __syncthreads();
    if (tid == 1023) carry = before + s;
    __syncthreads();
  }
  if (tid == 0) {
    meta[0] = carry;
    if (carry > icap) atomicOr(err, ERR_CAPACITY);
  }
}

// Sets the bits of cells [xa, xb] of row (j, k) in the coverage bitmap.
__device__ __forceinline__ void obj_mark_row(const GridD& g, unsigned* __restrict__ cover,
                                             int xa, int xb, int j, int k) {
  const long long m0 = (long long)xa + (long long)j * g.sy + (long long)k * g.sz;
  const long long m1 = m0 + (xb - xa);
  for (long long w = m0 >> 5; w <= (m1 >> 5); ++w) {
    const long long a = max(m0, w << 5), b = min(m1, (w << 5) + 31);
    const unsigned lo = (unsigned)(a & 31), n = (unsigned)(b - a + 1);
    atomicOr(&cover[w], (unsigned)(((1ull << n) - 1ull) << lo));
  }
}

__device__ __forceinline__ bool boxes_meet(const int* a, const int* b) {
  return a[0] <= a[1] && b[0] <= b[1] && a[0] <= b[1] && a[1] >= b[0] && a[2] <= b[3] &&
         a[3] >= b[2] && a[4] <= b[5] && a[5] >= b[4];
}

// One workgroup per capsule p: its items (slots off[p] ... below icap; the items of an
// object partition its box, c fastest), its boxes' cells in the coverage bitmap, and its
// neighbour mask: bit q of nbm[p][0..nwd) is set iff a box of capsule q meets a box of p
// (p itself included when it has a box); the candidates q are the capsules registered in
// the bins that p's boxes overlap.
__global__ void __launch_bounds__(256) k_obj_items(GridD g, long long ncap, int nwd,
                                                   const int* __restrict__ boxes,
                                                   const unsigned long long* __restrict__ binm,
                                                   const int* __restrict__ off,
                                                   ObjItem* __restrict__ items, int icap,
                                                   unsigned* __restrict__ cover,
                                                   unsigned long long* __restrict__ nbm,
                                                   const int* __restrict__ err) {
  if (offsets_bad(err)) return;
  __shared__ int B[9][6];
  const int tid = threadIdx.x;
  const long long p = blockIdx.x;
  if (tid < 54) B[tid / 6][tid % 6] = boxes[p * 54 + tid];
  __syncthreads();
  int at = off[p];
  for (int nb = 0; nb < 9; ++nb) {
    if (B[nb][0] > B[nb][1]) continue;  // uniform
    const ObjGrid og(B[nb]);
    for (int t = tid; t < og.count(); t += 256) {
      if (at + t >= icap) break;  // flagged by k_obj_scan
      const int a = t / (og.cy * og.cz), b = (t / og.cz) % og.cy, c = t % og.cz;
      ObjItem it;
      it.p = (int)p, it.nb = nb;
      it.x0 = B[nb][0] + a * og.nx, it.y0 = B[nb][2] + b * og.ny, it.z0 = B[nb][4] + c * og.nz;
      it.nx = min(og.nx, B[nb][1] - it.x0 + 1);
      it.ny = min(og.ny, B[nb][3] - it.y0 + 1);
      it.nz = min(og.nz, B[nb][5] - it.z0 + 1);
      items[at + t] = it;
    }
    at += og.count();
    const int Y = B[nb][3] - B[nb][2] + 1, Z = B[nb][5] - B[nb][4] + 1;
    for (int r = tid; r < Y * Z; r += 256)
      obj_mark_row(g, cover, B[nb][0], B[nb][1], B[nb][2] + r % Y, B[nb][4] + r / Y);
  }
  // neighbour mask, in chunks of OB_NWC words: the capsules registered in the bins of p's
  // boxes (one (bin, word) pair per thread), then one (candidate, image slot) box test per
  // thread; a bin list or candidate list that would overflow LDS takes the serial path
  const ObjBins bins(g);
  __shared__ int blist[OB_NBL];
  __shared__ int qlist[OB_NQL];
  __shared__ unsigned long long cand[OB_NWC], res[OB_NWC];
  __shared__ int nbl, nq;
  if (tid == 0) {
    int n = 0;
    for (int nb = 0; nb < 9; ++nb) {
      if (B[nb][0] > B[nb][1]) continue;
      int r[6];
      bins.range(B[nb], r);
      for (int bz = r[4]; bz <= r[5]; ++bz)
        for (int by = r[2]; by <= r[3]; ++by)
          for (int bx = r[0]; bx <= r[1]; ++bx) {
            if (n < OB_NBL) blist[n] = (int)bins.index(bx, by, bz);
            ++n;
          }
    }
    nbl = n;
  }
  __syncthreads();
  for (int w0 = 0; w0 < nwd; w0 += OB_NWC) {
    const int wn = min(OB_NWC, nwd - w0);
    for (int t = tid; t < wn; t += 256) cand[t] = 0ull, res[t] = 0ull;
    if (tid == 0) nq = 0;
    __syncthreads();
    if (nbl <= OB_NBL) {
      for (int pr = tid; pr < nbl * wn; pr += 256) {
        const unsigned long long v = binm[(long long)blist[pr / wn] * nwd + w0 + pr % wn];
        if (v) atomicOr(&cand[pr % wn], v);
      }
    } else {
      for (int t = tid; t < wn; t += 256) cand[t] = ~0ull;  // every capsule is a candidate
    }
    __syncthreads();
    for (int t = tid; t < wn; t += 256) {
      unsigned long long m = cand[t];
      while (m) {
        const long long q = (long long)(w0 + t) * 64 + __builtin_ctzll(m);
        m &= m - 1;
        if (q >= ncap) break;
        const int k = atomicAdd(&nq, 1);
        if (k < OB_NQL) qlist[k] = (int)q;
      }
    }
    __syncthreads();
    if (nq <= OB_NQL) {
      for (int pr = tid; pr < 9 * nq; pr += 256) {
        const int q = qlist[pr / 9];
        const int* bq = boxes + (long long)q * 54 + (pr % 9) * 6;
        const int v[6] = {bq[0], bq[1], bq[2], bq[3], bq[4], bq[5]};
        bool hit = false;
        for (int n1 = 0; n1 < 9 && !hit; ++n1) hit = boxes_meet(B[n1], v);
        if (hit) atomicOr(&res[q / 64 - w0], 1ull << (q & 63));
      }
    } else {
      for (int t = tid; t < wn; t += 256) {
        unsigned long long m = cand[t], r = 0ull;
        while (m) {
          const int l = __builtin_ctzll(m);
          m &= m - 1;
          const long long q = (long long)(w0 + t) * 64 + l;
          if (q >= ncap) break;
          const int* bq = boxes + q * 54;
          bool hit = false;
          for (int n2 = 0; n2 < 9 && !hit; ++n2)
            for (int n1 = 0; n1 < 9 && !hit; ++n1) hit = boxes_meet(B[n1], bq + n2 * 6);
          if (hit) r |= 1ull << l;
        }
        res[t] = r;
      }
    }
    __syncthreads();
    for (int t = tid; t < wn; t += 256) nbm[p * nwd + w0 + t] = res[t];
    __syncthreads();
  }
}

// One wavefront per item (grid-stride): the objects whose boxes meet the item's box, in
// ascending o, into src[item][0..nsrc[item]).  Candidates: the neighbours of the item's
// capsule (the item's box lies in its capsule's union box).
__global__ void __launch_bounds__(256) k_obj_sources(int nwd, const int* __restrict__ boxes,
                                                     const unsigned long long* __restrict__ nbm,
                                                     const ObjItem* __restrict__ items,
                                                     int icap, int* __restrict__ src,
                                                     int* __restrict__ nsrc,
                                                     int* __restrict__ meta,
                                                     int* __restrict__ err) {
  if (offsets_bad(err)) return;
  const int lane = threadIdx.x & 63;
  const int nitems = min(meta[0], icap);
  const int waves = gridDim.x * (blockDim.x >> 6);
  for (int item = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); item < nitems;
       item += waves) {
    const ObjItem it = items[item];
    const int xa = it.x0, xb = it.x0 + it.nx - 1, ya = it.y0, yb = it.y0 + it.ny - 1,
              za = it.z0, zb = it.z0 + it.nz - 1;
    const unsigned long long* row = nbm + (long long)it.p * nwd;
    int n = 0;
    for (int w0 = 0; w0 < nwd; w0 += 64) {
      const unsigned long long word = w0 + lane < nwd ? row[w0 + lane] : 0ull;
      unsigned long long nzw = __ballot(word != 0ull);
      while (nzw) {  // words ascending, then bits ascending: capsules ascending
        const int l = __builtin_ctzll(nzw);
        nzw &= nzw - 1;
        const unsigned lo32 = (unsigned)__shfl((int)(unsigned)word, l),
                       hi32 = (unsigned)__shfl((int)(unsigned)(word >> 32), l);
        unsigned long long bits = ((unsigned long long)hi32 << 32) | lo32;
        while (bits) {
          const long long q = (long long)(w0 + l) * 64 + __builtin_ctzll(bits);
          bits &= bits - 1;
          bool h = false;
          if (lane < 9) {
            const int* b = boxes + (q * 9 + lane) * 6;
            h = b[0] <= b[1] && b[0] <= xb && b[1] >= xa && b[2] <= yb && b[3] >= ya &&
                b[4] <= zb && b[5] >= za;
          }
          const unsigned long long hb = __ballot(h);
          if (h) {
            const int pos = n + __popcll(hb & __lanemask_lt());
            if (pos < OB_NSRC) src[(long long)item * OB_NSRC + pos] = (int)(q * 9 + lane);
          }
          n += __popcll(hb);
        }
      }
    }
    if (lane == 0) {
      nsrc[item] = min(n, OB_NSRC);
      if (n > OB_NSRC) atomicOr(err, ERR_CAPACITY);
    }
  }
}

__device__ __forceinline__ bool obj_covered(const unsigned* __restrict__ cover, long long m) {
  return (cover[m >> 5] >> (m & 31)) & 1u;
}

// Zeros of the uncovered cells of the 16-cell group (linear index) that holds cell mc: a
// cell is covered when its bitmap bit is set, a halo cell i = 0 / i1 when the bit of its
// source itot / 1 is (the items write those duplicates).  The group's 16 bits lie in one
// bitmap word; a group holds at most one cell i = 0 and one i = i1 (i1 + 1 > 16).
__device__ __forceinline__ void obj_pad_group(const GridD& g, long long ncells,
                                              const unsigned* __restrict__ cover, long long mc,
                                              double* __restrict__ o1, double* __restrict__ o2,
                                              double* __restrict__ o3) {
  const long long g0 = mc & ~15LL;
  unsigned bits = (cover[g0 >> 5] >> (g0 & 16)) & 0xffffu;
  const int i0 = (int)(g0 % g.sy);
  const int e0 = i0 == 0 ? 0 : (int)g.sy - i0;  // offset of the cell i = 0
  const int e1 = (int)g.sy - 1 - i0;             // offset of the cell i = i1
  if (e0 < 16 && g0 + e0 < ncells)
    bits = (bits & ~(1u << e0)) | ((unsigned)obj_covered(cover, g0 + e0 + g.itot) << e0);
  if (e1 < 16 && g0 + e1 < ncells)
    bits = (bits & ~(1u << e1)) | ((unsigned)obj_covered(cover, g0 + e1 - g.itot) << e1);
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const long long m = g0 + e;
    if (!((bits >> e) & 1u) && m < ncells) o1[m] = 0.0, o2[m] = 0.0, o3[m] = 0.0;
  }
}

template <bool CC>
struct ObjSpread {
  struct Smem {
    unsigned long long mx[OB_BX], my[OB_BX], mz[OB_NZ];  // candidates covering x / y / z
    int pend[OB_PEND];          // point index | nb << OB_QBITS
    double4 fw[SP_CH][4];       // per candidate and x tap: f1*wx, f2*wx, f3*wx, area/dVeul
    double2 wyz[SP_CH][2][4];   // per candidate, y / z tap: (centred, staggered) weight
    int org[SP_CH];             // window origin relative to the item, x / y / z bytes, +8
    int rng[SP_CH][3];          // clamped window within the item, per direction: a | b << 8
    int src[OB_NSRC];
    int wsum[OB_THREADS / 64];
    int item, ns, zu;
    ObjItem it;
  };

  // The cells of one thread, tid + r OB_THREADS (r = 0 .. OB_CPT-1), as item coordinates
  // (li, lj, lk), x fastest: the decomposition of tid and of the OB_THREADS step are computed
  // once per item, and each step is an add with carries (li + sx < 2 nx, lj + sy + 1 < 2 ny),
  // which keeps integer divisions out of the per-chunk cell loop.
  struct CellWalk {
    int nx, ny, li0, lj0, lk0, sx, sy, sz;
    __device__ explicit CellWalk(const ObjItem& it) {
      nx = it.nx, ny = it.ny;
      const int tid = threadIdx.x, nxy = max(nx * ny, 1);
      li0 = nx ? tid % nx : 0, lj0 = nx && ny ? (tid / nx) % ny : 0, lk0 = tid / nxy;
      sx = nx ? OB_THREADS % nx : 0, sy = nx && ny ? (OB_THREADS / nx) % ny : 0;
      sz = OB_THREADS / nxy;
    }
    __device__ __forceinline__ void step(int& li, int& lj, int& lk) const {
      li += sx;
      const int cx = li >= nx;
      li -= cx ? nx : 0;
      lj += sy + cx;
      const int cy = lj >= ny;
      lj -= cy ? ny : 0;
      lk += sz + cy;
    }
  };

  // one chunk of m <= 64 candidates (sm.pend[0..m)), accumulated into acc
  __device__ static void chunk(Smem& sm, int m, const CellWalk& cw, const GridD& g, int klo,
                               int khi,
                               const double* __restrict__ x, const double* __restrict__ y,
                               const double* __restrict__ z, const double* __restrict__ f1,
                               const double* __restrict__ f2, const double* __restrict__ f3,
                               const double* __restrict__ area, double (&acc)[OB_CPT][3],
                               int* __restrict__ err) {
    const int tid = threadIdx.x;
    const ObjItem& it = sm.it;
    const int ncell = it.nx * it.ny * it.nz;
    for (int c = tid >> 2; c < m; c += OB_THREADS / 4) {
      const int d = tid & 3;
      if (d == 3) continue;
      const int e = sm.pend[c];
      const long long q = e & (int)(OB_MAXPTS - 1);
      const int s = nb_slot(e >> OB_QBITS);
      const double v1 = f1[q], v2 = f2[q], v3 = f3[q];
      const double adv = area[q] / g.dveul;
      const bool fin = isfinite(v1) && isfinite(v2) && isfinite(v3) && isfinite(adv);
      double cc_;
      int lo, hi, base_, ext;
      if (d == 0) {
        cc_ = shifted_coord(x[q], slot_xs(s), g.lx, g.bl, g.dxi);
        lo = 1, hi = g.itot, base_ = it.x0, ext = it.nx;
      } else if (d == 1) {
        cc_ = shifted_coord(y[q], slot_ys(s), g.ly, g.bf, g.dyi);
        lo = 1, hi = g.jtot, base_ = it.y0, ext = it.ny;
      } else {
        cc_ = z[q] * g.dzi;
        lo = klo, hi = khi, base_ = it.z0, ext = it.nz;
      }
      Axis ax;
      int w0 = 0, bl = 1, bh = 0;
      double wn[4] = {0.0, 0.0, 0.0, 0.0}, ws[4] = {0.0, 0.0, 0.0, 0.0};
      if (axis_of(cc_, lo, hi, ax)) {
        const bool n0 = ax.na == 0;
        wn[0] = n0 ? ax.ns[0] : 0.0;
        wn[1] = n0 ? ax.ns[1] : ax.ns[0];
        wn[2] = n0 ? ax.ns[2] : ax.ns[1];
        wn[3] = n0 ? 0.0 : ax.ns[2];
        ws[1] = ax.st[0], ws[2] = ax.st[1], ws[3] = ax.st[2];
        int a, b;
        window(cc_, lo, hi, w0, a, b);
        // tap w0 has weight 0 in all three components (ws[0] = 0, and wn[0] = 0 unless the
        // centred stencil starts there; the cell-centred form uses ws only), so its terms are
        // +-0 and leave every sum unchanged: it is left out.  (Exact for finite forces and
        // areas; non-finite ones are rejected below.)
        if ((CC || !n0) && a == w0) ++a;
        w0 -= base_;
        bl = max(a - base_, 0);
        bh = min(b - base_, ext - 1);
      }
      if (d == 0) {
        if (!fin) atomicOr(err, ERR_FORCE);
#pragma unroll
        for (int u = 0; u < 4; ++u)
          sm.fw[c][u] = CC ? make_double4(v1 * ws[u], v2 * ws[u], v3 * ws[u], adv)
                           : make_double4(v1 * wn[u], v2 * ws[u], v3 * ws[u], adv);
      } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) sm.wyz[c][d - 1][u] = make_double2(wn[u], ws[u]);
      }
      sm.rng[c][d] = bl | (bh << 8);
      reinterpret_cast<unsigned char*>(&sm.org[c])[d] = (unsigned char)(w0 + 8);
    }
    __syncthreads();
    // separable masks: candidate c touches cell (li, lj, lk) iff its clamped window covers li,
    // lj and lk, so the cell's mask is mx[li] & my[lj] & mz[lk]; wavefront w < 3 builds the
    // masks of direction w with one ballot per position (lane = candidate)
    {
      const int wv = tid >> 6, lane = tid & 63;
      if (wv < 3) {
        const int ext = wv == 0 ? it.nx : (wv == 1 ? it.ny : it.nz);
        unsigned long long* dst = wv == 0 ? sm.mx : (wv == 1 ? sm.my : sm.mz);
        const int r = lane < m ? sm.rng[lane][wv] : 0x00ff;  // a = 255 > b: covers nothing
        const int a = r & 0xff, b = r >> 8;
        for (int l = 0; l < ext; ++l) {
          const unsigned long long bal = __ballot(l >= a && l <= b);
          if (lane == 0) dst[l] = bal;
        }
      }
    }
    __syncthreads();
    // ordered accumulation: bits ascending = pending order = K order.  The empty asm makes the
    // walk's start opaque, so the compiler redoes the few adds per chunk instead of holding
    // all OB_CPT cell coordinates in registers across chunks.
    int li = cw.li0, lj = cw.lj0, lk = cw.lk0;
    asm volatile("" : "+v"(li), "+v"(lj), "+v"(lk));
#pragma unroll
    for (int r = 0; r < OB_CPT; ++r) {
      if (r) cw.step(li, lj, lk);
      const int cell = tid + r * OB_THREADS;
      if (cell >= ncell) continue;
      unsigned long long mk = sm.mx[li] & sm.my[lj] & sm.mz[lk];
      double a1 = acc[r][0], a2 = acc[r][1], a3 = acc[r][2];
      while (mk) {
        const int c = __builtin_ctzll(mk);
        mk &= mk - 1;
        const unsigned o = (unsigned)sm.org[c];
        const int a = li + 8 - (int)(o & 0xff), b = lj + 8 - (int)((o >> 8) & 0xff),
                  dd = lk + 8 - (int)((o >> 16) & 0xff);
        const double4 fw = sm.fw[c][a];
        const double2 wy = sm.wyz[c][0][b], wz = sm.wyz[c][1][dd];
        if (!CC) {
          a1 = a1 + fw.x * wy.y * wz.y * fw.w;
          a2 = a2 + fw.y * wy.x * wz.y * fw.w;
          a3 = a3 + fw.z * wy.y * wz.x * fw.w;
        } else {
          a1 = a1 + fw.x * wy.y * wz.y * fw.w;
          a2 = a2 + fw.y * wy.y * wz.y * fw.w;
          a3 = a3 + fw.z * wy.y * wz.y * fw.w;
        }
      }
      acc[r][0] = a1, acc[r][1] = a2, acc[r][2] = a3;
    }
    __syncthreads();
  }

  // The item in sm.it with its sources in sm.src[0..sm.ns): candidates of each source in
  // point order (order-preserving compaction of the points whose window meets the box),
  // accumulated in chunks of SP_CH; then the whole box is written.
  __device__ static void item(Smem& sm, const GridD& g, int klo, int khi,
                              const long long* __restrict__ cap_off,
                              const double* __restrict__ x, const double* __restrict__ y,
                              const double* __restrict__ z, const double* __restrict__ f1,
                              const double* __restrict__ f2, const double* __restrict__ f3,
                              const double* __restrict__ area,
                              const unsigned* __restrict__ cover, long long ncells,
                              double* __restrict__ o1, double* __restrict__ o2,
                              double* __restrict__ o3, int* __restrict__ err) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const ObjItem it = sm.it;
    const CellWalk cw(it);
    double acc[OB_CPT][3];
#pragma unroll
    for (int r = 0; r < OB_CPT; ++r) acc[r][0] = acc[r][1] = acc[r][2] = 0.0;
    int npend = 0;
    const int ns = sm.ns;
    for (int si = 0; si < ns; ++si) {
      const int o = sm.src[si];
      const int p = o / 9, nb = o - 9 * p;
      const int s = nb_slot(nb), xs = slot_xs(s), ys = slot_ys(s);
      const long long lo = cap_off[p], hi = cap_off[p + 1];
      for (long long b0 = lo; b0 < hi; b0 += OB_THREADS) {
        const long long q = b0 + tid;
        bool cand = false;
        if (q < hi) {
          const double xq = x[q], yq = y[q], zq = z[q];
          int w0, a, b;
          if (isfinite(xq) && isfinite(yq) && isfinite(zq) &&
              window(zq * g.dzi, klo, khi, w0, a, b) && a <= it.z0 + it.nz - 1 && b >= it.z0 &&
              window(shifted_coord(xq, xs, g.lx, g.bl, g.dxi), 1, g.itot, w0, a, b) &&
              a <= it.x0 + it.nx - 1 && b >= it.x0 &&
              window(shifted_coord(yq, ys, g.ly, g.bf, g.dyi), 1, g.jtot, w0, a, b) &&
              a <= it.y0 + it.ny - 1 && b >= it.y0)
            cand = true;
        }
        const unsigned long long bal = __ballot(cand);
        if (lane == 0) sm.wsum[wv] = __popcll(bal);
        __syncthreads();
        int before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < OB_THREADS / 64; ++w) {
          before += w < wv ? sm.wsum[w] : 0;
          total += sm.wsum[w];
        }
        if (cand) sm.pend[npend + before + __popcll(bal & __lanemask_lt())] =
            (int)q | (nb << OB_QBITS);
        npend += total;
        __syncthreads();
        while (npend >= SP_CH) {
          chunk(sm, SP_CH, cw, g, klo, khi, x, y, z, f1, f2, f3, area, acc, err);
          // move the rest (< OB_THREADS entries) to the front
          const int rest = npend - SP_CH;
          const int v = tid < rest ? sm.pend[SP_CH + tid] : 0;
          __syncthreads();
          if (tid < rest) sm.pend[tid] = v;
          __syncthreads();
          npend = rest;
        }
      }
    }
    if (npend > 0) chunk(sm, npend, cw, g, klo, khi, x, y, z, f1, f2, f3, area, acc, err);
    const int ncell = it.nx * it.ny * it.nz;
    int li = cw.li0, lj = cw.lj0, lk = cw.lk0;
#pragma unroll
    for (int r = 0; r < OB_CPT; ++r) {
      if (r) cw.step(li, lj, lk);
      const int cell = tid + r * OB_THREADS;
      if (cell >= ncell) continue;
      const int i = it.x0 + li;
      const long long mm =
          (long long)i + (long long)(it.y0 + lj) * g.sy + (long long)(it.z0 + lk) * g.sz;
      // the cell and, for i = itot or i = 1, its periodic x-halo duplicate (bound.f90:141-171:
      // var(0) = var(imax), var(i1) = var(1)); the y halos are copied later by k_jhalo
      // (ordinary stores: partial lines merge in L2 with the rest of their group)
      o1[mm] = acc[r][0], o2[mm] = acc[r][1], o3[mm] = acc[r][2];
      if (i == g.itot)
        o1[mm - g.itot] = acc[r][0], o2[mm - g.itot] = acc[r][1], o3[mm - g.itot] = acc[r][2];
      if (i == 1)
        o1[mm + g.itot] = acc[r][0], o2[mm + g.itot] = acc[r][1], o3[mm + g.itot] = acc[r][2];
    }
    // the uncovered cells of the 16-cell groups this item writes in part (zero units skip
    // every group that holds a covered cell): per row of the box, the groups of its first
    // and last cell and of its x-halo duplicates; the groups between lie inside the box
    const int rows = it.ny * it.nz;
    const bool h0 = it.x0 + it.nx - 1 == g.itot, h1 = it.x0 == 1;
    for (int t = tid; t < 4 * rows; t += OB_THREADS) {
      const int r = t >> 2, slot = t & 3;
      if ((slot == 2 && !h0) || (slot == 3 && !h1)) continue;
      const long long rb =
          (long long)(it.y0 + r % it.ny) * g.sy + (long long)(it.z0 + r / it.ny) * g.sz;
      const long long mc = slot == 0   ? rb + it.x0
                           : slot == 1 ? rb + it.x0 + it.nx - 1
                           : slot == 2 ? rb
                                       : rb + g.itot + 1;
      obj_pad_group(g, ncells, cover, mc, o1, o2, o3);
    }
  }
};

// Zero unit u: cells [u OB_ZCELLS, (u+1) OB_ZCELLS) of the linear index, as pairs
// (2 tid + 2 OB_THREADS e, +1), eight lanes per 16-cell group: the cells of every group
// without a covered cell get 0 (covered: the bitmap bit is set; for the halo cells i = 0 and
// i1, the bit of their sources itot and 1, whose duplicates the items write).  Groups with a
// covered cell are completed by the items, so that every 128-byte line is written whole by
// one side.  All bitmap words are read before the stores.
__device__ __forceinline__ void obj_zero_unit(const GridD& g, long long ncells, long long u,
                                              const unsigned* __restrict__ cover,
                                              double* __restrict__ o1,
                                              double* __restrict__ o2,
                                              double* __restrict__ o3) {
  const long long m0 = u * OB_ZCELLS + 2 * threadIdx.x;
  const int sy = (int)g.sy, step = (int)((2LL * OB_THREADS) % g.sy);
  int i = (int)(m0 % g.sy);
  bool za[OB_ZIT], zb[OB_ZIT];
#pragma unroll
  for (int e = 0; e < OB_ZIT; ++e) {
    const long long m = m0 + 2LL * OB_THREADS * e;
    const int i1 = i + 1 == sy ? 0 : i + 1;
    const long long sa = i == 0 ? m + g.itot : (i == sy - 1 ? m - g.itot : m);
    const long long sb = i1 == 0 ? m + 1 + g.itot : (i1 == sy - 1 ? m + 1 - g.itot : m + 1);
    // whole 16-cell groups only (8 lanes): a group with a covered cell is left to the items
    const bool ca = m < ncells && obj_covered(cover, sa);
    const bool cb = m + 1 < ncells && obj_covered(cover, sb);
    const unsigned long long bal = __ballot(ca || cb);
    const bool any = ((bal >> (threadIdx.x & 56)) & 0xffull) != 0ull;
    za[e] = m < ncells && !any;
    zb[e] = m + 1 < ncells && !any;
    i += step;
    if (i >= sy) i -= sy;
  }
#pragma unroll
  for (int e = 0; e < OB_ZIT; ++e) {
    const long long m = m0 + 2LL * OB_THREADS * e;
    if (za[e] && zb[e]) {
      // streaming stores: these lines are not read again by this launch
      typedef double dv2 __attribute__((ext_vector_type(2)));
      const dv2 zz = {0.0, 0.0};
      __builtin_nontemporal_store(zz, reinterpret_cast<dv2*>(o1 + m));
      __builtin_nontemporal_store(zz, reinterpret_cast<dv2*>(o2 + m));
      __builtin_nontemporal_store(zz, reinterpret_cast<dv2*>(o3 + m));
    } else {
      if (za[e]) o1[m] = 0.0, o2[m] = 0.0, o3[m] = 0.0;
      if (zb[e]) o1[m + 1] = 0.0, o2[m + 1] = 0.0, o3[m + 1] = 0.0;
    }
  }
}

// Persistent.  Each workgroup alternates: one item, then up to `per` of its own zero units (units blockIdx.x, blockIdx.x + gridDim.x, ...: a static
// share, no counter to contend for); once the items are gone, the rest of its units.  Items
// and zero units write disjoint cells, so their order does not matter.  MINW: the minimum wavefronts
// per SIMD the register allocation must allow (1: unconstrained).
template <bool CC, int MINW>
__global__ void __launch_bounds__(OB_THREADS, MINW)
    k_obj_spread(GridD g, int klo, int khi, const long long* __restrict__ cap_off,
                 const double* __restrict__ x, const double* __restrict__ y,
                 const double* __restrict__ z, const double* __restrict__ f1,
                 const double* __restrict__ f2, const double* __restrict__ f3,
                 const double* __restrict__ area, const ObjItem* __restrict__ items,
                 int icap, const int* __restrict__ src, const int* __restrict__ nsrc,
                 const unsigned* __restrict__ cover, long long ncells,
                 int* __restrict__ meta, double* __restrict__ o1, double* __restrict__ o2,
                 double* __restrict__ o3, int* __restrict__ err) {
  if (offsets_bad(err)) return;
  using O = ObjSpread<CC>;
  __shared__ typename O::Smem sm;
  const int tid = threadIdx.x;
  // an overflowed item list was flagged by k_obj_boxes; never read past its capacity
  const int nitems = min(meta[0], icap);
  const long long nzu = (ncells + OB_ZCELLS - 1) / OB_ZCELLS;
  const long long mine = nzu > blockIdx.x ? (nzu - blockIdx.x + gridDim.x - 1) / gridDim.x : 0;
  const int per = nitems > 0 ? (int)((mine * gridDim.x + nitems - 1) / nitems) : 0;
  long long zu = blockIdx.x;
  const int xcd = blockIdx.x & 7;
  int qi = 0;  // queues tried so far (thread 0's view)
  bool items_left = nitems > 0, zero_left = mine > 0;
  while (items_left || zero_left) {
    if (items_left) {
      // items: from the queue of this workgroup's XCD (a contiguous share, so that a
      // capsule's items and their points stay in one L2), then from the other XCDs' queues
      if (tid == 0) {
        int item = -1;
        while (qi < 8) {
          const int x = (xcd + qi) & 7;
          const int lo = (int)((long long)x * nitems / 8), hi = (int)((long long)(x + 1) * nitems / 8);
          const int i = lo + atomicAdd(&meta[OB_QSTRIDE * (1 + x)], 1);
          if (i < hi) {
            item = i;
            break;
          }
          ++qi;
        }
        sm.item = item;
        if (item >= 0) sm.it = items[item], sm.ns = nsrc[item];
      }
      __syncthreads();
      const int item = sm.item;
      if (item >= 0) {
        if (tid < sm.ns) sm.src[tid] = src[(long long)item * OB_NSRC + tid];
        __syncthreads();
        O::item(sm, g, klo, khi, cap_off, x, y, z, f1, f2, f3, area, cover, ncells, o1, o2,
                o3, err);
      } else {
        items_left = false;
      }
      __syncthreads();
    }
    if (zero_left) {
      for (int t = 0; !items_left || t < per; ++t) {
        if (zu >= nzu) {
          zero_left = false;
          break;
        }
        obj_zero_unit(g, ncells, zu, cover, o1, o2, o3);
        zu += gridDim.x;
      }
    }
  }
}

// Flags a non-finite force or area of any point (the tile spread's check; the object spread
// checks its candidates while preparing them).
__global__ void __launch_bounds__(256) k_check_forces(long long npts,
                                                      const double* __restrict__ f1,
                                                      const double* __restrict__ f2,
                                                      const double* __restrict__ f3,
                                                      const double* __restrict__ area,
                                                      int* __restrict__ err) {
  bool bad = false;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < npts;
       q += (long long)gridDim.x * blockDim.x)
    bad |= !(isfinite(f1[q]) && isfinite(f2[q]) && isfinite(f3[q]) && isfinite(area[q]));
  if (__ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicOr(err, ERR_FORCE);
}

// Wall halo planes k = 0 and k = k1 of lagr2eulr: zero (the scatter never reaches them and
// updthalos copies only x and y halos, bound.f90:141-190).  Two cells per thread.
__global__ void __launch_bounds__(256) k_zero_kplanes(GridD g, double* __restrict__ o1,
                                                      double* __restrict__ o2,
                                                      double* __restrict__ o3) {
  const long long n = g.sz;  // cells per plane
  const long long q = 2 * ((long long)blockIdx.x * blockDim.x + threadIdx.x);
  if (q >= 2 * n) return;
  const long long idx = q < n ? q : q - n + (long long)(g.ktot + 1) * g.sz;
  if ((idx & 1) == 0 && q + 1 < 2 * n && (q < n) == (q + 1 < n)) {
    *reinterpret_cast<double2*>(o1 + idx) = make_double2(0.0, 0.0);
    *reinterpret_cast<double2*>(o2 + idx) = make_double2(0.0, 0.0);
    *reinterpret_cast<double2*>(o3 + idx) = make_double2(0.0, 0.0);
  } else {
    for (long long e = q; e < q + 2 && e < 2 * n; ++e) {
      const long long id = e < n ? e : e - n + (long long)(g.ktot + 1) * g.sz;
      o1[id] = 0.0, o2[id] = 0.0, o3[id] = 0.0;
    }
  }
}

// j halo rows (bound.f90:172-177): var(:,j1,k) = var(:,1,k), var(:,0,k) = var(:,jmax,k) for
// every plane k in [kfirst, klast], whole rows i = 0..i1 (the tiles computed the i halos).
__global__ void __launch_bounds__(256) k_jhalo(GridD g, int kfirst, int klast,
                                               double* __restrict__ o1,
                                               double* __restrict__ o2,
                                               double* __restrict__ o3) {
  const long long ncol = g.sy;
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per = 2 * ncol;
  const long long k = kfirst + q / per;
  if (k > klast) return;
  const long long rem = q - (k - kfirst) * per;
  const bool top = rem >= ncol;  // row j1 <- row 1; else row 0 <- row jtot
  const long long i = top ? rem - ncol : rem;
  const long long dst = i + (top ? (long long)(g.jtot + 1) : 0) * g.sy + k * g.sz;
  const long long srcx = i + (top ? 1ll : (long long)g.jtot) * g.sy + k * g.sz;
  o1[dst] = o1[srcx], o2[dst] = o2[srcx], o3[dst] = o3[srcx];
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

// collisions.f90:10-32: the Morse wall force of the bottom wall added to fz of every point.
// a2 = 2 col / Rep^2, rad0 the cutoff and zero-force distance, betacol the Morse width.
// The plane sum over j, k in [-jm, jm], jm = nint(sqrt(max(rad0^2 - z^2, 0)) dyi), in the
// reference's loop order (j outer, k inner; j steps dy, k steps dx) and operation order.
// Non-finite z gives jm = 0 and no term (the comparison ss < rad0 fails).
__global__ void __launch_bounds__(256) k_collisions(GridD g, long long npts,
                                                    const double* __restrict__ z,
                                                    const double* __restrict__ area,
                                                    double* __restrict__ fz, double dx,
                                                    double dy, double a2, double rad0,
                                                    double betacol) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npts) return;
  const double zq = z[q];
  const double r2 = rad0 * rad0 - zq * zq;
  const double rj = sqrt(r2 > 0.0 ? r2 : 0.0) * g.dyi;
  const int jm = (int)round(rj);  // Fortran nint: half away from zero
  const double a = area[q], dx2 = dx * dx;
  const double lo = a < dx2 ? a : dx2, hi = a > dx2 ? a : dx2;
  double fc = 0.0;
  for (int j = -jm; j <= jm; ++j) {
    const double jy = (double)j * dy;
    for (int k = -jm; k <= jm; ++k) {
      const double kx = (double)k * dx;
      const double ss = sqrt(zq * zq + jy * jy + kx * kx);
      if (ss < rad0) {
        const double e2 = exp(2.0 * betacol * (rad0 - ss));
        const double e1 = exp(betacol * (rad0 - ss));
        fc = fc + a2 * (e2 - e1) * fabs(zq / ss) * (lo / hi);
      }
    }
  }
  fz[q] = fz[q] + fc;
}

// z[q] = z[q] + s for every point: the in-place shift of ap(p)%zfp around the indicator's
// spread (indicator.f90:25-31 with s = (ktot/4) dz, :37-43 with -s).
__global__ void __launch_bounds__(256) k_shift_z(long long npts, double* __restrict__ z,
                                                 double s) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < npts) z[q] = z[q] + s;
}

// One workgroup per capsule: the 256 threads stage 256 consecutive points at a time in LDS
// (coalesced loads), and thread c < 6 adds component c of them in point order, so the sums
// are the reference's sequential ones (:135-145); then the shift decision and the shift.
__global__ void __launch_bounds__(256) k_centroid(GridD g, long long ncap,
                                                  const long long* __restrict__ off,
                                                  double* __restrict__ x,
                                                  double* __restrict__ y,
                                                  const double* __restrict__ z,
                                                  const double* __restrict__ ul,
                                                  const double* __restrict__ vl,
                                                  const double* __restrict__ wl,
                                                  double* __restrict__ cen,
                                                  double* __restrict__ um,
                                                  const int* __restrict__ err) {
  if (offsets_bad(err)) return;
  __shared__ double st[6][256];
  __shared__ double sum[6];
  const long long p = blockIdx.x;
  const int tid = threadIdx.x;
  const long long a = off[p], b = off[p + 1];
  double acc = 0.0;  // thread c < 6: component c (x, y, z, u, v, w)
  for (long long c0 = a; c0 < b; c0 += 256) {
    const long long q = c0 + tid;
    if (q < b) {
      st[0][tid] = x[q], st[1][tid] = y[q], st[2][tid] = z[q];
      st[3][tid] = ul[q], st[4][tid] = vl[q], st[5][tid] = wl[q];
    }
    __syncthreads();
    const int cnt = (int)(b - c0 < 256 ? b - c0 : 256);
    if (tid < 6)
      for (int l = 0; l < cnt; ++l) acc = acc + st[tid][l];
    __syncthreads();
  }
  if (tid < 6) sum[tid] = acc;
  __syncthreads();
  double nl = (double)(b - a);
  double cx = sum[0], cy = sum[1], cz = sum[2], mu = sum[3], mv = sum[4], mw = sum[5];
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
  if (tid == 0) {
    cen[3 * p] = cx, cen[3 * p + 1] = cy, cen[3 * p + 2] = cz;
    um[3 * p] = mu, um[3 * p + 1] = mv, um[3 * p + 2] = mw;
  }
  // the periodic shift of the whole capsule (:147-186), by the workgroup that owns its points
  // (all loads above precede these stores); per point the same operation sequence as the
  // reference's two successive loops
  if (sx == 0 && sy == 0) return;  // uniform
  for (long long q = a + tid; q < b; q += 256) {
    if (sx == -1 || sx == 2) x[q] = x[q] - g.lx;
    if (sx == 1 || sx == 2) x[q] = x[q] + g.lx;
    if (sy == -1 || sy == 2) y[q] = y[q] - g.ly;
    if (sy == 1 || sy == 2) y[q] = y[q] + g.ly;
  }
}

}  // namespace ibx
