// fluid.hip -- libfluid: the FFT flow update of one rank on gfx950 (declarations and reference
// citations in include/fluid.h).
//
// Layout.  The caller's fields are Fortran (0:i1,0:j1,0:k1) arrays, i fastest.  The solves work
// in compact spectral arrays of itot x jtot x ktot doubles (i fastest), one per velocity
// component; the pressure solve reuses the first.  A solve is five passes:
//   1. x transform of every (j, k) line (k_rfft, lines contiguous),
//   2. y transform of every (i, k) line (k_rfft, 8 neighbouring columns per workgroup: row
//      segments of 64 bytes),
//   3. Gaussian elimination in k per (i, j) mode (k_tri_seg: one lane per mode, the k rows in
//      register-held segments joined through LDS, the per-mode reciprocal pivots read from a
//      table built once per context; k_tri, two passes through the array, above 257 planes),
//   4. y inverse, 5. x inverse (into the caller's field interior).
// Transforms.  A real line of length N = 2M is transformed as M complex values
// z_n = x_2n + i x_2n+1 (radix-2, in LDS, input bit-reversed on load), then split into the
// half spectrum X_0 .. X_M.  Packed real layout of a line: [X_0, X_M, Re X_1, Im X_1, ...,
// Re X_M-1, Im X_M-1]; position p carries wavenumber 0 (p = 0), M (p = 1) or p / 2.  The
// second-difference eigenvalue of a position depends only on its wavenumber, which is what the
// reference's xrt/yrt tables encode for VFFTPACK's layout (initsolver.f90:108-132).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fluid.h"

#pragma clang fp contract(off)

namespace {

constexpr int FFT_THREADS = 256;
// loads each thread keeps in flight while a transform reads its lines (8 for the forward
// transforms; 4 for the inverse ones, whose four loads per value would cost occupancy at 8)
constexpr int FFT_LB_FWD = 8;
constexpr int FFT_LB_INV = 4;
constexpr int TRI_MODES = 64;        // modes (lanes) per elimination workgroup
constexpr int Y_LINES_FWD = 8;       // columns per forward y-transform workgroup
constexpr int Y_LINES_INV = 8;       // columns per inverse y-transform workgroup
constexpr size_t LDS_MAX = 160 * 1024;

struct FGrid {
  int itot, jtot, ktot;
  long long sy, sz;  // Fortran field pitches: row itot+2, plane (itot+2)(jtot+2)
  double dxi, dyi, dzi;
};

__device__ __forceinline__ long long fidx(const FGrid& g, int i, int j, int k) {
  return (long long)i + g.sy * j + g.sz * k;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// ---- transforms ---------------------------------------------------------------------------
// One workgroup transforms `nl` lines of length N = 2M of one k plane; element e of line l
// lives at base + l * ls + e * es (source and destination each have their own geometry), the
// workgroup's base being plane * (block / groups_per_plane) + gstride * (block % groups).
// tw[k] = exp(-2 pi i k / N), k < M.  x lines: nl consecutive rows; y lines: nl consecutive
// columns.

// radix-2 decimation in time over nl lines of M complex values (input in bit-reversed order)
// Line l starts at l * (M + 1) complex values in LDS: the pad of one keeps the lines of a
// workgroup in different banks when neighbouring threads work on neighbouring lines.
// Decimation in time over nl lines of M complex values (input in bit-reversed order): one
// radix-2 stage when log2 M is odd, then radix-4 passes, each the two radix-2 stages of spans
// h and 2h on four values held in registers (half the LDS round trips and barriers).
template <bool INV>
__device__ void fft_stages(double2* s, int nl, int M, int logm, int logn,
                           const double2* __restrict__ tw) {
  int h = 1, lg = 0;
  if (logm & 1) {
    __syncthreads();
    const int half_m = M >> 1;
    for (int b = threadIdx.x; b < nl * half_m; b += blockDim.x) {
      const int line = b >> (logm - 1), r = b & (half_m - 1);
      const int i0 = line * (M + 1) + 2 * r, i1 = i0 + 1;
      const double2 a = s[i0], t = s[i1];  // twiddle W^0 = 1
      s[i0] = make_double2(a.x + t.x, a.y + t.y);
      s[i1] = make_double2(a.x - t.x, a.y - t.y);
    }
    h = 2, lg = 1;
  }
  const int quarter = M >> 2;
  for (; h < M; h <<= 2, lg += 2) {
    __syncthreads();
    const int ts1 = logn - 1 - lg, ts2 = logn - 2 - lg;  // N / (2h), N / (4h)
    for (int b = threadIdx.x; b < nl * quarter; b += blockDim.x) {
      const int line = b >> (logm - 2), r = b & (quarter - 1);
      const int pos = r & (h - 1), grp = r >> lg;
      const int base = line * (M + 1) + (grp << (lg + 2)) + pos;
      double2 w1 = tw[pos << ts1], w2 = tw[pos << ts2], w3 = tw[(pos + h) << ts2];
      if (INV) w1.y = -w1.y, w2.y = -w2.y, w3.y = -w3.y;
      const double2 a0 = s[base], a1 = s[base + h], a2 = s[base + 2 * h], a3 = s[base + 3 * h];
      const double2 t1 = cmul(w1, a1), t3 = cmul(w1, a3);
      const double2 b0 = make_double2(a0.x + t1.x, a0.y + t1.y);
      const double2 b1 = make_double2(a0.x - t1.x, a0.y - t1.y);
      const double2 b2 = make_double2(a2.x + t3.x, a2.y + t3.y);
      const double2 b3 = make_double2(a2.x - t3.x, a2.y - t3.y);
      const double2 u2 = cmul(w2, b2), u3 = cmul(w3, b3);
      s[base] = make_double2(b0.x + u2.x, b0.y + u2.y);
      s[base + 2 * h] = make_double2(b0.x - u2.x, b0.y - u2.y);
      s[base + h] = make_double2(b1.x + u3.x, b1.y + u3.y);
      s[base + 3 * h] = make_double2(b1.x - u3.x, b1.y - u3.y);
    }
  }
  __syncthreads();
}

// Workgroup b runs on XCD b % 8 (round-robin dispatch).  With a grid of a multiple of 8
// workgroups, the y transforms take logical block (b % 8) (grid / 8) + b / 8, so that each XCD
// works through one contiguous eighth of the lines: neighbouring groups of y lines share their
// 128-byte rows, and these are then fetched once, into one L2 (y transforms 941 -> 883 us
// forward, 874 -> 739 us inverse at 1024 x 512 x 256).  The x lines are whole rows and keep
// the dispatch order.
template <bool XMAJOR>
__device__ __forceinline__ int xcd_logical_block() {
  const unsigned nb = gridDim.x;
  if (XMAJOR || (nb & 7u)) return (int)blockIdx.x;
  return (int)((blockIdx.x & 7u) * (nb >> 3) + (blockIdx.x >> 3));
}

__device__ __forceinline__ int brev(int n, int logm) {
  return logm ? (int)(__brev((unsigned)n) >> (32 - logm)) : 0;
}

// The velocity divergence of fillps.f90:8-47 as the source of an x transform: the value of
// compact cell (i, j, k) is computed from the fields while the lines are loaded, with
// k_fillps's expression (cz = dzi / dt, cy = dyi / dt, cx = dxi / dt as k_fillps forms them).
struct FillSrc {
  const double* u;
  const double* v;
  const double* w;
  double cx, cy, cz;
  long long sy, sz;
};

// Forward: real lines -> packed half spectra.  XMAJOR: element index fastest across threads
// (x lines); otherwise line index fastest (y lines).  src and dst may be the same array (the
// workgroup reads all of its lines before the first barrier and writes them after the last).
// FILL (x lines of a compact array only): the lines are fillps's divergence (fs), not src.
template <bool XMAJOR, bool FILL = false>
__global__ void __launch_bounds__(FFT_THREADS)
    k_rfft_fwd(const double* src, double* dst, int N, int logm,
               int nl, int groups_per_plane, long long plane_src, long long plane_dst,
               long long gstride_src, long long gstride_dst, long long es_src, long long ls_src,
               long long es_dst, long long ls_dst, long long off_src,
               const double2* __restrict__ tw, FillSrc fs) {
  extern __shared__ double2 sm[];
  const int M = N >> 1, logn = logm + 1, lognl = 31 - __clz(nl);
  const int blk = xcd_logical_block<XMAJOR>();
  const int kp = blk / groups_per_plane, gp = blk - kp * groups_per_plane;
  const long long bs = off_src + kp * plane_src + gp * gstride_src;
  const long long bd = kp * plane_dst + gp * gstride_dst;
  double* smd = reinterpret_cast<double*>(sm);
  // the y transforms read their twiddles from an LDS copy after the lines (each radix-4
  // butterfly needs three per pass; from the caches they put a round trip in every pass);
  // the x transforms, whose forward pass was slower that way, read them from memory
  const double2* twl = tw;
  if (!XMAJOR) {
    double2* stw = sm + nl * (M + 1);
    for (int q = threadIdx.x; q < M; q += blockDim.x) stw[q] = tw[q];
    twl = stw;
  }
  // every load of a batch in flight before the LDS stores
  for (int q0 = threadIdx.x; q0 < nl * N; q0 += blockDim.x * FFT_LB_FWD) {
    double v[FFT_LB_FWD];
#pragma unroll
    for (int u = 0; u < FFT_LB_FWD; ++u) {
      const int q = q0 + u * blockDim.x;
      const int l = XMAJOR ? q >> logn : q & (nl - 1), e = XMAJOR ? q & (N - 1) : q >> lognl;
      v[u] = 0.0;
      if (q < nl * N) {
        if constexpr (FILL) {
          const long long c = (long long)(e + 1) + fs.sy * (gp * nl + l + 1) + fs.sz * (kp + 1);
          v[u] = ((fs.w[c] - fs.w[c - fs.sz]) * fs.cz + (fs.v[c] - fs.v[c - fs.sy]) * fs.cy +
                  (fs.u[c] - fs.u[c - 1]) * fs.cx);
        } else {
          v[u] = src[bs + l * ls_src + e * es_src];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < FFT_LB_FWD; ++u) {
      const int q = q0 + u * blockDim.x;
      const int l = XMAJOR ? q >> logn : q & (nl - 1), e = XMAJOR ? q & (N - 1) : q >> lognl;
      if (q < nl * N) smd[2 * (l * (M + 1) + brev(e >> 1, logm)) + (e & 1)] = v[u];
    }
  }
  fft_stages<false>(sm, nl, M, logm, logn, twl);
  // X_k = E_k + W^k O_k, E_k = (Z_k + conj Z_M-k) / 2, O_k = (Z_k - conj Z_M-k) / 2i
  for (int q = threadIdx.x; q < nl * M; q += blockDim.x) {
    const int l = XMAJOR ? q >> logm : q & (nl - 1), k = XMAJOR ? q & (M - 1) : q >> lognl;
    const double2* z = sm + l * (M + 1);
    double* out = dst + bd + l * ls_dst;
    if (k == 0) {
      const double2 z0 = z[0];
      out[0] = z0.x + z0.y;
      out[es_dst] = z0.x - z0.y;
    } else {
      const double2 a = z[k], b = z[M - k];
      const double2 E = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
      const double2 O = make_double2(0.5 * (a.y + b.y), -0.5 * (a.x - b.x));
      const double2 X = cmul(twl[k], O);
      out[(2 * k) * es_dst] = E.x + X.x;
      out[(2 * k + 1) * es_dst] = E.y + X.y;
    }
  }
}

// Inverse: packed half spectra -> real lines (scaled by 1/M here, 1/2 in the split).
template <bool XMAJOR>
__global__ void __launch_bounds__(FFT_THREADS)
    k_rfft_inv(const double* src, double* dst, int N, int logm,
               int nl, int groups_per_plane, long long plane_src, long long plane_dst,
               long long gstride_src, long long gstride_dst, long long es_src, long long ls_src,
               long long es_dst, long long ls_dst, long long off_dst,
               const double2* __restrict__ tw) {
  extern __shared__ double2 sm[];
  const int M = N >> 1, logn = logm + 1, lognl = 31 - __clz(nl);
  const int blk = xcd_logical_block<XMAJOR>();
  const int kp = blk / groups_per_plane, gp = blk - kp * groups_per_plane;
  const long long bs = kp * plane_src + gp * gstride_src;
  const long long bd = off_dst + kp * plane_dst + gp * gstride_dst;
  // twiddles from an LDS copy after the lines (as in k_rfft_fwd)
  double2* stw = sm + nl * (M + 1);
  for (int q = threadIdx.x; q < M; q += blockDim.x) stw[q] = tw[q];
  __syncthreads();
  // Z_k = E_k + i O_k, E_k = (X_k + conj X_M-k) / 2, O_k = (X_k - conj X_M-k) conj(W^k) / 2
  for (int q0 = threadIdx.x; q0 < nl * M; q0 += blockDim.x * FFT_LB_INV) {
    double2 va[FFT_LB_INV], vb[FFT_LB_INV];
#pragma unroll
    for (int u = 0; u < FFT_LB_INV; ++u) {
      const int q = q0 + u * blockDim.x;
      const int l = XMAJOR ? q >> logm : q & (nl - 1), k = XMAJOR ? q & (M - 1) : q >> lognl;
      const double* in = src + bs + l * ls_src;
      va[u] = make_double2(0.0, 0.0), vb[u] = va[u];
      if (q < nl * M) {
        if (k == 0) {
          va[u] = make_double2(in[0], in[es_src]);
        } else {
          va[u] = make_double2(in[(2 * k) * es_src], in[(2 * k + 1) * es_src]);
          vb[u] = make_double2(in[(2 * (M - k)) * es_src], in[(2 * (M - k) + 1) * es_src]);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < FFT_LB_INV; ++u) {
      const int q = q0 + u * blockDim.x;
      if (q >= nl * M) continue;
      const int l = XMAJOR ? q >> logm : q & (nl - 1), k = XMAJOR ? q & (M - 1) : q >> lognl;
      double2 Z;
      if (k == 0) {
        const double x0 = va[u].x, xm = va[u].y;
        Z = make_double2(0.5 * (x0 + xm), 0.5 * (x0 - xm));
      } else {
        const double2 a = va[u], b = vb[u];
        const double2 E = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
        const double2 D = make_double2(0.5 * (a.x - b.x), 0.5 * (a.y + b.y));
        const double2 wc = make_double2(stw[k].x, -stw[k].y);
        const double2 O = cmul(D, wc);
        Z = make_double2(E.x - O.y, E.y + O.x);
      }
      sm[l * (M + 1) + brev(k, logm)] = Z;
    }
  }
  fft_stages<true>(sm, nl, M, logm, logn, stw);
  const double sc = 1.0 / (double)M;
  const double* smd = reinterpret_cast<const double*>(sm);
  for (int q = threadIdx.x; q < nl * N; q += blockDim.x) {
    const int l = XMAJOR ? q >> logn : q & (nl - 1), e = XMAJOR ? q & (N - 1) : q >> lognl;
    dst[bd + l * ls_dst + e * es_dst] = smd[2 * (l * (M + 1) + (e >> 1)) + (e & 1)] * sc;
  }
}

// ---- Gaussian elimination in k (solver.f90:74-129) ------------------------------------------
// Wavenumber of packed position p of a length-n line: 0, n/2, or p/2.
__host__ __device__ __forceinline__ int wavenumber(int p, int n) {
  return p == 0 ? 0 : (p == 1 ? n / 2 : p >> 1);
}

// Reciprocal pivots per (wavenumber pair, k), built once: ztab[k][ky][kx] = z of
// solver.f90:80,89 (k < kmax), and at k = kmax the raw pivot bb - a d (:103-107).
__global__ void __launch_bounds__(256) k_ztab(int nkx, int nky, int ktot, int ne,
                                              const double* __restrict__ xrt,
                                              const double* __restrict__ yrt,
                                              const double* __restrict__ a,
                                              const double* __restrict__ b,
                                              const double* __restrict__ c,
                                              double* __restrict__ ztab) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nkx * nky) return;
  const int kx = (int)(t % nkx), ky = (int)(t / nkx);
  const double xy = xrt[kx] + yrt[ky];
  const long long st = (long long)nkx * nky;
  double d = 0.0;
  for (int k = 0; k < ktot - 1; ++k) {
    const double bb = b[k] + xy;
    const double z = 1.0 / (bb - a[k] * d);
    ztab[k * st + t] = z;
    d = c[k] * z;
  }
  const double bb = (ne == 3) ? b[ktot - 1] : b[ktot - 1] + xy;
  ztab[(long long)(ktot - 1) * st + t] = bb - a[ktot - 1] * d;
}

// One lane per (i, j) mode of the compact spectral array, the forward sweep's values stored
// back into the array and read again by the back substitution (the path above 257 planes;
// k_tri_seg below takes every smaller grid in one pass).
__global__ void __launch_bounds__(TRI_MODES)
    k_tri(double* __restrict__ p, int itot, int jtot, int ktot, const double* __restrict__ a,
          const double* __restrict__ c, const double* __restrict__ ztab) {
  const long long plane = (long long)itot * jtot;
  const long long m = (long long)blockIdx.x * TRI_MODES + threadIdx.x;
  if (m >= plane) return;
  const int i = (int)(m % itot), j = (int)(m / itot);
  const int nkx = itot / 2 + 1, nky = jtot / 2 + 1;
  const long long zb = (long long)wavenumber(j, jtot) * nkx + wavenumber(i, itot);
  const long long zs = (long long)nkx * nky;
  double prev = 0.0;
  for (int k = 0; k < ktot - 1; ++k) {
    const double z = ztab[k * zs + zb];
    prev = (p[k * plane + m] - a[k] * prev) * z;
    p[k * plane + m] = prev;
  }
  const double zl = ztab[(long long)(ktot - 1) * zs + zb];
  double nxt = zl != 0.0 ? (p[(long long)(ktot - 1) * plane + m] - a[ktot - 1] * prev) / zl : 0.0;
  p[(long long)(ktot - 1) * plane + m] = nxt;
  for (int k = ktot - 2; k >= 0; --k) {
    const double d = c[k] * ztab[k * zs + zb];
    nxt = p[k * plane + m] - d * nxt;
    p[k * plane + m] = nxt;
  }
}

// One pass over the array for the same elimination.  A workgroup takes TRI_MODES modes (the
// lanes) and splits rows 0..ktot-2 into segments of TRS_LSEG rows, one wavefront each, held in
// registers with their reciprocal pivots.  Each recurrence of the sweep is affine in the value
// it carries in from the previous segment, so
//   forward  y_k = (r_k - a_k y_k-1) z_k      over a segment: y_end = y0_end + G y_in,
//   backward x_k = y_k - d_k x_k+1             over a segment: x_start = x0_start + H x_in,
// where (y0, G) and (x0, H) come from a first pass started at 0 (G, H the products of the
// coefficients -a z and -d).  The carried values are composed through LDS, then every segment
// runs the sweep again from its true incoming value (the same expressions as k_tri inside a
// segment; only the carried values are re-associated).  The last wavefront also solves the
// last row (solver.f90:103-107, with its zero-pivot guard).  Reads the array and the pivots
// once and writes the solution once, against twice each for k_tri.
constexpr int TRS_LSEG = 16;                   // rows per segment (registers per lane: 2 x 16)
constexpr int TRS_MAXSEG = 16;                 // segments (wavefronts) per workgroup
__global__ void __launch_bounds__(TRI_MODES * TRS_MAXSEG)
    k_tri_seg(double* __restrict__ p, int itot, int jtot, int ktot, const double* __restrict__ a,
              const double* __restrict__ c, const double* __restrict__ ztab) {
  __shared__ double fy0[TRS_MAXSEG][TRI_MODES], fg[TRS_MAXSEG][TRI_MODES];
  __shared__ double bx0[TRS_MAXSEG][TRI_MODES], bh[TRS_MAXSEG][TRI_MODES];
  __shared__ double xlast[TRI_MODES];
  const int lane = threadIdx.x & (TRI_MODES - 1);
  // the segment is uniform over the wavefront: its rows and coefficients live in scalars
  const int s = __builtin_amdgcn_readfirstlane(threadIdx.x / TRI_MODES);
  const int nseg = blockDim.x / TRI_MODES;
  const long long plane = (long long)itot * jtot;
  const long long m = (long long)blockIdx.x * TRI_MODES + lane;
  const bool live = m < plane;
  const int nk = ktot - 1;  // rows of the two recurrences
  const int k0 = min(s * TRS_LSEG, nk), n = min(TRS_LSEG, nk - k0);
  long long zb = 0;
  if (live) {
    const int i = (int)(m % itot), j = (int)(m / itot);
    zb = (long long)wavenumber(j, jtot) * (itot / 2 + 1) + wavenumber(i, itot);
  }
  const long long zs = (long long)(itot / 2 + 1) * (jtot / 2 + 1);
  // rows t >= n of a short segment are padding: they leave every carried value unchanged
  // (selects, so that the arrays stay in registers)
  // (every load is in bounds: lanes past the last mode read mode 0, then discard it)
  const long long mm = live ? m : 0;
  double r[TRS_LSEG], z[TRS_LSEG];
#pragma unroll
  for (int t = 0; t < TRS_LSEG; ++t) {
    const int k = k0 + min(t, max(n - 1, 0));
    const double rv = p[k * plane + mm], zv = ztab[k * zs + zb];
    r[t] = live && n > 0 ? rv : 0.0;
    z[t] = live && n > 0 ? zv : 1.0;
  }
  // forward, first pass: the segment's map y_in -> y_end = y0 + G y_in
  {
    double y = 0.0, gm = 1.0;
#pragma unroll
    for (int t = 0; t < TRS_LSEG; ++t) {
      const double ak = a[k0 + min(t, max(n - 1, 0))];
      const double yn = (r[t] - ak * y) * z[t];
      const double gn = gm * (-(ak * z[t]));
      y = t < n ? yn : y;
      gm = t < n ? gn : gm;
    }
    fy0[s][lane] = y, fg[s][lane] = gm;
  }
  __syncthreads();
  double yin = 0.0;
  for (int q = 0; q < s; ++q) yin = fy0[q][lane] + fg[q][lane] * yin;
  // forward, second pass from the true incoming value (r becomes y)
  double prev = yin;
#pragma unroll
  for (int t = 0; t < TRS_LSEG; ++t) {
    const double yn = (r[t] - a[k0 + min(t, max(n - 1, 0))] * prev) * z[t];
    prev = t < n ? yn : prev;
    r[t] = prev;
  }
  // the last row, by the last wavefront: y_{ktot-2} is its carried-out value
  if (s == nseg - 1) {
    double nxt = 0.0;
    if (live) {
      const double zl = ztab[(long long)(ktot - 1) * zs + zb];
      nxt = zl != 0.0 ? (p[(long long)(ktot - 1) * plane + m] - a[ktot - 1] * prev) / zl : 0.0;
      p[(long long)(ktot - 1) * plane + m] = nxt;
    }
    xlast[lane] = nxt;
  }
  // backward, first pass: the segment's map x_in -> x_start = x0 + H x_in, d_k = c_k z_k
  {
    double x = 0.0, hm = 1.0;
#pragma unroll
    for (int t = TRS_LSEG - 1; t >= 0; --t) {
      const double d = c[k0 + min(t, max(n - 1, 0))] * z[t];
      const double xn = r[t] - d * x;
      const double hn = hm * (-d);
      x = t < n ? xn : x;
      hm = t < n ? hn : hm;
    }
    bx0[s][lane] = x, bh[s][lane] = hm;
  }
  __syncthreads();
  double xin = xlast[lane];
  for (int q = nseg - 1; q > s; --q) xin = bx0[q][lane] + bh[q][lane] * xin;
  // backward, second pass: the solution
  double nxt = xin;
#pragma unroll
  for (int t = TRS_LSEG - 1; t >= 0; --t) {
    const double d = c[k0 + min(t, max(n - 1, 0))] * z[t];
    const double xn = r[t] - d * nxt;
    nxt = t < n ? xn : nxt;
    r[t] = nxt;
  }
  if (live) {
#pragma unroll
    for (int t = 0; t < TRS_LSEG; ++t)
      if (t < n) p[(k0 + t) * plane + m] = r[t];
  }
}

// ---- stencils --------------------------------------------------------------------------------
// The advection terms read each velocity field at offsets of at most one cell from (i, j, k),
// through an accessor A(di, dj, dk) = field(i + di, j + dj, k + dk) (planes staged in LDS by
// the k-marching kernel), in the reference's expressions and order.

// mom.f90:9-50 (momxad)
template <class A>
__device__ __forceinline__ double momxad(const FGrid& g, const A& Uo, const A& Vo, const A& Wo,
                                         const A& Uoo, const A& Voo, const A& Woo) {
  const double uuip = 0.25 * (Uo(1, 0, 0) + Uo(0, 0, 0)) * (Uo(1, 0, 0) + Uo(0, 0, 0));
  const double uuim = 0.25 * (Uo(-1, 0, 0) + Uo(0, 0, 0)) * (Uo(-1, 0, 0) + Uo(0, 0, 0));
  const double uvjp = 0.25 * (Uo(0, 1, 0) + Uo(0, 0, 0)) * (Vo(1, 0, 0) + Vo(0, 0, 0));
  const double uvjm = 0.25 * (Uo(0, -1, 0) + Uo(0, 0, 0)) * (Vo(1, -1, 0) + Vo(0, -1, 0));
  const double uwkp = 0.25 * (Uo(0, 0, 1) + Uo(0, 0, 0)) * (Wo(1, 0, 0) + Wo(0, 0, 0));
  const double uwkm = 0.25 * (Uo(0, 0, -1) + Uo(0, 0, 0)) * (Wo(1, 0, -1) + Wo(0, 0, -1));
  const double uuipo = 0.25 * (Uoo(1, 0, 0) + Uoo(0, 0, 0)) * (Uoo(1, 0, 0) + Uoo(0, 0, 0));
  const double uuimo = 0.25 * (Uoo(-1, 0, 0) + Uoo(0, 0, 0)) * (Uoo(-1, 0, 0) + Uoo(0, 0, 0));
  const double uvjpo = 0.25 * (Uoo(0, 1, 0) + Uoo(0, 0, 0)) * (Voo(1, 0, 0) + Voo(0, 0, 0));
  const double uvjmo = 0.25 * (Uoo(0, -1, 0) + Uoo(0, 0, 0)) * (Voo(1, -1, 0) + Voo(0, -1, 0));
  const double uwkpo = 0.25 * (Uoo(0, 0, 1) + Uoo(0, 0, 0)) * (Woo(1, 0, 0) + Woo(0, 0, 0));
  const double uwkmo = 0.25 * (Uoo(0, 0, -1) + Uoo(0, 0, 0)) * (Woo(1, 0, -1) + Woo(0, 0, -1));
  return g.dxi * ((1.5 * (-uuip + uuim)) - (.5 * (-uuipo + uuimo))) +
         g.dyi * ((1.5 * (-uvjp + uvjm)) - (.5 * (-uvjpo + uvjmo))) +
         g.dzi * ((1.5 * (-uwkp + uwkm)) - (.5 * (-uwkpo + uwkmo)));
}

// mom.f90:158-199 (momyad)
template <class A>
__device__ __forceinline__ double momyad(const FGrid& g, const A& Uo, const A& Vo, const A& Wo,
                                         const A& Uoo, const A& Voo, const A& Woo) {
  const double uvip = 0.25 * (Uo(0, 0, 0) + Uo(0, 1, 0)) * (Vo(0, 0, 0) + Vo(1, 0, 0));
  const double uvim = 0.25 * (Uo(-1, 0, 0) + Uo(-1, 1, 0)) * (Vo(0, 0, 0) + Vo(-1, 0, 0));
  const double vvjp = 0.25 * (Vo(0, 0, 0) + Vo(0, 1, 0)) * (Vo(0, 0, 0) + Vo(0, 1, 0));
  const double vvjm = 0.25 * (Vo(0, 0, 0) + Vo(0, -1, 0)) * (Vo(0, 0, 0) + Vo(0, -1, 0));
  const double wvkp = 0.25 * (Wo(0, 0, 0) + Wo(0, 1, 0)) * (Vo(0, 0, 1) + Vo(0, 0, 0));
  const double wvkm = 0.25 * (Wo(0, 0, -1) + Wo(0, 1, -1)) * (Vo(0, 0, -1) + Vo(0, 0, 0));
  const double uvipo = 0.25 * (Uoo(0, 0, 0) + Uoo(0, 1, 0)) * (Voo(0, 0, 0) + Voo(1, 0, 0));
  const double uvimo = 0.25 * (Uoo(-1, 0, 0) + Uoo(-1, 1, 0)) * (Voo(0, 0, 0) + Voo(-1, 0, 0));
  const double vvjpo = 0.25 * (Voo(0, 0, 0) + Voo(0, 1, 0)) * (Voo(0, 0, 0) + Voo(0, 1, 0));
  const double vvjmo = 0.25 * (Voo(0, 0, 0) + Voo(0, -1, 0)) * (Voo(0, 0, 0) + Voo(0, -1, 0));
  const double wvkpo = 0.25 * (Woo(0, 0, 0) + Woo(0, 1, 0)) * (Voo(0, 0, 1) + Voo(0, 0, 0));
  const double wvkmo = 0.25 * (Woo(0, 0, -1) + Woo(0, 1, -1)) * (Voo(0, 0, -1) + Voo(0, 0, 0));
  return g.dxi * ((1.5 * (-uvip + uvim)) - (.5 * (-uvipo + uvimo))) +
         g.dyi * ((1.5 * (-vvjp + vvjm)) - (.5 * (-vvjpo + vvjmo))) +
         g.dzi * ((1.5 * (-wvkp + wvkm)) - (.5 * (-wvkpo + wvkmo)));
}

// mom.f90:286-327 (momzad)
template <class A>
__device__ __forceinline__ double momzad(const FGrid& g, const A& Uo, const A& Vo, const A& Wo,
                                         const A& Uoo, const A& Voo, const A& Woo) {
  const double uwip = 0.25 * (Wo(0, 0, 0) + Wo(1, 0, 0)) * (Uo(0, 0, 0) + Uo(0, 0, 1));
  const double uwim = 0.25 * (Wo(0, 0, 0) + Wo(-1, 0, 0)) * (Uo(-1, 0, 0) + Uo(-1, 0, 1));
  const double vwjp = 0.25 * (Wo(0, 0, 0) + Wo(0, 1, 0)) * (Vo(0, 0, 0) + Vo(0, 0, 1));
  const double vwjm = 0.25 * (Wo(0, 0, 0) + Wo(0, -1, 0)) * (Vo(0, -1, 0) + Vo(0, -1, 1));
  const double wwkp = 0.25 * (Wo(0, 0, 0) + Wo(0, 0, 1)) * (Wo(0, 0, 0) + Wo(0, 0, 1));
  const double wwkm = 0.25 * (Wo(0, 0, 0) + Wo(0, 0, -1)) * (Wo(0, 0, 0) + Wo(0, 0, -1));
  const double uwipo = 0.25 * (Woo(0, 0, 0) + Woo(1, 0, 0)) * (Uoo(0, 0, 0) + Uoo(0, 0, 1));
  const double uwimo = 0.25 * (Woo(0, 0, 0) + Woo(-1, 0, 0)) * (Uoo(-1, 0, 0) + Uoo(-1, 0, 1));
  const double vwjpo = 0.25 * (Woo(0, 0, 0) + Woo(0, 1, 0)) * (Voo(0, 0, 0) + Voo(0, 0, 1));
  const double vwjmo = 0.25 * (Woo(0, 0, 0) + Woo(0, -1, 0)) * (Voo(0, -1, 0) + Voo(0, -1, 1));
  const double wwkpo = 0.25 * (Woo(0, 0, 0) + Woo(0, 0, 1)) * (Woo(0, 0, 0) + Woo(0, 0, 1));
  const double wwkmo = 0.25 * (Woo(0, 0, 0) + Woo(0, 0, -1)) * (Woo(0, 0, 0) + Woo(0, 0, -1));
  return g.dxi * ((1.5 * (-uwip + uwim)) - (.5 * (-uwipo + uwimo))) +
         g.dyi * ((1.5 * (-vwjp + vwjm)) - (.5 * (-vwjpo + vwjmo))) +
         g.dzi * ((1.5 * (-wwkp + wwkm)) - (.5 * (-wwkpo + wwkmo)));
}

// field(i + di, j + dj, k + dk) from the k-1, k, k+1 planes staged in LDS (row pitch w)
struct LAcc {
  const double* pl[3];  // &plane(i, j) of planes k-1, k, k+1
  int w;
  __device__ __forceinline__ double operator()(int di, int dj, int dk) const {
    return pl[dk + 1][di + w * dj];
  }
};

// The three right-hand sides of fftuvw.f90:30-105 at one interior cell, scaled by Rep
// (dwdt(:,:,kmax) = 0): -(dp/dx) + f, minus the advection terms, minus uo/dt.
template <class A>
__device__ __forceinline__ void mom_cell(const FGrid& g, double rep, double dt, const A& U,
                                         const A& V, const A& W, const A& UU, const A& VV,
                                         const A& WW, const double* __restrict__ p,
                                         long long q, double fxq, double fyq, double fzq, int k,
                                         double& du, double& dv, double& dw) {
  const double durk = (-g.dxi * ((p[q + 1] - p[q]))) + fxq;
  const double dvrk = (-g.dyi * (p[q + g.sy] - p[q])) + fyq;
  const double dwrk = (-g.dzi * (p[q + g.sz] - p[q])) + fzq;
  du = -durk, dv = -dvrk, dw = -dwrk;
  du = du - momxad(g, U, V, W, UU, VV, WW);
  dv = dv - momyad(g, U, V, W, UU, VV, WW);
  dw = dw - momzad(g, U, V, W, UU, VV, WW);
  du = du - (U(0, 0, 0) / dt);
  dv = dv - (V(0, 0, 0) / dt);
  dw = dw - (W(0, 0, 0) / dt);
  du = du * rep;
  dv = dv * rep;
  dw = (k == g.ktot) ? 0.0 : dw * rep;
}

// fftuvw.f90:30-105 per interior cell, k-marching, into the compact arrays: a workgroup owns
// an MR_TI x MR_TJ tile of (i, j) columns and MR_KB consecutive planes.  The six velocity fields' planes k-1, k, k+1 of the tile plus a
// one-cell apron sit in an LDS ring of three; each step prefetches plane k+2 into registers
// while it computes plane k, then stores it over plane k-1.  Every velocity value is read
// from HBM once per tile (plus the apron and the two extra planes); p and the forces are read
// directly.
constexpr int MR_TI = 32, MR_TJ = 8, MR_KB = 32;
constexpr int MR_W = MR_TI + 2, MR_H = MR_TJ + 2, MR_PL = MR_W * MR_H;
constexpr int MR_T = MR_TI * MR_TJ;
constexpr int MR_LD = (MR_PL + MR_T - 1) / MR_T;  // apron-plane cells per thread

__global__ void __launch_bounds__(MR_T)
    k_mom_rhs_march(FGrid g, double rep, double dt, const double* __restrict__ uo,
                    const double* __restrict__ vo, const double* __restrict__ wo,
                    const double* __restrict__ uoo, const double* __restrict__ voo,
                    const double* __restrict__ woo, const double* __restrict__ p,
                    const double* __restrict__ fx, const double* __restrict__ fy,
                    const double* __restrict__ fz, double* __restrict__ cu,
                    double* __restrict__ cv, double* __restrict__ cw) {
  __shared__ double sp[6][3][MR_PL];
  const double* src[6] = {uo, vo, wo, uoo, voo, woo};
  const int tx = threadIdx.x % MR_TI, ty = threadIdx.x / MR_TI;
  const int i0 = blockIdx.x * MR_TI, j0 = blockIdx.y * MR_TJ;  // apron origin (halo index)
  const int kb0 = blockIdx.z * MR_KB + 1;
  const int kb1 = min(kb0 + MR_KB - 1, g.ktot);
  // the apron-plane cells this thread loads: c = tid + r MR_T
  long long off[MR_LD];
  bool ok[MR_LD];
#pragma unroll
  for (int r = 0; r < MR_LD; ++r) {
    const int c = threadIdx.x + r * MR_T;
    const int ii = c % MR_W, jj = c / MR_W;
    ok[r] = c < MR_PL && i0 + ii <= g.itot + 1 && j0 + jj <= g.jtot + 1;
    off[r] = ok[r] ? (long long)(i0 + ii) + g.sy * (j0 + jj) : 0;
  }
  // planes kb0-1, kb0, kb0+1 into ring slots 0, 1, 2 (plane kk lives in slot (kk-kb0+1) % 3)
  for (int s0 = 0; s0 < 3; ++s0) {
    const long long kz = g.sz * (kb0 - 1 + s0);
#pragma unroll
    for (int f = 0; f < 6; ++f)
#pragma unroll
      for (int r = 0; r < MR_LD; ++r)
        if (threadIdx.x + r * MR_T < MR_PL) sp[f][s0][threadIdx.x + r * MR_T] = ok[r] ? src[f][off[r] + kz] : 0.0;
  }
  __syncthreads();
  const int i = i0 + tx + 1, j = j0 + ty + 1;
  const bool live = i <= g.itot && j <= g.jtot;
  const int me = (tx + 1) + MR_W * (ty + 1);
  for (int k = kb0; k <= kb1; ++k) {
    const int sm = (k - kb0) % 3, s0 = (k - kb0 + 1) % 3, sp1 = (k - kb0 + 2) % 3;
    // prefetch plane k + 2 (the next step's k + 1)
    const bool more = k + 2 <= kb1 + 1;
    double pre[6][MR_LD];
    if (more) {
      const long long kz = g.sz * (k + 2);
#pragma unroll
      for (int f = 0; f < 6; ++f)
#pragma unroll
        for (int r = 0; r < MR_LD; ++r) pre[f][r] = ok[r] ? src[f][off[r] + kz] : 0.0;
    }
    if (live) {
      const long long q = fidx(g, i, j, k);
      const double fxq = fx[q], fyq = fy[q], fzq = fz[q];
      LAcc A[6];
#pragma unroll
      for (int f = 0; f < 6; ++f)
        A[f] = LAcc{{&sp[f][sm][me], &sp[f][s0][me], &sp[f][sp1][me]}, MR_W};
      double du, dv, dw;
      mom_cell(g, rep, dt, A[0], A[1], A[2], A[3], A[4], A[5], p, q, fxq, fyq, fzq, k, du, dv,
               dw);
      const long long t = (long long)(i - 1) + (long long)g.itot * ((j - 1) + (long long)g.jtot * (k - 1));
      cu[t] = du;
      cv[t] = dv;
      cw[t] = dw;
    }
    __syncthreads();  // plane k-1 is no longer read
    if (more) {
#pragma unroll
      for (int f = 0; f < 6; ++f)
#pragma unroll
        for (int r = 0; r < MR_LD; ++r)
          if (threadIdx.x + r * MR_T < MR_PL) sp[f][sm][threadIdx.x + r * MR_T] = pre[f][r];
    }
    __syncthreads();
  }
}

// fillps.f90:8-47 into the compact array
__global__ void __launch_bounds__(256) k_fillps(FGrid g, double dt, const double* __restrict__ u,
                                                const double* __restrict__ v,
                                                const double* __restrict__ w,
                                                double* __restrict__ cp) {
  const long long n = (long long)g.itot * g.jtot * g.ktot;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % g.itot) + 1;
  const long long r = t / g.itot;
  const int j = (int)(r % g.jtot) + 1, k = (int)(r / g.jtot) + 1;
  const double dti = 1.0 / dt;
  const double dtidxi = dti * g.dxi, dtidyi = dti * g.dyi, dtidzi = dti * g.dzi;
  const long long q = fidx(g, i, j, k);
  cp[t] = ((w[q] - w[q - g.sz]) * dtidzi + (v[q] - v[q - g.sy]) * dtidyi +
           (u[q] - u[q - 1]) * dtidxi);
}

// a compact array into the interior of a field
__global__ void __launch_bounds__(256) k_to_field(FGrid g, const double* __restrict__ c,
                                                  double* __restrict__ f) {
  const long long n = (long long)g.itot * g.jtot * g.ktot;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % g.itot) + 1;
  const long long r = t / g.itot;
  f[fidx(g, i, (int)(r % g.jtot) + 1, (int)(r / g.jtot) + 1)] = c[t];
}

// correc.f90:11-38
__global__ void __launch_bounds__(256) k_correc(FGrid g, double dt, double* __restrict__ u,
                                                double* __restrict__ v, double* __restrict__ w,
                                                const double* __restrict__ p) {
  const long long n = (long long)g.itot * g.jtot * g.ktot;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % g.itot) + 1;
  const long long r = t / g.itot;
  const int j = (int)(r % g.jtot) + 1, k = (int)(r / g.jtot) + 1;
  const double factori = dt * g.dxi, factorj = dt * g.dyi, factork = dt * g.dzi;
  const long long q = fidx(g, i, j, k);
  u[q] = u[q] - factori * (p[q + 1] - p[q]);
  v[q] = v[q] - factorj * (p[q + g.sy] - p[q]);
  w[q] = w[q] - factork * (p[q + g.sz] - p[q]);
}

// correc (correc.f90:11-38) and pnew += p (main.f90:406-408) in one pass over the cells: the
// two touch disjoint outputs and both read p, which is then read once
__global__ void __launch_bounds__(256) k_correc_addp(FGrid g, double dt, double* __restrict__ u,
                                                     double* __restrict__ v,
                                                     double* __restrict__ w,
                                                     const double* __restrict__ p,
                                                     double* __restrict__ pn) {
  const long long n = (long long)g.itot * g.jtot * g.ktot;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % g.itot) + 1;
  const long long r = t / g.itot;
  const int j = (int)(r % g.jtot) + 1, k = (int)(r / g.jtot) + 1;
  const double factori = dt * g.dxi, factorj = dt * g.dyi, factork = dt * g.dzi;
  const long long q = fidx(g, i, j, k);
  const double pq = p[q];
  u[q] = u[q] - factori * (p[q + 1] - pq);
  v[q] = v[q] - factorj * (p[q + g.sy] - pq);
  w[q] = w[q] - factork * (p[q + g.sz] - pq);
  pn[q] = pn[q] + pq;
}

// Final value of a wall ghost / top face after the k step of bounduvw (bound.f90:38-56) or
// boundp (:93-98), at interior column (i, j).  kind 0: u or v (ghost = -interior), 1: w
// (0 on both wall faces, k1 copies kmax-1), 2: p (ghost = interior).
__device__ __forceinline__ double kstep(const FGrid& g, const double* f, int kind, int i, int j,
                                        int k) {
  const int kmax = g.ktot, k1 = g.ktot + 1;
  if (kind == 0) {
    if (k == 0) return -f[fidx(g, i, j, 1)];
    if (k == k1) return -f[fidx(g, i, j, kmax)];
  } else if (kind == 1) {
    if (k == 0 || k == kmax) return 0.0;
    if (k == k1) return f[fidx(g, i, j, kmax - 1)];
  } else {
    if (k == 0) return f[fidx(g, i, j, 1)];
    if (k == k1) return f[fidx(g, i, j, kmax)];
  }
  return f[fidx(g, i, j, k)];
}

// bounduvw / boundp in one launch: the k step, then updthalos in x, then in y
// (bound.f90:61-76, 100-104; bound.f90:141-190).  Every halo or ghost cell takes its final
// value straight from the interior column the sequence copies it from, and those interior
// values are not written here, so every cell is independent.  Threads [0, nfields * itot *
// jtot) take an interior column (its ghost planes k = 0 and ktot+1, and k = ktot for w); the
// rest one cell of a halo column: x halos i = 0, itot+1 for j = 0..jtot+1, then y halos
// j = 0, jtot+1 for i = 1..itot, every k (halo index fastest).
__global__ void __launch_bounds__(256) k_bound(FGrid g, double* __restrict__ f0,
                                               double* __restrict__ f1, double* __restrict__ f2,
                                               int nfields, int kind0) {
  const long long nin = (long long)g.itot * g.jtot;
  const long long nxh = 2LL * (g.jtot + 2), nh = nxh + 2LL * g.itot, K = g.ktot + 2;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nin * nfields) {
    const int fi = (int)(t / nin);
    const long long r = t - fi * nin;
    const int i = 1 + (int)(r % g.itot), j = 1 + (int)(r / g.itot);
    double* f = fi == 0 ? f0 : (fi == 1 ? f1 : f2);
    const int kind = kind0 == 2 ? 2 : (fi == 2 ? 1 : 0);
    f[fidx(g, i, j, 0)] = kstep(g, f, kind, i, j, 0);
    f[fidx(g, i, j, g.ktot + 1)] = kstep(g, f, kind, i, j, g.ktot + 1);
    if (kind == 1) f[fidx(g, i, j, g.ktot)] = 0.0;
    return;
  }
  const long long e = t - nin * nfields;
  if (e >= nh * K * nfields) return;
  const int fi = (int)(e / (nh * K));
  const long long r = e - fi * nh * K;
  const int k = (int)(r / nh);
  const long long h = r - k * nh;
  int i, j;
  if (h < nxh) {
    i = h < g.jtot + 2 ? 0 : g.itot + 1;
    j = (int)(h % (g.jtot + 2));
  } else {
    const long long hy = h - nxh;
    j = hy < g.itot ? 0 : g.jtot + 1;
    i = 1 + (int)(hy % g.itot);
  }
  double* f = fi == 0 ? f0 : (fi == 1 ? f1 : f2);
  const int kind = kind0 == 2 ? 2 : (fi == 2 ? 1 : 0);
  const int si = i == 0 ? g.itot : (i == g.itot + 1 ? 1 : i);
  const int sj = j == 0 ? g.jtot : (j == g.jtot + 1 ? 1 : j);
  f[fidx(g, i, j, k)] = kstep(g, f, kind, si, sj, k);
}

// indicator.f90:51-67: the central divergence of G = (gx, gy, gz), ((g(+1) - g(-1)) .5) d?i
// summed x + y + z, for every interior cell, into the compact array cp (the right-hand side
// of the indicator's Poisson solve) or, with f != nullptr, into the interior of the field f.
__global__ void __launch_bounds__(256) k_ind_div(FGrid g, const double* __restrict__ gx,
                                                 const double* __restrict__ gy,
                                                 const double* __restrict__ gz,
                                                 double* __restrict__ cp,
                                                 double* __restrict__ f) {
  const long long n = (long long)g.itot * g.jtot * g.ktot;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % g.itot) + 1;
  const long long r = t / g.itot;
  const int j = (int)(r % g.jtot) + 1, k = (int)(r / g.jtot) + 1;
  const long long q = fidx(g, i, j, k);
  const double dgx = (gx[q + 1] - gx[q - 1]) * .5 * g.dxi;
  const double dgy = (gy[q + g.sy] - gy[q - g.sy]) * .5 * g.dyi;
  const double dgz = (gz[q + g.sz] - gz[q - g.sz]) * .5 * g.dzi;
  const double dg = dgx + dgy + dgz;
  if (f)
    f[q] = dg;
  else
    cp[t] = dg;
}

// indicator.f90:93-103: indic(i,j,k) = indi(i,j,k + ktot/4) for k <= kmax - ktot/4, else 0
// (the solution moved back down by the planes the capsules were shifted up)
__global__ void __launch_bounds__(256) k_ind_shift(FGrid g, const double* __restrict__ indi,
                                                   double* __restrict__ indic) {
  const long long n = (long long)g.itot * g.jtot * g.ktot;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % g.itot) + 1;
  const long long r = t / g.itot;
  const int j = (int)(r % g.jtot) + 1, k = (int)(r / g.jtot) + 1;
  const int kq = g.ktot / 4;
  indic[fidx(g, i, j, k)] = k <= g.ktot - kq ? indi[fidx(g, i, j, k + kq)] : 0.0;
}

// indicator.f90:105-112: < .01 -> 0, >= .9 -> 1 (idempotent)
__device__ __forceinline__ double ind_threshold(double v) {
  if (v < .01) v = 0.;
  if (v >= .9) v = 1.;
  return v;
}

// indicator.f90:105-168 after boundind(indic), fused per interior cell: the thresholds, the
// viscosity visc0 (1 + (lambda - 1) indic) (:137-143) and the centrifugal buoyancy
//   forcey += .5 (indic(j) + indic(j+1)) sin(thetainc) tcoef          (:164)
//   forcez += .5 (indic(k) + indic(k+1)) (-cos(thetainc)) tcoef       (:165)
// The reference thresholds the whole interior before the force loop, so an interior
// neighbour is read thresholded and a halo neighbour (j+1 = j1, k+1 = k1) as boundind left
// it.  A neighbour is read here and thresholded on the spot: the threshold is idempotent, so
// reading it before or after its own thread rewrote it gives the same value.
__global__ void __launch_bounds__(256) k_ind_finish(FGrid g, double* __restrict__ indic,
                                                    double* __restrict__ visc,
                                                    double* __restrict__ forcey,
                                                    double* __restrict__ forcez, double lambda,
                                                    double visc0, double sin_t, double mcos_t,
                                                    double tcoef) {
  const long long n = (long long)g.itot * g.jtot * g.ktot;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % g.itot) + 1;
  const long long r = t / g.itot;
  const int j = (int)(r % g.jtot) + 1, k = (int)(r / g.jtot) + 1;
  const long long q = fidx(g, i, j, k);
  const double v = ind_threshold(indic[q]);
  double vy = indic[q + g.sy], vz = indic[q + g.sz];
  if (j < g.jtot) vy = ind_threshold(vy);
  if (k < g.ktot) vz = ind_threshold(vz);
  indic[q] = v;
  visc[q] = visc0 * (1. + ((lambda - 1.) * v));
  forcey[q] = forcey[q] + (.5 * (v + vy) * sin_t * tcoef);
  forcez[q] = forcez[q] + (.5 * (v + vz) * mcos_t * tcoef);
}

// chkdiv.f90:7-61: per-block sums and maxima of the discrete divergence
__global__ void __launch_bounds__(256) k_div(FGrid g, const double* __restrict__ u,
                                             const double* __restrict__ v,
                                             const double* __restrict__ w,
                                             double* __restrict__ part) {
  __shared__ double ss[256], sm[256];
  const long long n = (long long)g.itot * g.jtot * g.ktot;
  double s = 0.0, m = 0.0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % g.itot) + 1;
    const long long r = t / g.itot;
    const int j = (int)(r % g.jtot) + 1, k = (int)(r / g.jtot) + 1;
    const long long q = fidx(g, i, j, k);
    const double div = (w[q] - w[q - g.sz]) * g.dzi + (v[q] - v[q - g.sy]) * g.dyi +
                       (u[q] - u[q - 1]) * g.dxi;
    s = s + div;
    m = fmax(m, fabs(div));
  }
  ss[threadIdx.x] = s, sm[threadIdx.x] = m;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      ss[threadIdx.x] = ss[threadIdx.x] + ss[threadIdx.x + h];
      sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[2 * blockIdx.x] = ss[0], part[2 * blockIdx.x + 1] = sm[0];
}

}  // namespace

struct fld_ctx {
  ibx_grid grid{};
  FGrid g{};
  double rep = 0, dt = 0;
  int device = 0;
  hipStream_t own = nullptr, stream = nullptr;
  std::string err;
  long long ncomp = 0;                  // itot * jtot * ktot
  double* work[3] = {nullptr, nullptr, nullptr};  // compact spectral arrays
  double* ztab[4] = {nullptr, nullptr, nullptr, nullptr};  // index ne = 1..3
  double* coef[4] = {nullptr, nullptr, nullptr, nullptr};  // a, c per ne (2 * ktot each)
  double2* tw_x = nullptr;
  double2* tw_y = nullptr;
  double* part = nullptr;  // chkdiv partials
  int logmx = 0, logmy = 0;
  size_t lds_yf = 0, lds_yi = 0;
  int ylines_f = Y_LINES_FWD, ylines_i = Y_LINES_INV, xlines = 1;
};

namespace {

int fail(fld_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPF(ctx, call)                                                                      \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return fail((ctx), FLD_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_) +     \
                                        " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
  } while (0)

inline unsigned nblocks(long long n, int b) { return (unsigned)((n + b - 1) / b); }

int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

bool pow2(int n) { return n >= 2 && (n & (n - 1)) == 0; }

// exp(-2 pi i k / n), k < n/2, from long-double sine and cosine
std::vector<double2> twiddles(int n) {
  std::vector<double2> t((size_t)std::max(1, n / 2));
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < n / 2; ++k) {
    const long double a = 2.0L * pi * (long double)k / (long double)n;
    t[(size_t)k] = make_double2((double)cosl(a), (double)(-sinl(a)));
  }
  return t;
}

// initsolver.f90:108-132 by wavenumber: 0, 1 .. n/2 - 1, and n/2
std::vector<double> eigen(int n, double di) {
  std::vector<double> e((size_t)(n / 2 + 1));
  const double pi = std::acos(-1.0);
  e[0] = 0.0;
  for (int m = 1; m < n / 2; ++m) {
    const double s = std::sin((double)(2 * m) * pi / (2.0 * n));
    e[(size_t)m] = -(4.0 * di * di * (s * s));
  }
  e[(size_t)(n / 2)] = -(4.0 * di * di);
  return e;
}

// initsolver.f90:23-99 for e = 1, 2, 3 (main mesh)
void tridiag(const fld_ctx* c, int e, std::vector<double>& a, std::vector<double>& b,
             std::vector<double>& cc) {
  const int kmax = c->grid.ktot;
  const double dzi = c->grid.dzi;
  a.assign((size_t)kmax, 0.0), b.assign((size_t)kmax, 0.0), cc.assign((size_t)kmax, 0.0);
  for (int k = 0; k < kmax; ++k) {
    a[(size_t)k] = dzi * dzi;
    cc[(size_t)k] = dzi * dzi;
    b[(size_t)k] = -(a[(size_t)k] + cc[(size_t)k]);
  }
  if (e == 1) {
    b[0] = b[0] + a[0];
    b[(size_t)kmax - 1] = b[(size_t)kmax - 1] + cc[(size_t)kmax - 1];
    a[0] = 0.0;
    cc[(size_t)kmax - 1] = 0.0;
  } else {
    if (e == 2) {
      b[0] = b[0] - a[0];
      b[(size_t)kmax - 1] = b[(size_t)kmax - 1] - cc[(size_t)kmax - 1];
      a[0] = 0.0;
      cc[(size_t)kmax - 1] = 0.0;
    } else {
      b[(size_t)kmax - 1] = 1.0;
      a[0] = 0.0;
      a[(size_t)kmax - 1] = 0.0;
      cc[(size_t)kmax - 1] = 0.0;
    }
    for (int k = 0; k < kmax; ++k) b[(size_t)k] = b[(size_t)k] - (c->rep / c->dt);
    if (e == 3) b[(size_t)kmax - 1] = 1.0;
  }
}

// x transform of the interior lines of a field (or of a compact array), forward or inverse
int fft_x(fld_ctx* c, bool inv, const double* src, bool src_field, double* dst, bool dst_field) {
  const FGrid& g = c->g;
  const int N = g.itot, nl = c->xlines;
  const int gpp = g.jtot / nl;  // workgroups per k plane
  const long long cp = (long long)g.itot * g.jtot;
  const long long row_s = src_field ? g.sy : g.itot, plane_s = src_field ? g.sz : cp;
  const long long row_d = dst_field ? g.sy : g.itot, plane_d = dst_field ? g.sz : cp;
  const long long off = 1 + g.sy + g.sz;  // (1, 1, 1)
  // lines, plus the inverse's LDS copy of the twiddles
  const size_t lds = (size_t)nl * (N / 2 + 1) * sizeof(double2) + (inv ? (size_t)(N / 2) * sizeof(double2) : 0);
  const unsigned nb = (unsigned)(gpp * g.ktot);
  if (!inv)
    k_rfft_fwd<true><<<nb, FFT_THREADS, lds, c->stream>>>(
        src, dst, N, c->logmx, nl, gpp, plane_s, plane_d, nl * row_s, nl * row_d, 1, row_s, 1,
        row_d, src_field ? off : 0, c->tw_x, FillSrc{});
  else
    k_rfft_inv<true><<<nb, FFT_THREADS, lds, c->stream>>>(
        src, dst, N, c->logmx, nl, gpp, plane_s, plane_d, nl * row_s, nl * row_d, 1, row_s, 1,
        row_d, dst_field ? off : 0, c->tw_x);
  HIPF(c, hipGetLastError());
  return FLD_OK;
}

// y transform in place on a compact array
int fft_y(fld_ctx* c, bool inv, double* a) {
  const FGrid& g = c->g;
  const int N = g.jtot, nl = inv ? c->ylines_i : c->ylines_f;
  const int gpp = g.itot / nl;
  const long long cp = (long long)g.itot * g.jtot;
  const unsigned nb = (unsigned)(gpp * g.ktot);
  if (!inv)
    k_rfft_fwd<false><<<nb, FFT_THREADS, c->lds_yf, c->stream>>>(
        a, a, N, c->logmy, nl, gpp, cp, cp, nl, nl, g.itot, 1, g.itot, 1, 0, c->tw_y, FillSrc{});
  else
    k_rfft_inv<false><<<nb, FFT_THREADS, c->lds_yi, c->stream>>>(
        a, a, N, c->logmy, nl, gpp, cp, cp, nl, nl, g.itot, 1, g.itot, 1, 0, c->tw_y);
  HIPF(c, hipGetLastError());
  return FLD_OK;
}

int tri(fld_ctx* c, double* a, int ne) {
  const FGrid& g = c->g;
  const long long modes = (long long)g.itot * g.jtot;
  const double* ac = c->coef[ne];
  const double* cc = c->coef[ne] + g.ktot;
  const int nseg = std::max(1, (g.ktot - 1 + TRS_LSEG - 1) / TRS_LSEG);
  if (nseg <= TRS_MAXSEG)
    k_tri_seg<<<nblocks(modes, TRI_MODES), TRI_MODES * nseg, 0, c->stream>>>(
        a, g.itot, g.jtot, g.ktot, ac, cc, c->ztab[ne]);
  else
    k_tri<<<nblocks(modes, TRI_MODES), TRI_MODES, 0, c->stream>>>(a, g.itot, g.jtot, g.ktot, ac,
                                                                  cc, c->ztab[ne]);
  HIPF(c, hipGetLastError());
  return FLD_OK;
}

// solver1d after the forward x transform: y transform, eliminations, inverses into the
// interior of `out`
int solve_after_x(fld_ctx* c, double* a, int ne, double* out) {
  int rc;
  if ((rc = fft_y(c, false, a)) || (rc = tri(c, a, ne)) || (rc = fft_y(c, true, a)) ||
      (rc = fft_x(c, true, a, false, out, true)))
    return rc;
  return FLD_OK;
}

// solver1d on a compact right-hand side, result into the interior of `out`
int solve_compact(fld_ctx* c, double* a, int ne, double* out) {
  int rc;
  if ((rc = fft_x(c, false, a, false, a, false))) return rc;
  return solve_after_x(c, a, ne, out);
}

// fillps (the divergence of u, v, w over dt) already transformed in x, into the compact
// array dst: k_fillps and the pressure solve's forward x transform in one pass
int fft_x_fill(fld_ctx* c, const double* u, const double* v, const double* w, double* dst) {
  const FGrid& g = c->g;
  const int N = g.itot, nl = c->xlines;
  const int gpp = g.jtot / nl;
  const long long cp = (long long)g.itot * g.jtot, row = g.itot;
  const size_t lds = (size_t)nl * (N / 2 + 1) * sizeof(double2);
  const unsigned nb = (unsigned)(gpp * g.ktot);
  const double dti = 1.0 / c->dt;
  const FillSrc fs{u, v, w, dti * g.dxi, dti * g.dyi, dti * g.dzi, g.sy, g.sz};
  k_rfft_fwd<true, true><<<nb, FFT_THREADS, lds, c->stream>>>(
      nullptr, dst, N, c->logmx, nl, gpp, cp, cp, nl * row, nl * row, 1, row, 1, row, 0, c->tw_x,
      fs);
  HIPF(c, hipGetLastError());
  return FLD_OK;
}

int bound(fld_ctx* c, double* f0, double* f1, double* f2, int nfields, int kind0) {
  const FGrid& g = c->g;
  const long long n = ((long long)g.itot * g.jtot +
                       (2LL * (g.jtot + 2) + 2LL * g.itot) * (g.ktot + 2)) * nfields;
  k_bound<<<nblocks(n, 256), 256, 0, c->stream>>>(g, f0, f1, f2, nfields, kind0);
  HIPF(c, hipGetLastError());
  return FLD_OK;
}

}  // namespace

extern "C" {

int fld_create(fld_ctx** out, const ibx_grid* gr, double rep, double dt, int device) {
  if (!out || !gr) return FLD_E_ARG;
  *out = nullptr;
  if (!pow2(gr->itot) || !pow2(gr->jtot) || gr->itot < 4 || gr->jtot < 4 || gr->ktot < 2)
    return FLD_E_ARG;
  // one line of each transform must fit in LDS
  if ((size_t)(gr->itot / 2 + 1) * sizeof(double2) > LDS_MAX ||
      (size_t)(gr->jtot / 2 + 1) * sizeof(double2) > LDS_MAX)
    return FLD_E_ARG;
  if (!(rep > 0 && dt > 0 && gr->dxi > 0 && gr->dyi > 0 && gr->dzi > 0)) return FLD_E_ARG;
  fld_ctx* c = new fld_ctx();
  c->grid = *gr;
  c->rep = rep, c->dt = dt, c->device = device;
  FGrid& g = c->g;
  g.itot = gr->itot, g.jtot = gr->jtot, g.ktot = gr->ktot;
  g.sy = gr->itot + 2;
  g.sz = (long long)(gr->itot + 2) * (gr->jtot + 2);
  g.dxi = gr->dxi, g.dyi = gr->dyi, g.dzi = gr->dzi;
  c->ncomp = (long long)g.itot * g.jtot * g.ktot;
  c->logmx = ilog2(g.itot / 2), c->logmy = ilog2(g.jtot / 2);
  c->xlines = std::min(g.jtot, std::max(1, 2048 / g.itot));
  // 8 columns per workgroup both ways (at 1024 x 512 x 256: forward 543 us against 604 with 4
  // and 744 with 2; inverse 602 us against 674 with 16)
  auto lines = [&](int want) {
    int nl = std::min(g.itot, want);
    while (nl > 1 && (size_t)nl * (g.jtot / 2 + 1) * sizeof(double2) > LDS_MAX / 2) nl >>= 1;
    return nl;
  };
  c->ylines_f = lines(Y_LINES_FWD), c->ylines_i = lines(Y_LINES_INV);
  // lines plus the LDS copy of the twiddles
  c->lds_yf = ((size_t)c->ylines_f * (g.jtot / 2 + 1) + g.jtot / 2) * sizeof(double2);
  c->lds_yi = ((size_t)c->ylines_i * (g.jtot / 2 + 1) + g.jtot / 2) * sizeof(double2);
  auto bail = [&](int code, const char* what) {
    fld_destroy(c);
    (void)what;
    return code;
  };
  if (hipSetDevice(device) != hipSuccess) return bail(FLD_E_HIP, "device");
  if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess)
    return bail(FLD_E_HIP, "stream");
  c->stream = c->own;
  for (auto& w : c->work)
    if (hipMalloc(&w, (size_t)c->ncomp * sizeof(double)) != hipSuccess) return bail(FLD_E_HIP, "work");
  // dynamic LDS above the default limit for the y transform and the elimination
  if (hipFuncSetAttribute((const void*)k_rfft_fwd<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)LDS_MAX) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_rfft_inv<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)LDS_MAX) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_rfft_fwd<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)LDS_MAX) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_rfft_fwd<true, true>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_MAX) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_rfft_inv<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)LDS_MAX) != hipSuccess)
    return bail(FLD_E_HIP, "lds");
  // twiddles, eigenvalues, coefficients and the pivot tables of e = 1, 2, 3
  const std::vector<double2> twx = twiddles(g.itot), twy = twiddles(g.jtot);
  const std::vector<double> xrt = eigen(g.itot, g.dxi), yrt = eigen(g.jtot, g.dyi);
  double *dx = nullptr, *dy = nullptr, *dabc = nullptr;
  bool ok = hipMalloc(&c->tw_x, twx.size() * sizeof(double2)) == hipSuccess &&
            hipMalloc(&c->tw_y, twy.size() * sizeof(double2)) == hipSuccess &&
            hipMalloc(&dx, xrt.size() * sizeof(double)) == hipSuccess &&
            hipMalloc(&dy, yrt.size() * sizeof(double)) == hipSuccess &&
            hipMalloc(&dabc, 3 * (size_t)g.ktot * sizeof(double)) == hipSuccess &&
            hipMalloc(&c->part, 2 * 1024 * sizeof(double)) == hipSuccess;
  ok = ok && hipMemcpy(c->tw_x, twx.data(), twx.size() * sizeof(double2), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(c->tw_y, twy.data(), twy.size() * sizeof(double2), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(dx, xrt.data(), xrt.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(dy, yrt.data(), yrt.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
  const int nkx = g.itot / 2 + 1, nky = g.jtot / 2 + 1;
  for (int ne = 1; ok && ne <= 3; ++ne) {
    std::vector<double> a, b, cc;
    tridiag(c, ne, a, b, cc);
    std::vector<double> abc(a);
    abc.insert(abc.end(), b.begin(), b.end());
    abc.insert(abc.end(), cc.begin(), cc.end());
    std::vector<double> ac(a);
    ac.insert(ac.end(), cc.begin(), cc.end());
    ok = hipMalloc(&c->coef[ne], 2 * (size_t)g.ktot * sizeof(double)) == hipSuccess &&
         hipMemcpy(c->coef[ne], ac.data(), ac.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(dabc, abc.data(), abc.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
         hipMalloc(&c->ztab[ne], (size_t)nkx * nky * g.ktot * sizeof(double)) == hipSuccess;
    if (ok) {
      k_ztab<<<nblocks((long long)nkx * nky, 256), 256, 0, c->own>>>(
          nkx, nky, g.ktot, ne, dx, dy, dabc, dabc + g.ktot, dabc + 2 * g.ktot, c->ztab[ne]);
      ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(c->own) == hipSuccess;
    }
  }
  if (dx) (void)hipFree(dx);
  if (dy) (void)hipFree(dy);
  if (dabc) (void)hipFree(dabc);
  if (!ok) return bail(FLD_E_HIP, "tables");
  *out = c;
  return FLD_OK;
}

void fld_destroy(fld_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& w : c->work)
    if (w) (void)hipFree(w);
  for (int ne = 1; ne <= 3; ++ne) {
    if (c->ztab[ne]) (void)hipFree(c->ztab[ne]);
    if (c->coef[ne]) (void)hipFree(c->coef[ne]);
  }
  if (c->tw_x) (void)hipFree(c->tw_x);
  if (c->tw_y) (void)hipFree(c->tw_y);
  if (c->part) (void)hipFree(c->part);
  if (c->own) (void)hipStreamDestroy(c->own);
  delete c;
}

const char* fld_last_error(const fld_ctx* c) { return c ? c->err.c_str() : "null context"; }

int fld_set_stream(fld_ctx* c, void* s) {
  if (!c) return FLD_E_ARG;
  c->stream = s ? (hipStream_t)s : c->own;
  return FLD_OK;
}

int fld_sync(fld_ctx* c) {
  if (!c) return FLD_E_ARG;
  HIPF(c, hipSetDevice(c->device));
  HIPF(c, hipStreamSynchronize(c->stream));
  return FLD_OK;
}

int fld_solver1d_dev(fld_ctx* c, double* p, int ne) {
  if (!c) return FLD_E_ARG;
  if (!p) return fail(c, FLD_E_ARG, "null field");
  if (ne < 1 || ne > 3) return fail(c, FLD_E_ARG, "ne must be 1, 2 or 3");
  HIPF(c, hipSetDevice(c->device));
  int rc = fft_x(c, false, p, true, c->work[0], false);
  if (rc) return rc;
  if ((rc = fft_y(c, false, c->work[0])) || (rc = tri(c, c->work[0], ne)) ||
      (rc = fft_y(c, true, c->work[0])) || (rc = fft_x(c, true, c->work[0], false, p, true)))
    return rc;
  return FLD_OK;
}

int fld_fftuvw_dev(fld_ctx* c, double* unew, double* vnew, double* wnew, const double* uo,
                   const double* vo, const double* wo, const double* uoo, const double* voo,
                   const double* woo, const double* pnew, const double* fx, const double* fy,
                   const double* fz) {
  if (!c) return FLD_E_ARG;
  if (!(unew && vnew && wnew && uo && vo && wo && uoo && voo && woo && pnew && fx && fy && fz))
    return fail(c, FLD_E_ARG, "null field");
  HIPF(c, hipSetDevice(c->device));
  const dim3 mgrid((unsigned)((c->grid.itot + MR_TI - 1) / MR_TI),
                   (unsigned)((c->grid.jtot + MR_TJ - 1) / MR_TJ),
                   (unsigned)((c->grid.ktot + MR_KB - 1) / MR_KB));
  k_mom_rhs_march<<<mgrid, MR_T, 0, c->stream>>>(c->g, c->rep, c->dt, uo, vo, wo, uoo, voo, woo,
                                                 pnew, fx, fy, fz, c->work[0], c->work[1],
                                                 c->work[2]);
  HIPF(c, hipGetLastError());
  int rc;
  if ((rc = solve_compact(c, c->work[2], 3, wnew)) || (rc = solve_compact(c, c->work[0], 2, unew)) ||
      (rc = solve_compact(c, c->work[1], 2, vnew)))
    return rc;
  return FLD_OK;
}

int fld_bounduvw_dev(fld_ctx* c, double* u, double* v, double* w) {
  if (!c) return FLD_E_ARG;
  if (!(u && v && w)) return fail(c, FLD_E_ARG, "null field");
  HIPF(c, hipSetDevice(c->device));
  return bound(c, u, v, w, 3, 0);
}

int fld_boundp_dev(fld_ctx* c, double* p) {
  if (!c) return FLD_E_ARG;
  if (!p) return fail(c, FLD_E_ARG, "null field");
  HIPF(c, hipSetDevice(c->device));
  return bound(c, p, nullptr, nullptr, 1, 2);
}

int fld_fillps_dev(fld_ctx* c, const double* u, const double* v, const double* w, double* p) {
  if (!c) return FLD_E_ARG;
  if (!(u && v && w && p)) return fail(c, FLD_E_ARG, "null field");
  HIPF(c, hipSetDevice(c->device));
  k_fillps<<<nblocks(c->ncomp, 256), 256, 0, c->stream>>>(c->g, c->dt, u, v, w, c->work[0]);
  k_to_field<<<nblocks(c->ncomp, 256), 256, 0, c->stream>>>(c->g, c->work[0], p);
  HIPF(c, hipGetLastError());
  return FLD_OK;
}

int fld_correc_dev(fld_ctx* c, double* u, double* v, double* w, const double* p) {
  if (!c) return FLD_E_ARG;
  if (!(u && v && w && p)) return fail(c, FLD_E_ARG, "null field");
  HIPF(c, hipSetDevice(c->device));
  k_correc<<<nblocks(c->ncomp, 256), 256, 0, c->stream>>>(c->g, c->dt, u, v, w, p);
  HIPF(c, hipGetLastError());
  return FLD_OK;
}

int fld_step_dev(fld_ctx* c, double* unew, double* vnew, double* wnew, double* pnew, double* p,
                 const double* uo, const double* vo, const double* wo, const double* uoo,
                 const double* voo, const double* woo, const double* fx, const double* fy,
                 const double* fz) {
  if (!c) return FLD_E_ARG;
  if (!(pnew && p)) return fail(c, FLD_E_ARG, "null field");
  int rc = fld_fftuvw_dev(c, unew, vnew, wnew, uo, vo, wo, uoo, voo, woo, pnew, fx, fy, fz);
  if (rc) return rc;
  if ((rc = bound(c, unew, vnew, wnew, 3, 0))) return rc;
  if ((rc = fft_x_fill(c, unew, vnew, wnew, c->work[0])) ||
      (rc = solve_after_x(c, c->work[0], 1, p)) || (rc = bound(c, p, nullptr, nullptr, 1, 2)))
    return rc;
  // correc and pnew += p fused (the reference's bounduvw between them reads neither p nor
  // pnew), then bounduvw and boundp(pnew)
  k_correc_addp<<<nblocks(c->ncomp, 256), 256, 0, c->stream>>>(c->g, c->dt, unew, vnew, wnew, p,
                                                               pnew);
  HIPF(c, hipGetLastError());
  if ((rc = bound(c, unew, vnew, wnew, 3, 0))) return rc;
  return bound(c, pnew, nullptr, nullptr, 1, 2);
}

int fld_indicator_div_dev(fld_ctx* c, const double* gx, const double* gy, const double* gz,
                          double* indi) {
  if (!c) return FLD_E_ARG;
  if (!(gx && gy && gz && indi)) return fail(c, FLD_E_ARG, "null field");
  HIPF(c, hipSetDevice(c->device));
  k_ind_div<<<nblocks(c->ncomp, 256), 256, 0, c->stream>>>(c->g, gx, gy, gz, nullptr, indi);
  HIPF(c, hipGetLastError());
  return FLD_OK;
}

static int ind_finish(fld_ctx* c, double* indi, double* indic, double* visc, double* forcey,
                      double* forcez, double lambda, double visc0, double sin_t, double cos_t,
                      double tcoef) {
  int rc;
  if ((rc = bound(c, indi, nullptr, nullptr, 1, 2))) return rc;  // boundind(indi)
  k_ind_shift<<<nblocks(c->ncomp, 256), 256, 0, c->stream>>>(c->g, indi, indic);
  HIPF(c, hipGetLastError());
  if ((rc = bound(c, indic, nullptr, nullptr, 1, 2))) return rc;  // boundind(indic)
  k_ind_finish<<<nblocks(c->ncomp, 256), 256, 0, c->stream>>>(c->g, indic, visc, forcey, forcez,
                                                              lambda, visc0, sin_t, -cos_t,
                                                              tcoef);
  HIPF(c, hipGetLastError());
  return bound(c, visc, nullptr, nullptr, 1, 2);  // boundp(visc)
}

static bool ind_args_ok(double lambda, double visc0, double sin_t, double cos_t, double tcoef) {
  return std::isfinite(lambda) && std::isfinite(visc0) && std::isfinite(sin_t) &&
         std::isfinite(cos_t) && std::isfinite(tcoef);
}

int fld_indicator_finish_dev(fld_ctx* c, double* indi, double* indic, double* visc,
                             double* forcey, double* forcez, double lambda, double visc0,
                             double sin_t, double cos_t, double tcoef) {
  if (!c) return FLD_E_ARG;
  if (!(indi && indic && visc && forcey && forcez)) return fail(c, FLD_E_ARG, "null field");
  if (!ind_args_ok(lambda, visc0, sin_t, cos_t, tcoef))
    return fail(c, FLD_E_ARG, "non-finite indicator constant");
  HIPF(c, hipSetDevice(c->device));
  return ind_finish(c, indi, indic, visc, forcey, forcez, lambda, visc0, sin_t, cos_t, tcoef);
}

int fld_indicator_dev(fld_ctx* c, const double* gx, const double* gy, const double* gz,
                      double* indi, double* indic, double* visc, double* forcey, double* forcez,
                      double lambda, double visc0, double sin_t, double cos_t, double tcoef) {
  if (!c) return FLD_E_ARG;
  if (!(gx && gy && gz && indi && indic && visc && forcey && forcez))
    return fail(c, FLD_E_ARG, "null field");
  if (!ind_args_ok(lambda, visc0, sin_t, cos_t, tcoef))
    return fail(c, FLD_E_ARG, "non-finite indicator constant");
  HIPF(c, hipSetDevice(c->device));
  // divergence straight into the solver's compact array, solve (initsolver(0) and
  // solver1d(indi, 0) are the pressure solve, ne = 1), solution into the interior of indi
  k_ind_div<<<nblocks(c->ncomp, 256), 256, 0, c->stream>>>(c->g, gx, gy, gz, c->work[0], nullptr);
  HIPF(c, hipGetLastError());
  int rc = solve_compact(c, c->work[0], 1, indi);
  if (rc) return rc;
  return ind_finish(c, indi, indic, visc, forcey, forcez, lambda, visc0, sin_t, cos_t, tcoef);
}

int fld_chkdiv_dev(fld_ctx* c, const double* u, const double* v, const double* w,
                   double* divtot, double* divmax) {
  if (!c) return FLD_E_ARG;
  if (!(u && v && w && divtot && divmax)) return fail(c, FLD_E_ARG, "null pointer");
  HIPF(c, hipSetDevice(c->device));
  const unsigned nb = std::min<unsigned>(1024, nblocks(c->ncomp, 256));
  k_div<<<nb, 256, 0, c->stream>>>(c->g, u, v, w, c->part);
  HIPF(c, hipGetLastError());
  std::vector<double> h(2 * (size_t)nb);
  HIPF(c, hipMemcpyAsync(h.data(), c->part, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                         c->stream));
  HIPF(c, hipStreamSynchronize(c->stream));
  double s = 0.0, m = 0.0;
  for (unsigned b = 0; b < nb; ++b) s += h[2 * b], m = std::max(m, h[2 * b + 1]);
  *divtot = s, *divmax = m;
  return FLD_OK;
}

}  // extern "C"
