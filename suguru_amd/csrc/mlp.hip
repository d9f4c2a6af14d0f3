// mlp.hip -- the learned membrane-traction surrogate: an MLP trained and run with hand-written
// BF16 MFMA kernels (v_mfma_f32_16x16x32_bf16) accumulating in FP32, on gfx950.
//
// Layers: din -> width (SiLU) -> ... (nhidden hidden layers) -> dout (linear).  Every GEMM
// dimension is padded to a multiple of 64 with zeros, so the tiles need no bounds checks;
// padded rows/columns of the weights stay zero (their gradients are zero and Adam keeps them
// at zero), so the padded network equals the unpadded one.
//
// GEMM: C[M x N] = A[M x K] B[K x N], 64 x 64 tile per 256-thread workgroup, each wavefront
// computes a 32 x 32 quarter as 2 x 2 MFMA 16x16x32 blocks; A and B are staged through LDS
// with k contiguous, so that each lane's MFMA operand (8 consecutive k) is one 16-byte read.
// A is either row-major [M x K] or its transpose ([K x M], for dW = dZ^T H); B is either
// [K x N] row-major (dX = dZ W) or [N x K] row-major (forward: B = W^T).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/memb.h"

namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TM = 64, TN = 64, TK = 32, THREADS = 256;
constexpr int LDK = TK + 8;  // LDS row stride in bf16 (pads rows apart by 16 bytes)

__device__ inline __bf16 to_bf16(float f) { return (__bf16)f; }  // round to nearest even
__device__ inline float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ inline float dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

enum Epi {
  EPI_BIAS_SILU = 0,  // Z = acc + b (fp32 out0), H = bf16(silu(Z)) (out1)
  EPI_BIAS = 1,       // Y = acc + b (fp32 out0)
  EPI_DSILU = 2,      // dZ = bf16(acc * silu'(Z)), Z read from aux (fp32)
  EPI_F32 = 3,        // plain fp32 store (weight gradients)
};

// A: if !TA, A[m][k] = Ag[m * lda + k]; if TA, A[m][k] = Ag[k * lda + m]
// B: if TB,  B[k][n] = Bg[n * ldb + k]; if !TB, B[k][n] = Bg[k * ldb + n]
// Tile BM x BN (64 x 64, or 32 x 32 when the 64-tile grid would leave compute units idle):
// four wavefronts in a 2 x 2 arrangement, each owning (BM/2) x (BN/2) as 16 x 16 MFMA blocks.
// Each output element sums its K-slices in the same order whatever the tile, so both tiles
// give the same bits.
template <int BM, int BN, bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(THREADS)
    k_gemm(int M, int N, int K, const __bf16* __restrict__ Ag, int lda,
           const __bf16* __restrict__ Bg, int ldb, const float* __restrict__ bias,
           const float* __restrict__ aux, float* __restrict__ out0, void* __restrict__ out1,
           int ldc) {
  constexpr int FM = BM / 32, FN = BN / 32;  // MFMA blocks per wavefront
  __shared__ __attribute__((aligned(16))) __bf16 As[BM * LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[BN * LDK];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int wm = (wv >> 1) * (BM / 2), wn = (wv & 1) * (BN / 2);
  f32x4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += TK) {
    // stage A (BM x 32) and B (32 x BN as [n][k]): 8 bf16 per thread, BM * 4 (BN * 4) threads
    if (tid < BM * 4) {
      if (!TA) {
        const int r = tid >> 2, kc = (tid & 3) * 8;
        *reinterpret_cast<bf16x8*>(&As[r * LDK + kc]) =
            *reinterpret_cast<const bf16x8*>(&Ag[(size_t)(m0 + r) * lda + k0 + kc]);
      } else {
        const int kk = tid / (BM / 8), mc = (tid % (BM / 8)) * 8;
        const bf16x8 v = *reinterpret_cast<const bf16x8*>(&Ag[(size_t)(k0 + kk) * lda + m0 + mc]);
#pragma unroll
        for (int e = 0; e < 8; ++e) As[(mc + e) * LDK + kk] = v[e];
      }
    }
    if (tid < BN * 4) {
      if (TB) {
        const int r = tid >> 2, kc = (tid & 3) * 8;
        *reinterpret_cast<bf16x8*>(&Bs[r * LDK + kc]) =
            *reinterpret_cast<const bf16x8*>(&Bg[(size_t)(n0 + r) * ldb + k0 + kc]);
      } else {
        const int kk = tid / (BN / 8), nc = (tid % (BN / 8)) * 8;
        const bf16x8 v = *reinterpret_cast<const bf16x8*>(&Bg[(size_t)(k0 + kk) * ldb + n0 + nc]);
#pragma unroll
        for (int e = 0; e < 8; ++e) Bs[(nc + e) * LDK + kk] = v[e];
      }
    }
    __syncthreads();
    // operands: lane holds row (lane % 16), k = 8 (lane / 16) .. +7
    bf16x8 a[FM], b[FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
      a[i] = *reinterpret_cast<const bf16x8*>(&As[(wm + i * 16 + (lane & 15)) * LDK + 8 * (lane >> 4)]);
#pragma unroll
    for (int j = 0; j < FN; ++j)
      b[j] = *reinterpret_cast<const bf16x8*>(&Bs[(wn + j * 16 + (lane & 15)) * LDK + 8 * (lane >> 4)]);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // result: lane holds rows 4 (lane / 16) + r, column lane % 16 of each 16 x 16 block
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
        const int col = n0 + wn + j * 16 + (lane & 15);
        const size_t o = (size_t)row * ldc + col;
        const float v = acc[i][j][r];
        if (EPI == EPI_BIAS_SILU) {
          const float zz = v + bias[col];
          out0[o] = zz;
          reinterpret_cast<__bf16*>(out1)[o] = to_bf16(silu(zz));
        } else if (EPI == EPI_BIAS) {
          out0[o] = v + bias[col];
        } else if (EPI == EPI_DSILU) {
          reinterpret_cast<__bf16*>(out1)[o] = to_bf16(v * dsilu(aux[o]));
        } else {
          out0[o] = v;
        }
      }
}

// Forward GEMM of the larger batches: C[M x N] = A[M x K] W^T (A row-major, W [N x K]
// row-major), 128 x 128 tile per 256-thread workgroup, each wavefront a 64 x 64 quarter as
// 4 x 4 MFMA 16x16x32 blocks.  Two LDS buffers: while the MFMAs consume K-slice s, the
// slice s + 1 is already in flight from HBM into registers and is stored to the other
// buffer afterwards, so one barrier per slice separates the two.  Each output element adds
// its K-slices in the same order as k_gemm, so both kernels give the same bits.
template <int EPI>
__global__ void __launch_bounds__(THREADS)
    k_gemm_fwd(int M, int N, int K, const __bf16* __restrict__ Ag, int lda,
               const __bf16* __restrict__ Bg, int ldb, const float* __restrict__ bias,
               float* __restrict__ out0, __bf16* __restrict__ out1, int ldc) {
  constexpr int BM = 128, BN = 128, F = 4;
  __shared__ __attribute__((aligned(16))) __bf16 As[2][BM * LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[2][BN * LDK];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int wm = (wv >> 1) * 64, wn = (wv & 1) * 64;
  // staging: thread t moves rows t >> 2 and 64 + (t >> 2), 8 k-values at (t & 3) * 8
  const int sr = tid >> 2, sk = (tid & 3) * 8;
  const __bf16* ga0 = Ag + (size_t)(m0 + sr) * lda + sk;
  const __bf16* ga1 = Ag + (size_t)(m0 + 64 + sr) * lda + sk;
  const __bf16* gb0 = Bg + (size_t)(n0 + sr) * ldb + sk;
  const __bf16* gb1 = Bg + (size_t)(n0 + 64 + sr) * ldb + sk;
  bf16x8 ra0 = *reinterpret_cast<const bf16x8*>(ga0), ra1 = *reinterpret_cast<const bf16x8*>(ga1);
  bf16x8 rb0 = *reinterpret_cast<const bf16x8*>(gb0), rb1 = *reinterpret_cast<const bf16x8*>(gb1);
  f32x4 acc[F][F];
#pragma unroll
  for (int i = 0; i < F; ++i)
#pragma unroll
    for (int j = 0; j < F; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  *reinterpret_cast<bf16x8*>(&As[0][sr * LDK + sk]) = ra0;
  *reinterpret_cast<bf16x8*>(&As[0][(64 + sr) * LDK + sk]) = ra1;
  *reinterpret_cast<bf16x8*>(&Bs[0][sr * LDK + sk]) = rb0;
  *reinterpret_cast<bf16x8*>(&Bs[0][(64 + sr) * LDK + sk]) = rb1;
  __syncthreads();
  int cur = 0;
  for (int k0 = 0; k0 < K; k0 += TK) {
    const bool more = k0 + TK < K;
    if (more) {
      ra0 = *reinterpret_cast<const bf16x8*>(ga0 + k0 + TK);
      ra1 = *reinterpret_cast<const bf16x8*>(ga1 + k0 + TK);
      rb0 = *reinterpret_cast<const bf16x8*>(gb0 + k0 + TK);
      rb1 = *reinterpret_cast<const bf16x8*>(gb1 + k0 + TK);
    }
    bf16x8 a[F], b[F];
#pragma unroll
    for (int i = 0; i < F; ++i)
      a[i] = *reinterpret_cast<const bf16x8*>(&As[cur][(wm + i * 16 + (lane & 15)) * LDK + 8 * (lane >> 4)]);
#pragma unroll
    for (int j = 0; j < F; ++j)
      b[j] = *reinterpret_cast<const bf16x8*>(&Bs[cur][(wn + j * 16 + (lane & 15)) * LDK + 8 * (lane >> 4)]);
#pragma unroll
    for (int i = 0; i < F; ++i)
#pragma unroll
      for (int j = 0; j < F; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    if (more) {
      const int nx = cur ^ 1;
      *reinterpret_cast<bf16x8*>(&As[nx][sr * LDK + sk]) = ra0;
      *reinterpret_cast<bf16x8*>(&As[nx][(64 + sr) * LDK + sk]) = ra1;
      *reinterpret_cast<bf16x8*>(&Bs[nx][sr * LDK + sk]) = rb0;
      *reinterpret_cast<bf16x8*>(&Bs[nx][(64 + sr) * LDK + sk]) = rb1;
    }
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int i = 0; i < F; ++i)
#pragma unroll
    for (int j = 0; j < F; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
        const int col = n0 + wn + j * 16 + (lane & 15);
        const size_t o = (size_t)row * ldc + col;
        const float v = acc[i][j][r];
        if (EPI == EPI_BIAS_SILU) {
          const float zz = v + bias[col];
          out0[o] = zz;
          out1[o] = to_bf16(silu(zz));
        } else {
          out0[o] = v + bias[col];
        }
      }
  (void)M, (void)N;
}

// x[batch][din] fp32 -> X[Mp][Dp] bf16 with (x - mean) * scale, zero padding
__global__ void k_input(int batch, int din, int Mp, int Dp, const float* __restrict__ x,
                        const float* __restrict__ mean, const float* __restrict__ scale,
                        __bf16* __restrict__ X) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long long)Mp * Dp) return;
  const int r = (int)(q / Dp), c = (int)(q % Dp);
  float v = 0.0f;
  if (r < batch && c < din) v = (x[(size_t)r * din + c] - mean[c]) * scale[c];
  X[q] = to_bf16(v);
}

// y[batch][dout] = omean + oscale * Y (fp32), per output
__global__ void k_output(int batch, int dout, int ldy, const float* __restrict__ omean,
                         const float* __restrict__ oscale, const float* __restrict__ Y,
                         float* __restrict__ y) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long long)batch * dout) return;
  const int r = (int)(q / dout), c = (int)(q % dout);
  y[q] = omean[c] + oscale[c] * Y[(size_t)r * ldy + c];
}

// loss = mean((Y - t')^2 w) over batch x dout with t' = (t - omean) / oscale;
// dY = bf16(2 (Y - t') w / (batch dout))
__global__ void k_loss(int batch, int dout, int Mp, int ldy, const float* __restrict__ omean,
                       const float* __restrict__ oscale, const float* __restrict__ Y,
                       const float* __restrict__ t,
                       const float* __restrict__ wt, __bf16* __restrict__ dY,
                       float* __restrict__ loss) {
  __shared__ float red[256];
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (q < (long long)Mp * ldy) {
    const int r = (int)(q / ldy), c = (int)(q % ldy);
    float g = 0.0f;
    if (r < batch && c < dout) {
      const float w = wt ? wt[c] : 1.0f;
      const float d = Y[q] - (t[(size_t)r * dout + c] - omean[c]) / oscale[c];
      const float inv = 1.0f / ((float)batch * (float)dout);
      l = d * d * w * inv;
      g = 2.0f * d * w * inv;
    }
    dY[q] = to_bf16(g);
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// db[c] = sum over rows of dZ[r][c] (bf16 in, fp32 out), one thread per column
__global__ void k_colsum(int rows, int cols, int ld, const __bf16* __restrict__ dZ,
                         float* __restrict__ db) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0.0f;
  for (int r = 0; r < rows; ++r) s += (float)dZ[(size_t)r * ld + c];
  db[c] = s;
}

// Adam on fp32 masters; refreshes the bf16 copy.  Padded entries have zero gradient and stay 0.
__global__ void k_adam(long long n, float lr, float b1, float b2, float eps, float c1, float c2,
                       float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, __bf16* __restrict__ wb) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const float gg = g[q];
  const float mm = b1 * m[q] + (1.0f - b1) * gg;
  const float vv = b2 * v[q] + (1.0f - b2) * gg * gg;
  m[q] = mm, v[q] = vv;
  const float ww = w[q] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
  w[q] = ww;
  if (wb) wb[q] = to_bf16(ww);
}

// zero-fill by a kernel (graph-capture safe, unlike a memset node)
__global__ void k_zero(float* __restrict__ a, int n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) a[q] = 0.0f;
}

__global__ void k_to_bf16(long long n, const float* __restrict__ a, __bf16* __restrict__ b) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) b[q] = to_bf16(a[q]);
}

inline int pad64(int v) { return (v + 63) / 64 * 64; }
inline int pad128(int v) { return (v + 127) / 128 * 128; }  // batch rows
inline unsigned nblk(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace mlp

using namespace mlp;

struct memb_mlp {
  int device = 0;
  int cus = 256;  // compute units of the device (GEMM tile choice)
  hipStream_t own = nullptr, stream = nullptr;
  std::string err;
  int din = 0, width = 0, nh = 0, dout = 0, maxb = 0;
  int Dp = 0, Wp = 0, Op = 0, Mp = 0, nl = 0;
  std::vector<int> in_p, out_p, in_r, out_r;  // padded and real layer shapes
  // per layer: fp32 master W [out_p][in_p], b [out_p]; bf16 W copy; Adam moments; grads
  std::vector<float*> W, b, mW, vW, mb, vb, gW, gb;
  std::vector<__bf16*> Wb;
  // activations
  __bf16* X = nullptr;             // [Mp][Dp]
  std::vector<float*> Z;           // hidden pre-activations [Mp][Wp]
  std::vector<__bf16*> H;          // hidden activations [Mp][Wp]
  float* Y = nullptr;              // [Mp][Op]
  __bf16* dY = nullptr;            // [Mp][Op]
  __bf16* dZa = nullptr;           // [Mp][Wp] (two buffers, ping-pong)
  __bf16* dZb = nullptr;
  float* loss = nullptr;
  float *mean = nullptr, *scale = nullptr;    // input normalisation [Dp]
  float *omean = nullptr, *oscale = nullptr;  // output de-normalisation [Op]
  long long step = 0;
  std::vector<void*> allocs;
};

namespace {
int mfail(memb_mlp* m, int code, const std::string& s) {
  if (m) m->err = s;
  return code;
}
#define XCHK(m, call)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return mfail((m), MEMB_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));  \
  } while (0)

template <typename T>
int dalloc(memb_mlp* m, T** p, size_t n) {
  void* d = nullptr;
  XCHK(m, hipMalloc(&d, n * sizeof(T)));
  XCHK(m, hipMemset(d, 0, n * sizeof(T)));
  m->allocs.push_back(d);
  *p = (T*)d;
  return MEMB_OK;
}

template <bool TA, bool TB, int EPI>
int gemm(memb_mlp* m, int M, int N, int K, const __bf16* A, int lda, const __bf16* B, int ldb,
         const float* bias, const float* aux, float* out0, void* out1, int ldc) {
  if (M % 64 || N % 64 || K % 32) return mfail(m, MEMB_E_ARG, "gemm shape not padded");
  if constexpr (!TA && TB && (EPI == EPI_BIAS_SILU || EPI == EPI_BIAS)) {
    // forward layers of batches large enough to fill the chip with 128 x 128 tiles
    if (M % 128 == 0 && N % 128 == 0 && (long long)(M / 128) * (N / 128) >= m->cus) {
      k_gemm_fwd<EPI><<<dim3(N / 128, M / 128), THREADS, 0, m->stream>>>(
          M, N, K, A, lda, B, ldb, bias, out0, reinterpret_cast<__bf16*>(out1), ldc);
      XCHK(m, hipGetLastError());
      return MEMB_OK;
    }
  }
  if ((long long)(M / TM) * (N / TN) < 2LL * m->cus) {
    // too few 64 x 64 tiles to fill the compute units (training batches): 32 x 32 tiles
    k_gemm<32, 32, TA, TB, EPI><<<dim3(N / 32, M / 32), THREADS, 0, m->stream>>>(
        M, N, K, A, lda, B, ldb, bias, aux, out0, out1, ldc);
  } else {
    k_gemm<TM, TN, TA, TB, EPI><<<dim3(N / TN, M / TM), THREADS, 0, m->stream>>>(
        M, N, K, A, lda, B, ldb, bias, aux, out0, out1, ldc);
  }
  XCHK(m, hipGetLastError());
  return MEMB_OK;
}

int forward(memb_mlp* m, int batch, const float* x) {
  const int Mp = pad128(batch);
  k_input<<<nblk((long long)Mp * m->Dp), 256, 0, m->stream>>>(batch, m->din, Mp, m->Dp, x, m->mean,
                                                            m->scale, m->X);
  const __bf16* in = m->X;
  int ldi = m->Dp;
  for (int i = 0; i < m->nl; ++i) {
    int rc;
    if (i < m->nl - 1)
      rc = gemm<false, true, EPI_BIAS_SILU>(m, Mp, m->out_p[i], m->in_p[i], in, ldi, m->Wb[i],
                                            m->in_p[i], m->b[i], nullptr, m->Z[i], m->H[i],
                                            m->out_p[i]);
    else
      rc = gemm<false, true, EPI_BIAS>(m, Mp, m->out_p[i], m->in_p[i], in, ldi, m->Wb[i],
                                       m->in_p[i], m->b[i], nullptr, m->Y, nullptr, m->out_p[i]);
    if (rc) return rc;
    if (i < m->nl - 1) in = m->H[i], ldi = m->out_p[i];
  }
  return MEMB_OK;
}

int backward(memb_mlp* m, int batch, const float* x, const float* t, const float* wt, float* loss) {
  int rc = forward(m, batch, x);
  if (rc) return rc;
  const int Mp = pad128(batch);
  k_zero<<<1, 64, 0, m->stream>>>(m->loss, 1);
  k_loss<<<nblk((long long)Mp * m->Op), 256, 0, m->stream>>>(batch, m->dout, Mp, m->Op,
                                                            m->omean, m->oscale, m->Y, t, wt,
                                                            m->dY, m->loss);
  const __bf16* dcur = m->dY;  // gradient w.r.t. the current layer's pre-activation
  int ldd = m->Op;
  for (int i = m->nl - 1; i >= 0; --i) {
    const __bf16* hin = i == 0 ? m->X : m->H[i - 1];
    const int ldh = i == 0 ? m->Dp : m->out_p[i - 1];
    // dW_i = dZ_i^T H_{i-1}: [out x in], K = Mp
    if ((rc = gemm<true, false, EPI_F32>(m, m->out_p[i], m->in_p[i], Mp, dcur, ldd, hin, ldh,
                                         nullptr, nullptr, m->gW[i], nullptr, m->in_p[i])))
      return rc;
    k_colsum<<<nblk(m->out_p[i]), 256, 0, m->stream>>>(Mp, m->out_p[i], ldd, dcur, m->gb[i]);
    if (i > 0) {
      // dZ_{i-1} = (dZ_i W_i) * silu'(Z_{i-1}): [Mp x in], K = out
      __bf16* dst = (i & 1) ? m->dZa : m->dZb;
      if ((rc = gemm<false, false, EPI_DSILU>(m, Mp, m->in_p[i], m->out_p[i], dcur, ldd, m->Wb[i],
                                              m->in_p[i], nullptr, m->Z[i - 1], nullptr, dst,
                                              m->in_p[i])))
        return rc;
      dcur = dst, ldd = m->in_p[i];
    }
  }
  XCHK(m, hipGetLastError());
  if (loss) {
    XCHK(m, hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    XCHK(m, hipStreamSynchronize(m->stream));
  }
  return MEMB_OK;
}
}  // namespace

extern "C" {

int memb_mlp_create(memb_mlp** out, int din, int width, int nhidden, int dout, int max_batch,
                    int device) {
  if (!out || din <= 0 || width <= 0 || nhidden <= 0 || dout <= 0 || max_batch <= 0)
    return MEMB_E_ARG;
  memb_mlp* m = new memb_mlp();
  *out = m;
  m->device = device;
  XCHK(m, hipSetDevice(device));
  XCHK(m, hipStreamCreate(&m->own));
  m->stream = m->own;
  (void)hipDeviceGetAttribute(&m->cus, hipDeviceAttributeMultiprocessorCount, device);
  m->din = din, m->width = width, m->nh = nhidden, m->dout = dout, m->maxb = max_batch;
  m->Dp = pad64(din), m->Wp = pad64(width), m->Op = pad64(dout), m->Mp = pad128(max_batch);
  m->nl = nhidden + 1;
  for (int i = 0; i < m->nl; ++i) {
    m->in_p.push_back(i == 0 ? m->Dp : m->Wp);
    m->in_r.push_back(i == 0 ? din : width);
    m->out_p.push_back(i == m->nl - 1 ? m->Op : m->Wp);
    m->out_r.push_back(i == m->nl - 1 ? dout : width);
  }
  int rc;
  for (int i = 0; i < m->nl; ++i) {
    const size_t nw = (size_t)m->out_p[i] * m->in_p[i], nb = m->out_p[i];
    float *w, *bb, *a1, *a2, *a3, *a4, *g1, *g2;
    __bf16* wb;
    if ((rc = dalloc(m, &w, nw)) || (rc = dalloc(m, &bb, nb)) || (rc = dalloc(m, &a1, nw)) ||
        (rc = dalloc(m, &a2, nw)) || (rc = dalloc(m, &a3, nb)) || (rc = dalloc(m, &a4, nb)) ||
        (rc = dalloc(m, &g1, nw)) || (rc = dalloc(m, &g2, nb)) || (rc = dalloc(m, &wb, nw)))
      return rc;
    m->W.push_back(w), m->b.push_back(bb), m->mW.push_back(a1), m->vW.push_back(a2);
    m->mb.push_back(a3), m->vb.push_back(a4), m->gW.push_back(g1), m->gb.push_back(g2);
    m->Wb.push_back(wb);
  }
  for (int i = 0; i < nhidden; ++i) {
    float* z;
    __bf16* h;
    if ((rc = dalloc(m, &z, (size_t)m->Mp * m->Wp)) || (rc = dalloc(m, &h, (size_t)m->Mp * m->Wp)))
      return rc;
    m->Z.push_back(z), m->H.push_back(h);
  }
  if ((rc = dalloc(m, &m->X, (size_t)m->Mp * m->Dp)) || (rc = dalloc(m, &m->Y, (size_t)m->Mp * m->Op)) ||
      (rc = dalloc(m, &m->dY, (size_t)m->Mp * m->Op)) ||
      (rc = dalloc(m, &m->dZa, (size_t)m->Mp * m->Wp)) ||
      (rc = dalloc(m, &m->dZb, (size_t)m->Mp * m->Wp)) || (rc = dalloc(m, &m->loss, 1)) ||
      (rc = dalloc(m, &m->mean, (size_t)m->Dp)) || (rc = dalloc(m, &m->scale, (size_t)m->Dp)) ||
      (rc = dalloc(m, &m->omean, (size_t)m->Op)) || (rc = dalloc(m, &m->oscale, (size_t)m->Op)))
    return rc;
  std::vector<float> one(std::max(m->Dp, m->Op), 1.0f);
  XCHK(m, hipMemcpy(m->scale, one.data(), m->Dp * sizeof(float), hipMemcpyHostToDevice));
  XCHK(m, hipMemcpy(m->oscale, one.data(), m->Op * sizeof(float), hipMemcpyHostToDevice));
  return MEMB_OK;
}

void memb_mlp_destroy(memb_mlp* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (void* p : m->allocs) (void)hipFree(p);
  if (m->own) (void)hipStreamDestroy(m->own);
  delete m;
}

const char* memb_mlp_last_error(const memb_mlp* m) { return m ? m->err.c_str() : "null mlp"; }

int memb_mlp_set_stream(memb_mlp* m, void* s) {
  if (!m) return MEMB_E_ARG;
  m->stream = s ? (hipStream_t)s : m->own;
  return MEMB_OK;
}

int memb_mlp_num_layers(const memb_mlp* m) { return m ? m->nl : 0; }

int memb_mlp_layer_shape(const memb_mlp* m, int i, int* out_dim, int* in_dim) {
  if (!m || i < 0 || i >= m->nl) return MEMB_E_ARG;
  if (out_dim) *out_dim = m->out_r[i];
  if (in_dim) *in_dim = m->in_r[i];
  return MEMB_OK;
}

int memb_mlp_set_norm(memb_mlp* m, const float* mean, const float* scale, const float* out_mean,
                      const float* out_scale) {
  if (!m || !mean || !scale || !out_mean || !out_scale) return MEMB_E_ARG;
  for (int c = 0; c < m->dout; ++c)
    if (!(out_scale[c] > 0.0f)) return mfail(m, MEMB_E_ARG, "out_scale must be positive");
  XCHK(m, hipSetDevice(m->device));
  XCHK(m, hipMemcpy(m->mean, mean, m->din * sizeof(float), hipMemcpyHostToDevice));
  XCHK(m, hipMemcpy(m->scale, scale, m->din * sizeof(float), hipMemcpyHostToDevice));
  XCHK(m, hipMemcpy(m->omean, out_mean, m->dout * sizeof(float), hipMemcpyHostToDevice));
  XCHK(m, hipMemcpy(m->oscale, out_scale, m->dout * sizeof(float), hipMemcpyHostToDevice));
  return MEMB_OK;
}

int memb_mlp_set_layer(memb_mlp* m, int i, const float* W, const float* b) {
  if (!m || i < 0 || i >= m->nl || !W || !b) return MEMB_E_ARG;
  XCHK(m, hipSetDevice(m->device));
  const int op = m->out_p[i], ip = m->in_p[i], orr = m->out_r[i], ir = m->in_r[i];
  std::vector<float> w((size_t)op * ip, 0.0f), bb(op, 0.0f);
  for (int r = 0; r < orr; ++r) {
    for (int c = 0; c < ir; ++c) w[(size_t)r * ip + c] = W[(size_t)r * ir + c];
    bb[r] = b[r];
  }
  XCHK(m, hipMemcpy(m->W[i], w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
  XCHK(m, hipMemcpy(m->b[i], bb.data(), bb.size() * sizeof(float), hipMemcpyHostToDevice));
  k_to_bf16<<<nblk((long long)w.size()), 256, 0, m->stream>>>((long long)w.size(), m->W[i], m->Wb[i]);
  XCHK(m, hipGetLastError());
  XCHK(m, hipStreamSynchronize(m->stream));
  return MEMB_OK;
}

static int get_padded(memb_mlp* m, int i, const float* dW, const float* db, float* W, float* b) {
  const int op = m->out_p[i], ip = m->in_p[i], orr = m->out_r[i], ir = m->in_r[i];
  std::vector<float> w((size_t)op * ip), bb(op);
  XCHK(m, hipSetDevice(m->device));
  XCHK(m, hipStreamSynchronize(m->stream));
  XCHK(m, hipMemcpy(w.data(), dW, w.size() * sizeof(float), hipMemcpyDeviceToHost));
  XCHK(m, hipMemcpy(bb.data(), db, bb.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (int r = 0; r < orr; ++r) {
    for (int c = 0; c < ir; ++c) W[(size_t)r * ir + c] = w[(size_t)r * ip + c];
    b[r] = bb[r];
  }
  return MEMB_OK;
}

int memb_mlp_get_layer(memb_mlp* m, int i, float* W, float* b) {
  if (!m || i < 0 || i >= m->nl || !W || !b) return MEMB_E_ARG;
  return get_padded(m, i, m->W[i], m->b[i], W, b);
}

int memb_mlp_get_grad(memb_mlp* m, int i, float* dW, float* db) {
  if (!m || i < 0 || i >= m->nl || !dW || !db) return MEMB_E_ARG;
  return get_padded(m, i, m->gW[i], m->gb[i], dW, db);
}

int memb_mlp_init(memb_mlp* m, uint64_t seed) {
  if (!m) return MEMB_E_ARG;
  std::mt19937_64 rng(seed);
  for (int i = 0; i < m->nl; ++i) {
    const int orr = m->out_r[i], ir = m->in_r[i];
    const float a = std::sqrt(6.0f / (float)(ir + orr));
    std::uniform_real_distribution<float> U(-a, a);
    std::vector<float> w((size_t)orr * ir), bb(orr, 0.0f);
    for (auto& v : w) v = U(rng);
    int rc = memb_mlp_set_layer(m, i, w.data(), bb.data());
    if (rc) return rc;
  }
  m->step = 0;
  return MEMB_OK;
}

int memb_mlp_forward(memb_mlp* m, int batch, const float* x, float* y) {
  if (!m || batch <= 0 || batch > m->maxb || !x || !y) return MEMB_E_ARG;
  XCHK(m, hipSetDevice(m->device));
  int rc = forward(m, batch, x);
  if (rc) return rc;
  k_output<<<nblk((long long)batch * m->dout), 256, 0, m->stream>>>(batch, m->dout, m->Op,
                                                                   m->omean, m->oscale, m->Y, y);
  XCHK(m, hipGetLastError());
  return MEMB_OK;
}

int memb_mlp_grad_only(memb_mlp* m, int batch, const float* x, const float* t, const float* wt,
                       float* loss) {
  if (!m || batch <= 0 || batch > m->maxb || !x || !t) return MEMB_E_ARG;
  XCHK(m, hipSetDevice(m->device));
  return backward(m, batch, x, t, wt, loss);
}

int memb_mlp_train_step(memb_mlp* m, int batch, const float* x, const float* t, const float* wt,
                        float lr, float* loss) {
  if (!m || batch <= 0 || batch > m->maxb || !x || !t) return MEMB_E_ARG;
  XCHK(m, hipSetDevice(m->device));
  int rc = backward(m, batch, x, t, wt, loss);
  if (rc) return rc;
  ++m->step;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
  for (int i = 0; i < m->nl; ++i) {
    const long long nw = (long long)m->out_p[i] * m->in_p[i];
    k_adam<<<nblk(nw), 256, 0, m->stream>>>(nw, lr, b1, b2, eps, c1, c2, m->W[i], m->gW[i],
                                             m->mW[i], m->vW[i], m->Wb[i]);
    k_adam<<<nblk(m->out_p[i]), 256, 0, m->stream>>>(m->out_p[i], lr, b1, b2, eps, c1, c2,
                                                      m->b[i], m->gb[i], m->mb[i], m->vb[i],
                                                      nullptr);
  }
  XCHK(m, hipGetLastError());
  return MEMB_OK;
}

}  // extern "C"
