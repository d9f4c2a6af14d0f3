/*
 * memb.h -- C ABI of the capsule-membrane library (libmemb.so, gfx950).
 *
 * Two ways to get the membrane load of every capsule:
 *
 *  1. The FP64 spherical-harmonic membrane law of the reference (neo-Hookean shear plus
 *     bending about the reference curvature, dealiased on the 2x grid):
 *       complagrforces      forcing.f90:11-114   (force grid = harmonic grid: TRSSPH is the
 *                                                 identity)
 *       cal_spectral_force  sph.f90:1523-1755
 *     -> memb_traction_f64_dev, one workgroup per capsule, every capsule in one launch.
 *
 *  2. A learned surrogate of that law: an MLP from the capsule's real spherical-harmonic
 *     shape coefficients to the traction at its points, run and trained with hand-written
 *     BF16 MFMA kernels that accumulate in FP32 (memb_mlp_*).  It has no counterpart in the
 *     reference; the FP64 law is its training-label source and its fallback.
 *
 * Capsule geometry: degree L, (L+1) Gauss-Legendre colatitudes (increasing) x 2(L+1) uniform
 * longitudes phi_j = 2 pi j / (2(L+1)); point l = i * 2(L+1) + j (latitude-major,
 * forcing.f90:37-43); positions x = xc + R sin(t) sin(p), y = yc + R sin(t) cos(p),
 * z = zc + R cos(t) for a sphere (sph.f90:84-86).  The reference (stress-free) state is the
 * unit sphere (sph.f90:56-70).  Point arrays are capsule-major: x[c * npts + l].
 *
 * Shape coefficients (SPHEREPACK convention, sph1.f90:29-79):
 *   f = sum_{m<=n<=L} Pbar_n^m(cos t) (a_mn cos m p - b_mn sin m p), m = 0 terms halved,
 * packed per coordinate as a_mn for m = 0..L, n = m..L, then b_mn for m = 1..L, n = m..L:
 * (L+1)^2 values per coordinate, 3 (L+1)^2 per capsule (x block, y block, z block).
 *
 * Errors: IBX-style negative codes (0 = OK) and memb_last_error().
 */
#ifndef MEMB_H
#define MEMB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MEMB_OK 0
#define MEMB_E_ARG (-1)
#define MEMB_E_HIP (-2)
#define MEMB_E_STATE (-5)

typedef struct memb_plan memb_plan;

/* Tables (Gauss grid, Legendre functions and their theta-derivatives up to 4th order, the
 * unit-sphere reference state on the 2x grid) for degree L (2 <= L <= 23) on HIP device
 * `device`.  Replaces capsule_init_cal_coeff (sph.f90:16-191). */
int memb_plan_create(memb_plan** out, int L, int device);
void memb_plan_destroy(memb_plan* p);
const char* memb_last_error(const memb_plan* p);
int memb_set_stream(memb_plan* p, void* hip_stream);
int memb_sync(memb_plan* p);
/* points per capsule (L+1) * 2(L+1), and coefficients per coordinate (L+1)^2 */
int memb_npts(const memb_plan* p);
int memb_ncoef(const memb_plan* p);

/* The FP64 law for ncap capsules (device pointers, capsule-major).  gs: shear modulus
 * (gscell, param.f90:88); eb: bending modulus (ebend, param.f90:91; 0 = no bending).
 * Writes the traction fx,fy,fz (load per unit area, forcing.f90:97-99), the filtered unit
 * normal nx,ny,nz (forcing.f90:100-102) and the point area (forcing.f90:91,96). */
int memb_traction_f64_dev(memb_plan* p, int64_t ncap, double gs, double eb, const double* x,
                          const double* y, const double* z, double* fx, double* fy, double* fz,
                          double* nx, double* ny, double* nz, double* area);

/* Shape coefficients of ncap capsules (device): coef[c * 3 (L+1)^2 + ...], packed as above.
 * The analysis of cal_3dcrd_spectral_coeff (sph1.f90:29-37). */
int memb_coeffs_f64_dev(memb_plan* p, int64_t ncap, const double* x, const double* y,
                        const double* z, double* coef);

/* Geometry only (no law): the unit normal (x_theta x x_phi)/|.| and the point area on the
 * force grid, for capsules whose load comes from the surrogate.  The area equals the law's;
 * the normal is taken on the force grid instead of being filtered from the 2x grid.  skip
 * (optional device int array of ncap): capsules with skip[c] != 0 are not written (the law
 * writes theirs). */
int memb_geometry_f64_dev(memb_plan* p, int64_t ncap, const double* x, const double* y,
                          const double* z, double* nx, double* ny, double* nz, double* area,
                          const int* skip);

/* The FP64 law for the capsules c with sel[c] != 0 only (device int array of ncap); the
 * outputs of the other capsules are left as they are.  The surrogate's fallback: capsules
 * its gate rejects are recomputed in place, with no host round trip (graph-capturable). */
int memb_traction_select_f64_dev(memb_plan* p, int64_t ncap, const int* sel, double gs,
                                 double eb, const double* x, const double* y, const double* z,
                                 double* fx, double* fy, double* fz, double* nx, double* ny,
                                 double* nz, double* area);

/* ---- co-rotated frame of the surrogate (no reference counterpart) --------------------------- */
/* The law is translation invariant and rotation covariant, f(R x + c) = R f(x), so the
 * surrogate works on the deformation only.  Per capsule (X: the 3 x npts stress-free points,
 * x block, y block, z block): centroid c, R = orthogonal polar factor of
 * sum_p (x_p - c) X_p^T, delta_p = R^T (x_p - c) - X_p written as delta[c][3][npts], its RMS
 * amplitude amp[c] (= +inf when the polar factor fails: det <= 0.5), the direction
 * delta / amp in dhx, dhy, dhz (capsule-major, like x), and R row-major in rot[c][9]. */
int memb_corot_f64_dev(memb_plan* p, int64_t ncap, const double* X, const double* x,
                       const double* y, const double* z, double* delta, double* dhx,
                       double* dhy, double* dhz, double* rot, double* amp);
/* f = R (flin + amp^2 resid) per point: flin[c][3][npts] FP64 (the linear response in the
 * co-rotated frame), resid[c][3][npts] FP32 (the learned residual per amp^2; NULL = 0). */
int memb_corot_apply_f64_dev(memb_plan* p, int64_t ncap, const double* flin, const float* resid,
                             const double* rot, const double* amp, double* fx, double* fy,
                             double* fz);

/* The anchored surrogate's step (no reference counterpart: the learned stand-in of
 * complagrforces, forcing.f90:11-114).  Rows are [ncap][3][npts] FP64 in the co-rotated frame.
 * Prediction and gate: c = <delta, delta0> / |delta0|^2 (0 if delta0 = 0), f' = (lin + c^2 m0)
 * + b (m0 = NULL: lin + b), f = R f' into fx, fy, fz; sel[c] = 0 when beta * rms(delta -
 * delta0) <= tol and every value of f is finite, else 1; drift (optional, [ncap]) the rms.
 * pre (optional): the drift test's result from memb_anchored_gate_f64_dev; capsules with
 * pre[c] != 0 get sel[c] = 1 and nothing else is written for them (the law computes them).
 * late (optional): late[c] = 1 for the other capsules rejected here (non-finite prediction),
 * else 0: the law's second call. */
int memb_anchored_predict_f64_dev(memb_plan* p, int64_t ncap, const double* lin,
                                  const double* m0, const double* b, const double* delta,
                                  const double* delta0, const double* rot, double beta,
                                  double tol, double* fx, double* fy, double* fz, int* sel,
                                  double* drift, const int* pre, int* late);
/* The drift test alone, as memb_anchored_predict_f64_dev computes it (same sums, same order):
 * pre[c] = 0 when beta * rms(delta - delta0) <= tol, else 1.  It needs no linear response, so
 * the law can start on the rejected capsules while J delta is computed for the others. */
int memb_anchored_gate_f64_dev(memb_plan* p, int64_t ncap, const double* delta,
                               const double* delta0, double beta, double tol, int* pre);
/* New anchors for the capsules with sel != 0, after the law has written their traction into
 * fx, fy, fz: refreshed[c] += 1 (int64); b = R^T f - lin (- m_now), delta0 = delta, m0 = m_now
 * and s0 = amp (when m_now is given).  slot (optional): m_now has a row only for the capsules
 * with slot[c] >= 0 (row slot[c]); a selected capsule with slot[c] < 0 keeps its old anchor.
 * Without slot, m_now (if given) has one row per capsule.  Other capsules are not touched. */
int memb_anchored_update_f64_dev(memb_plan* p, int64_t ncap, const int* sel, const int* slot,
                                 const double* rot, const double* fx, const double* fy,
                                 const double* fz, const double* lin, const double* m_now,
                                 const double* delta, const double* amp, double* b,
                                 double* delta0, double* m0, double* s0, int64_t* refreshed);
/* The first k capsules with sel != 0, in capsule order: idx[0..k) = their indices (padded
 * with 0), slot[c] = their rank, or -1 for every other capsule (int32 device arrays). */
int memb_compact_sel_dev(memb_plan* p, int64_t ncap, const int* sel, int k, int* idx,
                         int* slot);

/* ---- learned surrogate: MLP, BF16 MFMA with FP32 accumulation ----------------------------- */
/* Layers: din -> w (SiLU) -> w (SiLU) -> ... (nhidden layers of width w) -> dout, linear
 * output.  Weights are FP32 masters; the GEMMs read BF16 copies, refreshed whenever the
 * masters change (set_layer, init, each Adam update).  Every GEMM dimension is padded
 * internally to a multiple of 64.  The loss is taken on the normalised outputs y'. */
typedef struct memb_mlp memb_mlp;
int memb_mlp_create(memb_mlp** out, int din, int width, int nhidden, int dout, int max_batch,
                    int device);
void memb_mlp_destroy(memb_mlp* m);
const char* memb_mlp_last_error(const memb_mlp* m);
int memb_mlp_set_stream(memb_mlp* m, void* hip_stream);
/* number of layers (nhidden + 1) and the shape of layer i: W_i is out x in, row-major */
int memb_mlp_num_layers(const memb_mlp* m);
int memb_mlp_layer_shape(const memb_mlp* m, int i, int* out_dim, int* in_dim);
/* copy FP32 weights/biases of layer i in or out (host pointers) */
int memb_mlp_set_layer(memb_mlp* m, int i, const float* W, const float* b);
int memb_mlp_get_layer(memb_mlp* m, int i, float* W, float* b);
/* input normalisation x' = (x - mean) * scale (din values each) and output de-normalisation
 * y = out_mean + out_scale * y' (dout values each): the network maps x' to y' */
int memb_mlp_set_norm(memb_mlp* m, const float* mean, const float* scale, const float* out_mean,
                      const float* out_scale);
/* seeded init: W ~ U(-a, a), a = sqrt(6 / (in + out)) (SiLU layers) ; b = 0 */
int memb_mlp_init(memb_mlp* m, uint64_t seed);
/* forward for `batch` rows: x[batch][din] FP32 (device) -> y[batch][dout] FP32 (device) */
int memb_mlp_forward(memb_mlp* m, int batch, const float* x, float* y);
/* one training step on a batch: forward, loss = mean over rows and outputs of
 * (y' - t')^2 * wt (normalised outputs and targets; wt[dout] FP32 per-output weight,
 * NULL = 1), backward, Adam update
 * (lr, beta1 0.9, beta2 0.999, eps 1e-8).  *loss (host) receives the batch loss. */
int memb_mlp_train_step(memb_mlp* m, int batch, const float* x, const float* t,
                        const float* wt, float lr, float* loss);
/* gradients of the loss of the last train step (host copy, FP32), for checking */
int memb_mlp_get_grad(memb_mlp* m, int i, float* dW, float* db);
/* run the step without the Adam update (gradients only) */
int memb_mlp_grad_only(memb_mlp* m, int batch, const float* x, const float* t, const float* wt,
                       float* loss);

#ifdef __cplusplus
}
#endif

#endif /* MEMB_H */
