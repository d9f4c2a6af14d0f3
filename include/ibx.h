/*
 * ibx.h -- C ABI of the MI355X (gfx950) immersed-boundary transfer library (libibx.so).
 *
 * The library replaces the three module procedures of the reference's
 * `mod_interp_spread` (interp_spread.f90:1-9) and the regularized delta of
 * `mod_kernel` (kernel.f90:7-25).  The reference procedures take no arguments and
 * talk through module state (common.f90:6-9,67-88; param.f90:19-27); here the same
 * state is passed explicitly as plain pointers and sizes:
 *
 *   reference                                   | this ABI
 *   --------------------------------------------+------------------------------------
 *   kernel(r)            kernel.f90:7           | ibx_kernel
 *   eulr2lagr            interp_spread.f90:12   | ibx_eulr2lagr      (host arrays)
 *                                               | ibx_interp_f64_dev (HBM-resident)
 *   lagr2eulr            interp_spread.f90:510  | ibx_lagr2eulr      (host arrays)
 *                                               | ibx_spread_f64_dev (HBM-resident)
 *   lagr2eulrind         interp_spread.f90:978  | ibx_lagr2eulrind   (host arrays)
 *                                               | ibx_spread_cc_f64_dev
 *   updthalos(var,dir)   bound.f90:141-190      | fused into the spread outputs
 *   Update_Pos (arith.)  Update_Pos.f90:77-186  | ibx_update_pos / ibx_update_pos_dev
 *
 * Semantics are those of the reference on ONE rank (dims = 1x1): i (x) and j (y) are
 * periodic through the nine image copies of every capsule (interp_spread.f90:60-146,
 * slot table below), k (z) is wall-bounded (window clamped to [1,kmax], or to
 * [0,kmax+1] for lagr2eulrind).  All arithmetic is IEEE FP64 in the reference's
 * evaluation order: for every Lagrangian point and image slot, the clamped stencil is
 * summed k -> j -> i with products formed left to right; the per-slot partial
 * velocities are added in slot order 0..8 (interp_spread.f90:436-440); every grid cell
 * of a spread receives its contributions in the reference's loop order
 * (capsule p, image loop nb = 0..8, point l; interp_spread.f90:825-878).  Results are
 * therefore bit-identical to that loop nest on the same inputs.
 *
 * Field arrays are Fortran `real(8), dimension(0:i1,0:j1,0:k1)` (common.f90:6-9):
 * (itot+2)*(jtot+2)*(ktot+2) doubles, i fastest.
 *
 * Lagrangian arrays are structure-of-arrays over all points of all capsules, capsule
 * after capsule; cap_off[0..ncap] (int64, cap_off[0] = 0, non-decreasing) gives the
 * first point of each capsule, so capsule p owns points cap_off[p] .. cap_off[p+1]-1.
 * In the reference every capsule has nl points (param.f90:77) stored as
 * ap(p)%xfp(1:nl) (common.f90:67-81); uneven capsules are allowed here.
 *
 * Error convention: every int-returning call returns IBX_OK (0) or a negative
 * IBX_E_* code; ibx_last_error() gives the message.  Calls on one context are
 * stream-ordered and not re-entrant (the reference routines are not re-entrant either:
 * they work on module state).
 */
#ifndef IBX_H
#define IBX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IBX_OK 0
#define IBX_E_ARG (-1)      /* invalid argument (null pointer, bad size, bad offsets) */
#define IBX_E_HIP (-2)      /* HIP runtime error */
#define IBX_E_DOMAIN (-3)   /* non-finite position (or, in a spread, force or area), or a
                               point beyond the one-image range */
#define IBX_E_CAPACITY (-4) /* internal candidate list overflow in a captured graph: re-plan */
#define IBX_E_STATE (-5)    /* call out of order (e.g. batch run before bind) */

/* Grid of one rank.  Mirrors param.f90:19-27 and common_mpi boundleftmyid/boundfrontmyid
 * (common.f90:105).  With one rank imax = itot, jmax = jtot, kmax = ktot and both bounds are
 * 0.  dx = 1/dxi etc. are passed (not recomputed) so that the caller's constants are used
 * bit for bit. */
typedef struct ibx_grid {
  int32_t itot, jtot, ktot; /* interior cells in i (x, periodic), j (y, periodic), k (z, walls) */
  int32_t reserved;         /* must be 0 */
  double lx, ly, lz;        /* domain lengths */
  double dxi, dyi, dzi;     /* inverse spacings, itot/lx etc. */
  double dx, dy, dz;        /* spacings, 1/dxi etc. */
  double boundleft;         /* x of the rank's left boundary (0 on one rank) */
  double boundfront;        /* y of the rank's front boundary (0 on one rank) */
} ibx_grid;

typedef struct ibx_ctx ibx_ctx;

/* ---- context ----------------------------------------------------------------------- */

/* Create a context on HIP device `device`: validates the grid, creates a stream of default
 * priority, and owns all device work buffers.  No reference counterpart (module state). */
int ibx_create(ibx_ctx** out, const ibx_grid* grid, int device);
void ibx_destroy(ibx_ctx* ctx);
const char* ibx_last_error(const ibx_ctx* ctx);
/* Run subsequent device work on `hip_stream` (a hipStream_t; NULL = the context's own). */
int ibx_set_stream(ibx_ctx* ctx, void* hip_stream);
/* Wait for the context's stream; reports errors raised by device-side checks. */
int ibx_sync(ibx_ctx* ctx);
/* Spread configuration of the one-call entry points (ibx_lagr2eulr, ibx_lagr2eulrind,
 * ibx_spread_f64_dev, ibx_spread_cc_f64_dev): an index of ibx_batch_config_name's list.
 * Default: objects-w8-o4.  A batch chooses its own (ibx_batch_set_config / _autotune).  No
 * reference counterpart: the reference has one loop nest. */
int ibx_set_config(ibx_ctx* ctx, int choice);
/* Library version string. */
const char* ibx_version(void);

/* ---- the regularized delta ----------------------------------------------------------- */

/* Roma-Peskin-Berger 3-point delta multiplied by h: kernel.f90:7-25.  Host FP64. */
double ibx_kernel(double r);
/* The same kernel.f90:7-25 on the device: out[q] = kernel(r[q]), q < n (device arrays), with
 * the device function every interpolation and spread weight uses.  Asynchronous. */
int ibx_kernel_dev(ibx_ctx* ctx, int64_t n, const double* r, double* out);
/* kerneltest, kernel.f90:43-63, evaluated on the device with that device kernel, summed in
 * the reference's order (k outer, i inner); waits and writes the sum to *sum (host). */
int ibx_kerneltest_dev(ibx_ctx* ctx, double* sum);

/* ---- wall collisions and the indicator's position shift -------------------------------- */

/* collisions, collisions.f90:10-32: fz[q] += the Morse force of the bottom wall on point q,
 * sum over the wall-plane lattice j, k in [-jm, jm], jm = nint(sqrt(max(rad0^2 - z^2, 0))
 * dyi), of a2 (exp(2 betacol (rad0 - s)) - exp(betacol (rad0 - s))) |z / s|
 * min(area, dx^2) / max(area, dx^2) over the lattice distances s < rad0.  a2 = 2 col / Rep^2
 * (collisions.f90:24), rad0 = 2.5 dz and betacol from param.f90:64.  Device arrays of npts;
 * asynchronous.  The exponentials are the device's, so results agree with a host evaluation
 * to rounding (~1e-15 relative), not bit for bit.  rad0 must span at most 64 cells. */
int ibx_collisions_dev(ibx_ctx* ctx, int64_t npts, const double* z, const double* area,
                       double* fz, double a2, double rad0, double betacol);
/* z[q] = z[q] + s in place, q < npts: the indicator's shift of ap(p)%zfp up by (ktot/4) dz
 * and back (indicator.f90:25-31, 37-43), so that positions return as (z + s) - s exactly as
 * in the reference.  Asynchronous. */
int ibx_shift_z_dev(ibx_ctx* ctx, int64_t npts, double* z, double s);

/* ---- host-array entry points (the Fortran drop-in calls these) -------------------------- */

/* eulr2lagr, interp_spread.f90:12-475.  u,v,w: fields (0:i1,0:j1,0:k1).  x,y,z: positions of
 * cap_off[ncap] points.  Writes ul,vl,wl (the owner sums over all image slots). */
int ibx_eulr2lagr(ibx_ctx* ctx, const double* u, const double* v, const double* w,
                  int64_t ncap, const int64_t* cap_off, const double* x, const double* y,
                  const double* z, double* ul, double* vl, double* wl);

/* lagr2eulr, interp_spread.f90:510-922 (zeroing :811-815, scatter :852-878, halos
 * :904-912 -> bound.f90:141-190).  Writes every element of forcex/forcey/forcez. */
int ibx_lagr2eulr(ibx_ctx* ctx, int64_t ncap, const int64_t* cap_off, const double* x,
                  const double* y, const double* z, const double* fx, const double* fy,
                  const double* fz, const double* area, double* forcex, double* forcey,
                  double* forcez);

/* lagr2eulrind, interp_spread.f90:978-1387: cell-centred weights in all three directions,
 * window clamped to [0,max+1], k halos keep their contributions.  Writes gx,gy,gz. */
int ibx_lagr2eulrind(ibx_ctx* ctx, int64_t ncap, const int64_t* cap_off, const double* x,
                     const double* y, const double* z, const double* xn, const double* yn,
                     const double* zn, const double* area, double* gx, double* gy,
                     double* gz);

/* ---- HBM-resident entry points ---------------------------------------------------------- */
/* Same semantics as above; every pointer (cap_off included) is device memory; work is
 * enqueued on the context stream and returns without waiting.  Domain errors found on the
 * device are reported by the next ibx_sync().  The spread and Update_Pos entry points check
 * the device-held cap_off on the device first (cap_off[0] = 0, non-decreasing,
 * cap_off[ncap] = npts); a table that fails makes every kernel that indexes through it
 * return at once, and the next ibx_sync() reports IBX_E_ARG (the outputs are then
 * undefined). */

int ibx_interp_f64_dev(ibx_ctx* ctx, const double* u, const double* v, const double* w,
                       int64_t ncap, const int64_t* cap_off, int64_t npts, const double* x,
                       const double* y, const double* z, double* ul, double* vl,
                       double* wl);

int ibx_spread_f64_dev(ibx_ctx* ctx, int64_t ncap, const int64_t* cap_off, int64_t npts,
                       const double* x, const double* y, const double* z, const double* fx,
                       const double* fy, const double* fz, const double* area,
                       double* forcex, double* forcey, double* forcez);

/* The plain spread (ibx_spread_f64_dev) allocating nothing: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under hipGraph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Its work buffers (object boxes, items, sources,
 * coverage bitmap, capsule masks, queue counters) live in one fixed 64 MiB device array per
 * GPU, held until the process ends; a call whose buffers do not fit returns IBX_E_ARG before
 * any device work, and one that needs more work items than fit reports IBX_E_CAPACITY at the
 * next ibx_sync.  Every call on a GPU, from any context, uses that GPU's one array, which
 * keeps the last call's contents (each call clears what must start at zero): calls that
 * overlap on one GPU (two streams, two host threads) race.  Same bits as ibx_spread_f64_dev. */
int ibx_spread_f64_dev_noalloc(ibx_ctx* ctx, int64_t ncap, const int64_t* cap_off,
                               int64_t npts, const double* x, const double* y,
                               const double* z, const double* fx, const double* fy,
                               const double* fz, const double* area, double* forcex,
                               double* forcey, double* forcez);

int ibx_spread_cc_f64_dev(ibx_ctx* ctx, int64_t ncap, const int64_t* cap_off, int64_t npts,
                          const double* x, const double* y, const double* z,
                          const double* xn, const double* yn, const double* zn,
                          const double* area, double* gx, double* gy, double* gz);

/* ---- Lagrangian advance: Update_Pos.f90:77-186 (one rank, every capsule owned) ----------- */
/* Per capsule: unwrap xfpold/yfpold by +-lx/+-ly (:80-92), explicit Euler x += u*dt (:95-99),
 * z clamped to [0,lz] (:118-126), centroid and mean point velocity (:132-155), whole-capsule
 * periodic shift when the centroid leaves [0,lx] / [0,ly] (:158-183).
 * centroid[3*ncap] (x,y,z per capsule) and umean[3*ncap] are outputs; xold/yold may be NULL
 * (then the unwrap step is skipped).  Returns in *nclamp the number of clamped points. */
int ibx_update_pos(ibx_ctx* ctx, int64_t ncap, const int64_t* cap_off, double dt, double* x,
                   double* y, double* z, double* xold, double* yold, const double* ul,
                   const double* vl, const double* wl, double* centroid, double* umean,
                   int64_t* nclamp);
int ibx_update_pos_dev(ibx_ctx* ctx, int64_t ncap, const int64_t* cap_off, int64_t npts,
                       double dt, double* x, double* y, double* z, double* xold, double* yold,
                       const double* ul, const double* vl, const double* wl,
                       double* centroid, double* umean);

/* ---- batched, graph-captured path (thousands of uneven capsules) ------------------------ */
/* A batch binds fixed device buffers, chooses tile/launch shapes (optionally by timing
 * every candidate on the bound data), and captures interp + spread as one hipGraph that
 * ibx_batch_run() replays.  No reference counterpart: the reference runs the same two
 * routines eagerly once per step (main.f90:225,233). */
typedef struct ibx_batch ibx_batch;

typedef struct ibx_batch_bufs {
  const double *u, *v, *w;        /* velocity fields (device) */
  const double *x, *y, *z;        /* positions (device) */
  const double *fx, *fy, *fz;     /* traction (device) */
  const double* area;             /* point areas (device) */
  double *ul, *vl, *wl;           /* interpolated velocity (device, out) */
  double *forcex, *forcey, *forcez; /* spread force field (device, out) */
} ibx_batch_bufs;

int ibx_batch_create(ibx_ctx* ctx, ibx_batch** out, int64_t ncap, const int64_t* cap_off_host);
int ibx_batch_bind(ibx_batch* b, const ibx_batch_bufs* bufs);
/* Time every (tile shape, launch mode) candidate on the bound data and keep the fastest.
 * cache_path (may be NULL) is a text file of previous choices keyed by grid and point count.
 * Writes the chosen configuration id into *choice (may be NULL). */
int ibx_batch_autotune(ibx_batch* b, const char* cache_path, int* choice);
int ibx_batch_set_config(ibx_batch* b, int choice);
/* on != 0: run the interpolation on a second stream of the context, concurrently with the
 * spread's set-up (they share only read-only inputs); the spread's main kernel starts once
 * both are done.  The captured graph then has two branches. */
int ibx_batch_set_overlap(ibx_batch* b, int on);
int ibx_batch_num_configs(void);
const char* ibx_batch_config_name(int choice);
/* Capture interp + spread on the context stream into a graph (use_graph = 1) or run eagerly
 * (use_graph = 0) on every ibx_batch_run.  Several batches may live on one context: when a
 * later capture or call grows (frees and reallocates) a context work buffer, the graphs
 * captured before it are recaptured by their next ibx_batch_run, whose eager sizing pass
 * is then that run's work. */
int ibx_batch_capture(ibx_batch* b, int use_graph);
/* Timed runs: `nslots` sets of HIP events on the context stream (start, after the
 * interpolation, before and after the tile spread kernel, end); ibx_batch_run then enqueues
 * the step eagerly, recording run n into slot n % nslots, so that nslots consecutive runs
 * leave one timing record each.  (A graph's event-record nodes cannot be read back as
 * timings, so timed runs are not graph replays.) */
#define IBX_TIMED_EV 5
int ibx_batch_capture_timed(ibx_batch* b, int nslots);
int ibx_batch_run(ibx_batch* b);
/* Wait for the stream and restart the run count at slot 0. */
int ibx_batch_reset_slots(ibx_batch* b);
/* Wait for the stream; us[4*s + i] receives slot s's last replay, in microseconds:
 * [0] interpolation, [1] spread set-up (tile configurations: binning + scan + wall planes;
 * object configurations: bitmap clear, object boxes and items, sources), [2] the spread's
 * main kernel (tile kernel, or the persistent item / zero-unit kernel), [3] whole step. */
int ibx_batch_slot_times(ibx_batch* b, double* us);
/* Run interp + spread `reps` times eagerly with HIP events on the context stream around each
 * stage; us[0..IBX_PROF_N) receives, in microseconds: [0] interpolation kernel, [1] spread
 * set-up (as for ibx_batch_slot_times), [2] the spread's main kernel, [3] whole sequence,
 * [4] tile configurations: the tile kernel's tiles with candidates, objects: the main kernel,
 * [5] tile configurations: tiles without candidates, objects: the y-halo copy; and the work
 * counts of the last run: tile configurations [6] non-empty tiles, [7] tiles, [8] candidate
 * entries; object configurations [6] work items, [7] extra source objects (sum over items of
 * their source objects minus one), [8] 0. */
#define IBX_PROF_N 9
int ibx_batch_profile(ibx_batch* b, int reps, double* us);
/* Also done by ibx_destroy for every batch of the context that is still alive; a batch
 * must not be used or destroyed after its context. */
void ibx_batch_destroy(ibx_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* IBX_H */
