// ibx.hip -- host runtime and C ABI of libibx (declarations and reference citations in
// include/ibx.h).  One context per grid and device; all device work is stream-ordered on the
// context stream.  The host-array entry points stage through context-owned HBM buffers;
// the *_dev entry points take HBM pointers and return without waiting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/ibx.h"
#include "ibx_device.hpp"

#pragma clang fp contract(off)

using namespace ibx;

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

// ---- tile configurations (compile-time instances; the batch autotuner picks one) --------
using SpreadFn = void (*)(GridD, TileGeom, const unsigned*, int*, int, unsigned long long*,
                          unsigned long long, const double*, const double*, const double*,
                          const double*, const double*, const double*, const double*, double*,
                          double*, double*);

struct TileCfg {
  int tj, tk;     // rows per tile (each row: one aligned 16-double chunk)
  int wg_per_cu;  // workgroups launched per compute unit
  const char* name;
  SpreadFn plain;
  SpreadFn cc;
  bool obj;       // spread by capsule objects instead of tiles
  int minw;       // objects: minimum wavefronts per SIMD of the register allocation
};

#define IBX_CFG(TJ, TK, W, NAME) \
  {TJ, TK, W, NAME, &k_spread<TJ, TK, false>, &k_spread<TJ, TK, true>, false, 0}
#define IBX_OBJ(W, MINW, NAME) {0, 0, W, NAME, nullptr, nullptr, true, MINW}

constexpr TileCfg kCfgs[] = {
    IBX_OBJ(4, 1, "objects-w4"),
    IBX_OBJ(8, 1, "objects-w8"),
    IBX_OBJ(4, 4, "objects-w4-o4"),
    IBX_OBJ(8, 4, "objects-w8-o4"),
    IBX_CFG(8, 8, 8, "chunk16x8x8"),
    IBX_CFG(8, 4, 8, "chunk16x8x4"),
    IBX_CFG(4, 8, 8, "chunk16x4x8"),
    IBX_CFG(16, 4, 8, "chunk16x16x4"),
    IBX_CFG(8, 8, 4, "chunk16x8x8-w4"),
    IBX_CFG(8, 4, 12, "chunk16x8x4-w12"),
    IBX_CFG(8, 4, 16, "chunk16x8x4-w16"),
    IBX_CFG(4, 16, 8, "chunk16x4x16"),
};
constexpr int kNumCfgs = (int)(sizeof(kCfgs) / sizeof(kCfgs[0]));
constexpr int kTileFallback = 5;  // chunk16x8x4: the tile spread used when objects overflow
static_assert(!kCfgs[kTileFallback].obj, "the fallback must be a tile configuration");
// the one-call entry points' configuration (a batch autotunes its own): objects-w8-o4, the
// fastest object configuration on the 5 % workload and faster than objects-w4 at 30 %
constexpr int kDefaultCfg = 3;
static_assert(kCfgs[kDefaultCfg].obj, "the default must be an object configuration");

}  // namespace

struct ibx_ctx {
  ibx_grid grid{};
  GridD gd{};
  int device = 0;
  int num_cus = 256;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;           // second stream (default priority) for overlap
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  std::string err;
  int* d_err = nullptr;
  void* scratch = nullptr;  // this device's no-allocation work array (device address)
  // spread binning
  DevBuf tcount, toff, tpart, tlist, tmeta, entries;
  // spread by capsule objects
  DevBuf oboxes, oitems, osrc, onsrc, ocover, ometa, onbm, ocnt, ooff, obinm;
  // staging for the host-array entry points
  DevBuf s_f1, s_f2, s_f3, s_x, s_y, s_z, s_a1, s_a2, s_a3, s_a4, s_o1, s_o2, s_o3, s_off, s_kt;
  int cfg = kDefaultCfg;
  unsigned long long ecap = 0;  // entries capacity (elements)
  std::vector<ibx_batch*> batches;  // live batches, destroyed with the context
  // bumped whenever a work buffer is freed and reallocated: a batch graph captured under an
  // older generation holds freed pointers and is recaptured before its next replay
  unsigned long long gen = 0;
};

struct ibx_batch {
  ibx_ctx* ctx = nullptr;
  long long ncap = 0, npts = 0;
  DevBuf d_off;
  ibx_batch_bufs bufs{};
  bool bound = false;
  int cfg = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool use_graph = false;
  bool captured = false;
  bool overlap = false;  // interpolation on the side stream, concurrent with the spread
  unsigned long long gen = 0;  // context buffer generation the graph was captured under
  // timed slots: IBX_TIMED_EV events per slot; timed run n records into slot n % nslots
  std::vector<hipEvent_t> tev;
  long long runs = 0;
};

namespace {

int fail(ibx_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, call)                                                                 \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail((ctx), IBX_E_HIP,                                                       \
                  std::string(#call) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + ":" + \
                      std::to_string(__LINE__) + ")");                                    \
  } while (0)

int ensure(ibx_ctx* c, DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return IBX_OK;
  if (b.p) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
    ++c->gen;  // captured graphs that reference the old buffer are now stale
  }
  size_t nb = std::max<size_t>(bytes, 256);
  HIPCHK(c, hipMalloc(&b.p, nb));
  b.bytes = nb;
  return IBX_OK;
}

void release(DevBuf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

size_t field_elems(const ibx_grid& g) {
  return (size_t)(g.itot + 2) * (size_t)(g.jtot + 2) * (size_t)(g.ktot + 2);
}

TileGeom geom(const ibx_grid& g, const TileCfg& cf, bool cc) {
  TileGeom t{};
  t.ti = SP_CW, t.tj = cf.tj, t.tk = cf.tk;
  t.kbase = cc ? 0 : 1;
  t.nkp = cc ? g.ktot + 2 : g.ktot;
  t.klo = cc ? 0 : 1;
  t.khi = cc ? g.ktot + 1 : g.ktot;
  t.nti = (g.itot + SP_CW) / SP_CW + 1;  // a row's cells reach position itot+1+15
  t.ntj = (g.jtot + cf.tj - 1) / cf.tj;
  t.ntk = (t.nkp + cf.tk - 1) / cf.tk;
  t.ntiles = (long long)t.nti * t.ntj * t.ntk;
  return t;
}

inline unsigned blocks_for(long long n, int b) { return (unsigned)((n + b - 1) / b); }

int check_offsets(ibx_ctx* c, long long ncap, const int64_t* off) {
  if (ncap < 0) return fail(c, IBX_E_ARG, "ncap < 0");
  if (ncap > 0 && !off) return fail(c, IBX_E_ARG, "cap_off is NULL");
  if (ncap == 0) return IBX_OK;
  if (off[0] != 0) return fail(c, IBX_E_ARG, "cap_off[0] != 0");
  for (long long p = 0; p < ncap; ++p)
    if (off[p + 1] < off[p]) return fail(c, IBX_E_ARG, "cap_off is decreasing");
  if (off[ncap] >= (1ll << 28)) return fail(c, IBX_E_ARG, "more than 2^28 - 1 points");
  return IBX_OK;
}

// exclusive scan counts[0..n) -> off[0..n]; non-empty tiles -> tlist, their number -> tmeta[0]
int enqueue_scan(ibx_ctx* c, const unsigned* counts, long long n, unsigned* off) {
  long long nb = (n + SCAN_B - 1) / SCAN_B;
  int rc = ensure(c, c->tpart, (size_t)std::max<long long>(nb, 1) * sizeof(unsigned long long));
  if (rc) return rc;
  if ((rc = ensure(c, c->tlist, (size_t)n * sizeof(int)))) return rc;
  if ((rc = ensure(c, c->tmeta, 9 * SP_QSTRIDE * sizeof(int)))) return rc;
  unsigned long long* part = (unsigned long long*)c->tpart.p;
  k_scan_reduce<<<(unsigned)nb, 256, 0, c->stream>>>(counts, n, part);
  k_scan_parts<<<1, 256, 0, c->stream>>>(part, nb);
  k_scan_apply<<<(unsigned)nb, 256, 0, c->stream>>>(counts, n, part, off, (int*)c->tlist.p,
                                                    (int*)c->tmeta.p);
  HIPCHK(c, hipGetLastError());
  return IBX_OK;
}

// ---- the fixed work array of the no-allocation spread ------------------------------------
// ibx_spread_f64_dev_noalloc carves the object spread's work buffers out of this array
// instead of allocating them.  Each GPU holds its own copy for the life of the process.  In
// host code `scratch` is only a stand-in for the symbol: its device address is taken once per
// device (hipGetSymbolAddress, at the first ibx_create on that device, with the device
// current) and only that address is ever passed to kernels.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

// the device address of `scratch` on `device` (which must be current)
hipError_t scratch_address(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_addr[device] = p;
  }
  *out = g_scratch_addr[device];
  return hipSuccess;
}

// Work buffers of the spread by capsule objects for one call (device pointers).
struct ObjBufs {
  int* boxes;                // 9 objects per capsule, 6 ints each
  unsigned long long* nbm;   // capsule neighbour masks, ncap * nwd words
  int* cnt;                  // items per capsule
  unsigned long long* binm;  // coarse-bin capsule masks, nbins * nwd words
  int* off;                  // item offsets per capsule
  ObjItem* items;            // icap items
  int* src;                  // icap * OB_NSRC source objects
  int* nsrc;                 // icap source counts
  unsigned* cover;           // coverage bitmap, 1 bit per cell
  int* meta;                 // OB_META queue counters
  long long icap;            // item capacity
};

// Sizes (bytes) of the object-spread work buffers for ncap capsules on the context's grid;
// items are sized separately (ObjSizes::per_item bytes per item of capacity).
struct ObjSizes {
  size_t boxes, nbm, cnt, binm, off, cover, meta;
  long long icap_want;  // the allocating path's item capacity
  int nwd;
  long long nw, nbw;
  static constexpr size_t per_item = sizeof(ObjItem) + OB_NSRC * sizeof(int) + sizeof(int);
  size_t fixed() const { return boxes + nbm + cnt + binm + off + cover + meta; }
};

ObjSizes obj_sizes(const ibx_ctx* c, long long ncap) {
  ObjSizes z{};
  const long long ncells = c->gd.sz * (long long)(c->grid.ktot + 2);
  z.nw = (ncells + 31) / 32;
  z.nwd = (int)((ncap + 63) / 64);
  const long long nbins = (long long)((c->grid.itot + OB_BIN - 1) / OB_BIN) *
                          ((c->grid.jtot + OB_BIN - 1) / OB_BIN) *
                          ((c->grid.ktot + 2 + OB_BIN - 1) / OB_BIN);
  z.nbw = nbins * z.nwd;
  z.icap_want = std::max<long long>(1024, 64 * ncap + 1024);
  z.boxes = (size_t)std::max<long long>(1, 9 * ncap) * 6 * sizeof(int);
  z.nbm = (size_t)std::max<long long>(1, ncap * z.nwd) * 8;
  z.cnt = (size_t)std::max<long long>(1, ncap) * sizeof(int);
  z.binm = (size_t)std::max<long long>(1, z.nbw) * 8;
  z.off = (size_t)std::max<long long>(1, ncap) * sizeof(int);
  z.cover = (size_t)z.nw * sizeof(unsigned);
  z.meta = OB_META * sizeof(int);
  return z;
}

// The allocating path: context-owned buffers, grown on demand.
int obj_bufs_alloc(ibx_ctx* c, long long ncap, ObjBufs& ob) {
  const ObjSizes z = obj_sizes(c, ncap);
  const long long icap = z.icap_want;
  int rc;
  if ((rc = ensure(c, c->oboxes, z.boxes)) || (rc = ensure(c, c->onbm, z.nbm)) ||
      (rc = ensure(c, c->ocnt, z.cnt)) || (rc = ensure(c, c->obinm, z.binm)) ||
      (rc = ensure(c, c->ooff, z.off)) ||
      (rc = ensure(c, c->oitems, (size_t)icap * sizeof(ObjItem))) ||
      (rc = ensure(c, c->osrc, (size_t)icap * OB_NSRC * sizeof(int))) ||
      (rc = ensure(c, c->onsrc, (size_t)icap * sizeof(int))) ||
      (rc = ensure(c, c->ocover, z.cover)) || (rc = ensure(c, c->ometa, z.meta)))
    return rc;
  ob.boxes = (int*)c->oboxes.p, ob.nbm = (unsigned long long*)c->onbm.p;
  ob.cnt = (int*)c->ocnt.p, ob.binm = (unsigned long long*)c->obinm.p;
  ob.off = (int*)c->ooff.p, ob.items = (ObjItem*)c->oitems.p;
  ob.src = (int*)c->osrc.p, ob.nsrc = (int*)c->onsrc.p;
  ob.cover = (unsigned*)c->ocover.p, ob.meta = (int*)c->ometa.p;
  ob.icap = icap;
  return IBX_OK;
}

// The no-allocation path: the same buffers carved out of this device's fixed work array, each
// aligned to 256 bytes.  The item capacity is what the array has left (at most the allocating
// path's); a call that needs more items than that reports IBX_E_CAPACITY at the next ibx_sync.
int obj_bufs_scratch(ibx_ctx* c, long long ncap, ObjBufs& ob) {
  if (!c->scratch) return fail(c, IBX_E_STATE, "no work array address for this device");
  const ObjSizes z = obj_sizes(c, ncap);
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t fixed = up(z.boxes) + up(z.nbm) + up(z.cnt) + up(z.binm) + up(z.off) +
                       up(z.cover) + up(z.meta);
  const size_t min_items = 1024;
  if (fixed + 3 * 256 + min_items * ObjSizes::per_item > kScratchBytes)
    return fail(c, IBX_E_ARG,
                "no-allocation spread: " + std::to_string(ncap) + " capsules on this grid need " +
                    std::to_string((fixed + 3 * 256 + min_items * ObjSizes::per_item) >> 20) +
                    " MiB of work buffers, more than the fixed " +
                    std::to_string(kScratchBytes >> 20) + " MiB array");
  const long long fit = (long long)((kScratchBytes - fixed - 3 * 256) / ObjSizes::per_item);
  const long long icap = std::min<long long>(z.icap_want, fit);
  unsigned char* p = (unsigned char*)c->scratch;
  auto take = [&](size_t b) {
    unsigned char* q = p;
    p += up(b);
    return (void*)q;
  };
  ob.boxes = (int*)take(z.boxes), ob.nbm = (unsigned long long*)take(z.nbm);
  ob.cnt = (int*)take(z.cnt), ob.binm = (unsigned long long*)take(z.binm);
  ob.off = (int*)take(z.off), ob.cover = (unsigned*)take(z.cover), ob.meta = (int*)take(z.meta);
  ob.items = (ObjItem*)take((size_t)icap * sizeof(ObjItem));
  ob.src = (int*)take((size_t)icap * OB_NSRC * sizeof(int));
  ob.nsrc = (int*)take((size_t)icap * sizeof(int));
  ob.icap = icap;
  if ((size_t)(p - (unsigned char*)c->scratch) > kScratchBytes)
    return fail(c, IBX_E_ARG, "no-allocation spread: work buffers overrun the fixed array");
  return IBX_OK;
}

// The spread by capsule objects on the context stream: bitmap and counters, object boxes,
// items and bitmap, capsule neighbours, sources, the persistent item / zero-unit kernel,
// y halos (x halos are written with the cells).  ev_tile / ev_tile_end (optional) bracket the persistent
// kernel; ev_split is recorded with ev_tile_end.  Allocates nothing: ob holds the buffers.
int enqueue_spread_obj(ibx_ctx* c, const TileCfg& cf, bool cc, long long ncap,
                       const long long* d_off, long long npts, const double* x,
                       const double* y, const double* z, const double* f1, const double* f2,
                       const double* f3, const double* area, double* o1, double* o2, double* o3,
                       const ObjBufs& ob, hipEvent_t ev_tile, hipEvent_t ev_tile_end,
                       hipEvent_t ev_split, hipEvent_t ev_before) {
  if (npts >= OB_MAXPTS) return fail(c, IBX_E_ARG, "object spread: more than 2^27 points");
  const int klo = cc ? 0 : 1, khi = cc ? c->grid.ktot + 1 : c->grid.ktot;
  const long long ncells = c->gd.sz * (long long)(c->grid.ktot + 2);
  const ObjSizes zs = obj_sizes(c, ncap);
  const long long nw = zs.nw, nbw = zs.nbw, icap = ob.icap;
  const int nwd = zs.nwd;
  int* meta = ob.meta;
  unsigned* cover = ob.cover;
  k_obj_clear<<<blocks_for(std::max<long long>(std::max(nw, nbw), OB_META), 256), 256, 0,
                c->stream>>>(nw, cover, nbw, ob.binm, meta);
  if (npts > 0 && ncap > 0) {
    k_obj_boxes<<<(unsigned)ncap, 256, 0, c->stream>>>(
        c->gd, klo, khi, d_off, x, y, z, ob.boxes, ob.cnt,
        nwd, ob.binm, c->d_err);
    k_obj_scan<<<1, 1024, 0, c->stream>>>(ncap, ob.cnt, ob.off,
                                          (int)icap, meta, c->d_err);
    k_obj_items<<<(unsigned)ncap, 256, 0, c->stream>>>(
        c->gd, ncap, nwd, ob.boxes, ob.binm,
        ob.off, ob.items, (int)icap, cover,
        ob.nbm, c->d_err);
    k_obj_sources<<<(unsigned)(c->num_cus * 4), 256, 0, c->stream>>>(
        nwd, ob.boxes, ob.nbm,
        ob.items, (int)icap, ob.src, ob.nsrc, meta,
        c->d_err);
  }
  // another stream's work (the interpolation) joins here, before the persistent kernel
  // takes every compute unit
  if (ev_before) HIPCHK(c, hipStreamWaitEvent(c->stream, ev_before, 0));
  if (ev_tile) HIPCHK(c, hipEventRecord(ev_tile, c->stream));
  {
    const unsigned nblk = (unsigned)((long long)c->num_cus * cf.wg_per_cu);
#define OBJ_ARGS                                                                             \
  c->gd, klo, khi, d_off, x, y, z, f1, f2, f3, area, ob.items, (int)icap, \
      ob.src, ob.nsrc, (const unsigned*)cover, ncells, meta, o1, \
      o2, o3, c->d_err
    if (cf.minw == 4) {
      if (cc)
        k_obj_spread<true, 4><<<nblk, OB_THREADS, 0, c->stream>>>(OBJ_ARGS);
      else
        k_obj_spread<false, 4><<<nblk, OB_THREADS, 0, c->stream>>>(OBJ_ARGS);
    } else {
      if (cc)
        k_obj_spread<true, 1><<<nblk, OB_THREADS, 0, c->stream>>>(OBJ_ARGS);
      else
        k_obj_spread<false, 1><<<nblk, OB_THREADS, 0, c->stream>>>(OBJ_ARGS);
    }
#undef OBJ_ARGS
  }
  if (ev_tile_end) HIPCHK(c, hipEventRecord(ev_tile_end, c->stream));
  if (ev_split) HIPCHK(c, hipEventRecord(ev_split, c->stream));
  const int kf = cc ? 0 : 1, kl = cc ? c->grid.ktot + 1 : c->grid.ktot;
  const long long nj = (long long)(kl - kf + 1) * 2 * c->gd.sy;
  k_jhalo<<<blocks_for(nj, 256), 256, 0, c->stream>>>(c->gd, kf, kl, o1, o2, o3);
  HIPCHK(c, hipGetLastError());
  return IBX_OK;
}

// The whole spread (binning, scan + compaction, fill, the two-queue tile kernel) on the
// context stream.  ev_tile (optional) is recorded just before the tile kernel.
// size_entries: read the candidate count back and grow the list (eager path); otherwise
// the current capacity is used and overflow is reported through the device error word.
int enqueue_spread(ibx_ctx* c, int cfgid, bool cc, long long ncap, const long long* d_off,
                   long long npts, const double* x, const double* y, const double* z,
                   const double* f1, const double* f2, const double* f3, const double* area,
                   double* o1, double* o2, double* o3, bool size_entries,
                   hipEvent_t ev_tile = nullptr, hipEvent_t ev_split = nullptr,
                   hipEvent_t ev_tile_end = nullptr, hipEvent_t ev_before = nullptr) {
  const TileCfg& cf = kCfgs[cfgid];
  if (cf.obj) {
    ObjBufs ob;
    int rc = obj_bufs_alloc(c, ncap, ob);
    if (rc) return rc;
    return enqueue_spread_obj(c, cf, cc, ncap, d_off, npts, x, y, z, f1, f2, f3, area, o1, o2,
                              o3, ob, ev_tile, ev_tile_end, ev_split, ev_before);
  }
  TileGeom tg = geom(c->grid, cf, cc);
  int rc;
  if ((rc = ensure(c, c->tcount, (size_t)tg.ntiles * sizeof(unsigned)))) return rc;
  if ((rc = ensure(c, c->toff, (size_t)(tg.ntiles + 1) * sizeof(unsigned)))) return rc;
  unsigned* cnt = (unsigned*)c->tcount.p;
  unsigned* off = (unsigned*)c->toff.p;
  if (c->ecap == 0) {
    if ((rc = ensure(c, c->entries, 1024 * sizeof(unsigned long long)))) return rc;
    c->ecap = c->entries.bytes / sizeof(unsigned long long);
  }
  if (npts > 0)
    k_check_forces<<<(unsigned)std::min<long long>(blocks_for(npts, 256), 4LL * c->num_cus), 256,
                     0, c->stream>>>(npts, f1, f2, f3, area, c->d_err);
  k_fill_u32<<<blocks_for(tg.ntiles, 256), 256, 0, c->stream>>>(cnt, tg.ntiles, 0u);
  if (npts > 0)
    k_bin<0><<<blocks_for(npts, 256), 256, 0, c->stream>>>(
        c->gd, tg, ncap, d_off, npts, x, y, z, cnt, off, (unsigned long long*)c->entries.p,
        c->ecap, c->d_err);
  if ((rc = enqueue_scan(c, cnt, tg.ntiles, off))) return rc;
  if (size_entries) {
    unsigned total = 0;
    HIPCHK(c, hipMemcpyAsync(&total, off + tg.ntiles, sizeof(unsigned), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((unsigned long long)total > c->ecap) {
      size_t want = (size_t)total + total / 4 + 1024;
      if ((rc = ensure(c, c->entries, want * sizeof(unsigned long long)))) return rc;
      c->ecap = c->entries.bytes / sizeof(unsigned long long);
    }
  }
  if (npts > 0)
    k_check_forces<<<(unsigned)std::min<long long>(blocks_for(npts, 256), 4LL * c->num_cus), 256,
                     0, c->stream>>>(npts, f1, f2, f3, area, c->d_err);
  k_fill_u32<<<blocks_for(tg.ntiles, 256), 256, 0, c->stream>>>(cnt, tg.ntiles, 0u);
  if (npts > 0)
    k_bin<1><<<blocks_for(npts, 256), 256, 0, c->stream>>>(
        c->gd, tg, ncap, d_off, npts, x, y, z, cnt, off, (unsigned long long*)c->entries.p,
        c->ecap, c->d_err);
  SpreadFn fn = cc ? cf.cc : cf.plain;
  const unsigned nblk = (unsigned)((long long)c->num_cus * cf.wg_per_cu);
  if (!cc) {
    const long long n2 = 2 * c->gd.sz;
    k_zero_kplanes<<<blocks_for((n2 + 1) / 2, 256), 256, 0, c->stream>>>(c->gd, o1, o2, o3);
  }
  if (ev_before) HIPCHK(c, hipStreamWaitEvent(c->stream, ev_before, 0));
  if (ev_tile) HIPCHK(c, hipEventRecord(ev_tile, c->stream));
  if (!ev_split) {
    fn<<<nblk, SP_THREADS, 0, c->stream>>>(c->gd, tg, off, (int*)c->tmeta.p, 3,
                                            (unsigned long long*)c->entries.p, c->ecap, x, y, z,
                                            f1, f2, f3, area, o1, o2, o3);
  } else {
    // the same work as two launches: tiles with candidates, then tiles without (timing)
    fn<<<nblk, SP_THREADS, 0, c->stream>>>(c->gd, tg, off, (int*)c->tmeta.p, 1,
                                            (unsigned long long*)c->entries.p, c->ecap, x, y, z,
                                            f1, f2, f3, area, o1, o2, o3);
    HIPCHK(c, hipEventRecord(ev_split, c->stream));
    for (int q = 1; q <= 8; ++q)
      k_fill_i32<<<1, 64, 0, c->stream>>>((int*)c->tmeta.p + SP_QSTRIDE * q, 1, 0);
    fn<<<nblk, SP_THREADS, 0, c->stream>>>(c->gd, tg, off, (int*)c->tmeta.p, 2,
                                            (unsigned long long*)c->entries.p, c->ecap, x, y, z,
                                            f1, f2, f3, area, o1, o2, o3);
  }
  if (ev_tile_end) HIPCHK(c, hipEventRecord(ev_tile_end, c->stream));
  {
    const int kf = tg.kbase, kl = tg.kbase + tg.nkp - 1;
    const long long nj = (long long)(kl - kf + 1) * 2 * c->gd.sy;
    k_jhalo<<<blocks_for(nj, 256), 256, 0, c->stream>>>(c->gd, kf, kl, o1, o2, o3);
  }
  HIPCHK(c, hipGetLastError());
  return IBX_OK;
}

int enqueue_interp(ibx_ctx* c, const double* u, const double* v, const double* w,
                   long long npts, const double* x, const double* y, const double* z,
                   double* ul, double* vl, double* wl) {
  if (npts > 0)
    k_interp<<<dim3(blocks_for(npts, 256), 3), 256, 0, c->stream>>>(c->gd, u, v, w, npts, x, y,
                                                                      z, ul, vl, wl, c->d_err);
  HIPCHK(c, hipGetLastError());
  return IBX_OK;
}

// read and clear the device error word (synchronises the stream)
int collect_errors(ibx_ctx* c) {
  int h = 0;
  HIPCHK(c, hipMemcpyAsync(&h, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (h) {
    k_fill_i32<<<1, 64, 0, c->stream>>>(c->d_err, 1, 0);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (h & ERR_OFFSETS)
      return fail(c, IBX_E_ARG,
                  "cap_off (device) is not 0 at the start, non-decreasing and npts at the end");
    if (h & ERR_NONFINITE) return fail(c, IBX_E_DOMAIN, "non-finite Lagrangian position");
    if (h & ERR_FORCE) return fail(c, IBX_E_DOMAIN, "non-finite Lagrangian force or area");
    if (h & ERR_CAPACITY)
      return fail(c, IBX_E_CAPACITY,
                  "capacity exceeded (tile candidate list, or object work items / sources per "
                  "item); capture the batch again, or use a tile configuration");
  }
  return IBX_OK;
}

int upload(ibx_ctx* c, DevBuf& b, const void* src, size_t bytes) {
  int rc = ensure(c, b, bytes);
  if (rc) return rc;
  if (bytes) HIPCHK(c, hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, c->stream));
  return IBX_OK;
}

int download(ibx_ctx* c, void* dst, const DevBuf& b, size_t bytes) {
  if (bytes) HIPCHK(c, hipMemcpyAsync(dst, b.p, bytes, hipMemcpyDeviceToHost, c->stream));
  return IBX_OK;
}

}  // namespace

// ======================================================================================
// C ABI
// ======================================================================================
extern "C" {

const char* ibx_version(void) { return "ibx 0.1 (gfx950, FP64, bit-exact reference order)"; }

double ibx_kernel(double r) {
  double ar = fabs(r);
  if (ar > 1.5) return 0.0;
  if (ar > 0.5) {
    double t = 1.0 - ar;
    return (1.0 / 6.0) * ((5.0 - 3.0 * ar) - sqrt(-(3.0 * (t * t)) + 1.0));
  }
  return (1.0 / 3.0) * (1.0 + sqrt(-(3.0 * (r * r)) + 1.0));
}

int ibx_kernel_dev(ibx_ctx* c, int64_t n, const double* r, double* out) {
  if (!c) return IBX_E_ARG;
  if (n < 0) return fail(c, IBX_E_ARG, "negative size");
  if (n > 0 && !(r && out)) return fail(c, IBX_E_ARG, "null pointer");
  if (n == 0) return IBX_OK;
  HIPCHK(c, hipSetDevice(c->device));
  k_phi<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, r, out);
  HIPCHK(c, hipGetLastError());
  return IBX_OK;
}

int ibx_kerneltest_dev(ibx_ctx* c, double* sum) {
  if (!c) return IBX_E_ARG;
  if (!sum) return fail(c, IBX_E_ARG, "null pointer");
  HIPCHK(c, hipSetDevice(c->device));
  int rc = ensure(c, c->s_kt, sizeof(double));
  if (rc) return rc;
  k_kerneltest<<<1, 64, 0, c->stream>>>((double*)c->s_kt.p);
  HIPCHK(c, hipGetLastError());
  if ((rc = download(c, sum, c->s_kt, sizeof(double)))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return IBX_OK;
}

int ibx_collisions_dev(ibx_ctx* c, int64_t npts, const double* z, const double* area,
                       double* fz, double a2, double rad0, double betacol) {
  if (!c) return IBX_E_ARG;
  if (npts < 0) return fail(c, IBX_E_ARG, "negative size");
  if (npts > 0 && !(z && area && fz)) return fail(c, IBX_E_ARG, "null pointer");
  if (!(rad0 >= 0.0) || !std::isfinite(a2) || !std::isfinite(betacol))
    return fail(c, IBX_E_ARG, "rad0 must be >= 0, a2 and betacol finite");
  // the plane sum spans 2 nint(rad0 dyi) + 1 points per direction; keep it a stencil
  if (rad0 * c->grid.dyi > 64.0) return fail(c, IBX_E_ARG, "rad0 spans more than 64 cells");
  if (npts == 0) return IBX_OK;
  HIPCHK(c, hipSetDevice(c->device));
  k_collisions<<<blocks_for(npts, 256), 256, 0, c->stream>>>(c->gd, npts, z, area, fz, c->grid.dx,
                                                             c->grid.dy, a2, rad0, betacol);
  HIPCHK(c, hipGetLastError());
  return IBX_OK;
}

int ibx_shift_z_dev(ibx_ctx* c, int64_t npts, double* z, double s) {
  if (!c) return IBX_E_ARG;
  if (npts < 0) return fail(c, IBX_E_ARG, "negative size");
  if (npts > 0 && !z) return fail(c, IBX_E_ARG, "null pointer");
  if (npts == 0) return IBX_OK;
  HIPCHK(c, hipSetDevice(c->device));
  k_shift_z<<<blocks_for(npts, 256), 256, 0, c->stream>>>(npts, z, s);
  HIPCHK(c, hipGetLastError());
  return IBX_OK;
}

int ibx_create(ibx_ctx** out, const ibx_grid* g, int device) {
  if (!out || !g) return IBX_E_ARG;
  *out = nullptr;
  if (g->itot < 4 || g->jtot < 4 || g->ktot < 4 || g->reserved != 0) return IBX_E_ARG;
  if (!(g->dxi > 0 && g->dyi > 0 && g->dzi > 0 && g->lx > 0 && g->ly > 0 && g->lz > 0))
    return IBX_E_ARG;
  if ((double)(g->itot + 2) * (g->jtot + 2) * (g->ktot + 2) > 4.0e9) return IBX_E_ARG;
  ibx_ctx* c = new ibx_ctx();
  c->grid = *g;
  c->device = device;
  GridD& d = c->gd;
  d.itot = g->itot, d.jtot = g->jtot, d.ktot = g->ktot;
  d.sy = g->itot + 2;
  d.sz = (long long)(g->itot + 2) * (g->jtot + 2);
  d.lx = g->lx, d.ly = g->ly;
  d.dxi = g->dxi, d.dyi = g->dyi, d.dzi = g->dzi;
  d.bl = g->boundleft, d.bf = g->boundfront;
  d.dveul = g->dx * g->dy * g->dz;  // interp_spread.f90:534
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
  if (e == hipSuccess) e = scratch_address(device, &c->scratch);
  if (e == hipSuccess) e = hipMalloc(&c->d_err, sizeof(int));
  if (e == hipSuccess) e = hipMemsetAsync(c->d_err, 0, sizeof(int), c->own);
  if (e == hipSuccess) e = hipStreamSynchronize(c->own);
  if (e == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
      c->num_cus = prop.multiProcessorCount;
  }
  if (e != hipSuccess) {
    // release what was created, in the order of ibx_destroy
    if (c->d_err) (void)hipFree(c->d_err);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->own) (void)hipStreamDestroy(c->own);
    delete c;
    return IBX_E_HIP;
  }
  c->stream = c->own;
  *out = c;
  return IBX_OK;
}

void ibx_destroy(ibx_ctx* c) {
  if (!c) return;
  while (!c->batches.empty()) ibx_batch_destroy(c->batches.back());
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  DevBuf* all[] = {&c->oboxes, &c->oitems, &c->osrc, &c->onsrc, &c->ocover, &c->ometa, &c->onbm, &c->ocnt, &c->ooff, &c->obinm,
                   &c->tcount, &c->toff, &c->tpart, &c->tlist, &c->tmeta, &c->entries, &c->s_f1, &c->s_f2,
                   &c->s_f3,   &c->s_x,  &c->s_y,   &c->s_z,     &c->s_a1, &c->s_a2,
                   &c->s_a3,   &c->s_a4, &c->s_o1,  &c->s_o2,    &c->s_o3, &c->s_off,
                   &c->s_kt};
  for (DevBuf* b : all) release(*b);
  if (c->d_err) (void)hipFree(c->d_err);
  if (c->side) (void)hipStreamSynchronize(c->side);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->own) (void)hipStreamDestroy(c->own);
  delete c;
}

const char* ibx_last_error(const ibx_ctx* c) { return c ? c->err.c_str() : "null context"; }

int ibx_set_stream(ibx_ctx* c, void* s) {
  if (!c) return IBX_E_ARG;
  c->stream = s ? (hipStream_t)s : c->own;
  return IBX_OK;
}

int ibx_set_config(ibx_ctx* c, int choice) {
  if (!c) return IBX_E_ARG;
  if (choice < 0 || choice >= kNumCfgs) return fail(c, IBX_E_ARG, "no such configuration");
  c->cfg = choice;
  return IBX_OK;
}

int ibx_sync(ibx_ctx* c) {
  if (!c) return IBX_E_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  return collect_errors(c);
}

// ---- device-pointer entry points --------------------------------------------------------

int ibx_interp_f64_dev(ibx_ctx* c, const double* u, const double* v, const double* w,
                       int64_t ncap, const int64_t* cap_off, int64_t npts, const double* x,
                       const double* y, const double* z, double* ul, double* vl, double* wl) {
  if (!c) return IBX_E_ARG;
  (void)cap_off;
  if (ncap < 0 || npts < 0) return fail(c, IBX_E_ARG, "negative size");
  if (npts >= (1ll << 28)) return fail(c, IBX_E_ARG, "more than 2^28 - 1 points");
  if (npts > 0 && !(u && v && w && x && y && z && ul && vl && wl))
    return fail(c, IBX_E_ARG, "null pointer");
  HIPCHK(c, hipSetDevice(c->device));
  return enqueue_interp(c, u, v, w, npts, x, y, z, ul, vl, wl);
}

// The spread writes 16-byte pairs of cells; fields allocated by hipMalloc (or torch) are
// 256-byte aligned, and a 128-byte aligned field gets whole-cache-line stores.
static bool aligned16(const double* a, const double* b, const double* c) {
  return (((uintptr_t)a | (uintptr_t)b | (uintptr_t)c) & 15) == 0;
}

static int spread_dev_args(ibx_ctx* c, int64_t ncap, const int64_t* cap_off, int64_t npts,
                           const double* x, const double* y, const double* z, const double* f1,
                           const double* f2, const double* f3, const double* area,
                           const double* o1, const double* o2, const double* o3) {
  if (ncap < 0 || npts < 0) return fail(c, IBX_E_ARG, "negative size");
  if (npts >= (1ll << 28)) return fail(c, IBX_E_ARG, "more than 2^28 - 1 points");
  if (!(o1 && o2 && o3)) return fail(c, IBX_E_ARG, "null output field");
  if (!aligned16(o1, o2, o3)) return fail(c, IBX_E_ARG, "output fields not 16-byte aligned");
  if (npts > 0 && !(cap_off && x && y && z && f1 && f2 && f3 && area && ncap > 0))
    return fail(c, IBX_E_ARG, "null pointer or no capsules");
  return IBX_OK;
}

static int spread_dev(ibx_ctx* c, bool cc, int64_t ncap, const int64_t* cap_off, int64_t npts,
                      const double* x, const double* y, const double* z, const double* f1,
                      const double* f2, const double* f3, const double* area, double* o1,
                      double* o2, double* o3) {
  if (!c) return IBX_E_ARG;
  int rc = spread_dev_args(c, ncap, cap_off, npts, x, y, z, f1, f2, f3, area, o1, o2, o3);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  // the offset table lives in HBM: checked on the device before anything indexes through it
  if (ncap > 0)
    k_check_off<<<1, 256, 0, c->stream>>>(ncap, (const long long*)cap_off, npts, c->d_err);
  return enqueue_spread(c, c->cfg, cc, ncap, (const long long*)cap_off, npts, x, y, z, f1, f2,
                        f3, area, o1, o2, o3, true);
}

int ibx_spread_f64_dev(ibx_ctx* c, int64_t ncap, const int64_t* cap_off, int64_t npts,
                       const double* x, const double* y, const double* z, const double* fx,
                       const double* fy, const double* fz, const double* area, double* forcex,
                       double* forcey, double* forcez) {
  return spread_dev(c, false, ncap, cap_off, npts, x, y, z, fx, fy, fz, area, forcex, forcey,
                    forcez);
}

int ibx_spread_f64_dev_noalloc(ibx_ctx* c, int64_t ncap, const int64_t* cap_off, int64_t npts,
                               const double* x, const double* y, const double* z,
                               const double* fx, const double* fy, const double* fz,
                               const double* area, double* forcex, double* forcey,
                               double* forcez) {
  if (!c) return IBX_E_ARG;
  int rc = spread_dev_args(c, ncap, cap_off, npts, x, y, z, fx, fy, fz, area, forcex, forcey,
                           forcez);
  if (rc) return rc;
  ObjBufs ob;
  if ((rc = obj_bufs_scratch(c, ncap, ob))) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  if (ncap > 0)
    k_check_off<<<1, 256, 0, c->stream>>>(ncap, (const long long*)cap_off, npts, c->d_err);
  // the context's configuration when it is an object one, else the default object one (the
  // tile spread sizes its candidate list per call)
  const int cfg = kCfgs[c->cfg].obj ? c->cfg : kDefaultCfg;
  return enqueue_spread_obj(c, kCfgs[cfg], false, ncap, (const long long*)cap_off, npts, x, y,
                            z, fx, fy, fz, area, forcex, forcey, forcez, ob, nullptr, nullptr,
                            nullptr, nullptr);
}

int ibx_spread_cc_f64_dev(ibx_ctx* c, int64_t ncap, const int64_t* cap_off, int64_t npts,
                          const double* x, const double* y, const double* z, const double* xn,
                          const double* yn, const double* zn, const double* area, double* gx,
                          double* gy, double* gz) {
  return spread_dev(c, true, ncap, cap_off, npts, x, y, z, xn, yn, zn, area, gx, gy, gz);
}

// ---- host-array entry points --------------------------------------------------------------

int ibx_eulr2lagr(ibx_ctx* c, const double* u, const double* v, const double* w, int64_t ncap,
                  const int64_t* cap_off, const double* x, const double* y, const double* z,
                  double* ul, double* vl, double* wl) {
  if (!c) return IBX_E_ARG;
  int rc = check_offsets(c, ncap, cap_off);
  if (rc) return rc;
  long long npts = ncap ? cap_off[ncap] : 0;
  if (!(u && v && w)) return fail(c, IBX_E_ARG, "null field");
  if (npts > 0 && !(x && y && z && ul && vl && wl)) return fail(c, IBX_E_ARG, "null pointer");
  HIPCHK(c, hipSetDevice(c->device));
  size_t fb = field_elems(c->grid) * sizeof(double), pb = (size_t)npts * sizeof(double);
  if ((rc = upload(c, c->s_f1, u, fb)) || (rc = upload(c, c->s_f2, v, fb)) ||
      (rc = upload(c, c->s_f3, w, fb)) || (rc = upload(c, c->s_x, x, pb)) ||
      (rc = upload(c, c->s_y, y, pb)) || (rc = upload(c, c->s_z, z, pb)) ||
      (rc = ensure(c, c->s_a1, pb)) || (rc = ensure(c, c->s_a2, pb)) ||
      (rc = ensure(c, c->s_a3, pb)))
    return rc;
  if ((rc = enqueue_interp(c, (double*)c->s_f1.p, (double*)c->s_f2.p, (double*)c->s_f3.p, npts,
                           (double*)c->s_x.p, (double*)c->s_y.p, (double*)c->s_z.p,
                           (double*)c->s_a1.p, (double*)c->s_a2.p, (double*)c->s_a3.p)))
    return rc;
  if ((rc = download(c, ul, c->s_a1, pb)) || (rc = download(c, vl, c->s_a2, pb)) ||
      (rc = download(c, wl, c->s_a3, pb)))
    return rc;
  return collect_errors(c);
}

static int spread_host(ibx_ctx* c, bool cc, int64_t ncap, const int64_t* cap_off,
                       const double* x, const double* y, const double* z, const double* f1,
                       const double* f2, const double* f3, const double* area, double* o1,
                       double* o2, double* o3) {
  if (!c) return IBX_E_ARG;
  int rc = check_offsets(c, ncap, cap_off);
  if (rc) return rc;
  long long npts = ncap ? cap_off[ncap] : 0;
  if (!(o1 && o2 && o3)) return fail(c, IBX_E_ARG, "null output field");
  if (npts > 0 && !(x && y && z && f1 && f2 && f3 && area))
    return fail(c, IBX_E_ARG, "null pointer");
  HIPCHK(c, hipSetDevice(c->device));
  size_t fb = field_elems(c->grid) * sizeof(double), pb = (size_t)npts * sizeof(double);
  if ((rc = upload(c, c->s_x, x, pb)) || (rc = upload(c, c->s_y, y, pb)) ||
      (rc = upload(c, c->s_z, z, pb)) || (rc = upload(c, c->s_a1, f1, pb)) ||
      (rc = upload(c, c->s_a2, f2, pb)) || (rc = upload(c, c->s_a3, f3, pb)) ||
      (rc = upload(c, c->s_a4, area, pb)) ||
      (rc = upload(c, c->s_off, cap_off, (size_t)(ncap + 1) * sizeof(int64_t))) ||
      (rc = ensure(c, c->s_o1, fb)) || (rc = ensure(c, c->s_o2, fb)) ||
      (rc = ensure(c, c->s_o3, fb)))
    return rc;
  // an object configuration that outgrows its item or source capacity (capsules packed more
  // densely than it plans for) is retried with the tile spread, which sizes its candidate
  // list here; both give the same bits
  for (int cfg = c->cfg;;) {
    if ((rc = enqueue_spread(c, cfg, cc, ncap, (const long long*)c->s_off.p, npts,
                             (double*)c->s_x.p, (double*)c->s_y.p, (double*)c->s_z.p,
                             (double*)c->s_a1.p, (double*)c->s_a2.p, (double*)c->s_a3.p,
                             (double*)c->s_a4.p, (double*)c->s_o1.p, (double*)c->s_o2.p,
                             (double*)c->s_o3.p, true)))
      return rc;
    rc = collect_errors(c);
    if (rc == IBX_E_CAPACITY && kCfgs[cfg].obj) {
      cfg = kTileFallback;
      continue;
    }
    if (rc) return rc;
    break;
  }
  if ((rc = download(c, o1, c->s_o1, fb)) || (rc = download(c, o2, c->s_o2, fb)) ||
      (rc = download(c, o3, c->s_o3, fb)))
    return rc;
  return collect_errors(c);
}

int ibx_lagr2eulr(ibx_ctx* c, int64_t ncap, const int64_t* cap_off, const double* x,
                  const double* y, const double* z, const double* fx, const double* fy,
                  const double* fz, const double* area, double* forcex, double* forcey,
                  double* forcez) {
  return spread_host(c, false, ncap, cap_off, x, y, z, fx, fy, fz, area, forcex, forcey,
                     forcez);
}

int ibx_lagr2eulrind(ibx_ctx* c, int64_t ncap, const int64_t* cap_off, const double* x,
                     const double* y, const double* z, const double* xn, const double* yn,
                     const double* zn, const double* area, double* gx, double* gy,
                     double* gz) {
  return spread_host(c, true, ncap, cap_off, x, y, z, xn, yn, zn, area, gx, gy, gz);
}

// ---- Update_Pos --------------------------------------------------------------------------

int ibx_update_pos_dev(ibx_ctx* c, int64_t ncap, const int64_t* cap_off, int64_t npts,
                       double dt, double* x, double* y, double* z, double* xold, double* yold,
                       const double* ul, const double* vl, const double* wl, double* centroid,
                       double* umean) {
  if (!c) return IBX_E_ARG;
  if (ncap <= 0 || npts < 0) return fail(c, IBX_E_ARG, "bad sizes");
  if (!(cap_off && x && y && z && ul && vl && wl && centroid && umean))
    return fail(c, IBX_E_ARG, "null pointer");
  if ((xold == nullptr) != (yold == nullptr))
    return fail(c, IBX_E_ARG, "xold and yold must both be given or both be NULL");
  HIPCHK(c, hipSetDevice(c->device));
  k_check_off<<<1, 256, 0, c->stream>>>(ncap, (const long long*)cap_off, npts, c->d_err);
  if (npts > 0)
    k_advance<<<blocks_for(npts, 256), 256, 0, c->stream>>>(c->gd, c->grid.lz, dt, npts, x, y,
                                                            z, xold, yold, ul, vl, wl);
  k_centroid<<<(unsigned)ncap, 256, 0, c->stream>>>(c->gd, ncap, (const long long*)cap_off, x,
                                                    y, z, ul, vl, wl, centroid, umean, c->d_err);
  HIPCHK(c, hipGetLastError());
  return IBX_OK;
}

int ibx_update_pos(ibx_ctx* c, int64_t ncap, const int64_t* cap_off, double dt, double* x,
                   double* y, double* z, double* xold, double* yold, const double* ul,
                   const double* vl, const double* wl, double* centroid, double* umean,
                   int64_t* nclamp) {
  if (!c) return IBX_E_ARG;
  int rc = check_offsets(c, ncap, cap_off);
  if (rc) return rc;
  if (ncap == 0) return IBX_OK;
  long long npts = cap_off[ncap];
  if ((xold == nullptr) != (yold == nullptr))
    return fail(c, IBX_E_ARG, "xold and yold must both be given or both be NULL");
  // points below / above the walls before the clamp (Update_Pos.f90:118-126 prints these)
  if (nclamp) {
    long long nc = 0;
    for (long long q = 0; q < npts; ++q) {
      double zn = z[q] + (wl[q] * dt);
      if (zn < 0.0 || zn > c->grid.lz) ++nc;
    }
    *nclamp = nc;
  }
  HIPCHK(c, hipSetDevice(c->device));
  size_t pb = (size_t)npts * sizeof(double), cb = (size_t)ncap * 3 * sizeof(double);
  if ((rc = upload(c, c->s_x, x, pb)) || (rc = upload(c, c->s_y, y, pb)) ||
      (rc = upload(c, c->s_z, z, pb)) || (rc = upload(c, c->s_a1, ul, pb)) ||
      (rc = upload(c, c->s_a2, vl, pb)) || (rc = upload(c, c->s_a3, wl, pb)) ||
      (rc = upload(c, c->s_off, cap_off, (size_t)(ncap + 1) * sizeof(int64_t))) ||
      (rc = ensure(c, c->s_o1, cb)) || (rc = ensure(c, c->s_o2, cb)))
    return rc;
  if (xold) {
    if ((rc = upload(c, c->s_f1, xold, pb)) || (rc = upload(c, c->s_f2, yold, pb))) return rc;
  }
  if ((rc = ibx_update_pos_dev(c, ncap, (const int64_t*)c->s_off.p, npts, dt, (double*)c->s_x.p,
                               (double*)c->s_y.p, (double*)c->s_z.p,
                               xold ? (double*)c->s_f1.p : nullptr,
                               yold ? (double*)c->s_f2.p : nullptr, (double*)c->s_a1.p,
                               (double*)c->s_a2.p, (double*)c->s_a3.p, (double*)c->s_o1.p,
                               (double*)c->s_o2.p)))
    return rc;
  if ((rc = download(c, x, c->s_x, pb)) || (rc = download(c, y, c->s_y, pb)) ||
      (rc = download(c, z, c->s_z, pb)) || (rc = download(c, centroid, c->s_o1, cb)) ||
      (rc = download(c, umean, c->s_o2, cb)))
    return rc;
  if (xold) {
    if ((rc = download(c, xold, c->s_f1, pb)) || (rc = download(c, yold, c->s_f2, pb)))
      return rc;
  }
  return collect_errors(c);
}

// ---- batched / graph path -----------------------------------------------------------------

int ibx_batch_num_configs(void) { return kNumCfgs; }

const char* ibx_batch_config_name(int i) {
  return (i >= 0 && i < kNumCfgs) ? kCfgs[i].name : "invalid";
}

int ibx_batch_create(ibx_ctx* c, ibx_batch** out, int64_t ncap, const int64_t* cap_off_host) {
  if (!c || !out) return IBX_E_ARG;
  *out = nullptr;
  int rc = check_offsets(c, ncap, cap_off_host);
  if (rc) return rc;
  if (ncap <= 0) return fail(c, IBX_E_ARG, "a batch needs at least one capsule");
  HIPCHK(c, hipSetDevice(c->device));
  ibx_batch* b = new ibx_batch();
  b->ctx = c;
  b->ncap = ncap;
  b->npts = cap_off_host[ncap];
  b->cfg = c->cfg;
  if ((rc = upload(c, b->d_off, cap_off_host, (size_t)(ncap + 1) * sizeof(int64_t)))) {
    delete b;
    return rc;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->batches.push_back(b);
  *out = b;
  return IBX_OK;
}

int ibx_batch_bind(ibx_batch* b, const ibx_batch_bufs* bufs) {
  if (!b || !bufs) return IBX_E_ARG;
  const ibx_batch_bufs& q = *bufs;
  if (!(q.u && q.v && q.w && q.x && q.y && q.z && q.fx && q.fy && q.fz && q.area && q.ul &&
        q.vl && q.wl && q.forcex && q.forcey && q.forcez))
    return fail(b->ctx, IBX_E_ARG, "null buffer in ibx_batch_bufs");
  if (!aligned16(q.forcex, q.forcey, q.forcez))
    return fail(b->ctx, IBX_E_ARG, "force fields not 16-byte aligned");
  b->bufs = q;
  b->bound = true;
  b->captured = false;
  return IBX_OK;
}

// ev (optional): IBX_TIMED_EV events recorded at the start, after the interpolation, before
// and after the tile spread kernel, and at the end
static void drop_graph(ibx_batch* b);

// With overlap, the interpolation runs on the side stream, forked from and joined back to
// the context stream (the two stages share only read-only inputs); a captured graph then
// holds two independent branches.
static int batch_enqueue(ibx_batch* b, bool size_entries, const hipEvent_t* ev = nullptr) {
  ibx_ctx* c = b->ctx;
  const ibx_batch_bufs& q = b->bufs;
  const bool ov = b->overlap && !size_entries;
  if (ev) HIPCHK(c, hipEventRecord(ev[0], c->stream));
  int rc;
  if (ov) {
    HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
    hipStream_t keep = c->stream;
    c->stream = c->side;
    rc = enqueue_interp(c, q.u, q.v, q.w, b->npts, q.x, q.y, q.z, q.ul, q.vl, q.wl);
    if (!rc && ev) {
      hipError_t e = hipEventRecord(ev[1], c->side);
      if (e != hipSuccess) rc = fail(c, IBX_E_HIP, "hipEventRecord(side)");
    }
    c->stream = keep;
    if (rc) return rc;
    HIPCHK(c, hipEventRecord(c->ev_join, c->side));
  } else {
    rc = enqueue_interp(c, q.u, q.v, q.w, b->npts, q.x, q.y, q.z, q.ul, q.vl, q.wl);
    if (rc) return rc;
    if (ev) HIPCHK(c, hipEventRecord(ev[1], c->stream));
  }
  rc = enqueue_spread(c, b->cfg, false, b->ncap, (const long long*)b->d_off.p, b->npts, q.x,
                      q.y, q.z, q.fx, q.fy, q.fz, q.area, q.forcex, q.forcey, q.forcez,
                      size_entries, ev ? ev[2] : nullptr, nullptr, ev ? ev[3] : nullptr,
                      ov ? c->ev_join : nullptr);
  if (rc) return rc;
  if (ov) HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));  // (already joined)
  if (ev) HIPCHK(c, hipEventRecord(ev[4], c->stream));
  return IBX_OK;
}

int ibx_batch_set_overlap(ibx_batch* b, int on) {
  if (!b) return IBX_E_ARG;
  b->overlap = on != 0;
  drop_graph(b);
  return IBX_OK;
}

static void drop_graph(ibx_batch* b) {
  for (auto e : b->tev) (void)hipEventDestroy(e);
  b->tev.clear();
  b->runs = 0;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->captured = false;
}

int ibx_batch_set_config(ibx_batch* b, int choice) {
  if (!b || choice < 0 || choice >= kNumCfgs) return IBX_E_ARG;
  b->cfg = choice;
  drop_graph(b);
  return IBX_OK;
}

static int capture_graph(ibx_batch* b, const hipEvent_t* ev, hipGraph_t* graph,
                         hipGraphExec_t* exec) {
  ibx_ctx* c = b->ctx;
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = batch_enqueue(b, false, ev);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(c, IBX_E_HIP, std::string("capture: ") + hipGetErrorString(e));
  *graph = g;
  HIPCHK(c, hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
  return IBX_OK;
}

static int capture_common(ibx_batch* b, int use_graph, int nslots) {
  if (!b) return IBX_E_ARG;
  ibx_ctx* c = b->ctx;
  if (!b->bound) return fail(c, IBX_E_STATE, "batch not bound");
  if (nslots < 0 || nslots > 4096) return fail(c, IBX_E_ARG, "nslots out of range 0..4096");
  HIPCHK(c, hipSetDevice(c->device));
  drop_graph(b);
  b->captured = false;
  // one eager pass sizes every work buffer; the candidate list then gets room for the points
  // to move: 1.5x the measured count, and at least 4 entries per point (a point's window
  // lists it in ~2-3 tiles on average)
  int rc = batch_enqueue(b, true);
  if (rc) return rc;
  if ((rc = collect_errors(c))) return rc;
  if (!kCfgs[b->cfg].obj) {
    TileGeom tg = geom(c->grid, kCfgs[b->cfg], false);
    unsigned total = 0;
    HIPCHK(c, hipMemcpyAsync(&total, (unsigned*)c->toff.p + tg.ntiles, sizeof(unsigned),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    size_t want = std::max<size_t>((size_t)total + total / 2, (size_t)(4 * b->npts)) + 4096;
    if (want > c->ecap) {
      if ((rc = ensure(c, c->entries, want * sizeof(unsigned long long)))) return rc;
      c->ecap = c->entries.bytes / sizeof(unsigned long long);
    }
  }
  b->use_graph = nslots == 0 && use_graph != 0;
  if (nslots > 0) {
    // timing events cannot be read back from a graph's event-record nodes on this HIP, so
    // timed runs are eager launches with events between the stages
    b->tev.assign((size_t)nslots * IBX_TIMED_EV, nullptr);
    for (auto& e : b->tev) HIPCHK(c, hipEventCreate(&e));
  } else if (b->use_graph) {
    if ((rc = capture_graph(b, nullptr, &b->graph, &b->exec))) return rc;
  }
  b->runs = 0;
  b->captured = true;
  b->gen = c->gen;
  return IBX_OK;
}

int ibx_batch_capture(ibx_batch* b, int use_graph) { return capture_common(b, use_graph, 0); }

int ibx_batch_capture_timed(ibx_batch* b, int nslots) {
  if (nslots <= 0) return b ? fail(b->ctx, IBX_E_ARG, "nslots must be positive") : IBX_E_ARG;
  return capture_common(b, 1, nslots);
}

int ibx_batch_run(ibx_batch* b) {
  if (!b) return IBX_E_ARG;
  ibx_ctx* c = b->ctx;
  if (!b->captured) return fail(c, IBX_E_STATE, "batch not captured");
  HIPCHK(c, hipSetDevice(c->device));
  if (!b->tev.empty()) {
    const size_t ns = b->tev.size() / IBX_TIMED_EV;
    const size_t sl = (size_t)(b->runs % (long long)ns);
    ++b->runs;
    return batch_enqueue(b, false, &b->tev[sl * IBX_TIMED_EV]);
  }
  ++b->runs;
  if (b->use_graph) {
    if (b->gen != c->gen) {
      // another batch or call on this context reallocated a work buffer this graph points
      // to: capture again (its eager sizing pass is this run's work)
      const long long runs = b->runs;
      int rc = capture_common(b, 1, 0);
      b->runs = runs;
      return rc;
    }
    HIPCHK(c, hipGraphLaunch(b->exec, c->stream));
    return IBX_OK;
  }
  return batch_enqueue(b, false);
}

int ibx_batch_reset_slots(ibx_batch* b) {
  if (!b) return IBX_E_ARG;
  ibx_ctx* c = b->ctx;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  b->runs = 0;
  return IBX_OK;
}

int ibx_batch_slot_times(ibx_batch* b, double* us) {
  if (!b || !us) return IBX_E_ARG;
  ibx_ctx* c = b->ctx;
  if (b->tev.empty()) return fail(c, IBX_E_STATE, "batch not captured with timed slots");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t ns = b->tev.size() / IBX_TIMED_EV;
  for (size_t sl = 0; sl < ns; ++sl) {
    const hipEvent_t* e = &b->tev[sl * IBX_TIMED_EV];
    float t[4];
    HIPCHK(c, hipEventElapsedTime(&t[0], e[0], e[1]));  // interpolation
    // binning, scan, wall planes: from the step start (with overlap the interpolation runs
    // concurrently on the side stream)
    HIPCHK(c, hipEventElapsedTime(&t[1], b->overlap ? e[0] : e[1], e[2]));
    HIPCHK(c, hipEventElapsedTime(&t[2], e[2], e[3]));  // tile spread kernel
    HIPCHK(c, hipEventElapsedTime(&t[3], e[0], e[4]));  // whole step
    for (int i = 0; i < 4; ++i) us[sl * 4 + i] = 1000.0 * (double)t[i];
  }
  return collect_errors(c);
}

int ibx_batch_profile(ibx_batch* b, int reps, double* us) {
  // HIP events on the context stream bracket each stage of `reps` eager runs.
  if (!b || reps <= 0 || !us) return IBX_E_ARG;
  ibx_ctx* c = b->ctx;
  if (!b->captured) return fail(c, IBX_E_STATE, "batch not captured");
  HIPCHK(c, hipSetDevice(c->device));
  const ibx_batch_bufs& q = b->bufs;
  hipEvent_t ev[6];
  for (auto& e : ev) HIPCHK(c, hipEventCreate(&e));
  double acc[IBX_PROF_N] = {0};
  for (int pass = 0; pass < 2; ++pass) {
    for (int r = 0; r < reps; ++r) {
      HIPCHK(c, hipEventRecord(ev[0], c->stream));
      int rc = enqueue_interp(c, q.u, q.v, q.w, b->npts, q.x, q.y, q.z, q.ul, q.vl, q.wl);
      if (rc) return rc;
      HIPCHK(c, hipEventRecord(ev[1], c->stream));
      rc = enqueue_spread(c, b->cfg, false, b->ncap, (const long long*)b->d_off.p, b->npts, q.x,
                          q.y, q.z, q.fx, q.fy, q.fz, q.area, q.forcex, q.forcey, q.forcez,
                          false, ev[2], pass == 1 ? ev[4] : nullptr);
      if (rc) return rc;
      HIPCHK(c, hipEventRecord(ev[3], c->stream));
      HIPCHK(c, hipEventSynchronize(ev[3]));
      float a, bb, s, t;
      if (pass == 0) {
        HIPCHK(c, hipEventElapsedTime(&a, ev[0], ev[1]));
        HIPCHK(c, hipEventElapsedTime(&bb, ev[1], ev[2]));
        HIPCHK(c, hipEventElapsedTime(&s, ev[2], ev[3]));
        HIPCHK(c, hipEventElapsedTime(&t, ev[0], ev[3]));
        acc[0] += a, acc[1] += bb, acc[2] += s, acc[3] += t;
      } else {
        HIPCHK(c, hipEventElapsedTime(&a, ev[2], ev[4]));
        HIPCHK(c, hipEventElapsedTime(&bb, ev[4], ev[3]));
        acc[4] += a, acc[5] += bb;
      }
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  for (int i = 0; i < 6; ++i) us[i] = 1000.0 * acc[i] / reps;
  // work counts of the last run: tiles: non-empty tiles, tiles, candidate entries;
  // objects: work items, extra source objects (sum over items of sources - 1), 0
  if (kCfgs[b->cfg].obj) {
    int m0 = 0;
    HIPCHK(c, hipMemcpyAsync(&m0, c->ometa.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int ni = (int)std::min<long long>(m0, c->onsrc.bytes / sizeof(int));
    std::vector<int> ns((size_t)std::max(ni, 1));
    if (ni > 0)
      HIPCHK(c, hipMemcpy(ns.data(), c->onsrc.p, (size_t)ni * sizeof(int),
                          hipMemcpyDeviceToHost));
    long long extra = 0;
    for (int i = 0; i < ni; ++i) extra += ns[i] - 1;
    us[6] = (double)m0, us[7] = (double)extra, us[8] = 0.0;
    return collect_errors(c);
  }
  int meta0 = 0;
  unsigned tot = 0;
  TileGeom tg = geom(c->grid, kCfgs[b->cfg], false);
  HIPCHK(c, hipMemcpyAsync(&meta0, c->tmeta.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&tot, (unsigned*)c->toff.p + tg.ntiles, sizeof(unsigned),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  us[6] = (double)meta0;
  us[7] = (double)tg.ntiles;
  us[8] = (double)tot;
  return collect_errors(c);
}

int ibx_batch_autotune(ibx_batch* b, const char* cache_path, int* choice) {
  if (!b) return IBX_E_ARG;
  ibx_ctx* c = b->ctx;
  if (!b->bound) return fail(c, IBX_E_STATE, "batch not bound");
  HIPCHK(c, hipSetDevice(c->device));
  char key[256];
  snprintf(key, sizeof key, "grid=%dx%dx%d npts=%lld ncap=%lld", c->grid.itot, c->grid.jtot,
           c->grid.ktot, b->npts, b->ncap);
  if (cache_path) {
    std::ifstream in(cache_path);
    std::string line;
    while (std::getline(in, line)) {
      size_t bar = line.find(" | ");
      if (bar == std::string::npos) continue;
      if (line.substr(0, bar) == key) {
        std::string name = line.substr(bar + 3);
        for (int i = 0; i < kNumCfgs; ++i)
          if (name == kCfgs[i].name) {
            b->cfg = i;
            drop_graph(b);
            if (choice) *choice = i;
            return IBX_OK;
          }
      }
    }
  }
  int best = -1;
  double best_t = 1e300;
  for (int i = 0; i < kNumCfgs; ++i) {
    b->cfg = i;
    int rc = batch_enqueue(b, true);  // warm-up and sizing
    if (rc) return rc;
    double us[IBX_PROF_N];
    if (!(rc = collect_errors(c))) {
      b->captured = true;
      rc = ibx_batch_profile(b, 5, us);
      b->captured = false;
    }
    // an object configuration whose item or source capacity this batch exceeds is skipped
    if (rc == IBX_E_CAPACITY && kCfgs[i].obj) continue;
    if (rc) return rc;
    if (us[3] < best_t) {
      best_t = us[3];
      best = i;
    }
  }
  if (best < 0) return fail(c, IBX_E_CAPACITY, "no configuration fits this batch");
  b->cfg = best;
  drop_graph(b);
  if (choice) *choice = best;
  if (cache_path) {
    std::ofstream outf(cache_path, std::ios::app);
    if (outf) outf << key << " | " << kCfgs[best].name << "\n";
  }
  return IBX_OK;
}

void ibx_batch_destroy(ibx_batch* b) {
  if (!b) return;
  if (b->ctx) {
    (void)hipSetDevice(b->ctx->device);
    if (b->ctx->stream) (void)hipStreamSynchronize(b->ctx->stream);
    auto& v = b->ctx->batches;
    v.erase(std::remove(v.begin(), v.end(), b), v.end());
  }
  drop_graph(b);
  release(b->d_off);
  delete b;
}

}  // extern "C"
