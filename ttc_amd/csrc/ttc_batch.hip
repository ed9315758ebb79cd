// Batched grouped transpositions: many independent ops (uneven rank, size,
// dtype, perm, alpha/beta) in ONE persistent launch driven by a device-side
// work table, optionally replayed through a captured hipGraph.  The reference
// has no counterpart (one transposition per host call,
// ttc/CUDAtranspose.py:395-429).
//
// Work table: one record per op (its tile parameters, pointers, scalars and a
// kernel "kind"), the ops' tiles numbered consecutively, and a split of that
// global tile range into one contiguous slice per work-group, balanced by
// bytes moved.  A work-group walks its slice; when the slice crosses into the
// next op it re-derives its per-thread tile offsets for that op.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "ttc_launch.hpp"
#include "ttc_plan.h"
#include "ttc_runtime.h"

namespace ttc {

struct BatchOpDev {
  TileParams p;
  const void* A;
  void* B;
  double alpha, beta;
  int32_t kind;
  int32_t pad_;
  uint32_t tile_begin;
  uint32_t ntiles;
};

// Kinds handled by the grouped kernel: same-type pairs, vector width 1 or the
// widest on each side, beta read or not (non-temporal stores when not).
// kind = ((pair * 2 + wide_a) * 2 + wide_b) * 2 + beta_is_zero
enum BatchPair : int { kBS = 0, kBD = 1, kBBF16 = 2, kBF16 = 3, kBC = 4, kBZ = 5, kBPairs = 6 };

template <class OP, int VA, int VB, bool BETA>
__device__ __forceinline__ void run_op_range(const BatchOpDev& o, uint32_t lo, uint32_t hi, unsigned char* lds_raw) {
  TileWorker<OP, VA, VB, BETA, !BETA> w;
  w.setup(o.p, threadIdx.x);
  float amax = 0.0f;
  using L = typename OP::L;
  const typename OP::TA* A = static_cast<const typename OP::TA*>(o.A);
  typename OP::TB* B = static_cast<typename OP::TB*>(o.B);
  const typename OP::AL al = (typename OP::AL)o.alpha;
  const typename OP::BE be = (typename OP::BE)o.beta;
  for (uint32_t t = lo; t < hi; ++t) w.tile(o.p, A, B, reinterpret_cast<L*>(lds_raw), t, al, be, 1.0f, amax);
}

template <class OP, int WA, int WB>
__device__ __forceinline__ void run_op_beta(const BatchOpDev& o, uint32_t lo, uint32_t hi, int bz, unsigned char* l) {
  if (bz) run_op_range<OP, WA, WB, false>(o, lo, hi, l);
  else run_op_range<OP, WA, WB, true>(o, lo, hi, l);
}

template <class OP, int WIDE_A, int WIDE_B>
__device__ __forceinline__ void run_op_pair(const BatchOpDev& o, uint32_t lo, uint32_t hi, int sub,
                                            unsigned char* l) {
  const int bz = sub & 1, wb = (sub >> 1) & 1, wa = (sub >> 2) & 1;
  if (wa && wb) run_op_beta<OP, WIDE_A, WIDE_B>(o, lo, hi, bz, l);
  else if (wa) run_op_beta<OP, WIDE_A, 1>(o, lo, hi, bz, l);
  else if (wb) run_op_beta<OP, 1, WIDE_B>(o, lo, hi, bz, l);
  else run_op_beta<OP, 1, 1>(o, lo, hi, bz, l);
}

__global__ void __launch_bounds__(kThreads) batch_kernel(const BatchOpDev* __restrict__ ops, int nops,
                                                         const uint32_t* __restrict__ wg_begin) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const uint32_t t0 = wg_begin[blockIdx.x], t1 = wg_begin[blockIdx.x + 1];
  if (t0 >= t1) return;
  int lo = 0, hi = nops - 1;  // last op with tile_begin <= t0
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ops[mid].tile_begin <= t0) lo = mid;
    else hi = mid - 1;
  }
  uint32_t t = t0;
  for (int op = lo; op < nops && t < t1; ++op) {
    const BatchOpDev& o = ops[op];
    const uint32_t end = o.tile_begin + o.ntiles;
    const uint32_t stop = end < t1 ? end : t1;
    if (t < stop) {
      const uint32_t a = t - o.tile_begin, b = stop - o.tile_begin;
      const int pair = o.kind >> 3, sub = o.kind & 7;
      switch (pair) {
        case kBS: run_op_pair<OpF32<float, float>, 4, 4>(o, a, b, sub, lds_raw); break;
        case kBD: run_op_pair<OpDD, 2, 2>(o, a, b, sub, lds_raw); break;
        case kBBF16: run_op_pair<OpF32<bf16s, bf16s>, 8, 8>(o, a, b, sub, lds_raw); break;
        case kBF16: run_op_pair<OpF32<f16s, f16s>, 8, 8>(o, a, b, sub, lds_raw); break;
        case kBC: run_op_pair<OpCC, 2, 2>(o, a, b, sub, lds_raw); break;
        default: run_op_pair<OpZZ, 1, 1>(o, a, b, sub, lds_raw); break;
      }
      t = stop;
    }
  }
}

static int batch_pair(int ta, int tb) {
  if (ta != tb) return -1;
  switch (ta) {
    case TTC_S: return kBS;
    case TTC_D: return kBD;
    case TTC_BF16: return kBBF16;
    case TTC_F16: return kBF16;
    case TTC_C: return kBC;
    case TTC_Z: return kBZ;
    default: return -1;
  }
}

}  // namespace ttc

using namespace ttc;

struct ttc_batch_s {
  int nops;
  std::vector<ttc_plan_t> plans;      // per-op plans (also used by the per-op loop)
  std::vector<ttc_batch_op_t> ops;
  BatchOpDev* d_ops;
  uint32_t* d_wg;
  int nwg;
  int lds_bytes;
  int64_t tiles, bytes;
  hipGraph_t graph[2];
  hipGraphExec_t exec[2];
};

static void batch_free(ttc_batch_s* b) {
  for (ttc_plan_t p : b->plans)
    if (p) ttc_plan_destroy(p);
  if (b->d_ops) (void)hipFree(b->d_ops);
  if (b->d_wg) (void)hipFree(b->d_wg);
  for (int i = 0; i < 2; ++i) {
    if (b->exec[i]) (void)hipGraphExecDestroy(b->exec[i]);
    if (b->graph[i]) (void)hipGraphDestroy(b->graph[i]);
  }
  delete b;
}

extern "C" ttc_status_t ttc_batch_create(ttc_batch_t* out, int nops, const ttc_batch_op_t* ops, int wg_per_cu) {
  if (!out || nops < 1 || !ops || wg_per_cu < 0) return TTC_ERROR_INVALID_VALUE;
  *out = nullptr;
  ttc_batch_s* b = new (std::nothrow) ttc_batch_s();
  if (!b) return TTC_ERROR_ALLOC;
  b->nops = nops;
  b->ops.assign(ops, ops + nops);
  b->plans.assign(nops, nullptr);
  std::vector<BatchOpDev> h(nops);
  std::vector<double> weight(nops);
  uint64_t tiles = 0;
  int lds_max = 16;
  b->bytes = 0;
  for (int i = 0; i < nops; ++i) {
    const ttc_batch_op_t& op = ops[i];
    const int pair = batch_pair(op.type_a, op.type_b);
    if (pair < 0) {
      batch_free(b);
      return TTC_ERROR_UNSUPPORTED;
    }
    if (op.dim < 1 || op.dim > TTC_MAX_INPUT_DIM) {
      batch_free(b);
      return TTC_ERROR_INVALID_VALUE;
    }
    bool any_lda = false, any_ldb = false;
    for (int k = 0; k < op.dim; ++k) {
      any_lda |= op.lda[k] != 0;
      any_ldb |= op.ldb[k] != 0;
    }
    // vector widths: 1 or the widest of the type (the kinds the grouped
    // kernel instantiates)
    ttc_config_t cfg = op.cfg;
    const int wide = std::max(1, 16 / dtype_bytes(op.type_a));
    for (int pass = 0; pass < 2; ++pass) {
      ttc_plan_t pl = nullptr;
      ttc_status_t st = ttc_plan_create(&pl, op.type_a, op.type_b, op.dim, op.perm, op.size,
                                        any_lda ? op.lda : nullptr, any_ldb ? op.ldb : nullptr,
                                        op.beta_is_zero, &cfg);
      if (st != TTC_SUCCESS) {
        batch_free(b);
        return st;
      }
      const bool ok_a = pl->va == 1 || pl->va == wide;
      const bool ok_b = pl->vb == 1 || pl->vb == wide;
      if ((ok_a && ok_b) || pass == 1) {
        b->plans[i] = pl;
        break;
      }
      ttc_plan_destroy(pl);
      if (!ok_a) cfg.vec_a = 1;
      if (!ok_b) cfg.vec_b = 1;
      for (int k = 0; k < TTC_MAX_DIM; ++k) cfg.tile[k] = 0;  // re-derive the tile for the new widths
    }
    const ttc_plan_s* pl = b->plans[i];
    if ((pl->va != 1 && pl->va != wide) || (pl->vb != 1 && pl->vb != wide)) {
      batch_free(b);
      return TTC_ERROR_UNSUPPORTED;
    }
    const ttc_plan_s* use = variant_for(pl, op.A, op.B);
    if (!use || (pl->g.elements > 0 && (!op.A || !op.B))) {
      batch_free(b);
      return TTC_ERROR_INVALID_VALUE;
    }
    BatchOpDev& d = h[i];
    std::memset(&d, 0, sizeof(d));
    d.p = use->p;
    d.A = op.A;
    d.B = op.B;
    d.alpha = op.alpha;
    d.beta = op.beta;
    const int wa = use->va > 1 ? 1 : 0, wb = use->vb > 1 ? 1 : 0;
    d.kind = ((pair * 2 + wa) * 2 + wb) * 2 + (pl->beta_is_zero ? 1 : 0);
    d.tile_begin = (uint32_t)tiles;
    d.ntiles = pl->g.elements > 0 ? use->p.ntiles : 0;
    tiles += d.ntiles;
    if (tiles >= (1ull << 31)) {
      batch_free(b);
      return TTC_ERROR_UNSUPPORTED;
    }
    lds_max = std::max(lds_max, use->lds_bytes);
    const int64_t sa = dtype_bytes(op.type_a), sb = dtype_bytes(op.type_b);
    const int64_t bytes = pl->g.elements * (sa + sb * (pl->beta_is_zero ? 1 : 2));
    b->bytes += bytes;
    weight[i] = d.ntiles ? (double)bytes / d.ntiles : 0.0;
  }
  b->tiles = (int64_t)tiles;
  b->lds_bytes = lds_max;
  // work-groups: persistent, as many as fit on the device
  int dev = 0;
  (void)hipGetDevice(&dev);
  const int cus = device_cus(dev);
  int wg = wg_per_cu > 0 ? wg_per_cu : occupancy(reinterpret_cast<const void*>(&batch_kernel), lds_max);
  if (wg < 1) wg = 1;
  int nwg = (int)std::min<uint64_t>((uint64_t)cus * wg, std::max<uint64_t>(tiles, 1));
  b->nwg = nwg;
  // slice the global tile range by bytes
  std::vector<uint32_t> wgb(nwg + 1, (uint32_t)tiles);
  {
    double total = 0;
    for (int i = 0; i < nops; ++i) total += weight[i] * h[i].ntiles;
    const double per = total / nwg;
    int w = 1;
    double acc = 0;
    wgb[0] = 0;
    for (int i = 0; i < nops && w < nwg; ++i) {
      const double wt = weight[i];
      const uint32_t n = h[i].ntiles;
      if (n == 0) continue;
      // tiles of op i where the cumulative byte count crosses w * per
      while (w < nwg && acc + wt * n >= per * w) {
        double need = per * w - acc;
        uint32_t k = wt > 0 ? (uint32_t)std::min<double>(n, std::max(0.0, std::ceil(need / wt))) : 0;
        wgb[w] = h[i].tile_begin + k;
        ++w;
      }
      acc += wt * n;
    }
    for (; w < nwg; ++w) wgb[w] = (uint32_t)tiles;
    wgb[nwg] = (uint32_t)tiles;
    for (int i = 1; i <= nwg; ++i) wgb[i] = std::max(wgb[i], wgb[i - 1]);
  }
  if (hipMalloc(&b->d_ops, sizeof(BatchOpDev) * nops) != hipSuccess ||
      hipMalloc(&b->d_wg, sizeof(uint32_t) * (nwg + 1)) != hipSuccess) {
    batch_free(b);
    return TTC_ERROR_ALLOC;
  }
  if (hipMemcpy(b->d_ops, h.data(), sizeof(BatchOpDev) * nops, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_wg, wgb.data(), sizeof(uint32_t) * (nwg + 1), hipMemcpyHostToDevice) != hipSuccess) {
    batch_free(b);
    return TTC_ERROR_HIP;
  }
  *out = b;
  return TTC_SUCCESS;
}

extern "C" ttc_status_t ttc_batch_execute(ttc_batch_t b, ttc_stream_t stream) {
  if (!b) return TTC_ERROR_INVALID_VALUE;
  if (b->tiles == 0) return TTC_SUCCESS;
  const BatchOpDev* ops = b->d_ops;
  int nops = b->nops;
  const uint32_t* wg = b->d_wg;
  void* args[] = {&ops, &nops, &wg};
  hipError_t e = hipLaunchKernel(reinterpret_cast<const void*>(&batch_kernel), dim3(b->nwg), dim3(kThreads), args,
                                 (size_t)b->lds_bytes, static_cast<hipStream_t>(stream));
  return e == hipSuccess ? TTC_SUCCESS : TTC_ERROR_HIP;
}

extern "C" ttc_status_t ttc_batch_execute_loop(ttc_batch_t b, ttc_stream_t stream) {
  if (!b) return TTC_ERROR_INVALID_VALUE;
  for (int i = 0; i < b->nops; ++i) {
    const ttc_batch_op_t& op = b->ops[i];
    ttc_status_t st = launch_plan(b->plans[i], op.A, op.B, op.alpha, op.beta, static_cast<hipStream_t>(stream));
    if (st != TTC_SUCCESS) return st;
  }
  return TTC_SUCCESS;
}

extern "C" ttc_status_t ttc_batch_graph_launch(ttc_batch_t b, ttc_stream_t stream, int per_op_loop) {
  if (!b || !stream) return TTC_ERROR_INVALID_VALUE;
  const int w = per_op_loop ? 1 : 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!b->exec[w]) {
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return TTC_ERROR_HIP;
    ttc_status_t st = w ? ttc_batch_execute_loop(b, stream) : ttc_batch_execute(b, stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (st != TTC_SUCCESS || e != hipSuccess || !g) {
      if (g) (void)hipGraphDestroy(g);
      return st != TTC_SUCCESS ? st : TTC_ERROR_HIP;
    }
    hipGraphExec_t x = nullptr;
    if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(g);
      return TTC_ERROR_HIP;
    }
    b->graph[w] = g;
    b->exec[w] = x;
  }
  return hipGraphLaunch(b->exec[w], s) == hipSuccess ? TTC_SUCCESS : TTC_ERROR_HIP;
}

extern "C" ttc_status_t ttc_batch_get_info(ttc_batch_t b, int64_t* tiles, int* workgroups, int64_t* bytes) {
  if (!b) return TTC_ERROR_INVALID_VALUE;
  if (tiles) *tiles = b->tiles;
  if (workgroups) *workgroups = b->nwg;
  if (bytes) *bytes = b->bytes;
  return TTC_SUCCESS;
}

extern "C" ttc_status_t ttc_batch_destroy(ttc_batch_t b) {
  if (!b) return TTC_ERROR_INVALID_VALUE;
  batch_free(b);
  return TTC_SUCCESS;
}
