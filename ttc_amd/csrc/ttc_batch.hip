// Batched grouped transpositions: many independent ops (uneven rank, size,
// dtype, perm, alpha/beta) driven by a device-side work table, optionally
// replayed through a captured hipGraph.  The reference has no counterpart
// (one transposition per host call, ttc/CUDAtranspose.py:395-429).
//
// Work table: one record per op (its tile parameters, pointers, scalars),
// ops grouped by kernel kind (type pair, vector widths, beta read or not),
// each group's tiles numbered consecutively, and a split of each group's tile
// range into one contiguous slice per work-group, balanced by bytes moved.
// One persistent launch per kind present walks that kind's slices; a
// work-group whose slice crosses into the next op re-derives its per-thread
// tile offsets for that op.  (One kernel for all kinds would carry the
// register needs of its most demanding kind into every op: measured on
// gfx950 it either ran at one wave per SIMD or spilled.)
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
  uint32_t tile_begin;  // first tile of this op in its group
  uint32_t ntiles;
  const TileBase* bases;  // this op's tile records (ntiles of them)
};

// Host version of the tile-origin decomposition (tiles numbered in A order
// over the tile grid), filling one TileBase per tile.
static void fill_bases(const TileParams& p, TileBase* out) {
  for (uint32_t t = 0; t < p.ntiles; ++t) {
    uint32_t rem = t;
    TileBase b{0, 0, 0xffffu, 0xffffu};
    for (int k = 0; k < p.rank; ++k) {
      const uint32_t g = p.gdiv[k].d;
      const uint32_t x = rem % g;
      rem /= g;
      b.a += (int64_t)x * p.tsa[k];
      b.b += (int64_t)x * p.tsb[k];
      const uint32_t left = (uint32_t)p.n[k] - x * (uint32_t)p.t[k];
      if (k == p.split0) b.lim0 = left;
      if (k == p.split1) b.lim1 = left;
    }
    out[t] = b;
  }
}

// Kinds handled by the grouped kernel: same-type pairs, vector width 1 or the
// widest on each side, beta read or not (non-temporal stores when not).
// kind = ((pair * 2 + wide_a) * 2 + wide_b) * 2 + beta_is_zero
enum BatchPair : int { kBS = 0, kBD = 1, kBBF16 = 2, kBF16 = 3, kBC = 4, kBZ = 5, kBPairs = 6 };

__device__ __forceinline__ int32_t uni(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint32_t uni(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)x); }

// The parameters the per-thread setup and the tile loop read, copied out of
// the op record with each value marked wave-uniform, so that they live in
// scalar registers as kernel arguments do in the single-op kernel.
__device__ __forceinline__ void uniform_params(const TileParams& src, TileParams& p) {
  p.rank = uni(src.rank);
  p.V = uni(src.V);
  p.split0 = uni(src.split0);
  p.split1 = uni(src.split1);
  p.tsplit0 = uni(src.tsplit0);
  p.tsplit1 = uni(src.tsplit1);
  p.splitB0 = uni(src.splitB0);
  p.splitB1 = uni(src.splitB1);
  p.lds_stride_b = uni(src.lds_stride_b);
#pragma unroll
  for (int k = 0; k < kMaxDim; ++k) {
    p.t[k] = uni(src.t[k]);
    p.sa[k] = uni(src.sa[k]);
    p.ls[k] = uni(src.ls[k]);
    p.tdiv[k].m = uni(src.tdiv[k].m);
    p.tdiv[k].s = uni(src.tdiv[k].s);
    p.tB[k] = uni(src.tB[k]);
    p.sbB[k] = uni(src.sbB[k]);
    p.lsB[k] = uni(src.lsB[k]);
    p.tdivB[k].m = uni(src.tdivB[k].m);
    p.tdivB[k].s = uni(src.tdivB[k].s);
  }
  const FastDiv* sd[6] = {&src.pA0, &src.pA1, &src.pB0, &src.pB1, &src.ts0, &src.ts1};
  FastDiv* dd[6] = {&p.pA0, &p.pA1, &p.pB0, &p.pB1, &p.ts0, &p.ts1};
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    dd[i]->d = uni(sd[i]->d);
    dd[i]->m = uni(sd[i]->m);
    dd[i]->s = uni(sd[i]->s);
  }
}

template <class OP, int VA, int VB, bool BETA>
__device__ __forceinline__ void run_op_range(const BatchOpDev& o, uint32_t lo, uint32_t hi, unsigned char* lds_raw) {
  TileWorker<OP, VA, VB, BETA, !BETA> w;
  TileParams up;
  uniform_params(o.p, up);
  w.setup(up, threadIdx.x);
  float amax = 0.0f;
  using L = typename OP::L;
  const typename OP::TA* A = static_cast<const typename OP::TA*>(o.A);
  typename OP::TB* B = static_cast<typename OP::TB*>(o.B);
  const typename OP::AL al = (typename OP::AL)o.alpha;
  const typename OP::BE be = (typename OP::BE)o.beta;
  w.run_table(up, A, B, reinterpret_cast<L*>(lds_raw), o.bases, lo, hi, al, be, 1.0f, amax);
}

template <class OP, int VA, int VB, bool BETA>
__global__ void __launch_bounds__(kThreads) batch_group_kernel(const BatchOpDev* __restrict__ ops, int nops,
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
      run_op_range<OP, VA, VB, BETA>(o, t - o.tile_begin, stop - o.tile_begin, lds_raw);
      t = stop;
    }
  }
}

// The kinds: kind = ((pair * 2 + wide_a) * 2 + wide_b) * 2 + beta_is_zero.
template <class OP, int WA, int WB>
static void add_pair(const void** tab, int pair) {
  const void* k[8] = {
      reinterpret_cast<const void*>(&batch_group_kernel<OP, 1, 1, true>),
      reinterpret_cast<const void*>(&batch_group_kernel<OP, 1, 1, false>),
      reinterpret_cast<const void*>(&batch_group_kernel<OP, 1, WB, true>),
      reinterpret_cast<const void*>(&batch_group_kernel<OP, 1, WB, false>),
      reinterpret_cast<const void*>(&batch_group_kernel<OP, WA, 1, true>),
      reinterpret_cast<const void*>(&batch_group_kernel<OP, WA, 1, false>),
      reinterpret_cast<const void*>(&batch_group_kernel<OP, WA, WB, true>),
      reinterpret_cast<const void*>(&batch_group_kernel<OP, WA, WB, false>),
  };
  for (int i = 0; i < 8; ++i) tab[pair * 8 + i] = k[i];
}

struct GroupKernels {
  const void* k[kBPairs * 8];
  GroupKernels() {
    add_pair<OpF32<float, float>, 4, 4>(k, kBS);
    add_pair<OpDD, 2, 2>(k, kBD);
    add_pair<OpF32<bf16s, bf16s>, 8, 8>(k, kBBF16);
    add_pair<OpF32<f16s, f16s>, 8, 8>(k, kBF16);
    add_pair<OpCC, 2, 2>(k, kBC);
    add_pair<OpZZ, 1, 1>(k, kBZ);
  }
};
static const void* group_kernel(int kind) {
  static const GroupKernels tab;  // thread-safe initialisation
  return tab.k[kind];
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

// One launch of the grouped execution: the ops of one kind.
struct BatchGroup {
  int kind;
  int op0, nops;   // slice of the device op table
  int wg0, nwg;    // slice of the device work-group table (nwg + 1 entries)
  int lds_bytes;
};

struct ttc_batch_s {
  int nops;
  std::vector<ttc_plan_t> plans;  // per-op plans (also used by the per-op loop)
  std::vector<ttc_batch_op_t> ops;
  std::vector<BatchGroup> groups;
  BatchOpDev* d_ops;
  uint32_t* d_wg;
  TileBase* d_bases;
  int nwg_total;
  int64_t tiles, bytes;
  hipGraph_t graph[2];
  hipGraphExec_t exec[2];
};

static void batch_free(ttc_batch_s* b) {
  for (ttc_plan_t p : b->plans)
    if (p) ttc_plan_destroy(p);
  if (b->d_ops) (void)hipFree(b->d_ops);
  if (b->d_wg) (void)hipFree(b->d_wg);
  if (b->d_bases) (void)hipFree(b->d_bases);
  for (int i = 0; i < 2; ++i) {
    if (b->exec[i]) (void)hipGraphExecDestroy(b->exec[i]);
    if (b->graph[i]) (void)hipGraphDestroy(b->graph[i]);
  }
  delete b;
}

// Work-group slices of one group's tile range, balanced by bytes: returns
// nwg + 1 boundaries (global tile indices of the group).
static std::vector<uint32_t> split_by_bytes(const std::vector<BatchOpDev>& h, const std::vector<double>& weight,
                                            int op0, int nops, uint32_t tiles, int nwg) {
  std::vector<uint32_t> wgb(nwg + 1, tiles);
  double total = 0;
  for (int i = op0; i < op0 + nops; ++i) total += weight[i] * h[i].ntiles;
  const double per = total / nwg;
  int w = 1;
  double acc = 0;
  wgb[0] = 0;
  for (int i = op0; i < op0 + nops && w < nwg; ++i) {
    const double wt = weight[i];
    const uint32_t n = h[i].ntiles;
    if (n == 0) continue;
    while (w < nwg && acc + wt * n >= per * w) {
      const double need = per * w - acc;
      const uint32_t k = wt > 0 ? (uint32_t)std::min<double>(n, std::max(0.0, std::ceil(need / wt))) : 0;
      wgb[w] = h[i].tile_begin + k;
      ++w;
    }
    acc += wt * n;
  }
  for (; w < nwg; ++w) wgb[w] = tiles;
  wgb[nwg] = tiles;
  for (int i = 1; i <= nwg; ++i) wgb[i] = std::max(wgb[i], wgb[i - 1]);
  return wgb;
}

extern "C" ttc_status_t ttc_batch_create(ttc_batch_t* out, int nops, const ttc_batch_op_t* ops, int wg_per_cu) {
  if (!out || nops < 1 || !ops || wg_per_cu < 0) return TTC_ERROR_INVALID_VALUE;
  *out = nullptr;
  ttc_batch_s* b = new (std::nothrow) ttc_batch_s();
  if (!b) return TTC_ERROR_ALLOC;
  b->nops = nops;
  b->ops.assign(ops, ops + nops);
  b->plans.assign(nops, nullptr);
  std::vector<BatchOpDev> rec(nops);
  std::vector<int> kinds(nops);
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
    // vector widths: 1 or the widest of the type (the kinds instantiated);
    // the grouped kernels run box tiles
    ttc_config_t cfg = op.cfg;
    cfg.family = kFamilyBox;
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
    if (pl->g.elements > 0 && (!op.A || !op.B)) {
      batch_free(b);
      return TTC_ERROR_INVALID_VALUE;
    }
    const ttc_plan_s* use = pl->g.elements > 0 ? variant_for(pl, op.A, op.B) : pl;
    if (!use) {
      batch_free(b);
      return TTC_ERROR_INVALID_VALUE;
    }
    BatchOpDev& d = rec[i];
    std::memset(&d, 0, sizeof(d));
    d.p = use->p;
    d.A = op.A;
    d.B = op.B;
    d.alpha = op.alpha;
    d.beta = op.beta;
    const int wa = use->va > 1 ? 1 : 0, wb = use->vb > 1 ? 1 : 0;
    d.kind = ((pair * 2 + wa) * 2 + wb) * 2 + (pl->beta_is_zero ? 1 : 0);
    d.ntiles = pl->g.elements > 0 ? use->p.ntiles : 0;
    kinds[i] = d.kind;
    const int64_t sa = dtype_bytes(op.type_a), sb = dtype_bytes(op.type_b);
    b->bytes += pl->g.elements * (sa + sb * (pl->beta_is_zero ? 1 : 2));
  }
  // order the table by kind (stable: ops keep their order inside a kind)
  std::vector<int> order(nops);
  for (int i = 0; i < nops; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return kinds[x] < kinds[y]; });
  std::vector<BatchOpDev> h(nops);
  std::vector<double> weight(nops);
  for (int j = 0; j < nops; ++j) h[j] = rec[order[j]];
  int dev = 0;
  (void)hipGetDevice(&dev);
  const int cus = device_cus(dev);
  std::vector<uint32_t> wg_all;
  b->tiles = 0;
  for (int j = 0; j < nops;) {
    int e = j;
    uint64_t tiles = 0;
    int lds = 16;
    while (e < nops && h[e].kind == h[j].kind) {
      h[e].tile_begin = (uint32_t)tiles;
      tiles += h[e].ntiles;
      lds = std::max(lds, h[e].p.lds_elems * lds_elem_bytes(ops[order[e]].type_a, ops[order[e]].type_b));
      const double bytes = (double)(dtype_bytes(ops[order[e]].type_a) +
                                    dtype_bytes(ops[order[e]].type_b) * ((h[e].kind & 1) ? 1 : 2));
      weight[e] = h[e].ntiles ? bytes * (double)h[e].p.V : 0.0;
      ++e;
    }
    if (tiles >= (1ull << 31)) {
      batch_free(b);
      return TTC_ERROR_UNSUPPORTED;
    }
    lds = (lds + 15) / 16 * 16;
    if (tiles > 0) {
      BatchGroup grp;
      grp.kind = h[j].kind;
      grp.op0 = j;
      grp.nops = e - j;
      grp.lds_bytes = lds;
      const void* fn = group_kernel(grp.kind);
      const int occ = occupancy(fn, lds);
      int wg = wg_per_cu > 0 ? wg_per_cu : std::min(occ, 4);
      if (wg < 1) wg = 1;
      grp.nwg = (int)std::min<uint64_t>((uint64_t)cus * wg, tiles);
      grp.wg0 = (int)wg_all.size();
      std::vector<uint32_t> wgb = split_by_bytes(h, weight, j, e - j, (uint32_t)tiles, grp.nwg);
      wg_all.insert(wg_all.end(), wgb.begin(), wgb.end());
      b->groups.push_back(grp);
      b->tiles += (int64_t)tiles;
    }
    j = e;
  }
  b->nwg_total = 0;
  for (const BatchGroup& g : b->groups) b->nwg_total += g.nwg;
  // per-tile records of every op, in table order
  std::vector<TileBase> bases((size_t)std::max<int64_t>(b->tiles, 1));
  std::vector<size_t> base_off(nops);
  {
    size_t off = 0;
    for (int j = 0; j < nops; ++j) {
      base_off[j] = off;
      if (h[j].ntiles) fill_bases(h[j].p, bases.data() + off);
      off += h[j].ntiles;
    }
  }
  if (hipMalloc(&b->d_ops, sizeof(BatchOpDev) * nops) != hipSuccess ||
      hipMalloc(&b->d_wg, sizeof(uint32_t) * std::max<size_t>(wg_all.size(), 1)) != hipSuccess ||
      hipMalloc(&b->d_bases, sizeof(TileBase) * bases.size()) != hipSuccess) {
    batch_free(b);
    return TTC_ERROR_ALLOC;
  }
  for (int j = 0; j < nops; ++j) h[j].bases = b->d_bases + base_off[j];
  if (hipMemcpy(b->d_ops, h.data(), sizeof(BatchOpDev) * nops, hipMemcpyHostToDevice) != hipSuccess ||
      (!wg_all.empty() && hipMemcpy(b->d_wg, wg_all.data(), sizeof(uint32_t) * wg_all.size(),
                                    hipMemcpyHostToDevice) != hipSuccess) ||
      hipMemcpy(b->d_bases, bases.data(), sizeof(TileBase) * bases.size(), hipMemcpyHostToDevice) != hipSuccess) {
    batch_free(b);
    return TTC_ERROR_HIP;
  }
  *out = b;
  return TTC_SUCCESS;
}

extern "C" ttc_status_t ttc_batch_execute(ttc_batch_t b, ttc_stream_t stream) {
  if (!b) return TTC_ERROR_INVALID_VALUE;
  for (const BatchGroup& g : b->groups) {
    const BatchOpDev* ops = b->d_ops + g.op0;
    int nops = g.nops;
    const uint32_t* wg = b->d_wg + g.wg0;
    void* args[] = {&ops, &nops, &wg};
    hipError_t e = hipLaunchKernel(group_kernel(g.kind), dim3(g.nwg), dim3(kThreads), args, (size_t)g.lds_bytes,
                                   static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return TTC_ERROR_HIP;
  }
  return TTC_SUCCESS;
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
  if (workgroups) *workgroups = b->nwg_total;
  if (bytes) *bytes = b->bytes;
  return TTC_SUCCESS;
}

extern "C" ttc_status_t ttc_batch_destroy(ttc_batch_t b) {
  if (!b) return TTC_ERROR_INVALID_VALUE;
  batch_free(b);
  return TTC_SUCCESS;
}
