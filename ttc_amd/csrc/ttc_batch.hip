// Batched grouped transpositions: many independent ops (uneven rank, size,
// dtype, perm, alpha/beta) driven by a device-side work table, optionally
// replayed through a captured hipGraph.  The reference has no counterpart
// (one transposition per host call, ttc/CUDAtranspose.py:395-429).
//
// Work table: one record per op (its tile parameters, pointers, scalars) and
// one record per tile (origin offsets, partial-tile limits), ops grouped by
// kernel family and type pair and ordered by variant (vector widths, beta
// read or not) inside a group, each group's tiles numbered consecutively,
// and a split of each group's tile range into one contiguous slice per
// work-group, balanced by bytes moved.  One persistent launch per group
// walks that group's slices, each op through its variant's walk (the
// kernels are in ttc_batch_kernels.hpp).  Each op takes the family its own
// plan chose: row x column tiles (a work-group rebuilds its LDS offset
// tables when its slice crosses into the next op), row copies (dimension 0
// in place; the unit of work is a chunk of vectors) or box tiles.  One
// launch per group rather than per (group, variant): a graph of many small
// independent launches runs them a few at a time (the hardware queues),
// each with its own tail.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "ttc_batch_kernels.hpp"
#include "ttc_launch.hpp"
#include "ttc_plan.h"
#include "ttc_runtime.h"

namespace ttc {

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

// Kernel of group key = family * kBPairs + pair (family 0 box tiles, 1 row
// x column tiles, 2 row copies).
enum GroupFamily { kGroupBox = 0, kGroupRC = 1, kGroupRow = 2 };
static int group_of(int kind) { return kind / kVariantStride; }
static int group_family(int group) { return group / kBPairs; }

static const BatchKernelSet& pair_kernels(int pair) {
  static const BatchKernelSet tab[kBPairs] = {batch_kernels_s(),   batch_kernels_d(), batch_kernels_bf16(),
                                              batch_kernels_f16(), batch_kernels_c(), batch_kernels_z()};
  return tab[pair];
}
static const void* group_kernel(int group) {
  const BatchKernelSet& k = pair_kernels(group % kBPairs);
  switch (group_family(group)) {
    case kGroupRC: return k.rc;
    case kGroupRow: return k.row;
    default: return k.box;
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

// One launch of the grouped execution: the ops of one kind.
struct BatchGroup {
  int group;       // family * kBPairs + type pair
  int op0, nops;   // slice of the op table of the group's family
  int wg0, nwg;    // slice of the work-group table (nwg + 1 entries)
  int lds_bytes;
};

// Host copy of a batch's device work table.  Its sections (box op records,
// row x column op records, work-group slices, box tile records, row x
// column tile records, offset tables) sit 256-byte aligned in one image, so
// that one allocation (or the fixed workspace) and one copy hold all of it;
// op records point at their tile records once the image's device address is
// known (render).
struct BatchImage {
  enum { kBox, kRC, kRow, kWg, kBases, kRcTiles, kRcTables, kSections };
  std::vector<BatchOpDev> box;
  std::vector<BatchRCOpDev> rcs;
  std::vector<BatchRowOpDev> rows;
  std::vector<uint32_t> wg;
  std::vector<TileBase> bases;
  std::vector<RCTile> rct;
  std::vector<int32_t> rtab;
  std::vector<size_t> box_rec;          // per box op: index of its first tile record
  std::vector<size_t> rc_rec, rc_tab;   // per row x column op: first tile record, offset tables
  size_t off[kSections] = {0};
  size_t bytes = 0;

  void layout() {
    const size_t len[kSections] = {sizeof(BatchOpDev) * box.size(),   sizeof(BatchRCOpDev) * rcs.size(),
                                   sizeof(BatchRowOpDev) * rows.size(), sizeof(uint32_t) * wg.size(),
                                   sizeof(TileBase) * bases.size(),     sizeof(RCTile) * rct.size(),
                                   sizeof(int32_t) * rtab.size()};
    size_t o = 0;
    for (int s = 0; s < kSections; ++s) {
      off[s] = o;
      o += (len[s] + 255) / 256 * 256;
    }
    bytes = o;
  }

  // The image as it must sit at device address `base`.
  void render(uintptr_t base, std::vector<unsigned char>& out) const {
    out.assign(bytes, 0);
    std::vector<BatchOpDev> b2(box);
    for (size_t i = 0; i < b2.size(); ++i)
      b2[i].bases = reinterpret_cast<const TileBase*>(base + off[kBases] + box_rec[i] * sizeof(TileBase));
    std::vector<BatchRCOpDev> r2(rcs);
    for (size_t i = 0; i < r2.size(); ++i) {
      r2[i].tiles = reinterpret_cast<const RCTile*>(base + off[kRcTiles] + rc_rec[i] * sizeof(RCTile));
      r2[i].tables = reinterpret_cast<const int32_t*>(base + off[kRcTables] + rc_tab[i] * sizeof(int32_t));
    }
    auto put = [&](int s, const void* src, size_t n) {
      if (n) std::memcpy(out.data() + off[s], src, n);
    };
    put(kBox, b2.data(), sizeof(BatchOpDev) * b2.size());
    put(kRC, r2.data(), sizeof(BatchRCOpDev) * r2.size());
    put(kRow, rows.data(), sizeof(BatchRowOpDev) * rows.size());
    put(kWg, wg.data(), sizeof(uint32_t) * wg.size());
    put(kBases, bases.data(), sizeof(TileBase) * bases.size());
    put(kRcTiles, rct.data(), sizeof(RCTile) * rct.size());
    put(kRcTables, rtab.data(), sizeof(int32_t) * rtab.size());
  }
};

// Everything a batch decides on the host: per-op plans, groups, the image.
struct BatchPlan {
  std::vector<ttc_plan_t> plans;  // per-op plans (also used by the per-op loop)
  std::vector<ttc_batch_op_t> ops;
  std::vector<BatchGroup> groups;
  BatchImage img;
  int nwg_total = 0;
  int64_t tiles = 0, bytes = 0;

  ~BatchPlan() {
    for (ttc_plan_t p : plans)
      if (p) ttc_plan_destroy(p);
  }
};

struct ttc_batch_s {
  BatchPlan bp;
  void* d_image = nullptr;  // the work table on the device
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
};

static void batch_free(ttc_batch_s* b) {
  if (b->d_image) (void)hipFree(b->d_image);
  for (int i = 0; i < 2; ++i) {
    if (b->exec[i]) (void)hipGraphExecDestroy(b->exec[i]);
    if (b->graph[i]) (void)hipGraphDestroy(b->graph[i]);
  }
  delete b;
}

// Host version of the row x column grid walk (the kernel's tile()): one
// record per tile, tiles numbered over the grid positions in plan order.
static void fill_rc_tiles(const RCParams& p, RCTile* out) {
  for (uint32_t t = 0; t < p.ntiles; ++t) {
    uint32_t rem = t;
    RCTile r{0, 0, 0, 0};
    for (int g = 0; g < p.grank; ++g) {
      const uint32_t n = p.gdiv[g].d;
      const uint32_t x = rem % n;
      rem /= n;
      r.a += (int64_t)x * p.tsa[g];
      r.b += (int64_t)x * p.tsb[g];
      const bool partial = (uint32_t)p.n[g] - x * (uint32_t)p.t[g] < (uint32_t)p.t[g];
      if (g == p.splitA && partial) r.flags |= 1u;
      if (g == p.splitB && partial) r.flags |= 2u;
    }
    out[t] = r;
  }
}

// One op of the work table while it is built.
struct OpRec {
  int kind;
  uint32_t ntiles, tile_begin;
  double weight;  // bytes moved per tile
  int src;        // index into the caller's op list
};

// Work-group slices of one group's tile range, balanced by bytes: returns
// nwg + 1 boundaries (tile indices of the group).
static std::vector<uint32_t> split_by_bytes(const std::vector<OpRec>& h, int op0, int nops, uint32_t tiles,
                                            int nwg) {
  std::vector<uint32_t> wgb(nwg + 1, tiles);
  double total = 0;
  for (int i = op0; i < op0 + nops; ++i) total += h[i].weight * h[i].ntiles;
  const double per = total / nwg;
  int w = 1;
  double acc = 0;
  wgb[0] = 0;
  for (int i = op0; i < op0 + nops && w < nwg; ++i) {
    const double wt = h[i].weight;
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

constexpr int kBatchSlicesPerCu = 96;  // default work-group slices per CU in a group's grid

// Plans every op and builds the work table on the host (no device memory).
static ttc_status_t plan_batch(int nops, const ttc_batch_op_t* ops, int wg_per_cu, BatchPlan& bp) {
  if (nops < 1 || !ops || wg_per_cu < 0) return TTC_ERROR_INVALID_VALUE;
  bp.ops.assign(ops, ops + nops);
  bp.plans.assign(nops, nullptr);
  std::vector<OpRec> rec(nops);
  bp.bytes = 0;
  for (int i = 0; i < nops; ++i) {
    const ttc_batch_op_t& op = ops[i];
    const int pair = batch_pair(op.type_a, op.type_b);
    if (pair < 0) return TTC_ERROR_UNSUPPORTED;
    if (op.dim < 1 || op.dim > TTC_MAX_INPUT_DIM) return TTC_ERROR_INVALID_VALUE;
    bool any_lda = false, any_ldb = false;
    for (int k = 0; k < op.dim; ++k) {
      any_lda |= op.lda[k] != 0;
      any_ldb |= op.ldb[k] != 0;
    }
    // vector widths the group kernels instantiate: 1, 2 and the widest on
    // row x column tiles and row copies, 1 and the widest on box tiles
    ttc_config_t cfg = op.cfg;
    if (cfg.tile_order > 2) cfg.tile_order = 2;  // grouped launches walk host-built tile records in B order
    if (cfg.lds_buffers == 0) cfg.lds_buffers = 1;  // one LDS buffer unless the op asks for two (measured faster)
    const int wide = pair_kernels(pair).widest;
    auto width_ok = [wide](const ttc_plan_s* q, int v) {
      return v == 1 || v == wide || (v == 2 && wide > 2 && q->family != kFamilyBox);
    };
    for (int pass = 0; pass < 2; ++pass) {
      ttc_plan_t pl = nullptr;
      ttc_status_t st = ttc_plan_create(&pl, op.type_a, op.type_b, op.dim, op.perm, op.size,
                                        any_lda ? op.lda : nullptr, any_ldb ? op.ldb : nullptr,
                                        op.beta_is_zero, &cfg);
      if (st != TTC_SUCCESS) return st;
      const bool ok_a = width_ok(pl, pl->va), ok_b = width_ok(pl, pl->vb);
      if ((ok_a && ok_b) || pass == 1) {
        bp.plans[i] = pl;
        break;
      }
      ttc_plan_destroy(pl);
      if (!ok_a) cfg.vec_a = 1;
      if (!ok_b) cfg.vec_b = 1;
      for (int k = 0; k < TTC_MAX_DIM; ++k) cfg.tile[k] = 0;  // re-derive the tile for the new widths
    }
    const ttc_plan_s* pl = bp.plans[i];
    if (!width_ok(pl, pl->va) || !width_ok(pl, pl->vb)) return TTC_ERROR_UNSUPPORTED;
    if (pl->g.elements > 0 && (!op.A || !op.B)) return TTC_ERROR_INVALID_VALUE;
    const ttc_plan_s* use = pl->g.elements > 0 ? variant_for(pl, op.A, op.B) : pl;
    if (!use) return TTC_ERROR_INVALID_VALUE;
    if (!width_ok(use, use->va) || !width_ok(use, use->vb)) return TTC_ERROR_UNSUPPORTED;
    const int fam = use->family == kFamilyRC ? kGroupRC : (use->family == kFamilyRow ? kGroupRow : kGroupBox);
    const int bz = pl->beta_is_zero ? 1 : 0;
    const int ia = index_of_width(use->va, wide), ib = index_of_width(use->vb, wide);
    OpRec& r = rec[i];
    // a row copy has one width for both sides
    const int variant = fam == kGroupRow ? ia * 2 + bz : (ia * 3 + ib) * 2 + bz;
    r.kind = (fam * kBPairs + pair) * kVariantStride + variant;
    r.ntiles = pl->g.elements > 0 ? use->p.ntiles : 0;
    r.src = i;
    const int64_t sa = dtype_bytes(op.type_a), sb = dtype_bytes(op.type_b);
    r.weight = r.ntiles ? (double)pl->g.elements * (double)(sa + sb * (pl->beta_is_zero ? 1 : 2)) / r.ntiles : 0.0;
    bp.bytes += pl->g.elements * (sa + sb * (pl->beta_is_zero ? 1 : 2));
  }
  // order the table by kind (stable: ops keep their order inside a kind)
  std::stable_sort(rec.begin(), rec.end(), [](const OpRec& x, const OpRec& y) { return x.kind < y.kind; });
  int dev = 0;
  (void)hipGetDevice(&dev);
  const int cus = device_cus(dev);
  BatchImage& im = bp.img;
  bp.tiles = 0;
  for (int j = 0; j < nops;) {
    int e = j;
    uint64_t tiles = 0;
    int lds = 16;
    const int group = group_of(rec[j].kind);
    const int family = group_family(group);
    const bool rc = family == kGroupRC;
    const size_t op0 = rc ? im.rcs.size() : (family == kGroupRow ? im.rows.size() : im.box.size());
    while (e < nops && group_of(rec[e].kind) == group) {
      OpRec& r = rec[e];
      r.tile_begin = (uint32_t)tiles;
      tiles += r.ntiles;
      const ttc_batch_op_t& op = ops[r.src];
      const ttc_plan_s* pl = bp.plans[r.src];
      const ttc_plan_s* use = pl->g.elements > 0 ? variant_for(pl, op.A, op.B) : pl;
      lds = std::max(lds, use->lds_bytes);
      if (family == kGroupRow) {
        BatchRowOpDev d;
        std::memset(&d, 0, sizeof(d));
        d.p = use->row;
        d.A = op.A;
        d.B = op.B;
        d.alpha = op.alpha;
        d.beta = op.beta;
        d.variant = r.kind % kVariantStride;
        d.tile_begin = r.tile_begin;
        d.ntiles = r.ntiles;  // chunks of kThreads * kRowUnroll vectors
        im.rows.push_back(d);
      } else if (rc) {
        BatchRCOpDev d;
        std::memset(&d, 0, sizeof(d));
        d.p = use->rc;
        if (d.p.fast) {
          // the grouped kernels run the general walk (its LDS region holds
          // the lane-fixed walk's, so the plan's LDS size still covers it)
          d.p.fast = 0;
          d.p.Rw = d.p.R;
          d.p.Cw = d.p.C;
          d.p.ntab = 2 * d.p.R + 2 * d.p.C;
        }
        d.A = op.A;
        d.B = op.B;
        d.alpha = op.alpha;
        d.beta = op.beta;
        d.variant = r.kind % kVariantStride;
        d.tile_begin = r.tile_begin;
        d.ntiles = r.ntiles;
        const size_t first = im.rct.size();
        im.rct.resize(first + r.ntiles);
        if (r.ntiles) fill_rc_tiles(use->rc, im.rct.data() + first);
        const size_t tab = im.rtab.size();
        im.rtab.resize(tab + (size_t)d.p.ntab);
        if (r.ntiles) rc_fill_tables(d.p, im.rtab.data() + tab);
        im.rc_rec.push_back(first);
        im.rc_tab.push_back(tab);
        im.rcs.push_back(d);
      } else {
        BatchOpDev d;
        std::memset(&d, 0, sizeof(d));
        d.p = use->p;
        d.A = op.A;
        d.B = op.B;
        d.alpha = op.alpha;
        d.beta = op.beta;
        d.variant = r.kind % kVariantStride;
        d.tile_begin = r.tile_begin;
        d.ntiles = r.ntiles;
        const size_t first = im.bases.size();
        im.bases.resize(first + r.ntiles);
        if (r.ntiles) fill_bases(use->p, im.bases.data() + first);
        im.box_rec.push_back(first);
        im.box.push_back(d);
      }
      ++e;
    }
    if (tiles >= (1ull << 31)) return TTC_ERROR_UNSUPPORTED;
    lds = (lds + 15) / 16 * 16;
    if (tiles > 0) {
      BatchGroup grp;
      grp.group = group;
      grp.op0 = (int)op0;
      grp.nops = e - j;
      grp.lds_bytes = lds;
      // Many more slices than resident work-groups: the dispatcher hands
      // the next slice to whichever CU frees a slot first, which balances
      // the uneven ops better than one long slice per resident work-group
      // (512 mixed ops, 4 KiB-64 MiB: 2.42 ms at the occupancy limit, 2.24
      // at 16 per CU, 2.05-2.06 at 96-128, 2.08 at 256).
      const int wg = wg_per_cu > 0 ? wg_per_cu : kBatchSlicesPerCu;
      grp.nwg = (int)std::min<uint64_t>((uint64_t)cus * wg, tiles);
      grp.wg0 = (int)im.wg.size();
      std::vector<uint32_t> wgb = split_by_bytes(rec, j, e - j, (uint32_t)tiles, grp.nwg);
      im.wg.insert(im.wg.end(), wgb.begin(), wgb.end());
      bp.groups.push_back(grp);
      bp.tiles += (int64_t)tiles;
    }
    j = e;
  }
  bp.nwg_total = 0;
  for (const BatchGroup& g : bp.groups) bp.nwg_total += g.nwg;
  im.layout();
  return TTC_SUCCESS;
}

// Kernel arguments of group g with the work table at device address base.
struct GroupArgs {
  const void* ops;
  int nops;
  const uint32_t* wg;
};
static GroupArgs group_args(const BatchPlan& bp, const BatchGroup& g, uintptr_t base) {
  const BatchImage& im = bp.img;
  GroupArgs a;
  switch (group_family(g.group)) {
    case kGroupRC:
      a.ops = reinterpret_cast<const void*>(base + im.off[BatchImage::kRC] + g.op0 * sizeof(BatchRCOpDev));
      break;
    case kGroupRow:
      a.ops = reinterpret_cast<const void*>(base + im.off[BatchImage::kRow] + g.op0 * sizeof(BatchRowOpDev));
      break;
    default:
      a.ops = reinterpret_cast<const void*>(base + im.off[BatchImage::kBox] + g.op0 * sizeof(BatchOpDev));
      break;
  }
  a.nops = g.nops;
  a.wg = reinterpret_cast<const uint32_t*>(base + im.off[BatchImage::kWg] + g.wg0 * sizeof(uint32_t));
  return a;
}

static ttc_status_t launch_groups(const BatchPlan& bp, uintptr_t base, hipStream_t s) {
  for (const BatchGroup& g : bp.groups) {
    GroupArgs a = group_args(bp, g, base);
    void* args[] = {&a.ops, &a.nops, &a.wg};
    const hipError_t e = hipLaunchKernel(group_kernel(g.group), dim3(g.nwg),
                                         dim3(group_family(g.group) == kGroupRC ? kRcThreads : kThreads), args,
                                         (size_t)g.lds_bytes, s);
    if (e != hipSuccess) return TTC_ERROR_HIP;
  }
  return TTC_SUCCESS;
}

extern "C" ttc_status_t ttc_batch_create(ttc_batch_t* out, int nops, const ttc_batch_op_t* ops, int wg_per_cu) {
  if (!out || nops < 1 || !ops || wg_per_cu < 0) return TTC_ERROR_INVALID_VALUE;
  *out = nullptr;
  ttc_batch_s* b = new (std::nothrow) ttc_batch_s();
  if (!b) return TTC_ERROR_ALLOC;
  ttc_status_t st = plan_batch(nops, ops, wg_per_cu, b->bp);
  if (st != TTC_SUCCESS) {
    batch_free(b);
    return st;
  }
  const BatchImage& im = b->bp.img;
  if (hipMalloc(&b->d_image, std::max<size_t>(im.bytes, 256)) != hipSuccess) {
    b->d_image = nullptr;
    batch_free(b);
    return TTC_ERROR_ALLOC;
  }
  std::vector<unsigned char> host;
  im.render(reinterpret_cast<uintptr_t>(b->d_image), host);
  if (im.bytes && hipMemcpy(b->d_image, host.data(), im.bytes, hipMemcpyHostToDevice) != hipSuccess) {
    batch_free(b);
    return TTC_ERROR_HIP;
  }
  *out = b;
  return TTC_SUCCESS;
}

extern "C" ttc_status_t ttc_batch_execute(ttc_batch_t b, ttc_stream_t stream) {
  if (!b) return TTC_ERROR_INVALID_VALUE;
  return launch_groups(b->bp, reinterpret_cast<uintptr_t>(b->d_image), static_cast<hipStream_t>(stream));
}

extern "C" ttc_status_t ttc_batch_execute_loop(ttc_batch_t b, ttc_stream_t stream) {
  if (!b) return TTC_ERROR_INVALID_VALUE;
  const BatchPlan& bp = b->bp;
  for (size_t i = 0; i < bp.ops.size(); ++i) {
    const ttc_batch_op_t& op = bp.ops[i];
    ttc_status_t st = launch_plan(bp.plans[i], op.A, op.B, op.alpha, op.beta, static_cast<hipStream_t>(stream));
    if (st != TTC_SUCCESS) return st;
  }
  return TTC_SUCCESS;
}

extern "C" ttc_status_t ttc_batch_graph_launch(ttc_batch_t b, ttc_stream_t stream, int per_op_loop) {
  if (!b || !stream) return TTC_ERROR_INVALID_VALUE;
  const int w = per_op_loop ? 1 : 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!b->exec[w]) {
    hipGraph_t g = nullptr;
    if (w) {
      // the per-op loop as a caller would capture it: one launch after another
      if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return TTC_ERROR_HIP;
      ttc_status_t st = ttc_batch_execute_loop(b, stream);
      hipError_t e = hipStreamEndCapture(s, &g);
      if (st != TTC_SUCCESS || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        return st != TTC_SUCCESS ? st : TTC_ERROR_HIP;
      }
    } else {
      // the grouped launch: one kernel node per kind, no edges between them
      // (they touch disjoint ops), so the graph may run kinds side by side
      if (hipGraphCreate(&g, 0) != hipSuccess) return TTC_ERROR_HIP;
      for (const BatchGroup& gr : b->bp.groups) {
        GroupArgs a = group_args(b->bp, gr, reinterpret_cast<uintptr_t>(b->d_image));
        void* args[] = {&a.ops, &a.nops, &a.wg};
        hipKernelNodeParams kp;
        std::memset(&kp, 0, sizeof(kp));
        kp.func = const_cast<void*>(group_kernel(gr.group));
        kp.gridDim = dim3(gr.nwg);
        kp.blockDim = dim3(group_family(gr.group) == kGroupRC ? kRcThreads : kThreads);
        kp.sharedMemBytes = (unsigned)gr.lds_bytes;
        kp.kernelParams = args;
        hipGraphNode_t node;
        if (hipGraphAddKernelNode(&node, g, nullptr, 0, &kp) != hipSuccess) {
          (void)hipGraphDestroy(g);
          return TTC_ERROR_HIP;
        }
      }
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
  if (tiles) *tiles = b->bp.tiles;
  if (workgroups) *workgroups = b->bp.nwg_total;
  if (bytes) *bytes = b->bp.bytes;
  return TTC_SUCCESS;
}

extern "C" ttc_status_t ttc_batch_destroy(ttc_batch_t b) {
  if (!b) return TTC_ERROR_INVALID_VALUE;
  batch_free(b);
  return TTC_SUCCESS;
}

// ------------------------------------------------------- no device allocation
// The fixed workspace of ttc_batch_execute_noalloc: every GPU the library
// runs on holds its own copy until the process ends.  In host code the name
// `scratch` is only a stand-in, never a device address: scratch_address()
// looks up the current device's copy once (hipGetSymbolAddress) and all
// later uses (the table copy, the kernels' arguments) go through that.
static __device__ __align__(256) unsigned char scratch[TTC_BATCH_WORKSPACE_BYTES];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

static ttc_status_t scratch_address(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return TTC_ERROR_NO_DEVICE;
  if (dev < 0 || dev >= kMaxDevices) return TTC_ERROR_NO_DEVICE;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return TTC_ERROR_HIP;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return TTC_SUCCESS;
}

extern "C" ttc_status_t ttc_batch_execute_noalloc(int nops, const ttc_batch_op_t* ops, int wg_per_cu,
                                                  ttc_stream_t stream) {
  if (nops < 1 || !ops || wg_per_cu < 0) return TTC_ERROR_INVALID_VALUE;
  BatchPlan bp;
  ttc_status_t st = plan_batch(nops, ops, wg_per_cu, bp);
  if (st != TTC_SUCCESS) return st;
  if (bp.img.bytes > (size_t)TTC_BATCH_WORKSPACE_BYTES) return TTC_ERROR_WORKSPACE;
  if (bp.groups.empty()) return TTC_SUCCESS;  // every op is empty
  void* base = nullptr;
  st = scratch_address(&base);
  if (st != TTC_SUCCESS) return st;
  std::vector<unsigned char> host;
  bp.img.render(reinterpret_cast<uintptr_t>(base), host);
  hipStream_t s = static_cast<hipStream_t>(stream);
  // Ordered on the stream after the previous call's kernels; the pageable
  // host table is consumed before the copy call returns.
  if (hipMemcpyAsync(base, host.data(), bp.img.bytes, hipMemcpyHostToDevice, s) != hipSuccess)
    return TTC_ERROR_HIP;
  return launch_groups(bp, reinterpret_cast<uintptr_t>(base), s);
}
