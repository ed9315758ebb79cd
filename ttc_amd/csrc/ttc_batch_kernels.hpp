// Device side of the grouped batch: the work-table records and the three
// persistent group kernels (box tiles, row x column tiles, row copies).
// The kernels of each type pair are instantiated in a file of their own
// (ttc_batch_k_<pair>.hip), so that the six pairs compile in parallel; the
// host side (ttc_batch.hip) reaches them through batch_kernels(pair).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ttc_kernels.hpp"
#include "ttc_rc.hpp"
#include "ttc_row.hpp"

namespace ttc {

// Type pairs of the grouped kernels (same type on both sides).
enum BatchPair : int { kBS = 0, kBD = 1, kBBF16 = 2, kBF16 = 3, kBC = 4, kBZ = 5, kBPairs = 6 };

// Variants: each op runs one instance of its family's walk, chosen per op
// by a work-group-uniform branch, so that the ops of one (family, type
// pair) share one grid balanced by bytes, whatever their widths and beta
// class.  A vector width is coded by its index in {1, 2, W} (W = the
// pair's widest, 16 bytes): the three widths the planner ever picks.
//   row x column: variant = (index_a * 3 + index_b) * 2 + beta_is_zero  (18)
//   row copy:     variant = index * 2 + beta_is_zero                    (6)
//   box tiles:    variant = (index_a * 3 + index_b) * 2 + beta_is_zero,
//                 widths 1 and W only (index 1 is never used)
// kind = (family * kBPairs + pair) * kVariantStride + variant.
constexpr int kVariantStride = 18;

__host__ __device__ constexpr int width_of_index(int idx, int W) { return idx == 0 ? 1 : (idx == 1 ? 2 : W); }
// Index of width v among {1, 2, W}: W itself when W == 2.
__host__ __device__ constexpr int index_of_width(int v, int W) { return v == 1 ? 0 : (v == W ? 2 : 1); }
// The indices a pair with widest width W can hold (2 exists only below W).
__host__ __device__ constexpr bool index_used(int idx, int W) {
  return idx == 0 || (idx == 1 && W > 2) || (idx == 2 && W > 1);
}

struct BatchOpDev {
  TileParams p;
  const void* A;
  void* B;
  double alpha, beta;
  int32_t variant;
  int32_t pad_;
  uint32_t tile_begin;  // first tile of this op in its group
  uint32_t ntiles;
  const TileBase* bases;  // this op's tile records (ntiles of them)
};

// Op record of a row x column group.
struct BatchRCOpDev {
  RCParams p;
  const void* A;
  void* B;
  double alpha, beta;
  int32_t variant;
  int32_t pad_;
  uint32_t tile_begin;    // first tile of this op in its group
  uint32_t ntiles;
  const RCTile* tiles;    // this op's tile records
  const int32_t* tables;  // its LDS offset tables (rc_fill_tables: RCParams::ntab entries)
};

// Op record of a row-copy group (perm[0] == 0 after fusion, dimension 0
// contiguous on both sides): rows move without LDS staging, as in the
// single-op row kernel.  Its work unit is a chunk of kThreads * kRowUnroll
// vectors of the op's row walk.
struct BatchRowOpDev {
  RowParams p;
  const void* A;
  void* B;
  double alpha, beta;
  int32_t variant;
  int32_t pad_;
  uint32_t tile_begin;  // first chunk of this op in its group
  uint32_t ntiles;      // chunks
};

// Parameter records are copied out of the work table whole, one 32-bit word
// at a time, so that no field can be left behind when a field is added (a
// field-by-field copy once missed one, and a kernel then read an
// uninitialised value).  The compiler drops the words a kernel never reads.
static_assert(sizeof(TileParams) % 4 == 0 && std::is_trivially_copyable<TileParams>::value,
              "TileParams is copied as 32-bit words");
static_assert(sizeof(RCParams) % 4 == 0 && std::is_trivially_copyable<RCParams>::value,
              "RCParams is copied as 32-bit words");
static_assert(sizeof(RowParams) % 4 == 0 && std::is_trivially_copyable<RowParams>::value,
              "RowParams is copied as 32-bit words");

using KRCOp = const __attribute__((address_space(4))) BatchRCOpDev;
using KRowOp = const __attribute__((address_space(4))) BatchRowOpDev;
using KWord = const __attribute__((address_space(4))) uint32_t;

__device__ __forceinline__ int32_t uni(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// The work-group's slice [t0, t1) of its group's units starts in the last
// op whose first unit is at or before t0.
template <class R>
__device__ __forceinline__ int first_op(const R* __restrict__ ops, int nops, uint32_t t0) {
  int lo = 0, hi = nops - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ops[mid].tile_begin <= t0) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// ------------------------------------------------------------- box tiles

// A box op's parameters, each word marked wave-uniform (readfirstlane), so
// that they live in scalar registers as kernel arguments do in the
// single-op kernel.
__device__ __forceinline__ void uniform_params(const TileParams& src, TileParams& p) {
  const uint32_t* s = reinterpret_cast<const uint32_t*>(&src);
  uint32_t* d = reinterpret_cast<uint32_t*>(&p);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(TileParams) / 4); ++i) d[i] = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)s[i]);
}

template <class OP, int VA, int VB, bool BETA>
__device__ __forceinline__ void box_op_slice(const BatchOpDev& o, uint32_t lo, uint32_t hi, unsigned char* lds_raw) {
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

template <class OP, int W, int I = 0>
__device__ __forceinline__ void box_dispatch(int v, const BatchOpDev& o, uint32_t lo, uint32_t hi,
                                             unsigned char* lds) {
  if constexpr (I < kVariantStride) {
    constexpr int ia = I / 6, ib = (I / 2) % 3;
    if constexpr (ia != 1 && ib != 1 && index_used(ia, W) && index_used(ib, W)) {
      if (v == I) {
        box_op_slice<OP, width_of_index(ia, W), width_of_index(ib, W), (I & 1) == 0>(o, lo, hi, lds);
        return;
      }
    }
    box_dispatch<OP, W, I + 1>(v, o, lo, hi, lds);
  }
}

template <class OP, int W>
__global__ void __launch_bounds__(kThreads) batch_box_kernel(const BatchOpDev* __restrict__ ops, int nops,
                                                             const uint32_t* __restrict__ wg_begin) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const uint32_t t0 = wg_begin[blockIdx.x], t1 = wg_begin[blockIdx.x + 1];
  if (t0 >= t1) return;
  uint32_t t = t0;
  for (int op = first_op(ops, nops, t0); op < nops && t < t1; ++op) {
    const BatchOpDev& o = ops[op];
    const uint32_t end = o.tile_begin + o.ntiles;
    const uint32_t stop = end < t1 ? end : t1;
    if (t < stop) {
      box_dispatch<OP, W>(uni(o.variant), o, t - o.tile_begin, stop - o.tile_begin, lds_raw);
      t = stop;
    }
  }
}

// ---------------------------------------------------- row x column tiles

// A row x column op's parameters, read through the constant address space
// (scalar loads), word by word.
__device__ __forceinline__ void scalar_rc(KRCOp& o, RCParams& p) {
  KWord* s = (KWord*)&o.p;
  uint32_t* d = reinterpret_cast<uint32_t*>(&p);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(RCParams) / 4); ++i) d[i] = s[i];
}

// One op's slice [lo, hi) of its tiles on row x column tiles.
template <class OP, int VA, int VB, bool BETA>
__device__ __forceinline__ void rc_op_slice(KRCOp& o, uint32_t lo, uint32_t hi, unsigned char* lds_raw) {
  RCWorker<OP, VA, VB, BETA, false> w;
  RCParams up;
  scalar_rc(o, up);
  // ends with a barrier; the previous op's last tile ended with one
  w.setup_tables(up, o.tables, lds_raw, threadIdx.x);
  rc_run_table(w, up, (KTile*)o.tiles, lo, hi, static_cast<const typename OP::TA*>(o.A),
               static_cast<typename OP::TB*>(o.B), (typename OP::AL)o.alpha, (typename OP::BE)o.beta);
}

template <class OP, int W, int I = 0>
__device__ __forceinline__ void rc_dispatch(int v, KRCOp& o, uint32_t lo, uint32_t hi, unsigned char* lds) {
  if constexpr (I < kVariantStride) {
    constexpr int ia = I / 6, ib = (I / 2) % 3;
    if constexpr (index_used(ia, W) && index_used(ib, W)) {
      if (v == I) {
        rc_op_slice<OP, width_of_index(ia, W), width_of_index(ib, W), (I & 1) == 0>(o, lo, hi, lds);
        return;
      }
    }
    rc_dispatch<OP, W, I + 1>(v, o, lo, hi, lds);
  }
}

template <class OP, int W>
__global__ void __launch_bounds__(kRcThreads) batch_rc_kernel(const BatchRCOpDev* __restrict__ ops, int nops,
                                                              const uint32_t* __restrict__ wg_begin) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const uint32_t t0 = wg_begin[blockIdx.x], t1 = wg_begin[blockIdx.x + 1];
  if (t0 >= t1) return;
  KRCOp* kops = (KRCOp*)ops;
  uint32_t t = t0;
  for (int op = first_op(ops, nops, t0); op < nops && t < t1; ++op) {
    KRCOp& o = kops[op];
    const uint32_t end = o.tile_begin + o.ntiles;
    const uint32_t stop = end < t1 ? end : t1;
    if (t < stop) {
      rc_dispatch<OP, W>(o.variant, o, t - o.tile_begin, stop - o.tile_begin, lds_raw);
      t = stop;
    }
  }
}

// ------------------------------------------------------------ row copies

// Chunks [lo, hi) of one row op: lane l of chunk c moves vectors
// c * kThreads * U + u * kThreads + l (u < U) of the op's walk (rows of
// dimension 0 in B order, V elements per vector).
template <class OP, int V, bool BETA>
__device__ __forceinline__ void row_op_slice(KRowOp& o, uint32_t lo, uint32_t hi) {
  using TA = typename OP::TA;
  using TB = typename OP::TB;
  constexpr int U = kRowUnroll;
  RowParams p;
  {
    KWord* s = (KWord*)&o.p;
    uint32_t* d = reinterpret_cast<uint32_t*>(&p);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(RowParams) / 4); ++i) d[i] = s[i];
  }
  const typename OP::AL al = (typename OP::AL)o.alpha;
  const typename OP::BE be = (typename OP::BE)o.beta;
  const char* a = static_cast<const char*>(o.A);
  char* b = static_cast<char*>(o.B);
  for (uint32_t c = lo; c < hi; ++c) {
    const uint32_t g0 = c * (uint32_t)(kThreads * U) + threadIdx.x;
    int64_t ob[U];
    bool ok[U];
    TA va[U][V];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t g = g0 + (uint32_t)(u * kThreads);
      ok[u] = g < p.total;
      const uint32_t gg = ok[u] ? g : 0u;
      const uint32_t row = fdiv(gg, p.dnv);
      const uint32_t v = gg - row * p.nv;
      int64_t xa = (int64_t)v * (int64_t)(V * sizeof(TA));
      int64_t xb = (int64_t)v * (int64_t)(V * sizeof(TB));
      uint32_t rem = row;
#pragma unroll
      for (int k = 0; k < kMaxDim; ++k) {
        if (k < p.rank) {
          const uint32_t q = fdiv(rem, p.dn[k]);
          const uint32_t x = rem - q * p.dn[k].d;
          rem = q;
          xa += (int64_t)x * p.sa[k];
          xb += (int64_t)x * p.sb[k];
        }
      }
      ob[u] = xb;
      if (ok[u]) vload<TA, V>(va[u], reinterpret_cast<const TA*>(a + xa));
    }
    TB old[U][V];
    if constexpr (BETA) {
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (ok[u]) vload<TB, V>(old[u], reinterpret_cast<const TB*>(b + ob[u]));
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!ok[u]) continue;
      TB out[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        if constexpr (BETA) out[v] = OP::outb(OP::prod(al, va[u][v]), be, old[u][v]);
        else out[v] = OP::out(OP::prod(al, va[u][v]));
      }
      vstore<TB, V, false>(reinterpret_cast<TB*>(b + ob[u]), out);
    }
  }
}

template <class OP, int W, int I = 0>
__device__ __forceinline__ void row_dispatch(int v, KRowOp& o, uint32_t lo, uint32_t hi) {
  if constexpr (I < 6) {
    if constexpr (index_used(I / 2, W)) {
      if (v == I) {
        row_op_slice<OP, width_of_index(I / 2, W), (I & 1) == 0>(o, lo, hi);
        return;
      }
    }
    row_dispatch<OP, W, I + 1>(v, o, lo, hi);
  }
}

template <class OP, int W>
__global__ void __launch_bounds__(kThreads) batch_row_kernel(const BatchRowOpDev* __restrict__ ops, int nops,
                                                             const uint32_t* __restrict__ wg_begin) {
  const uint32_t t0 = wg_begin[blockIdx.x], t1 = wg_begin[blockIdx.x + 1];
  if (t0 >= t1) return;
  KRowOp* kops = (KRowOp*)ops;
  uint32_t t = t0;
  for (int op = first_op(ops, nops, t0); op < nops && t < t1; ++op) {
    KRowOp& o = kops[op];
    const uint32_t end = o.tile_begin + o.ntiles;
    const uint32_t stop = end < t1 ? end : t1;
    if (t < stop) {
      row_dispatch<OP, W>(o.variant, o, t - o.tile_begin, stop - o.tile_begin);
      t = stop;
    }
  }
}

// ------------------------------------------------------------- registry

// The three group kernels of one type pair, W its widest vector width.
struct BatchKernelSet {
  const void* box;
  const void* rc;
  const void* row;
  int widest;
};

template <class OP, int W>
BatchKernelSet batch_kernel_set() {
  return BatchKernelSet{reinterpret_cast<const void*>(&batch_box_kernel<OP, W>),
                        reinterpret_cast<const void*>(&batch_rc_kernel<OP, W>),
                        reinterpret_cast<const void*>(&batch_row_kernel<OP, W>), W};
}

// Defined one per file, ttc_batch_k_<pair>.hip.
BatchKernelSet batch_kernels_s();
BatchKernelSet batch_kernels_d();
BatchKernelSet batch_kernels_bf16();
BatchKernelSet batch_kernels_f16();
BatchKernelSet batch_kernels_c();
BatchKernelSet batch_kernels_z();

}  // namespace ttc
