// Row x column tile kernels: the common transposition shape where the tile
// splits into an "A group" of dimensions (A's leading dimensions, the tile's
// rows, contiguous in A) and a disjoint "B group" (B's leading dimensions,
// the tile's columns, contiguous in B).  Element (i, j) of a tile sits at
//   A: base_a + rowA[i] + colA[j]      B: base_b + rowB[i] + colB[j]
//   LDS: i + j * RP
// with the four offset tables computed once per work-group into LDS.  Each
// element's (i, j) is recomputed from its tile-linear index with one
// multiply-high division, so no per-element state stays in registers: tiles
// can be twice the box kernel's volume with half its registers.
//
// Read phase: lane-consecutive vectors walk the tile row-major (i fastest:
// contiguous in A).  Write phase: column-major (j fastest: contiguous in B).
// The arithmetic (alpha in A's type before LDS, beta epilogue in B's type) is
// the Op's, exactly as in the box kernel.  Global memory is accessed through
// buffer resources with 32-bit tile-relative byte offsets; lanes past a
// partial tile get an out-of-range offset instead of a branch.
#pragma once
#include "ttc_kernels.hpp"

namespace ttc {

// The kernel parameters seen through the kernel-argument segment (constant
// address space: loads from it are scalar and uniform).
using KParams = const __attribute__((address_space(4))) RCParams;

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uniform64(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// FAST: the instance that runs full tiles of a p.fast launch on the
// lane-fixed walk (partial tiles on the general one); the general instance
// carries none of its code or state.
template <class OP, int VA, int VB, bool BETA, bool NT, bool FAST = false>
struct RCWorker {
  using TA = typename OP::TA;
  using TB = typename OP::TB;
  using L = typename OP::L;
  static constexpr int NIA = rc_elems((int)sizeof(TA)) / VA;  // read vectors per thread per tile
  static constexpr int NIB = rc_elems((int)sizeof(TB)) / VB;  // write vectors per thread per tile
  static constexpr uint32_t SA = kRcThreads * VA;               // read-walk step between a thread's slots
  static constexpr uint32_t SB = kRcThreads * VB;               // write-walk step
  static constexpr bool FP8 = is_fp8<OP>::value;                // fp8 cast: beta carries the scale

  mutable float amax = 0.0f;  // fp8 cast: max |alpha * a| over this thread's elements

  L* data;              // LDS tile
  const int32_t* rowA;  // LDS tables, byte offsets
  const int32_t* rowB;
  const int32_t* colA;
  const int32_t* colB;
  uint32_t tid;
  // Lane-fixed walk (full tiles of a launch with p.fast): this lane's row
  // offset and slot-ordered column offsets for the read walk, its column
  // offset and slot-ordered row offsets for the write walk, and the LDS
  // index of its slot 0 and the step between slots on each side.
  int32_t fa_row;
  const int32_t* fa_col;
  uint32_t fa_lds, fa_step;
  int32_t fb_col;
  const int32_t* fb_row;
  uint32_t fb_lds, fb_step;

  // One tile: origin, shape, and each walk's start and per-slot step.  The
  // read walk visits (i, j) with i fastest, the write walk with j fastest;
  // slot `it` of a thread is element l = tid * V + it * S, and stepping by S
  // is an add with one wrap ((i, j) never needs a division after the start).
  struct Tile {
    int64_t bA, bB;        // element offsets of the tile origin
    bool fast;             // full tile of a fast launch: lane-fixed walk (the fields below are unset)
    uint32_t rows, cols;
    int32_t limA, limB;    // slot it is inside the tile iff it * S < lim (per lane)
    uint32_t ai, aj;       // read walk start (per lane)
    uint32_t adi, adj;     // SA = adj * rows + adi
    uint32_t bi, bj;       // write walk start (per lane)
    uint32_t bdi, bdj;     // SB = bdi * cols + bdj
  };

  // LDS table layout (see kRcTabOOB): rowA[Rw] rowB[R] colA[C] colB[Cw]
  // colS[kA * NIA] rowS[kB * NIB].
  __device__ __forceinline__ static int32_t* tables(const RCParams& p, unsigned char* lds_raw) {
    return reinterpret_cast<int32_t*>(lds_raw + (size_t)p.nbuf * (size_t)p.lds_data_elems * sizeof(L));
  }

  // Offsets of tile row i (A, B) and tile column j (A, B): mixed-radix
  // digits over the group's tile extents times their byte strides.
  __device__ __forceinline__ static void row_offsets(const RCParams& p, uint32_t l, int32_t& oa, int32_t& ob) {
    oa = ob = 0;
    for (int m = 0; m < p.na; ++m) {
      const uint32_t q = fdiv(l, p.dta[m]);
      const uint32_t d = l - q * (uint32_t)p.ta[m];
      l = q;
      oa += (int32_t)d * p.saA[m];
      ob += (int32_t)d * p.sbA[m];
    }
  }
  __device__ __forceinline__ static void col_offsets(const RCParams& p, uint32_t l, int32_t& oa, int32_t& ob) {
    oa = ob = 0;
    for (int m = 0; m < p.nb; ++m) {
      const uint32_t q = fdiv(l, p.dtb[m]);
      const uint32_t d = l - q * (uint32_t)p.tb[m];
      l = q;
      oa += (int32_t)d * p.saB[m];
      ob += (int32_t)d * p.sbB[m];
    }
  }

  // Table pointers and this lane's constants of the lane-fixed walk
  // (fa_row and fb_col are set by the caller).
  __device__ __forceinline__ void bind_tables(const RCParams& p, int32_t* tab, uint32_t& ai, uint32_t& bj) {
    rowA = tab;
    rowB = tab + p.Rw;
    colA = rowB + p.R;
    colB = colA + p.C;
    if constexpr (!FAST) return;
    const int32_t* cS = colB + p.Cw;
    const int32_t* rS = cS + p.kA * NIA;
    const uint32_t la = tid * VA, lb = tid * VB;
    ai = la & (uint32_t)(p.Rw - 1);
    bj = lb & (uint32_t)(p.Cw - 1);
    const uint32_t jl = la >> p.Rw_shift, il = lb >> p.Cw_shift;
    // a padding row stages every slot into the spare slot; a padding column
    // reads element 0 (its stores are dropped)
    const bool row_ok = ai < (uint32_t)p.R, col_ok = bj < (uint32_t)p.C;
    fa_col = cS + jl * NIA;
    fa_lds = row_ok ? ai + jl * (uint32_t)p.RP : (uint32_t)p.lds_spare;
    fa_step = row_ok ? (uint32_t)p.kA * (uint32_t)p.RP : 0u;
    fb_row = rS + il * NIB;
    fb_lds = col_ok ? il + bj * (uint32_t)p.RP : 0u;
    fb_step = col_ok ? (uint32_t)p.kB : 0u;
  }

  // Offset tables into LDS (after the data region), once per launch, in
  // one pass and one barrier (layout: see kRcTabOOB).
  __device__ __forceinline__ void setup(const RCParams& p, unsigned char* lds_raw, uint32_t tid_) {
    tid = tid_;
    data = reinterpret_cast<L*>(lds_raw);
    int32_t* tab = tables(p, lds_raw);
    int32_t* rA = tab;
    int32_t* rB = tab + p.Rw;
    int32_t* cA = rB + p.R;
    int32_t* cB = cA + p.C;
    // (the general instance needs no padding entries)
    const int nr = FAST ? p.Rw : p.R, nc = FAST ? p.Cw : p.C;
    for (int i = (int)tid; i < nr; i += kRcThreads) {
      int32_t oa = (int32_t)kRcTabOOB, ob = 0;
      if (i < p.R) row_offsets(p, (uint32_t)i, oa, ob);
      rA[i] = oa;
      if (i < p.R) rB[i] = ob;
    }
    for (int j = (int)tid; j < nc; j += kRcThreads) {
      int32_t oa = 0, ob = (int32_t)kRcTabOOB;
      if (j < p.C) col_offsets(p, (uint32_t)j, oa, ob);
      if (j < p.C) cA[j] = oa;
      cB[j] = ob;
    }
    uint32_t ai = 0, bj = 0;
    bind_tables(p, tab, ai, bj);
    if (FAST && p.fast) {
      int32_t* cS = cB + p.Cw;
      int32_t* rS = cS + p.kA * NIA;
      for (int e = (int)tid; e < p.kA * NIA; e += kRcThreads) {
        const int j = e / NIA + (e % NIA) * p.kA;  // NIA is a power of two
        int32_t oa = (int32_t)kRcTabOOB, ob = 0;
        if (j < p.C) col_offsets(p, (uint32_t)j, oa, ob);
        cS[e] = oa;
      }
      for (int e = (int)tid; e < p.kB * NIB; e += kRcThreads) {
        const int i = e / NIB + (e % NIB) * p.kB;
        int32_t oa = 0, ob = (int32_t)kRcTabOOB;
        if (i < p.R) row_offsets(p, (uint32_t)i, oa, ob);
        rS[e] = ob;
      }
      int32_t oa = 0, ob = 0;
      fa_row = (int32_t)kRcTabOOB;
      if (ai < (uint32_t)p.R) row_offsets(p, ai, fa_row, ob);
      fb_col = (int32_t)kRcTabOOB;
      if (bj < (uint32_t)p.C) col_offsets(p, bj, oa, fb_col);
    }
    __syncthreads();
  }

  // Tile `id` of the grid (kp: the parameters read from the kernel-argument
  // segment with a runtime index, so that the grid arrays are not pinned in
  // scalar registers across the tile loop).  empty: a tile with no elements.
  __device__ __forceinline__ Tile tile(const RCParams& p, KParams* kp, uint32_t id, bool empty) const {
    Tile c;
    uint32_t rem = id;
    c.bA = 0;
    c.bB = 0;
    bool lastA = false, lastB = false, pad = false;
    uint32_t x0 = 0;
    const int gr = p.grank;
#pragma unroll 1
    for (int g = 0; g < gr; ++g) {
      FastDiv d;  // field by field: no struct copy across address spaces
      d.d = kp->gdiv[g].d;
      d.m = kp->gdiv[g].m;
      d.s = kp->gdiv[g].s;
      const uint32_t q = fdiv(rem, d);
      const uint32_t x = rem - q * d.d;
      rem = q;
      c.bA += (int64_t)x * kp->tsa[g];
      c.bB += (int64_t)x * kp->tsb[g];
      // tile index along the dimension: a split dimension's outer position
      // adds the inner one (position 0) back; past its tile count: padding
      uint32_t xa = x;
      if (g == 0) x0 = x;
      if (g == p.gouter) {
        xa = x0 + p.group * x;
        pad |= xa >= p.gcount;
      }
      const bool partial = (uint32_t)kp->n[g] - xa * (uint32_t)kp->t[g] < (uint32_t)kp->t[g];
      lastA |= g == p.splitA && partial;
      lastB |= g == p.splitB && partial;
    }
    c.bA = uniform64(c.bA);
    c.bB = uniform64(c.bB);
    c.fast = FAST && p.fast && !lastA && !lastB && !empty && !pad;
    if (c.fast) return c;
    c.rows = lastA ? (uint32_t)p.Rl : (uint32_t)p.R;
    c.cols = lastB ? (uint32_t)p.Cl : (uint32_t)p.C;
    const FastDiv dr = lastA ? p.dRl : p.dR;
    const FastDiv dc = lastB ? p.dCl : p.dC;
    finish(c, dr, dc, empty || pad);
    return c;
  }

  // Offset tables copied from host-computed ones (grouped launches: the
  // p.ntab entries of the LDS layout, rc_fill_tables).  The caller has
  // passed a barrier since the previous tile's last LDS read.
  __device__ __forceinline__ void setup_tables(const RCParams& p, const int32_t* __restrict__ tab,
                                               unsigned char* lds_raw, uint32_t tid_) {
    tid = tid_;
    data = reinterpret_cast<L*>(lds_raw);
    int32_t* t = tables(p, lds_raw);
    for (int i = (int)tid; i < p.ntab; i += kRcThreads) t[i] = tab[i];
    uint32_t ai = 0, bj = 0;
    bind_tables(p, t, ai, bj);
    __syncthreads();
    if (FAST && p.fast) {
      fa_row = t[ai];
      fb_col = colB[bj];
    }
  }

  // Tile from a host-computed record (grouped launches): origin offsets and
  // whether it is the last, partial tile along the row / column group.
  template <class Rec>
  __device__ __forceinline__ Tile tile_rec(const RCParams& p, Rec& rec, bool empty) const {
    Tile c;
    c.bA = uniform64(rec.a);
    c.bB = uniform64(rec.b);
    const uint32_t f = uniform(rec.flags);
    const bool lastA = (f & 1u) != 0, lastB = (f & 2u) != 0;
    c.fast = FAST && p.fast && !lastA && !lastB && !empty;
    if (c.fast) return c;
    c.rows = lastA ? (uint32_t)p.Rl : (uint32_t)p.R;
    c.cols = lastB ? (uint32_t)p.Cl : (uint32_t)p.C;
    const FastDiv dr = lastA ? p.dRl : p.dR;
    const FastDiv dc = lastB ? p.dCl : p.dC;
    finish(c, dr, dc, empty);
    return c;
  }

  // Per-lane walk starts and the uniform steps of a tile whose shape is set.
  __device__ __forceinline__ void finish(Tile& c, const FastDiv& dr, const FastDiv& dc, bool empty) const {
    const int32_t n = empty ? 0 : (int32_t)(c.rows * c.cols);
    c.limA = n - (int32_t)(tid * VA);
    c.limB = n - (int32_t)(tid * VB);
    c.aj = fdiv(tid * VA, dr);
    c.ai = tid * VA - c.aj * c.rows;
    c.adj = fdiv(SA, dr);
    c.adi = SA - c.adj * c.rows;
    c.bi = fdiv(tid * VB, dc);
    c.bj = tid * VB - c.bi * c.cols;
    c.bdi = fdiv(SB, dc);
    c.bdj = SB - c.bdi * c.cols;
  }

  template <bool DENSE>
  __device__ __forceinline__ void load_impl(const RCParams& p, const Tile& c, const TA* __restrict__ A,
                                            TA (&va)[NIA][VA]) const {
    const BufRes r = make_res(A + c.bA);
    const uint32_t cmax = (uint32_t)p.C - 1;
    uint32_t i = c.ai, j = c.aj;
#pragma unroll
    for (int it = 0; it < NIA; ++it) {
      const bool ok = (int32_t)(it * SA) < c.limA;
      const uint32_t jc = j < cmax ? j : cmax;  // past the tile: any table entry
      const uint32_t off = (DENSE ? i * (uint32_t)sizeof(TA) : (uint32_t)rowA[i]) + (uint32_t)colA[jc];
      bload<TA, VA>(va[it], r, ok ? off : kBufOOB);
      i += c.adi;
      j += c.adj;
      if (i >= c.rows) {
        i -= c.rows;
        ++j;
      }
    }
  }
  __device__ __forceinline__ void load(const RCParams& p, const Tile& c, const TA* __restrict__ A,
                                       TA (&va)[NIA][VA]) const {
    if (FAST && c.fast) {
      const BufRes r = make_res(A + c.bA);
#pragma unroll
      for (int it = 0; it < NIA; ++it) bload<TA, VA>(va[it], r, (uint32_t)fa_row + (uint32_t)fa_col[it]);
    } else if (p.rowA_dense) {
      load_impl<true>(p, c, A, va);
    } else {
      load_impl<false>(p, c, A, va);
    }
  }

  __device__ __forceinline__ void stage(const RCParams& p, const Tile& c, const TA (&va)[NIA][VA],
                                        typename OP::AL alpha) const {
    if (FAST && c.fast) {
      // a slot past column C lands at or beyond RP * C: clamped to the spare
      // slot, as are padding rows (fa_step 0)
      const uint32_t spare = (uint32_t)p.lds_spare;
#pragma unroll
      for (int it = 0; it < NIA; ++it) {
        const uint32_t li = min(fa_lds + (uint32_t)it * fa_step, spare);
#pragma unroll
        for (int v = 0; v < VA; ++v) {
          const L x = OP::prod(alpha, va[it][v]);
          if constexpr (FP8) amax = fmaxf(amax, fabsf(x));
          data[li + v] = x;
        }
      }
      return;
    }
    const uint32_t RP = (uint32_t)p.RP;
    const uint32_t dli = c.adi + c.adj * RP, wrap = RP - c.rows;
    uint32_t i = c.ai, li = c.ai + c.aj * RP;
#pragma unroll
    for (int it = 0; it < NIA; ++it) {
      const bool ok = (int32_t)(it * SA) < c.limA;
      const uint32_t a = ok ? li : (uint32_t)p.lds_spare;
#pragma unroll
      for (int v = 0; v < VA; ++v) {
        const L x = OP::prod(alpha, va[it][v]);
        if constexpr (FP8) amax = fmaxf(amax, fabsf(x));  // lanes past the tile loaded zeros
        data[a + v] = x;
      }
      i += c.adi;
      li += dli;
      if (i >= c.rows) {
        i -= c.rows;
        li += wrap;
      }
    }
  }

  // The write walk: for each slot, its byte offset in B (out of range past
  // the tile) and its LDS index; f(it, off, li).
  template <bool DENSE, class F>
  __device__ __forceinline__ void walk_b(const RCParams& p, const Tile& c, F&& f) const {
    const uint32_t RP = (uint32_t)p.RP;
    const uint32_t rmax = (uint32_t)p.R - 1;
    const uint32_t dli = c.bdj * RP + c.bdi, wrap = 1u - c.cols * RP;
    uint32_t i = c.bi, j = c.bj, li = c.bi + c.bj * RP;
#pragma unroll
    for (int it = 0; it < NIB; ++it) {
      const bool ok = (int32_t)(it * SB) < c.limB;
      const uint32_t ic = i < rmax ? i : rmax;
      const uint32_t off = (uint32_t)rowB[ic] + (DENSE ? j * (uint32_t)sizeof(TB) : (uint32_t)colB[j]);
      f(it, ok ? off : kBufOOB, ok ? li : 0u);
      j += c.bdj;
      i += c.bdi;
      li += dli;
      if (j >= c.cols) {
        j -= c.cols;
        ++i;
        li += wrap;
      }
    }
  }

  // The lane-fixed write walk: slot it is row fb_row's it-th entry of this
  // lane's column (offset kRcTabOOB past R or C).
  template <class F>
  __device__ __forceinline__ void walk_fast(F&& f) const {
#pragma unroll
    for (int it = 0; it < NIB; ++it) f(it, (uint32_t)fb_row[it] + (uint32_t)fb_col, fb_lds + (uint32_t)it * fb_step);
  }

  // Old values of B for the beta epilogue.
  __device__ __forceinline__ void load_b(const RCParams& p, const Tile& c, const TB* __restrict__ B,
                                         TB (&ob)[NIB][VB]) const {
    const BufRes r = make_res(B + c.bB);
    auto f = [&](int it, uint32_t off, uint32_t) { bload<TB, VB>(ob[it], r, off); };
    if (FAST && c.fast) walk_fast(f);
    else if (p.colB_dense) walk_b<true>(p, c, f);
    else walk_b<false>(p, c, f);
  }

  __device__ __forceinline__ void store(const RCParams& p, const Tile& c, TB* __restrict__ B,
                                        typename OP::BE beta, const TB (&ob)[NIB][VB]) const {
    const BufRes r = make_res(B + c.bB);
    const uint32_t RP = (uint32_t)p.RP;
    auto f = [&](int it, uint32_t off, uint32_t li) {
      TB o[VB];
#pragma unroll
      for (int v = 0; v < VB; ++v) {
        if constexpr (FP8) {
          if constexpr (VB == 1) {
            o[0] = OP::outs(data[li], beta);
          } else if (v % 2 == 0) {
            OP::outs2(data[li + v * RP], data[li + (v + 1) * RP], beta, o[v], o[v + 1]);
          }
        } else if constexpr (BETA) {
          o[v] = OP::outb(data[li + v * RP], beta, ob[it][v]);
        } else {
          o[v] = OP::out(data[li + v * RP]);
        }
      }
      bstore<TB, VB, NT>(r, off, o);
    };
    if (FAST && c.fast) walk_fast(f);
    else if (p.colB_dense) walk_b<true>(p, c, f);
    else walk_b<false>(p, c, f);
  }

  // Software pipeline over this work-group's tiles: the next tile's reads
  // are in flight while the current tile is written.  Memory operations per
  // iteration, in issue order: old B of tile t (beta only), A of tile t+1,
  // B stores of tile t; staging tile t+1 then waits only for its own reads.
  // The last iteration prefetches an empty tile (all lanes out of range).
  __device__ __forceinline__ void run(const RCParams& p, KParams* kp, const TA* __restrict__ A,
                                      TB* __restrict__ B, uint32_t first, uint32_t end, uint32_t step,
                                      typename OP::AL alpha, typename OP::BE beta) {
    if (first >= end) return;
    TA va[NIA][VA];
    TB ob[NIB][VB];
    Tile c = tile(p, kp, first, false);
    load(p, c, A, va);
    // Two LDS buffers (nbuf == 2): tile k stages into buffer k mod 2, so the
    // barrier after staging also orders the previous use of that buffer
    // (the reads of tile k-2, finished before barrier k-1) and the second
    // barrier goes.  One buffer: the next staging waits for this tile's reads.
    L* const base = data;
    const uint32_t alt = p.nbuf == 2 ? (uint32_t)p.lds_data_elems : 0u;
    uint32_t off = 0;
    // Wave priority falls per quarter of the work-group's tiles, so that a
    // work-group behind its neighbours on the CU (dispatched later: older
    // waves win arbitration) keeps the higher priority (kPrioMinRounds).
    const uint32_t total = (end - first + step - 1) / step, q1 = total / 4, q2 = total / 2, q3 = total - total / 4;
    const bool sched = total >= kPrioMinRounds;
    uint32_t k = 0;
    for (uint32_t t = first; t < end; t += step, ++k) {
      if (sched) {
        if (k == 0) __builtin_amdgcn_s_setprio(3);
        else if (k == q1) __builtin_amdgcn_s_setprio(2);
        else if (k == q2) __builtin_amdgcn_s_setprio(1);
        else if (k == q3) __builtin_amdgcn_s_setprio(0);
      }
      data = base + off;
      stage(p, c, va, alpha);
      __syncthreads();
      if constexpr (BETA) load_b(p, c, B, ob);
      const uint32_t tn = t + step;
      const Tile cn = tile(p, kp, tn < end ? tn : t, tn >= end);
      load(p, cn, A, va);
      store(p, c, B, beta, ob);
      if (alt == 0) __syncthreads();
      off = alt - off;
      c = cn;
    }
    data = base;
  }
};

// Host-computed tile records read through the constant address space
// (scalar loads; the table does not change during a launch).
using KTile = const __attribute__((address_space(4))) RCTile;

// run() over tiles [lo, hi) of a host-computed record table.
template <class OP, int VA, int VB, bool BETA, bool NT>
__device__ __forceinline__ void rc_run_table(RCWorker<OP, VA, VB, BETA, NT>& w, const RCParams& p,
                                             KTile* recs, uint32_t lo, uint32_t hi,
                                             const typename OP::TA* __restrict__ A, typename OP::TB* __restrict__ B,
                                             typename OP::AL alpha, typename OP::BE beta) {
  using W = RCWorker<OP, VA, VB, BETA, NT>;
  if (lo >= hi) return;
  typename OP::TA va[W::NIA][VA];
  typename OP::TB ob[W::NIB][VB];
  typename W::Tile c = w.tile_rec(p, recs[lo], false);
  w.load(p, c, A, va);
  // LDS buffers as in RCWorker::run; with two, one barrier at the end orders
  // this op's last reads before the next op rewrites the offset tables.
  typename OP::L* const base = w.data;
  const uint32_t alt = p.nbuf == 2 ? (uint32_t)p.lds_data_elems : 0u;
  uint32_t off = 0;
  for (uint32_t t = lo; t < hi; ++t) {
    w.data = base + off;
    w.stage(p, c, va, alpha);
    __syncthreads();
    if constexpr (BETA) w.load_b(p, c, B, ob);
    const uint32_t tn = t + 1 < hi ? t + 1 : t;
    const typename W::Tile cn = w.tile_rec(p, recs[tn], t + 1 >= hi);
    w.load(p, cn, A, va);
    w.store(p, c, B, beta, ob);
    if (alt == 0) __syncthreads();
    off = alt - off;
    c = cn;
  }
  w.data = base;
  if (alt != 0) __syncthreads();
}

// The parameters are the kernel's first argument, so they sit at the start
// of the kernel-argument segment.
template <class OP, int VA, int VB, bool BETA, bool NT, bool FAST>
__global__ void __launch_bounds__(kRcThreads) transpose_rc_kernel(const RCParams p, const typename OP::TA* __restrict__ A,
                                                                typename OP::TB* __restrict__ B, double alpha,
                                                                double beta) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  KParams* kp = (KParams*)__builtin_amdgcn_kernarg_segment_ptr();
  RCWorker<OP, VA, VB, BETA, NT, FAST> w;
  w.setup(p, lds_raw, threadIdx.x);
  const TileRange r = tile_range(p.ntiles, blockIdx.x, gridDim.x, p.grid_order);
  // Work-group-uniform by construction; readfirstlane keeps them (and all
  // tile coordinates derived from them) in scalar registers.
  w.run(p, kp, A, B, uniform(r.first), uniform(r.end), uniform(r.step), (typename OP::AL)alpha,
        (typename OP::BE)beta);
}

// fp8 cast-transpose on row x column tiles: q = sat_rne(fl32(fl32(alpha*a)
// * scale)), amax = max |fl32(alpha*a)| (one atomic per work-group).  The
// scale comes from *scale_in (delayed scaling), from fp8_max / *amax_in
// (current scaling, amax_in from a previous pass), or from the host value.
template <class OP, int VA, int VB, bool FAST>
__global__ void __launch_bounds__(kRcThreads)
    cast_fp8_rc_kernel(const RCParams p, const typename OP::TA* __restrict__ A, typename OP::TB* __restrict__ B,
                       float alpha, const float* __restrict__ scale_in, const float* __restrict__ amax_in,
                       float fp8_max, float scale_host, float* __restrict__ amax_out,
                       float* __restrict__ scale_inv_out) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  KParams* kp = (KParams*)__builtin_amdgcn_kernarg_segment_ptr();
  float scale = scale_host;
  if (scale_in) {
    scale = *scale_in;
  } else if (amax_in) {
    const float m = *amax_in;
    scale = (m > 0.0f && m <= 3.402823466e38f) ? fp8_max / m : 1.0f;
  }
  if (scale_inv_out && blockIdx.x == 0 && threadIdx.x == 0) *scale_inv_out = 1.0f / scale;
  RCWorker<OP, VA, VB, false, false, FAST> w;
  w.setup(p, lds_raw, threadIdx.x);
  const TileRange r = tile_range(p.ntiles, blockIdx.x, gridDim.x, p.grid_order);
  w.run(p, kp, A, B, uniform(r.first), uniform(r.end), uniform(r.step), alpha, scale);
  if (amax_out) block_amax_commit<kRcThreads>(w.amax, amax_out);
}

}  // namespace ttc
