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

template <class OP, int VA, int VB, bool BETA, bool NT>
struct RCWorker {
  using TA = typename OP::TA;
  using TB = typename OP::TB;
  using L = typename OP::L;
  static constexpr int NIA = rc_elems((int)sizeof(TA)) / VA;  // read vectors per thread per tile
  static constexpr int NIB = rc_elems((int)sizeof(TB)) / VB;  // write vectors per thread per tile

  L* data;             // LDS tile
  const int32_t* rowA; // LDS tables
  const int32_t* rowB;
  const int32_t* colA;
  const int32_t* colB;
  int tid;

  struct Coord {
    int64_t bA, bB;
    uint32_t rows, cols;   // this tile's R / C (rows * cols == 0: no tile, every access out of range)
    FastDiv dr, dc;        // division by rows / cols
  };

  // Offset tables into LDS (after the data region), once per launch.
  __device__ __forceinline__ void setup(const RCParams& p, unsigned char* lds_raw, int tid_) {
    tid = tid_;
    data = reinterpret_cast<L*>(lds_raw);
    int32_t* tab = reinterpret_cast<int32_t*>(lds_raw + (size_t)p.lds_data_elems * sizeof(L));
    int32_t* rA = tab;
    int32_t* rB = tab + p.R;
    int32_t* cA = tab + 2 * p.R;
    int32_t* cB = tab + 2 * p.R + p.C;
    for (int i = tid; i < p.R; i += kThreads) {
      uint32_t l = (uint32_t)i;
      int32_t oa = 0, ob = 0;
      for (int m = 0; m < p.na; ++m) {
        const uint32_t q = fdiv(l, p.dta[m]);
        const uint32_t d = l - q * (uint32_t)p.ta[m];
        l = q;
        oa += (int32_t)d * p.saA[m];
        ob += (int32_t)d * p.sbA[m];
      }
      rA[i] = oa;
      rB[i] = ob;
    }
    for (int j = tid; j < p.C; j += kThreads) {
      uint32_t l = (uint32_t)j;
      int32_t oa = 0, ob = 0;
      for (int m = 0; m < p.nb; ++m) {
        const uint32_t q = fdiv(l, p.dtb[m]);
        const uint32_t d = l - q * (uint32_t)p.tb[m];
        l = q;
        oa += (int32_t)d * p.saB[m];
        ob += (int32_t)d * p.sbB[m];
      }
      cA[j] = oa;
      cB[j] = ob;
    }
    rowA = rA;
    rowB = rB;
    colA = cA;
    colB = cB;
    __syncthreads();
  }

  __device__ __forceinline__ Coord coord(const RCParams& p, uint32_t tile_id) const {
    Coord c;
    uint32_t rem = tile_id;
    c.bA = 0;
    c.bB = 0;
    bool lastA = false, lastB = false;
#pragma unroll
    for (int k = 0; k < kMaxDim; ++k) {
      if (k < p.rank) {
        const uint32_t q = fdiv(rem, p.gdiv[k]);
        const uint32_t x = rem - q * p.gdiv[k].d;
        rem = q;
        c.bA += (int64_t)x * p.tsa[k];
        c.bB += (int64_t)x * p.tsb[k];
        const bool partial = (uint32_t)p.n[k] - x * (uint32_t)p.t[k] < (uint32_t)p.t[k];
        if (k == p.splitA) lastA = partial;
        if (k == p.splitB) lastB = partial;
      }
    }
    c.rows = lastA ? (uint32_t)p.Rl : (uint32_t)p.R;
    c.cols = lastB ? (uint32_t)p.Cl : (uint32_t)p.C;
    c.dr = lastA ? p.dRl : p.dR;
    c.dc = lastB ? p.dCl : p.dC;
    return c;
  }

  // Read-side slot `it` of this thread: element (i, j) with i fastest.
  // Lanes past the tile compute (0, 0) and are flagged invalid.
  __device__ __forceinline__ bool slot_a(const Coord& c, int it, uint32_t& i, uint32_t& j) const {
    const uint32_t l = (uint32_t)(it * kThreads + tid) * VA;
    const bool ok = l < c.rows * c.cols;
    const uint32_t lc = ok ? l : 0u;
    j = fdiv(lc, c.dr);
    i = lc - j * c.rows;
    return ok;
  }
  // Write-side slot: element (i, j) with j fastest.
  __device__ __forceinline__ bool slot_b(const Coord& c, int it, uint32_t& i, uint32_t& j) const {
    const uint32_t l = (uint32_t)(it * kThreads + tid) * VB;
    const bool ok = l < c.rows * c.cols;
    const uint32_t lc = ok ? l : 0u;
    i = fdiv(lc, c.dc);
    j = lc - i * c.cols;
    return ok;
  }

  __device__ __forceinline__ void load(const RCParams& p, const Coord& c, const TA* __restrict__ A,
                                       TA (&va)[NIA][VA]) const {
    const BufRes r = make_res(A + c.bA);
#pragma unroll
    for (int it = 0; it < NIA; ++it) {
      uint32_t i, j;
      const bool ok = slot_a(c, it, i, j);
      const uint32_t off = (uint32_t)((p.rowA_dense ? (int32_t)i : rowA[i]) + colA[j]) * (uint32_t)sizeof(TA);
      bload<TA, VA>(va[it], r, ok ? off : kBufOOB);
    }
  }

  __device__ __forceinline__ void stage(const RCParams& p, const Coord& c, const TA (&va)[NIA][VA],
                                        typename OP::AL alpha) const {
#pragma unroll
    for (int it = 0; it < NIA; ++it) {
      uint32_t i, j;
      const bool ok = slot_a(c, it, i, j);
      const uint32_t li = ok ? i + j * (uint32_t)p.RP : (uint32_t)p.lds_spare;
#pragma unroll
      for (int v = 0; v < VA; ++v) data[li + v] = OP::prod(alpha, va[it][v]);
    }
  }

  __device__ __forceinline__ uint32_t off_b(const RCParams& p, uint32_t i, uint32_t j) const {
    return (uint32_t)(rowB[i] + (p.colB_dense ? (int32_t)j : colB[j])) * (uint32_t)sizeof(TB);
  }

  // Old values of B for the beta epilogue.
  __device__ __forceinline__ void load_b(const RCParams& p, const Coord& c, const TB* __restrict__ B,
                                         TB (&ob)[NIB][VB]) const {
    const BufRes r = make_res(B + c.bB);
#pragma unroll
    for (int it = 0; it < NIB; ++it) {
      uint32_t i, j;
      const bool ok = slot_b(c, it, i, j);
      bload<TB, VB>(ob[it], r, ok ? off_b(p, i, j) : kBufOOB);
    }
  }

  __device__ __forceinline__ void store(const RCParams& p, const Coord& c, TB* __restrict__ B,
                                        typename OP::BE beta, const TB (&ob)[NIB][VB]) const {
    const BufRes r = make_res(B + c.bB);
    const uint32_t RP = (uint32_t)p.RP;
#pragma unroll
    for (int it = 0; it < NIB; ++it) {
      uint32_t i, j;
      const bool ok = slot_b(c, it, i, j);
      const uint32_t li = i + j * RP;
      TB o[VB];
#pragma unroll
      for (int v = 0; v < VB; ++v) {
        if constexpr (BETA) o[v] = OP::outb(data[li + v * RP], beta, ob[it][v]);
        else o[v] = OP::out(data[li + v * RP]);
      }
      bstore<TB, VB, NT>(r, ok ? off_b(p, i, j) : kBufOOB, o);
    }
  }

  // Software pipeline over this work-group's tiles: the next tile's reads
  // are in flight while the current tile is written.  Memory operations per
  // iteration, in issue order: old B of tile t (beta only), A of tile t+1,
  // B stores of tile t; staging tile t+1 then waits only for its own reads.
  // The last iteration prefetches an empty tile (all lanes out of range).
  __device__ __forceinline__ void run(const RCParams& p, const TA* __restrict__ A, TB* __restrict__ B,
                                      uint32_t first, uint32_t end, uint32_t step, typename OP::AL alpha,
                                      typename OP::BE beta) {
    if (first >= end) return;
    TA va[NIA][VA];
    TB ob[NIB][VB];
    Coord c = coord(p, first);
    load(p, c, A, va);
    for (uint32_t t = first; t < end; t += step) {
      stage(p, c, va, alpha);
      __syncthreads();
      if constexpr (BETA) load_b(p, c, B, ob);
      const uint32_t tn = t + step;
      Coord cn = coord(p, tn < end ? tn : t);
      if (tn >= end) cn.rows = 0;
      load(p, cn, A, va);
      store(p, c, B, beta, ob);
      __syncthreads();
      c = cn;
    }
  }
};

template <class OP, int VA, int VB, bool BETA, bool NT>
__global__ void __launch_bounds__(kThreads) transpose_rc_kernel(const RCParams p, const typename OP::TA* __restrict__ A,
                                                                typename OP::TB* __restrict__ B, double alpha,
                                                                double beta) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  RCWorker<OP, VA, VB, BETA, NT> w;
  w.setup(p, lds_raw, threadIdx.x);
  const TileRange r = tile_range(p.ntiles, blockIdx.x, gridDim.x, p.grid_order);
  w.run(p, A, B, r.first, r.end, r.step, (typename OP::AL)alpha, (typename OP::BE)beta);
}

}  // namespace ttc
