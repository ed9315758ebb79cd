// Row-copy kernel for permutations that keep dimension 0 in place
// (perm[0] == 0 after index fusion): A and B share the stride-1 dimension,
// so every row of dimension 0 is contiguous on both sides and moves without
// an LDS round trip.  The reference handles this case with its own loop
// (ttc/CUDAtranspose.py:179-243, threads along i0, rows in a cyclic block
// distribution); here each lane moves a V-element vector, the rows are
// walked in B order so that consecutive lanes write consecutive bytes of B,
// and every lane keeps U vectors in flight.
#pragma once
#include "ttc_kernels.hpp"

namespace ttc {

template <class OP, int V, int MODE>
__global__ void __launch_bounds__(kThreads)
    row_kernel(const RowParams p, const typename OP::TA* __restrict__ A, typename OP::TB* __restrict__ B,
               double alpha, double beta, const float* __restrict__ scale_in, const float* __restrict__ amax_in,
               float fp8_max, float scale_host, float* __restrict__ amax_out, float* __restrict__ scale_inv_out) {
  using TA = typename OP::TA;
  using TB = typename OP::TB;
  using L = typename OP::L;
  constexpr bool FP8 = MODE == kModeFp8;
  constexpr bool BETA = MODE == kModeBeta;
  constexpr bool NT = MODE == kModeBzNt;
  constexpr int U = kRowUnroll;  // vectors in flight per lane
  float scale = 1.0f;
  if constexpr (FP8) {
    scale = scale_host;
    if (scale_in) {
      scale = *scale_in;
    } else if (amax_in) {
      const float m = *amax_in;
      scale = (m > 0.0f && m <= 3.402823466e38f) ? fp8_max / m : 1.0f;
    }
    if (scale_inv_out && blockIdx.x == 0 && threadIdx.x == 0) *scale_inv_out = 1.0f / scale;
  }
  const typename OP::AL al = (typename OP::AL)alpha;
  const typename OP::BE be = (typename OP::BE)beta;
  const char* a = reinterpret_cast<const char*>(A);
  char* b = reinterpret_cast<char*>(B);
  const uint32_t stride = gridDim.x * (uint32_t)kThreads;  // the host keeps total + U * stride < 2^32
  float amax = 0.0f;
  for (uint32_t g0 = blockIdx.x * (uint32_t)kThreads + threadIdx.x; g0 < p.total; g0 += U * stride) {
    int64_t ob[U];
    bool ok[U];
    TA va[U][V];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t g = g0 + (uint32_t)u * stride;
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
      TB o[V];
      if constexpr (FP8) {
        L x[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
          x[v] = OP::prod(al, va[u][v]);
          amax = fmaxf(amax, fabsf(x[v]));
        }
        if constexpr (V == 1) {
          o[0] = OP::outs(x[0], scale);
        } else {
#pragma unroll
          for (int v = 0; v < V; v += 2) OP::outs2(x[v], x[v + 1], scale, o[v], o[v + 1]);
        }
      } else if constexpr (BETA) {
#pragma unroll
        for (int v = 0; v < V; ++v) o[v] = OP::outb(OP::prod(al, va[u][v]), be, old[u][v]);
      } else {
#pragma unroll
        for (int v = 0; v < V; ++v) o[v] = OP::out(OP::prod(al, va[u][v]));
      }
      vstore<TB, V, NT>(reinterpret_cast<TB*>(b + ob[u]), o);
    }
  }
  if constexpr (FP8) {
    if (amax_out) block_amax_commit(amax, amax_out);
  }
}

}  // namespace ttc
