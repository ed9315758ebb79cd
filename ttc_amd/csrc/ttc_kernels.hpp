// Device code of libttc_hip.so: one LDS-staged tile kernel family for every
// transposition shape on gfx950 (CDNA4, wave64).
//
// What the reference generates per shape (ttc/CUDAtranspose.py) — a 32x32
// LDS tile for perm[0] != 0 (289-323), remainder tiles (326-388), a strided
// row loop for perm[0] == 0 (179-243) and a plain copy for the identity
// (781-799) — is one kernel here.  A work-group stages a tile that covers
// whole runs of A's leading dimensions on the read side and of B's leading
// dimensions on the write side, so both the HBM read and the HBM write are
// contiguous runs, whatever the rank:
//   * read phase: lane-consecutive vectors walk the tile in A order
//     (coalesced loads), alpha is applied in A's type (as the reference does
//     before its shared-memory store, ttc/CUDAtranspose.py:907), the values
//     land in LDS at their A-order position;
//   * write phase: lane-consecutive vectors walk the tile in B order, gather
//     from LDS, apply the beta epilogue in B's type and store coalesced
//     (non-temporal when beta == 0: B is not read then).
// Every thread precomputes, once per kernel, the tile-relative A offset, B
// offset and LDS index of each vector it owns; a tile then costs one 64-bit
// base per side.  Work-groups are persistent and walk the tile grid.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ttc_hip.h"
#include "ttc_params.h"

namespace ttc {

// Kernel modes: beta != 0, beta == 0, beta == 0 with non-temporal stores,
// fp8 cast-transpose.
enum : int { kModeBeta = 0, kModeBz = 1, kModeBzNt = 2, kModeFp8 = 3 };

// ----------------------------------------------------------------- types
struct alignas(8) c32 { float re, im; };
struct alignas(16) c64 { double re, im; };
struct bf16s { uint16_t u; };
struct f16s { uint16_t u; };
struct e4m3s { uint8_t u; };
struct e5m2s { uint8_t u; };

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f) { return (__umulhi(n, f.m) + n) >> f.s; }

// ------------------------------------------------------------ conversions
__device__ __forceinline__ float to_f32(float a) { return a; }
__device__ __forceinline__ float to_f32(bf16s a) { return __uint_as_float((uint32_t)a.u << 16); }
__device__ __forceinline__ float to_f32(f16s a) {
  _Float16 h;
  __builtin_memcpy(&h, &a.u, 2);
  return (float)h;
}

// binary32 -> bfloat16, round to nearest even; NaN stays NaN (quiet bit set).
__device__ __forceinline__ bf16s to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  bf16s r;
  if ((u & 0x7fffffffu) > 0x7f800000u) {
    r.u = (uint16_t)((u >> 16) | 0x40u);
  } else {
    r.u = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
  }
  return r;
}
// binary32 -> binary16 with the hardware conversion (round to nearest even,
// gradual underflow, overflow to Inf).
__device__ __forceinline__ f16s to_f16(float x) {
  // x is a rounded binary32 value: keep the conversion its own instruction,
  // or the compiler may fold the multiply that produced x into a mixed
  // multiply-convert (v_fma_mixlo_f16) that rounds only once.
  asm volatile("" : "+v"(x));
  _Float16 h = (_Float16)x;
  f16s r;
  __builtin_memcpy(&r.u, &h, 2);
  return r;
}

// Two binary32 values -> OCP fp8 (e4m3fn, or e5m2 with E5), round to
// nearest even with gradual underflow, saturating to +-448 (e5m2: +-57344,
// Inf included), NaN -> sign | 0x7f; byte 0 from a, byte 1 from b.  The
// hardware conversion rounds; the clamp in front of it gives the saturation,
// and NaN is fixed up after it.  Over all 2^32 binary32 inputs this equals
// rounding the mantissa to 3 (e5m2: 2) bits in the integer domain with the
// exponent rebiased, and m * 2^-9 (e5m2: 2^-16) with m = rne(|x| * 2^9
// (2^16)) below the normal range.
template <bool E5> __device__ __forceinline__ uint32_t fp8_pair(float a, float b) {
  constexpr float lim = E5 ? 57344.0f : 448.0f;
  const float ca = __builtin_amdgcn_fmed3f(a, -lim, lim), cb = __builtin_amdgcn_fmed3f(b, -lim, lim);
  const uint32_t r = E5 ? (uint32_t)__builtin_amdgcn_cvt_pk_bf8_f32(ca, cb, 0, false)
                        : (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(ca, cb, 0, false);
  uint32_t qa = r & 0xffu, qb = (r >> 8) & 0xffu;
  if (a != a) qa = ((__float_as_uint(a) >> 24) & 0x80u) | 0x7fu;
  if (b != b) qb = ((__float_as_uint(b) >> 24) & 0x80u) | 0x7fu;
  return qa | (qb << 8);
}

template <class T> __device__ __forceinline__ T from_f32(float x);
template <> __device__ __forceinline__ float from_f32<float>(float x) { return x; }
template <> __device__ __forceinline__ bf16s from_f32<bf16s>(float x) { return to_bf16(x); }
template <> __device__ __forceinline__ f16s from_f32<f16s>(float x) { return to_f16(x); }

// ---------------------------------------------------------------- epilogues
// Each Op fixes: the LDS type L (alpha * a in A's type), the real types of
// alpha (AL) and beta (BE), prod() applied in the read phase and out()/outb()
// applied in the write phase.  The expressions are the C expressions of the
// reference's generated code with its declared types (ttc/transpose.py:
// 126-138 and 402-407; GPU: ttc/CUDAtranspose.py:801-961), evaluated without
// contraction (the library is built with -ffp-contract=off).
template <class TA_, class TB_> struct OpF32 {  // A, B in {float, bf16, f16}: arithmetic in fp32
  using TA = TA_; using TB = TB_; using L = float; using AL = float; using BE = float;
  __device__ static L prod(AL al, TA a) { return al * to_f32(a); }
  __device__ static TB out(L x) { return from_f32<TB>(x); }
  __device__ static TB outb(L x, BE be, TB b) { const float y = be * to_f32(b); return from_f32<TB>(x + y); }
};
struct OpDD {
  using TA = double; using TB = double; using L = double; using AL = double; using BE = double;
  __device__ static L prod(AL al, TA a) { return al * a; }
  __device__ static TB out(L x) { return x; }
  __device__ static TB outb(L x, BE be, TB b) { const double y = be * b; return x + y; }
};
struct OpSD {  // A float, B double: fl64((double)fl32(al*a) + fl64(be*b))
  using TA = float; using TB = double; using L = float; using AL = float; using BE = double;
  __device__ static L prod(AL al, TA a) { return al * a; }
  __device__ static TB out(L x) { return (double)x; }
  __device__ static TB outb(L x, BE be, TB b) { const double y = be * b; return (double)x + y; }
};
struct OpDS {  // A double, B float: fl32(fl64(fl64(al*a) + (double)fl32(be*b)))
  using TA = double; using TB = float; using L = double; using AL = double; using BE = float;
  __device__ static L prod(AL al, TA a) { return al * a; }
  __device__ static TB out(L x) { return (float)x; }
  __device__ static TB outb(L x, BE be, TB b) { const float y = be * b; return (float)(x + (double)y); }
};
// Complex: real alpha/beta scale each component (g++ "real * complex" in the
// generated CPU code), so the complex ops are the real ops per component.
template <class R, class CA, class CB, class CL> struct OpCplx {
  using TA = CA; using TB = CB; using L = CL; using AL = typename R::AL; using BE = typename R::BE;
  __device__ static L prod(AL al, TA a) { L r; r.re = R::prod(al, a.re); r.im = R::prod(al, a.im); return r; }
  __device__ static TB out(L x) { TB r; r.re = R::out(x.re); r.im = R::out(x.im); return r; }
  __device__ static TB outb(L x, BE be, TB b) {
    TB r; r.re = R::outb(x.re, be, b.re); r.im = R::outb(x.im, be, b.im); return r;
  }
};
using OpCC = OpCplx<OpF32<float, float>, c32, c32, c32>;
using OpZZ = OpCplx<OpDD, c64, c64, c64>;
using OpCZ = OpCplx<OpSD, c32, c64, c32>;
using OpZC = OpCplx<OpDS, c64, c32, c64>;

// fp8 outputs: q = sat_rne(fl32(fl32(al*a) * scale)), converted in pairs.
template <class TA_, class TB_> struct OpF8 {
  using TA = TA_; using TB = TB_; using L = float; using AL = float; using BE = float;
  static constexpr bool E5 = __is_same(TB, e5m2s);
  __device__ static L prod(AL al, TA a) { return al * to_f32(a); }
  __device__ static TB outs(L x, float scale) {
    TB r;
    const float y = x * scale;
    r.u = (uint8_t)fp8_pair<E5>(y, y);
    return r;
  }
  __device__ static void outs2(L x0, L x1, float scale, TB& o0, TB& o1) {
    const uint32_t q = fp8_pair<E5>(x0 * scale, x1 * scale);
    o0.u = (uint8_t)q;
    o1.u = (uint8_t)(q >> 8);
  }
};

template <class OP> struct is_fp8 { static constexpr bool value = false; };
template <class A, class B> struct is_fp8<OpF8<A, B>> { static constexpr bool value = true; };

// ----------------------------------------------------------- vector access
template <int BYTES> struct Raw;
template <> struct Raw<1> { using T = uint8_t; };
template <> struct Raw<2> { using T = uint16_t; };
template <> struct Raw<4> { using T = uint32_t; };
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
template <> struct Raw<8> { using T = u32x2; };
template <> struct Raw<16> { using T = u32x4; };

template <class T, int N> __device__ __forceinline__ void vload(T (&dst)[N], const T* p) {
  using R = typename Raw<sizeof(T) * N>::T;
  const R r = *reinterpret_cast<const R*>(p);
  __builtin_memcpy(&dst[0], &r, sizeof(R));
}
template <class T, int N, bool NT> __device__ __forceinline__ void vstore(T* p, const T (&src)[N]) {
  using R = typename Raw<sizeof(T) * N>::T;
  R r;
  __builtin_memcpy(&r, &src[0], sizeof(R));
  if constexpr (NT) __builtin_nontemporal_store(r, reinterpret_cast<R*>(p));
  else *reinterpret_cast<R*>(p) = r;
}

// ------------------------------------------------------------ buffer access
// Buffer resources: the base address lives in scalar registers and each lane
// supplies a 32-bit byte offset.  An offset at or above the record count
// (2^31) fails the hardware range check: the load returns zeros and the store
// is dropped, so edge lanes are predicated without branches and every memory
// instruction is issued unconditionally (which keeps the compiler's vmcnt
// accounting exact across the software pipeline).  0x00020000 is word 3 of a
// raw (untyped, unswizzled) buffer resource on the gfx9 family.
using BufRes = __amdgpu_buffer_rsrc_t;
constexpr uint32_t kBufOOB = 0x80000000u;

__device__ __forceinline__ BufRes make_res(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)kBufOOB, 0x00020000);
}

template <class T, int N> __device__ __forceinline__ void bload(T (&dst)[N], BufRes r, uint32_t off) {
  constexpr int kB = (int)sizeof(T) * N;
  using R = typename Raw<kB>::T;
  R v;
  if constexpr (kB == 1) v = __builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0);
  else if constexpr (kB == 2) v = __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
  else if constexpr (kB == 4) v = __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
  else if constexpr (kB == 8) v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
  else v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
  __builtin_memcpy(&dst[0], &v, sizeof(R));
}

// NT: non-temporal (streaming) store.
template <class T, int N, bool NT> __device__ __forceinline__ void bstore(BufRes r, uint32_t off, const T (&src)[N]) {
  constexpr int kB = (int)sizeof(T) * N;
  constexpr int kAux = NT ? 2 : 0;
  using R = typename Raw<kB>::T;
  R v;
  __builtin_memcpy(&v, &src[0], sizeof(R));
  if constexpr (kB == 1) __builtin_amdgcn_raw_buffer_store_b8(v, r, off, 0, kAux);
  else if constexpr (kB == 2) __builtin_amdgcn_raw_buffer_store_b16(v, r, off, 0, kAux);
  else if constexpr (kB == 4) __builtin_amdgcn_raw_buffer_store_b32(v, r, off, 0, kAux);
  else if constexpr (kB == 8) __builtin_amdgcn_raw_buffer_store_b64(v, r, off, 0, kAux);
  else __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, kAux);
}

// Where one tile of an op sits (host-precomputed for grouped launches):
// element offsets of its origin in A and B and its extent along the split
// dimensions (0xffff when the op has no such dimension).
struct TileBase {
  int64_t a, b;
  uint32_t lim0, lim1;
};

// ------------------------------------------------------------- the tile
// VA / VB: elements per lane vector on the read / write side.  Vectors never
// cross a tile row: the planner guarantees t[0] % VA == 0 and
// t[perm[0]] % VB == 0 (and matching strides / extents for alignment).
template <class OP, int VA, int VB, bool BETA, bool NT>
struct TileWorker {
  using TA = typename OP::TA;
  using TB = typename OP::TB;
  using L = typename OP::L;
  static constexpr int NSA = slots_for(VA, (int)sizeof(TA));
  static constexpr int NSB = slots_for(VB, (int)sizeof(TB));
  static constexpr bool FP8 = is_fp8<OP>::value;

  // Per slot: the tile-relative byte offset in A (resp. B) — unsigned 32-bit,
  // so that loads and stores use the scalar-base + 32-bit vector-offset
  // address form — and the LDS index, two 16-bit indices per register (LDS
  // holds < 65536 elements).
  uint32_t offA[NSA];
  uint32_t ldsA2[(NSA + 1) / 2];
  uint32_t offB[NSB];
  uint32_t ldsB2[(NSB + 1) / 2];
  int nA, nB;  // slots this thread owns (the last may be partially used)
  int tid;

  __device__ __forceinline__ int lds_a(int s) const { return (int)((ldsA2[s >> 1] >> ((s & 1) * 16)) & 0xffffu); }
  __device__ __forceinline__ int lds_b(int s) const { return (int)((ldsB2[s >> 1] >> ((s & 1) * 16)) & 0xffffu); }

  // Per-thread, once per launch (or per op in a grouped launch): decompose
  // this thread's vector positions inside a tile.
  __device__ __forceinline__ void setup(const TileParams& p, int tid_) {
    tid = tid_;
    const int nvA = p.V / VA;
    const int nvB = p.V / VB;
    nA = (nvA - tid + kThreads - 1) / kThreads;
    nB = (nvB - tid + kThreads - 1) / kThreads;
#pragma unroll
    for (int s = 0; s < NSA; ++s) {
      uint32_t l = (uint32_t)(s * kThreads + tid) * VA;
      int32_t oa = 0, ol = 0;
#pragma unroll
      for (int k = 0; k < kMaxDim; ++k) {
        if (k < p.rank) {
          const uint32_t q = fdiv(l, p.tdiv[k]);
          const uint32_t j = l - q * (uint32_t)p.t[k];
          l = q;
          oa += (int32_t)j * p.sa[k];
          ol += (int32_t)j * p.ls[k];
        }
      }
      offA[s] = (uint32_t)oa * (uint32_t)sizeof(TA);
      if (s & 1) ldsA2[s >> 1] |= (uint32_t)ol << 16;
      else ldsA2[s >> 1] = (uint32_t)ol;
    }
#pragma unroll
    for (int s = 0; s < NSB; ++s) {
      uint32_t l = (uint32_t)(s * kThreads + tid) * VB;
      int32_t ob = 0, ol = 0;
#pragma unroll
      for (int q = 0; q < kMaxDim; ++q) {
        if (q < p.rank) {
          const uint32_t qq = fdiv(l, p.tdivB[q]);
          const uint32_t j = l - qq * (uint32_t)p.tB[q];
          l = qq;
          ob += (int32_t)j * p.sbB[q];
          ol += (int32_t)j * p.lsB[q];
        }
      }
      offB[s] = (uint32_t)ob * (uint32_t)sizeof(TB);
      if (s & 1) ldsB2[s >> 1] |= (uint32_t)ol << 16;
      else ldsB2[s >> 1] = (uint32_t)ol;
    }
  }

  // Where a tile sits: base offsets in A and B and its extent along the
  // split dimensions (lim0/lim1; 0xffff when there is no such dimension).
  struct Coord {
    int64_t bA, bB;
    uint32_t lim0, lim1;
    bool full;
  };

  __device__ __forceinline__ Coord coord(const TileParams& p, uint32_t tile_id) const {
    Coord c;
    uint32_t rem = tile_id;
    c.bA = 0;
    c.bB = 0;
    c.lim0 = 0xffffu;
    c.lim1 = 0xffffu;
#pragma unroll
    for (int k = 0; k < kMaxDim; ++k) {
      if (k < p.rank) {
        const uint32_t q = fdiv(rem, p.gdiv[k]);
        const uint32_t x = rem - q * p.gdiv[k].d;
        rem = q;
        c.bA += (int64_t)x * p.tsa[k];
        c.bB += (int64_t)x * p.tsb[k];
        const uint32_t left = (uint32_t)p.n[k] - x * (uint32_t)p.t[k];
        if (k == p.split0) c.lim0 = left;
        if (k == p.split1) c.lim1 = left;
      }
    }
    c.full = c.lim0 >= (uint32_t)p.tsplit0 && c.lim1 >= (uint32_t)p.tsplit1;
    return c;
  }

  // Digit (l / P) mod t of a split dimension in tile-linear index l.
  static __device__ __forceinline__ uint32_t digit(uint32_t l, const FastDiv& P, const FastDiv& t) {
    const uint32_t x = fdiv(l, P);
    return x - fdiv(x, t) * t.d;
  }
  // Whether slot s holds data of the tile at c.  Full tiles need no test;
  // in an edge tile the slot's coordinates along the split dimensions are
  // recomputed from its tile-linear index (cheaper than keeping them live).
  __device__ __forceinline__ bool in_a(const TileParams& p, const Coord& c, int s) const {
    if (s >= nA) return false;
    if (c.full) return true;
    uint32_t l = (uint32_t)(s * kThreads + tid) * VA;
    asm volatile("" : "+v"(l));  // recompute per edge tile; do not hoist out of the tile loop
    return (p.split0 < 0 || digit(l, p.pA0, p.ts0) < c.lim0) && (p.split1 < 0 || digit(l, p.pA1, p.ts1) < c.lim1);
  }
  __device__ __forceinline__ bool in_b(const TileParams& p, const Coord& c, int s) const {
    if (s >= nB) return false;
    if (c.full) return true;
    uint32_t l = (uint32_t)(s * kThreads + tid) * VB;
    asm volatile("" : "+v"(l));  // recompute per edge tile; do not hoist out of the tile loop
    return (p.split0 < 0 || digit(l, p.pB0, p.ts0) < c.lim0) && (p.split1 < 0 || digit(l, p.pB1, p.ts1) < c.lim1);
  }

  // Global -> registers for the tile at c.
  __device__ __forceinline__ void load(const TileParams& p, const Coord& c, const TA* __restrict__ A,
                                       TA (&va)[NSA][VA]) const {
    const char* a = reinterpret_cast<const char*>(A + c.bA);
#pragma unroll
    for (int s = 0; s < NSA; ++s)
      if (in_a(p, c, s)) vload<TA, VA>(va[s], reinterpret_cast<const TA*>(a + offA[s]));
  }

  // Registers -> (alpha) -> LDS.
  __device__ __forceinline__ void stage(const TileParams& p, const Coord& c, const TA (&va)[NSA][VA],
                                        L* __restrict__ lds, typename OP::AL alpha, float& amax) const {
#pragma unroll
    for (int s = 0; s < NSA; ++s) {
      if (in_a(p, c, s)) {
        const int li = lds_a(s);
#pragma unroll
        for (int v = 0; v < VA; ++v) {
          const L x = OP::prod(alpha, va[s][v]);
          if constexpr (FP8) amax = fmaxf(amax, fabsf(x));
          lds[li + v] = x;
        }
      }
    }
  }

  // LDS -> (beta epilogue) -> global.
  __device__ __forceinline__ void store(const TileParams& p, const Coord& c, TB* __restrict__ B,
                                        const L* __restrict__ lds, typename OP::BE beta, float scale) const {
    char* b = reinterpret_cast<char*>(B + c.bB);
    const int S = p.lds_stride_b;
#pragma unroll
    for (int s = 0; s < NSB; ++s) {
      if (in_b(p, c, s)) {
        const int li = lds_b(s);
        TB o[VB];
        if constexpr (FP8) {
          if constexpr (VB == 1) {
            o[0] = OP::outs(lds[li], scale);
          } else {
#pragma unroll
            for (int v = 0; v < VB; v += 2) OP::outs2(lds[li + v * S], lds[li + (v + 1) * S], scale, o[v], o[v + 1]);
          }
        } else if constexpr (BETA) {
          TB ob[VB];
          vload<TB, VB>(ob, reinterpret_cast<const TB*>(b + offB[s]));
#pragma unroll
          for (int v = 0; v < VB; ++v) o[v] = OP::outb(lds[li + v * S], beta, ob[v]);
        } else {
#pragma unroll
          for (int v = 0; v < VB; ++v) o[v] = OP::out(lds[li + v * S]);
        }
        vstore<TB, VB, NT>(reinterpret_cast<TB*>(b + offB[s]), o);
      }
    }
  }

  // The same walk over tiles whose Coord comes from a table (one record per
  // tile, precomputed on the host) instead of being derived from `p`.
  __device__ __forceinline__ Coord coord_of(const TileParams& p, const TileBase& b) const {
    Coord c;
    c.bA = b.a;
    c.bB = b.b;
    c.lim0 = b.lim0;
    c.lim1 = b.lim1;
    c.full = c.lim0 >= (uint32_t)p.tsplit0 && c.lim1 >= (uint32_t)p.tsplit1;
    return c;
  }
  __device__ __forceinline__ void run_table(const TileParams& p, const TA* __restrict__ A, TB* __restrict__ B,
                                            L* __restrict__ lds, const TileBase* __restrict__ bases, uint32_t first,
                                            uint32_t end, typename OP::AL alpha, typename OP::BE beta, float scale,
                                            float& amax) {
    if (first >= end) return;
    TA va[NSA][VA];
    Coord c = coord_of(p, bases[first]);
    load(p, c, A, va);
    for (uint32_t t = first; t < end; ++t) {
      stage(p, c, va, lds, alpha, amax);
      __syncthreads();
      const bool more = t + 1 < end;
      const Coord cn = coord_of(p, bases[more ? t + 1 : t]);
      if (more) load(p, cn, A, va);
      store(p, c, B, lds, beta, scale);
      __syncthreads();
      c = cn;
    }
  }

  // All tiles first, first+step, ... < end.  Software pipelined: the loads
  // of the next tile are issued before the write phase of the current one,
  // so each work-group keeps a tile of reads in flight while it writes.
  __device__ __forceinline__ void run(const TileParams& p, const TA* __restrict__ A, TB* __restrict__ B,
                                      L* __restrict__ lds, uint32_t first, uint32_t end, uint32_t step,
                                      typename OP::AL alpha, typename OP::BE beta, float scale, float& amax) {
    if (first >= end) return;
    TA va[NSA][VA];
    Coord c = coord(p, first);
    load(p, c, A, va);
    for (uint32_t t = first; t < end; t += step) {
      stage(p, c, va, lds, alpha, amax);
      __syncthreads();
      const uint32_t tn = t + step;
      const bool more = tn < end;
      const Coord cn = coord(p, more ? tn : t);
      if (more) load(p, cn, A, va);
      store(p, c, B, lds, beta, scale);
      __syncthreads();
      c = cn;
    }
  }
};

// Tile index walk of work-group `wg` out of `nwg`: strided (tile = wg + i*nwg)
// or contiguous ranges.
struct TileRange {
  uint32_t first, end, step;
};
// Work-groups are dispatched round-robin over the 8 XCDs (work-group w runs
// on XCD w % 8), each XCD with its own L2.  order 2 renumbers them so that
// the work-groups of one XCD take consecutive tiles of each round of the
// strided walk: neighbouring tiles share the cache lines at their edges, and
// those then meet in one L2 instead of being fetched (and partially written
// back) by two.
constexpr uint32_t kXcds = 8;

__device__ __forceinline__ TileRange tile_range(uint32_t ntiles, uint32_t wg, uint32_t nwg, int order) {
  TileRange r;
  if (order == 3 && nwg % kXcds == 0) {
    // XCD x (work-groups wg % kXcds == x) owns a contiguous 1/kXcds share of
    // the tiles and its work-groups stride through it together: consecutive
    // tiles still meet in one L2, and the last, partial round is spread over
    // every XCD instead of landing on the first ones.
    const uint32_t x = wg % kXcds, j = wg / kXcds, per = nwg / kXcds;
    const uint32_t q = ntiles / kXcds, rem = ntiles % kXcds;
    const uint32_t lo = x * q + (x < rem ? x : rem);
    r.first = lo + j; r.end = lo + q + (x < rem ? 1u : 0u); r.step = per;
    return r;
  }
  if (order == 0 || order == 2 || order == 3) {
    if (order == 2 && nwg % kXcds == 0) wg = (wg % kXcds) * (nwg / kXcds) + wg / kXcds;
    r.first = wg; r.end = ntiles; r.step = nwg;
  } else {
    const uint32_t per = ntiles / nwg, extra = ntiles % nwg;
    r.first = wg * per + (wg < extra ? wg : extra);
    r.end = r.first + per + (wg < extra ? 1u : 0u);
    r.step = 1;
  }
  return r;
}

// Work-groups of the row x column walk with at least this many tiles step
// their wave priority down as they advance: those dispatched later onto a CU
// lose issue arbitration to the older ones beside them, and the falling
// priority lets them catch up.  With fewer tiles per work-group the schedule
// measured slower, so it stays off.
constexpr uint32_t kPrioMinRounds = 8;

// Block-wide max of a non-negative float, then one atomicMax on its bits
// (non-negative IEEE floats order like their unsigned bit patterns); NT is
// the work-group size.  The value only grows during a launch, so a block
// whose max does not exceed the one already there skips its atomic: the
// work-groups of a launch end together, and their same-address atomics
// would queue at one L2 channel.
template <int NT = kThreads>
__device__ __forceinline__ void block_amax_commit(float v, float* out) {
  __shared__ float red[NT / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = red[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) m = fmaxf(m, red[w]);
    unsigned int* const o = reinterpret_cast<unsigned int*>(out);
    const unsigned int bits = __float_as_uint(m);
    if (bits > __hip_atomic_load(o, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(o, bits);
  }
}

// --------------------------------------------------------------- kernels
template <class OP, int VA, int VB, bool BETA, bool NT>
__global__ void __launch_bounds__(kThreads) transpose_kernel(const TileParams p, const typename OP::TA* __restrict__ A,
                                                             typename OP::TB* __restrict__ B, double alpha,
                                                             double beta) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  using L = typename OP::L;
  L* lds = reinterpret_cast<L*>(lds_raw);
  TileWorker<OP, VA, VB, BETA, NT> w;
  w.setup(p, threadIdx.x);
  const TileRange r = tile_range(p.ntiles, blockIdx.x, gridDim.x, p.grid_order);
  float amax = 0.0f;
  w.run(p, A, B, lds, r.first, r.end, r.step, (typename OP::AL)alpha, (typename OP::BE)beta, 1.0f, amax);
}

// fp8 cast-transpose.  The scale comes from *scale_in (delayed scaling),
// from fp8_max / *amax_in (current scaling, amax_in produced by a previous
// pass), or from the host value.
template <class OP, int VA, int VB>
__global__ void __launch_bounds__(kThreads) cast_fp8_kernel(const TileParams p, const typename OP::TA* __restrict__ A,
                                                            typename OP::TB* __restrict__ B, float alpha,
                                                            const float* __restrict__ scale_in,
                                                            const float* __restrict__ amax_in, float fp8_max,
                                                            float scale_host, float* __restrict__ amax_out,
                                                            float* __restrict__ scale_inv_out) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  float* lds = reinterpret_cast<float*>(lds_raw);
  float scale = scale_host;
  if (scale_in) {
    scale = *scale_in;
  } else if (amax_in) {
    const float m = *amax_in;
    scale = (m > 0.0f && m <= 3.402823466e38f) ? fp8_max / m : 1.0f;
  }
  if (scale_inv_out && blockIdx.x == 0 && threadIdx.x == 0) *scale_inv_out = 1.0f / scale;
  TileWorker<OP, VA, VB, false, true> w;
  w.setup(p, threadIdx.x);
  const TileRange r = tile_range(p.ntiles, blockIdx.x, gridDim.x, p.grid_order);
  float amax = 0.0f;
  w.run(p, A, B, lds, r.first, r.end, r.step, alpha, 0.0f, scale, amax);
  if (amax_out) block_amax_commit(amax, amax_out);
}


// amax of alpha*A, current scaling pass 1.  Dense A: 16-byte vector loads
// over the flat array, kAmaxUnroll of them in flight per lane (a grid of two
// work-groups per CU: measured 5.6-5.9 TB/s on 235-537 MB where one load per
// lane over eight work-groups per CU gave 4.1-5.4).  Strided A (padded lda):
// element e of the A-order walk is (i0 = e mod n0, row = e div n0), row
// decomposed over dims 1..rank-1.
constexpr int kAmaxUnroll = 4;
struct AmaxParams {
  int32_t rank;
  int32_t n0;
  int32_t dense;
  FastDiv dn0;
  FastDiv dn[kMaxDim];
  int64_t sa[kMaxDim];
};
template <class TA, int VEC>
__global__ void __launch_bounds__(kThreads) amax_kernel(const AmaxParams q, const TA* __restrict__ A, float alpha,
                                                        float* __restrict__ amax_out, uint64_t N) {
  float m = 0.0f;
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  if (q.dense) {
    const uint64_t vstride = stride * VEC;
    uint64_t e = ((uint64_t)blockIdx.x * kThreads + threadIdx.x) * VEC;
    for (; e + (kAmaxUnroll - 1) * vstride + VEC <= N; e += kAmaxUnroll * vstride) {
      TA v[kAmaxUnroll][VEC];
#pragma unroll
      for (int u = 0; u < kAmaxUnroll; ++u) vload<TA, VEC>(v[u], A + e + u * vstride);
#pragma unroll
      for (int u = 0; u < kAmaxUnroll; ++u)
#pragma unroll
        for (int i = 0; i < VEC; ++i) m = fmaxf(m, fabsf(alpha * to_f32(v[u][i])));
    }
    for (; e < N; e += vstride) {
      if (e + VEC <= N) {
        TA v[VEC];
        vload<TA, VEC>(v, A + e);
#pragma unroll
        for (int i = 0; i < VEC; ++i) m = fmaxf(m, fabsf(alpha * to_f32(v[i])));
      } else {
        for (uint64_t i = e; i < N; ++i) m = fmaxf(m, fabsf(alpha * to_f32(A[i])));
      }
    }
  } else {
    for (uint64_t e = (uint64_t)blockIdx.x * kThreads + threadIdx.x; e < N; e += stride) {
      uint32_t row = fdiv((uint32_t)e, q.dn0);
      const uint32_t i0 = (uint32_t)e - row * (uint32_t)q.n0;
      int64_t off = i0;
#pragma unroll
      for (int k = 1; k < kMaxDim; ++k) {
        if (k < q.rank) {
          const uint32_t qq = fdiv(row, q.dn[k]);
          const uint32_t c = row - qq * q.dn[k].d;
          row = qq;
          off += (int64_t)c * q.sa[k];
        }
      }
      m = fmaxf(m, fabsf(alpha * to_f32(A[off])));
    }
  }
  block_amax_commit(m, amax_out);
}

}  // namespace ttc
