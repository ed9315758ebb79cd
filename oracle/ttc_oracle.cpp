// TEST INFRASTRUCTURE ONLY — the oracle.
//
// A CPU restatement of the scalar reference loop nest that HPAC/TTC generates
// for  B[pi(i)] = alpha * A[i] + beta * B[pi(i)]  (column-major, index 0 is
// stride-1).  Only tests/, the package smoke check and bench.py's CPU-baseline
// leg may load this library, and only as the checker or the timed CPU
// baseline; the product path (libttc_hip.so) never calls it.
//
// What each function follows in the reference:
//   * loop nest, loop order (perm reversed: the innermost loop runs over
//     i_{perm[0]}, i.e. B's stride-1 index) and the OpenMP
//     "parallel for collapse(max(1,dim-1))" pragma:
//       ttc/transpose.py:141-156, 570-573; configured by
//       ttc/transposeGenerator.py:174-180 (reference=1, loopPerm=perm[::-1]).
//   * element update "B = alpha*A + beta*B" / "B = alpha*A" when beta == 0
//     (B is not read):                       ttc/transpose.py:126-138.
//   * alpha has A's real type, beta has B's: ttc/transpose.py:402-407.
//   * lda/ldb -> strides, NULL means dense:  ttc/transpose.py:443-473.
//   * offsets offA = sum_k i_k*sA_k, offB = sum_k i_k*sB_{invperm[k]}:
//       ttc/transpose.py:101-124.
//   * complex types: "real * complex" scales each component (g++ semantics of
//     the generated code's C99 complex arithmetic).
// Additions with no counterpart in the reference (new output types):
//   * bf16 / fp16 outputs: the update is evaluated in fp32, then rounded to
//     nearest-even once (IEEE 754 binary16; bfloat16 = top half of binary32).
//   * fp8 e4m3fn / e5m2 outputs (OCP OFP8 v1.0): q = sat_rne(fl32(fl32(alpha*a)
//     * scale)); amax = max |fl32(alpha*a)| over the tensor.
// Build flags: -O3 -ffp-contract=off -fopenmp (no FMA contraction, IEEE).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <omp.h>

namespace {

enum { T_S = 0, T_D = 1, T_C = 2, T_Z = 3, T_BF16 = 4, T_F16 = 5, T_E4M3 = 6, T_E5M2 = 7 };

struct cf { float re, im; };
struct cd { double re, im; };

// ---------------------------------------------------------------- rounding
static inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// binary32 -> bfloat16, round to nearest even; NaN stays NaN (quiet bit set).
static inline uint16_t f32_to_bf16(float x) {
  uint32_t u = f2u(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  uint32_t lsb = (u >> 16) & 1u;
  return (uint16_t)((u + 0x7fffu + lsb) >> 16);
}
static inline float bf16_to_f32(uint16_t h) { return u2f((uint32_t)h << 16); }

// Generic binary32 -> small binary format (1 sign bit, E exponent bits, M
// mantissa bits, exponent bias BIAS), round to nearest even, with gradual
// underflow.  Written from the IEEE 754 definition: the result is the
// representable value nearest |x| (ties to even mantissa), computed by scaling
// |x| by the quantum of its binade (an exact power of two) and rounding to an
// integer with nearbyint() in the default rounding mode.
//   has_inf: format has Inf (fp16, e5m2); e4m3fn has none.
//   sat:     clamp finite overflow to the largest finite value.
//   max_bits: encoding (without sign) of the largest finite value.
//   nan_bits: encoding (without sign) used for NaN.
static inline uint32_t round_small(float x, int E, int M, int BIAS, bool has_inf, bool sat,
                                   uint32_t max_bits, uint32_t nan_bits) {
  uint32_t u = f2u(x);
  uint32_t sign = (u >> 31) << (E + M);
  float ax = std::fabs(x);
  if (std::isnan(x)) return sign | nan_bits;
  const int emin = 1 - BIAS;                     // exponent of the smallest normal
  // largest finite value
  int emax_field = (int)(max_bits >> M);
  uint32_t mmax = max_bits & ((1u << M) - 1);
  double maxv = std::ldexp(1.0 + (double)mmax / (double)(1u << M), emax_field - BIAS);
  if (std::isinf(x)) {
    if (sat) return sign | max_bits;
    return has_inf ? (sign | (((1u << E) - 1) << M)) : (sign | nan_bits);
  }
  int e;
  if (ax == 0.0f) return sign;
  std::frexp(ax, &e);                            // ax = f * 2^e, f in [0.5,1)
  int ex = e - 1;                                // ax in [2^ex, 2^(ex+1))
  if (ex < emin) ex = emin;                      // subnormal range: fixed quantum
  double q = std::ldexp(1.0, ex - M);            // quantum of this binade
  double r = std::nearbyint((double)ax / q) * q; // exact: ax/q < 2^(M+1)
  if (r > maxv) {
    // above the largest finite value after rounding
    if (sat) return sign | max_bits;
    // RNE overflow threshold: values >= maxv + half-quantum(top binade) round to Inf
    return has_inf ? (sign | (((1u << E) - 1) << M)) : (sign | nan_bits);
  }
  if (r == 0.0) return sign;
  int er;
  double fr = std::frexp(r, &er);                // r = fr*2^er
  int eb = er - 1;                               // r in [2^eb, 2^(eb+1))
  uint32_t field, mant;
  if (eb < emin) {                               // subnormal
    field = 0;
    mant = (uint32_t)std::ldexp(r, M - emin);
  } else {
    field = (uint32_t)(eb + BIAS);
    mant = (uint32_t)std::ldexp(fr * 2.0 - 1.0, M);
  }
  return sign | (field << M) | mant;
}
static inline uint16_t f32_to_f16(float x) {
  return (uint16_t)round_small(x, 5, 10, 15, true, false, 0x7bffu, 0x7e00u);
}
static inline uint8_t f32_to_e4m3(float x) {  // saturating, OCP e4m3fn
  return (uint8_t)round_small(x, 4, 3, 7, false, true, 0x7eu, 0x7fu);
}
static inline uint8_t f32_to_e5m2(float x) {  // saturating, OCP e5m2
  return (uint8_t)round_small(x, 5, 2, 15, true, true, 0x7bu, 0x7fu);
}
static inline float f16_to_f32(uint16_t h) {
  uint32_t s = (h >> 15) & 1u, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  float v;
  if (e == 0) v = std::ldexp((float)m, -24);
  else if (e == 31) v = m ? NAN : INFINITY;
  else v = std::ldexp((float)(m | 0x400u), (int)e - 25);
  return s ? -v : v;
}

// ---------------------------------------------------------------- strides
struct Geo {
  int dim;
  int64_t n[16];
  int64_t sA[16];   // stride of A-dim k in A
  int64_t sB[16];   // stride of A-dim k in B  (= B stride of B-position invperm[k])
  int perm[16];
};

static int make_geo(Geo& g, int dim, const int* perm, const int* size, const int* lda, const int* ldb) {
  if (dim < 1 || dim > 16) return 1;
  g.dim = dim;
  int seen[16] = {0};
  for (int k = 0; k < dim; ++k) {
    if (perm[k] < 0 || perm[k] >= dim || seen[perm[k]]) return 2;
    seen[perm[k]] = 1;
    g.perm[k] = perm[k];
    g.n[k] = size[k];
  }
  // ttc/transpose.py:449-460: lda1 = lda[0] (or size0), lda_k = lda[k-1]*lda_{k-1}
  int64_t strideA[16], strideBpos[16];
  strideA[0] = 1;
  for (int k = 1; k < dim; ++k) strideA[k] = strideA[k - 1] * (lda ? lda[k - 1] : size[k - 1]);
  // ttc/transpose.py:462-473: ldb1 = ldb[0] (or size[perm[0]]), ldb_k = ldb[k-1]*ldb_{k-1}
  strideBpos[0] = 1;
  for (int k = 1; k < dim; ++k) strideBpos[k] = strideBpos[k - 1] * (ldb ? ldb[k - 1] : size[perm[k - 1]]);
  for (int k = 0; k < dim; ++k) g.sA[k] = strideA[k];
  for (int q = 0; q < dim; ++q) g.sB[perm[q]] = strideBpos[q];
  return 0;
}

// Visit every element in the reference's loop order: outermost loop over
// i_{perm[dim-1]}, innermost over i_{perm[0]}.  The outer dim-1 loops are
// collapsed into one parallel loop, as "#pragma omp parallel for
// collapse(max(1,dim-1))" does; for dim == 1 the single loop is parallel.
template <class F>
static void visit(const Geo& g, int nthreads, F&& f) {
  const int d = g.dim;
  if (d == 1) {
    const int64_t n0 = g.n[0];
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n0; ++i) f(i * g.sA[0], i * g.sB[0]);
    return;
  }
  int64_t outer = 1;
  for (int q = 1; q < d; ++q) outer *= g.n[g.perm[q]];
  const int p0 = g.perm[0];
  const int64_t n_in = g.n[p0], sa_in = g.sA[p0], sb_in = g.sB[p0];
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t o = 0; o < outer; ++o) {
    // decompose o over B positions 1..d-1 (position 1 fastest)
    int64_t rem = o, offA = 0, offB = 0;
    for (int q = 1; q < d; ++q) {
      const int k = g.perm[q];
      const int64_t i = rem % g.n[k];
      rem /= g.n[k];
      offA += i * g.sA[k];
      offB += i * g.sB[k];
    }
    for (int64_t i = 0; i < n_in; ++i) f(offA + i * sa_in, offB + i * sb_in);
  }
}

// ---------------------------------------------------------------- updates
// Real updates, exactly the C expression of the generated code with the
// declared types (no excess precision on x86-64 SSE, contraction off).
template <class TA, class TB, class AL, class BE>
static void run_real(const Geo& g, const TA* A, TB* B, AL alpha, BE beta, bool bz, int nt) {
  if (bz) visit(g, nt, [&](int64_t oa, int64_t ob) { B[ob] = alpha * A[oa]; });
  else visit(g, nt, [&](int64_t oa, int64_t ob) { B[ob] = alpha * A[oa] + beta * B[ob]; });
}
template <class TA, class TB, class AL, class BE>
static void run_cplx(const Geo& g, const TA* A, TB* B, AL alpha, BE beta, bool bz, int nt) {
  if (bz)
    visit(g, nt, [&](int64_t oa, int64_t ob) {
      B[ob].re = alpha * A[oa].re;
      B[ob].im = alpha * A[oa].im;
    });
  else
    visit(g, nt, [&](int64_t oa, int64_t ob) {
      B[ob].re = alpha * A[oa].re + beta * B[ob].re;
      B[ob].im = alpha * A[oa].im + beta * B[ob].im;
    });
}

template <class LoadA>
static void run_to_half(const Geo& g, LoadA loadA, void* Bv, int tB, float alpha, float beta, bool bz, int nt) {
  uint16_t* B = (uint16_t*)Bv;
  if (tB == T_BF16) {
    visit(g, nt, [&](int64_t oa, int64_t ob) {
      float x = alpha * loadA(oa);
      if (!bz) { float y = beta * bf16_to_f32(B[ob]); x = x + y; }
      B[ob] = f32_to_bf16(x);
    });
  } else {
    visit(g, nt, [&](int64_t oa, int64_t ob) {
      float x = alpha * loadA(oa);
      if (!bz) { float y = beta * f16_to_f32(B[ob]); x = x + y; }
      B[ob] = f32_to_f16(x);
    });
  }
}

template <class LoadA>
static void run_from_half_to_f32(const Geo& g, LoadA loadA, float* B, float alpha, float beta, bool bz, int nt) {
  if (bz) visit(g, nt, [&](int64_t oa, int64_t ob) { B[ob] = alpha * loadA(oa); });
  else visit(g, nt, [&](int64_t oa, int64_t ob) { float x = alpha * loadA(oa); float y = beta * B[ob]; B[ob] = x + y; });
}

}  // namespace

extern "C" {

// Returns 0 on success, non-zero for invalid arguments / unsupported types.
int ttc_oracle_transpose(int tA, int tB, const void* A, void* B, double alpha, double beta, int beta_is_zero,
                         int dim, const int* perm, const int* size, const int* lda, const int* ldb, int nthreads) {
  Geo g;
  if (int e = make_geo(g, dim, perm, size, lda, ldb)) return 10 + e;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  const bool bz = beta_is_zero != 0;
  const float af = (float)alpha, bf = (float)beta;
  const double ad = alpha, bd = beta;
  if (tA == T_S && tB == T_S) run_real(g, (const float*)A, (float*)B, af, bf, bz, nthreads);
  else if (tA == T_D && tB == T_D) run_real(g, (const double*)A, (double*)B, ad, bd, bz, nthreads);
  else if (tA == T_S && tB == T_D) run_real(g, (const float*)A, (double*)B, af, bd, bz, nthreads);
  else if (tA == T_D && tB == T_S) run_real(g, (const double*)A, (float*)B, ad, bf, bz, nthreads);
  else if (tA == T_C && tB == T_C) run_cplx(g, (const cf*)A, (cf*)B, af, bf, bz, nthreads);
  else if (tA == T_Z && tB == T_Z) run_cplx(g, (const cd*)A, (cd*)B, ad, bd, bz, nthreads);
  else if (tA == T_C && tB == T_Z) run_cplx(g, (const cf*)A, (cd*)B, af, bd, bz, nthreads);
  else if (tA == T_Z && tB == T_C) run_cplx(g, (const cd*)A, (cf*)B, ad, bf, bz, nthreads);
  else if (tA == T_S && (tB == T_BF16 || tB == T_F16)) {
    const float* a = (const float*)A;
    run_to_half(g, [&](int64_t o) { return a[o]; }, B, tB, af, bf, bz, nthreads);
  } else if (tA == T_BF16 && (tB == T_BF16 || tB == T_F16)) {
    const uint16_t* a = (const uint16_t*)A;
    run_to_half(g, [&](int64_t o) { return bf16_to_f32(a[o]); }, B, tB, af, bf, bz, nthreads);
  } else if (tA == T_F16 && (tB == T_BF16 || tB == T_F16)) {
    const uint16_t* a = (const uint16_t*)A;
    run_to_half(g, [&](int64_t o) { return f16_to_f32(a[o]); }, B, tB, af, bf, bz, nthreads);
  } else if (tA == T_BF16 && tB == T_S) {
    const uint16_t* a = (const uint16_t*)A;
    run_from_half_to_f32(g, [&](int64_t o) { return bf16_to_f32(a[o]); }, (float*)B, af, bf, bz, nthreads);
  } else if (tA == T_F16 && tB == T_S) {
    const uint16_t* a = (const uint16_t*)A;
    run_from_half_to_f32(g, [&](int64_t o) { return f16_to_f32(a[o]); }, (float*)B, af, bf, bz, nthreads);
  } else {
    return 2;
  }
  return 0;
}

// fp8 cast-transpose: B[pi(i)] = sat_rne_fp8(fl32(fl32(alpha*a)*scale)),
// *amax_out = max |fl32(alpha*a)| (NaN ignored, as fmaxf does).
int ttc_oracle_cast_fp8(int tA, int tB, const void* A, void* Bv, double alpha, float scale, float* amax_out,
                        int dim, const int* perm, const int* size, const int* lda, const int* ldb, int nthreads) {
  Geo g;
  if (int e = make_geo(g, dim, perm, size, lda, ldb)) return 10 + e;
  if (tB != T_E4M3 && tB != T_E5M2) return 2;
  if (tA != T_S && tA != T_BF16 && tA != T_F16) return 2;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  uint8_t* B = (uint8_t*)Bv;
  const float af = (float)alpha;
  float amax = 0.0f;
  const float* as = (const float*)A;
  const uint16_t* ah = (const uint16_t*)A;
  auto load = [&](int64_t o) -> float {
    if (tA == T_S) return as[o];
    if (tA == T_BF16) return bf16_to_f32(ah[o]);
    return f16_to_f32(ah[o]);
  };
  // amax in its own sweep; max is exact, so the order of the reduction does
  // not matter.  One partial per thread (padded against false sharing).
  {
    float* part = new float[(size_t)nthreads * 16]();
    visit(g, nthreads, [&](int64_t oa, int64_t) {
      const float ax = std::fabs(af * load(oa));
      float& p = part[(size_t)omp_get_thread_num() * 16];
      p = std::fmax(p, ax);
    });
    for (int t = 0; t < nthreads; ++t) amax = std::fmax(amax, part[(size_t)t * 16]);
    delete[] part;
  }
  if (tB == T_E4M3)
    visit(g, nthreads, [&](int64_t oa, int64_t ob) { float x = af * load(oa); B[ob] = f32_to_e4m3(x * scale); });
  else
    visit(g, nthreads, [&](int64_t oa, int64_t ob) { float x = af * load(oa); B[ob] = f32_to_e5m2(x * scale); });
  if (amax_out) *amax_out = amax;
  return 0;
}

// Elementwise conversions, for checking the rounding routines against an
// independent implementation (torch on the CPU).
//   op 0: f32 -> bf16, 1: f32 -> f16, 2: f32 -> e4m3 (sat), 3: f32 -> e5m2 (sat)
int ttc_oracle_convert(int op, const float* in, void* out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    switch (op) {
      case 0: ((uint16_t*)out)[i] = f32_to_bf16(in[i]); break;
      case 1: ((uint16_t*)out)[i] = f32_to_f16(in[i]); break;
      case 2: ((uint8_t*)out)[i] = f32_to_e4m3(in[i]); break;
      case 3: ((uint8_t*)out)[i] = f32_to_e5m2(in[i]); break;
      default: return 2;
    }
  }
  return 0;
}

// Write-sweep over a scratch buffer larger than the last-level cache, as the
// reference harness's trashCache() does between repetitions
// (ttc/transposeGenerator.py:1493-1502).
double ttc_oracle_trash_cache(double* buf, int64_t n, int nthreads) {
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  double s = 0.0;
#pragma omp parallel for num_threads(nthreads) reduction(+ : s)
  for (int64_t i = 0; i < n; ++i) {
    buf[i] = 1.0001 * buf[i] + 1e-9;
    s += buf[i];
  }
  return s;
}

int ttc_oracle_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
