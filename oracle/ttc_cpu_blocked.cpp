// TEST / BASELINE INFRASTRUCTURE ONLY — a blocked, vectorised CPU
// transposition: the "TTC-best CPU" baseline beside the scalar oracle.
//
// HPAC/TTC's CPU generator emits, per transposition, a loop nest blocked
// over A's and B's stride-1 indices with an in-register micro-kernel that
// transposes a W x W block (W = 8 floats / 4 doubles with AVX, 16 floats
// with AVX-512: ttc/transposeGenerator.py:846-1474, micro-kernels
// ttc/micro-kernels/transpose_8x8_float_avx.kernel,
// transpose_4x4_double_avx.kernel, transpose_16x16_float_avx512.kernel; loop
// nests ttc/transpose.py:204-285, 496-569), OpenMP-parallel over the outer
// loops.  This file restates that method in code written here:
//   * blocks of W x W elements over (A's dimension 0, B's stride-1
//     dimension perm[0]); W rows of A (each W contiguous elements of
//     dimension 0) are loaded, transposed in registers, and stored as W rows
//     of B (each W contiguous elements of dimension perm[0]);
//   * AVX-512 (when the host has it): a butterfly of 2-source permutes, four
//     rounds for 16 x 16 floats and three for 8 x 8 doubles (each round
//     swaps the off-diagonal s x s sub-blocks of every 2s x 2s block);
//     AVX2 otherwise: the unpack / shuffle / 128-bit-lane permute network
//     for 8 x 8 floats and unpack / lane permute for 4 x 4 doubles;
//   * edges that do not fill a block, and everything when perm[0] == 0 (a
//     scaled row copy), in plain loops the compiler vectorises;
//   * the outer indices and the block indices form one index space split
//     statically over the OpenMP threads.
// Arithmetic per element is the oracle's (ttc/transpose.py:126-138):
// B = fl(alpha*a) when beta == 0 (B not read), else fl(fl(alpha*a) +
// fl(beta*b)), built with -ffp-contract=off, so results are bit-identical to
// the oracle for the same inputs (s and d, same type in A and B).
#include <immintrin.h>
#include <omp.h>

#include <cstdint>
#include <cstring>

namespace {

constexpr int kMaxDim = 16;

struct Geo {
  int r;
  int64_t n[kMaxDim];   // extent of A dimension k
  int64_t sa[kMaxDim];  // A stride of A dimension k
  int64_t sb[kMaxDim];  // B stride of A dimension k
  int p0;               // B's stride-1 dimension (an A dimension)
};

static bool make_geo(int dim, const int* perm, const int* size, const int* lda, const int* ldb, Geo& g) {
  if (dim < 1 || dim > kMaxDim) return false;
  bool seen[kMaxDim] = {false};
  for (int k = 0; k < dim; ++k) {
    if (perm[k] < 0 || perm[k] >= dim || seen[perm[k]] || size[k] < 0) return false;
    seen[perm[k]] = true;
  }
  g.r = dim;
  int64_t s = 1;
  for (int k = 0; k < dim; ++k) {
    g.n[k] = size[k];
    g.sa[k] = s;
    s *= lda ? lda[k] : size[k];
  }
  s = 1;
  for (int q = 0; q < dim; ++q) {
    g.sb[perm[q]] = s;
    s *= ldb ? ldb[q] : size[perm[q]];
  }
  g.p0 = perm[0];
  return true;
}

// ------------------------------------------------------------ micro-kernels
// Each loads W rows of A (row j: W elements of dimension 0 at a + j * lda_p0),
// transposes them, and stores row i (W elements of dimension p0) at
// b + i * ldb_0, with the element update applied in registers.

struct F32 {
  using T = float;
};
struct F64 {
  using T = double;
};

__attribute__((target("avx2"))) static inline __m256 upd8(__m256 x, float alpha, const float* bp, float beta,
                                                         bool bz) {
  __m256 y = _mm256_mul_ps(x, _mm256_set1_ps(alpha));
  if (!bz) y = _mm256_add_ps(y, _mm256_mul_ps(_mm256_loadu_ps(bp), _mm256_set1_ps(beta)));
  return y;
}

__attribute__((target("avx2"))) static void block8x8_f32(const float* a, int64_t lda, float* b, int64_t ldb,
                                                         float alpha, float beta, bool bz) {
  __m256 r0 = _mm256_loadu_ps(a + 0 * lda), r1 = _mm256_loadu_ps(a + 1 * lda);
  __m256 r2 = _mm256_loadu_ps(a + 2 * lda), r3 = _mm256_loadu_ps(a + 3 * lda);
  __m256 r4 = _mm256_loadu_ps(a + 4 * lda), r5 = _mm256_loadu_ps(a + 5 * lda);
  __m256 r6 = _mm256_loadu_ps(a + 6 * lda), r7 = _mm256_loadu_ps(a + 7 * lda);
  // pairs of rows interleaved: (r0 r1) -> a0 b0 a1 b1 | a4 b4 a5 b5, ...
  const __m256 t0 = _mm256_unpacklo_ps(r0, r1), t1 = _mm256_unpackhi_ps(r0, r1);
  const __m256 t2 = _mm256_unpacklo_ps(r2, r3), t3 = _mm256_unpackhi_ps(r2, r3);
  const __m256 t4 = _mm256_unpacklo_ps(r4, r5), t5 = _mm256_unpackhi_ps(r4, r5);
  const __m256 t6 = _mm256_unpacklo_ps(r6, r7), t7 = _mm256_unpackhi_ps(r6, r7);
  // quads: columns c, c+4 of rows 0..3 (and 4..7)
  const __m256 u0 = _mm256_shuffle_ps(t0, t2, 0x44), u1 = _mm256_shuffle_ps(t0, t2, 0xEE);
  const __m256 u2 = _mm256_shuffle_ps(t1, t3, 0x44), u3 = _mm256_shuffle_ps(t1, t3, 0xEE);
  const __m256 u4 = _mm256_shuffle_ps(t4, t6, 0x44), u5 = _mm256_shuffle_ps(t4, t6, 0xEE);
  const __m256 u6 = _mm256_shuffle_ps(t5, t7, 0x44), u7 = _mm256_shuffle_ps(t5, t7, 0xEE);
  // join the 128-bit halves: column c of rows 0..7
  r0 = _mm256_permute2f128_ps(u0, u4, 0x20);
  r1 = _mm256_permute2f128_ps(u1, u5, 0x20);
  r2 = _mm256_permute2f128_ps(u2, u6, 0x20);
  r3 = _mm256_permute2f128_ps(u3, u7, 0x20);
  r4 = _mm256_permute2f128_ps(u0, u4, 0x31);
  r5 = _mm256_permute2f128_ps(u1, u5, 0x31);
  r6 = _mm256_permute2f128_ps(u2, u6, 0x31);
  r7 = _mm256_permute2f128_ps(u3, u7, 0x31);
  const __m256 out[8] = {r0, r1, r2, r3, r4, r5, r6, r7};
  for (int i = 0; i < 8; ++i) _mm256_storeu_ps(b + i * ldb, upd8(out[i], alpha, b + i * ldb, beta, bz));
}

__attribute__((target("avx2"))) static inline __m256d upd4(__m256d x, double alpha, const double* bp, double beta,
                                                           bool bz) {
  __m256d y = _mm256_mul_pd(x, _mm256_set1_pd(alpha));
  if (!bz) y = _mm256_add_pd(y, _mm256_mul_pd(_mm256_loadu_pd(bp), _mm256_set1_pd(beta)));
  return y;
}

__attribute__((target("avx2"))) static void block4x4_f64(const double* a, int64_t lda, double* b, int64_t ldb,
                                                         double alpha, double beta, bool bz) {
  const __m256d r0 = _mm256_loadu_pd(a), r1 = _mm256_loadu_pd(a + lda);
  const __m256d r2 = _mm256_loadu_pd(a + 2 * lda), r3 = _mm256_loadu_pd(a + 3 * lda);
  const __m256d t0 = _mm256_unpacklo_pd(r0, r1), t1 = _mm256_unpackhi_pd(r0, r1);
  const __m256d t2 = _mm256_unpacklo_pd(r2, r3), t3 = _mm256_unpackhi_pd(r2, r3);
  const __m256d out[4] = {_mm256_permute2f128_pd(t0, t2, 0x20), _mm256_permute2f128_pd(t1, t3, 0x20),
                          _mm256_permute2f128_pd(t0, t2, 0x31), _mm256_permute2f128_pd(t1, t3, 0x31)};
  for (int i = 0; i < 4; ++i) _mm256_storeu_pd(b + i * ldb, upd4(out[i], alpha, b + i * ldb, beta, bz));
}

// AVX-512: in-register transpose of W rows by butterfly rounds s = 1, 2, ...:
// for each pair of rows (i, i + s) with (i & s) == 0, lanes j with (j & s)
// == 0 of the new row i come from row i (lane j) and the others from row
// i + s (lane j - s); the new row i + s takes row i's lane j + s and row
// i + s's lane j.  After log2(W) rounds row i holds column i.
__attribute__((target("avx512f"))) static void block16x16_f32(const float* a, int64_t lda, float* b, int64_t ldb,
                                                              float alpha, float beta, bool bz) {
  __m512 r[16];
  for (int j = 0; j < 16; ++j) r[j] = _mm512_loadu_ps(a + j * lda);
  for (int s = 1; s < 16; s <<= 1) {
    alignas(64) int lo[16], hi[16];
    for (int j = 0; j < 16; ++j) {
      lo[j] = (j & s) ? 16 + j - s : j;
      hi[j] = (j & s) ? 16 + j : j + s;
    }
    const __m512i ilo = _mm512_load_si512(lo), ihi = _mm512_load_si512(hi);
    for (int i = 0; i < 16; ++i) {
      if (i & s) continue;
      const __m512 x = r[i], y = r[i + s];
      r[i] = _mm512_permutex2var_ps(x, ilo, y);
      r[i + s] = _mm512_permutex2var_ps(x, ihi, y);
    }
  }
  const __m512 al = _mm512_set1_ps(alpha), be = _mm512_set1_ps(beta);
  for (int i = 0; i < 16; ++i) {
    __m512 y = _mm512_mul_ps(r[i], al);
    if (!bz) y = _mm512_add_ps(y, _mm512_mul_ps(_mm512_loadu_ps(b + i * ldb), be));
    _mm512_storeu_ps(b + i * ldb, y);
  }
}

__attribute__((target("avx512f"))) static void block8x8_f64(const double* a, int64_t lda, double* b, int64_t ldb,
                                                            double alpha, double beta, bool bz) {
  __m512d r[8];
  for (int j = 0; j < 8; ++j) r[j] = _mm512_loadu_pd(a + j * lda);
  for (int s = 1; s < 8; s <<= 1) {
    alignas(64) int64_t lo[8], hi[8];
    for (int j = 0; j < 8; ++j) {
      lo[j] = (j & s) ? 8 + j - s : j;
      hi[j] = (j & s) ? 8 + j : j + s;
    }
    const __m512i ilo = _mm512_load_si512(lo), ihi = _mm512_load_si512(hi);
    for (int i = 0; i < 8; ++i) {
      if (i & s) continue;
      const __m512d x = r[i], y = r[i + s];
      r[i] = _mm512_permutex2var_pd(x, ilo, y);
      r[i + s] = _mm512_permutex2var_pd(x, ihi, y);
    }
  }
  const __m512d al = _mm512_set1_pd(alpha), be = _mm512_set1_pd(beta);
  for (int i = 0; i < 8; ++i) {
    __m512d y = _mm512_mul_pd(r[i], al);
    if (!bz) y = _mm512_add_pd(y, _mm512_mul_pd(_mm512_loadu_pd(b + i * ldb), be));
    _mm512_storeu_pd(b + i * ldb, y);
  }
}

template <class T>
static inline void scalar_update(const T* a, T* b, T alpha, T beta, bool bz) {
  const T x = alpha * *a;
  if (bz) *b = x;
  else {
    const T y = beta * *b;
    *b = x + y;
  }
}

// The block size W and the micro-kernel for T on this host.
template <class T>
struct Kernel {
  int W;
  void (*fn)(const T*, int64_t, T*, int64_t, T, T, bool);
};

static Kernel<float> pick(float*, int isa) {
  if (isa >= 2) return {16, block16x16_f32};
  return {8, block8x8_f32};
}
static Kernel<double> pick(double*, int isa) {
  if (isa >= 2) return {8, block8x8_f64};
  return {4, block4x4_f64};
}

// Work item = one macro block (MB x MB elements of dimensions 0 and p0) for
// one position of the other dimensions; items are numbered with the macro
// block of dimension p0 fastest, then dimension 0, then the others in B order.
template <class T>
static void run(const Geo& g, const T* A, T* B, T alpha, T beta, bool bz, int nthreads, int isa) {
  const int r = g.r, p0 = g.p0;
  int64_t total = 1;
  for (int k = 0; k < r; ++k) total *= g.n[k];
  if (total == 0) return;
  if (p0 == 0) {
    // dimension 0 stride-1 on both sides: scaled row copies
    const int64_t n0 = g.n[0];
    const int64_t rows = total / n0;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t row = 0; row < rows; ++row) {
      int64_t rem = row, oa = 0, ob = 0;
      for (int k = 1; k < r; ++k) {
        const int64_t x = rem % g.n[k];
        rem /= g.n[k];
        oa += x * g.sa[k];
        ob += x * g.sb[k];
      }
      const T* a = A + oa;
      T* b = B + ob;
      if (bz) {
#pragma omp simd
        for (int64_t i = 0; i < n0; ++i) b[i] = alpha * a[i];
      } else {
#pragma omp simd
        for (int64_t i = 0; i < n0; ++i) {
          const T x = alpha * a[i];
          const T y = beta * b[i];
          b[i] = x + y;
        }
      }
    }
    return;
  }
  const Kernel<T> kern = pick((T*)nullptr, isa);
  const int W = kern.W;
  const int64_t MB = 8 * W;  // macro block edge: 8 x 8 micro-blocks
  const int64_t n0 = g.n[0], np = g.n[p0];
  const int64_t b0 = (n0 + MB - 1) / MB, bp = (np + MB - 1) / MB;
  int others[kMaxDim], no = 0;
  // the other dimensions in B order (innermost first)
  int perm_b[kMaxDim];
  {
    // B order of A dimensions: by increasing B stride
    int idx[kMaxDim];
    for (int k = 0; k < r; ++k) idx[k] = k;
    for (int i = 1; i < r; ++i)
      for (int j = i; j > 0 && g.sb[idx[j]] < g.sb[idx[j - 1]]; --j) {
        const int t = idx[j];
        idx[j] = idx[j - 1];
        idx[j - 1] = t;
      }
    for (int k = 0; k < r; ++k) perm_b[k] = idx[k];
  }
  for (int q = 0; q < r; ++q)
    if (perm_b[q] != 0 && perm_b[q] != p0) others[no++] = perm_b[q];
  int64_t nother = 1;
  for (int m = 0; m < no; ++m) nother *= g.n[others[m]];
  const int64_t items = nother * b0 * bp;
  const int64_t sa_p = g.sa[p0], sb_0 = g.sb[0];
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int64_t it = 0; it < items; ++it) {
    const int64_t jp = it % bp;
    const int64_t j0 = (it / bp) % b0;
    int64_t rem = it / bp / b0, oa = 0, ob = 0;
    for (int m = 0; m < no; ++m) {
      const int k = others[m];
      const int64_t x = rem % g.n[k];
      rem /= g.n[k];
      oa += x * g.sa[k];
      ob += x * g.sb[k];
    }
    const int64_t i0 = j0 * MB, e0 = i0 + MB < n0 ? i0 + MB : n0;
    const int64_t ip = jp * MB, ep = ip + MB < np ? ip + MB : np;
    const int64_t f0 = i0 + (e0 - i0) / W * W, fp = ip + (ep - ip) / W * W;  // end of whole micro-blocks
    for (int64_t x0 = i0; x0 < f0; x0 += W)
      for (int64_t xp = ip; xp < fp; xp += W)
        kern.fn(A + oa + x0 + xp * sa_p, sa_p, B + ob + xp + x0 * sb_0, sb_0, alpha, beta, bz);
    // edges: the rows of dimension 0 past f0 and the columns of p0 past fp
    for (int64_t x0 = i0; x0 < e0; ++x0)
      for (int64_t xp = (x0 < f0 ? fp : ip); xp < ep; ++xp)
        scalar_update(A + oa + x0 + xp * sa_p, B + ob + xp + x0 * sb_0, alpha, beta, bz);
  }
}

}  // namespace

extern "C" {

// Highest instruction set the micro-kernels use on this host: 2 AVX-512F,
// 1 AVX2, 0 none (the entry point then refuses).
int ttc_cpu_blocked_isa(void) {
  __builtin_cpu_init();
  if (__builtin_cpu_supports("avx512f")) return 2;
  if (__builtin_cpu_supports("avx2")) return 1;
  return 0;
}

// B = alpha*A + beta*B (beta_is_zero: B not read) for same-type s (0) or d
// (1) tensors, arguments as the oracle's ttc_oracle_transpose; isa: 0 =
// the best the host has, 1 = AVX2, 2 = AVX-512.  Returns 0 on success.
int ttc_cpu_blocked_transpose(int type, const void* A, void* B, double alpha, double beta, int beta_is_zero,
                              int dim, const int* perm, const int* size, const int* lda, const int* ldb,
                              int nthreads, int isa) {
  Geo g;
  if (!make_geo(dim, perm, size, lda, ldb, g)) return 1;
  const int have = ttc_cpu_blocked_isa();
  if (have == 0) return 3;
  if (isa == 0 || isa > have) isa = have;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  if (type == 0)
    run<float>(g, static_cast<const float*>(A), static_cast<float*>(B), (float)alpha, (float)beta,
               beta_is_zero != 0, nthreads, isa);
  else if (type == 1)
    run<double>(g, static_cast<const double*>(A), static_cast<double*>(B), alpha, beta, beta_is_zero != 0,
                nthreads, isa);
  else
    return 2;
  return 0;
}

}  // extern "C"
