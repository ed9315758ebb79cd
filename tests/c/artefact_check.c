/*
 * A C caller of emitted drop-in artefacts (the way TTC-C, a C wrapper of
 * TTC, or user C code calls them): includes the generated headers, links the
 * artefacts and libttc_hip.so with the Cflags/Libs of its pkg-config file,
 * and checks on the GPU
 *   1. sCuTranspose_102_11x9x2 against the golden fixture (tests/golden/s.npz
 *      case 13, extracted to golden_s_102_11x9x2.bin: A, B before, B after);
 *   2. sCuTranspose_102_1000x768x16_bz with the reference harness's input
 *      A[i] = i (ttc/transposeGenerator.py:612-621) against the oracle.
 * Arguments follow the reference wrapper (ttc/CUDAtranspose.py:395-451):
 * device A and B, host int arrays size, lda, ldb (lda/ldb = outer extents).
 * Exit status 0 and "artefact_check ok" on success.
 */
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sCuTranspose_102_11x9x2.h"
#include "sCuTranspose_102_1000x768x16_bz.h"

/* the oracle (test infrastructure), oracle/ttc_oracle.cpp */
int ttc_oracle_transpose(int type_a, int type_b, const void* A, void* B, double alpha, double beta,
                         int beta_is_zero, int dim, const int* perm, const int* size, const int* lda,
                         const int* ldb, int nthreads);

#define CHECK(x)                                                             \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));        \
      exit(2);                                                               \
    }                                                                        \
  } while (0)

static int same_bits(const float* x, const float* y, size_t n, size_t* first) {
  for (size_t i = 0; i < n; ++i)
    if (memcmp(&x[i], &y[i], sizeof(float)) != 0) {
      *first = i;
      return 0;
    }
  return 1;
}

static int check_fixture(const char* path) {
  enum { N = 11 * 9 * 2 };
  static float host[3 * N];
  FILE* f = fopen(path, "rb");
  if (!f || fread(host, sizeof(float), 3 * N, f) != 3 * N) {
    fprintf(stderr, "cannot read %s\n", path);
    return 0;
  }
  fclose(f);
  const int size[3] = {11, 9, 2}, lda[3] = {11, 9, 2}, ldb[3] = {9, 11, 2};
  float *dA, *dB, got[N];
  CHECK(hipMalloc((void**)&dA, N * sizeof(float)));
  CHECK(hipMalloc((void**)&dB, N * sizeof(float)));
  CHECK(hipMemcpy(dA, host, N * sizeof(float), hipMemcpyHostToDevice));
  CHECK(hipMemcpy(dB, host + N, N * sizeof(float), hipMemcpyHostToDevice));
  sCuTranspose_102_11x9x2(dA, dB, 1.0f, 1.0f, size, lda, ldb);
  CHECK(hipMemcpy(got, dB, N * sizeof(float), hipMemcpyDeviceToHost));
  CHECK(hipFree(dA));
  CHECK(hipFree(dB));
  size_t first = 0;
  if (!same_bits(got, host + 2 * N, N, &first)) {
    fprintf(stderr, "sCuTranspose_102_11x9x2: element %zu differs from the golden fixture\n", first);
    return 0;
  }
  printf("sCuTranspose_102_11x9x2: %d elements bit-exact against the golden fixture\n", N);
  return 1;
}

static int check_c1_bz(void) {
  const int perm[3] = {1, 0, 2}, size[3] = {1000, 768, 16}, lda[3] = {1000, 768, 16}, ldb[3] = {768, 1000, 16};
  const size_t n = (size_t)1000 * 768 * 16;
  float* A = (float*)malloc(n * sizeof(float));
  float* B = (float*)malloc(n * sizeof(float));
  float* want = (float*)malloc(n * sizeof(float));
  if (!A || !B || !want) return 0;
  uint32_t h = 12345u;
  for (size_t i = 0; i < n; ++i) {
    A[i] = (float)i; /* exact: n < 2^24 */
    h = h * 1664525u + 1013904223u;
    B[i] = (float)(h >> 8) * (1.0f / 16777216.0f);
  }
  memcpy(want, B, n * sizeof(float));
  if (ttc_oracle_transpose(0, 0, A, want, 1.5, 0.0, 1, 3, perm, size, NULL, NULL, 0) != 0) return 0;
  float *dA, *dB;
  CHECK(hipMalloc((void**)&dA, n * sizeof(float)));
  CHECK(hipMalloc((void**)&dB, n * sizeof(float)));
  CHECK(hipMemcpy(dA, A, n * sizeof(float), hipMemcpyHostToDevice));
  CHECK(hipMemcpy(dB, B, n * sizeof(float), hipMemcpyHostToDevice));
  sCuTranspose_102_1000x768x16_bz(dA, dB, 1.5f, size, lda, ldb);
  CHECK(hipMemcpy(B, dB, n * sizeof(float), hipMemcpyDeviceToHost));
  CHECK(hipFree(dA));
  CHECK(hipFree(dB));
  size_t first = 0;
  const int ok = same_bits(B, want, n, &first);
  if (!ok) fprintf(stderr, "sCuTranspose_102_1000x768x16_bz: element %zu differs from the oracle\n", first);
  else printf("sCuTranspose_102_1000x768x16_bz: %zu elements bit-exact against the oracle\n", n);
  free(A);
  free(B);
  free(want);
  return ok;
}

int main(int argc, char** argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: %s golden_s_102_11x9x2.bin\n", argv[0]);
    return 2;
  }
  printf("artefact_check: fixture %s\n", argv[1]);
  fflush(stdout);
  if (!check_fixture(argv[1])) return 1;
  fflush(stdout);
  if (!check_c1_bz()) return 1;
  printf("artefact_check ok\n");
  return 0;
}
