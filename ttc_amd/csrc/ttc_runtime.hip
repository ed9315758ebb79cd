// Execution side of libttc_hip.so: kernel registry, launch shape, the
// asynchronous entry points, fp8 cast-transpose and the synchronous drop-in
// that keeps the reference wrapper's semantics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "gen/inst_all.h"
#include "ttc_launch.hpp"
#include "ttc_plan.h"
#include "ttc_runtime.h"

namespace ttc {

// ---------------------------------------------------------------- registry
static const std::vector<KernelEntry>& registry() {
  static const std::vector<KernelEntry> all = [] {
    std::vector<KernelEntry> v;
    for (InstanceList f : kInstanceLists) {
      int n = 0;
      const KernelEntry* e = f(&n);
      v.insert(v.end(), e, e + n);
    }
    return v;
  }();
  return all;
}

const KernelEntry* find_kernel(int ta, int tb, int va, int vb, int mode, int family) {
  for (const KernelEntry& e : registry())
    if (e.ta == ta && e.tb == tb && e.va == va && e.vb == vb && e.mode == mode && e.family == family) return &e;
  return nullptr;
}

// ------------------------------------------------------------- device info
static std::mutex g_mu;

int device_cus(int dev) {
  static std::map<int, int> cache;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = cache.find(dev);
  if (it != cache.end()) return it->second;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
  cache[dev] = cus;
  return cus;
}

int occupancy(const void* func, int lds_bytes, int threads) {
  static std::map<std::pair<const void*, int>, int> cache;
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_pair(func, lds_bytes);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  // Work-groups above 64 KiB of LDS must say so once per kernel (gfx950 has 160 KiB per CU).
  if (lds_bytes > 65536) (void)hipFuncSetAttribute(func, hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, func, threads, (size_t)lds_bytes) != hipSuccess || n <= 0)
    n = 1;
  cache[key] = n;
  return n;
}

// Persistent work-groups per CU when the plan does not say: measured on
// MI355X, 4 resident 256-thread work-groups per CU stream faster than the
// occupancy limit (5-9), which only adds concurrent tile streams.
static constexpr int kDefaultWgPerCu = 4;

unsigned grid_size(uint32_t ntiles, int wg_per_cu, const void* func, int lds_bytes, int threads) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  const int cus = device_cus(dev);
  const int occ = occupancy(func, lds_bytes, threads);
  const int def = kDefaultWgPerCu * kThreads / threads;
  int wg = wg_per_cu > 0 ? wg_per_cu : (occ < def ? occ : def);
  if (wg < 1) wg = 1;
  const uint64_t g = (uint64_t)cus * (uint64_t)wg;
  return (unsigned)(ntiles < g ? ntiles : g);
}

// Row-copy launches keep no LDS and stream with more resident work-groups:
// 8 per CU unless the plan says otherwise (or occupancy allows fewer).
static unsigned row_grid(const ttc_plan_s* use, int wg_per_cu, const void* func) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  const int occ = occupancy(func, 0, kThreads);
  int wg = wg_per_cu > 0 ? wg_per_cu : std::min(occ, 8);
  if (wg < 1) wg = 1;
  const uint64_t g = (uint64_t)device_cus(dev) * (uint64_t)wg;
  const uint64_t n = use->p.ntiles;
  return (unsigned)(n < g ? n : g);
}

static bool aligned(const void* p, int bytes) { return ((uintptr_t)p % (uintptr_t)bytes) == 0; }

// Picks the plan variant whose vector widths the pointers allow.
const ttc_plan_s* variant_for(const ttc_plan_s* pl, const void* A, const void* B) {
  const int sa = dtype_bytes(pl->type_a), sb = dtype_bytes(pl->type_b);
  if (aligned(A, sa * pl->va) && aligned(B, sb * pl->vb)) return pl;
  const ttc_plan_s* fb = pl->fallback;
  if (fb && aligned(A, sa) && aligned(B, sb)) return fb;
  return nullptr;
}

static bool is_fp8_type(int t) { return t == TTC_E4M3 || t == TTC_E5M2; }

ttc_status_t launch_plan(const ttc_plan_s* pl, const void* A, void* B, double alpha, double beta, hipStream_t s) {
  if (pl->g.elements == 0 || pl->p.ntiles == 0) return TTC_SUCCESS;
  const ttc_plan_s* use = variant_for(pl, A, B);
  if (!use) return TTC_ERROR_INVALID_VALUE;
  const int mode = pl->beta_is_zero ? (pl->nt_store ? kModeBzNt : kModeBz) : kModeBeta;
  const KernelEntry* e = find_kernel(pl->type_a, pl->type_b, use->va, use->vb, mode, use->family);
  if (!e) return TTC_ERROR_UNSUPPORTED;
  if (use->family == kFamRow) {
    const unsigned grid = row_grid(use, pl->wg_per_cu, e->func);
    return e->launch_row(use->row, A, B, alpha, beta, nullptr, nullptr, 0.0f, 1.0f, nullptr, nullptr, grid, s) ==
                   hipSuccess
               ? TTC_SUCCESS
               : TTC_ERROR_HIP;
  }
  const bool fast = use->family == kFamRC && use->rc.fast;
  const unsigned grid = grid_size(use->p.ntiles, pl->wg_per_cu, fast ? e->func_fast : e->func, use->lds_bytes,
                                  use->family == kFamRC ? kRcThreads : kThreads);
  const hipError_t err =
      use->family == kFamRC
          ? (fast ? e->launch_rc_fast : e->launch_rc)(use->rc, A, B, alpha, beta, grid, (size_t)use->lds_bytes, s)
          : e->launch(use->p, A, B, alpha, beta, grid, (size_t)use->lds_bytes, s);
  return err == hipSuccess ? TTC_SUCCESS : TTC_ERROR_HIP;
}

// ------------------------------------------------------------------- amax
template <class TA>
static hipError_t launch_amax(const Geometry& g, const void* A, float alpha, float* out, hipStream_t s) {
  AmaxParams q;
  std::memset(&q, 0, sizeof(q));
  q.rank = g.rank;
  q.n0 = g.size[0];
  q.dn0 = make_fastdiv((uint32_t)g.size[0]);
  bool dense = true;
  int64_t prod = 1;
  for (int k = 0; k < g.rank; ++k) {
    if (g.sa[k] != prod) dense = false;
    prod *= g.size[k];
    q.dn[k] = make_fastdiv((uint32_t)g.size[k]);
    q.sa[k] = g.sa[k];
  }
  constexpr int VEC = 16 / sizeof(TA);
  q.dense = dense && aligned(A, 16) ? 1 : 0;
  const uint64_t N = (uint64_t)g.elements;
  if (!q.dense && N >= (1ull << 31)) return hipErrorInvalidValue;
  int dev = 0;
  (void)hipGetDevice(&dev);
  const uint64_t per = (uint64_t)kThreads * (q.dense ? VEC * kAmaxUnroll : 1);
  uint64_t blocks = (N + per - 1) / per;
  const uint64_t cap = (uint64_t)device_cus(dev) * (q.dense ? 2 : 8);
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  const TA* a = static_cast<const TA*>(A);
  void* args[] = {&q, &a, &alpha, &out, (void*)&N};
  return hipLaunchKernel(reinterpret_cast<const void*>(&amax_kernel<TA, VEC>), dim3((unsigned)blocks),
                         dim3(kThreads), args, 0, s);
}

// Sets *v = 0 (graph-safe replacement of a memset).
__global__ void zero_f32_kernel(float* v) {
  if (threadIdx.x == 0) *v = 0.0f;
}

static hipError_t launch_zero(float* p, hipStream_t s) {
  void* args[] = {&p};
  return hipLaunchKernel(reinterpret_cast<const void*>(&zero_f32_kernel), dim3(1), dim3(64), args, 0, s);
}

}  // namespace ttc

using namespace ttc;

extern "C" ttc_status_t ttc_execute(ttc_plan_t pl, const void* A, void* B, double alpha, double beta,
                                    ttc_stream_t stream) {
  if (!pl) return TTC_ERROR_INVALID_VALUE;
  if (is_fp8_type(pl->type_b)) return TTC_ERROR_UNSUPPORTED;
  if (pl->g.elements == 0) return TTC_SUCCESS;
  if (!A || !B) return TTC_ERROR_INVALID_VALUE;
  return launch_plan(pl, A, B, alpha, beta, static_cast<hipStream_t>(stream));
}

extern "C" ttc_status_t ttc_execute_fp8(ttc_plan_t pl, const void* A, void* B, double alpha, const float* d_scale_in,
                                        float scale, float* d_amax_out, float* d_scale_inv_out,
                                        ttc_scale_mode_t mode, ttc_stream_t stream) {
  if (!pl) return TTC_ERROR_INVALID_VALUE;
  if (!is_fp8_type(pl->type_b)) return TTC_ERROR_UNSUPPORTED;
  if (mode != TTC_SCALE_DELAYED && mode != TTC_SCALE_CURRENT) return TTC_ERROR_INVALID_VALUE;
  if (mode == TTC_SCALE_CURRENT && !d_amax_out) return TTC_ERROR_INVALID_VALUE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (pl->g.elements == 0) {
    if (d_amax_out && launch_zero(d_amax_out, s) != hipSuccess) return TTC_ERROR_HIP;
    return TTC_SUCCESS;
  }
  if (!A || !B) return TTC_ERROR_INVALID_VALUE;
  const ttc_plan_s* use = variant_for(pl, A, B);
  if (!use) return TTC_ERROR_INVALID_VALUE;
  const KernelEntry* e = find_kernel(pl->type_a, pl->type_b, use->va, use->vb, kModeFp8, use->family);
  if (!e) return TTC_ERROR_UNSUPPORTED;
  const float fp8_max = pl->type_b == TTC_E4M3 ? 448.0f : 57344.0f;
  const float al = (float)alpha;
  if (d_amax_out && launch_zero(d_amax_out, s) != hipSuccess) return TTC_ERROR_HIP;
  const float* amax_in = nullptr;
  float* amax_out = d_amax_out;
  if (mode == TTC_SCALE_CURRENT) {
    hipError_t err;
    switch (pl->type_a) {
      case TTC_S: err = launch_amax<float>(pl->g, A, al, d_amax_out, s); break;
      case TTC_BF16: err = launch_amax<bf16s>(pl->g, A, al, d_amax_out, s); break;
      default: err = launch_amax<f16s>(pl->g, A, al, d_amax_out, s); break;
    }
    if (err != hipSuccess) return TTC_ERROR_HIP;
    amax_in = d_amax_out;
    amax_out = nullptr;
    d_scale_in = nullptr;
  }
  if (use->family == kFamRow) {
    const unsigned grid = row_grid(use, pl->wg_per_cu, e->func);
    return e->launch_row(use->row, A, B, alpha, 0.0, d_scale_in, amax_in, fp8_max, scale, amax_out, d_scale_inv_out,
                         grid, s) == hipSuccess
               ? TTC_SUCCESS
               : TTC_ERROR_HIP;
  }
  if (use->family == kFamRC) {
    const bool fast = use->rc.fast != 0;
    const unsigned grid = grid_size(use->p.ntiles, pl->wg_per_cu, fast ? e->func_fast : e->func, use->lds_bytes,
                                    kRcThreads);
    return (fast ? e->launch_fp8_rc_fast : e->launch_fp8_rc)(use->rc, A, B, al, d_scale_in, amax_in, fp8_max, scale,
                                                             amax_out, d_scale_inv_out, grid, (size_t)use->lds_bytes,
                                                             s) == hipSuccess
               ? TTC_SUCCESS
               : TTC_ERROR_HIP;
  }
  const unsigned grid = grid_size(use->p.ntiles, pl->wg_per_cu, e->func, use->lds_bytes, kThreads);
  return e->launch_fp8(use->p, A, B, al, d_scale_in, amax_in, fp8_max, scale, amax_out, d_scale_inv_out, grid,
                       (size_t)use->lds_bytes, s) == hipSuccess
             ? TTC_SUCCESS
             : TTC_ERROR_HIP;
}

// ------------------------------------------------------ synchronous drop-in
extern "C" void ttc_legacy_transpose(ttc_dtype_t type_a, ttc_dtype_t type_b, const void* A, void* B, double alpha,
                                     double beta, int beta_is_zero, int dim, const int* perm, const int* size,
                                     const int* lda, const int* ldb, const ttc_config_t* cfg, const char* name) {
  static std::mutex mu;
  static std::map<std::string, ttc_plan_t> plans;
  const char* who = name ? name : "ttc_legacy_transpose";
  if (dim < 1 || dim > TTC_MAX_INPUT_DIM || !perm || !size) {
    printf("\nKernel ERROR in %s: %s (line: %d)\n", who, ttc_status_string(TTC_ERROR_INVALID_VALUE), __LINE__);
    exit(-1);
  }
  std::string key;
  key.reserve(256);
  auto add = [&key](long long v) { key += std::to_string(v); key += ','; };
  add(type_a); add(type_b); add(beta_is_zero ? 1 : 0); add(dim);
  for (int k = 0; k < dim; ++k) { add(perm[k]); add(size[k]); add(lda ? lda[k] : -1); add(ldb ? ldb[k] : -1); }
  if (cfg) {
    const int* c = reinterpret_cast<const int*>(cfg);
    for (size_t i = 0; i < sizeof(ttc_config_t) / sizeof(int); ++i) add(c[i]);
  }
  ttc_plan_t plan = nullptr;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = plans.find(key);
    if (it != plans.end()) {
      plan = it->second;
    } else {
      ttc_status_t st = ttc_plan_create(&plan, type_a, type_b, dim, perm, size, lda, ldb, beta_is_zero, cfg);
      if (st != TTC_SUCCESS && cfg) {
        // A configuration baked into an emitted artefact was chosen for the
        // lda/ldb it was tuned on; the wrapper takes lda/ldb at call time and
        // the reference kernels accept any.  Keep the launch settings and
        // let the planner re-derive the tile and vector widths, then drop
        // the configuration altogether.
        ttc_config_t relaxed = *cfg;
        for (int k = 0; k < TTC_MAX_DIM; ++k) relaxed.tile[k] = 0;
        relaxed.vec_a = relaxed.vec_b = 0;
        relaxed.run_a = relaxed.run_b = 0;
        st = ttc_plan_create(&plan, type_a, type_b, dim, perm, size, lda, ldb, beta_is_zero, &relaxed);
        if (st != TTC_SUCCESS)
          st = ttc_plan_create(&plan, type_a, type_b, dim, perm, size, lda, ldb, beta_is_zero, nullptr);
      }
      if (st != TTC_SUCCESS) {
        printf("\nKernel ERROR in %s: %s (line: %d)\n", who, ttc_status_string(st), __LINE__);
        exit(-1);
      }
      plans[key] = plan;
    }
  }
  ttc_status_t st = ttc_execute(plan, A, B, alpha, beta, nullptr);
  if (st != TTC_SUCCESS) {
    const char* msg = st == TTC_ERROR_HIP ? hipGetErrorString(hipGetLastError()) : ttc_status_string(st);
    printf("\nKernel ERROR in %s: %s (line: %d)\n", who, msg, __LINE__);
    exit(-1);
  }
  hipError_t err = hipDeviceSynchronize();
  if (err == hipSuccess) err = hipGetLastError();
  if (err != hipSuccess) {
    printf("\nKernel ERROR in %s: %s (line: %d)\n", who, hipGetErrorString(err), __LINE__);
    exit(-1);
  }
}
