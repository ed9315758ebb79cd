// Kernel registry: one entry per instantiated (type pair, vector widths, mode)
// kernel, with a host launcher.  The instance files under gen/ are written by
// ttc_amd/gen_instances.py and hold only lists of entries.
#pragma once
#include <hip/hip_runtime.h>

#include "ttc_kernels.hpp"
#include "ttc_rc.hpp"
#include "ttc_row.hpp"

namespace ttc {

// kernel family (matches ttc_config_t.family)
enum : int { kFamBox = 1, kFamRC = 2, kFamRow = 3 };

using LaunchFn = hipError_t (*)(const TileParams&, const void*, void*, double, double, unsigned, size_t, hipStream_t);
using LaunchFp8Fn = hipError_t (*)(const TileParams&, const void*, void*, float, const float*, const float*, float,
                                   float, float*, float*, unsigned, size_t, hipStream_t);

using LaunchRCFn = hipError_t (*)(const RCParams&, const void*, void*, double, double, unsigned, size_t,
                                  hipStream_t);

using LaunchFp8RCFn = hipError_t (*)(const RCParams&, const void*, void*, float, const float*, const float*, float,
                                     float, float*, float*, unsigned, size_t, hipStream_t);

using LaunchRowFn = hipError_t (*)(const RowParams&, const void*, void*, double, double, const float*, const float*,
                                   float, float, float*, float*, unsigned, hipStream_t);

struct KernelEntry {
  int ta, tb, va, vb, mode, family;
  LaunchFn launch;
  LaunchFp8Fn launch_fp8;
  LaunchRCFn launch_rc;
  const void* func;
  LaunchRowFn launch_row;
  LaunchFp8RCFn launch_fp8_rc;
  // row x column family: the instance for RCParams::fast launches
  LaunchRCFn launch_rc_fast;
  LaunchFp8RCFn launch_fp8_rc_fast;
  const void* func_fast;
};

template <class OP, int VA, int VB, bool BETA, bool NT>
hipError_t launch_transpose(const TileParams& p, const void* A, void* B, double alpha, double beta, unsigned grid,
                            size_t lds, hipStream_t s) {
  const typename OP::TA* a = static_cast<const typename OP::TA*>(A);
  typename OP::TB* b = static_cast<typename OP::TB*>(B);
  TileParams pp = p;
  void* args[] = {&pp, &a, &b, &alpha, &beta};
  return hipLaunchKernel(reinterpret_cast<const void*>(&transpose_kernel<OP, VA, VB, BETA, NT>), dim3(grid),
                         dim3(kThreads), args, lds, s);
}

template <class OP, int VA, int VB>
hipError_t launch_fp8(const TileParams& p, const void* A, void* B, float alpha, const float* scale_in,
                      const float* amax_in, float fp8_max, float scale_host, float* amax_out, float* scale_inv_out,
                      unsigned grid, size_t lds, hipStream_t s) {
  const typename OP::TA* a = static_cast<const typename OP::TA*>(A);
  typename OP::TB* b = static_cast<typename OP::TB*>(B);
  TileParams pp = p;
  void* args[] = {&pp, &a, &b, &alpha, &scale_in, &amax_in, &fp8_max, &scale_host, &amax_out, &scale_inv_out};
  return hipLaunchKernel(reinterpret_cast<const void*>(&cast_fp8_kernel<OP, VA, VB>), dim3(grid), dim3(kThreads),
                         args, lds, s);
}

template <class OP, int VA, int VB, bool BETA, bool NT, bool FAST>
hipError_t launch_rc(const RCParams& p, const void* A, void* B, double alpha, double beta, unsigned grid, size_t lds,
                     hipStream_t s) {
  const typename OP::TA* a = static_cast<const typename OP::TA*>(A);
  typename OP::TB* b = static_cast<typename OP::TB*>(B);
  RCParams pp = p;
  void* args[] = {&pp, &a, &b, &alpha, &beta};
  return hipLaunchKernel(reinterpret_cast<const void*>(&transpose_rc_kernel<OP, VA, VB, BETA, NT, FAST>), dim3(grid),
                         dim3(kRcThreads), args, lds, s);
}

template <class OP, int VA, int VB, bool FAST>
hipError_t launch_fp8_rc(const RCParams& p, const void* A, void* B, float alpha, const float* scale_in,
                         const float* amax_in, float fp8_max, float scale_host, float* amax_out, float* scale_inv_out,
                         unsigned grid, size_t lds, hipStream_t s) {
  const typename OP::TA* a = static_cast<const typename OP::TA*>(A);
  typename OP::TB* b = static_cast<typename OP::TB*>(B);
  RCParams pp = p;
  void* args[] = {&pp, &a, &b, &alpha, &scale_in, &amax_in, &fp8_max, &scale_host, &amax_out, &scale_inv_out};
  return hipLaunchKernel(reinterpret_cast<const void*>(&cast_fp8_rc_kernel<OP, VA, VB, FAST>), dim3(grid),
                         dim3(kRcThreads), args, lds, s);
}

template <class OP, int VA, int VB, int MODE>
KernelEntry make_rc_entry(int ta, int tb) {
  KernelEntry e{ta, tb, VA, VB, MODE, kFamRC, nullptr, nullptr, nullptr, nullptr};
  if constexpr (MODE == kModeFp8) {
    e.launch_fp8_rc = &launch_fp8_rc<OP, VA, VB, false>;
    e.func = reinterpret_cast<const void*>(&cast_fp8_rc_kernel<OP, VA, VB, false>);
    e.launch_fp8_rc_fast = &launch_fp8_rc<OP, VA, VB, true>;
    e.func_fast = reinterpret_cast<const void*>(&cast_fp8_rc_kernel<OP, VA, VB, true>);
  } else {
    constexpr bool beta = MODE == kModeBeta;
    constexpr bool nt = MODE == kModeBzNt;
    e.launch_rc = &launch_rc<OP, VA, VB, beta, nt, false>;
    e.func = reinterpret_cast<const void*>(&transpose_rc_kernel<OP, VA, VB, beta, nt, false>);
    e.launch_rc_fast = &launch_rc<OP, VA, VB, beta, nt, true>;
    e.func_fast = reinterpret_cast<const void*>(&transpose_rc_kernel<OP, VA, VB, beta, nt, true>);
  }
  return e;
}

template <class OP, int V, int MODE>
hipError_t launch_row(const RowParams& p, const void* A, void* B, double alpha, double beta, const float* scale_in,
                      const float* amax_in, float fp8_max, float scale_host, float* amax_out, float* scale_inv_out,
                      unsigned grid, hipStream_t s) {
  const typename OP::TA* a = static_cast<const typename OP::TA*>(A);
  typename OP::TB* b = static_cast<typename OP::TB*>(B);
  RowParams pp = p;
  void* args[] = {&pp, &a, &b, &alpha, &beta, &scale_in, &amax_in, &fp8_max, &scale_host, &amax_out, &scale_inv_out};
  return hipLaunchKernel(reinterpret_cast<const void*>(&row_kernel<OP, V, MODE>), dim3(grid), dim3(kThreads), args,
                         0, s);
}

template <class OP, int V, int MODE>
KernelEntry make_row_entry(int ta, int tb) {
  KernelEntry e{ta, tb, V, V, MODE, kFamRow, nullptr, nullptr, nullptr, nullptr, nullptr};
  e.launch_row = &launch_row<OP, V, MODE>;
  e.func = reinterpret_cast<const void*>(&row_kernel<OP, V, MODE>);
  return e;
}

template <class OP, int VA, int VB, int MODE>
KernelEntry make_entry(int ta, int tb) {
  KernelEntry e{ta, tb, VA, VB, MODE, kFamBox, nullptr, nullptr, nullptr, nullptr};
  if constexpr (MODE == kModeFp8) {
    e.launch_fp8 = &launch_fp8<OP, VA, VB>;
    e.func = reinterpret_cast<const void*>(&cast_fp8_kernel<OP, VA, VB>);
  } else {
    constexpr bool beta = MODE == kModeBeta;
    constexpr bool nt = MODE == kModeBzNt;
    e.launch = &launch_transpose<OP, VA, VB, beta, nt>;
    e.func = reinterpret_cast<const void*>(&transpose_kernel<OP, VA, VB, beta, nt>);
  }
  return e;
}

// Registry access (ttc_runtime.hip).
const KernelEntry* find_kernel(int ta, int tb, int va, int vb, int mode, int family);

}  // namespace ttc
