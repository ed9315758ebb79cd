// Internal helpers shared by the runtime translation units.
#pragma once
#include <hip/hip_runtime.h>

#include "ttc_plan.h"

namespace ttc {

int device_cus(int dev);
int occupancy(const void* func, int lds_bytes, int threads);
unsigned grid_size(uint32_t ntiles, int wg_per_cu, const void* func, int lds_bytes, int threads);
const ttc_plan_s* variant_for(const ttc_plan_s* pl, const void* A, const void* B);
ttc_status_t launch_plan(const ttc_plan_s* pl, const void* A, void* B, double alpha, double beta, hipStream_t s);

}  // namespace ttc
