// Grouped batch kernels of the complex fp32 -> complex fp32 pair (one file per pair so
// that the pairs compile in parallel).
#include "ttc_batch_kernels.hpp"

namespace ttc {
BatchKernelSet batch_kernels_c() { return batch_kernel_set<OpCC, 2>(); }
}  // namespace ttc
