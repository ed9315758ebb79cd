// Grouped batch kernels of the complex fp64 -> complex fp64 pair (one file per pair so
// that the pairs compile in parallel).
#include "ttc_batch_kernels.hpp"

namespace ttc {
BatchKernelSet batch_kernels_z() { return batch_kernel_set<OpZZ, 1>(); }
}  // namespace ttc
