// Grouped batch kernels of the fp64 -> fp64 pair (one file per pair so
// that the pairs compile in parallel).
#include "ttc_batch_kernels.hpp"

namespace ttc {
BatchKernelSet batch_kernels_d() { return batch_kernel_set<OpDD, 2>(); }
}  // namespace ttc
