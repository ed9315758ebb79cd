// Grouped batch kernels of the bf16 -> bf16 pair (one file per pair so
// that the pairs compile in parallel).
#include "ttc_batch_kernels.hpp"

namespace ttc {
BatchKernelSet batch_kernels_bf16() { return batch_kernel_set<OpF32<bf16s, bf16s>, 8>(); }
}  // namespace ttc
