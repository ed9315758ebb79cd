// Grouped batch kernels of the fp16 -> fp16 pair (one file per pair so
// that the pairs compile in parallel).
#include "ttc_batch_kernels.hpp"

namespace ttc {
BatchKernelSet batch_kernels_f16() { return batch_kernel_set<OpF32<f16s, f16s>, 8>(); }
}  // namespace ttc
