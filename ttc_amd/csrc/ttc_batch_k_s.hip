// Grouped batch kernels of the fp32 -> fp32 pair (one file per pair so
// that the pairs compile in parallel).
#include "ttc_batch_kernels.hpp"

namespace ttc {
BatchKernelSet batch_kernels_s() { return batch_kernel_set<OpF32<float, float>, 4>(); }
}  // namespace ttc
