// X/Y/Yi/Xi pass kernels of the complex64 matvec, X/Y shape (256 threads, 16 elements per thread).
#include "vh_matvec_impl.h"

namespace vh {
template void run_passes_t<float, 16, 256>(VH_RUN_PASSES_ARGS(float));
}  // namespace vh
