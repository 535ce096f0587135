// X/Y/Yi/Xi pass kernels of the complex128 matvec, every X/Y shape.
#include "vh_matvec_impl.h"

namespace vh {
template void run_passes_t<double, 8, 256>(VH_RUN_PASSES_ARGS(double));
template void run_passes_t<double, 4, 256>(VH_RUN_PASSES_ARGS(double));
template void run_passes_t<double, 8, 128>(VH_RUN_PASSES_ARGS(double));
}  // namespace vh
