// Pass-Z kernels of the FFT matvec in complex64 (a translation unit of its
// own so the precisions compile in parallel).
#include "vh_matvec_impl.h"

namespace vh {
template void pass_z<float>(Geometry&, int, float2*);
}  // namespace vh
