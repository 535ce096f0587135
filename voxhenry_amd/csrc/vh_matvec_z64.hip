// Pass-Z kernels of the FFT matvec in complex128 (a translation unit of its
// own so the precisions compile in parallel).
#include "vh_matvec_impl.h"

namespace vh {
template void pass_z<double>(Geometry&, int, double2*);
}  // namespace vh
