// Matvec drivers (pass order, workspaces) in complex64.
#include "vh_matvec_impl.h"

namespace vh {
template void matvec_ws<float>(Geometry&, int, const float2*, int64_t, float2*, int64_t, double, double, int, char*);
template void matvec<float>(Geometry&, int, const float2*, int64_t, float2*, int64_t, double, double, int, const int*,
                          const double2*, const float2* const*);
}  // namespace vh
