// Matvec drivers (pass order, workspaces) in complex128.
#include "vh_matvec_impl.h"

namespace vh {
template void matvec_ws<double>(Geometry&, int, const double2*, int64_t, double2*, int64_t, double, double, int, char*);
template void matvec<double>(Geometry&, int, const double2*, int64_t, double2*, int64_t, double, double, int, const int*,
                          const double2*, const double2* const*);
}  // namespace vh
