// Preconditioner and Krylov solvers on the device.
#pragma once
#include "vh_geom.h"

namespace vh {

struct Precon {
  Geometry* G = nullptr;
  bool ready = false;
};


}  // namespace vh
