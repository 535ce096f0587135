// Preconditioner and Krylov solvers on the device.
#pragma once
#include "vh_geom.h"
#include "vh_ldl.h"
#include <memory>

namespace vh {

// Pattern data of the Schur factor, shared by every frequency of a geometry:
// S pattern, ordering, elimination tree and L pattern, level schedules, and
// their device copies.
struct PreconSymbolic {
  std::vector<int> Sp, Si;            // pattern of S (full symmetric CSR)
  LdlFactor F;                        // perm/iperm/parent/Lp/Li (values unused)
  std::vector<int> csc2csr;           // position of each CSC entry of L in its CSR copy
  std::vector<int> fl_ptr, fl_rows, bl_ptr, bl_rows;
  int64_t nnzL = 0;
  DevBuf<int> d_perm, d_fl_rows, d_bl_rows, d_fl_ptr, d_bl_ptr;
  DevBuf<int> Rp, Ri, Cp, Ci, ArTp_col;
};

struct Precon {
  Geometry* G = nullptr;
  bool ready = false;
  bool geometric_order = true;        // nested dissection by voxel-face planes
  std::shared_ptr<PreconSymbolic> sym;
  LdlFactor F;                        // numeric values (Lx, D) of this frequency
  std::vector<double> ainv_h;
  // tuning
  int64_t wide_items = 2048;          // level launched on the whole grid above this many (row, rhs) items
  int chain_rpb = 16;                 // right-hand sides per workgroup in narrow-level chains
  // device values
  DevBuf<double> Rx64, Cx64, D64, ainv64;
  DevBuf<float> Rx32, Cx32, D32, ainv32;
  DevBuf<char> ws;
};

// A_inv, S, LDL^T for jwmu = j*omega*mu (lse_sparse_precon_prepare.m).
void precon_prepare(Precon& P, double jwmu_re, double jwmu_im);
// y = M^-1 x for nrhs vectors (device pointers), lse_sparse_precon_multiply.m:84-88.
template <typename R>
void precon_apply(Precon& P, int nrhs, const cplx<R>* x, int64_t ldx, cplx<R>* y, int64_t ldy,
                  const int* slots = nullptr);

}  // namespace vh
