// Krylov building blocks (fused Arnoldi reductions) and the batched solvers.
#pragma once
#include "vh_solver.h"
#include <complex>
#include <functional>

namespace vh {

// Device vector operations on a batch of slot vectors of length n (stride ld).
template <typename R> struct VecOps {
  Context* ctx;
  int64_t n, ld;
  int kmax;
  int nchunk;
  double2* part;   // reduction partials, part_size(kmax, nslot) entries (owned by the caller's workspace)
  VecOps(Context* c, int64_t n, int64_t ld, int kmax, int nslot, double2* part);
  static size_t part_size(int kmax, int nslot);
  // h[a][j] = V_j^H w (j < k); with_norm: also h[a][k] = ||w||^2.  need
  // (device, optional): systems with need[a] == 0 are skipped and get h = 0.
  void mdot(const cplx<R>* V, int64_t vstride, const cplx<R>* w, int k, int nact, const int* slots, double2* h,
            const int* need = nullptr, bool with_norm = false);
  // need (optional): skipped systems keep out and nrm2 unchanged
  void combine(const cplx<R>* V, int64_t vstride, cplx<R>* out, int k, int nact, const int* slots,
               const double2* coef, double beta, double sign, double2* nrm2, const int* need = nullptr);
  // need[a] = 1 where a second Gram-Schmidt pass is due: ||w'||^2 (nrm2) <
  // ||w||^2 / 2 (h[a][k], from mdot with_norm)
  void reorth_flags(const double2* h, int k, const double2* nrm2, int* need, int nact);
  // Fused two-pass classical Gram-Schmidt (k <= kFusedMax): combine_dot does
  // w <- w - V h1 in place and h2[a][j] = V_j^H w (j < k), h2[a][k] = ||w||^2
  // in one pass over the basis; combine_scale writes out = (w - V h2) / nu
  // with nu^2 = ||w||^2 - sum_j |h2_j|^2 and nrm2[a] = nu^2.
  static constexpr int kFusedMax = 16;
  void combine_dot(const cplx<R>* V, int64_t vstride, cplx<R>* w, int k, int nact, const int* slots,
                   const double2* h1, double2* h2);
  void combine_scale(const cplx<R>* V, int64_t vstride, const cplx<R>* w, cplx<R>* out, int k, int nact,
                     const int* slots, const double2* h2, double2* nrm2);
  void norm2(const cplx<R>* x, int nact, const int* slots, double2* nrm2);
  void scale(const cplx<R>* src, cplx<R>* dst, int nact, const int* slots, const double2* nrm2);
  void sub_from(const cplx<R>* b, int64_t ldb, cplx<R>* y, int nact, const int* slots);
  // v[slot] = (RD)(b[slot] - y[slot]) (b, y, v of stride ld) and nrm2[a] =
  // ||b - y||^2 before the rounding, in one pass
  template <typename RD>
  void residual_to(const cplx<R>* b, const cplx<R>* y, cplx<RD>* v, int nact, const int* slots, double2* nrm2);
  // x[slot] = 0 for the nact slots (one launch; no memset node in a captured graph)
  void zero(cplx<R>* x, int nact, const int* slots);
};

template <typename RD, typename RS>
void convert_vectors(Context* ctx, const cplx<RS>* src, int64_t lds, cplx<RD>* dst, int64_t ldd, int64_t n, int nact,
                     const int* slots_src, const int* slots_dst);
// dst[slot] += src[slot] (src converted to dst's precision) for the listed slots
template <typename RD, typename RS>
void accumulate_vectors(Context* ctx, const cplx<RS>* src, int64_t lds, cplx<RD>* dst, int64_t ldd, int64_t n,
                        int nact, const int* slots);

// Least squares of an Arnoldi cycle on the device, one thread per system:
// the (k+1) x k Hessenberg matrix H, its QR factor R by Givens rotations
// (cosines c, sines s) and the rotated right-hand side g = Q^H beta e1 of
// every slot stay in device memory across the steps of a cycle.
struct LsDev {
  double2* H;      // [slot][(m+1) * m], row i, column j at i*m + j
  double2* R;      // [slot][m * m], upper triangle
  double2* g;      // [slot][m + 1]
  double2* s;      // [slot][m]
  double* c;       // [slot][m]
  double* beta;    // [slot]
  int m;
};
// g = beta e1 with beta = sqrt(nrm2[a]) for the nact listed slots.
void ls_reset(Context* ctx, const LsDev& L, const int* slots, int nact, const double2* nrm2);
// Appends column j (h[a][i] (+ h2[a][i]) for i <= j, H(j+1, j) = sqrt(nrm2[a]))
// and rotates it; res[a] = |g_{j+1}|, the least-squares residual norm.
void ls_push(Context* ctx, const LsDev& L, const int* slots, int nact, int j, const double2* h, const double2* h2,
             int kmax, const double2* nrm2, double* res);
// coef[a][0..k) = y = R^{-1} g (back substitution); rcoef (optional)
// [a][0..k] = beta e1 - H y, the reference's restart residual (rhs - Hc*y).
void ls_solve(Context* ctx, const LsDev& L, const int* slots, int nl, int k, int kmax, double2* coef, double2* rcoef);

// y_slot = Op(x_slot) for the systems listed in `act` (host) / `slots` (the
// same list on the device).
template <typename R>
using OpFn = std::function<void(const std::vector<int>& act, const int* slots, const cplx<R>* x, cplx<R>* y)>;

struct SysStats {
  int flag = 1;               // 0 converged (true residual <= tol), 1 not
  double relres = 0;          // ||b - A x|| / ||b||
  int inner = 0, outer = 0;   // iter = [p outit] of pgmres.m:178
  int iters = 0;              // total Arnoldi steps
  int inner_iters = 0;        // mixed precision: total inner (FP32) Arnoldi steps
  int matvecs = 0;            // operator applications (incl. the final residual)
  std::vector<double> resvec; // resvec of pgmres.m (relative)
};

struct SolveOpts {
  int restart = 100, maxit = 10;
  double tol = 1e-8;
  bool final_residual = true;     // compute relres with one more matvec
  bool flexible = false;          // fpgmres semantics (right preconditioner)
  bool reorth = false;            // classical Gram-Schmidt with a second pass where the first one cancelled
                                  // more than half of ||w||^2 (selective CGS2)
  const std::vector<double>* tol_slot = nullptr;   // per-slot tolerance (overrides tol)
  bool x0 = false;                // x holds an initial guess (flexible or unpreconditioned solves only)
  // per-slot stop request (nullable): a system whose entry is set when a step
  // ends takes the update of the steps done so far and leaves the iteration
  // (its flag then follows from its residual)
  const std::vector<char>* stop_slot = nullptr;
  // flexible solves only: a system whose true residual (final_residual) is
  // above its tolerance although the Arnoldi estimate met it restarts from
  // that residual while outer iterations remain (the reference's fpgmres
  // reports such a system with flag 1; the mixed solver and the FP64
  // continuations, which count only systems at their tolerance, continue)
  bool true_restart = false;
};

// Krylov workspace for nslot systems of length ld with restart m.  The large
// vectors (basis V, flexible basis Z, w, t, r) are views into an arena owned
// by a KrylovCache; the small per-system arrays are owned here.
template <typename R> struct KrylovWork {
  int64_t n = 0, ld = 0;
  int m = 0, nslot = 0;
  bool flexible = false;
  cplx<R>* V = nullptr;   // (m+1) basis vectors, vector j of slot s at V + j*nslot*ld + s*ld
  cplx<R>* Z = nullptr;   // m preconditioned basis vectors (flexible only)
  cplx<R>* w = nullptr;
  cplx<R>* t = nullptr;
  cplx<R>* r = nullptr;
  DevBuf<double2> h, h2, nrm, coef, rcoef;
  DevBuf<double2> lsH, lsR, lsg, lss;     // least-squares state per slot (LsDev)
  DevBuf<double> lsc, lsbeta, lsres;
  DevBuf<double2> part;   // partials of the reductions (VecOps), sized for m + 1 vectors and nslot systems
  DevBuf<int> slots_dev, aux_slots;
  DevBuf<int> need;       // per active system: second Gram-Schmidt pass due (selective CGS2)
  // pinned host staging: slot lists, least-squares residuals, norms
  PinnedBuf<int> h_slots, h_aux;
  PinnedBuf<double> h_res;
  PinnedBuf<double2> h_nrm, h_bn;
  LsDev ls() { return LsDev{lsH.p, lsR.p, lsg.p, lss.p, lsc.p, lsbeta.p, m}; }
  // device bytes of the per-slot small arrays (Hessenberg columns, least-squares state)
  static size_t small_bytes_per_slot(int m) {
    return sizeof(double2) * (size_t)(4 * (m + 1) + 1 + (size_t)(m + 1) * m + (size_t)m * m + (m + 1) + m) +
           sizeof(double) * (size_t)(m + 2);
  }
  static size_t bytes_needed(int64_t ld, int m, int nslot, bool flexible) {
    return sizeof(cplx<R>) * (size_t)ld * nslot * ((m + 1) + (flexible ? m : 0) + 3);
  }
  // Points the large vectors at base (bytes_needed bytes) and sizes the small arrays.
  void bind(char* base, int64_t n, int64_t ld, int m, int nslot, bool flexible);
  void unbind() { V = Z = w = t = r = nullptr; m = 0; nslot = 0; }
};

// Batched (f)pgmres with the reference's semantics for the systems listed in
// sys (slot ids); b and x are slot arrays of stride ld (device).
template <typename R>
void gmres_batch(Context* ctx, int64_t n, const OpFn<R>& A, const OpFn<R>* M, const cplx<R>* b, cplx<R>* x,
                 const std::vector<int>& sys, const SolveOpts& o, KrylovWork<R>& W, std::vector<SysStats>& st);

// Mixed precision: FP64 flexible GMRES (outer, true residual) preconditioned by
// an FP32 left-preconditioned GMRES (inner).
struct MixedOpts {
  double tol = 1e-8;
  int outer_restart = 8, outer_maxit = 4;
  int inner_restart = 40, inner_maxit = 1;
  double inner_tol = 1e-5;          // tightest inner tolerance (the first inner solve)
  bool adaptive_inner = true;       // relax the inner tolerance as the outer residual falls
  const std::vector<double>* tol_slot = nullptr;   // per-slot outer tolerance (overrides tol)
  bool x0 = false;                  // x holds an initial guess (the outer FGMRES starts from b - A x)
  bool inner_flexible = false;      // inner solve right-preconditioned (fpgmres semantics): any M32 works
  // a system whose inner solves have taken escalate_inner FP32 steps in all
  // and whose last inner solve ended above max(escalate_resid, 100x its
  // tolerance) leaves the mixed iteration after that outer step (0: never),
  // for the caller to finish in FP64: its FP32 Krylov process stalls
  int escalate_inner = 0;
  double escalate_resid = 1e-2;
  // outer iteration: false -- FP64 FGMRES (basis of the FP32 corrections,
  // one more FP64 matvec for the final true residual); true -- iterative
  // refinement x += z, z the FP32 inner solution for r = b - A x, with r
  // recomputed in FP64 after every correction (GMRES-IR: the residual each
  // step tests is the true one; at most outer_restart * outer_maxit steps, a
  // system whose residual falls by less than half in a step stops there
  // unconverged, for the caller to finish in FP64)
  bool refine = false;
};
// Workspaces of the solvers, kept by a geometry across calls.  All large
// vectors live in one device arena that only grows: releasing and allocating
// a basis of hundreds of systems on every call would cost more than the solve
// (the driver clears a reused allocation).  layout() carves the FP64 Krylov
// workspace (m64 > 0), the FP32 one (m32 > 0) and, for the mixed solver, the
// FP32 copies of the outer vector and the inner solution.
struct KrylovCache {
  DevBuf<char> arena;
  KrylovWork<double> w64;
  KrylovWork<float> w32;
  float2* v32 = nullptr;
  float2* z32 = nullptr;
  static size_t bytes(int64_t n, int nslot, int m64, bool flex64, int m32, bool flex32, bool mixed_vectors);
  void layout(int64_t n, int nslot, int m64, bool flex64, int m32, bool flex32, bool mixed_vectors);
};

void fgmres_mixed(Context* ctx, int64_t n, const OpFn<double>& A64, const OpFn<float>& A32, const OpFn<float>& M32,
                  const double2* b, double2* x, const std::vector<int>& sys, const MixedOpts& o,
                  KrylovCache& K, std::vector<SysStats>& st);   // K laid out for the slots of sys

}  // namespace vh
