/*
 * voxhenry.h -- C ABI of the MI355X-native VoxHenry solve path.
 *
 * Conventions (all entry points):
 *   - return 0 on success, a positive VH_ERR_* code on failure; the message of
 *     the last failure on the calling thread is returned by vh_last_error();
 *     no C++ exception crosses this boundary;
 *   - complex data are interleaved (re, im) doubles or floats;
 *   - indices are 0-based int64; voxel order is MATLAB column-major order of
 *     the non-empty voxels (idxS of pre_define_structure_params.m:8-13);
 *   - the unknown vector of one right-hand side has n = num_curr + num_node
 *     entries ordered [Jx; Jy; Jz; J2d; J3d; phi] (idxS5, :16);
 *   - pointers flagged "device" are HIP device pointers, all others are host;
 *   - one context per thread; calls on one context are not thread-safe.
 */
#ifndef VOXHENRY_H
#define VOXHENRY_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VH_C128 0 /* complex double */
#define VH_C64 1  /* complex float */

#define VH_ERR_ARG 1
#define VH_ERR_HIP 2
#define VH_ERR_STATE 3
#define VH_ERR_NUMERIC 4
#define VH_ERR_ALLOC 5
#define VH_ERR_INTERNAL 6

typedef struct vh_ctx_s vh_ctx;
typedef struct vh_geom_s vh_geom;
typedef struct vh_precon_s vh_precon;

/* ---- context ------------------------------------------------------------ */
/* Process-wide device context (the MEX gateway keeps one alive with mexLock). */
int vh_ctx_create(int device, vh_ctx** out);
int vh_ctx_destroy(vh_ctx* ctx);
/* Launch all work of this context on `hip_stream` (a hipStream_t, NULL = default stream). */
int vh_ctx_set_stream(vh_ctx* ctx, void* hip_stream);
int vh_ctx_synchronize(vh_ctx* ctx);
const char* vh_last_error(void);
int vh_version(void);

/* ---- geometry -------------------------------------------------------------
 * Grid L x M x N of voxels of side dx with nvox non-empty voxels at the
 * 0-based column-major linear indices idx (strictly increasing).
 * pad_mode 0 pads to exactly (2L,2M,2N) like lse_generate_circulant_tensor.m:776
 * (then 2L,2M,2N must be {2,3,5,7}-smooth); pad_mode 1 picks the smallest
 * smooth length >= 2n-1 per axis (same result on the voxel grid). */
int vh_geom_create(vh_ctx* ctx, int64_t L, int64_t M, int64_t N, double dx, int64_t nvox, const int64_t* idx,
                   int pad_mode, vh_geom** out);
int vh_geom_destroy(vh_geom* g);
/* info[0..9] = nv, num_curr, num_node, Px, Py, Pz, nlines, dense, has_tensor, has_precon-ready */
int vh_geom_info(const vh_geom* g, int64_t* info);

/* Green's tensor, replaces lse_generate_circulant_tensor(dx,1,L,M,N,0)
 * (lse_generate_circulant_tensor.m:1-833): generated on the GPU.  st2 receives
 * st_sparse_precon2 (3 complex, :772). */
int vh_geom_generate_green(vh_geom* g, double* st2);
/* Drop-in alternative: the reference's padded spectrum fN (Px*Py*Pz*7 complex,
 * column-major) with fN*scale equal to the ko=1 spectrum fN_all2. */
int vh_geom_set_green_spectrum(vh_geom* g, const double* fN, int64_t P0, int64_t P1, int64_t P2, double scale_re,
                               double scale_im, const double* st2);
/* Real-space normalised tensor G_mn (L*M*N*7 complex, column-major, host). */
int vh_geom_set_green_realspace(vh_geom* g, const double* G, const double* st2);
/* Copy the real-space tensor generated on the GPU back to the host (L*M*N*7 complex). */
int vh_geom_get_green_realspace(vh_geom* g, double* G);

/* Node-incidence matrix Ar (num_node x 5*nvox) in CSC form (MATLAB sparse),
 * port rows already removed (VoxHenry_executer.m:181-192). */
int vh_geom_set_incidence(vh_geom* g, int64_t num_node, int64_t nnz, const int64_t* colptr, const int64_t* rowind,
                          const double* val);
/* Build the node numbering and incidence matrix of the geometry itself
 * (lse_generate_panelIDs.m, lse_compute_Ae_matrix.m), look up the port faces
 * (lse_assign_exc_grnd_nodes.m), remove the port rows and form the port
 * right-hand sides rhs = -sum of the excitation rows of Ae (VoxHenry_executer.m:
 * 165-192).  Port p's excitation faces are rows P[pP_ptr[p]..pP_ptr[p+1]) of
 * (x, y, z, side) 1-based quadruples (side 1..6 = -x,+x,-y,+y,-z,+z), its ground
 * faces N[pN_ptr[p]..pN_ptr[p+1]); `well` lists extra grounded faces.
 * num_node receives the node count after removal. */
int vh_geom_build_topology(vh_geom* g, int64_t nports, const int64_t* pP_ptr, const int64_t* P, const int64_t* pN_ptr,
                           const int64_t* N, int64_t nwell, const int64_t* well, int64_t* num_node);
/* Port right-hand sides from vh_geom_build_topology: nports rows of n = num_curr + num_node. */
int vh_geom_get_rhs(vh_geom* g, double* rhs);

/* z = z_real + z_imag per non-empty voxel (nvox complex), VoxHenry_executer.m:231-259. */
int vh_geom_set_impedance(vh_geom* g, const double* z);

/* Current coefficients on the voxel grid, replaces
 * post_obtain_curr_coefs_on_grid(x, Mc) (post_obtain_curr_coefs_on_grid.m:1-45)
 * for nrhs solutions x (complex128, stride ldx >= 5*nv; entries 0..5nv-1 are
 * Jx, Jy, Jz, J2d, J3d per non-empty voxel): grid receives, per solution at
 * stride ldg >= 5*L*M*N, the five L x M x N column-major grids Jx_currs_grid ..
 * J3d_currs_grid (zero at empty voxels); minmax (optional, host, 2*nrhs)
 * receives [cmin, cmax] = min / max of |J| over the Jx, Jy and Jz entries
 * (:24-25).  The voxel set is the geometry's idx (the reference's |Mc| >=
 * 1e-12 test selects the same voxels for conductors).  on_device: x and
 * grid are device pointers. */
int vh_geom_currents_on_grid(vh_geom* g, int64_t nrhs, const void* x, int64_t ldx, void* grid, int64_t ldg,
                             double* minmax, int on_device);

/* ---- matvec ---------------------------------------------------------------
 * y = A x, replaces lse_matvect_mult(J, fN_all, Ae, z_real, z_imag, dx, idx, ...)
 * (lse_matvect_mult.m:1-153) for nrhs vectors with leading dimensions ldx, ldy
 * (in complex elements).  jwmu = j*omega*mu scales the Green's product
 * (VoxHenry_executer.m:266).  prec: VH_C128 / VH_C64.  mode 0: full operator;
 * 1: FFT block + resistive diagonal (current rows only); 2: convolution only.
 * on_device != 0: x and y are device pointers. */
int vh_matvec(vh_geom* g, double jwmu_re, double jwmu_im, int prec, int mode, int64_t nrhs, const void* x,
              int64_t ldx, void* y, int64_t ldy, int on_device);

/* Allocation-free variant of vh_matvec.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  x and y
 * are device pointers (nrhs vectors, leading dimensions ldx, ldy); the pass
 * intermediates live in a fixed device array of VH_MATVEC_NOALLOC_BYTES per
 * GPU, held from library load until the process ends.  A call whose
 * intermediates need more (matvec_ws: 5*nrhs*(Px*M*N + Px*Py*N) complex
 * elements on the padded grid) fails with VH_ERR_ARG.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (several streams or
 * threads) race and must be serialised by the caller.  Same operator,
 * arguments and results as vh_matvec with on_device = 1. */
#define VH_MATVEC_NOALLOC_BYTES (64u << 20)
int vh_matvec_noalloc(vh_geom* g, double jwmu_re, double jwmu_im, int prec, int mode, int64_t nrhs, const void* x,
                      int64_t ldx, void* y, int64_t ldy);

/* Timing helper: run the matvec `reps` times on device-resident data and
 * report the average time of one call (ms); if ms_pass is given, a second
 * sweep records events around each fused pass and reports the average time of
 * the passes X, Y, Z, Yi, Xi and the node rows (npass = 6). */
int vh_matvec_timed(vh_geom* g, double jwmu_re, double jwmu_im, int prec, int mode, int64_t nrhs, const void* x,
                    int64_t ldx, void* y, int64_t ldy, int reps, double* ms_total, double* ms_pass, int* npass);

/* Autotuner of the FFT pass shapes (no counterpart in the reference, whose
 * FFT is MATLAB's fftn): times every compiled pass-Z shape (threads per
 * workgroup x elements per thread, which fixes the radices and the kx tile
 * width) on nrhs device vectors of this geometry (convolution only, `reps`
 * calls each) and keeps the fastest for precision prec; then, with that
 * pass-Z shape, every compiled X/Y-pass shape (threads per workgroup x
 * element bound of the x and y FFT passes, which fixes their radices and
 * lines per workgroup).  Shapes that cannot hold the grid's pencils are not
 * timed.  ms_shape (optional, room for *nshape entries on input) receives
 * the average time of one matvec per pass-Z shape followed by one per X/Y
 * shape (HUGE_VAL where not timed); *nshape receives the number of entries
 * written; *best receives the chosen pass-Z index (vh_geom_get_shapes gives
 * both choices). */
int vh_geom_autotune(vh_geom* g, int prec, int64_t nrhs, int reps, int* best, double* ms_shape, int* nshape);
/* Select a pass-Z shape by index (as reported by vh_geom_autotune). */
int vh_geom_set_zshape(vh_geom* g, int prec, int shape);
/* Select an X/Y-pass shape by index (refused when it cannot hold the grid's x or y pencils). */
int vh_geom_set_xyshape(vh_geom* g, int prec, int shape);
/* shapes[0..3] = pass-Z shape (FP32, FP64), X/Y-pass shape (FP32, FP64) in use. */
int vh_geom_get_shapes(vh_geom* g, int* shapes);

/* ---- preconditioner ---------------------------------------------------------
 * Replaces lse_sparse_precon_prepare.m (schur_invert + ldlt_decomp + CHOLMOD)
 * and lse_sparse_precon_multiply.m:61-97.  prepare() computes A_inv =
 * |1/(z/(s dx) + jwmu g_self)| (s = 1,1,1,6,2), S = Ar diag(A_inv) Ar', its
 * host LDL^T (nested-dissection order, symbolic analysis kept across calls) and
 * uploads the factors; apply() computes y = M^-1 x on the device.
 * stats (optional, 4 entries): nnz(L), forward levels, backward levels, n_node. */
int vh_precon_create(vh_geom* g, vh_precon** out);
int vh_precon_destroy(vh_precon* p);
int vh_precon_prepare(vh_precon* p, double jwmu_re, double jwmu_im, int64_t* stats);
int vh_precon_apply(vh_precon* p, int prec, int64_t nrhs, const void* x, int64_t ldx, void* y, int64_t ldy,
                    int on_device);
int vh_precon_get_ainv(vh_precon* p, double* ainv);
/* Preconditioner type, the fl_precon_type switch of lse_sparse_precon_prepare.m
 * (:83-88) and lse_sparse_precon_multiply.m (:26-46):
 *   VH_PC_SCHUR_INVERT (default): exact inverse through the LDL^T factor;
 *   VH_PC_SCHUR_GMRES: the Schur system is solved by an inner GMRES on
 *     S = Ar diag(A_inv) Ar' (restart min(100, num_curr-1), at most 10 cycles,
 *     Jacobi preconditioner diag(S)) to the relative tolerance prectol
 *     (VoxHenry_executer.m:38 uses 0.1); FP64 only, and a nonlinear operator, so
 *     the outer solve is flexible (vh_gmres with flexible = 1, as the executer's
 *     fpgmres at :297). */
#define VH_PC_SCHUR_INVERT 0
#define VH_PC_SCHUR_GMRES 1
int vh_precon_set_type(vh_precon* p, int type, double prectol);
/* Inner Arnoldi steps (summed over the systems) of the last schur_gmres apply. */
int vh_precon_inner_steps(vh_precon* p, int64_t* steps);
/* y = S x = Ar diag(A_inv) Ar' x for nrhs complex128 node vectors (num_node
 * entries, strides ldx / ldy), lse_sparse_schur_compl_multiply.m:1-5. */
int vh_precon_schur_multiply(vh_precon* p, int64_t nrhs, const void* x, int64_t ldx, void* y, int64_t ldy,
                             int on_device);
/* Prepare n preconditioners of one geometry, one per frequency: jwmu holds n
 * interleaved complex values; z (optional) n impedance vectors of nv
 * interleaved complex values each (superconductors: z depends on f), NULL
 * for the geometry's impedance.  The host factorisations run on up to
 * nthreads threads (0: one per core, at most 16); the symbolic analysis is
 * done once.  The per-frequency loop of lse_sparse_precon_prepare.m calls. */
int vh_precon_prepare_many(vh_precon* const* ps, int64_t n, const double* jwmu, const double* z, int nthreads);

/* Host-only sparse LDL^T solve of a symmetric positive definite matrix given in
 * CSC with its full pattern (both triangles), the stand-in for CHOLMOD
 * [LD,p,q] = ldlchol(S) followed by ldlsolve(LD, b) (lse_sparse_precon_prepare.m:154,
 * lse_sparse_precon_multiply.m:86): nested-dissection order, elimination tree,
 * up-looking numeric factorisation.  b, x: n x nrhs real, column-major.
 * xyz (3n) / type (n), optional: voxel-face centre and normal axis of every
 * node; when given, the ordering cuts along face planes.
 * stats (optional, 3 entries): nnz(L), forward levels, backward levels. */
int vh_ldl_solve_host(int64_t n, const int64_t* Ap, const int64_t* Ai, const double* Ax, int64_t nrhs,
                      const double* b, double* x, const float* xyz, const int* type, int64_t* stats);

/* ---- solvers -----------------------------------------------------------------
 * vh_gmres replaces pgmres(A, b, restart, tol, maxit, L1) (pgmres.m:70-197,
 * left preconditioner) or, with flexible != 0, fpgmres(A, b, restart, tol,
 * maxit, R1) (fpgmres.m:58-152), for nrhs right-hand sides b (stride n) that
 * share the operator (same frequency).  p may be NULL (no preconditioner).
 * Outputs per right-hand side: flag, relres, iter[2] = [p outit], niters =
 * Arnoldi steps; resvec (optional) receives up to resvec_cap relative residuals
 * per right-hand side (row stride resvec_cap), nresvec their counts. */
int vh_gmres(vh_geom* g, vh_precon* p, double jwmu_re, double jwmu_im, int prec, int flexible, int64_t nrhs,
             const void* b, void* x, int on_device, int restart, double tol, int maxit, int* flag, double* relres,
             int* iter, int* niters, double* resvec, int64_t resvec_cap, int* nresvec);

/* Mixed precision solve: FP64 flexible GMRES (outer, true FP64 residual)
 * preconditioned by an FP32 left-preconditioned GMRES using the FP32 matvec and
 * the FP32 preconditioner apply.  b, x complex128 (stride n). */
int vh_solve_mixed(vh_geom* g, vh_precon* p, double jwmu_re, double jwmu_im, int64_t nrhs, const double* b,
                   double* x, int on_device, double tol, int outer_restart, int outer_maxit, int inner_restart,
                   double inner_tol, int* flag, double* relres, int* outer_iters, int* inner_iters);

/* Batched solve of nsys (port, frequency) systems of one geometry: right-hand
 * side s is b + s*n (complex128), its j*omega*mu is jwmu[s] (complex) and its
 * preconditioner precons[pidx[s]] (prepared for that frequency, or for a
 * nearby one: any preconditioner gives the same solution).
 * method 0: FP64 left-preconditioned GMRES with the pgmres semantics
 *           (restart, maxit, tol) for every system;
 * method 2: method 0, then every system it leaves with a true residual above
 *           its tolerance (pgmres stops on the preconditioned residual)
 *           continues from its solution in FP64 flexible GMRES with the FP64
 *           preconditioner (restart 50, up to 40 cycles), which
 *           stops on the true residual: the plain FP64 path carried to the
 *           same true-residual tolerance as method 1;
 * method 1: mixed precision, iterative refinement x += z (up to
 *           outer_restart * maxit steps) where z is an FP32 flexible GMRES
 *           solution (inner_restart, inner_tol, relaxed as the residual
 *           falls) of A z = r on the FP32 matvec, preconditioned by the FP32
 *           factors applied with FP32 work vectors, and r = b - A x is
 *           recomputed with the FP64 matvec after every step, so each step
 *           tests the true residual (with the environment variable
 *           VH_MIXED_OUTER=fgmres the outer iteration is instead FP64
 *           flexible GMRES with that inner solve as its preconditioner, as
 *           in vh_solve_mixed).  A system whose inner
 *           solves end a restart cycle above max(1e-2, 100x their tolerance)
 *           (its FP32 Krylov process stalls), a system whose residual falls
 *           by less than half in a step, and any system the outer
 *           iteration leaves above tolerance, continues from its solution in
 *           FP64 flexible GMRES with the FP64 preconditioner (restart 50
 *           within the budget, up to 40 cycles); a system whose
 *           preconditioner was prepared with a kinetic (two-fluid)
 *           impedance runs that FP64 solve from the start (its FP32 process
 *           stalls).  In these FGMRES solves a system whose recomputed
 *           residual misses the tolerance its estimate met restarts from
 *           that residual.
 * All systems run in lock step and converged ones leave the batch; if the
 * Krylov workspaces exceed hbm_budget bytes (0: 80% of free memory) the batch
 * is processed in chunks.  tols (nullable, host) gives system s its own
 * tolerance tols[s] in place of tol: a sweep then follows the frequency rule
 * of VoxHenry_executer.m:219-223 (1e-12 below 1 MHz, 1e-8 above) in one
 * batch.  x_init != 0 (method 1 only): x holds initial guesses and the
 * outer iteration starts from b - A x (e.g. to tighten the tolerance of some
 * systems after a first solve); otherwise x0 = 0 as in pgmres.m.
 * Outputs per system (each nullable): flag, relres (true residual
 * ||b - A x|| / ||b|| of the returned x), iters (outer steps: Arnoldi
 * steps, or refinement steps of method 1, plus any FP64 continuation),
 * inner_iters (FP32 steps, method 1), resid_est (the solver's own residual
 * estimate at exit: method 0 the reference-style preconditioned
 * ||M r|| / ||M b|| that pgmres.m:126 stops on, method 1 the last
 * refinement residual ||r|| / ||b|| (the outer FGMRES estimate with
 * VH_MIXED_OUTER=fgmres)). */
int vh_solve_batch(vh_geom* g, int64_t nprecon, vh_precon* const* precons, int64_t nsys, const double* jwmu,
                   const int* pidx, const double* b, double* x, int on_device, int x_init, int method, double tol,
                   const double* tols, int restart,
                   int maxit, int inner_restart, double inner_tol, double hbm_budget, int* flag, double* relres,
                   int* iters, int* inner_iters, double* resid_est);

/* Device bytes per system that vh_solve_batch budgets for this geometry
 * (Krylov workspace of the method plus the system's share of the matvec
 * workspaces); a caller running several geometries at once passes each its
 * own hbm_budget from these figures instead of letting every call take 80%
 * of the free memory.  No counterpart in the reference (MATLAB allocates per
 * call). */
int vh_solve_batch_bytes(vh_geom* g, int method, int restart, int inner_restart, double* bytes_per_system);

/* Test hook: in-place 3-D FFT (dir -1 forward, +1 inverse unnormalised) of ncomp
 * device grids of the geometry's padded size. */
int vh_fft3d(vh_geom* g, void* data, int ncomp, int prec, int dir);

#ifdef __cplusplus
}
#endif
#endif /* VOXHENRY_H */
