// Preconditioner and Krylov solvers on the device.
#pragma once
#include "vh_geom.h"
#include "vh_ldl.h"
#include <memory>

namespace vh {

// Device pointers to the factor values of one preconditioner, in one precision.
template <typename R> struct PcTab {
  const R* Fx;     // off-block entries of L, packed row by row (forward solve)
  const R* Bx;     // off-block entries of L, packed column by column (backward solve)
  const R* D;
  const R* ainv;
  const R* Li;     // inverses of the diagonal blocks, packed row-major lower
  const R* LiT;    // their transposes, packed row-major upper
  const cplx<R>* z;  // impedance per voxel at this preconditioner's frequency
};

// Device table of the column groups of the last preconditioner apply
// (npg, ng, gtab[ng], gslot[ng * npg]) with its pinned staging.
struct PcGroups {
  std::vector<int> key;
  PinnedBuf<int> h;
  DevBuf<int> d;
  hipEvent_t ev = nullptr;   // completion of the last upload from h
  PcGroups() = default;
  PcGroups(const PcGroups&) = delete;
  PcGroups& operator=(const PcGroups&) = delete;
  ~PcGroups() { if (ev) (void)hipEventDestroy(ev); }
};

// Blocks of the levels of one solve direction: blk[bptr[l] .. bptr[l+1]) are
// the blocks of level l in ascending key = max(size, column union),
// rows[lptr[l] .. lptr[l+1]) their rows in the same order (rowoff[i]: first
// row of blk[i] in rows); smax[i] / umax[i] / emax[i] are the largest size /
// union / entry span among the blocks of its level up to and including blk[i].
// What the block kernel needs of one block: rows [c0, c0+s), column union
// ucol[u0 .. u0+nu), packed inverse at offset off, and the span [e0, e0+ne)
// of the L entries of its rows (forward: CSR rows c0..c0+s-1; backward: CSC
// columns), which holds every off-block entry the block's update reads.
struct BlkInfo {
  int c0, s, u0, nu;
  int64_t off;
  int e0, ne;
};

struct LevelBlocks {
  std::vector<int> bptr, blk, lptr, rows, rowoff, key, smax, umax, emax;
};

// Pattern data of the Schur factor, shared by every frequency of a geometry:
// S pattern, ordering, elimination tree and L pattern, level schedules, and
// their device copies.
struct PreconSymbolic {
  std::vector<int> Sp, Si;            // pattern of S (full symmetric CSR)
  LdlFactor F;                        // perm/iperm/parent/Lp/Li (values unused)
  std::vector<int> csc2csr;           // position of each CSC entry of L in its CSR copy
  int64_t nnzL = 0;
  // supernodes = the dissection blocks: columns [sn_ptr[k], sn_ptr[k+1])
  std::vector<int> blocks, sn_ptr, sn_of_row;
  std::vector<int64_t> sn_off;        // offset of block k in the packed dense inverses
  int64_t dense = 0;                  // total packed entries, sum s(s+1)/2
  std::vector<int> row_split;         // CSR: first entry of row r in its own block's columns
  std::vector<int> col_split;         // CSC: first entry of column r below its own block
  std::vector<int> f_lptr, f_rows, b_lptr, b_rows;   // rows of the blocks of each level
  LevelBlocks fl, bl;                 // forward / backward levels by block
  // column unions per block (forward: the columns left of the block its
  // rows reach; backward: the rows below it its columns reach) and the
  // position of every entry's index in its block's union
  std::vector<int> f_up, f_ucol, f_loc, b_up, b_ucol, b_loc;
  // off-block entries of L (outside the diagonal blocks, the only ones the
  // solves read): forward packed row by row, fp[r] .. fp[r+1] with column fi
  // and union position floc; backward packed column by column (bp, bi,
  // bloc); fsrc / bsrc: CSC positions of the packed entries (value upload)
  std::vector<int> fp, fi, floc, fsrc, bp, bi, bloc, bsrc;
  // lanes per (row, column) item of each level's kernels, from the mean row
  // length (update: entries outside the block; gemv: entries of the inverse)
  std::vector<int> f_lpr_u, f_lpr_g, b_lpr_u, b_lpr_g;
  std::vector<int64_t> f_upd_nnz;     // update entries of each forward level (0: no update launch)
  DevBuf<char> ws;                    // apply workspace (two [group][node][port] vectors), shared
  PcGroups groups;
  bool plan_traced = false;           // VH_TRACE: the launch plan was printed
  DevBuf<int> d_perm, d_iperm, d_fp, d_fi, d_floc, d_bp, d_bi, d_bloc, ArTp_col;
  DevBuf<int> d_sn_ptr, d_sn_of_row, d_f_rows, d_b_rows;
  DevBuf<BlkInfo> d_f_info, d_b_info;  // per sorted block position (LevelBlocks order)
  DevBuf<int> d_f_ucol, d_b_ucol;
  DevBuf<int64_t> d_sn_off;
};

struct SchurGmres;   // workspaces of the schur_gmres apply (vh_schur.hip)

// Preconditioner types of lse_sparse_precon_prepare.m (fl_precon_type):
// schur_invert (default) applies the exact inverse through the LDL^T factor;
// schur_gmres solves the Schur system by an inner GMRES on S with a Jacobi
// preconditioner to the relative tolerance prectol.
enum PreconType { PC_SCHUR_INVERT = 0, PC_SCHUR_GMRES = 1 };

struct Precon {
  Geometry* G = nullptr;
  bool ready = false;
  int type = PC_SCHUR_INVERT;
  double prectol = 0.1;               // VoxHenry_executer.m:38
  std::shared_ptr<SchurGmres> schur;
  int64_t inner_steps = 0;            // schur_gmres: inner Arnoldi steps of the last apply (all systems)
  bool geometric_order = true;        // nested dissection by voxel-face planes
  std::shared_ptr<PreconSymbolic> sym;
  LdlFactor F;                        // numeric values (Lx, D) of this frequency
  std::vector<double> ainv_h;
  // device values: L (off-block entries packed per solve direction), D,
  // A_inv, inverses of the diagonal blocks
  // (row-major lower: Linv; the transpose row-major: LinvT)
  DevBuf<double> Fx64, Bx64, D64, ainv64, Li64, LiT64;
  DevBuf<float> Fx32, Bx32, D32, ainv32, Li32, LiT32;
  std::vector<std::complex<double>> zh;   // impedance per voxel it was prepared with
  bool kinetic = false;               // some voxel's impedance has a kinetic (imaginary) part: a two-fluid conductor
  DevBuf<double2> z64;                // device copies of zh (the batched matvec reads them per system)
  DevBuf<float2> z32;
  DevBuf<PcTab<double>> tab64;        // one-entry tables for the single apply
  DevBuf<PcTab<float>> tab32;
  PcTab<double> tab64_h{};
  PcTab<float> tab32_h{};
};

// A_inv, S, LDL^T for jwmu = j*omega*mu (lse_sparse_precon_prepare.m) with
// the impedance z (nv values; null: the geometry's current impedance).
// Several preconditioners of one geometry may be prepared concurrently once
// the geometry's symbolic analysis exists (precon_prepare_many does that).
void precon_prepare(Precon& P, double jwmu_re, double jwmu_im, const std::complex<double>* z = nullptr);
// Prepare n preconditioners (jwmu[k], impedance z + k*nv or the geometry's)
// on up to nthreads host threads.
void precon_prepare_many(Precon* const* P, int n, const std::complex<double>* jwmu, const std::complex<double>* z,
                         int nthreads);
// y = M^-1 x for the vectors in the given slots (vector s at x + s*ldx,
// device memory), lse_sparse_precon_multiply.m:84-88.
template <typename R>
void precon_apply(Precon& P, const std::vector<int>& slots, const cplx<R>* x, int64_t ldx, cplx<R>* y, int64_t ldy);
// schur_gmres apply (lse_sparse_precon_multiply.m:26-46) for the vectors in
// the given slots, FP64; inner_steps (optional) receives the inner Arnoldi
// steps summed over the systems.
void schur_gmres_apply(Precon& P, const std::vector<int>& slots, const double2* x, int64_t ldx, double2* y,
                       int64_t ldy, int64_t* inner_steps = nullptr);
// y = S x = Ar diag(A_inv) Ar' x for nrhs node vectors (stride ldx / ldy,
// device), lse_sparse_schur_compl_multiply.m:1-5.
void schur_multiply(Precon& P, int nrhs, const double2* x, int64_t ldx, double2* y, int64_t ldy);
// Several preconditioners of one geometry (sharing Y) applied in one pass:
// slot s uses tab[pidx[s]] (tab[0] when pidx is null); tab is a device
// array, pidx a host array indexed by slot.
// The same with work vectors of type WT (precon_apply_cols: double).  With
// FP32 factors and WT = float the apply is exact only to FP32 rounding at
// every level, so it is not a fixed linear operator: callers use it inside a
// flexible Krylov solve.
template <typename R, typename WT>
void precon_apply_cols_wt(Geometry& G, PreconSymbolic& Y, const PcTab<R>* tab, const int* pidx,
                          const std::vector<int>& slots, const cplx<R>* x, int64_t ldx, cplx<R>* y, int64_t ldy);
template <typename R>
void precon_apply_cols(Geometry& G, PreconSymbolic& Y, const PcTab<R>* tab, const int* pidx,
                       const std::vector<int>& slots, const cplx<R>* x, int64_t ldx, cplx<R>* y, int64_t ldy);

}  // namespace vh
