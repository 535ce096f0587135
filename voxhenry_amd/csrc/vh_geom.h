// Device-resident description of one voxel geometry: grid, padded FFT grid,
// voxel map, Green's-tensor spectrum (one octant), impedance diagonal and the
// node-incidence matrix Ar in CSR form (rows = nodes) and its transpose.
#pragma once
#include "vh_common.h"
#include "vh_fft.h"
#include "vh_fft2.h"
#include <memory>
#include <mutex>

namespace vh {

constexpr int TENSOR_STRIDE = 8;   // stored components per spectral point (7 + 1 pad)

struct GridDims {
  int L, M, N;          // voxel grid
  int Px, Py, Pz;       // padded FFT grid (>= 2L-1, 2M-1, 2N-1)
  int Hx, Hy, Hz;       // stored spectral octant: P/2 + 1 per axis
  int nv;               // non-empty voxels
};

// Sparse matrix in CSR with 32-bit indices, values in both precisions.
struct DevCsr {
  int rows = 0, cols = 0, nnz = 0;
  DevBuf<int> ptr, col;
  DevBuf<double> v64;
  DevBuf<float> v32;
};

// Per-pass timing: when Geometry::timer is set, matvec records an event before
// and after every pass (X, Y, Z, Yi, Xi, node rows) and accumulates the times.
// Layout of the packed Ar' rows (per current component: offset, entries per current).
struct PackedInc {
  int off[5];
  int k[5];
};

struct PassTimer {
  static constexpr int NP = 6;
  hipEvent_t ev[NP + 1] = {};
  double ms[NP] = {0, 0, 0, 0, 0, 0};
  int calls = 0;
  bool pending = false;
  void create();
  void destroy();
  void collect();   // waits for the last recorded call and adds its pass times
};

struct PreconSymbolic;
struct KrylovCache;

struct Geometry {
  Context* ctx = nullptr;
  GridDims g{};
  double dx = 0;
  int64_t num_curr = 0, num_node = 0;
  bool dense = false;                 // every voxel of the grid is non-empty
  std::vector<int64_t> vox;           // host copy: 0-based column-major linear index per voxel
  DevBuf<int> vmap;                   // L*M*N -> voxel id or -1
  DevBuf<int> lines;                  // non-empty (y + M z) lines
  // per non-empty line: {x0, x1, vid(x0) - x0, 0} when its voxels are the one
  // run x0 <= x < x1 (voxel id = x + vid(x0) - x0), else {-1, 0, 0, 0}
  DevBuf<int4> runs;
  // per z plane, 4 words: bit y set when the x line (y, z) holds voxels; used
  // when M <= 256 and some line is empty (row_bits_on)
  DevBuf<unsigned long long> rowbits;
  bool row_bits_on = false;
  int nlines = 0;
  FftAxis ax, ay, az;                 // Green's tensor transforms
  FftAxis2 fx, fy, fz;                // matvec passes
  // Green's tensor spectrum of the normalised real-space tensor (ko = 1),
  // divided by Px*Py*Pz: one octant [kz][ky][kx] of Hx*Hy*Hz points, the 7
  // components of a point contiguous (padded to TENSOR_STRIDE).
  DevBuf<double2> t64;                // complex, exact
  DevBuf<float> t32;                  // real part of even comps / imag part of odd comps
  bool has_tensor = false;
  double st2[6] = {0, 0, 0, 0, 0, 0}; // self terms G(1,1,1,{1,4,7}) as complex pairs
  // impedance z = z_real + z_imag per voxel (complex)
  DevBuf<double2> z64;
  DevBuf<float2> z32;
  bool has_z = false;
  DevCsr Ar;    // num_node x num_curr
  DevCsr ArT;   // num_curr x num_node
  // Ar' rows packed for pass Xi: component c holds pinc.k[c] entries per
  // current from pinc.off[c], entry = node << 3 | code (code 0 none, 1 +1,
  // 2 -1, 3 +1/2, 4 -1/2); empty when Ar does not fit that form
  DevBuf<int> atpk;
  PackedInc pinc{};
  // host copies used by the preconditioner set-up
  std::vector<int> arp_h, arc_h, atp_h, atc_h;
  std::vector<double> arv_h, atv_h;
  std::vector<std::complex<double>> zh;
  bool has_inc = false;
  int64_t num_nodes_full = 0;         // nodes before the port rows were removed
  std::vector<double> rhs_h;          // port right-hand sides (nports x n), from build_topology
  int nports = 0;
  PassTimer* timer = nullptr;
  int zshape[2] = {0, 0};             // pass-Z shape index per precision [FP32, FP64] (autotuned)
  int xyshape[2] = {0, 0};            // X/Y-pass shape index per precision [FP32, FP64] (autotuned)
  bool use_p2z = true;                // compile-time pass-Z kernels for power-of-two Pz
  std::shared_ptr<PreconSymbolic> precon_sym;   // shared by all preconditioners of this geometry
  std::mutex precon_sym_mu;                     // guards its construction
  std::shared_ptr<KrylovCache> krylov;          // solver workspaces kept across calls
  // matvec workspaces, grown on demand.  A (pass X output) relies on its
  // never-written empty lines reading as zero, so each precision has its own.
  DevBuf<char> wsA64, wsA32, wsB;

  size_t octant() const { return (size_t)g.Hx * g.Hy * g.Hz; }
  int64_t n() const { return num_curr + num_node; }
};

// Matvec y = A x for nrhs vectors (x, y device pointers, leading dimension ld
// in complex elements).  alpha = j*omega*mu multiplies the Green's product.
// mode 0: full operator [Z, -Ar'; Ar, 0]; mode 1: the FFT block with the
// resistive diagonal only (current rows); mode 2: convolution only.
// slots (device, optional): vector k of the batch is x + slots[k]*ldx.
// alphas (device, optional): j*omega*mu per vector slot (overrides alpha).
// zslot (device, optional): impedance vector per vector slot (overrides the
// geometry's impedance; frequency-dependent conductors in a batch).
template <typename R>
void matvec(Geometry& G, int nrhs, const cplx<R>* x, int64_t ldx, cplx<R>* y, int64_t ldy,
            double alpha_re, double alpha_im, int mode = 0, const int* slots = nullptr,
            const double2* alphas = nullptr, const cplx<R>* const* zslot = nullptr);

// The same operator on a caller-owned workspace of matvec_ws_bytes(G, prec,
// nrhs) bytes (256-byte aligned device memory; prec 0 = FP64, 1 = FP32): no
// allocation, unit slots, the geometry's j*omega*mu scaling by (are, aim) and
// impedance.  The workspace is cleared where the passes rely on zeros.
size_t matvec_ws_bytes(const Geometry& G, int prec, int nrhs);
template <typename R>
void matvec_ws(Geometry& G, int nrhs, const cplx<R>* x, int64_t ldx, cplx<R>* y, int64_t ldy, double are,
               double aim, int mode, char* ws);
// Device address of this library's fixed scratch array on `device` (looked
// up once per device by scratch_init, with that device current) and its size.
void scratch_init(int device);
char* scratch_ptr(int device);
size_t scratch_bytes();

// Build the spectral tensor on the device from a real-space normalised tensor
// G (L*M*N*7 complex, device, column-major [x][y][z][c]).
void tensor_from_realspace(Geometry& G, const double2* Greal_dev);
// Compute the real-space tensor on the device (quadrature + near table).
void generate_realspace_green(Geometry& G, DevBuf<double2>& out);
// In-place 3-D FFT of ncomp padded grids (test hook); prec 0 = double, 1 = float.
void fft3d_test(Geometry& Gm, void* data, int ncomp, int prec, int dir);




// geometry setup helpers
void geometry_init(Geometry& G, Context* ctx, int L, int M, int N, double dx, int64_t nv,
                   const int64_t* idx, int pad_mode);
void geometry_set_incidence_csc(Geometry& G, int64_t num_node, int64_t nnz, const int64_t* colptr,
                                const int64_t* rowind, const double* val);
void geometry_set_impedance(Geometry& G, const double* z_interleaved);
// Panel ids, Ae, port rows removed (Ar, set as the incidence) and the port
// right-hand sides (rows of length num_curr + num_node).  Face lists hold
// (x, y, z, side) 1-based quadruples, side 1..6 = -x,+x,-y,+y,-z,+z.
void geometry_build_topology(Geometry& G, const std::vector<std::vector<int64_t>>& Pfaces,
                             const std::vector<std::vector<int64_t>>& Nfaces, const std::vector<int64_t>& well,
                             std::vector<double>& rhs_out);
// Current coefficients of nrhs solutions x (device, stride ldx) on the voxel
// grid: grid (device, stride ldg >= 5*L*M*N) receives Jx, Jy, Jz, J2d, J3d as
// five column-major L x M x N blocks, zero at empty voxels; minmax (host,
// optional) the min and max of |J| over the Jx, Jy, Jz entries per solution.
void currents_on_grid(Geometry& G, int nrhs, const double2* x, int64_t ldx, double2* grid, int64_t ldg,
                      double* minmax);
// Number of compiled pass-Z shapes of a precision (prec 0 = FP64, 1 = FP32).
int matvec_num_zshapes(int prec);
// Whether pass-Z shape `shape` changes anything on this grid: false when its
// threads x elements cannot hold one z pencil, or when the grid's z length is
// served by the one-radix kernel, which has no shape.
bool matvec_zshape_applies(const Geometry& G, int prec, int shape);
// Compiled X/Y-pass shapes (threads x element bound of the x and y FFT
// passes) and whether one can hold the grid's x and y pencils.
int matvec_num_xyshapes(int prec);
bool matvec_xyshape_applies(const Geometry& G, int prec, int shape);
void geometry_set_spectrum_full(Geometry& G, const double* fN, int64_t P0, int64_t P1, int64_t P2,
                                double scale_re, double scale_im);

}  // namespace vh
