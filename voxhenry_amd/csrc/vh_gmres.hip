// Krylov solvers on the device.
//
// * pgmres  -- src_iterative_solver/pgmres.m:70-197 + pgmres_iter.m:22-153 with
//   the same semantics: x0 = 0, left preconditioner, bnorm = ||M b||, one-pass
//   classical Gram-Schmidt, stop when the least-squares residual drops below
//   tol*bnorm, restart from r = V(:,1:k+1)*res, final true relres.  The small
//   least-squares problem is solved on the device by Givens rotations (same
//   residuals as the reference's Hc\rhs up to rounding).  Runs a batch of
//   independent systems in lock step; a system that converges is retired from
//   the batch (slot map).
// * fpgmres -- fpgmres.m:58-152 + fpgmres_iter.m:26-121 (flexible, right
//   preconditioned, bnorm = ||b||).
// * fgmres_mixed -- FP32 flexible GMRES solves (FP32 matvec and
//   preconditioner) inside an FP64 outer iteration that recovers the
//   accuracy: iterative refinement on the FP64 residual b - A x, recomputed
//   after every correction, or FP64 flexible GMRES with the FP32 solve as
//   its preconditioner (vh_solve.hip).
//
// Vectors of one batch live in slots of length ld: vector s of basis j is at
// V + j*vstride + s*ld.
#include "vh_krylov.h"
#include <algorithm>
#include <cmath>

namespace vh {

constexpr int KT = 256;       // threads per block of the vector kernels
constexpr int EPT = 8;        // elements per thread
constexpr int CHUNK = KT * EPT;
constexpr int KMAX = 128;     // largest Krylov basis of one cycle (restart + 1)

template <typename R> __device__ __forceinline__ double2 to_d2(cplx<R> v) { double2 d; d.x = v.x; d.y = v.y; return d; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// part[(a*kmax + j)*nblk + b] = sum over the chunks of block b of conj(V_j) w.
// Blocks stride over chunks; each wave accumulates its own LDS row, so the
// summation order is fixed (deterministic, no atomics).
// need (optional): systems with need[a] == 0 are skipped (their partials are
// not written; the reduction gives them zeros).  with_norm: entry k receives
// ||w||^2 from the same pass over w.
template <typename R>
__global__ __launch_bounds__(KT) void k_mdot(const cplx<R>* __restrict__ V, int64_t vstride,
                                              const cplx<R>* __restrict__ w, int64_t ld, int64_t n, int k, int kmax,
                                              const int* __restrict__ slots, double2* __restrict__ part, int nchunk,
                                              const int* __restrict__ need, int with_norm) {
  __shared__ double2 acc[4][KMAX];
  const int a = blockIdx.y;
  if (need && !need[a]) return;
  const int s = slots[a];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int j = tid; j < 4 * KMAX; j += KT) { acc[j / KMAX][j % KMAX].x = 0; acc[j / KMAX][j % KMAX].y = 0; }
  __syncthreads();
  for (int c = blockIdx.x; c < nchunk; c += gridDim.x) {
    const int64_t base = (int64_t)c * CHUNK;
    cplx<R> wv[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int64_t i = base + e * KT + tid;
      if (i < n) wv[e] = w[(int64_t)s * ld + i]; else { wv[e].x = 0; wv[e].y = 0; }
    }
    for (int j = 0; j < k; ++j) {
      const cplx<R>* vj = V + j * vstride + (int64_t)s * ld;
      double sx = 0, sy = 0;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int64_t i = base + e * KT + tid;
        if (i < n) {
          const cplx<R> v = vj[i];
          sx += (double)v.x * wv[e].x + (double)v.y * wv[e].y;
          sy += (double)v.x * wv[e].y - (double)v.y * wv[e].x;
        }
      }
      sx = wave_sum(sx);
      sy = wave_sum(sy);
      if (lane == 0) { acc[wave][j].x += sx; acc[wave][j].y += sy; }
    }
    if (with_norm) {
      double sq = 0;
#pragma unroll
      for (int e = 0; e < EPT; ++e) sq += (double)wv[e].x * wv[e].x + (double)wv[e].y * wv[e].y;
      sq = wave_sum(sq);
      if (lane == 0) acc[wave][k].x += sq;
    }
  }
  __syncthreads();
  for (int j = tid; j < k + with_norm; j += KT) {
    double2 r;
    r.x = ((acc[0][j].x + acc[1][j].x) + acc[2][j].x) + acc[3][j].x;
    r.y = ((acc[0][j].y + acc[1][j].y) + acc[2][j].y) + acc[3][j].y;
    part[((int64_t)a * kmax + j) * gridDim.x + blockIdx.x] = r;
  }
}

// out[o] = sum_{q < np} part[o*np + q] with o = blockIdx.y*ostride + blockIdx.x
// (fixed summation order: deterministic).  need (optional, per system a =
// blockIdx.y, or blockIdx.x when ostride is 0): skipped systems get zero
// (keep == 0) or keep their out value (keep != 0).
__global__ __launch_bounds__(256) void k_reduce2(const double2* __restrict__ part, int np, double2* __restrict__ out,
                                                 int ostride, const int* __restrict__ need, int keep) {
  __shared__ double sx[256], sy[256];
  const int o = blockIdx.y * ostride + blockIdx.x;
  if (need && !need[ostride ? blockIdx.y : blockIdx.x]) {
    if (!keep && threadIdx.x == 0) out[o] = make_double2(0.0, 0.0);
    return;
  }
  double ax = 0, ay = 0;
  for (int q = threadIdx.x; q < np; q += 256) { const double2 v = part[(int64_t)o * np + q]; ax += v.x; ay += v.y; }
  sx[threadIdx.x] = ax; sy[threadIdx.x] = ay;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) { sx[threadIdx.x] += sx[threadIdx.x + h]; sy[threadIdx.x] += sy[threadIdx.x + h]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { double2 r; r.x = sx[0]; r.y = sy[0]; out[o] = r; }
}

// out[s] = beta*out[s] + sign * sum_j V_j[s] coef[a][j]; optional per-block
// partial of ||out||^2 at npart[a*nblk + b].
template <typename R>
__global__ __launch_bounds__(KT) void k_combine(const cplx<R>* __restrict__ V, int64_t vstride, cplx<R>* __restrict__ out,
                                                 int64_t ld, int64_t n, int k, int kmax, const int* __restrict__ slots,
                                                 const double2* __restrict__ coef, double beta, double sign,
                                                 double2* __restrict__ npart, int nchunk, const int* __restrict__ need) {
  __shared__ double2 cf[KMAX];
  __shared__ double wn[4];
  const int a = blockIdx.y;
  if (need && !need[a]) return;
  const int s = slots[a];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int j = tid; j < k; j += KT) { double2 v = coef[(int64_t)a * kmax + j]; v.x *= sign; v.y *= sign; cf[j] = v; }
  __syncthreads();
  double nrm = 0;
  for (int c = blockIdx.x; c < nchunk; c += gridDim.x) {
    const int64_t base = (int64_t)c * CHUNK;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int64_t i = base + e * KT + tid;
      if (i < n) {
        double ax = 0, ay = 0;
        if (beta != 0) { const cplx<R> o = out[(int64_t)s * ld + i]; ax = beta * o.x; ay = beta * o.y; }
        for (int j = 0; j < k; ++j) {
          const cplx<R> v = V[j * vstride + (int64_t)s * ld + i];
          ax += (double)v.x * cf[j].x - (double)v.y * cf[j].y;
          ay += (double)v.x * cf[j].y + (double)v.y * cf[j].x;
        }
        cplx<R> r; r.x = (R)ax; r.y = (R)ay;
        out[(int64_t)s * ld + i] = r;
        nrm += (double)r.x * r.x + (double)r.y * r.y;
      }
    }
  }
  if (npart) {
    nrm = wave_sum(nrm);
    if (lane == 0) wn[wave] = nrm;
    __syncthreads();
    if (tid == 0) { double2 r; r.x = ((wn[0] + wn[1]) + wn[2]) + wn[3]; r.y = 0; npart[(int64_t)a * gridDim.x + blockIdx.x] = r; }
  }
}

// Rows a thread of the fused Gram-Schmidt kernels loads at once: UE rows of
// KB basis elements (32 registers of complex64 data), all loads issued before
// the arithmetic so every wave keeps UE*(KB+1) loads in flight.
template <int KB> constexpr int fused_rows() { return KB >= 16 ? 1 : (KB >= 8 ? 2 : 4); }

// First half of a fused two-pass classical Gram-Schmidt step (CGS2):
// w <- w - sum_j V_j h1_j in place and, from the same pass over the basis,
// the partials of h2_j = V_j^H w (the second pass's projections, j < k) and
// of ||w||^2 (entry k), taken on the stored (rounded) w.  The KB + 1 running
// sums of a thread stay in registers and meet in the wave and the block once,
// at the end, in a fixed order (deterministic).
template <typename R, int KB>
__global__ __launch_bounds__(KT) void k_combine_dot(const cplx<R>* __restrict__ V, int64_t vstride,
                                                     cplx<R>* __restrict__ w, int64_t ld, int64_t n, int k, int kmax,
                                                     const int* __restrict__ slots, const double2* __restrict__ h1,
                                                     double2* __restrict__ part, int nchunk) {
  constexpr int UE = fused_rows<KB>();
  static_assert(EPT % UE == 0, "rows per load group must divide the rows per thread");
  __shared__ double2 cf[KB];
  __shared__ double2 red[4][KB + 1];
  const int a = blockIdx.y;
  const int s = slots[a];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid < KB) cf[tid] = tid < k ? h1[(int64_t)a * kmax + tid] : make_double2(0.0, 0.0);
  __syncthreads();
  double ax_[KB + 1], ay_[KB + 1];
#pragma unroll
  for (int j = 0; j <= KB; ++j) { ax_[j] = 0; ay_[j] = 0; }
  const cplx<R>* Vs = V + (int64_t)s * ld;
  cplx<R>* ws = w + (int64_t)s * ld;
  for (int c = blockIdx.x; c < nchunk; c += gridDim.x) {
    const int64_t base = (int64_t)c * CHUNK;
#pragma unroll 1
    for (int e0 = 0; e0 < EPT; e0 += UE) {
      cplx<R> vr[UE][KB], wv[UE];
#pragma unroll
      for (int u = 0; u < UE; ++u) {
        const int64_t i = base + (e0 + u) * KT + tid;
        if (i < n) {
          // 32-bit element offset from a per-vector uniform base (n < 2^31):
          // scalar base + vector offset addressing, no 64-bit address per lane
          const uint32_t ii = (uint32_t)i;
          wv[u] = ws[ii];
#pragma unroll
          for (int j = 0; j < KB; ++j)
            if (j < k) vr[u][j] = (Vs + j * vstride)[ii];
        }
      }
#pragma unroll
      for (int u = 0; u < UE; ++u) {
        const int64_t i = base + (e0 + u) * KT + tid;
        if (i < n) {
          double x = wv[u].x, y = wv[u].y;
#pragma unroll
          for (int j = 0; j < KB; ++j)
            if (j < k) {
              x -= (double)vr[u][j].x * cf[j].x - (double)vr[u][j].y * cf[j].y;
              y -= (double)vr[u][j].x * cf[j].y + (double)vr[u][j].y * cf[j].x;
            }
          cplx<R> r; r.x = (R)x; r.y = (R)y;
          ws[i] = r;
          const double rx = r.x, ry = r.y;
#pragma unroll
          for (int j = 0; j < KB; ++j)
            if (j < k) {
              ax_[j] += (double)vr[u][j].x * rx + (double)vr[u][j].y * ry;
              ay_[j] += (double)vr[u][j].x * ry - (double)vr[u][j].y * rx;
            }
          ax_[KB] += rx * rx + ry * ry;
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j <= KB; ++j) {
    if (j < k || j == KB) {
      const double sx = wave_sum(ax_[j]), sy = wave_sum(ay_[j]);
      if (lane == 0) red[wave][j] = make_double2(sx, sy);
    }
  }
  __syncthreads();
  for (int j = tid; j <= k; j += KT) {
    const int q = j < k ? j : KB;
    double2 r;
    r.x = ((red[0][q].x + red[1][q].x) + red[2][q].x) + red[3][q].x;
    r.y = ((red[0][q].y + red[1][q].y) + red[2][q].y) + red[3][q].y;
    part[((int64_t)a * kmax + j) * gridDim.x + blockIdx.x] = r;
  }
}

// Second half of the fused CGS2 step: out = (w - sum_j V_j h2_j) / nu with
// nu^2 = ||w||^2 - sum_j |h2_j|^2 (h2[a][k] = ||w||^2 from k_combine_dot; the
// basis is orthonormal, so the norm of the corrected vector follows without
// another pass over it); nrm2[a] = nu^2.  Every block derives nu from the
// same values in the same order; nu = 0 (breakdown) gives a zero vector.
template <typename R, int KB>
__global__ __launch_bounds__(KT) void k_combine_scale(const cplx<R>* __restrict__ V, int64_t vstride,
                                                       const cplx<R>* __restrict__ w, cplx<R>* __restrict__ out,
                                                       int64_t ld, int64_t n, int k, int kmax,
                                                       const int* __restrict__ slots, const double2* __restrict__ h2,
                                                       double2* __restrict__ nrm2, int nchunk) {
  constexpr int UE = fused_rows<KB>();
  __shared__ double2 cf[KB];
  __shared__ double inv_s;
  const int a = blockIdx.y;
  const int s = slots[a];
  const int tid = threadIdx.x;
  if (tid < KB) cf[tid] = tid < k ? h2[(int64_t)a * kmax + tid] : make_double2(0.0, 0.0);
  __syncthreads();
  if (tid == 0) {
    double nu2 = h2[(int64_t)a * kmax + k].x;
    for (int j = 0; j < k; ++j) nu2 -= cf[j].x * cf[j].x + cf[j].y * cf[j].y;
    nu2 = nu2 > 0 ? nu2 : 0.0;
    inv_s = nu2 > 0 ? 1.0 / sqrt(nu2) : 0.0;
    if (blockIdx.x == 0) nrm2[a] = make_double2(nu2, 0.0);
  }
  __syncthreads();
  const double inv = inv_s;
  const cplx<R>* Vs = V + (int64_t)s * ld;
  const cplx<R>* ws = w + (int64_t)s * ld;
  cplx<R>* os = out + (int64_t)s * ld;
  for (int c = blockIdx.x; c < nchunk; c += gridDim.x) {
    const int64_t base = (int64_t)c * CHUNK;
#pragma unroll 1
    for (int e0 = 0; e0 < EPT; e0 += UE) {
      cplx<R> vr[UE][KB], wv[UE];
#pragma unroll
      for (int u = 0; u < UE; ++u) {
        const int64_t i = base + (e0 + u) * KT + tid;
        if (i < n) {
          // 32-bit element offset from a per-vector uniform base (n < 2^31):
          // scalar base + vector offset addressing, no 64-bit address per lane
          const uint32_t ii = (uint32_t)i;
          wv[u] = ws[ii];
#pragma unroll
          for (int j = 0; j < KB; ++j)
            if (j < k) vr[u][j] = (Vs + j * vstride)[ii];
        }
      }
#pragma unroll
      for (int u = 0; u < UE; ++u) {
        const int64_t i = base + (e0 + u) * KT + tid;
        if (i < n) {
          double x = wv[u].x, y = wv[u].y;
#pragma unroll
          for (int j = 0; j < KB; ++j)
            if (j < k) {
              x -= (double)vr[u][j].x * cf[j].x - (double)vr[u][j].y * cf[j].y;
              y -= (double)vr[u][j].x * cf[j].y + (double)vr[u][j].y * cf[j].x;
            }
          cplx<R> r; r.x = (R)(x * inv); r.y = (R)(y * inv);
          os[i] = r;
        }
      }
    }
  }
}

// Selective second Gram-Schmidt pass: system a needs it when the first pass
// removed more than half of ||w||^2 (||w'||^2 < eta2 ||w||^2, the
// Kahan-Parlett criterion with eta = 1/sqrt(2)): then w' carries too little
// of w for its rounding error to stay small against it.
__global__ void k_reorth_flags(const double2* __restrict__ h, int kmax, int k, const double2* __restrict__ nrm2,
                               int* __restrict__ need, int nact, double eta2) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nact) return;
  need[a] = nrm2[a].x < eta2 * h[(int64_t)a * kmax + k].x ? 1 : 0;
}

// dst[s] = src[s] * (1/sqrt(nrm2[a]))  (scale by the inverse norm computed on the device)
template <typename R>
__global__ void k_scale(const cplx<R>* __restrict__ src, cplx<R>* __restrict__ dst, int64_t ld, int64_t n,
                        const int* __restrict__ slots, const double2* __restrict__ nrm2) {
  const int a = blockIdx.y;
  const int s = slots[a];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double nv = sqrt(nrm2[a].x);
  const R inv = nv > 0 ? (R)(1.0 / nv) : R(0);
  cplx<R> v = src[(int64_t)s * ld + i];
  v.x *= inv; v.y *= inv;
  dst[(int64_t)s * ld + i] = v;
}

template <typename RD, typename RS>
__global__ void k_convert(const cplx<RS>* __restrict__ src, int64_t lds, cplx<RD>* __restrict__ dst, int64_t ldd,
                          int64_t n, const int* __restrict__ slots_src, const int* __restrict__ slots_dst) {
  const int a = blockIdx.y;
  const int ss = slots_src ? slots_src[a] : a, sd = slots_dst ? slots_dst[a] : a;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cplx<RS> v = src[(int64_t)ss * lds + i];
  cplx<RD> o; o.x = (RD)v.x; o.y = (RD)v.y;
  dst[(int64_t)sd * ldd + i] = o;
}

// v[s] = (RD)(b[s] - y[s]) with per-block partials of ||b - y||^2 taken in
// the source precision before the rounding: npart[a*nblk + b] (the
// residual of an iterative-refinement step handed to a lower-precision
// solve in one pass over b and y)
template <typename RD, typename RS>
__global__ __launch_bounds__(KT) void k_residual_to(const cplx<RS>* __restrict__ b, const cplx<RS>* __restrict__ y,
                                                     cplx<RD>* __restrict__ v, int64_t ld, int64_t n,
                                                     const int* __restrict__ slots, double2* __restrict__ npart,
                                                     int nchunk) {
  __shared__ double wn[4];
  const int a = blockIdx.y;
  const int s = slots[a];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const cplx<RS>* bs = b + (int64_t)s * ld;
  const cplx<RS>* ys = y + (int64_t)s * ld;
  cplx<RD>* vs = v + (int64_t)s * ld;
  double nrm = 0;
  for (int c = blockIdx.x; c < nchunk; c += gridDim.x) {
    const int64_t base = (int64_t)c * CHUNK;
    cplx<RS> bb[EPT], yy[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int64_t i = base + e * KT + tid;
      if (i < n) { bb[e] = bs[i]; yy[e] = ys[i]; }
    }
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int64_t i = base + e * KT + tid;
      if (i < n) {
        const double rx = (double)bb[e].x - (double)yy[e].x, ry = (double)bb[e].y - (double)yy[e].y;
        nrm += rx * rx + ry * ry;
        cplx<RD> o; o.x = (RD)rx; o.y = (RD)ry;
        vs[i] = o;
      }
    }
  }
  nrm = wave_sum(nrm);
  if (lane == 0) wn[wave] = nrm;
  __syncthreads();
  if (tid == 0) { double2 r; r.x = ((wn[0] + wn[1]) + wn[2]) + wn[3]; r.y = 0; npart[(int64_t)a * gridDim.x + blockIdx.x] = r; }
}

template <typename RD, typename RS>
__global__ void k_accumulate(const cplx<RS>* __restrict__ src, int64_t lds, cplx<RD>* __restrict__ dst, int64_t ldd,
                             int64_t n, const int* __restrict__ slots) {
  const int s = slots[blockIdx.y];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cplx<RS> v = src[(int64_t)s * lds + i];
  cplx<RD> o = dst[(int64_t)s * ldd + i];
  o.x += (RD)v.x; o.y += (RD)v.y;
  dst[(int64_t)s * ldd + i] = o;
}

template <typename R>
__global__ void k_zero_slots(cplx<R>* __restrict__ x, int64_t ld, int64_t n, const int* __restrict__ slots) {
  const int s = slots[blockIdx.y];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cplx<R> z;
  z.x = 0;
  z.y = 0;
  x[(int64_t)s * ld + i] = z;
}

// y[s] = x[s] - y[s]   (residual r = b - A x given y = A x)
template <typename R>
__global__ void k_sub_from(const cplx<R>* __restrict__ b, int64_t ldb, cplx<R>* __restrict__ y, int64_t ldy, int64_t n,
                           const int* __restrict__ slots) {
  const int a = blockIdx.y;
  const int s = slots[a];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cplx<R> bb = b[(int64_t)s * ldb + i];
  cplx<R> v = y[(int64_t)s * ldy + i];
  v.x = bb.x - v.x; v.y = bb.y - v.y;
  y[(int64_t)s * ldy + i] = v;
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
static constexpr int BLOCK_BUDGET = 32768;   // blocks per reduction launch (all systems together)

static int blocks_for(int nchunk, int nact) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(nchunk, std::max(8, BLOCK_BUDGET / std::max(1, nact))));
}

// nact * nb <= max(BLOCK_BUDGET, 8 * nact) partials per basis vector
template <typename R>
size_t VecOps<R>::part_size(int kmax, int nslot) {
  return (size_t)kmax * std::max<int64_t>(BLOCK_BUDGET, 8 * (int64_t)nslot) + 64;
}

template <typename R>
VecOps<R>::VecOps(Context* c, int64_t n_, int64_t ld_, int kmax_, int nslot_, double2* part_)
    : ctx(c), n(n_), ld(ld_), kmax(kmax_), part(part_) {
  VH_REQUIRE(kmax <= KMAX, "Krylov basis larger than 128 vectors");
  VH_REQUIRE(part != nullptr, "internal: no reduction partials buffer");
  (void)nslot_;
  nchunk = (int)ceil_div(n, CHUNK);
}

// h[a][j] = V_j^H w for j < k (device, double2 [a][kmax])
template <typename R>
void VecOps<R>::mdot(const cplx<R>* V, int64_t vstride, const cplx<R>* w, int k, int nact, const int* slots,
                     double2* h, const int* need, bool with_norm) {
  if (k <= 0 || nact <= 0) return;
  VH_REQUIRE(k + (with_norm ? 1 : 0) <= kmax, "internal: Krylov basis larger than the reduction layout");
  hipStream_t st = ctx->stream;
  const int nb = blocks_for(nchunk, nact);
  const int kk = k + (with_norm ? 1 : 0);
  hipLaunchKernelGGL((k_mdot<R>), dim3(nb, nact), dim3(KT), 0, st, V, vstride, w, ld, n, k, kmax, slots, part,
                     nchunk, need, with_norm ? 1 : 0);
  VH_LAUNCHED();
  // reduce each (a, j < kk) over nb partials into h[a][kmax]
  hipLaunchKernelGGL(k_reduce2, dim3(kk, nact), dim3(256), 0, st, part, nb, h, kmax, need, 0);
  VH_LAUNCHED();
}

template <typename R>
void VecOps<R>::reorth_flags(const double2* h, int k, const double2* nrm2, int* need, int nact) {
  if (nact <= 0) return;
  hipLaunchKernelGGL(k_reorth_flags, dim3(ceil_div(nact, 256)), dim3(256), 0, ctx->stream, h, kmax, k, nrm2, need,
                     nact, 0.5);
  VH_LAUNCHED();
}

// out = beta*out + sign*sum_j V_j coef_j ; if nrm2 != nullptr also ||out||^2 per system
template <typename R>
void VecOps<R>::combine(const cplx<R>* V, int64_t vstride, cplx<R>* out, int k, int nact, const int* slots,
                        const double2* coef, double beta, double sign, double2* nrm2, const int* need) {
  if (nact <= 0) return;
  VH_REQUIRE(k <= KMAX, "Krylov basis larger than 128 vectors");
  hipStream_t st = ctx->stream;
  const int nb = blocks_for(nchunk, nact);
  hipLaunchKernelGGL((k_combine<R>), dim3(nb, nact), dim3(KT), 0, st, V, vstride, out, ld, n, k, kmax, slots, coef,
                     beta, sign, nrm2 ? part : nullptr, nchunk, need);
  VH_LAUNCHED();
  if (nrm2) {
    hipLaunchKernelGGL(k_reduce2, dim3(nact), dim3(256), 0, st, part, nb, nrm2, 0, need, 1);
    VH_LAUNCHED();
  }
}

template <typename R>
void VecOps<R>::combine_dot(const cplx<R>* V, int64_t vstride, cplx<R>* w, int k, int nact, const int* slots,
                            const double2* h1, double2* h2) {
  if (nact <= 0) return;
  VH_REQUIRE(k >= 1 && k <= kFusedMax && k + 1 <= kmax, "internal: fused Gram-Schmidt step outside its basis range");
  hipStream_t st = ctx->stream;
  const int nb = blocks_for(nchunk, nact);
  const dim3 grid(nb, nact);
  if (k <= 4)
    hipLaunchKernelGGL((k_combine_dot<R, 4>), grid, dim3(KT), 0, st, V, vstride, w, ld, n, k, kmax, slots, h1, part,
                       nchunk);
  else if (k <= 8)
    hipLaunchKernelGGL((k_combine_dot<R, 8>), grid, dim3(KT), 0, st, V, vstride, w, ld, n, k, kmax, slots, h1, part,
                       nchunk);
  else
    hipLaunchKernelGGL((k_combine_dot<R, 16>), grid, dim3(KT), 0, st, V, vstride, w, ld, n, k, kmax, slots, h1, part,
                       nchunk);
  VH_LAUNCHED();
  hipLaunchKernelGGL(k_reduce2, dim3(k + 1, nact), dim3(256), 0, st, part, nb, h2, kmax, nullptr, 0);
  VH_LAUNCHED();
}

template <typename R>
void VecOps<R>::combine_scale(const cplx<R>* V, int64_t vstride, const cplx<R>* w, cplx<R>* out, int k, int nact,
                              const int* slots, const double2* h2, double2* nrm2) {
  if (nact <= 0) return;
  VH_REQUIRE(k >= 1 && k <= kFusedMax && k + 1 <= kmax, "internal: fused Gram-Schmidt step outside its basis range");
  hipStream_t st = ctx->stream;
  const int nb = blocks_for(nchunk, nact);
  const dim3 grid(nb, nact);
  if (k <= 4)
    hipLaunchKernelGGL((k_combine_scale<R, 4>), grid, dim3(KT), 0, st, V, vstride, w, out, ld, n, k, kmax, slots, h2,
                       nrm2, nchunk);
  else if (k <= 8)
    hipLaunchKernelGGL((k_combine_scale<R, 8>), grid, dim3(KT), 0, st, V, vstride, w, out, ld, n, k, kmax, slots, h2,
                       nrm2, nchunk);
  else
    hipLaunchKernelGGL((k_combine_scale<R, 16>), grid, dim3(KT), 0, st, V, vstride, w, out, ld, n, k, kmax, slots,
                       h2, nrm2, nchunk);
  VH_LAUNCHED();
}

template <typename R>
void VecOps<R>::norm2(const cplx<R>* x, int nact, const int* slots, double2* nrm2) {
  hipStream_t st = ctx->stream;
  const int nb = blocks_for(nchunk, nact);
  hipLaunchKernelGGL((k_mdot<R>), dim3(nb, nact), dim3(KT), 0, st, x, (int64_t)0, x, ld, n, 1, 1, slots, part,
                     nchunk, nullptr, 0);
  VH_LAUNCHED();
  hipLaunchKernelGGL(k_reduce2, dim3(nact), dim3(256), 0, st, part, nb, nrm2, 0, nullptr, 0);
  VH_LAUNCHED();
}

template <typename R>
void VecOps<R>::scale(const cplx<R>* src, cplx<R>* dst, int nact, const int* slots, const double2* nrm2) {
  if (nact <= 0) return;
  hipLaunchKernelGGL((k_scale<R>), dim3(ceil_div(n, 256), nact), dim3(256), 0, ctx->stream, src, dst, ld, n, slots,
                     nrm2);
  VH_LAUNCHED();
}

template <typename R>
template <typename RD>
void VecOps<R>::residual_to(const cplx<R>* b, const cplx<R>* y, cplx<RD>* v, int nact, const int* slots,
                            double2* nrm2) {
  if (nact <= 0) return;
  hipStream_t st = ctx->stream;
  const int nb = blocks_for(nchunk, nact);
  hipLaunchKernelGGL((k_residual_to<RD, R>), dim3(nb, nact), dim3(KT), 0, st, b, y, v, ld, n, slots, part, nchunk);
  VH_LAUNCHED();
  hipLaunchKernelGGL(k_reduce2, dim3(nact), dim3(256), 0, st, part, nb, nrm2, 0, nullptr, 0);
  VH_LAUNCHED();
}

template <typename R>
void VecOps<R>::zero(cplx<R>* x, int nact, const int* slots) {
  if (nact <= 0) return;
  hipLaunchKernelGGL((k_zero_slots<R>), dim3(ceil_div(n, 256), nact), dim3(256), 0, ctx->stream, x, ld, n, slots);
  VH_LAUNCHED();
}

template <typename R>
void VecOps<R>::sub_from(const cplx<R>* b, int64_t ldb, cplx<R>* y, int nact, const int* slots) {
  if (nact <= 0) return;
  hipLaunchKernelGGL((k_sub_from<R>), dim3(ceil_div(n, 256), nact), dim3(256), 0, ctx->stream, b, ldb, y, ld, n, slots);
  VH_LAUNCHED();
}

template <typename RD, typename RS>
void convert_vectors(Context* ctx, const cplx<RS>* src, int64_t lds, cplx<RD>* dst, int64_t ldd, int64_t n, int nact,
                     const int* slots_src, const int* slots_dst) {
  if (nact <= 0) return;
  hipLaunchKernelGGL((k_convert<RD, RS>), dim3(ceil_div(n, 256), nact), dim3(256), 0, ctx->stream, src, lds, dst, ldd,
                     n, slots_src, slots_dst);
  VH_LAUNCHED();
}

template struct VecOps<double>;
template struct VecOps<float>;
template void VecOps<double>::residual_to<float>(const double2*, const double2*, float2*, int, const int*, double2*);
template void convert_vectors<float, double>(Context*, const double2*, int64_t, float2*, int64_t, int64_t, int,
                                             const int*, const int*);
template void convert_vectors<double, float>(Context*, const float2*, int64_t, double2*, int64_t, int64_t, int,
                                             const int*, const int*);
template void convert_vectors<double, double>(Context*, const double2*, int64_t, double2*, int64_t, int64_t, int,
                                              const int*, const int*);
template void convert_vectors<float, float>(Context*, const float2*, int64_t, float2*, int64_t, int64_t, int,
                                            const int*, const int*);

template <typename RD, typename RS>
void accumulate_vectors(Context* ctx, const cplx<RS>* src, int64_t lds, cplx<RD>* dst, int64_t ldd, int64_t n,
                        int nact, const int* slots) {
  if (nact <= 0) return;
  hipLaunchKernelGGL((k_accumulate<RD, RS>), dim3(ceil_div(n, 256), nact), dim3(256), 0, ctx->stream, src, lds, dst,
                     ldd, n, slots);
  VH_LAUNCHED();
}
template void accumulate_vectors<double, float>(Context*, const float2*, int64_t, double2*, int64_t, int64_t, int,
                                                const int*);

// ---------------------------------------------------------------------------
// Least squares of the Arnoldi cycle (device, one thread per system)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 zcmul(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 zdiv(double2 a, double2 b) {
  // Smith's algorithm (no overflow for |b| near the range limits)
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, d = b.x + b.y * r;
    return make_double2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  }
  const double r = b.x / b.y, d = b.x * r + b.y;
  return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}

__global__ void k_ls_reset(LsDev L, const int* __restrict__ slots, int nact, const double2* __restrict__ nrm2) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nact) return;
  const int s = slots[a];
  const double beta = sqrt(nrm2[a].x);
  L.beta[s] = beta;
  double2* g = L.g + (size_t)s * (L.m + 1);
  g[0] = make_double2(beta, 0.0);
  for (int i = 1; i <= L.m; ++i) g[i] = make_double2(0.0, 0.0);
}

// Column j of H: h_ij (i <= j) from the Gram-Schmidt projections, h_{j+1,j} =
// ||w||.  The earlier rotations are applied to it in order (each acts on
// rows i, i+1), then the rotation that annihilates h_{j+1,j}:
//   c = |a| / t, s = (a / |a|) conj(b) / t, r = (a / |a|) t,  t = hypot(|a|, |b|),
// and g_{j+1} = -conj(s) g_j, g_j = c g_j (pgmres_iter.m:115-123 solves the
// same least-squares problem with Hc\rhs).
__global__ void k_ls_push(LsDev L, const int* __restrict__ slots, int nact, int j, const double2* __restrict__ h,
                          const double2* __restrict__ h2, int kmax, const double2* __restrict__ nrm2,
                          double* __restrict__ res) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nact) return;
  const int s = slots[a], m = L.m;
  double2* H = L.H + (size_t)s * (m + 1) * m;
  double2* R = L.R + (size_t)s * m * m;
  double2* g = L.g + (size_t)s * (m + 1);
  double2* sn = L.s + (size_t)s * m;
  double* cs = L.c + (size_t)s * m;
  const double2* ha = h + (size_t)a * kmax;
  const double2* h2a = h2 ? h2 + (size_t)a * kmax : nullptr;
  auto col = [&](int i) {
    double2 v = ha[i];
    if (h2a) { v.x += h2a[i].x; v.y += h2a[i].y; }
    return v;
  };
  const double hn = sqrt(nrm2[a].x);
  double2 x = col(0);
  H[j] = x;
  for (int i = 0; i < j; ++i) {
    const double2 y = col(i + 1);
    H[(size_t)(i + 1) * m + j] = y;
    const double c = cs[i];
    const double2 sv = sn[i];
    const double2 sy = zmul(sv, y), csx = zcmul(sv, x);
    R[(size_t)i * m + j] = make_double2(c * x.x + sy.x, c * x.y + sy.y);
    x = make_double2(c * y.x - csx.x, c * y.y - csx.y);
  }
  H[(size_t)(j + 1) * m + j] = make_double2(hn, 0.0);
  const double na = hypot(x.x, x.y), nb = hn;
  double cc;
  double2 ss, rr;
  if (nb == 0.0) {
    cc = 1.0; ss = make_double2(0.0, 0.0); rr = x;
  } else if (na == 0.0) {
    cc = 0.0; ss = make_double2(1.0, 0.0); rr = make_double2(nb, 0.0);
  } else {
    const double t = hypot(na, nb);
    const double2 u = make_double2(x.x / na, x.y / na);
    cc = na / t;
    ss = make_double2(u.x * (nb / t), u.y * (nb / t));
    rr = make_double2(u.x * t, u.y * t);
  }
  cs[j] = cc;
  sn[j] = ss;
  R[(size_t)j * m + j] = rr;
  const double2 gj = g[j];
  const double2 cg = zcmul(ss, gj);
  g[j + 1] = make_double2(-cg.x, -cg.y);
  g[j] = make_double2(cc * gj.x, cc * gj.y);
  res[a] = hypot(cg.x, cg.y);
}

// y = R^{-1} g (k unknowns) into coef[a]; rcoef[a][i] = beta e1 - H y (i <= k).
__global__ void k_ls_solve(LsDev L, const int* __restrict__ slots, int nl, int k, int kmax, double2* __restrict__ coef,
                           double2* __restrict__ rcoef) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nl) return;
  const int s = slots[a], m = L.m;
  const double2* H = L.H + (size_t)s * (m + 1) * m;
  const double2* R = L.R + (size_t)s * m * m;
  const double2* g = L.g + (size_t)s * (m + 1);
  double2* y = coef + (size_t)a * kmax;
  for (int i = k - 1; i >= 0; --i) {
    double2 t = g[i];
    for (int l = i + 1; l < k; ++l) {
      const double2 p = zmul(R[(size_t)i * m + l], y[l]);
      t.x -= p.x; t.y -= p.y;
    }
    y[i] = zdiv(t, R[(size_t)i * m + i]);
  }
  if (!rcoef) return;
  double2* r = rcoef + (size_t)a * kmax;
  for (int i = 0; i <= k; ++i) {
    double2 v = make_double2(i == 0 ? L.beta[s] : 0.0, 0.0);
    for (int l = i > 0 ? i - 1 : 0; l < k; ++l) {   // H is upper Hessenberg: H(i, l) = 0 for l < i - 1
      const double2 p = zmul(H[(size_t)i * m + l], y[l]);
      v.x -= p.x; v.y -= p.y;
    }
    r[i] = v;
  }
}

void ls_reset(Context* ctx, const LsDev& L, const int* slots, int nact, const double2* nrm2) {
  if (nact <= 0) return;
  hipLaunchKernelGGL(k_ls_reset, dim3(ceil_div(nact, 64)), dim3(64), 0, ctx->stream, L, slots, nact, nrm2);
  VH_LAUNCHED();
}

void ls_push(Context* ctx, const LsDev& L, const int* slots, int nact, int j, const double2* h, const double2* h2,
             int kmax, const double2* nrm2, double* res) {
  if (nact <= 0) return;
  VH_REQUIRE(j >= 0 && j < L.m && j < kmax, "internal: least-squares column outside the cycle");
  hipLaunchKernelGGL(k_ls_push, dim3(ceil_div(nact, 64)), dim3(64), 0, ctx->stream, L, slots, nact, j, h, h2, kmax,
                     nrm2, res);
  VH_LAUNCHED();
}

void ls_solve(Context* ctx, const LsDev& L, const int* slots, int nl, int k, int kmax, double2* coef,
              double2* rcoef) {
  if (nl <= 0) return;
  VH_REQUIRE(k >= 1 && k <= L.m && k + 1 <= kmax, "internal: least-squares solve outside the cycle");
  hipLaunchKernelGGL(k_ls_solve, dim3(ceil_div(nl, 64)), dim3(64), 0, ctx->stream, L, slots, nl, k, kmax, coef,
                     rcoef);
  VH_LAUNCHED();
}

}  // namespace vh
