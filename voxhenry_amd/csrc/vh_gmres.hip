// Krylov solvers on the device.
//
// * pgmres  -- src_iterative_solver/pgmres.m:70-197 + pgmres_iter.m:22-153 with
//   the same semantics: x0 = 0, left preconditioner, bnorm = ||M b||, one-pass
//   classical Gram-Schmidt, stop when the least-squares residual drops below
//   tol*bnorm, restart from r = V(:,1:k+1)*res, final true relres.  The small
//   least-squares problem is solved by Givens rotations (same residuals as the
//   reference's Hc\rhs up to rounding).  Runs a batch of independent systems in
//   lock step; a system that converges is retired from the batch (slot map).
// * fpgmres -- fpgmres.m:58-152 + fpgmres_iter.m:26-121 (flexible, right
//   preconditioned, bnorm = ||b||).
// * fgmres_mixed -- FP64 flexible GMRES whose preconditioner is an FP32
//   left-preconditioned GMRES with the FP32 matvec and FP32 preconditioner:
//   each outer step is one FP64 matvec, so the outer residual is the true
//   FP64 residual and accuracy is recovered by the outer iteration.
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
template void convert_vectors<float, double>(Context*, const double2*, int64_t, float2*, int64_t, int64_t, int,
                                             const int*, const int*);
template void convert_vectors<double, float>(Context*, const float2*, int64_t, double2*, int64_t, int64_t, int,
                                             const int*, const int*);
template void convert_vectors<double, double>(Context*, const double2*, int64_t, double2*, int64_t, int64_t, int,
                                              const int*, const int*);
template void convert_vectors<float, float>(Context*, const float2*, int64_t, float2*, int64_t, int64_t, int,
                                            const int*, const int*);

// ---------------------------------------------------------------------------
// Givens least squares for one system (host)
// ---------------------------------------------------------------------------
using cd = std::complex<double>;

void Givens::reset(int m, double beta) {
  R.assign((size_t)m * m, 0.0);
  H.assign((size_t)(m + 1) * m, 0.0);
  g.assign(m + 1, 0.0);
  c.assign(m, 0.0);
  s.assign(m, 0.0);
  g[0] = beta;
  this->m = m;
  k = 0;
  beta0 = beta;
}

// append column k (h[0..k], h[k] = ||w||) -> returns |g_{k+1}|
double Givens::push(const cd* h, double hnext) {
  const int j = k;
  for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = h[i];
  H[(size_t)(j + 1) * m + j] = hnext;
  std::vector<cd> col(j + 2);
  for (int i = 0; i <= j; ++i) col[i] = h[i];
  col[j + 1] = hnext;
  for (int i = 0; i < j; ++i) {
    const cd t1 = col[i], t2 = col[i + 1];
    col[i] = c[i] * t1 + s[i] * t2;
    col[i + 1] = -std::conj(s[i]) * t1 + c[i] * t2;
  }
  const cd a = col[j], b = col[j + 1];
  const double na = std::abs(a), nb = std::abs(b);
  double cc;
  cd ss, rr;
  if (nb == 0.0) { cc = 1.0; ss = 0.0; rr = a; }
  else if (na == 0.0) { cc = 0.0; ss = std::conj(b) / nb; rr = nb; }
  else {
    const double t = std::hypot(na, nb);
    cc = na / t;
    ss = (a / na) * std::conj(b) / t;
    rr = (a / na) * t;
  }
  c[j] = cc; s[j] = ss;
  for (int i = 0; i < j; ++i) R[(size_t)i * m + j] = col[i];
  R[(size_t)j * m + j] = rr;
  g[j + 1] = -std::conj(ss) * g[j];
  g[j] = cc * g[j];
  k = j + 1;
  return std::abs(g[j + 1]);
}

void Givens::solve(std::vector<cd>& y) const {
  y.assign(k, 0.0);
  for (int i = k - 1; i >= 0; --i) {
    cd t = g[i];
    for (int l = i + 1; l < k; ++l) t -= R[(size_t)i * m + l] * y[l];
    y[i] = t / R[(size_t)i * m + i];
  }
}

// res = beta e1 - H y  (k+1 entries), the reference's rhs - Hc*y
void Givens::residual(const std::vector<cd>& y, std::vector<cd>& res) const {
  res.assign(k + 1, 0.0);
  res[0] = beta0;
  for (int i = 0; i <= k; ++i)
    for (int l = 0; l < k; ++l) res[i] -= H[(size_t)i * m + l] * y[l];
}

}  // namespace vh
