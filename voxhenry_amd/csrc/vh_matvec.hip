// Circulant-tensor FFT matvec of the VoxHenry JVIE operator on MI355X.
//
// Replaces src_lse_formation/lse_matvect_mult.m:34-140.  Five fused passes over
// the padded grid (Px,Py,Pz) with every intermediate stored x-fastest:
//   X  : gather the compact currents through the voxel map, zero-pad along x,
//        forward FFT along x                        -> A[rc][z][y][kx]  (Px*M*N)
//   Y  : forward FFT along y (M -> Py)              -> B[rc][z][ky][kx] (Px*Py*N)
//   Z  : forward FFT along z (N -> Pz), 5x5 Green's product with the 7-component
//        tensor octant, inverse FFT along z, keep z < N  (in place in B)
//   Yi : inverse FFT along y, keep y < M            -> A
//   Xi : inverse FFT along x, keep x < L, add the resistive diagonal
//        w_c z / dx J and subtract Ar' phi, scatter to the compact output.
// plus Ar*J for the node rows.  rc = rhs*5 + component.  Lines of the voxel grid
// without any conductor are skipped in X and Xi (their A rows stay zero).
#include "vh_geom.h"

namespace vh {

// XCD-aware tile order: the hardware deals consecutive workgroups round-robin
// over the 8 XCDs; give each XCD a contiguous range of tiles so neighbouring
// tiles (which share 128-byte lines at tile edges) meet in the same L2.
__device__ __forceinline__ int xcd_tile(int bid, int nblk) {
  const int per = nblk >> 3, rem = nblk & 7;
  const int x = bid & 7, local = bid >> 3;
  return x * per + (x < rem ? x : rem) + local;
}

template <typename R> __device__ __forceinline__ cplx<R> czero() { cplx<R> c; c.x = 0; c.y = 0; return c; }

// ---------------------------------------------------------------------------
// pass X: gather + forward FFT along x.  One block = T lines of one (rhs, comp).
// ---------------------------------------------------------------------------
template <typename R, int EMAX>
__global__ __launch_bounds__(FFT_THREADS) void k_pass_x_fwd(
    const cplx<R>* __restrict__ x, int64_t ldx, const int* __restrict__ vmap,
    const int* __restrict__ lines, int nlines, int T, cplx<R>* __restrict__ A, GridDims g, FftPlan pl,
    const int* __restrict__ slots) {
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<R>* buf = (cplx<R>*)smem;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int rc = blockIdx.y, rhs = rc / 5, comp = rc - rhs * 5;
  const int slot = slots ? slots[rhs] : rhs;
  const int l0 = tile * T;
  const int Px = g.Px, L = g.L;
  const cplx<R>* xs = x + (int64_t)slot * ldx + (int64_t)comp * g.nv;
  for (int e = tid; e < T * Px; e += nthr) {
    const int p = e / Px, i = e - p * Px;
    cplx<R> v = czero<R>();
    if (i < L && l0 + p < nlines) {
      const int line = lines[l0 + p];
      const int vid = vmap[i + L * line];
      if (vid >= 0) v = xs[vid];
    }
    buf[e] = v;
  }
  __syncthreads();
  lds_fft<R, -1, EMAX>(buf, T, Px, 1, pl, tid, nthr, false);
  cplx<R>* As = A + (int64_t)rc * Px * g.M * g.N;
  for (int e = tid; e < T * Px; e += nthr) {
    const int p = e / Px, i = e - p * Px;
    if (l0 + p < nlines) As[(int64_t)lines[l0 + p] * Px + i] = buf[e];
  }
}

// ---------------------------------------------------------------------------
// pass Y (forward M -> Py or inverse Py -> M): one block = T consecutive kx of
// one z plane of one (rhs, comp).
// ---------------------------------------------------------------------------
template <typename R, int DIR, int EMAX>
__global__ __launch_bounds__(FFT_THREADS) void k_pass_y(
    const cplx<R>* in, int in_rows, int nin, cplx<R>* out, int out_rows, int nout, int T, GridDims g, FftPlan pl) {
  // in/out may alias (inverse pass works in place): each block reads its whole
  // tile before the FFT and writes it back afterwards.
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<R>* buf = (cplx<R>*)smem;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int tiles_x = (g.Px + T - 1) / T;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = tile % tiles_x, z = tile / tiles_x;
  const int kx0 = tx * T;
  const int rc = blockIdx.y;
  const int Px = g.Px, Py = pl.n;
  const cplx<R>* is = in + ((int64_t)rc * g.N + z) * in_rows * Px;
  for (int e = tid; e < T * Py; e += nthr) {
    const int y = e / T, t = e - y * T;
    const int kx = kx0 + t;
    cplx<R> v = czero<R>();
    if (y < nin && kx < Px) v = is[(int64_t)y * Px + kx];
    buf[e] = v;
  }
  __syncthreads();
  lds_fft<R, DIR, EMAX>(buf, T, 1, T, pl, tid, nthr, true);
  cplx<R>* os = out + ((int64_t)rc * g.N + z) * out_rows * Px;
  for (int e = tid; e < T * nout; e += nthr) {
    const int y = e / T, t = e - y * T;
    const int kx = kx0 + t;
    if (kx < Px) os[(int64_t)y * Px + kx] = buf[e];
  }
}

// ---------------------------------------------------------------------------
// pass Z: forward FFT along z, Green's product, inverse FFT, in place.
// One block = T consecutive kx of one ky column, looping over the rhs.
// LDS holds 5T pencils: element (pencil p = c*T + t, kz) at kz*5T + p.
// ---------------------------------------------------------------------------
struct TensorIdx {
  int idx;        // octant index
  float sx, sy, sz;
};

__device__ __forceinline__ TensorIdx tensor_index(const GridDims& g, int kx, int ky, int kz) {
  TensorIdx t;
  int ax = kx, ay = ky, az = kz;
  t.sx = 1.f; t.sy = 1.f; t.sz = 1.f;
  if (ax >= g.Hx) { ax = g.Px - ax; t.sx = -1.f; }
  if (ay >= g.Hy) { ay = g.Py - ay; t.sy = -1.f; }
  if (az >= g.Hz) { az = g.Pz - az; t.sz = -1.f; }
  t.idx = ax + g.Hx * (ay + g.Hy * az);
  return t;
}

// 5x5 block product of lse_matvect_mult.m:57-86 with complex tensor entries.
template <typename C>
__device__ __forceinline__ void green_product(C* f, const C* G, C alpha) {
  const C a34 = cadd(f[3], f[4]), d34 = csub(f[3], f[4]);
  C o0 = cadd(cmul(G[0], f[0]), cmul(G[1], a34));
  C o1 = cadd(cmul(G[0], f[1]), cmul(G[2], d34));
  C o2 = cadd(cmul(G[0], f[2]), cmul(G[4], f[4]));
  C o3 = csub(cadd(cmul(G[3], f[3]), cmul(G[5], f[4])), cadd(cmul(G[1], f[0]), cmul(G[2], f[1])));
  C o4 = cadd(csub(cmul(G[2], f[1]), cadd(cmul(G[1], f[0]), cmul(G[4], f[2]))),
              cadd(cmul(G[5], f[3]), cmul(G[6], f[4])));
  f[0] = cmul(alpha, o0); f[1] = cmul(alpha, o1); f[2] = cmul(alpha, o2);
  f[3] = cmul(alpha, o3); f[4] = cmul(alpha, o4);
}

// Same product with the compressed real tensor: even components g0,g3,g5,g6 are
// real, odd components G1 = i g1, G2 = i g2, G4 = i g4.
template <typename C, typename R>
__device__ __forceinline__ void green_product_real(C* f, const R* g, C alpha) {
  auto im = [](R s, C a) { C r; r.x = -s * a.y; r.y = s * a.x; return r; };   // i*s*a
  const C a34 = cadd(f[3], f[4]), d34 = csub(f[3], f[4]);
  C o0 = cadd(cscale(f[0], g[0]), im(g[1], a34));
  C o1 = cadd(cscale(f[1], g[0]), im(g[2], d34));
  C o2 = cadd(cscale(f[2], g[0]), im(g[4], f[4]));
  C o3 = csub(cadd(cscale(f[3], g[3]), cscale(f[4], g[5])), cadd(im(g[1], f[0]), im(g[2], f[1])));
  C o4 = cadd(csub(im(g[2], f[1]), cadd(im(g[1], f[0]), im(g[4], f[2]))),
              cadd(cscale(f[3], g[5]), cscale(f[4], g[6])));
  f[0] = cmul(alpha, o0); f[1] = cmul(alpha, o1); f[2] = cmul(alpha, o2);
  f[3] = cmul(alpha, o3); f[4] = cmul(alpha, o4);
}

template <typename R> struct TensorRef;
template <> struct TensorRef<double> { const double2* t; };
template <> struct TensorRef<float> { const float* t; };

template <typename R>
__device__ __forceinline__ void apply_point(cplx<R>* f, const TensorRef<R>& T, size_t O, const TensorIdx& ti,
                                            cplx<R> alpha);
template <>
__device__ __forceinline__ void apply_point<double>(double2* f, const TensorRef<double>& T, size_t O,
                                                    const TensorIdx& ti, double2 alpha) {
  double2 G[7];
#pragma unroll
  for (int c = 0; c < 7; ++c) G[c] = T.t[c * O + ti.idx];
  G[1] = cscale(G[1], (double)ti.sx);
  G[2] = cscale(G[2], (double)ti.sy);
  G[4] = cscale(G[4], (double)ti.sz);
  green_product(f, G, alpha);
}
template <>
__device__ __forceinline__ void apply_point<float>(float2* f, const TensorRef<float>& T, size_t O,
                                                   const TensorIdx& ti, float2 alpha) {
  float g[7];
#pragma unroll
  for (int c = 0; c < 7; ++c) g[c] = T.t[c * O + ti.idx];
  g[1] *= ti.sx; g[2] *= ti.sy; g[4] *= ti.sz;
  green_product_real(f, g, alpha);
}

template <typename R, int EMAX>
__global__ __launch_bounds__(FFT_THREADS) void k_pass_z(
    cplx<R>* __restrict__ B, int nrhs, int T, GridDims g, FftPlan pl, TensorRef<R> tens, cplx<R> alpha0,
    const double2* __restrict__ alphas, const int* __restrict__ slots) {
  // alphas (optional): j*omega*mu per vector slot; otherwise alpha0 for all
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<R>* buf = (cplx<R>*)smem;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int tiles_x = (g.Px + T - 1) / T;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = tile % tiles_x, ky = tile / tiles_x;
  const int kx0 = tx * T;
  const int Px = g.Px, Pz = g.Pz, N = g.N;
  const int T5 = 5 * T;
  const int64_t plane = (int64_t)Px * g.Py;     // elements per z-plane of one rc
  const size_t O = (size_t)g.Hx * g.Hy * g.Hz;
  for (int rhs = 0; rhs < nrhs; ++rhs) {
    cplx<R> alpha = alpha0;
    if (alphas) {
      const double2 a = alphas[slots ? slots[rhs] : rhs];
      alpha.x = (R)a.x; alpha.y = (R)a.y;
    }
    cplx<R>* base = B + (int64_t)rhs * 5 * plane * N + (int64_t)ky * Px + kx0;
    for (int e = tid; e < T5 * Pz; e += nthr) {
      const int kz = e / T5, p = e - kz * T5;
      const int c = p / T, t = p - c * T;
      cplx<R> v = czero<R>();
      if (kz < N && kx0 + t < Px) v = base[((int64_t)c * N + kz) * plane + t];
      buf[e] = v;
    }
    __syncthreads();
    lds_fft<R, -1, EMAX>(buf, T5, 1, T5, pl, tid, nthr, true);
    for (int e = tid; e < T * Pz; e += nthr) {
      const int kz = e / T, t = e - kz * T;
      const int kx = kx0 + t;
      if (kx < Px) {
        cplx<R>* f0 = buf + kz * T5 + t;
        cplx<R> f[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) f[c] = f0[c * T];
        const TensorIdx ti = tensor_index(g, kx, ky, kz);
        apply_point<R>(f, tens, O, ti, alpha);
#pragma unroll
        for (int c = 0; c < 5; ++c) f0[c * T] = f[c];
      }
    }
    __syncthreads();
    lds_fft<R, +1, EMAX>(buf, T5, 1, T5, pl, tid, nthr, true);
    for (int e = tid; e < T5 * N; e += nthr) {
      const int kz = e / T5, p = e - kz * T5;
      const int c = p / T, t = p - c * T;
      if (kx0 + t < Px) base[((int64_t)c * N + kz) * plane + t] = buf[e];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// pass Xi: inverse FFT along x, crop, diagonal, -Ar' phi, scatter.
// ---------------------------------------------------------------------------
template <typename R, int EMAX>
__global__ __launch_bounds__(FFT_THREADS) void k_pass_x_inv(
    const cplx<R>* __restrict__ A, const cplx<R>* __restrict__ x, int64_t ldx, cplx<R>* __restrict__ y, int64_t ldy,
    const int* __restrict__ vmap, const int* __restrict__ lines, int nlines, int T, GridDims g, FftPlan pl,
    const cplx<R>* __restrict__ z, R inv_dx, const int* __restrict__ atp, const int* __restrict__ atc,
    const R* __restrict__ atv, int64_t num_curr, int mode, const int* __restrict__ slots) {
  // mode 0: conv + diag - Ar'phi ; mode 1: conv + diag ; mode 2: conv only
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<R>* buf = (cplx<R>*)smem;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int rc = blockIdx.y, rhs = rc / 5, comp = rc - rhs * 5;
  const int slot = slots ? slots[rhs] : rhs;
  const int l0 = tile * T;
  const int Px = g.Px, L = g.L;
  const cplx<R>* As = A + (int64_t)rc * Px * g.Py * g.N;
  for (int e = tid; e < T * Px; e += nthr) {
    const int p = e / Px, i = e - p * Px;
    cplx<R> v = czero<R>();
    if (l0 + p < nlines) {
      const int line = lines[l0 + p];
      const int yy = line % g.M, zz = line / g.M;
      v = As[((int64_t)zz * g.Py + yy) * Px + i];
    }
    buf[e] = v;
  }
  __syncthreads();
  lds_fft<R, +1, EMAX>(buf, T, Px, 1, pl, tid, nthr, false);
  const R w = inv_dx * (comp < 3 ? R(1) : comp == 3 ? R(1) / R(6) : R(1) / R(2));
  const cplx<R>* xs = x + (int64_t)slot * ldx;
  cplx<R>* ys = y + (int64_t)slot * ldy;
  for (int e = tid; e < T * L; e += nthr) {
    const int p = e / L, i = e - p * L;
    if (l0 + p >= nlines) continue;
    const int vid = vmap[i + L * lines[l0 + p]];
    if (vid < 0) continue;
    const int64_t row = (int64_t)comp * g.nv + vid;
    cplx<R> acc = buf[p * Px + i];
    if (mode <= 1) {
      const cplx<R> J = xs[row];
      const cplx<R> zz = z[vid];
      acc = cadd(acc, cscale(cmul(zz, J), w));
    }
    if (mode == 0) {
      cplx<R> s = czero<R>();
      for (int k = atp[row]; k < atp[row + 1]; ++k) {
        const cplx<R> ph = xs[num_curr + atc[k]];
        s.x += atv[k] * ph.x; s.y += atv[k] * ph.y;
      }
      acc = csub(acc, s);
    }
    ys[row] = acc;
  }
}

// Node rows: y_node = Ar * J.
template <typename R>
__global__ __launch_bounds__(256) void k_node_rows(
    const cplx<R>* __restrict__ x, int64_t ldx, cplx<R>* __restrict__ y, int64_t ldy, const int* __restrict__ arp,
    const int* __restrict__ arc, const R* __restrict__ arv, int num_node, int64_t num_curr,
    const int* __restrict__ slots) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int rhs = slots ? slots[blockIdx.y] : blockIdx.y;
  if (i >= num_node) return;
  const cplx<R>* xs = x + (int64_t)rhs * ldx;
  cplx<R> s = czero<R>();
  for (int k = arp[i]; k < arp[i + 1]; ++k) {
    const cplx<R> v = xs[arc[k]];
    s.x += arv[k] * v.x; s.y += arv[k] * v.y;
  }
  y[(int64_t)rhs * ldy + num_curr + i] = s;
}

// ---------------------------------------------------------------------------
// host launch logic
// ---------------------------------------------------------------------------
template <typename R> constexpr int emax_of() { return sizeof(R) == 8 ? 16 : 32; }

// Pick the number of pencils per block: T*n*npen_factor fits EMAX*threads and
// the LDS budget; prefer >= 128 contiguous bytes per row.
static int pick_T(int n, int factor, size_t elem, int emax, size_t lds_budget, int want) {
  int T = want;
  while (T > 1 && ((int64_t)T * n * factor > (int64_t)emax * FFT_THREADS ||
                   (size_t)T * n * factor * elem > lds_budget))
    T >>= 1;
  if ((int64_t)T * n * factor > (int64_t)emax * FFT_THREADS || (size_t)T * n * factor * elem > lds_budget)
    throw Error(ERR_ARG, "padded FFT length " + std::to_string(n) + " exceeds the LDS pencil kernels");
  return T;
}

template <typename R>
static void ensure_ws(Geometry& G, int nrhs) {
  const GridDims& g = G.g;
  // A is written only by pass X and only on non-empty lines: its empty lines
  // must read as zero, so it is cleared once when (re)allocated.
  const size_t a = (size_t)nrhs * 5 * g.Px * g.M * g.N * sizeof(cplx<R>);
  const size_t b = (size_t)nrhs * 5 * g.Px * g.Py * g.N * sizeof(cplx<R>);
  DevBuf<char>& wsA = sizeof(R) == 8 ? G.wsA64 : G.wsA32;
  if (wsA.n < a) { wsA.alloc(a); VH_HIP(hipMemsetAsync(wsA.p, 0, a, G.ctx->stream)); }
  if (G.wsB.n < b) G.wsB.alloc(b);
}

template <typename R>
static void run_passes(Geometry& G, int nrhs, const cplx<R>* x, int64_t ldx, cplx<R>* y, int64_t ldy,
                       double are, double aim, int mode, const int* slots, const double2* alphas) {
  VH_REQUIRE(G.has_tensor, "matvec: Green's tensor not set");
  VH_REQUIRE(nrhs >= 1, "matvec: nrhs must be >= 1");
  VH_REQUIRE(nrhs * 5 <= 65535, "matvec: too many right-hand sides in one call");
  const GridDims& g = G.g;
  ensure_ws<R>(G, nrhs);
  cplx<R>* A = (cplx<R>*)(sizeof(R) == 8 ? G.wsA64.p : G.wsA32.p);
  cplx<R>* B = (cplx<R>*)G.wsB.p;
  hipStream_t st = G.ctx->stream;
  constexpr int EM = emax_of<R>();
  const size_t elem = sizeof(cplx<R>);
  const size_t budget = 64 * 1024;
  const int want = sizeof(R) == 8 ? 8 : 16;
  FftPlan px = G.ax.plan<R>(), py = G.ay.plan<R>(), pz = G.az.plan<R>();
  PassTimer* tm = G.timer;
  if (tm) { tm->collect(); VH_HIP(hipEventRecord(tm->ev[0], st)); }
  auto mark = [&](int k) { if (tm) VH_HIP(hipEventRecord(tm->ev[k], st)); };

  // X
  if (G.nlines > 0) {
    const int T = pick_T(g.Px, 1, elem, EM, budget, 4);
    dim3 grid(ceil_div(G.nlines, T), nrhs * 5);
    hipLaunchKernelGGL((k_pass_x_fwd<R, EM>), grid, dim3(FFT_THREADS), T * g.Px * elem, st, x, ldx, G.vmap.p,
                       G.lines.p, G.nlines, T, A, g, px, slots);
    VH_LAUNCHED();
  }
  mark(1);
  // Y
  {
    const int T = pick_T(g.Py, 1, elem, EM, budget, want);
    dim3 grid(ceil_div(g.Px, T) * g.N, nrhs * 5);
    hipLaunchKernelGGL((k_pass_y<R, -1, EM>), grid, dim3(FFT_THREADS), T * g.Py * elem, st, A, g.M, g.M, B, g.Py,
                       g.Py, T, g, py);
    VH_LAUNCHED();
  }
  mark(2);
  // Z
  {
    const int T = pick_T(g.Pz, 5, elem, EM, budget, want);
    dim3 grid(ceil_div(g.Px, T) * g.Py);
    cplx<R> alpha;
    const double scale = 1.0;
    alpha.x = (R)(are * scale);
    alpha.y = (R)(aim * scale);
    TensorRef<R> tr;
    if constexpr (sizeof(R) == 8) tr.t = G.t64.p; else tr.t = G.t32.p;
    hipLaunchKernelGGL((k_pass_z<R, EM>), grid, dim3(FFT_THREADS), 5 * T * g.Pz * elem, st, B, nrhs, T, g, pz, tr,
                       alpha, alphas, slots);
    VH_LAUNCHED();
  }
  mark(3);
  // Yi
  {
    const int T = pick_T(g.Py, 1, elem, EM, budget, want);
    dim3 grid(ceil_div(g.Px, T) * g.N, nrhs * 5);
    hipLaunchKernelGGL((k_pass_y<R, +1, EM>), grid, dim3(FFT_THREADS), T * g.Py * elem, st, B, g.Py, g.Py, B, g.Py,
                       g.M, T, g, py);
    VH_LAUNCHED();
  }
  mark(4);
  // Xi
  if (G.nlines > 0) {
    const int T = pick_T(g.Px, 1, elem, EM, budget, 4);
    dim3 grid(ceil_div(G.nlines, T), nrhs * 5);
    const R* atv;
    if constexpr (sizeof(R) == 8) atv = G.ArT.v64.p; else atv = G.ArT.v32.p;
    const cplx<R>* zz;
    if constexpr (sizeof(R) == 8) zz = G.z64.p; else zz = G.z32.p;
    if (mode <= 1) VH_REQUIRE(G.has_z, "matvec: impedance not set");
    if (mode == 0) VH_REQUIRE(G.has_inc, "matvec: incidence not set");
    hipLaunchKernelGGL((k_pass_x_inv<R, EM>), grid, dim3(FFT_THREADS), T * g.Px * elem, st, B, x, ldx, y, ldy,
                       G.vmap.p, G.lines.p, G.nlines, T, g, px, zz, (R)(1.0 / G.dx), G.ArT.ptr.p, G.ArT.col.p, atv,
                       G.num_curr, mode, slots);
    VH_LAUNCHED();
  }
  mark(5);
  if (mode == 0 && G.num_node > 0) {
    const R* arv;
    if constexpr (sizeof(R) == 8) arv = G.Ar.v64.p; else arv = G.Ar.v32.p;
    dim3 grid(ceil_div(G.num_node, 256), nrhs);
    hipLaunchKernelGGL((k_node_rows<R>), grid, dim3(256), 0, st, x, ldx, y, ldy, G.Ar.ptr.p, G.Ar.col.p, arv,
                       (int)G.num_node, G.num_curr, slots);
    VH_LAUNCHED();
  }
  mark(6);
  if (tm) tm->pending = true;
}

void PassTimer::create() {
  for (auto& e : ev) VH_HIP(hipEventCreate(&e));
}
void PassTimer::destroy() {
  for (auto& e : ev) if (e) (void)hipEventDestroy(e);
}
void PassTimer::collect() {
  if (!pending) return;
  VH_HIP(hipEventSynchronize(ev[NP]));
  for (int k = 0; k < NP; ++k) {
    float t = 0;
    VH_HIP(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
    ms[k] += t;
  }
  calls++;
  pending = false;
}

template <typename R>
void matvec(Geometry& G, int nrhs, const cplx<R>* x, int64_t ldx, cplx<R>* y, int64_t ldy, double are, double aim,
            int mode, const int* slots, const double2* alphas) {
  run_passes<R>(G, nrhs, x, ldx, y, ldy, are, aim, mode, slots, alphas);
}

template void matvec<double>(Geometry&, int, const double2*, int64_t, double2*, int64_t, double, double, int,
                             const int*, const double2*);
template void matvec<float>(Geometry&, int, const float2*, int64_t, float2*, int64_t, double, double, int,
                            const int*, const double2*);

}  // namespace vh
