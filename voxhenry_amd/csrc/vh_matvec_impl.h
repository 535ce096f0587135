// Kernels and launch logic of the FFT matvec passes, shared by the
// translation units that instantiate them per precision (the pass-Z
// kernels in vh_matvec_z32.hip / vh_matvec_z64.hip, the X/Y passes and the
// matvec drivers in vh_matvec_f32.hip / vh_matvec_f64.hip); the data flow is
// described in vh_matvec.hip.
#pragma once
#include "vh_geom.h"
#include <type_traits>

namespace vh {

// Pass-Z tile order.  One row of the tensor octant, (|kx| tile, |ky|), serves
// the four tiles (±kx, ±ky).  Rows ky and Py-ky are interleaved tile by tile
// (row 0, and row Py/2 for even Py, are their own mirrors), and within a row
// the kx tiles go 0, last, 1, last-1, ... so the tiles that read one octant
// slice run back to back (in one XCD's L2, with xcd_tile).  A bijection of
// [0, tiles_x*Py) onto the (tx, ky) pairs.  Used for the complex128 kernels,
// whose 16-byte tensor elements are re-read across the mirrored tiles
// otherwise (128^3 cube: 0.928 -> 0.910 ms); in complex64 the plain row order
// measured faster (0.285 vs 0.290 ms) and is kept (MIRROR = false: row u = ky
// order 0, 1, Py-1, 2, Py-2, ..., kx tiles in order).
template <bool MIRROR>
__device__ __forceinline__ void z_tile(int tile, int tiles_x, int Py, int& tx, int& ky) {
  int u = tile / tiles_x, p = tile - u * tiles_x;
  if (!MIRROR) {
    tx = p;
    const int v = u - 1;
    ky = u == 0 ? 0 : ((v & 1) ? Py - 1 - (v >> 1) : 1 + (v >> 1));
    return;
  }
  if (u != 0 && !(u == Py - 1 && (Py & 1) == 0)) {   // rows 2q+1, 2q+2: ky = 1+q and Py-1-q
    const int q = (u - 1) >> 1;
    p = tile - (1 + 2 * q) * tiles_x;
    u = 1 + 2 * q + (p & 1);
    p >>= 1;
  }
  tx = (p & 1) ? tiles_x - 1 - (p >> 1) : (p >> 1);
  const int v = u - 1;
  ky = u == 0 ? 0 : ((v & 1) ? Py - 1 - (v >> 1) : 1 + (v >> 1));
}

template <typename R> __device__ __forceinline__ cplx<R> czero() { cplx<R> c; c.x = 0; c.y = 0; return c; }

// Pass shapes: threads per workgroup and the bound E on elements per thread
// and pencil.  X/Xi and Y/Yi hold one pencil element set per thread; Z holds
// the five components' spectra of its pencils in registers between the
// forward and the inverse transform, hence smaller bounds (ZShape below).
// X/Y-pass shapes the autotuner chooses from (threads per workgroup, element
// bound); entry 0 is the default.
struct XYShape { int nt, e; };
constexpr XYShape kXYShapes32[] = {{256, 16}, {256, 8}, {128, 16}};
constexpr XYShape kXYShapes64[] = {{256, 8}, {256, 4}, {128, 8}};
template <typename R> constexpr int num_xyshapes() {
  return sizeof(R) == 8 ? (int)(sizeof(kXYShapes64) / sizeof(XYShape)) : (int)(sizeof(kXYShapes32) / sizeof(XYShape));
}
template <typename R> static XYShape xyshape(int i) { return sizeof(R) == 8 ? kXYShapes64[i] : kXYShapes32[i]; }
template <int N> constexpr int ilog2() { return N <= 1 ? 0 : 1 + ilog2<N / 2>(); }

// ---------------------------------------------------------------------------
// pass X: gather + forward FFT along x.  A workgroup transforms T lines of one
// (rhs, comp); the tpp threads of a line are consecutive, so the first stage's
// loads of a line are contiguous.  LDS: line t at t*LP, element i at
// i + i/16 (one pad slot every 16 elements keeps the stride-R writes of the
// early stages on distinct banks).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }
inline int lpad_host(int i) { return i + (i >> 4); }

template <typename R, int E, int NT>
__global__ __launch_bounds__(NT) void k_pass_x_fwd(
    const cplx<R>* __restrict__ x, int64_t ldx, const int* __restrict__ vmap, const int* __restrict__ lines,
    const int4* __restrict__ runs, int nlines, int ltpp, cplx<R>* __restrict__ A, GridDims g, Plan2 pl,
    const int* __restrict__ slots, int dense) {
  using C = cplx<R>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x;
  const int tpp = 1 << ltpp, T = NT >> ltpp;
  const int jl = tid & (tpp - 1), t = tid >> ltpp;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int rc = blockIdx.y, rhs = rc / 5, comp = rc - rhs * 5;
  const int slot = slots ? slots[rhs] : rhs;
  const int l = tile * T + t;
  const bool valid = l < nlines;
  const int line = valid ? lines[l] : 0;
  const int L = g.L, Px = g.Px;
  const C* xs = x + (int64_t)slot * ldx + (int64_t)comp * g.nv;
  const int* vm = vmap + (int64_t)L * line;
  const int64_t lbase = (int64_t)L * line;
  // a line whose voxels are one run needs no voxel-map lookup (no dependent load)
  const int4 run = valid && !dense ? runs[l] : make_int4(-1, 0, 0, 0);
  C* lds = (C*)smem + t * (lpad(Px - 1) + 1);
  auto lx = [](int i) { return lpad(i); };
  auto gld = [&](int i) -> C {
    C v = czero<R>();
    if (valid && i < L) {
      if (dense) v = xs[lbase + i];
      else if (run.x >= 0) { if (i >= run.x && i < run.y) v = xs[run.z + i]; }
      else { const int vid = vm[i]; if (vid >= 0) v = xs[vid]; }
    }
    return v;
  };
  C* As = A + ((int64_t)rc * g.M * g.N + line) * Px;
  auto gst = [&](int o, C v) { if (valid) As[o] = v; };
  fft_pencil<R, -1, E>(pl, lds, lx, jl, tpp, gld, gst);
}

// ---------------------------------------------------------------------------
// pass Y (forward M -> Py or inverse Py -> M): a workgroup transforms T
// consecutive kx of one z plane of one (rhs, comp); consecutive threads take
// consecutive kx, so every load and store moves T contiguous elements.  in and
// out may alias (the inverse works in place): all loads of a workgroup precede
// its first store.  rowbits (optional, M <= 256: 4 words per z plane) marks
// the y rows whose x line holds voxels: the inverse writes only those (pass Xi
// reads no other row).  The plane is uniform across the workgroup, so its
// words are loaded once and the per-row test is arithmetic only.  (Skipping
// the forward pass's reads of the empty, zero rows the same way measured
// slower: that pass is bound by instruction issue, not by its loads.)
// ---------------------------------------------------------------------------
template <typename R, int DIR, int E, int NT>
__global__ __launch_bounds__(NT) void k_pass_y(const cplx<R>* in, int in_rows, int nin, cplx<R>* out, int out_rows,
                                               int nout, int ltpp, GridDims g, Plan2 pl,
                                               const unsigned long long* __restrict__ rowbits) {
  using C = cplx<R>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x;
  const int lT = ilog2<NT>() - ltpp;
  const int T = 1 << lT, tpp = 1 << ltpp;
  const int t = tid & (T - 1), jl = tid >> lT;
  const int tiles_x = (g.Px + T - 1) >> lT;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = tile % tiles_x, z = tile / tiles_x;
  const int kx = (tx << lT) + t;
  const bool valid = kx < g.Px;
  const int rc = blockIdx.y;
  const int Px = g.Px;
  const C* is = in + ((int64_t)rc * g.N + z) * in_rows * Px + kx;
  C* os = out + ((int64_t)rc * g.N + z) * out_rows * Px + kx;
  C* lds = (C*)smem + t;
  unsigned long long w0 = ~0ull, w1 = ~0ull, w2 = ~0ull, w3 = ~0ull;
  if (rowbits) {
    const unsigned long long* rb = rowbits + (int64_t)z * 4;
    w0 = rb[0]; w1 = rb[1]; w2 = rb[2]; w3 = rb[3];
  }
  auto live = [&](int i) -> bool {
    const unsigned long long w = i < 64 ? w0 : i < 128 ? w1 : i < 192 ? w2 : w3;
    return (w >> (i & 63)) & 1ull;
  };
  auto lx = [T](int i) { return i * T; };
  auto gld = [&](int i) -> C { return (valid && i < nin) ? is[i * Px] : czero<R>(); };
  auto gst = [&](int o, C v) { if (valid && o < nout && (DIR < 0 || live(o))) os[o * Px] = v; };
  fft_pencil<R, DIR, E>(pl, lds, lx, jl, tpp, gld, gst);
}

// ---------------------------------------------------------------------------
// pass Z: forward FFT along z, Green's product, inverse FFT, in place.
// A workgroup owns T consecutive kx of one ky column of one rhs.
// The five components are transformed one after the other through the same
// LDS array; each leaves its spectrum in registers (last forward stage), the
// 5x5 product is applied in registers, and the inverse transforms start from
// registers.  Columns ky and Py-ky read the same tensor octant rows, so the
// tile order visits them back to back.
// ---------------------------------------------------------------------------
struct TensorIdx {
  int idx;        // octant point index
  float sx, sy, sz;
};

// 5x5 block product of lse_matvect_mult.m:57-86 with complex tensor entries.
template <typename C>
__device__ __forceinline__ void green_product(C* f, const C* G) {
  const C a34 = cadd(f[3], f[4]), d34 = csub(f[3], f[4]);
  C o0 = cadd(cmul(G[0], f[0]), cmul(G[1], a34));
  C o1 = cadd(cmul(G[0], f[1]), cmul(G[2], d34));
  C o2 = cadd(cmul(G[0], f[2]), cmul(G[4], f[4]));
  C o3 = csub(cadd(cmul(G[3], f[3]), cmul(G[5], f[4])), cadd(cmul(G[1], f[0]), cmul(G[2], f[1])));
  C o4 = cadd(csub(cmul(G[2], f[1]), cadd(cmul(G[1], f[0]), cmul(G[4], f[2]))),
              cadd(cmul(G[5], f[3]), cmul(G[6], f[4])));
  f[0] = o0; f[1] = o1; f[2] = o2; f[3] = o3; f[4] = o4;
}

// Same product with the compressed real tensor: even components g0,g3,g5,g6 are
// real, odd components G1 = i g1, G2 = i g2, G4 = i g4.
template <typename C, typename R>
__device__ __forceinline__ void green_product_real(C* f, const R* g) {
  auto im = [](R s, C a) { C r; r.x = -s * a.y; r.y = s * a.x; return r; };   // i*s*a
  const C a34 = cadd(f[3], f[4]), d34 = csub(f[3], f[4]);
  C o0 = cadd(cscale(f[0], g[0]), im(g[1], a34));
  C o1 = cadd(cscale(f[1], g[0]), im(g[2], d34));
  C o2 = cadd(cscale(f[2], g[0]), im(g[4], f[4]));
  C o3 = csub(cadd(cscale(f[3], g[3]), cscale(f[4], g[5])), cadd(im(g[1], f[0]), im(g[2], f[1])));
  C o4 = cadd(csub(im(g[2], f[1]), cadd(im(g[1], f[0]), im(g[4], f[2]))),
              cadd(cscale(f[3], g[5]), cscale(f[4], g[6])));
  f[0] = o0; f[1] = o1; f[2] = o2; f[3] = o3; f[4] = o4;
}

template <typename R> struct TensorRef;
template <> struct TensorRef<double> { const double2* t; };
template <> struct TensorRef<float> { const float* t; };

template <typename R>
__device__ __forceinline__ void apply_point(cplx<R>* f, const TensorRef<R>& T, const TensorIdx& ti);
template <>
__device__ __forceinline__ void apply_point<double>(double2* f, const TensorRef<double>& T, const TensorIdx& ti) {
  const double2* p = T.t + (size_t)ti.idx * TENSOR_STRIDE;
  double2 G[7];
#pragma unroll
  for (int c = 0; c < 7; ++c) G[c] = p[c];
  G[1] = cscale(G[1], (double)ti.sx);
  G[2] = cscale(G[2], (double)ti.sy);
  G[4] = cscale(G[4], (double)ti.sz);
  green_product(f, G);
}
template <>
__device__ __forceinline__ void apply_point<float>(float2* f, const TensorRef<float>& T, const TensorIdx& ti) {
  const float4* p = (const float4*)(T.t + (size_t)ti.idx * TENSOR_STRIDE);
  const float4 a = p[0], b = p[1];
  float g[7] = {a.x, a.y * ti.sx, a.z * ti.sy, a.w, b.x * ti.sz, b.y, b.z};
  green_product_real(f, g);
}

template <typename C, int BL, int RL>
__device__ __forceinline__ void copy_regs(C (&d)[BL][RL], const C (&s)[BL][RL]) {
#pragma unroll
  for (int m = 0; m < BL; ++m)
#pragma unroll
    for (int r = 0; r < RL; ++r) d[m][r] = s[m][r];
}

template <typename R, int E, int NT, int RL>
__global__ __launch_bounds__(NT) void k_pass_z(cplx<R>* __restrict__ B, int ltpp, GridDims g, Plan2 pf,
                                               Plan2 pi, TensorRef<R> tens) {
  using C = cplx<R>;
  constexpr int BL = (E + RL - 1) / RL;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x;
  const int lT = ilog2<NT>() - ltpp;
  const int T = 1 << lT, tpp = 1 << ltpp;
  const int t = tid & (T - 1), jl = tid >> lT;
  const int Px = g.Px, Py = g.Py, Pz = g.Pz, N = g.N;
  const int tiles_x = (Px + T - 1) >> lT;
  int tx, ky;
  z_tile<sizeof(R) == 8>(xcd_tile(blockIdx.x, gridDim.x), tiles_x, Py, tx, ky);
  const int kx = (tx << lT) + t;
  const bool valid = kx < Px;
  TensorIdx ti;
  const int ax = kx < g.Hx ? kx : Px - kx;
  const int ay = ky < g.Hy ? ky : Py - ky;
  ti.sx = kx < g.Hx ? 1.f : -1.f;
  ti.sy = ky < g.Hy ? 1.f : -1.f;
  const int axy = valid ? ax + g.Hx * ay : 0;
  const int HxHy = g.Hx * g.Hy;
  const int plane = Px * Py;
  const int nbL = pf.st[pf.nstage - 1].nb;
  C* lds = (C*)smem + t;
  auto lx = [T](int i) { return i * T; };
  {
    const int rhs = blockIdx.y;
    C* base = B + (int64_t)rhs * 5 * N * plane + (int64_t)ky * Px + kx;
    C h0[BL][RL], h1[BL][RL], h2[BL][RL], h3[BL][RL], h4[BL][RL];
#pragma unroll 1
    for (int c = 0; c < 5; ++c) {
      const C* bc = base + (int64_t)c * N * plane;
      auto gld = [&](int i) -> C { return (valid && i < N) ? bc[(int64_t)i * plane] : czero<R>(); };
      C o[BL][RL];
      fft_pencil_to_regs<R, -1, E, RL>(pf, lds, lx, jl, tpp, gld, o);
      switch (c) {
        case 0: copy_regs(h0, o); break;
        case 1: copy_regs(h1, o); break;
        case 2: copy_regs(h2, o); break;
        case 3: copy_regs(h3, o); break;
        default: copy_regs(h4, o); break;
      }
      __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < BL; ++m) {
      const int j = jl + m * tpp;
      if (j < nbL && valid) {
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          const int kz = j + r * nbL;
          const bool lo = kz < g.Hz;
          TensorIdx tk = ti;
          tk.sz = lo ? 1.f : -1.f;
          tk.idx = axy + HxHy * (lo ? kz : Pz - kz);
          C f[5] = {h0[m][r], h1[m][r], h2[m][r], h3[m][r], h4[m][r]};
          apply_point<R>(f, tens, tk);
          h0[m][r] = f[0]; h1[m][r] = f[1]; h2[m][r] = f[2]; h3[m][r] = f[3]; h4[m][r] = f[4];
        }
      }
    }
#pragma unroll 1
    for (int c = 0; c < 5; ++c) {
      C* bc = base + (int64_t)c * N * plane;
      auto gst = [&](int o, C val) { if (valid && o < N) bc[(int64_t)o * plane] = val; };
      C in[BL][RL];
      switch (c) {
        case 0: copy_regs(in, h0); break;
        case 1: copy_regs(in, h1); break;
        case 2: copy_regs(in, h2); break;
        case 3: copy_regs(in, h3); break;
        default: copy_regs(in, h4); break;
      }
      fft_pencil_from_regs<R, +1, E, RL>(pi, lds, lx, jl, tpp, in, gst);
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// pass Z for a one-radix Pz (2, 3, 4, 5, 7, 8: thin grids such as 4 voxels in
// z), one thread per (kx, ky) pencil: the thread loads the N <= (Pz+1)/2
// voxels of all five components first (every load in flight at once; HBM
// latency is hidden by memory-level parallelism, not by occupancy), takes the
// five Pz-point DFTs in registers (the zero padding is compile-time zero),
// applies the Green product at each kz and stores the first N points of the
// inverse DFTs.  A workgroup covers NT consecutive kx of one ky row, so each
// access is an NT-element contiguous run.  Same data flow and results as
// k_pass_z with a one-stage plan.
// ---------------------------------------------------------------------------
template <typename R, int PZ, int NT>
__global__ __launch_bounds__(NT, 1) void k_pass_z1(cplx<R>* __restrict__ B, GridDims g, TensorRef<R> tens) {
  using C = cplx<R>;
  constexpr int NM = (PZ + 1) / 2;   // most voxels a circulant embedding of length PZ holds
  const int Px = g.Px, Py = g.Py, N = g.N;
  const int tiles_x = (Px + NT - 1) / NT;
  int tx, ky;
  z_tile<sizeof(R) == 8>(xcd_tile(blockIdx.x, gridDim.x), tiles_x, Py, tx, ky);
  const int kx = tx * NT + (int)threadIdx.x;
  if (kx >= Px) return;   // no barriers below
  const int ax = kx < g.Hx ? kx : Px - kx;
  const int ay = ky < g.Hy ? ky : Py - ky;
  TensorIdx ti;
  ti.sx = kx < g.Hx ? 1.f : -1.f;
  ti.sy = ky < g.Hy ? 1.f : -1.f;
  const int axy = ax + g.Hx * ay;
  const int HxHy = g.Hx * g.Hy;
  const int rhs = blockIdx.y;
  const int64_t plane = (int64_t)Px * Py;
  C* base = B + (int64_t)rhs * 5 * N * plane + (int64_t)ky * Px + kx;
  auto point = [&](int k) {
    const bool lo = k < g.Hz;
    TensorIdx tk = ti;
    tk.sz = lo ? 1.f : -1.f;
    tk.idx = axy + HxHy * (lo ? k : g.Pz - k);
    return tk;
  };
  C f[5][PZ];
#pragma unroll
  for (int c = 0; c < 5; ++c)
#pragma unroll
    for (int i = 0; i < PZ; ++i)
      f[c][i] = (i < NM && i < N) ? base[(int64_t)(c * N + i) * plane] : czero<R>();
#pragma unroll
  for (int c = 0; c < 5; ++c) Dft<PZ, -1, C>::run(f[c]);
#pragma unroll
  for (int k = 0; k < PZ; ++k) {
    C p[5] = {f[0][k], f[1][k], f[2][k], f[3][k], f[4][k]};
    apply_point<R>(p, tens, point(k));
#pragma unroll
    for (int c = 0; c < 5; ++c) f[c][k] = p[c];
  }
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    Dft<PZ, +1, C>::run(f[c]);
#pragma unroll
    for (int o = 0; o < NM; ++o)
      if (o < N) base[(int64_t)(c * N + o) * plane] = f[c][o];
  }
}

// One-radix pass Z with two lanes per pencil (adjacent lanes 2p, 2p+1): one
// thread holding five Pz-point FP64 spectra plus the tensor needs more
// registers than two waves per SIMD allow.  Lane h owns components 3h + i
// (lane 0: 0,1,2; lane 1: 3,4) for the loads, the forward DFTs and the
// inverse DFTs; the Green product is split by kz (lane 0: kz < KA, lane 1:
// kz >= KA) and the lanes exchange the spectra they need by shuffles.  Same
// sums as k_pass_z1.
template <typename C>
__device__ __forceinline__ C lane_xor1(C v) {
  C r;
  r.x = __shfl_xor(v.x, 1, 64);
  r.y = __shfl_xor(v.y, 1, 64);
  return r;
}

template <typename R, int PZ, int NT, int OCC>
__global__ __launch_bounds__(NT, OCC) void k_pass_z1_pair(cplx<R>* __restrict__ B, GridDims g, TensorRef<R> tens) {
  using C = cplx<R>;
  constexpr int NM = (PZ + 1) / 2;   // most voxels an embedding of length PZ holds
  constexpr int KA = (PZ + 1) / 2;   // kz points of lane 0
  constexpr int TP = NT / 2;         // pencils per workgroup
  const int h = threadIdx.x & 1, pl = threadIdx.x >> 1;
  const int Px = g.Px, Py = g.Py, N = g.N;
  const int tiles_x = (Px + TP - 1) / TP;
  int tx, ky;
  z_tile<sizeof(R) == 8>(xcd_tile(blockIdx.x, gridDim.x), tiles_x, Py, tx, ky);
  const int kx = tx * TP + pl;
  if (kx >= Px) return;   // both lanes of a pencil leave together; no barriers below
  const int ax = kx < g.Hx ? kx : Px - kx;
  const int ay = ky < g.Hy ? ky : Py - ky;
  TensorIdx ti;
  ti.sx = kx < g.Hx ? 1.f : -1.f;
  ti.sy = ky < g.Hy ? 1.f : -1.f;
  const int axy = ax + g.Hx * ay;
  const int HxHy = g.Hx * g.Hy;
  const int rhs = blockIdx.y;
  const int64_t plane = (int64_t)Px * Py;
  C* base = B + (int64_t)rhs * 5 * N * plane + (int64_t)ky * Px + kx;
  C S[3][PZ];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int c = 3 * h + i;
#pragma unroll
    for (int n = 0; n < PZ; ++n)
      S[i][n] = (n < NM && n < N && c < 5) ? base[(int64_t)(c * N + n) * plane] : czero<R>();
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) Dft<PZ, -1, C>::run(S[i]);
  C Rs[3][PZ];   // product spectra of the lane's components, all kz
#pragma unroll
  for (int j = 0; j < KA; ++j) {
    const int j2c = KA + j < PZ ? KA + j : PZ - 1;   // lane 1's kz (clamped; unused past PZ - 1)
    const int k = h ? KA + j : j;
    // each lane sends what its partner needs at the partner's kz
    C rv[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) rv[i] = lane_xor1(h ? S[i][j] : S[i][j2c]);
    C f[5];
#pragma unroll
    for (int c = 0; c < 5; ++c)
      f[c] = h ? (c < 3 ? rv[c] : S[c - 3][j2c]) : (c < 3 ? S[c][j] : rv[c - 3]);
    if (k < PZ) {
      const bool lo = k < g.Hz;
      TensorIdx tk = ti;
      tk.sz = lo ? 1.f : -1.f;
      tk.idx = axy + HxHy * (lo ? k : g.Pz - k);
      apply_point<R>(f, tens, tk);
    }
    // back: lane 0 keeps components 0..2 at kz j and receives them at kz
    // KA + j; lane 1 keeps 3,4 at KA + j and receives them at j
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const C mine_other = i < 2 ? f[i < 2 ? 3 + i : 0] : czero<R>();   // components 3, 4
      const C back = lane_xor1(h ? f[i] : mine_other);
      Rs[i][j] = h ? back : f[i];
      if (KA + j < PZ) Rs[i][KA + j < PZ ? KA + j : 0] = h ? mine_other : back;
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int c = 3 * h + i;
    Dft<PZ, +1, C>::run(Rs[i]);
    if (c < 5) {
#pragma unroll
      for (int o = 0; o < NM; ++o)
        if (o < N) base[(int64_t)(c * N + o) * plane] = Rs[i][o];
    }
  }
}

// ---------------------------------------------------------------------------
// pass Z for power-of-two Pz (16..512), compile-time plan: radices, thread
// mapping and twiddle offsets are constants, the middle stages ping-pong
// between two LDS buffers (no barrier between a stage's reads and writes),
// and global loads run ahead of the transforms (see zp2_all_ahead).  HALF
// (N = Pz/2, the usual padding) makes the
// upper half of the first stage's inputs and of the last stage's outputs
// compile-time zero / unused.  Same data flow and results as k_pass_z.
// ---------------------------------------------------------------------------
template <typename R, int DIR, class P, int PZ, int TPP, int T, int S, int LAST> struct MidStages {
  __device__ __forceinline__ static void run(const cplx<R>* __restrict__ tw, int jl, cplx<R>* b0, cplx<R>* b1, int t) {
    if constexpr (S < LAST) {
      const cplx<R>* in = ((S - 1) & 1) ? b1 : b0;
      cplx<R>* out = (S & 1) ? b1 : b0;
      auto ld = [&](int i) { return in[i * T + t]; };
      auto st = [&](int o, cplx<R> v) { out[o * T + t] = v; };
      sstage<R, DIR, PZ, P::rad(S), P::ns(S), TPP>(tw + P::tw_off(S), jl, ld, st);
      __syncthreads();
      MidStages<R, DIR, P, PZ, TPP, T, S + 1, LAST>::run(tw, jl, b0, b1, t);
    }
  }
};

// The twiddles of a stage depend only on the thread, so the optimiser hoists
// their loads out of the five-component loops and keeps every stage's table
// in registers (54 VGPRs in complex64, 72 in complex128), which costs a wave
// slot per SIMD or spills.  Offsetting the table pointer by a zero the
// optimiser cannot see through, once per component, makes each iteration load
// them where they are used (from L1).
template <typename T> __device__ __forceinline__ const T* opaque_ptr(const T* p) {
  int z = 0;
  asm volatile("" : "+s"(z));
  return p + z;
}

// Rows of the B grid of one right-hand side through a buffer resource: the
// lane's byte offset (its (kx, ky) column and first z row) is one VGPR and
// each row offset is a uniform SGPR, where 64-bit flat addresses of the five
// components' rows would hold 40 VGPRs for the whole kernel.  The host checks
// that the right-hand side's five component grids span less than 2^31 bytes.
constexpr uint32_t kBufWord3 = 0x00020000u;   // raw buffer, 32-bit data format (CDNA)
template <typename R> struct ZRows {
  __amdgpu_buffer_rsrc_t rs;
  uint32_t vo;   // lane byte offset
  uint32_t pb;   // bytes per z row (plane)
  __device__ __forceinline__ ZRows(cplx<R>* rhs_base, uint32_t bytes, uint32_t lane_elem, uint32_t plane_elems) {
    rs = __builtin_amdgcn_make_buffer_rsrc(rhs_base, 0, (int)bytes, (int)kBufWord3);
    vo = lane_elem * (uint32_t)sizeof(cplx<R>);
    pb = plane_elems * (uint32_t)sizeof(cplx<R>);
  }
  __device__ __forceinline__ cplx<R> ld(int row) const {
    if constexpr (sizeof(R) == 4)
      return __builtin_bit_cast(cplx<R>, __builtin_amdgcn_raw_buffer_load_b64(rs, vo, (uint32_t)row * pb, 0));
    else
      return __builtin_bit_cast(cplx<R>, __builtin_amdgcn_raw_buffer_load_b128(rs, vo, (uint32_t)row * pb, 0));
  }
  __device__ __forceinline__ void st(int row, cplx<R> v) const {
    if constexpr (sizeof(R) == 4)
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(rs, 0, 0, 0)), v),
                                            rs, vo, (uint32_t)row * pb, 0);
    else
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b128(rs, 0, 0, 0)), v),
                                             rs, vo, (uint32_t)row * pb, 0);
  }
};

// Register budget of the compile-time pass-Z kernel, as waves per SIMD (the
// launch bound asks the compiler to fit it): complex64 with 8 elements per
// thread 4 (122 VGPRs; 5 spills), complex64 with 4 elements 3, complex128 2
// (3 spills 40+ VGPRs and runs slower).
template <typename R, int E> constexpr int zp2_waves() { return sizeof(R) == 8 ? 2 : (E == 8 ? 4 : 3); }
// FP32 with 4 elements per thread loads all five components before the first
// transform (each wave has its whole input in flight); with 8 elements one
// component ahead measured faster at 4 waves per SIMD (0.248 vs 0.262 ms on
// the 128^3 cube).
template <typename R, int E> constexpr bool zp2_all_ahead() { return sizeof(R) == 4 && E == 4; }
template <typename R, int E, int NT> constexpr int zp2_minb() { return zp2_waves<R, E>() * 256 / NT > 0 ? zp2_waves<R, E>() * 256 / NT : 1; }

template <typename R, int PZ, int NT, int E, bool HALF>
__global__ __launch_bounds__(NT, (zp2_minb<R, E, NT>())) void k_pass_z_p2(cplx<R>* __restrict__ B, GridDims g, const cplx<R>* __restrict__ twf,
                                                  const cplx<R>* __restrict__ twi, TensorRef<R> tens) {
  using C = cplx<R>;
  constexpr int TPP = PZ / E, T = NT / TPP;
  static_assert(T >= 1 && (T & (T - 1)) == 0, "pencils per workgroup must be a power of two");
  using PF = P2Plan<PZ, E, false>;
  using PI = P2Plan<PZ, E, true>;
  constexpr int S = PF::S;
  static_assert(S >= 2, "one-stage lengths use k_pass_z");
  constexpr int R0 = PF::rad(0);
  constexpr int BM0 = SStage<PZ, R0, TPP>::BM, NB0 = SStage<PZ, R0, TPP>::NB;
  constexpr int RL = PF::rad(S - 1), NSL = PF::ns(S - 1);
  constexpr int BML = SStage<PZ, RL, TPP>::BM, NBL = SStage<PZ, RL, TPP>::NB;
  static_assert(PI::rad(0) == RL, "inverse plan starts with the last forward radix");
  extern __shared__ __align__(16) unsigned char smem[];
  C* const buf0 = (C*)smem;
  C* const buf1 = (C*)smem + T * PZ;
  const int tid = threadIdx.x;
  const int t = tid & (T - 1), jl = tid >> clog2(T);
  const int Px = g.Px, Py = g.Py, N = g.N;
  const int tiles_x = (Px + T - 1) / T;
  int tx, ky;
  z_tile<sizeof(R) == 8>(xcd_tile(blockIdx.x, gridDim.x), tiles_x, Py, tx, ky);
  const int kx = tx * T + t;
  const bool valid = kx < Px;
  TensorIdx ti;
  const int ax = kx < g.Hx ? kx : Px - kx;
  const int ay = ky < g.Hy ? ky : Py - ky;
  ti.sx = kx < g.Hx ? 1.f : -1.f;
  ti.sy = ky < g.Hy ? 1.f : -1.f;
  const int axy = valid ? ax + g.Hx * ay : 0;
  const int HxHy = g.Hx * g.Hy;
  const int rhs = blockIdx.y;
  const int64_t plane = (int64_t)Px * Py;
  const ZRows<R> zr(B + (int64_t)rhs * 5 * N * plane, (uint32_t)(5 * N * plane * (int64_t)sizeof(C)),
                    (uint32_t)(ky * Px + kx + (int64_t)jl * plane), (uint32_t)plane);
  auto gload = [&](int c, C (&p)[BM0][R0]) {
#pragma unroll
    for (int m = 0; m < BM0; ++m)
#pragma unroll
      for (int r = 0; r < R0; ++r) {
        const int iu = m * TPP + r * NB0;   // row i = jl + iu (jl is in the lane offset)
        if constexpr (HALF) p[m][r] = (r < R0 / 2 && valid) ? zr.ld(c * N + iu) : czero<R>();   // i < Pz/2 iff r < R0/2
        else p[m][r] = (valid && jl + iu < N) ? zr.ld(c * N + iu) : czero<R>();
      }
  };
  C held[5][BML][RL];
  // Forward transforms.  With zp2_all_ahead (N = Pz/2) all five components
  // are loaded before the first transform and component c's input registers
  // die as its spectrum forms (the upper half of each input is the constant
  // zero padding and takes no registers); otherwise one component is loaded
  // ahead of the transform.
  constexpr bool ALL = HALF && zp2_all_ahead<R, E>();
  C in[ALL ? 5 : 2][BM0][R0];
  if constexpr (ALL) {
#pragma unroll
    for (int c = 0; c < 5; ++c) gload(c, in[c]);
  } else {
    gload(0, in[0]);
  }
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    C (&cur)[BM0][R0] = in[ALL ? c : (c & 1)];
    const C* const twc = opaque_ptr(twf);
    if constexpr (!ALL) {
      if (c < 4) gload(c + 1, in[(c + 1) & 1]);
    }
    {
      auto st = [&](int o, C val) { buf0[o * T + t] = val; };
      sstage_regs<R, -1, PZ, R0, 1, TPP>(twc, jl, cur, st);
    }
    __syncthreads();
    MidStages<R, -1, PF, PZ, TPP, T, 1, S - 1>::run(twc, jl, buf0, buf1, t);
    {
      const C* lin = (((S - 2) & 1) ? buf1 : buf0);
#pragma unroll
      for (int m = 0; m < BML; ++m)
#pragma unroll
        for (int r = 0; r < RL; ++r) held[c][m][r] = lin[(jl + m * TPP + r * NBL) * T + t];
      // last stage: Ns * RL = Pz, so j < Ns and j mod Ns = j
#pragma unroll
      for (int m = 0; m < BML; ++m)
        butterfly<R, -1, RL>(held[c][m], twc + PF::tw_off(S - 1) + (jl + m * TPP) * (RL - 1), NSL > 1);
    }
    __syncthreads();
  }
  if (valid) {
#pragma unroll
    for (int m = 0; m < BML; ++m)
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const int kz = jl + m * TPP + r * NBL;
        const bool lo = kz < g.Hz;
        TensorIdx tk = ti;
        tk.sz = lo ? 1.f : -1.f;
        tk.idx = axy + HxHy * (lo ? kz : g.Pz - kz);
        C f[5] = {held[0][m][r], held[1][m][r], held[2][m][r], held[3][m][r], held[4][m][r]};
        apply_point<R>(f, tens, tk);
#pragma unroll
        for (int c = 0; c < 5; ++c) held[c][m][r] = f[c];
      }
  }
  constexpr int RI = PI::rad(S - 1), NSI = PI::ns(S - 1);
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    const C* const tic = opaque_ptr(twi);
    {
      auto st = [&](int o, C val) { buf0[o * T + t] = val; };
      sstage_regs<R, +1, PZ, RL, 1, TPP>(tic, jl, held[c], st);
    }
    __syncthreads();
    MidStages<R, +1, PI, PZ, TPP, T, 1, S - 1>::run(tic, jl, buf0, buf1, t);
    {
      // last inverse stage: Ns * RI = Pz, so output o = j + r*Ns with j < Ns
      const C* in = (((S - 2) & 1) ? buf1 : buf0);
      constexpr int BMI = SStage<PZ, RI, TPP>::BM, NBI = SStage<PZ, RI, TPP>::NB;
      static_assert(NBI == NSI, "last stage");
#pragma unroll
      for (int m = 0; m < BMI; ++m) {
        const int j = jl + m * TPP;
        C w[RI];
#pragma unroll
        for (int r = 0; r < RI; ++r) w[r] = in[(j + r * NBI) * T + t];
        butterfly<R, +1, RI>(w, tic + PI::tw_off(S - 1) + j * (RI - 1), NSI > 1);
#pragma unroll
        for (int r = 0; r < RI; ++r) {
          const int ou = m * TPP + r * NSI;   // row o = jl + ou
          if constexpr (HALF) { if (r < RI / 2 && valid) zr.st(c * N + ou, w[r]); }   // o < Pz/2 iff r < RI/2
          else { if (valid && jl + ou < N) zr.st(c * N + ou, w[r]); }
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// pass Xi: inverse FFT along x, crop, diagonal, -Ar' phi, scatter.
// ---------------------------------------------------------------------------
template <typename R, int E, int NT, int mode>
__global__ __launch_bounds__(NT) void k_pass_x_inv(
    const cplx<R>* __restrict__ B, const cplx<R>* __restrict__ x, int64_t ldx, cplx<R>* __restrict__ y, int64_t ldy,
    const int* __restrict__ vmap, const int* __restrict__ lines, const int4* __restrict__ runs, int nlines, int ltpp,
    GridDims g, Plan2 pl, const cplx<R>* __restrict__ z, R inv_dx, const int* __restrict__ atp, const int* __restrict__ atc,
    const R* __restrict__ atv, const int* __restrict__ atpk, PackedInc pinc, int64_t num_curr,
    const int* __restrict__ slots, int dense,
    const cplx<R>* const* __restrict__ zslot, cplx<R> alpha0, const double2* __restrict__ alphas) {
  // mode 0: conv + diag - Ar'phi ; mode 1: conv + diag ; mode 2: conv only.
  // conv = j*omega*mu (alpha, per system or one for all) times the cropped
  // inverse transform: the scale is applied here, once per output element,
  // instead of at every spectral point of pass Z.
  using C = cplx<R>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x;
  const int tpp = 1 << ltpp, T = NT >> ltpp;
  const int jl = tid & (tpp - 1), t = tid >> ltpp;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int rc = blockIdx.y, rhs = rc / 5, comp = rc - rhs * 5;
  const int slot = slots ? slots[rhs] : rhs;
  const int l = tile * T + t;
  const bool valid = l < nlines;
  const int line = valid ? lines[l] : 0;
  const int L = g.L, Px = g.Px;
  const int yy = line % g.M, zz = line / g.M;
  const C* Bs = B + (((int64_t)rc * g.N + zz) * g.Py + yy) * Px;
  C* lds = (C*)smem + t * (lpad(Px - 1) + 1);
  auto lx = [](int i) { return lpad(i); };
  auto gld = [&](int i) -> C { return valid ? Bs[i] : czero<R>(); };
  const R w = inv_dx * (comp < 3 ? R(1) : comp == 3 ? R(1) / R(6) : R(1) / R(2));
  const C* xs = x + (int64_t)slot * ldx;
  C* ys = y + (int64_t)slot * ldy;
  const C* __restrict__ zv = zslot ? zslot[slot] : z;
  C alpha = alpha0;
  if (alphas) {
    const double2 a = alphas[slot];
    alpha.x = (R)a.x; alpha.y = (R)a.y;
  }
  const int* vm = vmap + (int64_t)L * line;
  const int64_t lbase = (int64_t)L * line;
  const int4 run = valid && !dense ? runs[l] : make_int4(-1, 0, 0, 0);
  // voxel id of output o of this line (-1: no voxel)
  auto vid_of = [&](int o) -> int {
    if (dense) return (int)(lbase + o);
    if (run.x >= 0) return (o >= run.x && o < run.y) ? run.z + o : -1;
    return vm[o];
  };
  // Modes 0 and 1: the diagonal term w z J and the incidence term Ar' phi of
  // every output of the line do not depend on the transform, so they are
  // gathered before it (their loads overlap each other instead of forming a
  // chain per output element in the last stage) into the line's row `pre`
  // of shared memory behind the transform's buffer.
  C* const pre = (C*)smem + T * (lpad(Px - 1) + 1) + t * L;
  if constexpr (mode <= 1) {
    constexpr int U = 4;   // outputs per round: all their loads in flight together
    auto coef = [](int c) { return c == 1 ? R(1) : c == 2 ? R(-1) : c == 3 ? R(0.5) : R(-0.5); };
    const int K = atpk ? pinc.k[comp] : 0;
    const int* const pk = atpk ? atpk + pinc.off[comp] : nullptr;
    for (int o0 = jl; o0 < L; o0 += U * tpp) {
      int vid[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int o = o0 + u * tpp;
        vid[u] = (valid && o < L) ? vid_of(o) : -1;
      }
      C zx[U], jv[U];
      int2 e[U][3];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int vi = vid[u] < 0 ? 0 : vid[u];
        zx[u] = vid[u] >= 0 ? zv[vi] : czero<R>();
        jv[u] = vid[u] >= 0 ? xs[(int64_t)comp * g.nv + vi] : czero<R>();
        if constexpr (mode == 0) {
          // packed row: up to three pairs of (node, coefficient code)
          // without the row-pointer chain; K is uniform across the workgroup
#pragma unroll
          for (int q = 0; q < 3; ++q)
            e[u][q] = (pk && vid[u] >= 0 && 2 * q < K) ? ((const int2*)(pk + (int64_t)vi * K))[q] : make_int2(0, 0);
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (vid[u] < 0) continue;
        C acc = cscale(cmul(zx[u], jv[u]), w);
        if constexpr (mode == 0) {
          C sum = czero<R>();
          if (pk) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
              const int ca = e[u][q].x & 7, cb = e[u][q].y & 7;
              if (ca) { const C ph = xs[num_curr + (e[u][q].x >> 3)]; const R v = coef(ca); sum.x += v * ph.x; sum.y += v * ph.y; }
              if (cb) { const C ph = xs[num_curr + (e[u][q].y >> 3)]; const R v = coef(cb); sum.x += v * ph.x; sum.y += v * ph.y; }
            }
          } else {
            const int64_t row = (int64_t)comp * g.nv + vid[u];
            for (int k = atp[row]; k < atp[row + 1]; ++k) {
              const C ph = xs[num_curr + atc[k]];
              sum.x += atv[k] * ph.x; sum.y += atv[k] * ph.y;
            }
          }
          acc = csub(acc, sum);
        }
        pre[o0 + u * tpp] = acc;
      }
    }
    __syncthreads();
  }
  auto gst = [&](int o, C acc) {
    if (!valid || o >= L) return;
    const int vid = vid_of(o);
    if (vid < 0) return;
    acc = cmul(alpha, acc);
    if constexpr (mode <= 1) acc = cadd(acc, pre[o]);
    ys[(int64_t)comp * g.nv + vid] = acc;
  };
  fft_pencil<R, +1, E>(pl, lds, lx, jl, tpp, gld, gst);
}

// Node rows: y_node = Ar * J.
template <typename R>
__global__ __launch_bounds__(256) void k_node_rows(
    const cplx<R>* __restrict__ x, int64_t ldx, cplx<R>* __restrict__ y, int64_t ldy, const int* __restrict__ arp,
    const int* __restrict__ arc, const R* __restrict__ arv, int num_node, int64_t num_curr,
    const int* __restrict__ slots, int nrhs) {
  // One thread per node serves a group of up to RB right-hand sides: the
  // row's indices and values are loaded once for all of them, and the
  // currents of RC right-hand sides at a time are loaded before any is used.
  // A node touches at most 6 currents; a longer row takes the looping path.
  // One-dimensional grid of node blocks x right-hand-side groups in XCD
  // order: the currents of consecutive y- and z-type nodes lie a grid row or
  // plane apart, so a wave touches 64 lines that the next node blocks use
  // again; those blocks run in the same L2.
  constexpr int RB = 8, RC = sizeof(R) == 4 ? 8 : 4, D = 6;
  const int nbx = (num_node + 255) >> 8;
  const int lt = xcd_tile(blockIdx.x, gridDim.x);
  const int by = lt / nbx, bx = lt - by * nbx;
  const int i = bx * 256 + threadIdx.x;
  if (i >= num_node) return;
  const int r0 = by * RB, r1 = min(nrhs, r0 + RB);
  const int p0 = arp[i], p1 = arp[i + 1];
  if (p1 - p0 <= D) {
    int kk[D];
    R vv[D];
#pragma unroll
    for (int u = 0; u < D; ++u) {
      kk[u] = p0 + u < p1 ? arc[p0 + u] : 0;
      vv[u] = p0 + u < p1 ? arv[p0 + u] : R(0);
    }
#pragma unroll
    for (int q0 = 0; q0 < RB; q0 += RC) {
      cplx<R> a[RC][D];
      int rs[RC];
#pragma unroll
      for (int c = 0; c < RC; ++c) {
        const int q = r0 + q0 + c;
        rs[c] = q < r1 ? (slots ? slots[q] : q) : -1;
        if (rs[c] >= 0) {
          const cplx<R>* xs = x + (int64_t)rs[c] * ldx;
#pragma unroll
          for (int u = 0; u < D; ++u)
            if (p0 + u < p1) a[c][u] = xs[kk[u]];
        }
      }
#pragma unroll
      for (int c = 0; c < RC; ++c) {
        if (rs[c] < 0) continue;
        cplx<R> s = czero<R>();
#pragma unroll
        for (int u = 0; u < D; ++u)
          if (p0 + u < p1) { s.x += vv[u] * a[c][u].x; s.y += vv[u] * a[c][u].y; }
        y[(int64_t)rs[c] * ldy + num_curr + i] = s;
      }
    }
  } else {
    for (int q = r0; q < r1; ++q) {
      const int rhs = slots ? slots[q] : q;
      const cplx<R>* xs = x + (int64_t)rhs * ldx;
      cplx<R> s = czero<R>();
      for (int k = p0; k < p1; ++k) {
        const cplx<R> v = xs[arc[k]];
        s.x += arv[k] * v.x; s.y += arv[k] * v.y;
      }
      y[(int64_t)rhs * ldy + num_curr + i] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// host launch logic
// ---------------------------------------------------------------------------
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

inline int log2i(int v) { int l = 0; while ((1 << l) < v) ++l; return l; }

// Threads per pencil for a pass, or an error naming the unsupported length.
// Starts from the smallest power of two that keeps ceil(n/tpp) <= E and halves
// it while some stage has fewer butterflies than threads and every stage's
// butterflies are still covered (ceil(E/R) * tpp >= n/R), so short
// transforms such as a single radix-7 stage use one thread per pencil.
inline int pass_tpp(const FftAxis2& ax, int E, int NT, const char* axis) {
  int tpp = fft_tpp(ax.n, E, NT);
  if (tpp < 0)
    throw Error(ERR_ARG, std::string("padded FFT length ") + std::to_string(ax.n) + " along " + axis +
                         " exceeds the pencil kernels (max " + std::to_string(E * NT) + ")");
  const auto& st = ax.stages(rmax_for(E));
  while (tpp > 1) {
    const int h = tpp / 2;
    bool idle = false, covered = true;
    for (const Stage2& S : st) {
      if (S.nb < tpp) idle = true;
      if ((E + S.rad - 1) / S.rad * h < S.nb) covered = false;
    }
    if (!idle || !covered) break;
    tpp = h;
  }
  return tpp;
}

// Z-pass shapes (threads per workgroup, element bound per thread) the
// autotuner chooses from; entry 0 is the default.  Fewer elements per thread
// hold fewer spectra in registers (higher occupancy) at the price of more
// stages (radix <= rmax_for(E)) and narrower kx tiles.
struct ZShape { int nt, e; };
constexpr ZShape kZShapes32[] = {{512, 8}, {256, 8}, {512, 4}, {256, 4}};
constexpr ZShape kZShapes64[] = {{512, 4}, {256, 4}};
template <typename R> constexpr int num_zshapes() {
  return sizeof(R) == 8 ? (int)(sizeof(kZShapes64) / sizeof(ZShape)) : (int)(sizeof(kZShapes32) / sizeof(ZShape));
}
template <typename R> static ZShape zshape(int i) { return sizeof(R) == 8 ? kZShapes64[i] : kZShapes32[i]; }

// A one-stage plan (Pz = RL) with one thread per pencil holds exactly RL
// elements per component: its kernel is built with E = RL, so no register is
// reserved for butterflies that do not exist.
template <typename R, int NT, int E, int RL>
static void launch_z_k(dim3 grid, size_t lds, hipStream_t st, cplx<R>* B, int ltpp, const GridDims& g,
                       const Plan2& pf, const Plan2& pi, TensorRef<R> tr) {
  if (pf.nstage == 1 && ltpp == 0)
    hipLaunchKernelGGL((k_pass_z<R, RL, NT, RL>), grid, dim3(NT), lds, st, B, ltpp, g, pf, pi, tr);
  else
    hipLaunchKernelGGL((k_pass_z<R, E, NT, RL>), grid, dim3(NT), lds, st, B, ltpp, g, pf, pi, tr);
}

template <typename R, int NT, int E>
static void launch_z_rl(int rl, dim3 grid, size_t lds, hipStream_t st, cplx<R>* B, int ltpp, const GridDims& g,
                        const Plan2& pf, const Plan2& pi, TensorRef<R> tr) {
  switch (rl) {
    case 2: launch_z_k<R, NT, E, 2>(grid, lds, st, B, ltpp, g, pf, pi, tr); break;
    case 3: launch_z_k<R, NT, E, 3>(grid, lds, st, B, ltpp, g, pf, pi, tr); break;
    case 4: launch_z_k<R, NT, E, 4>(grid, lds, st, B, ltpp, g, pf, pi, tr); break;
    case 5: launch_z_k<R, NT, E, 5>(grid, lds, st, B, ltpp, g, pf, pi, tr); break;
    case 7: launch_z_k<R, NT, E, 7>(grid, lds, st, B, ltpp, g, pf, pi, tr); break;
    case 8:
      if constexpr (rmax_for(E) >= 8) {
        launch_z_k<R, NT, E, 8>(grid, lds, st, B, ltpp, g, pf, pi, tr);
        break;
      }
      [[fallthrough]];
    default: throw Error(ERR_INTERNAL, "pass Z: unexpected last radix");
  }
}

// One-radix Pz: the one-thread-per-pencil kernel (returns false when it does
// not apply).
template <typename R>
static bool launch_z1(int nrhs, hipStream_t st, cplx<R>* B, const GridDims& g, TensorRef<R> tr) {
  // pencils per workgroup: the width (64, 128 or 256 kx) that leaves the
  // fewest idle pencils in the last kx tile (Px = 800: 64 wide, 4% idle,
  // instead of 256 wide, 22% idle); the wider one on a tie
  auto width = [&](int wmax) {
    int best = wmax;
    for (int w = wmax; w >= 64; w >>= 1)
      if (ceil_div(g.Px, w) * w < ceil_div(g.Px, best) * best) best = w;
    return best;
  };
  auto go = [&](auto pzc) {
    constexpr int PZ = decltype(pzc)::value;
    if (g.N > (PZ + 1) / 2) return false;
    if constexpr (sizeof(R) == 8 && PZ >= 4) {
      // FP64 at four or more kz points: two lanes per pencil (in FP32 one
      // thread per pencil at three waves per SIMD measured faster)
      auto pair = [&](auto ntc) {
        constexpr int NT = decltype(ntc)::value;
        const dim3 grid((unsigned)(ceil_div(g.Px, NT / 2) * g.Py), nrhs);
        hipLaunchKernelGGL((k_pass_z1_pair<R, PZ, NT, 2>), grid, dim3(NT), 0, st, B, g, tr);
      };
      if (width(128) == 64) pair(std::integral_constant<int, 128>());
      else pair(std::integral_constant<int, 256>());
    } else {
      auto one = [&](auto ntc) {
        constexpr int NT = decltype(ntc)::value;
        const dim3 grid((unsigned)(ceil_div(g.Px, NT) * g.Py), nrhs);
        hipLaunchKernelGGL((k_pass_z1<R, PZ, NT>), grid, dim3(NT), 0, st, B, g, tr);
      };
      const int w = width(256);
      if (w == 64) one(std::integral_constant<int, 64>());
      else if (w == 128) one(std::integral_constant<int, 128>());
      else one(std::integral_constant<int, 256>());
    }
    return true;
  };
  switch (g.Pz) {
    case 2: return go(std::integral_constant<int, 2>());
    case 3: return go(std::integral_constant<int, 3>());
    case 4: return go(std::integral_constant<int, 4>());
    case 5: return go(std::integral_constant<int, 5>());
    case 7: return go(std::integral_constant<int, 7>());
    case 8: return go(std::integral_constant<int, 8>());
    default: return false;
  }
}

// Power-of-two Pz in [16, 512] with the precision's default element bound:
// the compile-time kernel (returns false when it does not apply).
template <typename R, int NT, int E>
static bool launch_z_p2(int pz, int nrhs, hipStream_t st, cplx<R>* B, const GridDims& g, const Plan2& pf,
                        const Plan2& pi, TensorRef<R> tr) {
  // the kernel addresses one right-hand side's five component grids through
  // a buffer resource with 32-bit offsets
  if ((int64_t)5 * g.N * g.Px * g.Py * (int64_t)sizeof(cplx<R>) >= (int64_t(1) << 31)) return false;
  auto go = [&](auto pzc) {
    constexpr int PZ = decltype(pzc)::value;
    constexpr int TPP = PZ / E, T = NT / TPP;
    if constexpr (T < 1) {
      return false;
    } else {
      const dim3 grid((unsigned)(ceil_div(g.Px, T) * g.Py), nrhs);
      const size_t lds = 2 * (size_t)T * PZ * sizeof(cplx<R>);
      if (2 * g.N == PZ)
        hipLaunchKernelGGL((k_pass_z_p2<R, PZ, NT, E, true>), grid, dim3(NT), lds, st, B, g, (const cplx<R>*)pf.tw,
                           (const cplx<R>*)pi.tw, tr);
      else
        hipLaunchKernelGGL((k_pass_z_p2<R, PZ, NT, E, false>), grid, dim3(NT), lds, st, B, g, (const cplx<R>*)pf.tw,
                           (const cplx<R>*)pi.tw, tr);
      return true;
    }
  };
  switch (pz) {
    case 16: return go(std::integral_constant<int, 16>());
    case 32: return go(std::integral_constant<int, 32>());
    case 64: return go(std::integral_constant<int, 64>());
    case 128: return go(std::integral_constant<int, 128>());
    case 256: return go(std::integral_constant<int, 256>());
    case 512: return go(std::integral_constant<int, 512>());
    default: return false;
  }
}

template <typename R>
static void launch_z(int shape, int rl, dim3 grid, size_t lds, hipStream_t st, cplx<R>* B, int ltpp,
                     const GridDims& g, const Plan2& pf, const Plan2& pi, TensorRef<R> tr) {
  const ZShape z = zshape<R>(shape);
  if constexpr (sizeof(R) == 8) {
    if (z.nt == 512) launch_z_rl<R, 512, 4>(rl, grid, lds, st, B, ltpp, g, pf, pi, tr);
    else launch_z_rl<R, 256, 4>(rl, grid, lds, st, B, ltpp, g, pf, pi, tr);
  } else {
    if (z.nt == 512 && z.e == 8) launch_z_rl<R, 512, 8>(rl, grid, lds, st, B, ltpp, g, pf, pi, tr);
    else if (z.e == 8) launch_z_rl<R, 256, 8>(rl, grid, lds, st, B, ltpp, g, pf, pi, tr);
    else if (z.nt == 512) launch_z_rl<R, 512, 4>(rl, grid, lds, st, B, ltpp, g, pf, pi, tr);
    else launch_z_rl<R, 256, 4>(rl, grid, lds, st, B, ltpp, g, pf, pi, tr);
  }
}

// The passes on the pass-X output grid A (whose rows of empty x lines read
// as zero) and the B grid of the caller.
// Pass Z of one matvec on the B grid with the geometry's pass-Z shape of
// this precision (instantiated in its own translation unit per precision).
template <typename R>
void pass_z(Geometry& G, int nrhs, cplx<R>* B) {
  const GridDims& g = G.g;
  hipStream_t st = G.ctx->stream;
  const int zsh = std::min(std::max(G.zshape[sizeof(R) == 8 ? 1 : 0], 0), num_zshapes<R>() - 1);
  const ZShape Zs = zshape<R>(zsh);
  const int EZ = Zs.e, NTZ = Zs.nt;
  const size_t elem = sizeof(cplx<R>);
  const int tz = pass_tpp(G.fz, EZ, NTZ, "z");
  const Plan2 pzf = G.fz.plan<R>(rmax_for(EZ), false);
  const Plan2 pzi = G.fz.plan<R>(rmax_for(EZ), true);
  const int Tz = NTZ / tz;
  {
    dim3 grid(ceil_div(g.Px, Tz) * g.Py, nrhs);
    TensorRef<R> tr;
    if constexpr (sizeof(R) == 8) tr.t = G.t64.p; else tr.t = G.t32.p;
    const size_t lds = pzf.nstage > 1 ? (size_t)Tz * g.Pz * elem : 0;   // one stage: registers only
    const int ltz = log2i(tz);
    bool done = launch_z1<R>(nrhs, st, B, g, tr);
    if (!done && G.use_p2z) {
      if (EZ == 4)
        done = NTZ == 512 ? launch_z_p2<R, 512, 4>(g.Pz, nrhs, st, B, g, pzf, pzi, tr)
                          : launch_z_p2<R, 256, 4>(g.Pz, nrhs, st, B, g, pzf, pzi, tr);
      else if constexpr (sizeof(R) == 4)
        if (EZ == 8)
          done = NTZ == 512 ? launch_z_p2<R, 512, 8>(g.Pz, nrhs, st, B, g, pzf, pzi, tr)
                            : launch_z_p2<R, 256, 8>(g.Pz, nrhs, st, B, g, pzf, pzi, tr);
    }
    if (!done)
      launch_z<R>(zsh, G.fz.last_radix(rmax_for(EZ)), grid, lds, st, B, ltz, g, pzf, pzi, tr);
    VH_LAUNCHED();
  }
}
extern template void pass_z<double>(Geometry&, int, double2*);
extern template void pass_z<float>(Geometry&, int, float2*);

template <typename R, int EXY, int NT_XY>
void run_passes_t(Geometry& G, int nrhs, const cplx<R>* x, int64_t ldx, cplx<R>* y, int64_t ldy,
                         double are, double aim, int mode, const int* slots, const double2* alphas,
                         const cplx<R>* const* zslot, cplx<R>* A, cplx<R>* B) {
  const GridDims& g = G.g;
  hipStream_t st = G.ctx->stream;
  const int zsh = std::min(std::max(G.zshape[sizeof(R) == 8 ? 1 : 0], 0), num_zshapes<R>() - 1);
  const ZShape Zs = zshape<R>(zsh);
  const int EZ = Zs.e, NTZ = Zs.nt;
  const size_t elem = sizeof(cplx<R>);
  const int dense = G.dense ? 1 : 0;
  PassTimer* tm = G.timer;
  if (tm) { tm->collect(); VH_HIP(hipEventRecord(tm->ev[0], st)); }
  auto mark = [&](int k) { if (tm) VH_HIP(hipEventRecord(tm->ev[k], st)); };

  const int tx = pass_tpp(G.fx, EXY, NT_XY, "x");
  const int ty = pass_tpp(G.fy, EXY, NT_XY, "y");
  const int tz = pass_tpp(G.fz, EZ, NTZ, "z");
  const Plan2 px = G.fx.plan<R>(rmax_for(EXY), false);
  const Plan2 py = G.fy.plan<R>(rmax_for(EXY), false);
  const Plan2 pzf = G.fz.plan<R>(rmax_for(EZ), false);
  const Plan2 pzi = G.fz.plan<R>(rmax_for(EZ), true);
  const int Tx = NT_XY / tx, Ty = NT_XY / ty, Tz = NTZ / tz;
  const size_t lds_x = px.nstage > 1 ? (size_t)Tx * (lpad_host(g.Px - 1) + 1) * elem : 0;
  const size_t lds_y = py.nstage > 1 ? (size_t)Ty * g.Py * elem : 0;

  // X
  if (G.nlines > 0) {
    dim3 grid(ceil_div(G.nlines, Tx), nrhs * 5);
    hipLaunchKernelGGL((k_pass_x_fwd<R, EXY, NT_XY>), grid, dim3(NT_XY), lds_x, st, x, ldx, G.vmap.p, G.lines.p,
                       G.runs.p, G.nlines, log2i(tx), A, g, px, slots, dense);
    VH_LAUNCHED();
  }
  mark(1);
  // Y
  {
    dim3 grid(ceil_div(g.Px, Ty) * g.N, nrhs * 5);
    hipLaunchKernelGGL((k_pass_y<R, -1, EXY, NT_XY>), grid, dim3(NT_XY), lds_y, st, A, g.M, g.M,
                       B, g.Py, g.Py, log2i(ty), g, py, nullptr);
    VH_LAUNCHED();
  }
  mark(2);
  // Z
  pass_z<R>(G, nrhs, B);
  mark(3);
  // Yi (in place in B: rows y < M of every (rc, z) slab)
  {
    dim3 grid(ceil_div(g.Px, Ty) * g.N, nrhs * 5);
    hipLaunchKernelGGL((k_pass_y<R, +1, EXY, NT_XY>), grid, dim3(NT_XY), lds_y, st, B, g.Py,
                       g.Py, B, g.Py, g.M, log2i(ty), g, py, G.row_bits_on ? G.rowbits.p : nullptr);
    VH_LAUNCHED();
  }
  mark(4);
  // Xi
  if (G.nlines > 0) {
    dim3 grid(ceil_div(G.nlines, Tx), nrhs * 5);
    const R* atv;
    if constexpr (sizeof(R) == 8) atv = G.ArT.v64.p; else atv = G.ArT.v32.p;
    const cplx<R>* zz;
    if constexpr (sizeof(R) == 8) zz = G.z64.p; else zz = G.z32.p;
    if (mode <= 1) VH_REQUIRE(G.has_z || zslot, "matvec: impedance not set");
    if (mode == 0) VH_REQUIRE(G.has_inc, "matvec: incidence not set");
    cplx<R> alpha;
    alpha.x = (R)are;
    alpha.y = (R)aim;
    // the output mode is a template parameter: each variant is compiled for its own work
    // modes 0 and 1 keep each line's diagonal and incidence terms (L
    // elements) behind the transform's buffer
    const size_t lds_xi = mode <= 1 ? (size_t)Tx * (lpad_host(g.Px - 1) + 1 + g.L) * elem : lds_x;
    auto xi = [&](auto mc) {
      constexpr int MODE = decltype(mc)::value;
      hipLaunchKernelGGL((k_pass_x_inv<R, EXY, NT_XY, MODE>), grid, dim3(NT_XY), lds_xi, st, B, x, ldx, y, ldy,
                         G.vmap.p, G.lines.p, G.runs.p, G.nlines, log2i(tx), g, px, zz, (R)(1.0 / G.dx), G.ArT.ptr.p,
                         G.ArT.col.p, atv, G.atpk.p, G.pinc, G.num_curr, slots, dense, zslot, alpha, alphas);
    };
    if (mode == 0) xi(std::integral_constant<int, 0>{});
    else if (mode == 1) xi(std::integral_constant<int, 1>{});
    else xi(std::integral_constant<int, 2>{});
    VH_LAUNCHED();
  }
  mark(5);
  if (mode == 0 && G.num_node > 0) {
    const R* arv;
    if constexpr (sizeof(R) == 8) arv = G.Ar.v64.p; else arv = G.Ar.v32.p;
    VH_REQUIRE((int64_t)ceil_div(G.num_node, 256) * ceil_div(nrhs, 8) < (1LL << 31), "node rows: grid too large");
    dim3 grid((unsigned)(ceil_div(G.num_node, 256) * ceil_div(nrhs, 8)));
    hipLaunchKernelGGL((k_node_rows<R>), grid, dim3(256), 0, st, x, ldx, y, ldy, G.Ar.ptr.p, G.Ar.col.p, arv,
                       (int)G.num_node, G.num_curr, slots, nrhs);
    VH_LAUNCHED();
  }
  mark(6);
  if (tm) tm->pending = true;
}

// The X/Y-pass kernels of each shape are instantiated in translation units of
// their own (vh_matvec_xy32*.hip, vh_matvec_xy64.hip) so they compile in
// parallel.
#define VH_RUN_PASSES_ARGS(R)                                                                                   \
  Geometry&, int, const cplx<R>*, int64_t, cplx<R>*, int64_t, double, double, int, const int*, const double2*, \
      const cplx<R>* const*, cplx<R>*, cplx<R>*
extern template void run_passes_t<float, 16, 256>(VH_RUN_PASSES_ARGS(float));
extern template void run_passes_t<float, 8, 256>(VH_RUN_PASSES_ARGS(float));
extern template void run_passes_t<float, 16, 128>(VH_RUN_PASSES_ARGS(float));
extern template void run_passes_t<double, 8, 256>(VH_RUN_PASSES_ARGS(double));
extern template void run_passes_t<double, 4, 256>(VH_RUN_PASSES_ARGS(double));
extern template void run_passes_t<double, 8, 128>(VH_RUN_PASSES_ARGS(double));

// The passes with the geometry's X/Y shape of this precision.
template <typename R>
static void run_passes(Geometry& G, int nrhs, const cplx<R>* x, int64_t ldx, cplx<R>* y, int64_t ldy,
                       double are, double aim, int mode, const int* slots, const double2* alphas,
                       const cplx<R>* const* zslot, cplx<R>* A, cplx<R>* B) {
  const int sh = std::min(std::max(G.xyshape[sizeof(R) == 8 ? 1 : 0], 0), num_xyshapes<R>() - 1);
  if constexpr (sizeof(R) == 8) {
    if (sh == 1) run_passes_t<R, 4, 256>(G, nrhs, x, ldx, y, ldy, are, aim, mode, slots, alphas, zslot, A, B);
    else if (sh == 2) run_passes_t<R, 8, 128>(G, nrhs, x, ldx, y, ldy, are, aim, mode, slots, alphas, zslot, A, B);
    else run_passes_t<R, 8, 256>(G, nrhs, x, ldx, y, ldy, are, aim, mode, slots, alphas, zslot, A, B);
  } else {
    if (sh == 1) run_passes_t<R, 8, 256>(G, nrhs, x, ldx, y, ldy, are, aim, mode, slots, alphas, zslot, A, B);
    else if (sh == 2) run_passes_t<R, 16, 128>(G, nrhs, x, ldx, y, ldy, are, aim, mode, slots, alphas, zslot, A, B);
    else run_passes_t<R, 16, 256>(G, nrhs, x, ldx, y, ldy, are, aim, mode, slots, alphas, zslot, A, B);
  }
}

inline void check_matvec(const Geometry& G, int nrhs) {
  VH_REQUIRE(G.has_tensor, "matvec: Green's tensor not set");
  VH_REQUIRE(nrhs >= 1, "matvec: nrhs must be >= 1");
  VH_REQUIRE(nrhs * 5 <= 65535, "matvec: too many right-hand sides in one call");
}

template <typename R>
void matvec(Geometry& G, int nrhs, const cplx<R>* x, int64_t ldx, cplx<R>* y, int64_t ldy, double are, double aim,
            int mode, const int* slots, const double2* alphas, const cplx<R>* const* zslot) {
  check_matvec(G, nrhs);
  ensure_ws<R>(G, nrhs);
  run_passes<R>(G, nrhs, x, ldx, y, ldy, are, aim, mode, slots, alphas, zslot,
                (cplx<R>*)(sizeof(R) == 8 ? G.wsA64.p : G.wsA32.p), (cplx<R>*)G.wsB.p);
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Clears n16 16-byte words on the stream with a fill kernel (no memset node
// when the caller captures the stream into a graph).
void zero_fill16(void* p, int64_t n16, hipStream_t st);

template <typename R>
void matvec_ws(Geometry& G, int nrhs, const cplx<R>* x, int64_t ldx, cplx<R>* y, int64_t ldy, double are,
               double aim, int mode, char* ws) {
  check_matvec(G, nrhs);
  const GridDims& g = G.g;
  const size_t a = (size_t)nrhs * 5 * g.Px * g.M * g.N * sizeof(cplx<R>);
  // the rows of x lines without voxels are never written by pass X: clear
  // them (the whole grid, with a fill kernel: no memset node if captured)
  if ((int64_t)G.nlines < (int64_t)g.M * g.N) {
    const int64_t n16 = (int64_t)(a / 16);
    zero_fill16(ws, n16, G.ctx->stream);
  }
  run_passes<R>(G, nrhs, x, ldx, y, ldy, are, aim, mode, nullptr, nullptr, nullptr, (cplx<R>*)ws,
                (cplx<R>*)(ws + align256(a)));
}

}  // namespace vh
