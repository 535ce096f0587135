// Green's tensor producer on the GPU.
//
// Replaces src_lse_formation/lse_generate_circulant_tensor.m (default branches)
// for the matvec: real-space G_mn by Gauss-Legendre volume-volume quadrature
// (far: 3^6 points, medium block [0,5)^3: 5^6 points; volume_volume.m:1-117),
// the 2x2x2 near block from the pre-computed K table scaled by (dx/1mm)^{5,6,7}
// (:363-400, assembled with the rotation scatter of :405-490), normalisation
// (:761-768), signed circulant embedding (:776-816) and the forward FFT (:829).
// The spectrum's octant (kx<=Px/2, ky<=Py/2, kz<=Pz/2) is kept, divided by
// Px*Py*Pz; the other octants follow from the even/odd symmetry of each
// component (gperiodic_coeff_nop_linJVIE.m:13-23).
#include "vh_geom.h"
#include <cmath>

namespace vh {

// K_near_interactions_1mm.txt (6 rows x 3 complex): [pulse-pulse, linear(x)-pulse,
// linear(x)-linear(x')] for offsets (mx,my,mz) in {(1,1,1),(1,2,1),(1,2,2),(2,1,1),
// (2,2,1),(2,2,2)} of a 1 mm cube pair (lse_generate_circulant_tensor.m:342-363).
static const double KTAB[6][6] = {
    {1.497896808995041e-016, -5.003461427972696e-027, -1.884273158210909e-037, -2.012211605895784e-048,
     2.910077873947119e-024, -3.626541131608855e-051},
    {7.805636278778824e-017, -5.003461427972697e-027, 1.206473223753176e-035, -4.160008679876429e-049,
     6.11217610670709e-025, -2.004923164762494e-052},
    {5.638025079837469e-017, -5.003461427972694e-027, 2.412452305035321e-035, 2.98904327368899e-049,
     2.297463434675418e-025, -5.006665539022287e-052},
    {7.805636278778813e-017, -5.003461427972696e-027, -5.757414693177654e-021, -7.50496380728818e-048,
     -7.440943011730615e-025, -1.219881017509337e-051},
    {5.638025079837419e-017, -5.003461427972694e-027, -2.390823543614732e-021, -3.06145823959795e-048,
     -1.275734678285788e-025, 1.681050961223937e-051},
    {4.605920193990263e-017, -5.003461427972694e-027, -1.321546896796384e-021, 2.534523806813602e-049,
     -1.402863426594222e-026, -1.963545763613549e-051},
};

constexpr double PI = 3.141592653589793238462643383279502884;
constexpr double K0_ONE_OVER_R = 1e-30;   // lse_generate_circulant_tensor.m:23

// Gauss-Legendre nodes and weights of orders 3 and 5 (gauss_1d.m:80-115).
__constant__ double c_gl3x[3] = {-0.774596669241483377035853079956, 0.0, 0.774596669241483377035853079956};
__constant__ double c_gl3w[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
__constant__ double c_gl5x[5] = {-0.906179845938663992797626878299, -0.538469310105683091036314420700, 0.0,
                                 0.538469310105683091036314420700, 0.906179845938663992797626878299};
__constant__ double c_gl5w[5] = {0.236926885056189087514264040720, 0.478628670499366468041291514836,
                                 0.568888888888888888888888888889, 0.478628670499366468041291514836,
                                 0.236926885056189087514264040720};

// The 7 integrals [K1, K3, K5, K7, K8, K9, K10] of volume_volume.m:54-97 for an
// offset o (metres) with an NP^6 product rule, accumulated in complex FP64.
template <int NP>
__device__ void vv_sums(double ox, double oy, double oz, double dx, double2* K) {
  const double* gx = NP == 3 ? c_gl3x : c_gl5x;
  const double* gw = NP == 3 ? c_gl3w : c_gl5w;
  const double h = dx / 2;
  double re[7] = {0, 0, 0, 0, 0, 0, 0}, im[7] = {0, 0, 0, 0, 0, 0, 0};
  const double inv4pi = 1.0 / (4.0 * PI);
  for (int ix = 0; ix < NP; ++ix)
    for (int jx = 0; jx < NP; ++jx) {
      const double ax = ox + h * (gx[ix] - gx[jx]);
      const double wx = gw[ix] * gw[jx], tx = h * gx[ix], ttx = h * gx[jx] * h * gx[ix];
      for (int iy = 0; iy < NP; ++iy)
        for (int jy = 0; jy < NP; ++jy) {
          const double ay = oy + h * (gx[iy] - gx[jy]);
          const double wy = wx * gw[iy] * gw[jy], ty = h * gx[iy], tty = h * gx[jy] * h * gx[iy];
          const double axy2 = ax * ax + ay * ay;
          for (int iz = 0; iz < NP; ++iz)
            for (int jz = 0; jz < NP; ++jz) {
              const double az = oz + h * (gx[iz] - gx[jz]);
              const double w = wy * gw[iz] * gw[jz], tz = h * gx[iz], ttz = h * gx[jz] * h * gx[iz];
              const double R = sqrt(axy2 + az * az);
              double s, c;
              sincos(K0_ONE_OVER_R * R, &s, &c);
              const double kr = w * c / R * inv4pi, ki = -w * s / R * inv4pi;   // w * exp(-i k0 R)/(4 pi R)
              re[0] += kr; im[0] += ki;
              re[1] += kr * tx; im[1] += ki * tx;
              re[2] += kr * ty; im[2] += ki * ty;
              re[3] += kr * tz; im[3] += ki * tz;
              re[4] += kr * ttx; im[4] += ki * ttx;
              re[5] += kr * tty; im[5] += ki * tty;
              re[6] += kr * ttz; im[6] += ki * ttz;
            }
        }
    }
  const double J = h * h * h * h * h * h;
  for (int c = 0; c < 7; ++c) { K[c].x = re[c] * J; K[c].y = im[c] * J; }
}

// Normalisation factors of lse_generate_circulant_tensor.m:761-768 for each
// component (all divided by dx^4 first).
__host__ __device__ inline double comp_norm(int c, double dx) {
  const double d4 = dx * dx * dx * dx;
  switch (c) {
    case 1: case 2: case 4: return 1.0 / (d4 * dx);
    case 3: case 5: case 6: return 1.0 / (d4 * dx * dx);
    default: return 1.0 / d4;
  }
}

// One thread per offset; near block offsets take the host-assembled values.
__global__ void k_green_realspace(double2* __restrict__ G, int L, int M, int N, double dx,
                                  const double2* __restrict__ nearblk, int nmL, int nmM, int nmN) {
  const int64_t V = (int64_t)L * M * N;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= V) return;
  const int mx = (int)(id % L);
  const int my = (int)((id / L) % M);
  const int mz = (int)(id / ((int64_t)L * M));
  double2 I[7];
  if (mx < 2 && my < 2 && mz < 2) {
    for (int c = 0; c < 7; ++c) I[c] = nearblk[(mx + 2 * (my + 2 * mz)) * 7 + c];
  } else {
    double2 K[7];
    if (mx < nmL && my < nmM && mz < nmN)
      vv_sums<5>(mx * dx, my * dx, mz * dx, dx, K);
    else
      vv_sums<3>(mx * dx, my * dx, mz * dx, dx, K);
    // volume_volume.m:103-115
    I[0] = K[0];
    I[1].x = -K[1].x; I[1].y = -K[1].y;
    I[2] = K[2];
    I[3] = cadd(K[4], K[5]);
    I[4].x = 2 * K[3].x; I[4].y = 2 * K[3].y;
    I[5] = csub(K[4], K[5]);
    I[6] = cadd(cadd(K[4], K[5]), cscale(K[6], 4.0));
  }
  for (int c = 0; c < 7; ++c) G[id + V * c] = cscale(I[c], comp_norm(c, dx));
}

// Signed circulant embedding into the padded grid (lse_generate_circulant_tensor.m:799-816),
// generalised to any padded length P >= 2n-1: Gp[P-i] = s * G[i] for 1 <= i < n.
__global__ void k_embed(const double2* __restrict__ G, double2* __restrict__ Gp, GridDims g) {
  const int64_t P = (int64_t)g.Px * g.Py * g.Pz;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= P) return;
  const int px = (int)(id % g.Px);
  const int py = (int)((id / g.Px) % g.Py);
  const int pz = (int)(id / ((int64_t)g.Px * g.Py));
  int sx = px, sy = py, sz = pz;
  bool rx = false, ry = false, rz = false, zero = false;
  if (px >= g.L) { sx = g.Px - px; rx = true; if (sx >= g.L || sx < 1) zero = true; }
  if (py >= g.M) { sy = g.Py - py; ry = true; if (sy >= g.M || sy < 1) zero = true; }
  if (pz >= g.N) { sz = g.Pz - pz; rz = true; if (sz >= g.N || sz < 1) zero = true; }
  const int64_t V = (int64_t)g.L * g.M * g.N;
  const int64_t src = sx + (int64_t)g.L * (sy + (int64_t)g.M * sz);
  for (int c = 0; c < 7; ++c) {
    double2 v; v.x = 0; v.y = 0;
    if (!zero) {
      v = G[src + V * c];
      // gperiodic_coeff_nop_linJVIE.m: comp 2 odd in x, comp 3 odd in y, comp 5 odd in z
      double s = 1.0;
      if (c == 1 && rx) s = -s;
      if (c == 2 && ry) s = -s;
      if (c == 4 && rz) s = -s;
      v = cscale(v, s);
    }
    Gp[id + P * c] = v;
  }
}

// In-place FFT of contiguous pencils of length n (npen of them).
template <typename R, int DIR, int EMAX>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_lines(cplx<R>* data, int64_t npen, int T, FftPlan pl) {
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<R>* buf = (cplx<R>*)smem;
  const int n = pl.n;
  const int64_t p0 = (int64_t)blockIdx.x * T;
  for (int e = threadIdx.x; e < T * n; e += blockDim.x) {
    const int p = e / n;
    cplx<R> v; v.x = 0; v.y = 0;
    if (p0 + p < npen) v = data[p0 * n + e];
    buf[e] = v;
  }
  __syncthreads();
  lds_fft<R, DIR, EMAX>(buf, T, n, 1, pl, threadIdx.x, blockDim.x, false);
  for (int e = threadIdx.x; e < T * n; e += blockDim.x) {
    const int p = e / n;
    if (p0 + p < npen) data[p0 * n + e] = buf[e];
  }
}

// In-place FFT along a strided axis: pencils of length n with stride `stride`,
// tiles of T consecutive columns out of `ncols`; `nouter` independent slabs of
// size `slab`.
template <typename R, int DIR, int EMAX>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_strided(cplx<R>* data, int64_t stride, int ncols, int T,
                                                             int64_t slab, FftPlan pl) {
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<R>* buf = (cplx<R>*)smem;
  const int n = pl.n;
  const int tiles = (ncols + T - 1) / T;
  const int tile = blockIdx.x % tiles;
  const int64_t outer = blockIdx.x / tiles;
  cplx<R>* base = data + outer * slab + (int64_t)tile * T;
  const int c0 = tile * T;
  for (int e = threadIdx.x; e < T * n; e += blockDim.x) {
    const int i = e / T, t = e - i * T;
    cplx<R> v; v.x = 0; v.y = 0;
    if (c0 + t < ncols) v = base[(int64_t)i * stride + t];
    buf[e] = v;
  }
  __syncthreads();
  lds_fft<R, DIR, EMAX>(buf, T, 1, T, pl, threadIdx.x, blockDim.x, true);
  for (int e = threadIdx.x; e < T * n; e += blockDim.x) {
    const int i = e / T, t = e - i * T;
    if (c0 + t < ncols) base[(int64_t)i * stride + t] = buf[e];
  }
}

__global__ void k_extract_octant(const double2* __restrict__ F, double2* __restrict__ t64, float* __restrict__ t32,
                                 GridDims g, double invP) {
  const int64_t O = (int64_t)g.Hx * g.Hy * g.Hz;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= O * 7) return;
  const int c = (int)(id / O);
  const int64_t o = id - (int64_t)c * O;
  const int kx = (int)(o % g.Hx);
  const int ky = (int)((o / g.Hx) % g.Hy);
  const int kz = (int)(o / ((int64_t)g.Hx * g.Hy));
  const int64_t P = (int64_t)g.Px * g.Py * g.Pz;
  const double2 v = cscale(F[kx + (int64_t)g.Px * (ky + (int64_t)g.Py * kz) + P * c], invP);
  t64[o * TENSOR_STRIDE + c] = v;
  const bool odd = (c == 1 || c == 2 || c == 4);
  t32[o * TENSOR_STRIDE + c] = (float)(odd ? v.y : v.x);
}

// Near block (offsets in {0,1}^3), assembled on the host exactly as
// lse_generate_circulant_tensor.m:397-490 does from the K table.
static void near_block(double dx, std::vector<double2>& out /* 8*7 */) {
  std::vector<std::complex<double>> G(8 * 7, 0.0);
  auto at = [&](int x, int y, int z, int c) -> std::complex<double>& { return G[(x + 2 * (y + 2 * z)) * 7 + c]; };
  const double s = dx / 1e-3;
  const double sc[3] = {std::pow(s, 5), std::pow(s, 6), std::pow(s, 7)};
  int row = 0;
  for (int mx = 0; mx < 2; ++mx)
    for (int my = 0; my < 2; ++my)
      for (int mz = 0; mz <= my; ++mz, ++row) {
        std::complex<double> k1(KTAB[row][0], KTAB[row][1]), k2(KTAB[row][2], KTAB[row][3]), k3(KTAB[row][4], KTAB[row][5]);
        k1 *= sc[0]; k2 *= sc[1]; k3 *= sc[2];
        at(mx, my, mz, 0) = k1; at(mx, my, mz, 1) = -k2;
        at(mx, my, mz, 3) += k3; at(mx, my, mz, 5) += k3; at(mx, my, mz, 6) += k3;
        if (my != mz) {
          at(mx, mz, my, 0) = k1; at(mx, mz, my, 1) = -k2;
          at(mx, mz, my, 3) += k3; at(mx, mz, my, 5) += k3; at(mx, mz, my, 6) += k3;
        }
        at(mz, mx, my, 2) = k2;
        at(mz, mx, my, 3) += k3; at(mz, mx, my, 5) -= k3; at(mz, mx, my, 6) += k3;
        if (my != mz) {
          at(my, mx, mz, 2) = k2;
          at(my, mx, mz, 3) += k3; at(my, mx, mz, 5) -= k3; at(my, mx, mz, 6) += k3;
        }
        at(my, mz, mx, 4) = 2.0 * k2;
        at(my, mz, mx, 6) += 4.0 * k3;
        if (my != mz) {
          at(mz, my, mx, 4) = 2.0 * k2;
          at(mz, my, mx, 6) += 4.0 * k3;
        }
      }
  out.resize(56);
  // the kernel applies the component normalisation, so store raw values
  for (int i = 0; i < 56; ++i) { out[i].x = G[i].real(); out[i].y = G[i].imag(); }
}

void generate_realspace_green(Geometry& Gm, DevBuf<double2>& out) {
  const GridDims& g = Gm.g;
  const int64_t V = (int64_t)g.L * g.M * g.N;
  std::vector<double2> nb;
  near_block(Gm.dx, nb);
  DevBuf<double2> dnb(56);
  VH_HIP(hipMemcpy(dnb.p, nb.data(), 56 * sizeof(double2), hipMemcpyHostToDevice));
  out.alloc(V * 7);
  const int nm = 5;   // n_medium, lse_generate_circulant_tensor.m:43
  hipLaunchKernelGGL(k_green_realspace, dim3(ceil_div(V, 128)), dim3(128), 0, Gm.ctx->stream, out.p, g.L, g.M, g.N,
                     Gm.dx, dnb.p, std::min(g.L, nm), std::min(g.M, nm), std::min(g.N, nm));
  VH_LAUNCHED();
  VH_HIP(hipStreamSynchronize(Gm.ctx->stream));
  // self terms G(1,1,1,{1,4,7}) (:772)
  const double d4 = Gm.dx * Gm.dx * Gm.dx * Gm.dx;
  Gm.st2[0] = nb[0].x / d4;                 Gm.st2[1] = nb[0].y / d4;
  Gm.st2[2] = nb[3].x / (d4 * Gm.dx * Gm.dx); Gm.st2[3] = nb[3].y / (d4 * Gm.dx * Gm.dx);
  Gm.st2[4] = nb[6].x / (d4 * Gm.dx * Gm.dx); Gm.st2[5] = nb[6].y / (d4 * Gm.dx * Gm.dx);
}

template <typename R, int DIR>
static void fft3d_inplace(Geometry& Gm, cplx<R>* data, int ncomp) {
  const GridDims& g = Gm.g;
  hipStream_t st = Gm.ctx->stream;
  constexpr int EM = sizeof(R) == 8 ? 16 : 32;
  const size_t elem = sizeof(cplx<R>);
  auto pickT = [&](int n, int want) {
    int T = want;
    while (T > 1 && ((int64_t)T * n > (int64_t)EM * FFT_THREADS || (size_t)T * n * elem > 65536)) T >>= 1;
    VH_REQUIRE((int64_t)T * n <= (int64_t)EM * FFT_THREADS && (size_t)T * n * elem <= 65536, "FFT length too large");
    return T;
  };
  const int64_t P = (int64_t)g.Px * g.Py * g.Pz;
  {  // x: contiguous lines
    const int T = pickT(g.Px, 4);
    const int64_t npen = (int64_t)g.Py * g.Pz * ncomp;
    hipLaunchKernelGGL((k_fft_lines<R, DIR, EM>), dim3(ceil_div(npen, T)), dim3(FFT_THREADS), T * g.Px * elem, st,
                       data, npen, T, Gm.ax.plan<R>());
    VH_LAUNCHED();
  }
  {  // y: stride Px, slabs = z planes
    const int T = pickT(g.Py, 8);
    const int64_t nslab = (int64_t)g.Pz * ncomp;
    hipLaunchKernelGGL((k_fft_strided<R, DIR, EM>), dim3(ceil_div(g.Px, T) * nslab), dim3(FFT_THREADS),
                       T * g.Py * elem, st, data, (int64_t)g.Px, g.Px, T, (int64_t)g.Px * g.Py, Gm.ay.plan<R>());
    VH_LAUNCHED();
  }
  {  // z: stride Px*Py, columns = Px*Py, slabs = components
    const int T = pickT(g.Pz, 8);
    hipLaunchKernelGGL((k_fft_strided<R, DIR, EM>), dim3(ceil_div((int64_t)g.Px * g.Py, T) * ncomp),
                       dim3(FFT_THREADS), T * g.Pz * elem, st, data, (int64_t)g.Px * g.Py, g.Px * g.Py, T, P,
                       Gm.az.plan<R>());
    VH_LAUNCHED();
  }
}

void tensor_from_realspace(Geometry& Gm, const double2* Greal) {
  const GridDims& g = Gm.g;
  const int64_t P = (int64_t)g.Px * g.Py * g.Pz;
  DevBuf<double2> Gp(P * 7);
  hipStream_t st = Gm.ctx->stream;
  hipLaunchKernelGGL(k_embed, dim3(ceil_div(P, 256)), dim3(256), 0, st, Greal, Gp.p, g);
  VH_LAUNCHED();
  fft3d_inplace<double, -1>(Gm, Gp.p, 7);
  const size_t O = Gm.octant();
  Gm.t64.alloc(TENSOR_STRIDE * O);
  Gm.t32.alloc(TENSOR_STRIDE * O);
  VH_HIP(hipMemsetAsync(Gm.t64.p, 0, Gm.t64.bytes(), st));
  VH_HIP(hipMemsetAsync(Gm.t32.p, 0, Gm.t32.bytes(), st));
  hipLaunchKernelGGL(k_extract_octant, dim3(ceil_div(7 * O, 256)), dim3(256), 0, st, Gp.p, Gm.t64.p, Gm.t32.p, g,
                     1.0 / (double)P);
  VH_LAUNCHED();
  VH_HIP(hipStreamSynchronize(st));
  Gm.has_tensor = true;
}

// Public FFT test hook: in-place 3-D FFT of ncomp complex grids (Px,Py,Pz).
void fft3d_test(Geometry& Gm, void* data, int ncomp, int prec, int dir) {
  if (prec == 0) {
    if (dir < 0) fft3d_inplace<double, -1>(Gm, (double2*)data, ncomp); else fft3d_inplace<double, +1>(Gm, (double2*)data, ncomp);
  } else {
    if (dir < 0) fft3d_inplace<float, -1>(Gm, (float2*)data, ncomp); else fft3d_inplace<float, +1>(Gm, (float2*)data, ncomp);
  }
}

}  // namespace vh
