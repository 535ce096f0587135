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
#include "vh_matvec_impl.h"

namespace vh {

int matvec_num_zshapes(int prec) { return prec == 0 ? num_zshapes<double>() : num_zshapes<float>(); }
int matvec_num_xyshapes(int prec) { return prec == 0 ? num_xyshapes<double>() : num_xyshapes<float>(); }

template <typename R> static bool xyshape_applies(const Geometry& G, int shape) {
  const XYShape z = xyshape<R>(shape);
  return fft_tpp(G.fx.n, z.e, z.nt) >= 0 && fft_tpp(G.fy.n, z.e, z.nt) >= 0;
}

bool matvec_xyshape_applies(const Geometry& G, int prec, int shape) {
  if (shape < 0 || shape >= matvec_num_xyshapes(prec)) return false;
  return prec == 0 ? xyshape_applies<double>(G, shape) : xyshape_applies<float>(G, shape);
}

template <typename R> static bool zshape_applies(const Geometry& G, int shape) {
  const GridDims& g = G.g;
  switch (g.Pz) {   // the one-radix kernel (launch_z1) has no shape
    case 2: case 3: case 4: case 5: case 7: case 8:
      if (g.N <= (g.Pz + 1) / 2) return false;
      break;
    default: break;
  }
  const ZShape z = zshape<R>(shape);
  return fft_tpp(G.fz.n, z.e, z.nt) >= 0;
}

bool matvec_zshape_applies(const Geometry& G, int prec, int shape) {
  if (shape < 0 || shape >= matvec_num_zshapes(prec)) return false;
  return prec == 0 ? zshape_applies<double>(G, shape) : zshape_applies<float>(G, shape);
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

size_t matvec_ws_bytes(const Geometry& G, int prec, int nrhs) {
  const GridDims& g = G.g;
  const size_t c = prec == 0 ? 16 : 8;
  return align256((size_t)nrhs * 5 * g.Px * g.M * g.N * c) + (size_t)nrhs * 5 * g.Px * g.Py * g.N * c;
}

__global__ void k_zero16(uint4* __restrict__ p, int64_t n) {
  const uint4 z = make_uint4(0, 0, 0, 0);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = z;
}

void zero_fill16(void* p, int64_t n16, hipStream_t st) {
  hipLaunchKernelGGL(k_zero16, dim3((unsigned)std::min<int64_t>(ceil_div(n16, 256), 8192)), dim3(256), 0, st,
                     (uint4*)p, n16);
  VH_LAUNCHED();
}

}  // namespace vh
