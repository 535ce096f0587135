// Post-processing of solutions: current coefficients on the voxel grid,
// post_obtain_curr_coefs_on_grid.m:1-45.
//
// A solution's first 5*nv entries are the currents Jx, Jy, Jz, J2d, J3d of the
// non-empty voxels in column-major voxel order.  They are scattered into five
// L x M x N column-major grids (zero at empty voxels), and min/max of |J| over
// the Jx, Jy and Jz blocks are returned (the plot colour range, :24-25).
//
// Both kernels stream HBM: the scatter reads 5*nv and writes 5*L*M*N complex
// values per solution (one thread per grid cell, writes coalesced, reads in
// voxel order because the voxel map is monotonic); the min/max reads 3*nv.
#include "vh_geom.h"
#include <algorithm>
#include <cmath>
#include <vector>

namespace vh {

namespace {

constexpr int PT = 256;

__global__ __launch_bounds__(PT) void k_curr_grid(const double2* __restrict__ x, int64_t ldx,
                                                  const int* __restrict__ vmap, int64_t ncell, int nv,
                                                  double2* __restrict__ grid, int64_t ldg) {
  const int64_t cell = (int64_t)blockIdx.x * PT + threadIdx.x;
  if (cell >= ncell) return;
  const int64_t r = blockIdx.y;
  const int v = vmap[cell];
  const double2* xr = x + r * ldx;
  double2* gr = grid + r * ldg;
  const double2 zero = make_double2(0.0, 0.0);
#pragma unroll
  for (int c = 0; c < 5; ++c) gr[c * ncell + cell] = v >= 0 ? xr[(int64_t)c * nv + v] : zero;
}

// part[(r*nb + b)] = (min, max) of |x| over the block's share of x[r][0:cnt]
__global__ __launch_bounds__(PT) void k_abs_minmax(const double2* __restrict__ x, int64_t ldx, int64_t cnt,
                                                   double2* __restrict__ part) {
  __shared__ double smin[PT], smax[PT];
  const int64_t r = blockIdx.y;
  const double2* xr = x + r * ldx;
  double lo = INFINITY, hi = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * PT + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * PT) {
    const double a = hypot(xr[i].x, xr[i].y);
    lo = fmin(lo, a);
    hi = fmax(hi, a);
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int h = PT / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + h]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[r * gridDim.x + blockIdx.x] = make_double2(smin[0], smax[0]);
}

}  // namespace

void currents_on_grid(Geometry& G, int nrhs, const double2* x, int64_t ldx, double2* grid, int64_t ldg,
                      double* minmax) {
  const int64_t ncell = (int64_t)G.g.L * G.g.M * G.g.N;
  const int nv = G.g.nv;
  VH_REQUIRE(ldx >= 5 * (int64_t)nv, "leading dimension of x smaller than 5*nv");
  VH_REQUIRE(ldg >= 5 * ncell, "leading dimension of the grid smaller than 5*L*M*N");
  VH_REQUIRE(nrhs >= 1 && nrhs < (1 << 16), "nrhs out of range");
  hipStream_t st = G.ctx->stream;
  hipLaunchKernelGGL(k_curr_grid, dim3((unsigned)ceil_div(ncell, PT), nrhs), dim3(PT), 0, st, x, ldx, G.vmap.p,
                     ncell, nv, grid, ldg);
  VH_LAUNCHED();
  if (!minmax) return;
  const int64_t cnt = 3 * (int64_t)nv;
  const int nb = (int)std::min<int64_t>(256, ceil_div(cnt, PT));
  DevBuf<double2> part((size_t)nb * nrhs);
  hipLaunchKernelGGL(k_abs_minmax, dim3(nb, nrhs), dim3(PT), 0, st, x, ldx, cnt, part.p);
  VH_LAUNCHED();
  std::vector<double2> h((size_t)nb * nrhs);
  VH_HIP(hipMemcpyAsync(h.data(), part.p, h.size() * sizeof(double2), hipMemcpyDeviceToHost, st));
  VH_HIP(hipStreamSynchronize(st));
  for (int r = 0; r < nrhs; ++r) {
    double lo = INFINITY, hi = 0.0;
    for (int b = 0; b < nb; ++b) {
      lo = std::min(lo, h[(size_t)r * nb + b].x);
      hi = std::max(hi, h[(size_t)r * nb + b].y);
    }
    minmax[2 * r] = lo;
    minmax[2 * r + 1] = hi;
  }
}

}  // namespace vh
