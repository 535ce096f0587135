// Schur-complement preconditioner on the device.
//
// Replaces lse_sparse_precon_prepare.m:33-104,149-159 (A_inv, S, ldlchol) and
// lse_sparse_precon_multiply.m:84-88 (apply):
//     bb = b - Ar*(A_inv .* a);   y(q) = LD \ bb(q);   x = (a + Ar'*y) .* A_inv
// S = Ar*diag(A_inv)*Ar' is factored on the host (vh_ldl.cpp); the permuted
// L, L' and D go to the device.  The solves run level by level: a level with
// many rows is one grid launch (one thread per row and right-hand side); a run
// of consecutive narrow levels is one launch in which each workgroup owns a
// group of right-hand sides and walks the levels with a barrier in between.
// The work vector is stored node-major, right-hand side fastest ([node][rhs]),
// so a level touches contiguous memory for all right-hand sides of a row.
#include "vh_solver.h"
#include <algorithm>
#include <cmath>

namespace vh {

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// w[i][r] = b_r[perm[i]] - sum_k Ar(perm[i],k) * Ainv[k] * a_r[k]
template <typename R>
__global__ void k_pc_gather(const cplx<R>* __restrict__ x, int64_t ldx, cplx<R>* __restrict__ w, int nrhs, int nn,
                            const int* __restrict__ perm, const int* __restrict__ arp, const int* __restrict__ arc,
                            const R* __restrict__ arv, const R* __restrict__ ainv, int64_t nc,
                            const int* __restrict__ slots) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nn * nrhs) return;
  const int i = (int)(t / nrhs), r = (int)(t - (int64_t)i * nrhs);
  const int node = perm[i];
  const cplx<R>* xr = x + (int64_t)(slots ? slots[r] : r) * ldx;
  cplx<R> s = xr[nc + node];
  for (int p = arp[node]; p < arp[node + 1]; ++p) {
    const int k = arc[p];
    const R c = arv[p] * ainv[k];
    const cplx<R> a = xr[k];
    s.x -= c * a.x; s.y -= c * a.y;
  }
  w[t] = s;
}

// One wide level: rows[0..nrows), thread per (row, rhs).
// forward (BWD = 0): w_i -= sum_j L(i,j) w_j            (CSR of L)
// backward (BWD = 1): w_i = w_i / D_i - sum_k L(k,i) w_k (CSC of L = CSR of L')
template <typename R, int BWD>
__global__ void k_trsv_level(const int* __restrict__ rows, int nrows, const int* __restrict__ ptr,
                             const int* __restrict__ idx, const R* __restrict__ val, const R* __restrict__ D,
                             cplx<R>* __restrict__ w, int nrhs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nrows * nrhs) return;
  const int q = (int)(t / nrhs), r = (int)(t - (int64_t)q * nrhs);
  const int i = rows[q];
  cplx<R> s = w[(int64_t)i * nrhs + r];
  if (BWD) { const R inv = R(1) / D[i]; s.x *= inv; s.y *= inv; }
  for (int p = ptr[i]; p < ptr[i + 1]; ++p) {
    const R l = val[p];
    const cplx<R> v = w[(int64_t)idx[p] * nrhs + r];
    s.x -= l * v.x; s.y -= l * v.y;
  }
  w[(int64_t)i * nrhs + r] = s;
}

// A run of narrow levels [l0, l1): block b owns right-hand sides
// [b*rpb, min(nrhs,(b+1)*rpb)); one wave per (row, rhs) item with the row's
// entries split over the 64 lanes.
template <typename R, int BWD>
__global__ __launch_bounds__(256) void k_trsv_chain(const int* __restrict__ lptr, const int* __restrict__ rows, int l0,
                                                    int l1, const int* __restrict__ ptr, const int* __restrict__ idx,
                                                    const R* __restrict__ val, const R* __restrict__ D,
                                                    cplx<R>* __restrict__ w, int nrhs, int rpb) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const int r0 = blockIdx.x * rpb;
  const int nr = min(rpb, nrhs - r0);
  for (int l = l0; l < l1; ++l) {
    const int a = lptr[l], b = lptr[l + 1];
    const int items = (b - a) * nr;
    for (int it = wave; it < items; it += nwave) {
      const int q = it / nr, r = r0 + (it - q * nr);
      const int i = rows[a + q];
      R sx = 0, sy = 0;
      for (int p = ptr[i] + lane; p < ptr[i + 1]; p += 64) {
        const R lv = val[p];
        const cplx<R> v = w[(int64_t)idx[p] * nrhs + r];
        sx += lv * v.x; sy += lv * v.y;
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        sx += __shfl_xor(sx, o, 64);
        sy += __shfl_xor(sy, o, 64);
      }
      if (lane == 0) {
        cplx<R> s = w[(int64_t)i * nrhs + r];
        if (BWD) { const R inv = R(1) / D[i]; s.x *= inv; s.y *= inv; }
        s.x -= sx; s.y -= sy;
        w[(int64_t)i * nrhs + r] = s;
      }
    }
    __syncthreads();
  }
}

// y_r[nc + perm[i]] = w[i][r]
template <typename R>
__global__ void k_pc_scatter_nodes(const cplx<R>* __restrict__ w, cplx<R>* __restrict__ y, int64_t ldy, int nrhs,
                                   int nn, const int* __restrict__ perm, int64_t nc, const int* __restrict__ slots) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nn * nrhs) return;
  const int i = (int)(t / nrhs), r = (int)(t - (int64_t)i * nrhs);
  y[(int64_t)(slots ? slots[r] : r) * ldy + nc + perm[i]] = w[t];
}

// y_r[k] = (a_r[k] + sum_i ArTp(k,i) w[i][r]) * Ainv[k]   (ArTp columns are permuted node ids)
template <typename R>
__global__ void k_pc_currents(const cplx<R>* __restrict__ x, int64_t ldx, const cplx<R>* __restrict__ w,
                              cplx<R>* __restrict__ y, int64_t ldy, int nrhs, int64_t nc, const int* __restrict__ tp,
                              const int* __restrict__ tc, const R* __restrict__ tv, const R* __restrict__ ainv,
                              const int* __restrict__ slots) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  const int slot = slots ? slots[r] : r;
  if (k >= nc) return;
  cplx<R> s = x[(int64_t)slot * ldx + k];
  for (int p = tp[k]; p < tp[k + 1]; ++p) {
    const cplx<R> v = w[(int64_t)tc[p] * nrhs + r];
    s.x += tv[p] * v.x; s.y += tv[p] * v.y;
  }
  const R ai = ainv[k];
  s.x *= ai; s.y *= ai;
  y[(int64_t)slot * ldy + k] = s;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <typename T> static void upload(DevBuf<T>& d, const std::vector<T>& h) {
  d.alloc(std::max<size_t>(h.size(), 1));
  if (!h.empty()) VH_HIP(hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}
template <typename T> static void upload_f32(DevBuf<float>& d, const std::vector<T>& h) {
  std::vector<float> f(h.begin(), h.end());
  upload(d, f);
}

static void build_schedule(const std::vector<int>& lev, std::vector<int>& lptr, std::vector<int>& rows) {
  const int n = (int)lev.size();
  int nl = 0;
  for (int l : lev) nl = std::max(nl, l + 1);
  lptr.assign(nl + 1, 0);
  for (int l : lev) lptr[l + 1]++;
  for (int l = 0; l < nl; ++l) lptr[l + 1] += lptr[l];
  rows.assign(n, 0);
  std::vector<int> fill(lptr.begin(), lptr.end() - 1);
  for (int i = 0; i < n; ++i) rows[fill[lev[i]]++] = i;
}

static void build_symbolic(Geometry& G, bool geometric, PreconSymbolic& Y) {
  const int nv = G.g.nv;
  const int nn = (int)G.num_node;
  const auto& arp = G.arp_h; const auto& arc = G.arc_h;
  const auto& atp = G.atp_h; const auto& atc = G.atc_h; const auto& atv = G.atv_h;
  std::vector<int> Sp(nn + 1, 0), Si;
  std::vector<int> mark(nn, -1);
  Si.reserve((size_t)nn * 16);
  for (int i = 0; i < nn; ++i) {
    for (int p = arp[i]; p < arp[i + 1]; ++p) {
      const int k = arc[p];
      for (int q = atp[k]; q < atp[k + 1]; ++q) {
        const int j = atc[q];
        if (mark[j] != i) { mark[j] = i; Si.push_back(j); }
      }
    }
    std::sort(Si.begin() + Sp[i], Si.end());
    Sp[i + 1] = (int)Si.size();
  }
  Y.Sp = std::move(Sp);
  Y.Si = std::move(Si);
  // face coordinates of every node from the Jx/Jy/Jz columns of its voxels:
  // the -1 entry is the face at the voxel's lower side, +1 the upper side
  std::vector<float> xyz(3 * (size_t)nn, 0.f);
  std::vector<int> type(nn, -1);
  const int L = G.g.L, M = G.g.M;
  for (int v = 0; v < nv; ++v) {
    const int64_t lin = G.vox[v];
    const float c[3] = {(float)(lin % L), (float)((lin / L) % M), (float)(lin / ((int64_t)L * M))};
    for (int d = 0; d < 3; ++d) {
      const int64_t row = (int64_t)d * nv + v;
      for (int q = atp[row]; q < atp[row + 1]; ++q) {
        const int node = atc[q];
        float* pxyz = &xyz[3 * (size_t)node];
        for (int e = 0; e < 3; ++e) pxyz[e] = c[e] + 0.5f;
        pxyz[d] = c[d] + (atv[q] > 0 ? 1.f : 0.f);
        type[node] = d;
      }
    }
  }
  for (int i = 0; i < nn && geometric; ++i) geometric = type[i] >= 0;
  std::vector<int> perm = geometric ? nd_ordering_geometric(nn, Y.Sp.data(), Y.Si.data(), xyz.data(), type.data())
                                    : nd_ordering(nn, Y.Sp.data(), Y.Si.data());
  ldl_symbolic(nn, Y.Sp.data(), Y.Si.data(), perm, Y.F);
  Y.nnzL = (int64_t)Y.F.Lp[nn];
}

// CSR copy of the L pattern, CSC->CSR position map and level schedules
// (needs the row indices of L, i.e. one numeric factorisation).
static void build_levels(Geometry& G, const LdlFactor& F, PreconSymbolic& Y) {
  const int n = F.n;
  std::vector<int> Rp(n + 1, 0);
  for (int j = 0; j < n; ++j)
    for (int p = F.Lp[j]; p < F.Lp[j + 1]; ++p) Rp[F.Li[p] + 1]++;
  for (int i = 0; i < n; ++i) Rp[i + 1] += Rp[i];
  std::vector<int> Ri(Rp[n]);
  Y.csc2csr.assign(Rp[n], 0);
  std::vector<int> fill(Rp.begin(), Rp.end() - 1);
  for (int j = 0; j < n; ++j)
    for (int p = F.Lp[j]; p < F.Lp[j + 1]; ++p) {
      const int i = F.Li[p];
      Ri[fill[i]] = j;
      Y.csc2csr[p] = fill[i]++;
    }
  std::vector<int> lev_f(n, 0), lev_b(n, 0);
  for (int i = 0; i < n; ++i) {
    int l = 0;
    for (int p = Rp[i]; p < Rp[i + 1]; ++p) l = std::max(l, lev_f[Ri[p]] + 1);
    lev_f[i] = l;
  }
  for (int i = n - 1; i >= 0; --i) {
    int l = 0;
    for (int p = F.Lp[i]; p < F.Lp[i + 1]; ++p) l = std::max(l, lev_b[F.Li[p]] + 1);
    lev_b[i] = l;
  }
  build_schedule(lev_f, Y.fl_ptr, Y.fl_rows);
  build_schedule(lev_b, Y.bl_ptr, Y.bl_rows);
  upload(Y.d_fl_rows, Y.fl_rows);
  upload(Y.d_bl_rows, Y.bl_rows);
  upload(Y.d_fl_ptr, Y.fl_ptr);
  upload(Y.d_bl_ptr, Y.bl_ptr);
  upload(Y.d_perm, F.perm);
  upload(Y.Rp, Rp);
  upload(Y.Ri, Ri);
  upload(Y.Cp, F.Lp);
  upload(Y.Ci, F.Li);
  // Ar' with permuted node columns
  std::vector<int> tcp(G.atc_h.size());
  for (size_t q = 0; q < G.atc_h.size(); ++q) tcp[q] = F.iperm[G.atc_h[q]];
  upload(Y.ArTp_col, tcp);
}

void precon_prepare(Precon& P, double jre, double jim) {
  Geometry& G = *P.G;
  VH_REQUIRE(G.has_inc && G.has_z, "preconditioner: incidence and impedance must be set first");
  const int nv = G.g.nv;
  const int64_t nc = G.num_curr;
  const int nn = (int)G.num_node;
  // A_inv (lse_sparse_precon_prepare.m:36-62): 1 / |z/(s dx) + jwmu g_self|
  const std::complex<double> jw(jre, jim);
  std::complex<double> st[3];
  for (int k = 0; k < 3; ++k) st[k] = jw * std::complex<double>(G.st2[2 * k], G.st2[2 * k + 1]);
  std::vector<double> ainv(nc);
  for (int v = 0; v < nv; ++v) {
    const std::complex<double> z = G.zh[v];
    const double dp = 1.0 / std::abs(z / G.dx + st[0]);
    const double d2 = 1.0 / std::abs(z / (6.0 * G.dx) + st[1]);
    const double d3 = 1.0 / std::abs(z / (2.0 * G.dx) + st[2]);
    ainv[v] = dp; ainv[nv + v] = dp; ainv[2 * nv + v] = dp;
    ainv[3 * nv + v] = d2; ainv[4 * nv + v] = d3;
  }
  P.ainv_h = ainv;
  bool fresh = false;
  if (!G.precon_sym) {
    G.precon_sym = std::make_shared<PreconSymbolic>();
    build_symbolic(G, P.geometric_order, *G.precon_sym);
    fresh = true;
  }
  P.sym = G.precon_sym;
  PreconSymbolic& Y = *P.sym;
  // S = Ar diag(A_inv) Ar' (:102) on the shared pattern
  const auto& arp = G.arp_h; const auto& arc = G.arc_h; const auto& arv = G.arv_h;
  const auto& atp = G.atp_h; const auto& atc = G.atc_h; const auto& atv = G.atv_h;
  std::vector<double> Sx(Y.Si.size(), 0.0), acc(nn, 0.0);
  for (int i = 0; i < nn; ++i) {
    for (int p = arp[i]; p < arp[i + 1]; ++p) {
      const int k = arc[p];
      const double a = arv[p] * ainv[k];
      for (int q = atp[k]; q < atp[k + 1]; ++q) acc[atc[q]] += a * atv[q];
    }
    for (int p = Y.Sp[i]; p < Y.Sp[i + 1]; ++p) { Sx[p] = acc[Y.Si[p]]; acc[Y.Si[p]] = 0.0; }
  }
  // numeric factorisation on a private copy of the symbolic data
  P.F.n = Y.F.n;
  P.F.perm = Y.F.perm; P.F.iperm = Y.F.iperm; P.F.parent = Y.F.parent; P.F.Lp = Y.F.Lp;
  P.F.Li.assign(Y.nnzL, 0);
  P.F.Lx.assign(Y.nnzL, 0.0);
  P.F.D.assign(nn, 0.0);
  try {
    ldl_numeric(nn, Y.Sp.data(), Y.Si.data(), Sx.data(), P.F);
  } catch (const std::runtime_error& e) {
    throw Error(ERR_NUMERIC, e.what());
  }
  if (fresh) {
    Y.F.Li = P.F.Li;
    build_levels(G, P.F, Y);
  }
  std::vector<double> Rx(Y.nnzL);
  for (int64_t p = 0; p < Y.nnzL; ++p) Rx[Y.csc2csr[p]] = P.F.Lx[p];
  upload(P.Rx64, Rx); upload_f32(P.Rx32, Rx);
  upload(P.Cx64, P.F.Lx); upload_f32(P.Cx32, P.F.Lx);
  upload(P.D64, P.F.D); upload_f32(P.D32, P.F.D);
  upload(P.ainv64, ainv); upload_f32(P.ainv32, ainv);
  // the values are on the device; the host copies of L are not needed any more
  P.F.Li.clear(); P.F.Li.shrink_to_fit();
  P.F.Lx.clear(); P.F.Lx.shrink_to_fit();
  P.ready = true;
}

template <typename R>
static void run_levels(Precon& P, const std::vector<int>& lptr, const DevBuf<int>& dptr, const DevBuf<int>& drows,
                       const int* ptr, const int* idx, const R* val, const R* D, cplx<R>* w, int nrhs, bool bwd,
                       hipStream_t st) {
  const int nl = (int)lptr.size() - 1;
  const int64_t wide = P.wide_items;
  int l = 0;
  while (l < nl) {
    const int64_t items = (int64_t)(lptr[l + 1] - lptr[l]) * nrhs;
    if (items >= wide) {
      const int64_t nb = ceil_div(items, 256);
      if (bwd)
        hipLaunchKernelGGL((k_trsv_level<R, 1>), dim3(nb), dim3(256), 0, st, drows.p + lptr[l], lptr[l + 1] - lptr[l],
                           ptr, idx, val, D, w, nrhs);
      else
        hipLaunchKernelGGL((k_trsv_level<R, 0>), dim3(nb), dim3(256), 0, st, drows.p + lptr[l], lptr[l + 1] - lptr[l],
                           ptr, idx, val, D, w, nrhs);
      VH_LAUNCHED();
      ++l;
    } else {
      int l1 = l + 1;
      while (l1 < nl && (int64_t)(lptr[l1 + 1] - lptr[l1]) * nrhs < wide) ++l1;
      const int rpb = std::min(nrhs, P.chain_rpb);
      const int nb = (int)ceil_div(nrhs, rpb);
      if (bwd)
        hipLaunchKernelGGL((k_trsv_chain<R, 1>), dim3(nb), dim3(256), 0, st, dptr.p, drows.p, l, l1, ptr, idx, val, D,
                           w, nrhs, rpb);
      else
        hipLaunchKernelGGL((k_trsv_chain<R, 0>), dim3(nb), dim3(256), 0, st, dptr.p, drows.p, l, l1, ptr, idx, val, D,
                           w, nrhs, rpb);
      VH_LAUNCHED();
      l = l1;
    }
  }
}

template <typename R>
void precon_apply(Precon& P, int nrhs, const cplx<R>* x, int64_t ldx, cplx<R>* y, int64_t ldy, const int* slots) {
  VH_REQUIRE(P.ready, "preconditioner not prepared");
  Geometry& G = *P.G;
  hipStream_t st = G.ctx->stream;
  const int nn = (int)G.num_node;
  const int64_t nc = G.num_curr;
  const size_t need = (size_t)nn * nrhs * sizeof(cplx<R>);
  if (P.ws.n < need) P.ws.alloc(need);
  cplx<R>* w = (cplx<R>*)P.ws.p;
  const bool d = sizeof(R) == 8;
  const R* arv = d ? (const R*)G.Ar.v64.p : (const R*)G.Ar.v32.p;
  const R* ainv = d ? (const R*)P.ainv64.p : (const R*)P.ainv32.p;
  if (nn > 0) {
    const int64_t items = (int64_t)nn * nrhs;
    hipLaunchKernelGGL((k_pc_gather<R>), dim3(ceil_div(items, 256)), dim3(256), 0, st, x, ldx, w, nrhs, nn,
                       P.sym->d_perm.p, G.Ar.ptr.p, G.Ar.col.p, arv, ainv, nc, slots);
    VH_LAUNCHED();
    const R* Rx = d ? (const R*)P.Rx64.p : (const R*)P.Rx32.p;
    const R* Cx = d ? (const R*)P.Cx64.p : (const R*)P.Cx32.p;
    const R* D = d ? (const R*)P.D64.p : (const R*)P.D32.p;
    PreconSymbolic& Y = *P.sym;
    run_levels<R>(P, Y.fl_ptr, Y.d_fl_ptr, Y.d_fl_rows, Y.Rp.p, Y.Ri.p, Rx, D, w, nrhs, false, st);
    run_levels<R>(P, Y.bl_ptr, Y.d_bl_ptr, Y.d_bl_rows, Y.Cp.p, Y.Ci.p, Cx, D, w, nrhs, true, st);
    hipLaunchKernelGGL((k_pc_scatter_nodes<R>), dim3(ceil_div(items, 256)), dim3(256), 0, st, w, y, ldy, nrhs, nn,
                       P.sym->d_perm.p, nc, slots);
    VH_LAUNCHED();
  }
  const R* atv = d ? (const R*)G.ArT.v64.p : (const R*)G.ArT.v32.p;
  hipLaunchKernelGGL((k_pc_currents<R>), dim3(ceil_div(nc, 256), nrhs), dim3(256), 0, st, x, ldx, w, y, ldy, nrhs, nc,
                     G.ArT.ptr.p, P.sym->ArTp_col.p, atv, ainv, slots);
  VH_LAUNCHED();
}

template void precon_apply<double>(Precon&, int, const double2*, int64_t, double2*, int64_t, const int*);
template void precon_apply<float>(Precon&, int, const float2*, int64_t, float2*, int64_t, const int*);

}  // namespace vh
