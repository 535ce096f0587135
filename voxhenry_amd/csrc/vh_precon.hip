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

// Supernodal solves with inverted diagonal blocks.  Block k owns rows and
// columns [c0, c1).  Forward (L y = b), blocks in increasing level order:
//   t_r = b_r - sum_{j < c0} L(r,j) y_j   (k_sn_update_fwd, rows of the level)
//   y_k = Linv_k t_k                      (k_sn_gemv_fwd)
// Backward (L' x = D^-1 y), blocks in decreasing level order:
//   u_r = y_r / D_r - sum_{i >= c1} L(i,r) x_i   (k_sn_update_bwd)
//   x_k = Linv_k' u_k                             (k_sn_gemv_bwd)
// w holds b, t and finally x; w2 holds y then u.  Vectors are [row][rhs].
template <typename R>
__global__ void k_sn_update_fwd(const int* __restrict__ rows, int nrows, const int* __restrict__ rp,
                                const int* __restrict__ split, const int* __restrict__ ri, const R* __restrict__ rx,
                                cplx<R>* __restrict__ w, const cplx<R>* __restrict__ w2, int nrhs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nrows * nrhs) return;
  const int q = (int)(t / nrhs), c = (int)(t - (int64_t)q * nrhs);
  const int r = rows[q];
  cplx<R> s = w[(int64_t)r * nrhs + c];
  for (int p = rp[r]; p < split[r]; ++p) {
    const R l = rx[p];
    const cplx<R> v = w2[(int64_t)ri[p] * nrhs + c];
    s.x -= l * v.x; s.y -= l * v.y;
  }
  w[(int64_t)r * nrhs + c] = s;
}

template <typename R>
__global__ void k_sn_gemv_fwd(const int* __restrict__ rows, int nrows, const int* __restrict__ sn_of_row,
                              const int* __restrict__ sn_ptr, const int64_t* __restrict__ sn_off,
                              const R* __restrict__ Li, const cplx<R>* __restrict__ w, cplx<R>* __restrict__ w2,
                              int nrhs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nrows * nrhs) return;
  const int q = (int)(t / nrhs), c = (int)(t - (int64_t)q * nrhs);
  const int r = rows[q];
  const int k = sn_of_row[r];
  const int c0 = sn_ptr[k];
  const int i = r - c0;
  const R* row = Li + sn_off[k] + (int64_t)i * (i + 1) / 2;
  const cplx<R>* tt = w + (int64_t)c0 * nrhs + c;
  R sx = 0, sy = 0;
  for (int j = 0; j <= i; ++j) {
    const R l = row[j];
    const cplx<R> v = tt[(int64_t)j * nrhs];
    sx += l * v.x; sy += l * v.y;
  }
  cplx<R> o; o.x = sx; o.y = sy;
  w2[(int64_t)r * nrhs + c] = o;
}

template <typename R>
__global__ void k_sn_update_bwd(const int* __restrict__ rows, int nrows, const int* __restrict__ cp,
                                const int* __restrict__ split, const int* __restrict__ ci, const R* __restrict__ cx,
                                const R* __restrict__ D, const cplx<R>* __restrict__ w, cplx<R>* __restrict__ w2,
                                int nrhs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nrows * nrhs) return;
  const int q = (int)(t / nrhs), c = (int)(t - (int64_t)q * nrhs);
  const int r = rows[q];
  cplx<R> s = w2[(int64_t)r * nrhs + c];
  const R inv = R(1) / D[r];
  s.x *= inv; s.y *= inv;
  for (int p = split[r]; p < cp[r + 1]; ++p) {
    const R l = cx[p];
    const cplx<R> v = w[(int64_t)ci[p] * nrhs + c];
    s.x -= l * v.x; s.y -= l * v.y;
  }
  w2[(int64_t)r * nrhs + c] = s;
}

template <typename R>
__global__ void k_sn_gemv_bwd(const int* __restrict__ rows, int nrows, const int* __restrict__ sn_of_row,
                              const int* __restrict__ sn_ptr, const int64_t* __restrict__ sn_off,
                              const R* __restrict__ LiT, const cplx<R>* __restrict__ w2, cplx<R>* __restrict__ w,
                              int nrhs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nrows * nrhs) return;
  const int q = (int)(t / nrhs), c = (int)(t - (int64_t)q * nrhs);
  const int r = rows[q];
  const int k = sn_of_row[r];
  const int c0 = sn_ptr[k], sz = sn_ptr[k + 1] - c0;
  const int i = r - c0;
  // row i of the transpose holds Linv(i..sz-1, i): offset i*sz - i(i-1)/2
  const R* row = LiT + sn_off[k] + (int64_t)i * sz - (int64_t)i * (i - 1) / 2;
  const cplx<R>* uu = w2 + (int64_t)r * nrhs + c;
  R sx = 0, sy = 0;
  for (int j = 0; j < sz - i; ++j) {
    const R l = row[j];
    const cplx<R> v = uu[(int64_t)j * nrhs];
    sx += l * v.x; sy += l * v.y;
  }
  cplx<R> o; o.x = sx; o.y = sy;
  w[(int64_t)r * nrhs + c] = o;
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
  std::vector<int> perm = geometric
      ? nd_ordering_geometric(nn, Y.Sp.data(), Y.Si.data(), xyz.data(), type.data(), &Y.blocks)
      : nd_ordering(nn, Y.Sp.data(), Y.Si.data(), &Y.blocks);
  ldl_symbolic(nn, Y.Sp.data(), Y.Si.data(), perm, Y.F);
  Y.nnzL = (int64_t)Y.F.Lp[nn];
}

// Rows of the blocks of each level, ordered by level.
static void level_rows(const std::vector<int>& lev, const std::vector<int>& sn_ptr, std::vector<int>& lptr,
                       std::vector<int>& rows) {
  const int nsn = (int)lev.size();
  int nl = 0;
  for (int l : lev) nl = std::max(nl, l + 1);
  lptr.assign(nl + 1, 0);
  for (int k = 0; k < nsn; ++k) lptr[lev[k] + 1] += sn_ptr[k + 1] - sn_ptr[k];
  for (int l = 0; l < nl; ++l) lptr[l + 1] += lptr[l];
  rows.assign(lptr[nl], 0);
  std::vector<int> fill(lptr.begin(), lptr.end() - 1);
  for (int k = 0; k < nsn; ++k)
    for (int r = sn_ptr[k]; r < sn_ptr[k + 1]; ++r) rows[fill[lev[k]]++] = r;
}

// CSR copy of the L pattern, CSC->CSR map, blocks, their levels and the
// packed offsets of the dense inverses (needs the row indices of L).
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
  // blocks
  const int nsn = (int)Y.blocks.size();
  Y.sn_ptr.assign(nsn + 1, 0);
  for (int k = 0; k < nsn; ++k) Y.sn_ptr[k + 1] = Y.sn_ptr[k] + Y.blocks[k];
  VH_REQUIRE(Y.sn_ptr[nsn] == n, "dissection blocks do not cover the Schur matrix");
  Y.sn_of_row.assign(n, 0);
  for (int k = 0; k < nsn; ++k)
    for (int r = Y.sn_ptr[k]; r < Y.sn_ptr[k + 1]; ++r) Y.sn_of_row[r] = k;
  Y.sn_off.assign(nsn + 1, 0);
  for (int k = 0; k < nsn; ++k) {
    const int64_t sz = Y.blocks[k];
    Y.sn_off[k + 1] = Y.sn_off[k] + sz * (sz + 1) / 2;
  }
  Y.dense = Y.sn_off[nsn];
  // splits: CSR row r entries with column < c0 come first; CSC column r entries
  // with row >= c1 come last (indices are sorted)
  Y.row_split.assign(n, 0);
  Y.col_split.assign(n, 0);
  for (int r = 0; r < n; ++r) {
    const int k = Y.sn_of_row[r];
    const int c0 = Y.sn_ptr[k], c1 = Y.sn_ptr[k + 1];
    int p = Rp[r];
    while (p < Rp[r + 1] && Ri[p] < c0) ++p;
    Y.row_split[r] = p;
    int q = F.Lp[r];
    while (q < F.Lp[r + 1] && F.Li[q] < c1) ++q;
    Y.col_split[r] = q;
  }
  // block levels: forward from the blocks a block's rows reach, backward from
  // the blocks its columns reach
  std::vector<int> lf(nsn, 0), lb(nsn, 0);
  for (int k = 0; k < nsn; ++k) {
    int l = 0;
    for (int r = Y.sn_ptr[k]; r < Y.sn_ptr[k + 1]; ++r)
      for (int p = Rp[r]; p < Y.row_split[r]; ++p) l = std::max(l, lf[Y.sn_of_row[Ri[p]]] + 1);
    lf[k] = l;
  }
  for (int k = nsn - 1; k >= 0; --k) {
    int l = 0;
    for (int r = Y.sn_ptr[k]; r < Y.sn_ptr[k + 1]; ++r)
      for (int q = Y.col_split[r]; q < F.Lp[r + 1]; ++q) l = std::max(l, lb[Y.sn_of_row[F.Li[q]]] + 1);
    lb[k] = l;
  }
  level_rows(lf, Y.sn_ptr, Y.f_lptr, Y.f_rows);
  level_rows(lb, Y.sn_ptr, Y.b_lptr, Y.b_rows);
  upload(Y.d_perm, F.perm);
  upload(Y.Rp, Rp);
  upload(Y.Ri, Ri);
  upload(Y.Cp, F.Lp);
  upload(Y.Ci, F.Li);
  upload(Y.d_sn_ptr, Y.sn_ptr);
  upload(Y.d_sn_of_row, Y.sn_of_row);
  upload(Y.d_row_split, Y.row_split);
  upload(Y.d_col_split, Y.col_split);
  upload(Y.d_f_rows, Y.f_rows);
  upload(Y.d_b_rows, Y.b_rows);
  upload(Y.d_sn_off, Y.sn_off);
  // Ar' with permuted node columns
  std::vector<int> tcp(G.atc_h.size());
  for (size_t q = 0; q < G.atc_h.size(); ++q) tcp[q] = F.iperm[G.atc_h[q]];
  upload(Y.ArTp_col, tcp);
}

// Dense inverses of the unit lower diagonal blocks of L (packed row-major
// lower, and the transposes packed row-major upper).
static void block_inverses(const PreconSymbolic& Y, const LdlFactor& F, std::vector<double>& Li,
                           std::vector<double>& LiT) {
  const int nsn = (int)Y.blocks.size();
  Li.assign(Y.dense, 0.0);
  LiT.assign(Y.dense, 0.0);
  std::vector<double> B, X;
  for (int k = 0; k < nsn; ++k) {
    const int c0 = Y.sn_ptr[k], sz = Y.blocks[k];
    B.assign((size_t)sz * sz, 0.0);
    for (int j = 0; j < sz; ++j) {
      B[(size_t)j * sz + j] = 1.0;
      for (int q = F.Lp[c0 + j]; q < Y.col_split[c0 + j]; ++q) B[(size_t)(F.Li[q] - c0) * sz + j] = F.Lx[q];
    }
    // X = B^-1 by forward substitution, column by column (B unit lower)
    X.assign((size_t)sz * sz, 0.0);
    for (int c = 0; c < sz; ++c) {
      X[(size_t)c * sz + c] = 1.0;
      for (int r = c + 1; r < sz; ++r) {
        double acc = 0.0;
        const double* br = &B[(size_t)r * sz];
        for (int m = c; m < r; ++m) acc += br[m] * X[(size_t)m * sz + c];
        X[(size_t)r * sz + c] = -acc;
      }
    }
    const int64_t off = Y.sn_off[k];
    for (int r = 0; r < sz; ++r)
      for (int c = 0; c <= r; ++c) Li[off + (int64_t)r * (r + 1) / 2 + c] = X[(size_t)r * sz + c];
    for (int c = 0; c < sz; ++c) {
      const int64_t base = off + (int64_t)c * sz - (int64_t)c * (c - 1) / 2;
      for (int r = c; r < sz; ++r) LiT[base + (r - c)] = X[(size_t)r * sz + c];
    }
  }
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
  std::vector<double> Li, LiT;
  block_inverses(Y, P.F, Li, LiT);
  upload(P.Rx64, Rx); upload_f32(P.Rx32, Rx);
  upload(P.Cx64, P.F.Lx); upload_f32(P.Cx32, P.F.Lx);
  upload(P.D64, P.F.D); upload_f32(P.D32, P.F.D);
  upload(P.ainv64, ainv); upload_f32(P.ainv32, ainv);
  upload(P.Li64, Li); upload_f32(P.Li32, Li);
  upload(P.LiT64, LiT); upload_f32(P.LiT32, LiT);
  // the values are on the device; the host copies of L are not needed any more
  P.F.Li.clear(); P.F.Li.shrink_to_fit();
  P.F.Lx.clear(); P.F.Lx.shrink_to_fit();
  P.ready = true;
}

template <typename R>
void precon_apply(Precon& P, int nrhs, const cplx<R>* x, int64_t ldx, cplx<R>* y, int64_t ldy, const int* slots) {
  VH_REQUIRE(P.ready, "preconditioner not prepared");
  Geometry& G = *P.G;
  hipStream_t st = G.ctx->stream;
  const int nn = (int)G.num_node;
  const int64_t nc = G.num_curr;
  const size_t need = 2 * (size_t)nn * nrhs * sizeof(cplx<R>);
  if (P.ws.n < need) P.ws.alloc(need);
  cplx<R>* w = (cplx<R>*)P.ws.p;
  cplx<R>* w2 = w + (size_t)nn * nrhs;
  const bool d = sizeof(R) == 8;
  const R* arv = d ? (const R*)G.Ar.v64.p : (const R*)G.Ar.v32.p;
  const R* ainv = d ? (const R*)P.ainv64.p : (const R*)P.ainv32.p;
  if (nn > 0) {
    PreconSymbolic& Y = *P.sym;
    const int64_t items = (int64_t)nn * nrhs;
    hipLaunchKernelGGL((k_pc_gather<R>), dim3(ceil_div(items, 256)), dim3(256), 0, st, x, ldx, w, nrhs, nn,
                       Y.d_perm.p, G.Ar.ptr.p, G.Ar.col.p, arv, ainv, nc, slots);
    VH_LAUNCHED();
    const R* Rx = d ? (const R*)P.Rx64.p : (const R*)P.Rx32.p;
    const R* Cx = d ? (const R*)P.Cx64.p : (const R*)P.Cx32.p;
    const R* D = d ? (const R*)P.D64.p : (const R*)P.D32.p;
    const R* Li = d ? (const R*)P.Li64.p : (const R*)P.Li32.p;
    const R* LiT = d ? (const R*)P.LiT64.p : (const R*)P.LiT32.p;
    const int nlf = (int)Y.f_lptr.size() - 1;
    for (int l = 0; l < nlf; ++l) {
      const int a = Y.f_lptr[l], nr = Y.f_lptr[l + 1] - a;
      const dim3 grid(ceil_div((int64_t)nr * nrhs, 256));
      if (l > 0) {
        hipLaunchKernelGGL((k_sn_update_fwd<R>), grid, dim3(256), 0, st, Y.d_f_rows.p + a, nr, Y.Rp.p,
                           Y.d_row_split.p, Y.Ri.p, Rx, w, w2, nrhs);
        VH_LAUNCHED();
      }
      hipLaunchKernelGGL((k_sn_gemv_fwd<R>), grid, dim3(256), 0, st, Y.d_f_rows.p + a, nr, Y.d_sn_of_row.p,
                         Y.d_sn_ptr.p, Y.d_sn_off.p, Li, w, w2, nrhs);
      VH_LAUNCHED();
    }
    const int nlb = (int)Y.b_lptr.size() - 1;
    for (int l = 0; l < nlb; ++l) {
      const int a = Y.b_lptr[l], nr = Y.b_lptr[l + 1] - a;
      const dim3 grid(ceil_div((int64_t)nr * nrhs, 256));
      hipLaunchKernelGGL((k_sn_update_bwd<R>), grid, dim3(256), 0, st, Y.d_b_rows.p + a, nr, Y.Cp.p,
                         Y.d_col_split.p, Y.Ci.p, Cx, D, w, w2, nrhs);
      VH_LAUNCHED();
      hipLaunchKernelGGL((k_sn_gemv_bwd<R>), grid, dim3(256), 0, st, Y.d_b_rows.p + a, nr, Y.d_sn_of_row.p,
                         Y.d_sn_ptr.p, Y.d_sn_off.p, LiT, w2, w, nrhs);
      VH_LAUNCHED();
    }
    hipLaunchKernelGGL((k_pc_scatter_nodes<R>), dim3(ceil_div(items, 256)), dim3(256), 0, st, w, y, ldy, nrhs, nn,
                       Y.d_perm.p, nc, slots);
    VH_LAUNCHED();
  }
  const R* atv = d ? (const R*)G.ArT.v64.p : (const R*)G.ArT.v32.p;
  hipLaunchKernelGGL((k_pc_currents<R>), dim3(ceil_div(nc, 256), nrhs), dim3(256), 0, st, x, ldx, w, y, ldy, nrhs, nc,
                     G.ArT.ptr.p, nn > 0 ? P.sym->ArTp_col.p : nullptr, atv, ainv, slots);
  VH_LAUNCHED();
}

template void precon_apply<double>(Precon&, int, const double2*, int64_t, double2*, int64_t, const int*);
template void precon_apply<float>(Precon&, int, const float2*, int64_t, float2*, int64_t, const int*);

}  // namespace vh
