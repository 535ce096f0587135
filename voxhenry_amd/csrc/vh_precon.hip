// Schur-complement preconditioner on the device.
//
// Replaces lse_sparse_precon_prepare.m:33-104,149-159 (A_inv, S, ldlchol) and
// lse_sparse_precon_multiply.m:84-88 (apply):
//     bb = b - Ar*(A_inv .* a);   y(q) = LD \ bb(q);   x = (a + Ar'*y) .* A_inv
// S = Ar*diag(A_inv)*Ar' is factored on the host (vh_ldl.cpp) after a nested
// dissection ordering whose separator blocks become dense supernodes; L (CSR
// and CSC copies), D and the inverses of the diagonal blocks go to the device.
// The triangular solves run level by level over the block dependency levels,
// every column of the batch in the same launches: columns may belong to
// different frequencies (different factor values, same pattern) and are
// grouped by frequency; work vectors are [group][node][port].  Blocks whose
// column union fits the LDS go through the per-block kernel (update and
// inverse product in one workgroup), the rest through the row kernels.
#include "vh_solver.h"
#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <thread>
#include <type_traits>

namespace vh {

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// A batch is applied in groups: group g holds up to NPG columns (systems) that
// share one preconditioner (one frequency), tab[gtab[g]]; its column j is the
// vector in slot gslot[g*NPG + j] (-1: an unused position, computed as zero).
// Work vectors are [group][node][port] (index (g*nn + i)*NPG + j), so the
// ports of a group at one node are contiguous: one (row, group) item reads
// each entry of L and its index once for all ports of the group, and one
// line of the work vector per entry.  Work vectors are FP64 with FP64
// accumulation in both precisions: with FP32 factors the apply is then a
// fixed linear operator up to the rounding of its output (FP32 work vectors
// add rounding noise, amplified by the conditioning of L at every level,
// which stalls an FP32 Krylov solve).
struct PcGroupDev {
  int ng;
  const int* gtab;    // ng
  const int* gslot;   // ng * NPG
};

// Work-vector lines of one node: the NPG ports of a node are contiguous, so
// a line moves as 16-byte vector accesses.
template <typename WT, int NPG> __device__ __forceinline__ void line_store(cplx<WT>* p, const cplx<WT> (&v)[NPG]) {
#pragma unroll
  for (int j = 0; j < NPG; ++j) p[j] = v[j];
}

// w[g][iperm[v]][j] = b[v] - sum_k Ar(v,k) * Ainv[k] * a[k] of column (g, j).
// One thread per (node v, group g) serves all NPG ports: the
// row of Ar and the Ainv factors are loaded once, then every port's currents
// (a wave reads 64 consecutive nodes' currents of one port, which lie in a
// few contiguous runs) are in flight together, and the node's line of the
// work vector is written whole.
template <typename R, typename WT, int NPG>
__global__ __launch_bounds__(256) void k_pc_gather(const cplx<R>* __restrict__ x, int64_t ldx, cplx<WT>* __restrict__ w,
                                                   int nn, const int* __restrict__ iperm, const int* __restrict__ arp,
                                                   const int* __restrict__ arc, const R* __restrict__ arv,
                                                   const PcTab<R>* __restrict__ tab, PcGroupDev gr, int64_t nc) {
  // one-dimensional grid of node blocks x groups in XCD order (the currents
  // of consecutive y- and z-type nodes lie a grid row or plane apart: the
  // lines a wave touches are used again by the next node blocks, which run
  // in the same L2)
  const int nbx = (nn + 255) >> 8;
  const int lt = xcd_tile(blockIdx.x, gridDim.x);
  const int g = lt / nbx;
  const int v = (lt - g * nbx) * 256 + threadIdx.x;
  if (v >= nn) return;
  const R* __restrict__ ainv = tab[gr.gtab[g]].ainv;
  const int p0 = arp[v], p1 = arp[v + 1];
  // a node touches at most 6 currents: a short row is loaded all at once
  constexpr int D = 6;
  int kk[D];
  double cf[D];
  const bool short_row = p1 - p0 <= D;
  if (short_row) {
#pragma unroll
    for (int u = 0; u < D; ++u) kk[u] = p0 + u < p1 ? arc[p0 + u] : 0;
#pragma unroll
    for (int u = 0; u < D; ++u) cf[u] = p0 + u < p1 ? (double)arv[p0 + u] * (double)ainv[kk[u]] : 0.0;
  }
  cplx<WT> o[NPG];
  // ports in chunks of four: each chunk has all its loads in flight at once
  constexpr int PC = NPG < 4 ? NPG : 4;
#pragma unroll
  for (int j0 = 0; j0 < NPG; j0 += PC) {
    cplx<R> b[PC], a[PC][D];
    int sl[PC];
#pragma unroll
    for (int jj = 0; jj < PC; ++jj) {
      sl[jj] = gr.gslot[g * NPG + j0 + jj];
      const cplx<R>* xr = x + (int64_t)(sl[jj] >= 0 ? sl[jj] : 0) * ldx;
      if (sl[jj] >= 0) {
        b[jj] = xr[nc + v];
        if (short_row) {
#pragma unroll
          for (int u = 0; u < D; ++u)
            if (p0 + u < p1) a[jj][u] = xr[kk[u]];
        }
      }
    }
#pragma unroll
    for (int jj = 0; jj < PC; ++jj) {
      cplx<WT> r; r.x = 0; r.y = 0;
      if (sl[jj] >= 0) {
        double sx = b[jj].x, sy = b[jj].y;
        if (short_row) {
#pragma unroll
          for (int u = 0; u < D; ++u)
            if (p0 + u < p1) { sx -= cf[u] * a[jj][u].x; sy -= cf[u] * a[jj][u].y; }
        } else {
          const cplx<R>* xr = x + (int64_t)sl[jj] * ldx;
          for (int p = p0; p < p1; ++p) {
            const int k = arc[p];
            const double c = (double)arv[p] * (double)ainv[k];
            const cplx<R> av = xr[k];
            sx -= c * av.x; sy -= c * av.y;
          }
        }
        r.x = (WT)sx; r.y = (WT)sy;
      }
      o[j0 + jj] = r;
    }
  }
  line_store<WT, NPG>(w + ((int64_t)g * nn + iperm[v]) * NPG, o);
}

// Lanes of the level kernels: an item is (row q of the level, group g),
// item = q * ng + g (the groups of one row are neighbours: they share the
// index reads), served by NPG * LPR consecutive lanes; lane -> port
// j = lane mod NPG and part (lane div NPG) mod LPR.  The parts stride over
// the row's entries and are summed with butterflies inside the item's lanes.
template <int NPG, int LPR> struct PLanes {
  static_assert(NPG * LPR <= 64, "an item must fit in a wave");
  int64_t item;
  int j, part;
  __device__ __forceinline__ PLanes() {
    const int lane = threadIdx.x & 63;
    item = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / (NPG * LPR);
    j = lane % NPG;
    part = (lane / NPG) % LPR;
  }
  __device__ __forceinline__ double2 sum(double2 v) const {
#pragma unroll
    for (int o = NPG; o < NPG * LPR; o <<= 1) { v.x += __shfl_xor(v.x, o, 64); v.y += __shfl_xor(v.y, o, 64); }
    return v;
  }
};

// sum over p = p0, p0 + LPR, ... < p1 of val[p] * wg[idx[p] * NPG]: the entries
// of four strides are loaded before their lines, so a lane has four
// independent line loads in flight (rows of the top separators hold
// thousands of entries).
template <typename WT, int NPG, int LPR, typename R>
__device__ __forceinline__ double2 strided_dot(const R* __restrict__ val, const int* __restrict__ idx,
                                               const cplx<WT>* __restrict__ wg, int p0, int p1) {
  double2 s = make_double2(0.0, 0.0);
  int p = p0;
  for (; p + 3 * LPR < p1; p += 4 * LPR) {
    int i[4];
    double l[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) { i[u] = idx[p + u * LPR]; l[u] = (double)val[p + u * LPR]; }
    cplx<WT> v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = wg[(int64_t)i[u] * NPG];
#pragma unroll
    for (int u = 0; u < 4; ++u) { s.x += l[u] * v[u].x; s.y += l[u] * v[u].y; }
  }
  for (; p < p1; p += LPR) {
    const double l = (double)val[p];
    const cplx<WT> v = wg[(int64_t)idx[p] * NPG];
    s.x += l * v.x; s.y += l * v.y;
  }
  return s;
}

// Supernodal solves with inverted diagonal blocks.  Block k owns rows and
// columns [c0, c1).  Forward (L y = b), blocks in increasing level order:
//   t_r = b_r - sum_{j < c0} L(r,j) y_j   (k_sn_update_fwd, rows of the level)
//   y_k = Linv_k t_k                      (k_sn_gemv_fwd)
// Backward (L' x = D^-1 y), blocks in decreasing level order:
//   u_r = y_r / D_r - sum_{i >= c1} L(i,r) x_i   (k_sn_update_bwd)
//   x_k = Linv_k' u_k                             (k_sn_gemv_bwd)
// w holds b, t and finally x; w2 holds y then u.
template <typename R, typename WT, int NPG, int LPR>
__global__ __launch_bounds__(256) void k_sn_update_fwd(const int* __restrict__ rows, int nrows,
                                                       const int* __restrict__ fp, const int* __restrict__ fi,
                                                       const PcTab<R>* __restrict__ tab,
                                                       PcGroupDev gr, int nn, cplx<WT>* __restrict__ w,
                                                       const cplx<WT>* __restrict__ w2) {
  const PLanes<NPG, LPR> ln;
  if (ln.item >= (int64_t)nrows * gr.ng) return;   // all lanes of the item
  const int q = (int)(ln.item / gr.ng), g = (int)(ln.item - (int64_t)q * gr.ng);
  const int r = rows[q];
  const R* fx = tab[gr.gtab[g]].Fx;
  const cplx<WT>* wg = w2 + (int64_t)g * nn * NPG + ln.j;
  double2 s = strided_dot<WT, NPG, LPR>(fx, fi, wg, fp[r] + ln.part, fp[r + 1]);
  s = ln.sum(s);
  if (ln.part == 0) {
    cplx<WT>& o = w[((int64_t)g * nn + r) * NPG + ln.j];
    o.x = (WT)((double)o.x - s.x); o.y = (WT)((double)o.y - s.y);
  }
}

template <typename R, typename WT, int NPG, int LPR>
__global__ __launch_bounds__(256) void k_sn_gemv_fwd(const int* __restrict__ rows, int nrows,
                                                     const int* __restrict__ sn_of_row, const int* __restrict__ sn_ptr,
                                                     const int64_t* __restrict__ sn_off,
                                                     const PcTab<R>* __restrict__ tab, PcGroupDev gr, int nn,
                                                     const cplx<WT>* __restrict__ w, cplx<WT>* __restrict__ w2) {
  const PLanes<NPG, LPR> ln;
  if (ln.item >= (int64_t)nrows * gr.ng) return;
  const int q = (int)(ln.item / gr.ng), g = (int)(ln.item - (int64_t)q * gr.ng);
  const int r = rows[q];
  const int k = sn_of_row[r];
  const int c0 = sn_ptr[k];
  const int i = r - c0;
  const R* row = tab[gr.gtab[g]].Li + sn_off[k] + (int64_t)i * (i + 1) / 2;
  const cplx<WT>* tt = w + ((int64_t)g * nn + c0) * NPG + ln.j;
  double2 s; s.x = 0; s.y = 0;
  for (int jj = ln.part; jj <= i; jj += LPR) {
    const double l = row[jj];
    const cplx<WT> v = tt[(int64_t)jj * NPG];
    s.x += l * v.x; s.y += l * v.y;
  }
  s = ln.sum(s);
  if (ln.part == 0) {
    cplx<WT> o; o.x = (WT)s.x; o.y = (WT)s.y;
    w2[((int64_t)g * nn + r) * NPG + ln.j] = o;
  }
}

template <typename R, typename WT, int NPG, int LPR>
__global__ __launch_bounds__(256) void k_sn_update_bwd(const int* __restrict__ rows, int nrows,
                                                       const int* __restrict__ bp, const int* __restrict__ bi,
                                                       const PcTab<R>* __restrict__ tab,
                                                       PcGroupDev gr, int nn, const cplx<WT>* __restrict__ w,
                                                       cplx<WT>* __restrict__ w2) {
  const PLanes<NPG, LPR> ln;
  if (ln.item >= (int64_t)nrows * gr.ng) return;
  const int q = (int)(ln.item / gr.ng), g = (int)(ln.item - (int64_t)q * gr.ng);
  const int r = rows[q];
  const PcTab<R>& T = tab[gr.gtab[g]];
  const R* bx = T.Bx;
  const cplx<WT>* wg = w + (int64_t)g * nn * NPG + ln.j;
  double2 s = strided_dot<WT, NPG, LPR>(bx, bi, wg, bp[r] + ln.part, bp[r + 1]);
  s = ln.sum(s);
  if (ln.part == 0) {
    cplx<WT>& o = w2[((int64_t)g * nn + r) * NPG + ln.j];
    const double inv = 1.0 / (double)T.D[r];
    o.x = (WT)((double)o.x * inv - s.x);
    o.y = (WT)((double)o.y * inv - s.y);
  }
}

template <typename R, typename WT, int NPG, int LPR>
__global__ __launch_bounds__(256) void k_sn_gemv_bwd(const int* __restrict__ rows, int nrows,
                                                     const int* __restrict__ sn_of_row, const int* __restrict__ sn_ptr,
                                                     const int64_t* __restrict__ sn_off,
                                                     const PcTab<R>* __restrict__ tab, PcGroupDev gr, int nn,
                                                     const cplx<WT>* __restrict__ w2, cplx<WT>* __restrict__ w) {
  const PLanes<NPG, LPR> ln;
  if (ln.item >= (int64_t)nrows * gr.ng) return;
  const int q = (int)(ln.item / gr.ng), g = (int)(ln.item - (int64_t)q * gr.ng);
  const int r = rows[q];
  const int k = sn_of_row[r];
  const int c0 = sn_ptr[k], sz = sn_ptr[k + 1] - c0;
  const int i = r - c0;
  // row i of the transpose holds Linv(i..sz-1, i): offset i*sz - i(i-1)/2
  const R* row = tab[gr.gtab[g]].LiT + sn_off[k] + (int64_t)i * sz - (int64_t)i * (i - 1) / 2;
  const cplx<WT>* uu = w2 + ((int64_t)g * nn + r) * NPG + ln.j;
  double2 s; s.x = 0; s.y = 0;
  for (int jj = ln.part; jj < sz - i; jj += LPR) {
    const double l = row[jj];
    const cplx<WT> v = uu[(int64_t)jj * NPG];
    s.x += l * v.x; s.y += l * v.y;
  }
  s = ln.sum(s);
  if (ln.part == 0) {
    cplx<WT> o; o.x = (WT)s.x; o.y = (WT)s.y;
    w[((int64_t)g * nn + r) * NPG + ln.j] = o;
  }
}

// Block kernel: one workgroup per (block k of a level, group g) does the
// off-block update of all the block's rows and its dense inverse product:
//   forward   t_r = b_r - sum_j L(r,j) y_j   then  y_k = Linv_k t_k
//   backward  u_r = y_r/D_r - sum_i L(i,r) x_i  then  x_k = Linv_k' u_k
// Everything the two products read is first staged in shared memory, in one
// round of independent loads after the block's metadata: the work-vector
// lines of the block's column union (ucol[up[k] .. up[k+1]), sorted; an entry
// addresses its line through loc[p], the position of its index in the union),
// the block's own values, the row ranges, and -- where they fit (BlkLds) --
// the block's off-block entries (values and union positions, one contiguous
// run in the packed layout) and the packed dense inverse.  The products then run out of shared memory, so a
// workgroup waits on memory about three times instead of once per row pass.
// Lanes: NPG ports x 2^llu parts per row in the update, NPG x 2^llg in the
// product (each at most 64 and at most NT).  The workgroup size NT follows
// the tier's largest block (smaller workgroups for small blocks: more of
// them in flight per CU hide the staging latency).
__device__ __forceinline__ double2 lane_sum(double2 v, int from, int to) {
  for (int o = from; o < to; o <<= 1) { v.x += __shfl_xor(v.x, o, 64); v.y += __shfl_xor(v.y, o, 64); }
  return v;
}

// Shared-memory regions of one block-kernel launch, in elements: staged
// lines and own values (complex work type), packed inverse and entry span
// (0: read from global memory instead), and the row ranges (2 per row).
struct BlkLds {
  int stg_n, tt_n, inv_n, ent_n;
};

template <typename R, typename WT, int NPG, bool FWD, int NT>
__global__ __launch_bounds__(NT) void k_blk_solve(const BlkInfo* __restrict__ info, const int* __restrict__ pp,
                                                   const int* __restrict__ loc,
                                                   const int* __restrict__ ucol, const PcTab<R>* __restrict__ tab,
                                                   PcGroupDev gr, int nn, cplx<WT>* __restrict__ w,
                                                   cplx<WT>* __restrict__ w2, BlkLds lay, int llu, int llg) {
  using CW = cplx<WT>;
  extern __shared__ __align__(16) unsigned char smem_pc_raw[];
  CW* stg = (CW*)smem_pc_raw;
  CW* tt = stg + lay.stg_n;
  R* sinv = (R*)(tt + lay.tt_n);
  R* sval = sinv + lay.inv_n;
  int* sloc = (int*)(sval + lay.ent_n);
  int* rng = sloc + lay.ent_n;
  const int q = blockIdx.x / gr.ng, g = blockIdx.x - q * gr.ng;
  const BlkInfo bi = info[q];
  const int c0 = bi.c0, s = bi.s, u0 = bi.u0, nu = bi.nu, e0 = bi.e0, ne = bi.ne;
  const PcTab<R>& T = tab[gr.gtab[g]];
  const R* __restrict__ val = FWD ? T.Fx : T.Bx;
  const R* __restrict__ Lb = (FWD ? T.Li : T.LiT) + bi.off;
  // forward: staged lines are y (w2), the block's own values b come from w;
  // backward: staged lines are x (w), the own values y from w2
  const CW* __restrict__ src = FWD ? w2 : w;
  const CW* __restrict__ own = FWD ? w : w2;
  CW* __restrict__ dst = FWD ? w2 : w;
  const int64_t gb = (int64_t)g * nn * NPG;
  const int tid = threadIdx.x;
  const int j = tid % NPG;
  const bool inv_lds = lay.inv_n > 0, ent_lds = lay.ent_n > 0;
  // ---- staging: every load below depends on the block metadata only (the
  // lines also on their union index); each thread issues up to four loads of
  // a region before storing them
  {
    constexpr int U = 4;
    for (int b = tid; b < nu * NPG; b += U * NT) {
      int cq[U];
#pragma unroll
      for (int u = 0; u < U; ++u) cq[u] = b + u * NT < nu * NPG ? ucol[u0 + (b + u * NT) / NPG] : 0;
      CW v[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (b + u * NT < nu * NPG) v[u] = src[gb + (int64_t)cq[u] * NPG + (b + u * NT) % NPG];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (b + u * NT < nu * NPG) stg[b + u * NT] = v[u];
    }
    for (int b = tid; b < s * NPG; b += U * NT) {
      CW v[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (b + u * NT < s * NPG) v[u] = own[gb + (int64_t)c0 * NPG + b + u * NT];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = b + u * NT;
        if (e < s * NPG) {
          if (FWD) {
            tt[e] = v[u];
          } else {
            const double inv = 1.0 / (double)T.D[c0 + e / NPG];
            CW o; o.x = (WT)((double)v[u].x * inv); o.y = (WT)((double)v[u].y * inv);
            tt[e] = o;
          }
        }
      }
    }
    for (int i = tid; i < s; i += NT) {
      rng[2 * i] = pp[c0 + i] - e0;
      rng[2 * i + 1] = pp[c0 + i + 1] - e0;
    }
    if (inv_lds) {
      const int ni = s * (s + 1) / 2;
      for (int b = tid; b < ni; b += U * NT) {
        R v[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (b + u * NT < ni) v[u] = Lb[b + u * NT];
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (b + u * NT < ni) sinv[b + u * NT] = v[u];
      }
    }
    if (ent_lds) {
      for (int b = tid; b < ne; b += U * NT) {
        R v[U];
        int l[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (b + u * NT < ne) { v[u] = val[e0 + b + u * NT]; l[u] = loc[e0 + b + u * NT]; }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (b + u * NT < ne) { sval[b + u * NT] = v[u]; sloc[b + u * NT] = l[u]; }
      }
    }
  }
  __syncthreads();
  {
    const int lu = NPG << llu, lpr = 1 << llu;
    const int part = (tid / NPG) & (lpr - 1), rs = tid / lu, rpp = NT / lu;
    for (int i = rs; i < s; i += rpp) {
      const int lo = rng[2 * i], hi = rng[2 * i + 1];
      double2 a = make_double2(0.0, 0.0);
      int p = lo + part;
      // four strides at a time: their entries (from LDS or memory) are all
      // requested before any is used
      for (; p + 3 * lpr < hi; p += 4 * lpr) {
        double l[4];
        int c[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int q = p + u * lpr;
          l[u] = ent_lds ? (double)sval[q] : (double)val[e0 + q];
          c[u] = ent_lds ? sloc[q] : loc[e0 + q];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const CW v = stg[c[u] * NPG + j];
          a.x += l[u] * v.x; a.y += l[u] * v.y;
        }
      }
      for (; p < hi; p += lpr) {
        const double l = ent_lds ? (double)sval[p] : (double)val[e0 + p];
        const CW v = stg[(ent_lds ? sloc[p] : loc[e0 + p]) * NPG + j];
        a.x += l * v.x; a.y += l * v.y;
      }
      a = lane_sum(a, NPG, lu);
      if (part == 0) {
        CW& o = tt[i * NPG + j];
        o.x = (WT)((double)o.x - a.x);
        o.y = (WT)((double)o.y - a.y);
      }
    }
  }
  __syncthreads();
  {
    const int lg = NPG << llg, lpr = 1 << llg;
    const int part = (tid / NPG) & (lpr - 1), rs = tid / lg, rpp = NT / lg;
    const R* __restrict__ Lrow = inv_lds ? sinv : Lb;
    for (int i = rs; i < s; i += rpp) {
      // forward: row i of Linv holds Linv(i, 0..i) at i(i+1)/2; backward:
      // row i of the transpose holds Linv(i..s-1, i) at i*s - i(i-1)/2
      const R* __restrict__ row = FWD ? Lrow + (int64_t)i * (i + 1) / 2 : Lrow + (int64_t)i * s - (int64_t)i * (i - 1) / 2;
      const int len = FWD ? i + 1 : s - i;
      const CW* tv = FWD ? tt + j : tt + i * NPG + j;
      double2 a = make_double2(0.0, 0.0);
      int jj = part;
      for (; jj + 3 * lpr < len; jj += 4 * lpr) {
        double l[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) l[u] = (double)row[jj + u * lpr];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const CW v = tv[(jj + u * lpr) * NPG];
          a.x += l[u] * v.x; a.y += l[u] * v.y;
        }
      }
      for (; jj < len; jj += lpr) {
        const double l = (double)row[jj];
        const CW v = tv[jj * NPG];
        a.x += l * v.x; a.y += l * v.y;
      }
      a = lane_sum(a, NPG, lg);
      if (part == 0) {
        CW o; o.x = (WT)a.x; o.y = (WT)a.y;
        dst[gb + (int64_t)(c0 + i) * NPG + j] = o;
      }
    }
  }
}

// y[nc + v] = w[g][iperm[v]][j] of column (g, j): a workgroup reads the
// lines of NN = 256/NPG consecutive nodes (ports fastest) and writes each
// system's NN node entries contiguously (node fastest), through shared memory.
template <typename R, typename WT, int NPG>
__global__ __launch_bounds__(256) void k_pc_scatter_nodes(const cplx<WT>* __restrict__ w, cplx<R>* __restrict__ y,
                                                          int64_t ldy, int nn, const int* __restrict__ iperm,
                                                          int64_t nc, PcGroupDev gr) {
  constexpr int NN = 256 / NPG;
  __shared__ cplx<WT> tile[NN * NPG];
  const int tiles = (nn + NN - 1) / NN;
  const int g = blockIdx.x / tiles, v0 = (blockIdx.x - g * tiles) * NN;
  {
    const int j = threadIdx.x % NPG, nl = threadIdx.x / NPG;
    const int node = v0 + nl;
    cplx<WT> v; v.x = 0; v.y = 0;
    if (node < nn) v = w[((int64_t)g * nn + iperm[node]) * NPG + j];
    tile[nl * NPG + j] = v;
  }
  __syncthreads();
  {
    const int nl = threadIdx.x % NN, j = threadIdx.x / NN;
    const int node = v0 + nl;
    const int slot = gr.gslot[g * NPG + j];
    if (node < nn && slot >= 0) {
      const cplx<WT> v = tile[nl * NPG + j];
      cplx<R> o; o.x = (R)v.x; o.y = (R)v.y;
      y[(int64_t)slot * ldy + nc + node] = o;
    }
  }
}

// y[k] = (a[k] + sum_i ArTp(k,i) w[g][i][j]) * Ainv[k]   (ArTp columns are permuted node ids).
// One thread per (current k, group g) serves all NPG ports: the
// current's incidence row and Ainv[k] are loaded once, the lines of its nodes
// (at most 6) are read whole, and a wave reads and writes 64 consecutive
// currents of each port.
template <typename R, typename WT, int NPG>
__global__ __launch_bounds__(256) void k_pc_currents(const cplx<R>* __restrict__ x, int64_t ldx,
                                                     const cplx<WT>* __restrict__ w, cplx<R>* __restrict__ y,
                                                     int64_t ldy, int nn, int64_t nc, const int* __restrict__ tp,
                                                     const int* __restrict__ tc, const R* __restrict__ tv,
                                                     const PcTab<R>* __restrict__ tab, PcGroupDev gr) {
  // one-dimensional grid of current blocks x groups in XCD order
  const int nbx = (int)((nc + 255) >> 8);
  const int lt = xcd_tile(blockIdx.x, gridDim.x);
  const int g = lt / nbx;
  const int64_t k = (int64_t)(lt - g * nbx) * 256 + threadIdx.x;
  if (k >= nc) return;
  const double ai = (double)tab[gr.gtab[g]].ainv[k];
  const cplx<WT>* wg = w + (int64_t)g * nn * NPG;
  constexpr int D = 6;
  int p0 = 0, p1 = 0;
  if (tc) { p0 = tp[k]; p1 = tp[k + 1]; }
  const bool short_row = p1 - p0 <= D;
  int nd[D];
  double cv[D];
  if (short_row) {
#pragma unroll
    for (int u = 0; u < D; ++u) {
      nd[u] = p0 + u < p1 ? tc[p0 + u] : 0;
      cv[u] = p0 + u < p1 ? (double)tv[p0 + u] : 0.0;
    }
  }
  constexpr int PC = NPG < 4 ? NPG : 4;
#pragma unroll
  for (int j0 = 0; j0 < NPG; j0 += PC) {
    int sl[PC];
    cplx<R> a[PC];
#pragma unroll
    for (int jj = 0; jj < PC; ++jj) {
      sl[jj] = gr.gslot[g * NPG + j0 + jj];
      if (sl[jj] >= 0) a[jj] = x[(int64_t)sl[jj] * ldx + k];
    }
    double sx[PC], sy[PC];
#pragma unroll
    for (int jj = 0; jj < PC; ++jj) { sx[jj] = sl[jj] >= 0 ? (double)a[jj].x : 0.0; sy[jj] = sl[jj] >= 0 ? (double)a[jj].y : 0.0; }
    if (short_row) {
#pragma unroll
      for (int u = 0; u < D; ++u)
        if (p0 + u < p1) {
          const cplx<WT>* ln = wg + (int64_t)nd[u] * NPG + j0;
#pragma unroll
          for (int jj = 0; jj < PC; ++jj) { const cplx<WT> v = ln[jj]; sx[jj] += cv[u] * v.x; sy[jj] += cv[u] * v.y; }
        }
    } else {
      for (int p = p0; p < p1; ++p) {
        const cplx<WT>* ln = wg + (int64_t)tc[p] * NPG + j0;
        const double c = (double)tv[p];
#pragma unroll
        for (int jj = 0; jj < PC; ++jj) { const cplx<WT> v = ln[jj]; sx[jj] += c * v.x; sy[jj] += c * v.y; }
      }
    }
#pragma unroll
    for (int jj = 0; jj < PC; ++jj)
      if (sl[jj] >= 0) {
        cplx<R> o; o.x = (R)(sx[jj] * ai); o.y = (R)(sy[jj] * ai);
        y[(int64_t)sl[jj] * ldy + k] = o;
      }
  }
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

// Blocks of each level (bptr/blk, ascending key within a level) and their
// rows in the same order (lptr/rows), plus rowoff[i] = position in rows of
// the first row of blk[i] (rowoff has one more entry, = rows.size()).
static void level_blocks(const std::vector<int>& lev, const std::vector<int>& sn_ptr, const std::vector<int>& key,
                         LevelBlocks& B) {
  const int nsn = (int)lev.size();
  int nl = 0;
  for (int l : lev) nl = std::max(nl, l + 1);
  B.bptr.assign(nl + 1, 0);
  for (int k = 0; k < nsn; ++k) B.bptr[lev[k] + 1]++;
  for (int l = 0; l < nl; ++l) B.bptr[l + 1] += B.bptr[l];
  B.blk.assign(nsn, 0);
  {
    std::vector<int> fill(B.bptr.begin(), B.bptr.end() - 1);
    for (int k = 0; k < nsn; ++k) B.blk[fill[lev[k]]++] = k;
  }
  for (int l = 0; l < nl; ++l)
    std::stable_sort(B.blk.begin() + B.bptr[l], B.blk.begin() + B.bptr[l + 1],
                     [&](int a, int b) { return key[a] < key[b]; });
  B.lptr.assign(nl + 1, 0);
  B.rows.clear();
  B.rowoff.assign(nsn + 1, 0);
  for (int l = 0; l < nl; ++l) {
    for (int i = B.bptr[l]; i < B.bptr[l + 1]; ++i) {
      B.rowoff[i] = (int)B.rows.size();
      for (int r = sn_ptr[B.blk[i]]; r < sn_ptr[B.blk[i] + 1]; ++r) B.rows.push_back(r);
    }
    B.lptr[l + 1] = (int)B.rows.size();
  }
  B.rowoff[nsn] = (int)B.rows.size();
}

// Column union of every block: up/ucol list, for each block k, the sorted
// distinct indices idx[p] over the entries p in [beg(r), end(r)) of its rows
// r; loc[p] is the position of idx[p] in that list.
template <class BEG, class END>
static void block_unions(const std::vector<int>& sn_ptr, const int* idx, int64_t nent, BEG beg, END end,
                         std::vector<int>& up, std::vector<int>& ucol, std::vector<int>& loc) {
  const int nsn = (int)sn_ptr.size() - 1;
  up.assign(nsn + 1, 0);
  ucol.clear();
  loc.assign(nent, 0);
  std::vector<int> tmp;
  for (int k = 0; k < nsn; ++k) {
    tmp.clear();
    for (int r = sn_ptr[k]; r < sn_ptr[k + 1]; ++r)
      for (int p = beg(r); p < end(r); ++p) tmp.push_back(idx[p]);
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    for (int r = sn_ptr[k]; r < sn_ptr[k + 1]; ++r)
      for (int p = beg(r); p < end(r); ++p)
        loc[p] = (int)(std::lower_bound(tmp.begin(), tmp.end(), idx[p]) - tmp.begin());
    ucol.insert(ucol.end(), tmp.begin(), tmp.end());
    up[k + 1] = (int)ucol.size();
  }
  // the block kernel relies on: loc addresses the entry's own index, and
  // increases along each row (a chunk's entries are one run)
  for (int k = 0; k < nsn; ++k)
    for (int r = sn_ptr[k]; r < sn_ptr[k + 1]; ++r)
      for (int p = beg(r); p < end(r); ++p)
        VH_REQUIRE(loc[p] >= 0 && loc[p] < up[k + 1] - up[k] && ucol[up[k] + loc[p]] == idx[p] &&
                       (p == beg(r) || loc[p] > loc[p - 1]),
                   "internal: block union index");
}

// Sort key of a block for the block kernel's tiers: the larger of its size,
// its column union and a quarter of its off-block entries (a half, an
// eighth, a sixteenth measured: numex1 sweep 0.074 / 0.081 / 0.089 / 0.089 s
// mixed, bus step 0.616 / 0.599 / 0.600 / 0.607 s; FP64 0.632 / 0.620 /
// 0.634 / 0.646 s).  A block whose
// entries are many for its workgroup (the top separators of a compact
// conductor: 1e5 entries in one block) thereby falls to the row kernels,
// which spread its rows over many workgroups (numex1: one such block took
// 1.5 ms in one workgroup per group).
static int block_key(int sz, int un, int ne) { return std::max(std::max(sz, un), (ne + 3) / 4); }

// Per level, along the sorted blocks: the sort key (block_key) and the
// prefix maxima of sizes, unions and entry spans (the block kernel sizes its
// shared memory from them).  ptr: the row (forward, CSR) or column
// (backward, CSC) pointers of L.
static void level_extents(const std::vector<int>& sn_ptr, const std::vector<int>& up, const std::vector<int>& ptr,
                          LevelBlocks& B) {
  const int nb = (int)B.blk.size();
  B.key.assign(nb, 0);
  B.smax.assign(nb, 0);
  B.umax.assign(nb, 0);
  B.emax.assign(nb, 0);
  for (int l = 0; l + 1 < (int)B.bptr.size(); ++l) {
    int ms = 0, mu = 0, me = 0;
    for (int i = B.bptr[l]; i < B.bptr[l + 1]; ++i) {
      const int k = B.blk[i];
      const int sz = sn_ptr[k + 1] - sn_ptr[k], un = up[k + 1] - up[k];
      B.key[i] = block_key(sz, un, ptr[sn_ptr[k + 1]] - ptr[sn_ptr[k]]);
      ms = std::max(ms, sz);
      mu = std::max(mu, un);
      me = std::max(me, ptr[sn_ptr[k + 1]] - ptr[sn_ptr[k]]);
      B.smax[i] = ms;
      B.umax[i] = mu;
      B.emax[i] = me;
    }
  }
}

// Per sorted block position: rows, union range, inverse offset and entry span
// (what the block kernel reads first, in one load).
static std::vector<BlkInfo> block_info(const LevelBlocks& B, const std::vector<int>& sn_ptr,
                                       const std::vector<int64_t>& sn_off, const std::vector<int>& up,
                                       const std::vector<int>& ptr) {
  std::vector<BlkInfo> v(B.blk.size());
  for (size_t i = 0; i < B.blk.size(); ++i) {
    const int k = B.blk[i];
    v[i].c0 = sn_ptr[k];
    v[i].s = sn_ptr[k + 1] - sn_ptr[k];
    v[i].u0 = up[k];
    v[i].nu = up[k + 1] - up[k];
    v[i].off = sn_off[k];
    v[i].e0 = ptr[sn_ptr[k]];
    v[i].ne = ptr[sn_ptr[k + 1]] - ptr[sn_ptr[k]];
  }
  return v;
}

static std::vector<int> block_keys(const std::vector<int>& sn_ptr, const std::vector<int>& up, const std::vector<int>& ptr) {
  std::vector<int> key(sn_ptr.size() - 1);
  for (size_t k = 0; k + 1 < sn_ptr.size(); ++k)
    key[k] = block_key(sn_ptr[k + 1] - sn_ptr[k], up[k + 1] - up[k], ptr[sn_ptr[k + 1]] - ptr[sn_ptr[k]]);
  return key;
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
  // unions: forward over the CSR entries left of each row's block, backward
  // over the CSC entries below each column's block
  block_unions(Y.sn_ptr, Ri.data(), Rp[n], [&](int r) { return Rp[r]; }, [&](int r) { return Y.row_split[r]; },
               Y.f_up, Y.f_ucol, Y.f_loc);
  block_unions(Y.sn_ptr, F.Li.data(), F.Lp[n], [&](int r) { return Y.col_split[r]; },
               [&](int r) { return F.Lp[r + 1]; }, Y.b_up, Y.b_ucol, Y.b_loc);
  // off-block entries packed contiguously: forward row by row (CSR order),
  // backward column by column (CSC order); the values of every frequency
  // are stored in these orders (fsrc / bsrc: their CSC positions)
  {
    std::vector<int> csr2csc(Rp[n]);
    for (int p = 0; p < Rp[n]; ++p) csr2csc[Y.csc2csr[p]] = p;
    Y.fp.assign(n + 1, 0);
    Y.bp.assign(n + 1, 0);
    Y.fi.clear(); Y.floc.clear(); Y.fsrc.clear();
    Y.bi.clear(); Y.bloc.clear(); Y.bsrc.clear();
    for (int r = 0; r < n; ++r) {
      for (int p = Rp[r]; p < Y.row_split[r]; ++p) {
        Y.fi.push_back(Ri[p]); Y.floc.push_back(Y.f_loc[p]); Y.fsrc.push_back(csr2csc[p]);
      }
      Y.fp[r + 1] = (int)Y.fi.size();
      for (int q = Y.col_split[r]; q < F.Lp[r + 1]; ++q) {
        Y.bi.push_back(F.Li[q]); Y.bloc.push_back(Y.b_loc[q]); Y.bsrc.push_back(q);
      }
      Y.bp[r + 1] = (int)Y.bi.size();
    }
  }
  std::vector<int>().swap(Y.f_loc);     // (only the packed copies are used from here on)
  std::vector<int>().swap(Y.b_loc);
  std::vector<int>().swap(Y.csc2csr);
  level_blocks(lf, Y.sn_ptr, block_keys(Y.sn_ptr, Y.f_up, Y.fp), Y.fl);
  level_blocks(lb, Y.sn_ptr, block_keys(Y.sn_ptr, Y.b_up, Y.bp), Y.bl);
  Y.f_lptr = Y.fl.lptr; Y.f_rows = Y.fl.rows;
  Y.b_lptr = Y.bl.lptr; Y.b_rows = Y.bl.rows;
  level_extents(Y.sn_ptr, Y.f_up, Y.fp, Y.fl);
  level_extents(Y.sn_ptr, Y.b_up, Y.bp, Y.bl);
  if (trace_on()) {
    // block statistics: sizes, and the dense panels a block-wise update would use
    int64_t off_nnz = 0, fpanel = 0, bpanel = 0, fu = 0, bu = 0;
    int smax = 0, umax = 0;
    const int TH[5] = {128, 256, 512, 1024, 2048};
    int64_t fin[5] = {0, 0, 0, 0, 0}, bin[5] = {0, 0, 0, 0, 0}, bnnz = 0;
    std::vector<int> mark(n, -1);
    for (int k = 0; k < nsn; ++k) {
      const int c0 = Y.sn_ptr[k], c1 = Y.sn_ptr[k + 1];
      smax = std::max(smax, c1 - c0);
      int u = 0;
      int64_t e = 0;
      for (int r = c0; r < c1; ++r)
        for (int p = Rp[r]; p < Y.row_split[r]; ++p) {
          ++off_nnz; ++e;
          if (mark[Ri[p]] != k) { mark[Ri[p]] = k; ++u; }
        }
      umax = std::max(umax, u);
      fpanel += (int64_t)u * (c1 - c0);
      fu += u;
      for (int t = 0; t < 5; ++t) if (u <= TH[t]) fin[t] += e;
    }
    std::fill(mark.begin(), mark.end(), -1);
    for (int k = 0; k < nsn; ++k) {
      const int c0 = Y.sn_ptr[k], c1 = Y.sn_ptr[k + 1];
      int v = 0;
      int64_t e = 0;
      for (int r = c0; r < c1; ++r)
        for (int q = Y.col_split[r]; q < F.Lp[r + 1]; ++q) {
          ++e;
          if (mark[F.Li[q]] != k) { mark[F.Li[q]] = k; ++v; }
        }
      bpanel += (int64_t)v * (c1 - c0);
      bu += v; bnnz += e;
      for (int t = 0; t < 5; ++t) if (v <= TH[t]) bin[t] += e;
    }
    fprintf(stderr, "[vh] schur unions: fwd sum %ld, bwd sum %ld (bwd nnz %ld); entries in blocks with union <= "
                    "128/256/512/1024/2048: fwd %.3f %.3f %.3f %.3f %.3f bwd %.3f %.3f %.3f %.3f %.3f\n",
            (long)fu, (long)bu, (long)bnnz, fin[0] / (double)off_nnz, fin[1] / (double)off_nnz,
            fin[2] / (double)off_nnz, fin[3] / (double)off_nnz, fin[4] / (double)off_nnz, bin[0] / (double)bnnz,
            bin[1] / (double)bnnz, bin[2] / (double)bnnz, bin[3] / (double)bnnz, bin[4] / (double)bnnz);
    fprintf(stderr, "[vh] schur: n %d nnzL %ld blocks %d (max %d, mean %.1f) levels f %zu b %zu; "
                    "off-block nnz %ld, dense panels fwd %ld bwd %ld (max union %d), inverses %ld\n",
            n, (long)F.Lp[n], nsn, smax, (double)n / nsn, Y.f_lptr.size() - 1, Y.b_lptr.size() - 1, (long)off_nnz,
            (long)fpanel, (long)bpanel, umax, (long)Y.dense);
  }
  // lanes per item: the power of two nearest above half the mean row length, in [1, 64]
  auto lpr = [](int64_t tot, int64_t cnt) {
    const double avg = cnt > 0 ? (double)tot / (double)cnt : 0.0;
    int l = 1;
    while (l < 64 && 2.0 * l < avg) l <<= 1;
    return l;
  };
  const int nlf = (int)Y.f_lptr.size() - 1, nlb = (int)Y.b_lptr.size() - 1;
  Y.f_lpr_u.assign(nlf, 1); Y.f_lpr_g.assign(nlf, 1); Y.f_upd_nnz.assign(nlf, 0);
  Y.b_lpr_u.assign(nlb, 1); Y.b_lpr_g.assign(nlb, 1);
  for (int l = 0; l < nlf; ++l) {
    int64_t u = 0, g = 0;
    for (int q = Y.f_lptr[l]; q < Y.f_lptr[l + 1]; ++q) {
      const int r = Y.f_rows[q];
      u += Y.row_split[r] - Rp[r];
      g += r - Y.sn_ptr[Y.sn_of_row[r]] + 1;
    }
    const int64_t cnt = Y.f_lptr[l + 1] - Y.f_lptr[l];
    Y.f_upd_nnz[l] = u;
    Y.f_lpr_u[l] = lpr(u, cnt);
    Y.f_lpr_g[l] = lpr(g, cnt);
  }
  for (int l = 0; l < nlb; ++l) {
    int64_t u = 0, g = 0;
    for (int q = Y.b_lptr[l]; q < Y.b_lptr[l + 1]; ++q) {
      const int r = Y.b_rows[q];
      const int k = Y.sn_of_row[r];
      u += F.Lp[r + 1] - Y.col_split[r];
      g += Y.sn_ptr[k + 1] - r;
    }
    const int64_t cnt = Y.b_lptr[l + 1] - Y.b_lptr[l];
    Y.b_lpr_u[l] = lpr(u, cnt);
    Y.b_lpr_g[l] = lpr(g, cnt);
  }
  upload(Y.d_perm, F.perm);
  upload(Y.d_iperm, F.iperm);
  upload(Y.d_fp, Y.fp);
  upload(Y.d_fi, Y.fi);
  upload(Y.d_floc, Y.floc);
  upload(Y.d_bp, Y.bp);
  upload(Y.d_bi, Y.bi);
  upload(Y.d_bloc, Y.bloc);
  upload(Y.d_sn_ptr, Y.sn_ptr);
  upload(Y.d_sn_of_row, Y.sn_of_row);
  upload(Y.d_f_rows, Y.f_rows);
  upload(Y.d_b_rows, Y.b_rows);
  upload(Y.d_f_info, block_info(Y.fl, Y.sn_ptr, Y.sn_off, Y.f_up, Y.fp));
  upload(Y.d_b_info, block_info(Y.bl, Y.sn_ptr, Y.sn_off, Y.b_up, Y.bp));
  upload(Y.d_f_ucol, Y.f_ucol);
  upload(Y.d_b_ucol, Y.b_ucol);
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

void precon_prepare(Precon& P, double jre, double jim, const std::complex<double>* zin) {
  Geometry& G = *P.G;
  VH_REQUIRE(G.has_inc && (G.has_z || zin), "preconditioner: incidence and impedance must be set first");
  const int nv = G.g.nv;
  const int64_t nc = G.num_curr;
  const int nn = (int)G.num_node;
  // A_inv (lse_sparse_precon_prepare.m:36-62): 1 / |z/(s dx) + jwmu g_self|
  const std::complex<double> jw(jre, jim);
  std::complex<double> st[3];
  for (int k = 0; k < 3; ++k) st[k] = jw * std::complex<double>(G.st2[2 * k], G.st2[2 * k + 1]);
  P.zh.assign(zin ? zin : G.zh.data(), (zin ? zin : G.zh.data()) + nv);
  P.kinetic = std::any_of(P.zh.begin(), P.zh.end(), [](const std::complex<double>& z) { return z.imag() != 0.0; });
  std::vector<double> ainv(nc);
  for (int v = 0; v < nv; ++v) {
    const std::complex<double> z = P.zh[v];
    const double dp = 1.0 / std::abs(z / G.dx + st[0]);
    const double d2 = 1.0 / std::abs(z / (6.0 * G.dx) + st[1]);
    const double d3 = 1.0 / std::abs(z / (2.0 * G.dx) + st[2]);
    ainv[v] = dp; ainv[nv + v] = dp; ainv[2 * nv + v] = dp;
    ainv[3 * nv + v] = d2; ainv[4 * nv + v] = d3;
  }
  P.ainv_h = ainv;
  // The first preparation of a geometry builds the shared symbolic data (and,
  // after its numeric factorisation, the level schedules); concurrent
  // preparations wait on the lock until it is complete, later ones skip it.
  bool fresh = false;
  std::unique_lock<std::mutex> lk(G.precon_sym_mu);
  if (!G.precon_sym) {
    G.precon_sym = std::make_shared<PreconSymbolic>();
    build_symbolic(G, P.geometric_order, *G.precon_sym);
    fresh = true;
  } else {
    lk.unlock();
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
    lk.unlock();
  }
  std::vector<double> Fx(Y.fsrc.size()), Bx(Y.bsrc.size());
  for (size_t q = 0; q < Fx.size(); ++q) Fx[q] = P.F.Lx[Y.fsrc[q]];
  for (size_t q = 0; q < Bx.size(); ++q) Bx[q] = P.F.Lx[Y.bsrc[q]];
  std::vector<double> Li, LiT;
  block_inverses(Y, P.F, Li, LiT);
  upload(P.Fx64, Fx); upload_f32(P.Fx32, Fx);
  upload(P.Bx64, Bx); upload_f32(P.Bx32, Bx);
  upload(P.D64, P.F.D); upload_f32(P.D32, P.F.D);
  upload(P.ainv64, ainv); upload_f32(P.ainv32, ainv);
  upload(P.Li64, Li); upload_f32(P.Li32, Li);
  upload(P.LiT64, LiT); upload_f32(P.LiT32, LiT);
  {
    std::vector<float2> z32(nv);
    for (int v = 0; v < nv; ++v) { z32[v].x = (float)P.zh[v].real(); z32[v].y = (float)P.zh[v].imag(); }
    P.z64.alloc(nv);
    P.z32.alloc(nv);
    VH_HIP(hipMemcpy(P.z64.p, P.zh.data(), nv * sizeof(double2), hipMemcpyHostToDevice));
    VH_HIP(hipMemcpy(P.z32.p, z32.data(), nv * sizeof(float2), hipMemcpyHostToDevice));
  }
  P.tab64_h = PcTab<double>{P.Fx64.p, P.Bx64.p, P.D64.p, P.ainv64.p, P.Li64.p, P.LiT64.p, P.z64.p};
  P.tab32_h = PcTab<float>{P.Fx32.p, P.Bx32.p, P.D32.p, P.ainv32.p, P.Li32.p, P.LiT32.p, P.z32.p};
  P.tab64.alloc(1);
  P.tab32.alloc(1);
  VH_HIP(hipMemcpy(P.tab64.p, &P.tab64_h, sizeof(PcTab<double>), hipMemcpyHostToDevice));
  VH_HIP(hipMemcpy(P.tab32.p, &P.tab32_h, sizeof(PcTab<float>), hipMemcpyHostToDevice));
  // the values are on the device; the host copies of L are not needed any more
  P.F.Li.clear(); P.F.Li.shrink_to_fit();
  P.F.Lx.clear(); P.F.Lx.shrink_to_fit();
  P.ready = true;
}

void precon_prepare_many(Precon* const* P, int n, const std::complex<double>* jw, const std::complex<double>* z,
                         int nthreads) {
  if (n <= 0) return;
  const int nv = P[0]->G->g.nv;
  auto one = [&](int k) {
    precon_prepare(*P[k], jw[k].real(), jw[k].imag(), z ? z + (size_t)k * nv : nullptr);
  };
  one(0);   // builds the symbolic data once
  if (n == 1) return;
  if (nthreads <= 0) nthreads = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
  nthreads = std::min(nthreads, n - 1);
  const int dev = P[0]->G->ctx->device;
  std::atomic<int> next(1);
  std::exception_ptr err;
  std::mutex em;
  std::vector<std::thread> pool;
  for (int t = 0; t < nthreads; ++t)
    pool.emplace_back([&] {
      if (hipSetDevice(dev) != hipSuccess) return;
      for (;;) {
        const int k = next++;
        if (k >= n) break;
        try {
          one(k);
        } catch (...) {
          std::lock_guard<std::mutex> g(em);
          if (!err) err = std::current_exception();
        }
      }
    });
  for (auto& th : pool) th.join();
  if (err) std::rethrow_exception(err);
}

template <int V> using IC = std::integral_constant<int, V>;

// f(IC<NPG>, IC<LPR>) for the runtime lane counts, LPR clamped to 64 / NPG
template <int N, int L, typename F> static void lanes_if(F& f) {
  if constexpr (N * L <= 64) f(IC<N>{}, IC<L>{});
}
template <int N, typename F> static void lanes_l(int lpr, F& f) {
  switch (std::min(lpr, 64 / N)) {
    case 1: lanes_if<N, 1>(f); break;
    case 2: lanes_if<N, 2>(f); break;
    case 4: lanes_if<N, 4>(f); break;
    case 8: lanes_if<N, 8>(f); break;
    case 16: lanes_if<N, 16>(f); break;
    case 32: lanes_if<N, 32>(f); break;
    default: lanes_if<N, 64>(f); break;
  }
}
template <typename F> static void with_lanes(int npg, int lpr, F&& f) {
  switch (npg) {
    case 1: lanes_l<1>(lpr, f); break;
    case 2: lanes_l<2>(lpr, f); break;
    case 4: lanes_l<4>(lpr, f); break;
    default: lanes_l<8>(lpr, f); break;
  }
}

// grid for nitems items of lanes threads each, 256-thread blocks
static dim3 lane_grid(int64_t nitems, int lanes) { return dim3((unsigned)ceil_div(nitems * lanes, 256)); }

// Groups of the columns: the slots of each preconditioner in chunks of npg.
// npg (1, 2, 4 or 8) minimises an estimate of the level kernels' traffic:
// per group and entry of L, its value and index (12 bytes) and the line of
// the work vector (at least 64 bytes, 16 per port); on a tie the smaller npg
// (one system per preconditioner gets groups of one, not of four with three
// empty slots whose lanes idle).
static int build_groups(const int* pidx, const std::vector<int>& slots, std::vector<int>& key) {
  std::vector<int> order;            // preconditioners in order of first use
  std::vector<std::vector<int>> cols;
  {
    std::vector<int> where;
    for (int s : slots) {
      const int t = pidx ? pidx[s] : 0;
      if (t >= (int)where.size()) where.resize(t + 1, -1);
      if (where[t] < 0) { where[t] = (int)order.size(); order.push_back(t); cols.emplace_back(); }
      cols[where[t]].push_back(s);
    }
  }
  int best = 1;
  double best_cost = 0;
  for (int npg = 1; npg <= 8; npg <<= 1) {
    int64_t ng = 0;
    for (auto& c : cols) ng += ceil_div((int64_t)c.size(), (int64_t)npg);
    const double cost = (double)ng * (12.0 + std::max(64.0, 16.0 * npg));
    if (npg == 1 || cost < best_cost) { best = npg; best_cost = cost; }
  }
  // key: npg, ng, gtab[ng], gslot[ng * npg]
  int ng = 0;
  for (auto& c : cols) ng += (int)ceil_div((int64_t)c.size(), (int64_t)best);
  key.assign(2 + (size_t)ng * (1 + best), -1);
  key[0] = best; key[1] = ng;
  int g = 0;
  for (size_t t = 0; t < cols.size(); ++t)
    for (size_t q = 0; q < cols[t].size(); q += best, ++g) {
      key[2 + g] = order[t];
      for (int j = 0; j < best && q + j < cols[t].size(); ++j) key[2 + ng + (size_t)g * best + j] = cols[t][q + j];
    }
  return ng;
}

// Device copy of the group table; uploaded only when it changes.  The host
// staging is pinned and reused once its previous copy has completed.
static PcGroupDev group_table(PreconSymbolic& Y, const std::vector<int>& key, hipStream_t st) {
  PcGroups& C = Y.groups;
  if (key != C.key) {
    if (C.ev) VH_HIP(hipEventSynchronize(C.ev));
    else VH_HIP(hipEventCreateWithFlags(&C.ev, hipEventDisableTiming));
    if (C.d.n < key.size()) {
      VH_HIP(hipStreamSynchronize(st));   // kernels in flight may still read the old table
      C.d.alloc(key.size());
    }
    C.h.ensure(key.size());
    std::copy(key.begin(), key.end(), C.h.p);
    VH_HIP(hipMemcpyAsync(C.d.p, C.h.p, key.size() * sizeof(int), hipMemcpyHostToDevice, st));
    VH_HIP(hipEventRecord(C.ev, st));
    C.key = key;
  }
  PcGroupDev gr;
  gr.ng = key[1];
  gr.gtab = C.d.p + 2;
  gr.gslot = C.d.p + 2 + key[1];
  return gr;
}

// the largest tier of the block kernel, in staged values (lines or t) of
// max(size, union) x npg; the staging must also fit kBlkLdsMax bytes
// (blk_lds).  (A tier of 8192 values, which takes the bus's level-6/7
// separators into the block kernel, measured 1-2% slower per step than their
// row kernels.)
constexpr int kBlkCap = 2048;
static int ilog2_floor(int v) { int l = 0; while ((2 << l) <= v) ++l; return l; }

// Shared memory of a block-kernel tier: the staged lines and own values
// always; the packed inverses and the entry spans of its blocks when each
// fits kBlkStageCap bytes and the whole stays within kBlkLdsMax.
constexpr size_t kBlkStageCap = 16384, kBlkLdsMax = 65536;
template <typename R, typename WT> static size_t blk_lds_bytes_t(const BlkLds& a, int smax) {
  return (size_t)(a.stg_n + a.tt_n) * sizeof(cplx<WT>) + (size_t)a.inv_n * sizeof(R) +
         (size_t)a.ent_n * (sizeof(R) + sizeof(int)) + 2 * (size_t)smax * sizeof(int);
}
template <typename R, typename WT> static BlkLds blk_lds_t(int umax, int smax, int emax, int npg) {
  BlkLds a;
  a.stg_n = umax * npg;
  a.tt_n = smax * npg;
  const size_t ni = (size_t)smax * (smax + 1) / 2;
  a.inv_n = ni * sizeof(R) <= kBlkStageCap ? (int)ni : 0;
  a.ent_n = (size_t)emax * (sizeof(R) + sizeof(int)) <= kBlkStageCap ? emax : 0;
  if (blk_lds_bytes_t<R, WT>(a, smax) > kBlkLdsMax) a.ent_n = 0;
  if (blk_lds_bytes_t<R, WT>(a, smax) > kBlkLdsMax) a.inv_n = 0;
  return a;
}

template <typename R, typename WT>
void precon_apply_cols_wt(Geometry& G, PreconSymbolic& Y, const PcTab<R>* tab, const int* pidx,
                          const std::vector<int>& slots, const cplx<R>* x, int64_t ldx, cplx<R>* y, int64_t ldy) {
  VH_REQUIRE(!slots.empty(), "preconditioner: no columns");
  hipStream_t st = G.ctx->stream;
  const int nn = (int)G.num_node;
  const int64_t nc = G.num_curr;
  std::vector<int> key;
  build_groups(pidx, slots, key);
  const int npg = key[0];
  const PcGroupDev gr = group_table(Y, key, st);
  const int64_t cols = (int64_t)gr.ng * npg;
  VH_REQUIRE(cols < 65536, "preconditioner: too many columns");
  VH_REQUIRE((int64_t)ceil_div(std::max<int64_t>(nc, nn), 256) * gr.ng < (1LL << 31),
             "preconditioner: grid of the gather / current kernels too large");
  const size_t need = 2 * (size_t)nn * cols * sizeof(cplx<WT>);
  if (Y.ws.n < need) Y.ws.alloc(need);
  cplx<WT>* w = (cplx<WT>*)Y.ws.p;
  cplx<WT>* w2 = w + (size_t)nn * cols;
  const bool d = sizeof(R) == 8;
  const R* arv = d ? (const R*)G.Ar.v64.p : (const R*)G.Ar.v32.p;
  auto by_npg = [&](auto&& f) {
    switch (npg) {
      case 1: f(IC<1>{}); break;
      case 2: f(IC<2>{}); break;
      case 4: f(IC<4>{}); break;
      default: f(IC<8>{}); break;
    }
  };
  if (nn > 0) {
    by_npg([&](auto N) {
      constexpr int n = decltype(N)::value;
      hipLaunchKernelGGL((k_pc_gather<R, WT, n>), dim3((unsigned)(ceil_div(nn, 256) * gr.ng)), dim3(256), 0, st,
                         x, ldx, w, nn, Y.d_iperm.p, G.Ar.ptr.p, G.Ar.col.p, arv, tab, gr, nc);
    });
    VH_LAUNCHED();
    // Blocks whose t fits the shared memory (s * npg <= kBlkCap) go through
    // the block kernel; the rows of larger blocks through the row kernels.
    // (VH_PC_BLOCK: bit 0 forward, bit 1 backward; default both)
    static const int blk_on = [] { const char* e = std::getenv("VH_PC_BLOCK"); return e ? std::atoi(e) : 3; }();
    auto blk_lds = [](int umax, int smax, int emax, int np) { return blk_lds_t<R, WT>(umax, smax, emax, np); };
    auto blk_lds_bytes = [](const BlkLds& a, int smax) { return blk_lds_bytes_t<R, WT>(a, smax); };
    const bool trace_plan = trace_on() && !Y.plan_traced;
    Y.plan_traced = true;
    auto blk_launch = [&](const LevelBlocks& B, int l, bool fwd, int lpr_u, int lpr_g) -> int {
      const int b0 = B.bptr[l], b1 = B.bptr[l + 1];
      if (!(blk_on & (fwd ? 1 : 2))) return B.lptr[l];
      // half the row kernels' lanes per row: the block kernel keeps more rows
      // in flight per wave (measured best on the bus; 64/128-thread
      // workgroups and finer tiers were slower)
      const int llu = std::max(0, std::min(ilog2_floor(lpr_u), ilog2_floor(64 / npg)) - 1);
      const int llg = std::max(0, std::min(ilog2_floor(lpr_g), ilog2_floor(64 / npg)) - 1);
      const BlkInfo* info = fwd ? Y.d_f_info.p : Y.d_b_info.p;
      // eligible: the prefix of the blocks (sorted by max(size, union)) whose
      // t and union fit the shared memory; launched in tiers of that bound
      // so small blocks do not reserve the shared memory of large ones
      int i0 = b0;
      for (const int cap : {256, 1024, kBlkCap}) {
        int i1 = (int)(std::upper_bound(B.key.begin() + i0, B.key.begin() + b1, cap / npg) - B.key.begin());
        // (the prefix maxima grow along the sorted blocks: drop the tail whose
        // staging would not fit; those blocks take the row kernels)
        while (i1 > i0 && blk_lds_bytes(blk_lds(B.umax[i1 - 1], B.smax[i1 - 1], B.emax[i1 - 1], npg), B.smax[i1 - 1]) >
                              kBlkLdsMax)
          --i1;
        if (i1 > i0) {
          const BlkLds lay = blk_lds(B.umax[i1 - 1], B.smax[i1 - 1], B.emax[i1 - 1], npg);
          const size_t lds = blk_lds_bytes(lay, B.smax[i1 - 1]);
          if (trace_plan)
            fprintf(stderr, "[vh] pc %s level %d tier %d: %d blocks x %d groups, s<=%d union<=%d entries<=%d, lds %zu%s%s\n",
                    fwd ? "fwd" : "bwd", l, cap, i1 - i0, gr.ng, B.smax[i1 - 1], B.umax[i1 - 1], B.emax[i1 - 1], lds,
                    lay.inv_n ? " inv" : "", lay.ent_n ? " ent" : "");
          const dim3 grid((unsigned)((int64_t)(i1 - i0) * gr.ng));
          // workgroup size: one thread per value of the tier's largest t (s *
          // npg values), in [64, 256] (half or a quarter of that measured
          // slower on the bus: the staging then takes more rounds per thread)
          int nt = 64;
          while (nt < 256 && nt < B.smax[i1 - 1] * npg) nt <<= 1;
          const int lu = std::min(llu, ilog2_floor(nt / npg)), lg = std::min(llg, ilog2_floor(nt / npg));
          by_npg([&](auto N) {
            constexpr int n = decltype(N)::value;
            auto go = [&](auto ntc) {
              constexpr int NTB = decltype(ntc)::value;
              if (fwd)
                hipLaunchKernelGGL((k_blk_solve<R, WT, n, true, NTB>), grid, dim3(NTB), lds, st, info + i0, Y.d_fp.p,
                                   Y.d_floc.p, Y.d_f_ucol.p, tab, gr, nn, w, w2, lay, lu, lg);
              else
                hipLaunchKernelGGL((k_blk_solve<R, WT, n, false, NTB>), grid, dim3(NTB), lds, st, info + i0,
                                   Y.d_bp.p, Y.d_bloc.p, Y.d_b_ucol.p, tab, gr, nn, w, w2, lay, lu, lg);
            };
            if (nt == 64) go(IC<64>{});
            else if (nt == 128) go(IC<128>{});
            else go(IC<256>{});
          });
          VH_LAUNCHED();
        }
        i0 = i1;
      }
      return B.rowoff[i0];
    };
    const int nlf = (int)Y.f_lptr.size() - 1;
    for (int l = 0; l < nlf; ++l) {
      const int a = blk_launch(Y.fl, l, true, Y.f_lpr_u[l], Y.f_lpr_g[l]), nr = Y.f_lptr[l + 1] - a;
      if (nr == 0) continue;
      if (trace_plan) fprintf(stderr, "[vh] pc fwd level %d: %d rows by the row kernels\n", l, nr);
      const int64_t it = (int64_t)nr * gr.ng;
      if (Y.f_upd_nnz[l] > 0) {
        with_lanes(npg, Y.f_lpr_u[l], [&](auto N, auto L) {
          constexpr int n = decltype(N)::value, lp = decltype(L)::value;
          hipLaunchKernelGGL((k_sn_update_fwd<R, WT, n, lp>), lane_grid(it, n * lp), dim3(256), 0, st,
                             Y.d_f_rows.p + a, nr, Y.d_fp.p, Y.d_fi.p, tab, gr, nn, w, w2);
        });
        VH_LAUNCHED();
      }
      with_lanes(npg, Y.f_lpr_g[l], [&](auto N, auto L) {
        constexpr int n = decltype(N)::value, lp = decltype(L)::value;
        hipLaunchKernelGGL((k_sn_gemv_fwd<R, WT, n, lp>), lane_grid(it, n * lp), dim3(256), 0, st, Y.d_f_rows.p + a,
                           nr, Y.d_sn_of_row.p, Y.d_sn_ptr.p, Y.d_sn_off.p, tab, gr, nn, w, w2);
      });
      VH_LAUNCHED();
    }
    const int nlb = (int)Y.b_lptr.size() - 1;
    for (int l = 0; l < nlb; ++l) {
      const int a = blk_launch(Y.bl, l, false, Y.b_lpr_u[l], Y.b_lpr_g[l]), nr = Y.b_lptr[l + 1] - a;
      if (nr == 0) continue;
      if (trace_plan) fprintf(stderr, "[vh] pc bwd level %d: %d rows by the row kernels\n", l, nr);
      const int64_t it = (int64_t)nr * gr.ng;
      with_lanes(npg, Y.b_lpr_u[l], [&](auto N, auto L) {
        constexpr int n = decltype(N)::value, lp = decltype(L)::value;
        hipLaunchKernelGGL((k_sn_update_bwd<R, WT, n, lp>), lane_grid(it, n * lp), dim3(256), 0, st, Y.d_b_rows.p + a,
                           nr, Y.d_bp.p, Y.d_bi.p, tab, gr, nn, w, w2);
      });
      VH_LAUNCHED();
      with_lanes(npg, Y.b_lpr_g[l], [&](auto N, auto L) {
        constexpr int n = decltype(N)::value, lp = decltype(L)::value;
        hipLaunchKernelGGL((k_sn_gemv_bwd<R, WT, n, lp>), lane_grid(it, n * lp), dim3(256), 0, st, Y.d_b_rows.p + a,
                           nr, Y.d_sn_of_row.p, Y.d_sn_ptr.p, Y.d_sn_off.p, tab, gr, nn, w2, w);
      });
      VH_LAUNCHED();
    }
    by_npg([&](auto N) {
      constexpr int n = decltype(N)::value;
      hipLaunchKernelGGL((k_pc_scatter_nodes<R, WT, n>), dim3((unsigned)(gr.ng * ceil_div(nn, 256 / n))), dim3(256), 0, st,
                         w, y, ldy, nn, Y.d_iperm.p, nc, gr);
    });
    VH_LAUNCHED();
  }
  const R* atv = d ? (const R*)G.ArT.v64.p : (const R*)G.ArT.v32.p;
  by_npg([&](auto N) {
    constexpr int n = decltype(N)::value;
    hipLaunchKernelGGL((k_pc_currents<R, WT, n>), dim3((unsigned)(ceil_div(nc, 256) * gr.ng)), dim3(256), 0,
                       st, x, ldx, w, y, ldy, nn, nc, G.ArT.ptr.p, nn > 0 ? Y.ArTp_col.p : nullptr, atv, tab, gr);
  });
  VH_LAUNCHED();
}

template <typename R>
void precon_apply(Precon& P, const std::vector<int>& slots, const cplx<R>* x, int64_t ldx, cplx<R>* y, int64_t ldy) {
  VH_REQUIRE(P.ready, "preconditioner not prepared");
  if (P.type == PC_SCHUR_GMRES) {
    if constexpr (sizeof(R) == 8) {
      schur_gmres_apply(P, slots, x, ldx, y, ldy, &P.inner_steps);
      return;
    } else {
      throw Error(ERR_ARG, "the schur_gmres preconditioner is applied in FP64 only");
    }
  }
  const PcTab<R>* tab;
  if constexpr (sizeof(R) == 8) tab = P.tab64.p; else tab = P.tab32.p;
  precon_apply_cols<R>(*P.G, *P.sym, tab, nullptr, slots, x, ldx, y, ldy);
}

template <typename R>
void precon_apply_cols(Geometry& G, PreconSymbolic& Y, const PcTab<R>* tab, const int* pidx,
                       const std::vector<int>& slots, const cplx<R>* x, int64_t ldx, cplx<R>* y, int64_t ldy) {
  precon_apply_cols_wt<R, double>(G, Y, tab, pidx, slots, x, ldx, y, ldy);
}

template void precon_apply_cols_wt<float, float>(Geometry&, PreconSymbolic&, const PcTab<float>*, const int*,
                                                 const std::vector<int>&, const float2*, int64_t, float2*, int64_t);
template void precon_apply_cols<double>(Geometry&, PreconSymbolic&, const PcTab<double>*, const int*,
                                        const std::vector<int>&, const double2*, int64_t, double2*, int64_t);
template void precon_apply_cols<float>(Geometry&, PreconSymbolic&, const PcTab<float>*, const int*,
                                       const std::vector<int>&, const float2*, int64_t, float2*, int64_t);
template void precon_apply<double>(Precon&, const std::vector<int>&, const double2*, int64_t, double2*, int64_t);
template void precon_apply<float>(Precon&, const std::vector<int>&, const float2*, int64_t, float2*, int64_t);

}  // namespace vh
