// The schur_gmres preconditioner of lse_sparse_precon_multiply.m:26-46: the
// Schur system S d = b - Ar A_inv a is solved by an inner GMRES (the
// reference calls MATLAB gmres with restart min(100, num_curr-1), maxit 10,
// tolerance prectol and the Jacobi preconditioner diag(S),
// lse_sparse_precon_prepare.m:83-88) whose operator is the matrix-free Schur
// product S v = Ar diag(A_inv) Ar' v (lse_sparse_schur_compl_multiply.m:1-5);
// then c = A_inv (a + Ar' d).  All systems of a batch run the inner solve in
// lock step (gmres_batch).  FP64.
#include "vh_krylov.h"

namespace vh {

struct SchurGmres {
  DevBuf<double> sdiag;        // diag(S)
  DevBuf<double2> t;           // Ar' v scaled by A_inv, per system (num_curr)
  DevBuf<double2> bb, d;       // Schur right-hand sides and solutions, per system (num_node)
  DevBuf<int> oslot;           // outer slot of each system of the apply
  PinnedBuf<int> h_oslot;
  DevBuf<char> arena;          // inner Krylov vectors
  KrylovWork<double> W;
  int cap = 0;                 // systems the buffers hold
  const double* ainv_for = nullptr;   // A_inv the diagonal was computed from
};

// diag(S)_v = sum_k Ar(v,k)^2 A_inv(k)
__global__ void k_schur_diag(int nn, const int* __restrict__ arp, const int* __restrict__ arc,
                             const double* __restrict__ arv, const double* __restrict__ ainv, double* __restrict__ sd) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  double s = 0;
  for (int p = arp[v]; p < arp[v + 1]; ++p) s += arv[p] * arv[p] * ainv[arc[p]];
  sd[v] = s;
}

// t[s][k] = A_inv(k) * sum_v Ar'(k,v) x[s][v]   (system s = slots[blockIdx.y])
__global__ void k_schur_t(int64_t nc, const int* __restrict__ tp, const int* __restrict__ tc,
                          const double* __restrict__ tv, const double* __restrict__ ainv,
                          const double2* __restrict__ x, int64_t ldx, double2* __restrict__ t, int64_t ldt,
                          const int* __restrict__ slots) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nc) return;
  const int s = slots ? slots[blockIdx.y] : blockIdx.y;
  const double2* xs = x + (int64_t)s * ldx;
  double2 a = make_double2(0.0, 0.0);
  for (int p = tp[k]; p < tp[k + 1]; ++p) {
    const double2 v = xs[tc[p]];
    a.x += tv[p] * v.x; a.y += tv[p] * v.y;
  }
  t[(int64_t)s * ldt + k] = make_double2(a.x * ainv[k], a.y * ainv[k]);
}

// y[s][v] = sum_k Ar(v,k) t[s][k]
__global__ void k_schur_y(int nn, const int* __restrict__ arp, const int* __restrict__ arc,
                          const double* __restrict__ arv, const double2* __restrict__ t, int64_t ldt,
                          double2* __restrict__ y, int64_t ldy, const int* __restrict__ slots) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  const int s = slots ? slots[blockIdx.y] : blockIdx.y;
  const double2* ts = t + (int64_t)s * ldt;
  double2 a = make_double2(0.0, 0.0);
  for (int p = arp[v]; p < arp[v + 1]; ++p) {
    const double2 u = ts[arc[p]];
    a.x += arv[p] * u.x; a.y += arv[p] * u.y;
  }
  y[(int64_t)s * ldy + v] = a;
}

// y[s] = x[s] / diag(S)
__global__ void k_schur_jacobi(int nn, const double* __restrict__ sd, const double2* __restrict__ x,
                               double2* __restrict__ y, int64_t ld, const int* __restrict__ slots) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  const int s = slots ? slots[blockIdx.y] : blockIdx.y;
  const double2 a = x[(int64_t)s * ld + v];
  const double r = 1.0 / sd[v];
  y[(int64_t)s * ld + v] = make_double2(a.x * r, a.y * r);
}

// bb[i][v] = b[v] - sum_k Ar(v,k) A_inv(k) a[k] of the outer vector in slot oslot[i]
__global__ void k_schur_rhs(int nn, int64_t nc, const int* __restrict__ arp, const int* __restrict__ arc,
                            const double* __restrict__ arv, const double* __restrict__ ainv,
                            const double2* __restrict__ x, int64_t ldx, const int* __restrict__ oslot,
                            double2* __restrict__ bb) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  const int i = blockIdx.y;
  const double2* xs = x + (int64_t)oslot[i] * ldx;
  double2 a = xs[nc + v];
  for (int p = arp[v]; p < arp[v + 1]; ++p) {
    const int k = arc[p];
    const double c = arv[p] * ainv[k];
    a.x -= c * xs[k].x; a.y -= c * xs[k].y;
  }
  bb[(int64_t)i * nn + v] = a;
}

// y[curr] = (a + Ar' d) A_inv, y[node] = d, into the outer slot oslot[i]
__global__ void k_schur_out(int nn, int64_t nc, const int* __restrict__ tp, const int* __restrict__ tc,
                            const double* __restrict__ tv, const double* __restrict__ ainv,
                            const double2* __restrict__ x, int64_t ldx, const double2* __restrict__ d,
                            const int* __restrict__ oslot, double2* __restrict__ y, int64_t ldy) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (e >= nc + nn) return;
  const int s = oslot[i];
  const double2* ds = d + (int64_t)i * nn;
  if (e < nc) {
    double2 a = x[(int64_t)s * ldx + e];
    for (int p = tp[e]; p < tp[e + 1]; ++p) {
      const double2 u = ds[tc[p]];
      a.x += tv[p] * u.x; a.y += tv[p] * u.y;
    }
    y[(int64_t)s * ldy + e] = make_double2(a.x * ainv[e], a.y * ainv[e]);
  } else {
    y[(int64_t)s * ldy + e] = ds[e - nc];
  }
}

static SchurGmres& schur_state(Precon& P, int nsys) {
  Geometry& G = *P.G;
  if (!P.schur) P.schur = std::make_shared<SchurGmres>();
  SchurGmres& S = *P.schur;
  const int nn = (int)G.num_node;
  const int64_t nc = G.num_curr;
  if (S.ainv_for != P.ainv64.p) {   // prepared (again): diag(S) of the current A_inv
    S.sdiag.alloc(std::max(nn, 1));
    if (nn > 0) {
      hipLaunchKernelGGL(k_schur_diag, dim3(ceil_div(nn, 256)), dim3(256), 0, G.ctx->stream, nn, G.Ar.ptr.p,
                         G.Ar.col.p, G.Ar.v64.p, P.ainv64.p, S.sdiag.p);
      VH_LAUNCHED();
    }
    S.ainv_for = P.ainv64.p;
  }
  if (nsys > S.cap) {
    VH_HIP(hipStreamSynchronize(G.ctx->stream));   // buffers may still be read by queued work
    S.t.alloc((size_t)nsys * std::max<int64_t>(nc, 1));
    S.bb.alloc((size_t)nsys * std::max(nn, 1));
    S.d.alloc((size_t)nsys * std::max(nn, 1));
    S.oslot.alloc(nsys);
    S.h_oslot.alloc(nsys);
    S.cap = nsys;
    S.W.unbind();
  }
  return S;
}

void schur_multiply(Precon& P, int nrhs, const double2* x, int64_t ldx, double2* y, int64_t ldy) {
  VH_REQUIRE(P.ready, "preconditioner not prepared");
  Geometry& G = *P.G;
  VH_REQUIRE(G.has_inc, "geometry has no incidence");
  SchurGmres& S = schur_state(P, nrhs);
  const int nn = (int)G.num_node;
  const int64_t nc = G.num_curr;
  if (nn == 0 || nrhs == 0) return;
  VH_REQUIRE(nrhs < 65536, "too many vectors");
  hipStream_t st = G.ctx->stream;
  hipLaunchKernelGGL(k_schur_t, dim3(ceil_div(nc, 256), nrhs), dim3(256), 0, st, nc, G.ArT.ptr.p, G.ArT.col.p,
                     G.ArT.v64.p, P.ainv64.p, x, ldx, S.t.p, nc, nullptr);
  VH_LAUNCHED();
  hipLaunchKernelGGL(k_schur_y, dim3(ceil_div(nn, 256), nrhs), dim3(256), 0, st, nn, G.Ar.ptr.p, G.Ar.col.p,
                     G.Ar.v64.p, S.t.p, nc, y, ldy, nullptr);
  VH_LAUNCHED();
}

void schur_gmres_apply(Precon& P, const std::vector<int>& slots, const double2* x, int64_t ldx, double2* y,
                       int64_t ldy, int64_t* inner_steps) {
  VH_REQUIRE(P.ready, "preconditioner not prepared");
  Geometry& G = *P.G;
  const int nsys = (int)slots.size();
  VH_REQUIRE(nsys >= 1 && nsys < 65536, "schur_gmres: number of systems out of range");
  SchurGmres& S = schur_state(P, nsys);
  const int nn = (int)G.num_node;
  const int64_t nc = G.num_curr;
  hipStream_t st = G.ctx->stream;
  // outer slots of the systems (pinned staging, copied in stream order)
  VH_HIP(hipStreamSynchronize(st));   // the staging buffer may still feed a queued copy
  for (int i = 0; i < nsys; ++i) S.h_oslot[i] = slots[i];
  VH_HIP(hipMemcpyAsync(S.oslot.p, S.h_oslot.p, nsys * sizeof(int), hipMemcpyHostToDevice, st));
  if (nn > 0) {
    hipLaunchKernelGGL(k_schur_rhs, dim3(ceil_div(nn, 256), nsys), dim3(256), 0, st, nn, nc, G.Ar.ptr.p,
                       G.Ar.col.p, G.Ar.v64.p, P.ainv64.p, x, ldx, S.oslot.p, S.bb.p);
    VH_LAUNCHED();
    // inner solve, systems 0..nsys-1 of bb / d
    SolveOpts o;
    o.restart = (int)std::max<int64_t>(1, std::min<int64_t>(100, std::min<int64_t>(nc - 1, nn)));
    o.maxit = 10;
    o.tol = P.prectol;
    o.final_residual = false;
    o.flexible = false;
    const size_t need = KrylovWork<double>::bytes_needed(nn, o.restart, nsys, false);
    if (S.arena.n < need || S.W.m != o.restart || S.W.nslot < nsys) {
      if (S.arena.n < need) {
        VH_HIP(hipStreamSynchronize(st));
        S.W.unbind();
        S.arena.alloc(need);
      }
      S.W.bind(S.arena.p, nn, nn, o.restart, nsys, false);
    }
    OpFn<double> A = [&](const std::vector<int>& act, const int* sl, const double2* xx, double2* yy) {
      const int na = (int)act.size();
      hipLaunchKernelGGL(k_schur_t, dim3(ceil_div(nc, 256), na), dim3(256), 0, st, nc, G.ArT.ptr.p, G.ArT.col.p,
                         G.ArT.v64.p, P.ainv64.p, xx, (int64_t)nn, S.t.p, nc, sl);
      VH_LAUNCHED();
      hipLaunchKernelGGL(k_schur_y, dim3(ceil_div(nn, 256), na), dim3(256), 0, st, nn, G.Ar.ptr.p, G.Ar.col.p,
                         G.Ar.v64.p, S.t.p, nc, yy, (int64_t)nn, sl);
      VH_LAUNCHED();
    };
    OpFn<double> M = [&](const std::vector<int>& act, const int* sl, const double2* xx, double2* yy) {
      hipLaunchKernelGGL(k_schur_jacobi, dim3(ceil_div(nn, 256), (int)act.size()), dim3(256), 0, st, nn, S.sdiag.p,
                         xx, yy, (int64_t)nn, sl);
      VH_LAUNCHED();
    };
    std::vector<int> sys(nsys);
    for (int i = 0; i < nsys; ++i) sys[i] = i;
    std::vector<SysStats> stats;
    gmres_batch<double>(G.ctx, nn, A, &M, S.bb.p, S.d.p, sys, o, S.W, stats);
    if (inner_steps) {
      int64_t tot = 0;
      for (int i = 0; i < nsys; ++i) tot += stats[i].iters;
      *inner_steps = tot;
    }
  } else if (inner_steps) {
    *inner_steps = 0;
  }
  hipLaunchKernelGGL(k_schur_out, dim3(ceil_div(nc + nn, 256), nsys), dim3(256), 0, st, nn, nc, G.ArT.ptr.p,
                     G.ArT.col.p, G.ArT.v64.p, P.ainv64.p, x, ldx, S.d.p, S.oslot.p, y, ldy);
  VH_LAUNCHED();
}

}  // namespace vh
