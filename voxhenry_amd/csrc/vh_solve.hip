// Batched restarted GMRES drivers (host orchestration of the device kernels).
//
// gmres_batch follows src_iterative_solver/pgmres.m:70-197 / pgmres_iter.m:22-153
// (left preconditioner) or fpgmres.m:58-152 / fpgmres_iter.m:26-121 (flexible
// right preconditioner) for every system of the batch:
//   r = M b (left) | b (flexible); bnorm = ||r|| (left) | ||b||; resvec(1) = ||r||/bnorm
//   while resvec(end) > tol and outit < maxit:
//     V(:,1) = r/||r||; for k = 1..m: w = M A V(:,k) | A M V(:,k) ; CGS ; LS ;
//       stop when ||res|| < tol*bnorm
//     x += V y | Z y ;  r = V(:,1:k+1) res
//   relres = ||b - A x|| / ||b||;  flag = relres > tol
// All systems advance together; one that meets the stopping rule gets its
// update immediately and leaves the active set, so later steps only work on
// the systems still running.  The Hessenberg columns, the Givens QR and the
// least-squares solutions stay on the device; one host synchronisation per
// Arnoldi step brings back one residual norm per active system, from which
// the host decides which systems retire.
#include "vh_krylov.h"
#include <algorithm>
#include <chrono>
#include <cmath>

namespace vh {

template <typename R>
void KrylovWork<R>::bind(char* base, int64_t n_, int64_t ld_, int m_, int nslot_, bool flexible_) {
  const bool resize = (m_ != m || nslot_ != nslot || !h.p);
  n = n_; ld = ld_; m = m_; nslot = std::max(nslot_, 1); flexible = flexible_;
  const size_t vec = (size_t)nslot * ld;
  cplx<R>* p = (cplx<R>*)base;
  V = p; p += (size_t)(m + 1) * vec;
  Z = nullptr;
  if (flexible) { Z = p; p += (size_t)m * vec; }
  w = p; p += vec;
  t = p; p += vec;
  r = p;
  if (!resize) return;
  h.alloc((size_t)nslot * (m + 1));
  h2.alloc((size_t)nslot * (m + 1));
  coef.alloc((size_t)nslot * (m + 1));
  rcoef.alloc((size_t)nslot * (m + 1));
  lsH.alloc((size_t)nslot * (m + 1) * m);
  lsR.alloc((size_t)nslot * m * m);
  lsg.alloc((size_t)nslot * (m + 1));
  lss.alloc((size_t)nslot * m);
  lsc.alloc((size_t)nslot * m);
  lsbeta.alloc(nslot);
  lsres.alloc(nslot);
  part.alloc(VecOps<R>::part_size(m + 1, nslot));
  nrm.alloc(nslot);
  slots_dev.alloc(nslot);
  aux_slots.alloc(nslot);
  need.alloc(nslot);
  h_slots.alloc(nslot);
  h_aux.alloc(nslot);
  h_res.alloc(nslot);
  h_nrm.alloc(nslot);
  h_bn.alloc(nslot);
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t KrylovCache::bytes(int64_t n, int nslot, int m64, bool flex64, int m32, bool flex32, bool mixed_vectors) {
  size_t b = 0;
  if (m64 > 0) b += align256(KrylovWork<double>::bytes_needed(n, m64, nslot, flex64));
  if (m32 > 0) b += align256(KrylovWork<float>::bytes_needed(n, m32, nslot, flex32));
  if (mixed_vectors) b += 2 * align256(sizeof(float2) * (size_t)nslot * n);
  return b;
}

void KrylovCache::layout(int64_t n, int nslot, int m64, bool flex64, int m32, bool flex32, bool mixed_vectors) {
  const size_t need = bytes(n, nslot, m64, flex64, m32, flex32, mixed_vectors);
  if (arena.n < need) {
    w64.unbind(); w32.unbind();
    arena.alloc(need);
  }
  char* p = arena.p;
  if (m64 > 0) {
    w64.bind(p, n, n, m64, nslot, flex64);
    p += align256(KrylovWork<double>::bytes_needed(n, m64, nslot, flex64));
  } else {
    w64.unbind();
  }
  if (m32 > 0) {
    w32.bind(p, n, n, m32, nslot, flex32);
    p += align256(KrylovWork<float>::bytes_needed(n, m32, nslot, flex32));
  } else {
    w32.unbind();
  }
  v32 = z32 = nullptr;
  if (mixed_vectors) {
    v32 = (float2*)p; p += align256(sizeof(float2) * (size_t)nslot * n);
    z32 = (float2*)p;
  }
}

template struct KrylovWork<double>;
template struct KrylovWork<float>;

template <typename R>
void gmres_batch(Context* ctx, int64_t n, const OpFn<R>& A, const OpFn<R>* M, const cplx<R>* b, cplx<R>* x,
                 const std::vector<int>& sys, const SolveOpts& o, KrylovWork<R>& W, std::vector<SysStats>& st) {
  using C = cplx<R>;
  hipStream_t strm = ctx->stream;
  const int nsys = (int)sys.size();
  if (nsys == 0) return;
  const int m = o.restart;
  VH_REQUIRE(m >= 1 && m + 1 <= 128, "restart must be in [1, 127]");
  VH_REQUIRE(o.maxit >= 1, "maxit must be >= 1");
  int nslot = 0;
  for (int s : sys) { VH_REQUIRE(s >= 0, "negative slot"); nslot = std::max(nslot, s + 1); }
  using clk = std::chrono::steady_clock;
  const auto t_start = clk::now();
  double t_wait = 0, t_host = 0;
  int nsteps = 0;
  VH_REQUIRE(W.V && W.ld == n && W.m == m && W.nslot >= nslot && (!o.flexible || W.Z),
             "internal: Krylov workspace not laid out for this solve");
  const double t_alloc = std::chrono::duration<double>(clk::now() - t_start).count();
  VecOps<R> ops(ctx, n, n, m + 1, W.nslot, W.part.p);
  const int64_t vstr = (int64_t)W.nslot * n;
  if ((int)st.size() < nslot) st.resize(nslot);
  const bool left = (M != nullptr) && !o.flexible;
  auto tol_of = [&](int s) { return o.tol_slot ? (*o.tol_slot)[s] : o.tol; };

  // All host<->device transfers of the loop go through pinned staging buffers
  // of the workspace; a staging buffer is rewritten only after the stream has
  // been synchronised since its last copy (pending[] tracks the copies issued
  // since the last synchronisation; a previous call on this workspace may
  // have left one in flight).
  const double2* nbuf = W.h_nrm.p;
  bool pending[2] = {true, true};
  auto sync = [&]() {
    const auto t0 = clk::now();
    VH_HIP(hipStreamSynchronize(strm));
    t_wait += std::chrono::duration<double>(clk::now() - t0).count();
    pending[0] = pending[1] = false;
  };
  auto upload = [&](const std::vector<int>& v, int* dst) {
    const int q = dst == W.aux_slots.p ? 1 : 0;
    int* stage = q ? W.h_aux.p : W.h_slots.p;
    if (pending[q]) sync();
    std::copy(v.begin(), v.end(), stage);
    VH_HIP(hipMemcpyAsync(dst, stage, v.size() * sizeof(int), hipMemcpyHostToDevice, strm));
    pending[q] = true;
  };
  auto fetch_norms = [&](int cnt) {
    VH_HIP(hipMemcpyAsync(W.h_nrm.p, W.nrm.p, cnt * sizeof(double2), hipMemcpyDeviceToHost, strm));
    sync();
  };
  const LsDev ls = W.ls();

  std::vector<int> act = sys;
  upload(act, W.slots_dev.p);
  VH_REQUIRE(!o.x0 || !left, "internal: an initial guess needs a flexible or unpreconditioned solve");
  for (int s : sys) st[s] = SysStats();
  std::vector<double> bnorm(nslot, 0.0), bnorm_true(nslot, 0.0);
  // r = M b (pgmres.m:120-139) or b (fpgmres.m:91-95); with an initial guess
  // (flexible only) r = b - A x0, relative to ||b||
  std::vector<double> r0(nslot, 1.0);
  // the first cycle's residual: b itself for a flexible or unpreconditioned
  // solve from x0 = 0 (no copy), else W.r; later cycles always use W.r
  const C* rsrc = W.r;
  if (o.x0) {
    ops.norm2(b, nsys, W.slots_dev.p, W.nrm.p);
    VH_HIP(hipMemcpyAsync(W.h_bn.p, W.nrm.p, nsys * sizeof(double2), hipMemcpyDeviceToHost, strm));
    A(act, W.slots_dev.p, x, W.r);
    ops.sub_from(b, n, W.r, nsys, W.slots_dev.p);
    for (int s : sys) st[s].matvecs++;
  } else if (left) {
    (*M)(act, W.slots_dev.p, b, W.r);
  } else {
    rsrc = b;
  }
  ops.norm2(rsrc, nsys, W.slots_dev.p, W.nrm.p);
  fetch_norms(nsys);
  std::vector<int> outit(nslot, 0);
  {
    std::vector<int> keep;
    for (int a = 0; a < nsys; ++a) {
      const int s = sys[a];
      if (o.x0) {
        bnorm[s] = std::sqrt(W.h_bn[a].x);
        r0[s] = bnorm[s] > 0 ? std::sqrt(nbuf[a].x) / bnorm[s] : 0.0;
      } else {
        bnorm[s] = std::sqrt(nbuf[a].x);
        r0[s] = bnorm[s] > 0 ? 1.0 : 0.0;
      }
      st[s].resvec.push_back(r0[s]);
      if (bnorm[s] > 0 && r0[s] > tol_of(s)) keep.push_back(s);
    }
    act.swap(keep);
  }
  // x0 = 0: the systems that start no cycle are zeroed here; the others get
  // x = basis * y at their first update (no zero fill to read back)
  std::vector<char> xzero(nslot, 0);
  if (!o.x0) {
    for (int s : act) xzero[s] = 1;
    std::vector<int> idle;
    for (int s : sys) if (!xzero[s]) idle.push_back(s);
    if (!idle.empty()) {
      upload(idle, W.aux_slots.p);
      ops.zero(x, (int)idle.size(), W.aux_slots.p);
    }
  }
  // W.nrm holds ||r||^2 of the systems of sys in order: the first cycle
  // reuses it when every system starts one
  bool norms_fresh = act == sys;

  // x += basis * y for the listed systems (k vectors); optionally r = V res
  // (k+1 vectors), with y and res from the device least-squares state
  auto finish = [&](const std::vector<int>& L, int k, bool restart_r) {
    const int nl = (int)L.size();
    if (nl == 0) return;
    for (int s : L) {
      st[s].inner = k;
      st[s].outer = outit[s];
    }
    // a list is either all first updates (the first cycle, which every
    // system of it starts together) or all later ones
    const bool first = xzero[L[0]] != 0;
    for (int s : L) {
      VH_REQUIRE((xzero[s] != 0) == first, "internal: mixed first and later solution updates");
      xzero[s] = 0;
    }
    upload(L, W.aux_slots.p);
    ls_solve(ctx, ls, W.aux_slots.p, nl, k, m + 1, W.coef.p, restart_r ? W.rcoef.p : nullptr);
    const C* basis = o.flexible ? W.Z : W.V;
    ops.combine(basis, vstr, x, k, nl, W.aux_slots.p, W.coef.p, first ? 0.0 : 1.0, 1.0, nullptr);   // pgmres_iter.m:136-152
    if (restart_r) ops.combine(W.V, vstr, W.r, k + 1, nl, W.aux_slots.p, W.rcoef.p, 0.0, 1.0, nullptr);  // :153
  };

  VH_REQUIRE(!o.true_restart || (o.flexible && o.final_residual), "internal: true_restart needs a flexible solve");
  std::vector<int> fin = sys;   // systems whose true residual the next check computes
  for (;;) {
    while (!act.empty()) {
      for (int s : act) outit[s]++;
      int nact = (int)act.size();
      upload(act, W.slots_dev.p);
      // V(:,1) = r/||r||  (pgmres_iter.m:55); g = ||r|| e1
      if (!norms_fresh) ops.norm2(rsrc, nact, W.slots_dev.p, W.nrm.p);
      norms_fresh = false;
      ops.scale(rsrc, W.V, nact, W.slots_dev.p, W.nrm.p);
      ls_reset(ctx, ls, W.slots_dev.p, nact, W.nrm.p);
      rsrc = W.r;
      int k = 0;
      while (k < m && !act.empty()) {
        ++k;
        nact = (int)act.size();
        const C* vk = W.V + (int64_t)(k - 1) * vstr;
        if (o.flexible) {
          C* zk = W.Z + (int64_t)(k - 1) * vstr;
          if (M) (*M)(act, W.slots_dev.p, vk, zk);
          else convert_vectors<R, R>(ctx, vk, n, zk, n, n, nact, W.slots_dev.p, W.slots_dev.p);
          A(act, W.slots_dev.p, zk, W.w);
        } else {
          if (M) {
            A(act, W.slots_dev.p, vk, W.t);
            (*M)(act, W.slots_dev.p, W.t, W.w);
          } else {
            A(act, W.slots_dev.p, vk, W.w);
          }
        }
        for (int s : act) { st[s].iters++; st[s].matvecs++; }
        // classical Gram-Schmidt, one pass (pgmres_iter.m:107-111).  With reorth
        // (the FP32 inner solve: one pass loses orthogonality once the residual
        // nears sqrt(eps)) a second pass re-projects w.  Up to kFusedMax basis
        // vectors the two passes take three sweeps over the basis (projection;
        // update fused with the second projection; second update fused with
        // the normalisation); beyond, the second pass is selective: it runs for
        // the systems whose first pass cancelled more than half of ||w||^2, the
        // others get a zero correction and keep their norm.
        const bool fused = o.reorth && k <= VecOps<R>::kFusedMax;
        if (fused) {
          ops.mdot(W.V, vstr, W.w, k, nact, W.slots_dev.p, W.h.p);
          ops.combine_dot(W.V, vstr, W.w, k, nact, W.slots_dev.p, W.h.p, W.h2.p);
          ops.combine_scale(W.V, vstr, W.w, W.V + (int64_t)k * vstr, k, nact, W.slots_dev.p, W.h2.p, W.nrm.p);
        } else {
          ops.mdot(W.V, vstr, W.w, k, nact, W.slots_dev.p, W.h.p, nullptr, o.reorth);
          ops.combine(W.V, vstr, W.w, k, nact, W.slots_dev.p, W.h.p, 1.0, -1.0, W.nrm.p);
          if (o.reorth) {
            ops.reorth_flags(W.h.p, k, W.nrm.p, W.need.p, nact);
            ops.mdot(W.V, vstr, W.w, k, nact, W.slots_dev.p, W.h2.p, W.need.p);
            ops.combine(W.V, vstr, W.w, k, nact, W.slots_dev.p, W.h2.p, 1.0, -1.0, W.nrm.p, W.need.p);
          }
          ops.scale(W.w, W.V + (int64_t)k * vstr, nact, W.slots_dev.p, W.nrm.p);
        }
        // column k of H and its Givens rotation (pgmres_iter.m:112-123)
        ls_push(ctx, ls, W.slots_dev.p, nact, k - 1, W.h.p, o.reorth ? W.h2.p : nullptr, m + 1, W.nrm.p, W.lsres.p);
        VH_HIP(hipMemcpyAsync(W.h_res.p, W.lsres.p, (size_t)nact * sizeof(double), hipMemcpyDeviceToHost, strm));
        sync();
        ++nsteps;
        const auto th = clk::now();
        std::vector<int> conv, keep;
        for (int a = 0; a < nact; ++a) {
          const int s = act[a];
          const double rres = W.h_res[a];
          st[s].resvec.push_back(rres / bnorm[s]);
          if (!(rres >= tol_of(s) * bnorm[s])) conv.push_back(s);   // pgmres_iter.m:126 (NaN also stops)
          else if (o.stop_slot && (*o.stop_slot)[s]) conv.push_back(s);
          else keep.push_back(s);
        }
        t_host += std::chrono::duration<double>(clk::now() - th).count();
        if (!conv.empty()) {
          finish(conv, k, false);
          act.swap(keep);
          if (!act.empty()) upload(act, W.slots_dev.p);
        }
      }
      if (!act.empty()) {
        finish(act, k, true);
        std::vector<int> keep;
        for (int s : act) if (outit[s] < o.maxit) keep.push_back(s);
        act.swap(keep);
      }
    }

    // relres = ||b - A x|| / ||b||  (pgmres.m:181-186)
    if (!o.final_residual) {
      for (int s : sys) { st[s].relres = st[s].resvec.back(); st[s].flag = st[s].relres <= tol_of(s) ? 0 : 1; }
      break;
    }
    const int nf = (int)fin.size();
    upload(fin, W.slots_dev.p);
    A(fin, W.slots_dev.p, x, W.t);
    ops.sub_from(b, n, W.t, nf, W.slots_dev.p);
    ops.norm2(W.t, nf, W.slots_dev.p, W.nrm.p);
    fetch_norms(nf);
    const double2* bn = W.h_bn.p;
    ops.norm2(b, nf, W.slots_dev.p, W.nrm.p);
    VH_HIP(hipMemcpyAsync(W.h_bn.p, W.nrm.p, nf * sizeof(double2), hipMemcpyDeviceToHost, strm));
    VH_HIP(hipStreamSynchronize(strm));
    std::vector<int> redo;
    for (int a = 0; a < nf; ++a) {
      const int s = fin[a];
      const double nb = std::sqrt(bn[a].x);
      st[s].relres = nb > 0 ? std::sqrt(nbuf[a].x) / nb : 0.0;
      st[s].flag = (st[s].relres <= tol_of(s)) ? 0 : 1;
      st[s].matvecs++;
      if (o.true_restart && st[s].flag && outit[s] < o.maxit && !(o.stop_slot && (*o.stop_slot)[s]) &&
          std::isfinite(st[s].relres)) {
        redo.push_back(s);
        st[s].resvec.push_back(st[s].relres);
      }
    }
    if (redo.empty()) break;
    // restart those systems from their true residual r = b - A x (in W.t)
    upload(redo, W.slots_dev.p);
    convert_vectors<R, R>(ctx, W.t, n, W.r, n, n, (int)redo.size(), W.slots_dev.p, W.slots_dev.p);
    if (trace_on()) fprintf(stderr, "[vh] gmres: %zu systems restart from their true residual\n", redo.size());
    act = redo;
    fin = redo;
  }
  if (trace_on())
    fprintf(stderr, "[vh] gmres<%s> nsys %d m %d: %d steps, alloc %.3f s, gpu wait %.3f s, host %.3f s, total %.3f s\n",
            sizeof(R) == 8 ? "f64" : "f32", nsys, m, nsteps, t_alloc, t_wait, t_host,
            std::chrono::duration<double>(clk::now() - t_start).count());
}

template void gmres_batch<double>(Context*, int64_t, const OpFn<double>&, const OpFn<double>*, const double2*, double2*,
                                  const std::vector<int>&, const SolveOpts&, KrylovWork<double>&,
                                  std::vector<SysStats>&);
template void gmres_batch<float>(Context*, int64_t, const OpFn<float>&, const OpFn<float>*, const float2*, float2*,
                                 const std::vector<int>&, const SolveOpts&, KrylovWork<float>&, std::vector<SysStats>&);

// Iterative refinement with FP32 GMRES corrections (GMRES-IR): r = b - A x in
// FP64, z = FP32 solution of A z = r (inner_solve writes z32), x += z, until
// ||r|| <= tol ||b||.  One FP64 matvec per step, and the residual it tests is
// the true FP64 residual of the current x.  A system leaves unconverged when
// a step reduces its residual by less than half (its FP32 model of A or its
// inner solve no longer contracts), when its inner solve escalated (stop), or
// after outer_restart * outer_maxit steps.  Per system: iters = refinement
// steps, resvec = relative true residuals (the first one of x0).
template <class Inner>
static void refine_mixed(Context* ctx, int64_t n, const OpFn<double>& A64, const double2* b, double2* x,
                         const std::vector<int>& sys, const MixedOpts& o, KrylovWork<double>& W, float2* v32,
                         const float2* z32, Inner& inner_solve, const std::vector<char>& stop,
                         std::vector<SysStats>& st) {
  hipStream_t strm = ctx->stream;
  int nslot = 0;
  for (int s : sys) nslot = std::max(nslot, s + 1);
  VH_REQUIRE(W.r && W.ld == n && W.nslot >= nslot, "internal: refinement workspace not laid out for this solve");
  VecOps<double> ops(ctx, n, n, W.m + 1, W.nslot, W.part.p);
  if ((int)st.size() < nslot) st.resize(nslot);
  auto tol_of = [&](int s) { return o.tol_slot ? (*o.tol_slot)[s] : o.tol; };
  // slot lists reach the device through the pinned staging of W, rewritten
  // only after the stream has drained (the previous copy may be in flight)
  auto upload = [&](const std::vector<int>& v) {
    VH_HIP(hipStreamSynchronize(strm));
    std::copy(v.begin(), v.end(), W.h_slots.p);
    VH_HIP(hipMemcpyAsync(W.slots_dev.p, W.h_slots.p, v.size() * sizeof(int), hipMemcpyHostToDevice, strm));
  };
  const int ns = (int)sys.size();
  if (ns == 0) return;
  for (int s : sys) st[s] = SysStats();
  upload(sys);
  ops.norm2(b, ns, W.slots_dev.p, W.nrm.p);
  VH_HIP(hipMemcpyAsync(W.h_bn.p, W.nrm.p, ns * sizeof(double2), hipMemcpyDeviceToHost, strm));
  // right-hand side of the next correction: b while x = 0 (converted by the
  // inner solve), then r = b - A x, which residual_to leaves in v32
  const double2* src = b;
  if (o.x0) {
    A64(sys, W.slots_dev.p, x, W.r);
    ops.residual_to<float>(b, W.r, v32, ns, W.slots_dev.p, W.nrm.p);
    VH_HIP(hipMemcpyAsync(W.h_nrm.p, W.nrm.p, ns * sizeof(double2), hipMemcpyDeviceToHost, strm));
    src = nullptr;
  }
  VH_HIP(hipStreamSynchronize(strm));
  std::vector<double> bnorm(nslot, 0.0), rel(nslot, 0.0);
  std::vector<int> act;
  for (int a = 0; a < ns; ++a) {
    const int s = sys[a];
    bnorm[s] = std::sqrt(W.h_bn[a].x);
    if (o.x0) {
      rel[s] = bnorm[s] > 0 ? std::sqrt(W.h_nrm[a].x) / bnorm[s] : 0.0;
      st[s].matvecs++;
    } else {
      rel[s] = bnorm[s] > 0 ? 1.0 : 0.0;
    }
    st[s].resvec.push_back(rel[s]);
    if (rel[s] > tol_of(s)) act.push_back(s);
  }
  // x0 = 0: the first correction is written over x (no zero fill to read
  // back); the systems that take no step are zeroed here
  bool x_zero = !o.x0;
  if (x_zero && (int)act.size() < ns) {
    std::vector<char> on(nslot, 0);
    for (int s : act) on[s] = 1;
    std::vector<int> idle;
    for (int s : sys) if (!on[s]) idle.push_back(s);
    upload(idle);
    ops.zero(x, (int)idle.size(), W.slots_dev.p);
  }
  const int max_steps = std::max(1, o.outer_restart * o.outer_maxit);
  for (int step = 0; step < max_steps && !act.empty(); ++step) {
    const int na = (int)act.size();
    upload(act);
    inner_solve(act, W.slots_dev.p, src);
    if (x_zero) convert_vectors<double, float>(ctx, z32, n, x, n, n, na, W.slots_dev.p, W.slots_dev.p);
    else accumulate_vectors<double, float>(ctx, z32, n, x, n, n, na, W.slots_dev.p);
    x_zero = false;
    A64(act, W.slots_dev.p, x, W.r);
    ops.residual_to<float>(b, W.r, v32, na, W.slots_dev.p, W.nrm.p);
    VH_HIP(hipMemcpyAsync(W.h_nrm.p, W.nrm.p, na * sizeof(double2), hipMemcpyDeviceToHost, strm));
    VH_HIP(hipStreamSynchronize(strm));
    src = nullptr;
    std::vector<int> keep;
    for (int a = 0; a < na; ++a) {
      const int s = act[a];
      const double r = bnorm[s] > 0 ? std::sqrt(W.h_nrm[a].x) / bnorm[s] : 0.0;
      st[s].iters++;
      st[s].matvecs++;
      st[s].resvec.push_back(r);
      const bool contracts = r <= 0.5 * rel[s];   // false for NaN
      rel[s] = r;
      if (r > tol_of(s) && contracts && !stop[s]) keep.push_back(s);
    }
    act.swap(keep);
  }
  for (int s : sys) {
    st[s].relres = rel[s];
    st[s].flag = rel[s] <= tol_of(s) ? 0 : 1;
    st[s].outer = st[s].iters;
  }
  if (trace_on()) {
    int nc = 0;
    for (int s : sys) nc += st[s].flag == 0;
    fprintf(stderr, "[vh] refine: %d of %d systems at tolerance\n", nc, ns);
  }
}

void fgmres_mixed(Context* ctx, int64_t n, const OpFn<double>& A64, const OpFn<float>& A32, const OpFn<float>& M32,
                  const double2* b, double2* x, const std::vector<int>& sys, const MixedOpts& o, KrylovCache& K,
                  std::vector<SysStats>& st) {
  int nslot = 0;
  for (int s : sys) nslot = std::max(nslot, s + 1);
  VH_REQUIRE(K.v32 && K.z32 && K.w64.nslot >= nslot && K.w32.nslot >= nslot,
             "internal: mixed-precision workspace not laid out for this solve");
  float2* v32 = K.v32;
  float2* z32 = K.z32;
  KrylovWork<double>& W64 = K.w64;
  KrylovWork<float>& W32 = K.w32;
  std::vector<SysStats> inner_st;
  std::vector<int> inner_iters(nslot, 0);
  SolveOpts io;
  io.restart = o.inner_restart;
  io.maxit = o.inner_maxit;
  io.tol = o.inner_tol;
  io.final_residual = false;
  io.flexible = o.inner_flexible;
  io.reorth = true;
  // Inner tolerance per system: the outer residual only has to fall from its
  // current value r (relative) to tol, so an inner solve that reduces its
  // own residual by about tol / r suffices (inexact Krylov: the admissible
  // inner error grows as the outer residual falls).  The target keeps a
  // factor 5 of margin below tol: on the bus, margins 20 / 5 / 2 gave steps of
  // 0.641 / 0.605 / 0.580 s with R within 3.7e-8 / 3.4e-7 / 1.2e-6 of FP64
  // solved to 1e-12 (FP64 itself at 1e-8: 6e-7); systems whose R or omega L
  // is a small part of |Z| are refined separately (sweep_batched(rl_rtol)).
  // Clamped to [inner_tol, 0.1]; the first inner solve (r = 1) uses inner_tol.
  std::vector<double> itol(nslot, o.inner_tol);
  io.tol_slot = &itol;
  std::vector<char> stop(nslot, 0);
  // z32 = FP32 GMRES solution of A z = v for the listed systems, each to its
  // inner tolerance (from the outer relative residual, the last entry of
  // st[s].resvec); v == nullptr: the right-hand side is already in v32
  auto inner_solve = [&](const std::vector<int>& act, const int* slots, const double2* v) {
    const int na = (int)act.size();
    for (int s : act) {
      const double r = st[s].resvec.empty() ? 1.0 : st[s].resvec.back();
      const double t = o.tol_slot ? (*o.tol_slot)[s] : o.tol;
      itol[s] = o.adaptive_inner && r > 0 ? std::min(0.1, std::max(o.inner_tol, 0.2 * t / r)) : o.inner_tol;
    }
    if (v) convert_vectors<float, double>(ctx, v, n, v32, n, n, na, slots, slots);
    gmres_batch<float>(ctx, n, A32, &M32, v32, z32, act, io, W32, inner_st);
    for (int s : act) {
      inner_iters[s] += inner_st[s].iters;
      // escalate a system whose inner solve ended above escalate_resid (and
      // above 100x its own tolerance): its FP32 process is stalling, not just
      // slow (measured on the uneven batch: 1e-3 also escalated the thick
      // bars and wires at high frequency, whose FP32 solves finish faster
      // than FP64; 1e-1 left the superconducting strips to crawl in FP32)
      const double ir = inner_st[s].resvec.empty() ? 1.0 : inner_st[s].resvec.back();
      if (o.escalate_inner > 0 && inner_iters[s] >= o.escalate_inner &&
          !(ir <= std::max(100.0 * itol[s], o.escalate_resid)))
        stop[s] = 1;
    }
  };
  if (o.refine) {
    refine_mixed(ctx, n, A64, b, x, sys, o, W64, v32, z32, inner_solve, stop, st);
    for (int s : sys) st[s].inner_iters = inner_iters[s];
    return;
  }
  // preconditioner of the outer iteration: z = (FP32 GMRES solve of A z = v)
  OpFn<double> inner = [&](const std::vector<int>& act, const int* slots, const double2* v, double2* z) {
    inner_solve(act, slots, v);
    convert_vectors<double, float>(ctx, z32, n, z, n, n, (int)act.size(), slots, slots);
  };
  SolveOpts oo;
  oo.restart = o.outer_restart;
  oo.maxit = o.outer_maxit;
  oo.tol = o.tol;
  oo.tol_slot = o.tol_slot;
  oo.final_residual = true;
  oo.flexible = true;
  oo.x0 = o.x0;
  oo.stop_slot = &stop;
  oo.true_restart = true;
  gmres_batch<double>(ctx, n, A64, &inner, b, x, sys, oo, W64, st);
  for (int s : sys) st[s].inner_iters = inner_iters[s];
}

}  // namespace vh
