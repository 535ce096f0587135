// extern "C" boundary of the library (include/voxhenry.h).  Every entry point
// catches C++ exceptions and turns them into a status code plus a thread-local
// message.
#include "vh_geom.h"
#include "vh_krylov.h"
#include "../../include/voxhenry.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

struct vh_ctx_s { vh::Context c; };
struct vh_geom_s {
  vh::Geometry G;
  vh::DevBuf<double2> greal;   // real-space tensor from the last generate call
};
struct vh_precon_s { vh::Precon P; };

namespace {
// Caller host memory is pageable: copies to and from it are blocking copies
// made after the context stream has drained (device to host) or before any
// work that reads the buffer is queued (host to device).
vh::KrylovCache& krylov_cache(vh::Geometry& G) {
  if (!G.krylov) G.krylov = std::make_shared<vh::KrylovCache>();
  return *G.krylov;
}
// Device bytes one system of vh_solve_batch needs: its Krylov workspace and
// its share of the matvec workspaces on the padded grid (the pass-X output A
// of each precision used, Px*M*N per component, and the shared B grid,
// Px*Py*N per component, sized for complex128).
double batch_bytes_per_system(const vh::Geometry& G, int method, int restart, int inner_restart) {
  const bool mixed = method == 1;
  // mixed: the FP32 inner solve is flexible (its basis Z is stored too);
  // method 2 continues in a flexible FP64 solve laid out in the same bytes
  double b = (double)vh::KrylovCache::bytes(G.n(), 1, restart, mixed, mixed ? inner_restart : 0, mixed, mixed);
  // the per-system least-squares state and reduction columns of each Krylov workspace
  b += (double)vh::KrylovWork<double>::small_bytes_per_slot(restart);
  if (mixed) b += (double)vh::KrylovWork<float>::small_bytes_per_slot(inner_restart);
  const vh::GridDims& gd = G.g;
  const double a = 5.0 * gd.Px * (double)gd.M * gd.N, bgrid = 5.0 * gd.Px * (double)gd.Py * gd.N;
  return b + bgrid * 16 + a * (16 + (mixed ? 8 : 0));
}
void h2d(void* dst, const void* src, size_t bytes) { VH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); }
void d2h(void* dst, const void* src, size_t bytes, hipStream_t st) {
  VH_HIP(hipStreamSynchronize(st));
  VH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}
thread_local std::string g_last_error;

template <typename F> int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const vh::Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return vh::ERR_ALLOC;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return vh::ERR_INTERNAL;
  } catch (...) {
    g_last_error = "unknown error";
    return vh::ERR_INTERNAL;
  }
}

template <typename T> void need(T* p, const char* what) {
  if (!p) throw vh::Error(vh::ERR_ARG, std::string("null argument: ") + what);
}

// Stage host or device vectors into device memory of the given precision.
struct Staged {
  vh::DevBuf<char> buf;
  void* p = nullptr;
};
}  // namespace

extern "C" {

const char* vh_last_error(void) { return g_last_error.c_str(); }
int vh_version(void) { return 100; }

int vh_ctx_create(int device, vh_ctx** out) {
  return guard([&] {
    need(out, "out");
    int n = 0;
    VH_HIP(hipGetDeviceCount(&n));
    VH_REQUIRE(device >= 0 && device < n, "device index out of range");
    VH_HIP(hipSetDevice(device));
    auto c = std::make_unique<vh_ctx>();
    c->c.device = device;
    hipDeviceProp_t prop;
    VH_HIP(hipGetDeviceProperties(&prop, device));
    c->c.num_cus = prop.multiProcessorCount;
    c->c.lds_per_block = prop.sharedMemPerBlock;
    vh::scratch_init(device);   // device address of the allocation-free entry point's scratch (device current)
    *out = c.release();
  });
}

int vh_ctx_destroy(vh_ctx* ctx) {
  return guard([&] { delete ctx; });
}

int vh_ctx_set_stream(vh_ctx* ctx, void* s) {
  return guard([&] { need(ctx, "ctx"); ctx->c.stream = (hipStream_t)s; });
}

int vh_ctx_synchronize(vh_ctx* ctx) {
  return guard([&] { need(ctx, "ctx"); VH_HIP(hipStreamSynchronize(ctx->c.stream)); });
}

int vh_geom_create(vh_ctx* ctx, int64_t L, int64_t M, int64_t N, double dx, int64_t nvox, const int64_t* idx,
                   int pad_mode, vh_geom** out) {
  return guard([&] {
    need(ctx, "ctx"); need(idx, "idx"); need(out, "out");
    VH_REQUIRE(L < (1 << 20) && M < (1 << 20) && N < (1 << 20), "grid dims too large");
    auto g = std::make_unique<vh_geom>();
    vh::geometry_init(g->G, &ctx->c, (int)L, (int)M, (int)N, dx, nvox, idx, pad_mode);
    *out = g.release();
  });
}

int vh_geom_destroy(vh_geom* g) {
  return guard([&] { delete g; });
}

int vh_geom_info(const vh_geom* g, int64_t* info) {
  return guard([&] {
    need(g, "g"); need(info, "info");
    const auto& G = g->G;
    info[0] = G.g.nv; info[1] = G.num_curr; info[2] = G.num_node;
    info[3] = G.g.Px; info[4] = G.g.Py; info[5] = G.g.Pz;
    info[6] = G.nlines; info[7] = G.dense; info[8] = G.has_tensor; info[9] = G.has_inc && G.has_z;
  });
}

int vh_geom_generate_green(vh_geom* g, double* st2) {
  return guard([&] {
    need(g, "g");
    vh::generate_realspace_green(g->G, g->greal);
    vh::tensor_from_realspace(g->G, g->greal.p);
    if (st2) std::memcpy(st2, g->G.st2, 6 * sizeof(double));
  });
}

int vh_geom_set_green_spectrum(vh_geom* g, const double* fN, int64_t P0, int64_t P1, int64_t P2, double sre,
                               double sim, const double* st2) {
  return guard([&] {
    need(g, "g"); need(fN, "fN");
    vh::geometry_set_spectrum_full(g->G, fN, P0, P1, P2, sre, sim);
    if (st2) std::memcpy(g->G.st2, st2, 6 * sizeof(double));
  });
}

int vh_geom_set_green_realspace(vh_geom* g, const double* Gh, const double* st2) {
  return guard([&] {
    need(g, "g"); need(Gh, "G");
    const size_t n = (size_t)g->G.g.L * g->G.g.M * g->G.g.N * 7;
    g->greal.alloc(n);
    VH_HIP(hipMemcpy(g->greal.p, Gh, n * sizeof(double2), hipMemcpyHostToDevice));
    vh::tensor_from_realspace(g->G, g->greal.p);
    if (st2) std::memcpy(g->G.st2, st2, 6 * sizeof(double));
    else {
      // G(1,1,1,{1,4,7})
      const size_t V = (size_t)g->G.g.L * g->G.g.M * g->G.g.N;
      for (int k = 0; k < 3; ++k) {
        const int c = k == 0 ? 0 : k == 1 ? 3 : 6;
        g->G.st2[2 * k] = Gh[2 * (c * V)];
        g->G.st2[2 * k + 1] = Gh[2 * (c * V) + 1];
      }
    }
  });
}

int vh_geom_get_green_realspace(vh_geom* g, double* Gh) {
  return guard([&] {
    need(g, "g"); need(Gh, "G");
    VH_REQUIRE(g->greal.p, "no real-space tensor on the device");
    VH_HIP(hipMemcpy(Gh, g->greal.p, g->greal.bytes(), hipMemcpyDeviceToHost));
  });
}

int vh_geom_set_incidence(vh_geom* g, int64_t num_node, int64_t nnz, const int64_t* colptr, const int64_t* rowind,
                          const double* val) {
  return guard([&] {
    need(g, "g"); need(colptr, "colptr");
    VH_REQUIRE(nnz == 0 || (rowind && val), "null CSC arrays");
    vh::geometry_set_incidence_csc(g->G, num_node, nnz, colptr, rowind, val);
  });
}

int vh_geom_build_topology(vh_geom* g, int64_t nports, const int64_t* pP_ptr, const int64_t* P, const int64_t* pN_ptr,
                           const int64_t* N, int64_t nwell, const int64_t* well, int64_t* num_node) {
  return guard([&] {
    need(g, "g");
    VH_REQUIRE(nports >= 0 && nports < 100000, "nports out of range");
    VH_REQUIRE(nports == 0 || (pP_ptr && pN_ptr), "null port pointers");
    std::vector<std::vector<int64_t>> Pf(nports), Nf(nports);
    for (int64_t p = 0; p < nports; ++p) {
      VH_REQUIRE(pP_ptr[p + 1] >= pP_ptr[p] && pN_ptr[p + 1] >= pN_ptr[p], "port pointer arrays not increasing");
      Pf[p].assign(P + 4 * pP_ptr[p], P + 4 * pP_ptr[p + 1]);
      Nf[p].assign(N + 4 * pN_ptr[p], N + 4 * pN_ptr[p + 1]);
    }
    std::vector<int64_t> w;
    if (nwell > 0) { need(well, "well"); w.assign(well, well + 4 * nwell); }
    vh::geometry_build_topology(g->G, Pf, Nf, w, g->G.rhs_h);
    g->G.nports = (int)nports;
    if (num_node) *num_node = g->G.num_node;
  });
}

int vh_geom_get_rhs(vh_geom* g, double* rhs) {
  return guard([&] {
    need(g, "g"); need(rhs, "rhs");
    VH_REQUIRE(!g->G.rhs_h.empty() || g->G.nports == 0, "no topology built");
    std::memcpy(rhs, g->G.rhs_h.data(), g->G.rhs_h.size() * sizeof(double));
  });
}

int vh_geom_set_impedance(vh_geom* g, const double* z) {
  return guard([&] {
    need(g, "g"); need(z, "z");
    vh::geometry_set_impedance(g->G, z);
  });
}

int vh_matvec(vh_geom* g, double jre, double jim, int prec, int mode, int64_t nrhs, const void* x, int64_t ldx,
              void* y, int64_t ldy, int on_device) {
  return guard([&] {
    need(g, "g"); need(x, "x"); need(y, "y");
    auto& G = g->G;
    VH_REQUIRE(prec == VH_C128 || prec == VH_C64, "prec must be VH_C128 or VH_C64");
    VH_REQUIRE(mode >= 0 && mode <= 2, "mode must be 0, 1 or 2");
    const int64_t n = mode == 0 ? G.n() : G.num_curr;
    VH_REQUIRE(ldx >= n && ldy >= n, "leading dimension smaller than the vector length");
    VH_REQUIRE(nrhs >= 1 && nrhs <= 13107, "nrhs out of range");
    const size_t es = prec == VH_C128 ? 16 : 8;
    hipStream_t st = G.ctx->stream;
    const void* dx = x;
    void* dy = y;
    vh::DevBuf<char> bx, by;
    if (!on_device) {
      bx.alloc((size_t)nrhs * ldx * es);
      by.alloc((size_t)nrhs * ldy * es);
      h2d(bx.p, x, bx.bytes());
      VH_HIP(hipMemsetAsync(by.p, 0, by.bytes(), st));
      dx = bx.p;
      dy = by.p;
    }
    if (prec == VH_C128)
      vh::matvec<double>(G, (int)nrhs, (const double2*)dx, ldx, (double2*)dy, ldy, jre, jim, mode);
    else
      vh::matvec<float>(G, (int)nrhs, (const float2*)dx, ldx, (float2*)dy, ldy, jre, jim, mode);
    if (!on_device) {
      d2h(y, by.p, by.bytes(), st);
    }
  });
}

int vh_geom_autotune(vh_geom* g, int prec, int64_t nrhs, int reps, int* best, double* ms_shape, int* nshape) {
  return guard([&] {
    need(g, "g"); need(best, "best");
    auto& G = g->G;
    VH_REQUIRE(prec == VH_C128 || prec == VH_C64, "prec must be VH_C128 or VH_C64");
    VH_REQUIRE(nrhs >= 1 && nrhs <= 13107 && reps >= 1, "nrhs / reps out of range");
    VH_REQUIRE(G.has_tensor, "autotune: Green's tensor not set");
    const int p = prec == VH_C128 ? 0 : 1;
    const int q = prec == VH_C128 ? 1 : 0;    // index of the precision in G.zshape / G.xyshape
    const int nz = vh::matvec_num_zshapes(p), nxy = vh::matvec_num_xyshapes(p);
    const int64_t n = G.num_curr;
    const size_t es = prec == VH_C128 ? 16 : 8;
    vh::DevBuf<char> x((size_t)nrhs * n * es), y((size_t)nrhs * n * es);
    hipStream_t st = G.ctx->stream;
    VH_HIP(hipMemsetAsync(x.p, 0, x.bytes(), st));
    hipEvent_t e0, e1;
    VH_HIP(hipEventCreate(&e0));
    VH_HIP(hipEventCreate(&e1));
    auto run = [&] {
      if (prec == VH_C128)
        vh::matvec<double>(G, (int)nrhs, (const double2*)x.p, n, (double2*)y.p, n, 0.0, 1.0, 2);
      else
        vh::matvec<float>(G, (int)nrhs, (const float2*)x.p, n, (float2*)y.p, n, 0.0, 1.0, 2);
    };
    auto timed = [&]() -> double {
      run();   // warm-up
      VH_HIP(hipEventRecord(e0, st));
      for (int r = 0; r < reps; ++r) run();
      VH_HIP(hipEventRecord(e1, st));
      VH_HIP(hipEventSynchronize(e1));
      float ms = 0;
      VH_HIP(hipEventElapsedTime(&ms, e0, e1));
      return ms / reps;
    };
    auto put = [&](int i, double t) { if (ms_shape && nshape && i < *nshape) ms_shape[i] = t; };
    const int keep_z = G.zshape[q], keep_xy = G.xyshape[q];
    int arg = 0, arg_xy = 0;
    try {
      // pass-Z shapes with the default X/Y shape; shapes that cannot hold a z
      // pencil, or that the grid's kernel ignores, are not timed (a
      // shape-less grid keeps shape 0)
      G.xyshape[q] = 0;
      double tbest = 1e300;
      for (int i = 0; i < nz; ++i) {
        put(i, HUGE_VAL);
        if (!vh::matvec_zshape_applies(G, p, i)) continue;
        G.zshape[q] = i;
        const double t = timed();
        put(i, t);
        if (t < tbest) { tbest = t; arg = i; }
      }
      G.zshape[q] = arg;
      // then the X/Y-pass shapes with that pass-Z shape
      tbest = 1e300;
      for (int i = 0; i < nxy; ++i) {
        put(nz + i, HUGE_VAL);
        if (!vh::matvec_xyshape_applies(G, p, i)) continue;
        G.xyshape[q] = i;
        const double t = timed();
        put(nz + i, t);
        if (t < tbest) { tbest = t; arg_xy = i; }
      }
    } catch (...) {
      G.zshape[q] = keep_z;
      G.xyshape[q] = keep_xy;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    G.zshape[q] = arg;
    G.xyshape[q] = arg_xy;
    *best = arg;
    if (nshape) *nshape = nz + nxy;
  });
}

int vh_geom_set_zshape(vh_geom* g, int prec, int shape) {
  return guard([&] {
    need(g, "g");
    VH_REQUIRE(prec == VH_C128 || prec == VH_C64, "prec must be VH_C128 or VH_C64");
    const int p = prec == VH_C128 ? 0 : 1;
    VH_REQUIRE(shape >= 0 && shape < vh::matvec_num_zshapes(p), "shape index out of range");
    g->G.zshape[prec == VH_C128 ? 1 : 0] = shape;
  });
}

int vh_geom_set_xyshape(vh_geom* g, int prec, int shape) {
  return guard([&] {
    need(g, "g");
    VH_REQUIRE(prec == VH_C128 || prec == VH_C64, "prec must be VH_C128 or VH_C64");
    const int p = prec == VH_C128 ? 0 : 1;
    VH_REQUIRE(shape >= 0 && shape < vh::matvec_num_xyshapes(p), "shape index out of range");
    VH_REQUIRE(vh::matvec_xyshape_applies(g->G, p, shape), "this X/Y-pass shape cannot hold the grid's pencils");
    g->G.xyshape[prec == VH_C128 ? 1 : 0] = shape;
  });
}

int vh_geom_get_shapes(vh_geom* g, int* shapes) {
  return guard([&] {
    need(g, "g"); need(shapes, "shapes");
    shapes[0] = g->G.zshape[0];
    shapes[1] = g->G.zshape[1];
    shapes[2] = g->G.xyshape[0];
    shapes[3] = g->G.xyshape[1];
  });
}

int vh_matvec_noalloc(vh_geom* g, double jre, double jim, int prec, int mode, int64_t nrhs, const void* x,
                      int64_t ldx, void* y, int64_t ldy) {
  return guard([&] {
    need(g, "g"); need(x, "x"); need(y, "y");
    auto& G = g->G;
    VH_REQUIRE(prec == VH_C128 || prec == VH_C64, "prec must be VH_C128 or VH_C64");
    VH_REQUIRE(mode >= 0 && mode <= 2, "mode must be 0, 1 or 2");
    const int64_t n = mode == 0 ? G.n() : G.num_curr;
    VH_REQUIRE(ldx >= n && ldy >= n, "leading dimension smaller than the vector length");
    VH_REQUIRE(nrhs >= 1 && nrhs <= 13107, "nrhs out of range");
    const size_t need_b = vh::matvec_ws_bytes(G, prec == VH_C128 ? 0 : 1, (int)nrhs);
    if (need_b > vh::scratch_bytes())
      throw vh::Error(vh::ERR_ARG, "vh_matvec_noalloc: " + std::to_string(nrhs) + " vectors of this geometry need " +
                                       std::to_string(need_b) + " bytes of scratch, more than the fixed " +
                                       std::to_string(vh::scratch_bytes()) +
                                       " (VH_MATVEC_NOALLOC_BYTES): use fewer vectors per call or vh_matvec");
    char* ws = vh::scratch_ptr(G.ctx->device);
    if (prec == VH_C128)
      vh::matvec_ws<double>(G, (int)nrhs, (const double2*)x, ldx, (double2*)y, ldy, jre, jim, mode, ws);
    else
      vh::matvec_ws<float>(G, (int)nrhs, (const float2*)x, ldx, (float2*)y, ldy, jre, jim, mode, ws);
  });
}

int vh_matvec_timed(vh_geom* g, double jre, double jim, int prec, int mode, int64_t nrhs, const void* x, int64_t ldx,
                    void* y, int64_t ldy, int reps, double* ms_total, double* ms_pass, int* npass) {
  return guard([&] {
    need(g, "g"); need(x, "x"); need(y, "y"); need(ms_total, "ms_total");
    auto& G = g->G;
    VH_REQUIRE(reps >= 1, "reps must be >= 1");
    hipStream_t st = G.ctx->stream;
    hipEvent_t e0, e1;
    VH_HIP(hipEventCreate(&e0));
    VH_HIP(hipEventCreate(&e1));
    VH_HIP(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) {
      if (prec == VH_C128)
        vh::matvec<double>(G, (int)nrhs, (const double2*)x, ldx, (double2*)y, ldy, jre, jim, mode);
      else
        vh::matvec<float>(G, (int)nrhs, (const float2*)x, ldx, (float2*)y, ldy, jre, jim, mode);
    }
    VH_HIP(hipEventRecord(e1, st));
    VH_HIP(hipEventSynchronize(e1));
    float ms = 0;
    VH_HIP(hipEventElapsedTime(&ms, e0, e1));
    *ms_total = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_pass) {
      // second sweep with per-pass events (events between passes add no device work)
      vh::PassTimer tm;
      tm.create();
      G.timer = &tm;
      try {
        for (int r = 0; r < reps; ++r) {
          if (prec == VH_C128)
            vh::matvec<double>(G, (int)nrhs, (const double2*)x, ldx, (double2*)y, ldy, jre, jim, mode);
          else
            vh::matvec<float>(G, (int)nrhs, (const float2*)x, ldx, (float2*)y, ldy, jre, jim, mode);
        }
        tm.collect();
      } catch (...) {
        G.timer = nullptr;
        tm.destroy();
        throw;
      }
      G.timer = nullptr;
      for (int k = 0; k < vh::PassTimer::NP; ++k) ms_pass[k] = tm.ms[k] / std::max(1, tm.calls);
      tm.destroy();
      if (npass) *npass = vh::PassTimer::NP;
    } else if (npass) {
      *npass = 0;
    }
  });
}

int vh_precon_create(vh_geom* g, vh_precon** out) {
  return guard([&] {
    need(g, "g"); need(out, "out");
    auto p = std::make_unique<vh_precon>();
    p->P.G = &g->G;
    *out = p.release();
  });
}

int vh_precon_destroy(vh_precon* p) {
  return guard([&] { delete p; });
}

int vh_precon_prepare(vh_precon* p, double jre, double jim, int64_t* stats) {
  return guard([&] {
    need(p, "p");
    vh::precon_prepare(p->P, jre, jim);
    if (stats) {
      stats[0] = p->P.sym->nnzL;
      stats[1] = (int64_t)p->P.sym->f_lptr.size() - 1;
      stats[2] = (int64_t)p->P.sym->b_lptr.size() - 1;
      stats[3] = p->P.G->num_node;
    }
  });
}

int vh_precon_prepare_many(vh_precon* const* ps, int64_t n, const double* jwmu, const double* z, int nthreads) {
  return guard([&] {
    need(ps, "ps"); need(jwmu, "jwmu");
    VH_REQUIRE(n >= 1 && n <= (1 << 20), "n out of range");
    std::vector<vh::Precon*> P(n);
    for (int64_t k = 0; k < n; ++k) {
      VH_REQUIRE(ps[k] && ps[k]->P.G == ps[0]->P.G, "all preconditioners must belong to one geometry");
      P[k] = &ps[k]->P;
    }
    vh::precon_prepare_many(P.data(), (int)n, (const std::complex<double>*)jwmu, (const std::complex<double>*)z,
                            nthreads);
  });
}

int vh_precon_get_ainv(vh_precon* p, double* ainv) {
  return guard([&] {
    need(p, "p"); need(ainv, "ainv");
    VH_REQUIRE(p->P.ready, "preconditioner not prepared");
    std::memcpy(ainv, p->P.ainv_h.data(), p->P.ainv_h.size() * sizeof(double));
  });
}

namespace {
// Host or device staging of nrhs vectors of n complex values of es bytes.
struct Io {
  vh::DevBuf<char> bx, by;
  const void* x = nullptr;
  void* y = nullptr;
  void in(const void* hx, int64_t nrhs, int64_t ldx, size_t es, int on_device, hipStream_t st, int64_t ldy) {
    if (on_device) { x = hx; return; }
    bx.alloc((size_t)nrhs * ldx * es);
    h2d(bx.p, hx, bx.bytes());
    by.alloc((size_t)nrhs * ldy * es);
    VH_HIP(hipMemsetAsync(by.p, 0, by.bytes(), st));
    x = bx.p;
    y = by.p;
  }
};
}  // namespace

int vh_precon_apply(vh_precon* p, int prec, int64_t nrhs, const void* x, int64_t ldx, void* y, int64_t ldy,
                    int on_device) {
  return guard([&] {
    need(p, "p"); need(x, "x"); need(y, "y");
    auto& G = *p->P.G;
    VH_REQUIRE(prec == VH_C128 || prec == VH_C64, "prec must be VH_C128 or VH_C64");
    VH_REQUIRE(ldx >= G.n() && ldy >= G.n(), "leading dimension smaller than the vector length");
    VH_REQUIRE(nrhs >= 1 && nrhs < (1 << 16), "nrhs out of range");
    const size_t es = prec == VH_C128 ? 16 : 8;
    hipStream_t st = G.ctx->stream;
    Io io;
    io.in(x, nrhs, ldx, es, on_device, st, ldy);
    void* dy = on_device ? y : io.y;
    std::vector<int> slots(nrhs);
    for (int64_t i = 0; i < nrhs; ++i) slots[i] = (int)i;
    if (prec == VH_C128)
      vh::precon_apply<double>(p->P, slots, (const double2*)io.x, ldx, (double2*)dy, ldy);
    else
      vh::precon_apply<float>(p->P, slots, (const float2*)io.x, ldx, (float2*)dy, ldy);
    if (!on_device) {
      d2h(y, io.y, (size_t)nrhs * ldy * es, st);
    }
    VH_HIP(hipStreamSynchronize(st));
  });
}


int vh_precon_set_type(vh_precon* p, int type, double prectol) {
  return guard([&] {
    need(p, "p");
    VH_REQUIRE(type == VH_PC_SCHUR_INVERT || type == VH_PC_SCHUR_GMRES, "unknown preconditioner type");
    VH_REQUIRE(prectol > 0 && prectol < 1, "prectol must be in (0, 1)");
    p->P.type = type;
    p->P.prectol = prectol;
  });
}

int vh_precon_inner_steps(vh_precon* p, int64_t* steps) {
  return guard([&] {
    need(p, "p"); need(steps, "steps");
    *steps = p->P.inner_steps;
  });
}

int vh_precon_schur_multiply(vh_precon* p, int64_t nrhs, const void* x, int64_t ldx, void* y, int64_t ldy,
                             int on_device) {
  return guard([&] {
    need(p, "p"); need(x, "x"); need(y, "y");
    VH_REQUIRE(p->P.ready, "preconditioner not prepared");
    auto& G = *p->P.G;
    const int64_t nn = G.num_node;
    VH_REQUIRE(ldx >= nn && ldy >= nn, "leading dimension smaller than the number of nodes");
    VH_REQUIRE(nrhs >= 1 && nrhs < (1 << 16), "nrhs out of range");
    hipStream_t st = G.ctx->stream;
    Io io;
    io.in(x, nrhs, ldx, 16, on_device, st, ldy);
    void* dy = on_device ? y : io.y;
    vh::schur_multiply(p->P, (int)nrhs, (const double2*)io.x, ldx, (double2*)dy, ldy);
    if (!on_device) d2h(y, io.y, (size_t)nrhs * ldy * 16, st);
    VH_HIP(hipStreamSynchronize(st));
  });
}

int vh_geom_currents_on_grid(vh_geom* g, int64_t nrhs, const void* x, int64_t ldx, void* grid, int64_t ldg,
                             double* minmax, int on_device) {
  return guard([&] {
    need(g, "g"); need(x, "x"); need(grid, "grid");
    auto& G = g->G;
    hipStream_t st = G.ctx->stream;
    Io io;
    io.in(x, nrhs, ldx, 16, on_device, st, ldg);
    void* dg = on_device ? grid : io.y;
    vh::currents_on_grid(G, (int)nrhs, (const double2*)io.x, ldx, (double2*)dg, ldg, minmax);
    if (!on_device) d2h(grid, io.y, (size_t)nrhs * ldg * 16, st);
    VH_HIP(hipStreamSynchronize(st));
  });
}

int vh_gmres(vh_geom* g, vh_precon* p, double jre, double jim, int prec, int flexible, int64_t nrhs, const void* b,
             void* x, int on_device, int restart, double tol, int maxit, int* flag, double* relres, int* iter,
             int* niters, double* resvec, int64_t resvec_cap, int* nresvec) {
  return guard([&] {
    need(g, "g"); need(b, "b"); need(x, "x");
    auto& G = g->G;
    VH_REQUIRE(G.has_tensor && G.has_inc && G.has_z, "geometry not complete (tensor, incidence, impedance)");
    VH_REQUIRE(!p || p->P.ready, "preconditioner not prepared");
    VH_REQUIRE(prec == VH_C128 || prec == VH_C64, "prec must be VH_C128 or VH_C64");
    VH_REQUIRE(nrhs >= 1 && nrhs <= 13107, "nrhs out of range");
    VH_REQUIRE(tol > 0 && restart >= 1 && restart <= 127 && maxit >= 1, "bad solver parameters");
    const int64_t n = G.n();
    const size_t es = prec == VH_C128 ? 16 : 8;
    hipStream_t st = G.ctx->stream;
    Io io;
    io.in(b, nrhs, n, es, on_device, st, n);
    void* dx = on_device ? x : io.y;
    std::vector<int> sys(nrhs);
    for (int i = 0; i < nrhs; ++i) sys[i] = i;
    vh::SolveOpts o;
    o.restart = restart; o.maxit = maxit; o.tol = tol; o.flexible = flexible != 0; o.final_residual = true;
    std::vector<vh::SysStats> stats;
    vh::KrylovCache& K = krylov_cache(G);
    if (prec == VH_C128) {
      K.layout(n, (int)nrhs, restart, o.flexible, 0, false, false);
      vh::KrylovWork<double>& W = K.w64;
      vh::OpFn<double> A = [&](const std::vector<int>& act, const int* sl, const double2* xx, double2* yy) {
        vh::matvec<double>(G, (int)act.size(), xx, n, yy, n, jre, jim, 0, sl);
      };
      vh::OpFn<double> M = [&](const std::vector<int>& act, const int* sl, const double2* xx, double2* yy) {
        vh::precon_apply<double>(p->P, act, xx, n, yy, n);
      };
      vh::gmres_batch<double>(G.ctx, n, A, p ? &M : nullptr, (const double2*)io.x, (double2*)dx, sys, o, W, stats);
    } else {
      K.layout(n, (int)nrhs, 0, false, restart, o.flexible, false);
      vh::KrylovWork<float>& W = K.w32;
      vh::OpFn<float> A = [&](const std::vector<int>& act, const int* sl, const float2* xx, float2* yy) {
        vh::matvec<float>(G, (int)act.size(), xx, n, yy, n, jre, jim, 0, sl);
      };
      vh::OpFn<float> M = [&](const std::vector<int>& act, const int* sl, const float2* xx, float2* yy) {
        vh::precon_apply<float>(p->P, act, xx, n, yy, n);
      };
      vh::gmres_batch<float>(G.ctx, n, A, p ? &M : nullptr, (const float2*)io.x, (float2*)dx, sys, o, W, stats);
    }
    if (!on_device) d2h(x, io.y, (size_t)nrhs * n * es, st);
    VH_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < nrhs; ++i) {
      const auto& s = stats[i];
      if (flag) flag[i] = s.flag;
      if (relres) relres[i] = s.relres;
      if (iter) { iter[2 * i] = s.inner; iter[2 * i + 1] = s.outer; }
      if (niters) niters[i] = s.iters;
      if (nresvec) nresvec[i] = (int)s.resvec.size();
      if (resvec)
        for (int64_t k = 0; k < resvec_cap && k < (int64_t)s.resvec.size(); ++k) resvec[i * resvec_cap + k] = s.resvec[k];
    }
  });
}

int vh_solve_mixed(vh_geom* g, vh_precon* p, double jre, double jim, int64_t nrhs, const double* b, double* x,
                   int on_device, double tol, int outer_restart, int outer_maxit, int inner_restart, double inner_tol,
                   int* flag, double* relres, int* outer_iters, int* inner_iters) {
  return guard([&] {
    need(g, "g"); need(p, "p"); need(b, "b"); need(x, "x");
    auto& G = g->G;
    VH_REQUIRE(G.has_tensor && G.has_inc && G.has_z, "geometry not complete (tensor, incidence, impedance)");
    VH_REQUIRE(p->P.ready, "preconditioner not prepared");
    VH_REQUIRE(nrhs >= 1 && nrhs <= 13107, "nrhs out of range");
    VH_REQUIRE(outer_restart >= 1 && outer_restart <= 127 && inner_restart >= 1 && inner_restart <= 127, "bad restart");
    const int64_t n = G.n();
    hipStream_t st = G.ctx->stream;
    Io io;
    io.in(b, nrhs, n, 16, on_device, st, n);
    void* dx = on_device ? (void*)x : io.y;
    std::vector<int> sys(nrhs);
    for (int i = 0; i < nrhs; ++i) sys[i] = i;
    vh::MixedOpts o;
    o.tol = tol; o.outer_restart = outer_restart; o.outer_maxit = outer_maxit;
    o.inner_restart = inner_restart; o.inner_tol = inner_tol;
    vh::OpFn<double> A64 = [&](const std::vector<int>& act, const int* sl, const double2* xx, double2* yy) {
      vh::matvec<double>(G, (int)act.size(), xx, n, yy, n, jre, jim, 0, sl);
    };
    vh::OpFn<float> A32 = [&](const std::vector<int>& act, const int* sl, const float2* xx, float2* yy) {
      vh::matvec<float>(G, (int)act.size(), xx, n, yy, n, jre, jim, 0, sl);
    };
    vh::OpFn<float> M32 = [&](const std::vector<int>& act, const int* sl, const float2* xx, float2* yy) {
      vh::precon_apply<float>(p->P, act, xx, n, yy, n);
    };
    std::vector<vh::SysStats> stats;
    vh::KrylovCache& K = krylov_cache(G);
    K.layout(n, (int)nrhs, outer_restart, true, inner_restart, false, true);
    vh::fgmres_mixed(G.ctx, n, A64, A32, M32, (const double2*)io.x, (double2*)dx, sys, o, K, stats);
    if (!on_device) d2h(x, io.y, (size_t)nrhs * n * 16, st);
    VH_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < nrhs; ++i) {
      if (flag) flag[i] = stats[i].flag;
      if (relres) relres[i] = stats[i].relres;
      if (outer_iters) outer_iters[i] = stats[i].iters;
      if (inner_iters) inner_iters[i] = stats[i].inner_iters;
    }
  });
}

int vh_ldl_solve_host(int64_t n, const int64_t* Ap, const int64_t* Ai, const double* Ax, int64_t nrhs,
                      const double* b, double* x, const float* xyz, const int* type, int64_t* stats) {
  return guard([&] {
    need(Ap, "Ap"); need(Ai, "Ai"); need(Ax, "Ax");
    VH_REQUIRE(n >= 0 && n < (1LL << 31) && Ap[n] < (1LL << 31), "matrix too large for 32-bit indices");
    std::vector<int> p(n + 1), i(Ap[n]);
    for (int64_t k = 0; k <= n; ++k) p[k] = (int)Ap[k];
    for (int64_t k = 0; k < Ap[n]; ++k) {
      VH_REQUIRE(Ai[k] >= 0 && Ai[k] < n, "row index out of range");
      i[k] = (int)Ai[k];
    }
    vh::LdlFactor F;
    std::vector<int> perm = (xyz && type) ? vh::nd_ordering_geometric((int)n, p.data(), i.data(), xyz, type)
                                          : vh::nd_ordering((int)n, p.data(), i.data());
    vh::ldl_symbolic((int)n, p.data(), i.data(), perm, F);
    try {
      vh::ldl_numeric((int)n, p.data(), i.data(), Ax, F);
    } catch (const std::runtime_error& e) {
      throw vh::Error(vh::ERR_NUMERIC, e.what());
    }
    std::vector<int> Rp, Ri, lf, lb;
    std::vector<double> Rx;
    vh::ldl_levels(F, Rp, Ri, Rx, lf, lb);
    if (stats) {
      stats[0] = F.Lp[n];
      stats[1] = n ? *std::max_element(lf.begin(), lf.end()) + 1 : 0;
      stats[2] = n ? *std::max_element(lb.begin(), lb.end()) + 1 : 0;
    }
    if (b && x) {
      std::vector<double> w(n);
      for (int64_t r = 0; r < nrhs; ++r) {
        const double* br = b + r * n;
        for (int64_t k = 0; k < n; ++k) w[k] = br[F.perm[k]];
        for (int64_t j = 0; j < n; ++j)               // L w = b
          for (int q = F.Lp[j]; q < F.Lp[j + 1]; ++q) w[F.Li[q]] -= F.Lx[q] * w[j];
        for (int64_t k = 0; k < n; ++k) w[k] /= F.D[k];
        for (int64_t j = n - 1; j >= 0; --j)          // L' x = w
          for (int q = F.Lp[j]; q < F.Lp[j + 1]; ++q) w[j] -= F.Lx[q] * w[F.Li[q]];
        double* xr = x + r * n;
        for (int64_t k = 0; k < n; ++k) xr[F.perm[k]] = w[k];
      }
    }
  });
}

int vh_solve_batch(vh_geom* g, int64_t nprecon, vh_precon* const* precons, int64_t nsys, const double* jwmu,
                   const int* pidx, const double* b, double* x, int on_device, int x_init, int method, double tol,
                   const double* tols, int restart,
                   int maxit, int inner_restart, double inner_tol, double hbm_budget, int* flag, double* relres,
                   int* iters, int* inner_iters, double* resid_est) {
  return guard([&] {
    need(g, "g"); need(precons, "precons"); need(jwmu, "jwmu"); need(pidx, "pidx"); need(b, "b"); need(x, "x");
    auto& G = g->G;
    VH_REQUIRE(G.has_tensor && G.has_inc, "geometry not complete (tensor, incidence)");
    VH_REQUIRE(nprecon >= 1 && nsys >= 1 && nsys <= 13107, "nprecon / nsys out of range");
    VH_REQUIRE(method >= 0 && method <= 2, "method must be 0 (fp64), 1 (mixed) or 2 (fp64 to the true residual)");
    VH_REQUIRE(!x_init || method == 1, "an initial guess (x_init) needs method 1 (the pgmres semantics start from 0)");
    VH_REQUIRE(restart >= 1 && restart <= 127 && inner_restart >= 1 && inner_restart <= 127 && maxit >= 1,
               "bad restart / maxit");
    for (int64_t k = 0; k < nprecon; ++k) VH_REQUIRE(precons[k] && precons[k]->P.ready, "preconditioner not prepared");
    for (int64_t k = 0; k < nprecon; ++k)
      VH_REQUIRE(precons[k]->P.type == vh::PC_SCHUR_INVERT, "the batched solve takes schur_invert preconditioners");
    for (int64_t s = 0; s < nsys; ++s) VH_REQUIRE(pidx[s] >= 0 && pidx[s] < nprecon, "pidx out of range");
    VH_REQUIRE(tol > 0, "tol must be positive");
    if (tols)
      for (int64_t s = 0; s < nsys; ++s) VH_REQUIRE(tols[s] > 0, "tols must be positive");
    const int64_t n = G.n();
    hipStream_t st = G.ctx->stream;
    Io io;
    io.in(b, nsys, n, 16, on_device, st, n);
    double2* dx = (double2*)(on_device ? (void*)x : io.y);
    const double2* db = (const double2*)io.x;
    if (x_init && !on_device) {
      VH_HIP(hipStreamSynchronize(st));   // the clearing of io.y is queued on st
      h2d(dx, x, (size_t)nsys * n * 16);
    }
    // per-slot j*omega*mu on the device
    vh::DevBuf<double2> alphas(nsys);
    h2d(alphas.p, jwmu, nsys * sizeof(double2));
    // factor tables of the preconditioners and the preconditioner of every system
    vh::PreconSymbolic* sym = G.precon_sym.get();
    std::vector<vh::PcTab<double>> t64(nprecon);
    std::vector<vh::PcTab<float>> t32(nprecon);
    for (int64_t k = 0; k < nprecon; ++k) {
      VH_REQUIRE(precons[k]->P.G == &G && precons[k]->P.sym.get() == sym,
                 "every preconditioner must belong to this geometry");
      t64[k] = precons[k]->P.tab64_h;
      t32[k] = precons[k]->P.tab32_h;
    }
    vh::DevBuf<vh::PcTab<double>> tabs64(nprecon);
    vh::DevBuf<vh::PcTab<float>> tabs32(nprecon);
    // impedance of every system: the one its preconditioner was prepared with
    // (frequency-dependent for superconductors)
    std::vector<const double2*> zp64(nsys);
    std::vector<const float2*> zp32(nsys);
    for (int64_t k = 0; k < nsys; ++k) {
      zp64[k] = precons[pidx[k]]->P.z64.p;
      zp32[k] = precons[pidx[k]]->P.z32.p;
    }
    vh::DevBuf<const double2*> zs64(nsys);
    vh::DevBuf<const float2*> zs32(nsys);
    VH_HIP(hipMemcpy(zs64.p, zp64.data(), nsys * sizeof(zp64[0]), hipMemcpyHostToDevice));
    VH_HIP(hipMemcpy(zs32.p, zp32.data(), nsys * sizeof(zp32[0]), hipMemcpyHostToDevice));
    VH_HIP(hipMemcpy(tabs64.p, t64.data(), nprecon * sizeof(t64[0]), hipMemcpyHostToDevice));
    VH_HIP(hipMemcpy(tabs32.p, t32.data(), nprecon * sizeof(t32[0]), hipMemcpyHostToDevice));
    // HBM budget -> chunk size.  The geometry's workspace arena counts as
    // available (it is reused), so the chunk size, and with it the arena,
    // is the same on every call with the same sizes.
    vh::KrylovCache& K = krylov_cache(G);
    size_t freeb = 0, totalb = 0;
    VH_HIP(hipMemGetInfo(&freeb, &totalb));
    const double reusable = (double)K.arena.n + (double)G.wsA64.n + (double)G.wsA32.n + (double)G.wsB.n;
    const double budget = hbm_budget > 0 ? hbm_budget : 0.8 * ((double)freeb + reusable);
    const bool mixed = method == 1;
    const double per_sys = batch_bytes_per_system(G, method, restart, inner_restart);
    int chunk = (int)std::max(1.0, std::min((double)nsys, std::floor(budget / per_sys)));
    std::vector<vh::SysStats> stats(nsys);
    if (vh::trace_on()) fprintf(stderr, "[vh] solve_batch nsys %ld chunk %d budget %.1f GB per_sys %.3f GB\n",
                                (long)nsys, chunk, budget / 1e9, per_sys / 1e9);
    K.layout(n, chunk, restart, mixed, mixed ? inner_restart : 0, mixed, mixed);
    vh::KrylovWork<double>& W64 = K.w64;
    // each chunk is solved with chunk-relative slots 0..cnt-1 (workspaces laid
    // out for the chunk); coff is the chunk's first system
    int64_t coff = 0;
    vh::OpFn<double> A64 = [&](const std::vector<int>& act, const int* sl, const double2* xx, double2* yy) {
      vh::matvec<double>(G, (int)act.size(), xx, n, yy, n, 0.0, 0.0, 0, sl, alphas.p + coff, zs64.p + coff);
    };
    vh::OpFn<float> A32 = [&](const std::vector<int>& act, const int* sl, const float2* xx, float2* yy) {
      vh::matvec<float>(G, (int)act.size(), xx, n, yy, n, 0.0, 0.0, 0, sl, alphas.p + coff, zs32.p + coff);
    };
    // preconditioner of every active system of the chunk in one pass (the
    // system's preconditioner is pidx[coff + chunk-relative slot])
    vh::OpFn<double> M64 = [&](const std::vector<int>& act, const int* sl, const double2* xx, double2* yy) {
      vh::precon_apply_cols<double>(G, *sym, tabs64.p, pidx + coff, act, xx, n, yy, n);
    };
    // inner preconditioner of the mixed solve: FP32 factors on FP32 work
    // vectors, exact to FP32 rounding per level (not a fixed linear
    // operator), inside a flexible inner GMRES
    vh::OpFn<float> M32 = [&](const std::vector<int>& act, const int* sl, const float2* xx, float2* yy) {
      vh::precon_apply_cols_wt<float, float>(G, *sym, tabs32.p, pidx + coff, act, xx, n, yy, n);
    };
    for (int64_t s0 = 0; s0 < nsys; s0 += chunk) {
      coff = s0;
      const int cnt = (int)std::min<int64_t>(chunk, nsys - s0);
      std::vector<int> sys(cnt);
      for (int i = 0; i < cnt; ++i) sys[i] = i;
      std::vector<vh::SysStats> cs;
      std::vector<double> tolc;
      if (tols) tolc.assign(tols + s0, tols + s0 + cnt);
      const double2* bc = db + s0 * n;
      double2* xc = dx + s0 * n;
      if (method != 1) {
        vh::SolveOpts o;
        o.restart = restart; o.maxit = maxit; o.tol = tol; o.final_residual = true; o.flexible = false;
        if (tols) o.tol_slot = &tolc;
        if (method == 2) K.layout(n, chunk, restart, false, 0, false, false);
        vh::gmres_batch<double>(G.ctx, n, A64, &M64, bc, xc, sys, o, W64, cs);
        std::vector<int> again;
        for (int i = 0; i < cnt && method == 2; ++i)
          if (cs[i].flag != 0) again.push_back(i);
        if (!again.empty()) {
          // method 2: the systems the pgmres semantics left above their
          // tolerance (it stops on the preconditioned residual) continue from
          // their solutions in FP64 flexible GMRES with the FP64
          // preconditioner, which stops on the true residual (restart 50, its
          // workspace in the arena grown within the budget)
          int rc = 50;
          while (rc > 8 && (double)vh::KrylovCache::bytes(n, chunk, rc, true, 0, false, false) > budget) rc = rc * 3 / 4;
          K.layout(n, chunk, rc, true, 0, false, false);
          vh::SolveOpts o3;
          o3.restart = rc; o3.maxit = 40; o3.tol = tol; o3.final_residual = true; o3.flexible = true;
          o3.x0 = true;
          o3.true_restart = true;
          if (tols) o3.tol_slot = &tolc;
          std::vector<vh::SysStats> cs3;
          vh::gmres_batch<double>(G.ctx, n, A64, &M64, bc, xc, again, o3, K.w64, cs3);
          for (int i : again) {
            cs3[i].iters += cs[i].iters;
            cs3[i].matvecs += cs[i].matvecs;
            cs[i] = cs3[i];
          }
          if (vh::trace_on()) fprintf(stderr, "[vh] solve_batch: %zu systems continued to the true residual\n",
                                      again.size());
        }
      } else {
        if (s0 > 0) K.layout(n, chunk, restart, true, inner_restart, true, true);   // after a continuation
        // Systems of a two-fluid conductor (kinetic impedance) skip the FP32
        // inner solves: their FP32 Krylov process stalls (the preconditioner's
        // |A_inv| diagonal models the mostly inductive operator poorly and FP32
        // rounding slows it further; on the uneven batch's strips 30 FP32
        // steps left the residual near 1e-2), so they run the FP64 FGMRES
        // below from the start.
        std::vector<int> msys, dsys;
        for (int i = 0; i < cnt; ++i) (precons[pidx[s0 + i]]->P.kinetic ? dsys : msys).push_back(i);
        if (!x_init)
          for (int i : dsys) VH_HIP(hipMemsetAsync(xc + (int64_t)i * n, 0, (size_t)n * sizeof(double2), st));
        cs.resize(cnt);
        vh::MixedOpts o;
        o.tol = tol; o.outer_restart = restart; o.outer_maxit = maxit;
        o.inner_restart = inner_restart; o.inner_tol = inner_tol;
        o.x0 = x_init != 0;
        o.inner_flexible = true;
        o.escalate_inner = inner_restart;
        // outer iteration: FP32-GMRES iterative refinement (every step tests
        // the true FP64 residual with its one FP64 matvec), or FP64 FGMRES
        // with VH_MIXED_OUTER=fgmres (one more FP64 matvec per system for the
        // final true residual, and its basis and Gram-Schmidt)
        static const bool outer_fgmres = [] {
          const char* e = std::getenv("VH_MIXED_OUTER");
          return e && std::string(e) == "fgmres";
        }();
        o.refine = !outer_fgmres;
        if (tols) o.tol_slot = &tolc;
        if (!msys.empty()) vh::fgmres_mixed(G.ctx, n, A64, A32, M32, bc, xc, msys, o, K, cs);
        // The two-fluid systems, and the systems left above their tolerance
        // -- those whose first inner restart cycle missed its tolerance (the
        // FP32 Krylov process converges too slowly on them) and any the outer
        // iteration did not finish -- continue from their solutions in FP64
        // flexible GMRES with the FP64 preconditioner.  Its workspace takes
        // the arena (grown within the budget to restart 50).
        std::vector<int> last = dsys;
        for (int i : msys)
          if (cs[i].flag != 0) last.push_back(i);
        if (!last.empty()) {
          int rc = 50;
          while (rc > 8 && (double)vh::KrylovCache::bytes(n, chunk, rc, true, 0, false, false) > budget) rc = rc * 3 / 4;
          K.layout(n, chunk, rc, true, 0, false, false);
          vh::SolveOpts o3;
          o3.restart = rc; o3.maxit = 40; o3.tol = tol; o3.final_residual = true; o3.flexible = true;
          o3.x0 = true;
          o3.true_restart = true;
          if (tols) o3.tol_slot = &tolc;
          std::vector<vh::SysStats> cs3;
          vh::gmres_batch<double>(G.ctx, n, A64, &M64, bc, xc, last, o3, K.w64, cs3);
          for (int i : last) {
            cs3[i].iters += cs[i].iters;
            cs3[i].inner_iters = cs[i].inner_iters;
            cs3[i].matvecs += cs[i].matvecs;
            cs[i] = cs3[i];
          }
          if (vh::trace_on()) fprintf(stderr, "[vh] solve_batch: %zu systems finished in FP64 (restart %d)\n", last.size(), rc);
        }
      }
      for (int i = 0; i < cnt; ++i) stats[s0 + i] = cs[i];
    }
    if (!on_device) d2h(x, io.y, (size_t)nsys * n * 16, st);
    VH_HIP(hipStreamSynchronize(st));
    for (int64_t s = 0; s < nsys; ++s) {
      if (flag) flag[s] = stats[s].flag;
      if (relres) relres[s] = stats[s].relres;
      if (iters) iters[s] = stats[s].iters;
      if (inner_iters) inner_iters[s] = stats[s].inner_iters;
      if (resid_est) resid_est[s] = stats[s].resvec.empty() ? 1.0 : stats[s].resvec.back();
    }
  });
}

int vh_solve_batch_bytes(vh_geom* g, int method, int restart, int inner_restart, double* bytes_per_system) {
  return guard([&] {
    need(g, "g"); need(bytes_per_system, "bytes_per_system");
    VH_REQUIRE(method >= 0 && method <= 2, "method must be 0 (fp64), 1 (mixed) or 2 (fp64 to the true residual)");
    VH_REQUIRE(restart >= 1 && restart <= 127 && inner_restart >= 1 && inner_restart <= 127, "bad restart");
    *bytes_per_system = batch_bytes_per_system(g->G, method, restart, inner_restart);
  });
}

int vh_fft3d(vh_geom* g, void* data, int ncomp, int prec, int dir) {
  return guard([&] {
    need(g, "g"); need(data, "data");
    VH_REQUIRE(dir == -1 || dir == 1, "dir must be -1 or +1");
    vh::fft3d_test(g->G, data, ncomp, prec, dir);
  });
}

}  // extern "C"
