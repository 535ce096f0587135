// extern "C" boundary of the library (include/voxhenry.h).  Every entry point
// catches C++ exceptions and turns them into a status code plus a thread-local
// message.
#include "vh_geom.h"
#include "vh_solver.h"
#include "../../include/voxhenry.h"

#include <cstring>
#include <memory>

struct vh_ctx_s { vh::Context c; };
struct vh_geom_s {
  vh::Geometry G;
  vh::DevBuf<double2> greal;   // real-space tensor from the last generate call
};
struct vh_precon_s { vh::Precon P; };

namespace {
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
      VH_HIP(hipMemcpyAsync(bx.p, x, bx.bytes(), hipMemcpyHostToDevice, st));
      VH_HIP(hipMemsetAsync(by.p, 0, by.bytes(), st));
      dx = bx.p;
      dy = by.p;
    }
    if (prec == VH_C128)
      vh::matvec<double>(G, (int)nrhs, (const double2*)dx, ldx, (double2*)dy, ldy, jre, jim, mode);
    else
      vh::matvec<float>(G, (int)nrhs, (const float2*)dx, ldx, (float2*)dy, ldy, jre, jim, mode);
    if (!on_device) {
      VH_HIP(hipMemcpyAsync(y, by.p, by.bytes(), hipMemcpyDeviceToHost, st));
      VH_HIP(hipStreamSynchronize(st));
    }
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
    if (npass) *npass = 0;
    (void)ms_pass;
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
