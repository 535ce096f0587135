// Shared definitions of the VoxHenry MI355X library: error handling, complex
// types, device buffers, the per-process context.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <stdexcept>
#include <vector>
#include <complex>
#include <cstdlib>

namespace vh {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum Status {
  OK = 0,
  ERR_ARG = 1,
  ERR_HIP = 2,
  ERR_STATE = 3,
  ERR_NUMERIC = 4,
  ERR_ALLOC = 5,
  ERR_INTERNAL = 6,
};

#define VH_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      throw ::vh::Error(::vh::ERR_HIP, std::string(#call " -> ") +            \
                        hipGetErrorString(e_) + " at " __FILE__ ":" +            \
                        std::to_string(__LINE__));                                \
  } while (0)

#define VH_REQUIRE(cond, msg)                                                     \
  do {                                                                            \
    if (!(cond)) throw ::vh::Error(::vh::ERR_ARG, std::string(msg));           \
  } while (0)

// Launch check: catches launch-configuration errors right at the launch site.
#define VH_LAUNCHED() VH_HIP(hipGetLastError())

template <typename R> struct cplx_t;
template <> struct cplx_t<double> { using type = double2; };
template <> struct cplx_t<float> { using type = float2; };
template <typename R> using cplx = typename cplx_t<R>::type;

// XCD-aware workgroup order: the hardware deals consecutive workgroups of a
// launch round-robin over the 8 XCDs; xcd_tile maps the linear workgroup id
// bid of nblk to a logical tile so that each XCD takes a contiguous range of
// tiles, and neighbouring tiles (which share 128-byte lines) meet in the same
// L2.  A bijection of [0, nblk).
__device__ __forceinline__ int xcd_tile(int bid, int nblk) {
  const int per = nblk >> 3, rem = nblk & 7;
  const int x = bid & 7, local = bid >> 3;
  return x * per + (x < rem ? x : rem) + local;
}

template <typename R> __host__ __device__ inline cplx<R> mk(R x, R y) { cplx<R> c; c.x = x; c.y = y; return c; }
template <typename C> __host__ __device__ inline C cadd(C a, C b) { a.x += b.x; a.y += b.y; return a; }
template <typename C> __host__ __device__ inline C csub(C a, C b) { a.x -= b.x; a.y -= b.y; return a; }
template <typename C> __host__ __device__ inline C cmul(C a, C b) {
  C c; c.x = a.x * b.x - a.y * b.y; c.y = a.x * b.y + a.y * b.x; return c;
}
template <typename C> __host__ __device__ inline C cconjmul(C a, C b) {  // conj(a)*b
  C c; c.x = a.x * b.x + a.y * b.y; c.y = a.x * b.y - a.y * b.x; return c;
}
template <typename C, typename S> __host__ __device__ inline C cscale(C a, S s) { a.x *= s; a.y *= s; return a; }

// Debug aid: with VH_POISON_ALLOC=1 in the environment every new device
// allocation is filled with 0xFF bytes (NaN in floating point), so a read of
// memory that was never written shows up in the results.
inline bool poison_alloc() {
  static const bool on = [] { const char* e = std::getenv("VH_POISON_ALLOC"); return e && e[0] == '1'; }();
  return on;
}

// Debug aid: with VH_TRACE=1 the solvers print host-side phase times to stderr.
inline bool trace_on() {
  static const bool on = [] { const char* e = std::getenv("VH_TRACE"); return e && e[0] == '1'; }();
  return on;
}

// RAII device allocation.
template <typename T> struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept { p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; return *this; }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      throw Error(ERR_ALLOC, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed: " + hipGetErrorString(e));
    }
    n = count;
    if (poison_alloc()) {
      e = hipMemset(p, 0xFF, count * sizeof(T));
      if (e != hipSuccess) throw Error(ERR_HIP, std::string("poison fill failed: ") + hipGetErrorString(e));
      e = hipDeviceSynchronize();
      if (e != hipSuccess) throw Error(ERR_HIP, std::string("poison fill failed: ") + hipGetErrorString(e));
    }
  }
  void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
  size_t bytes() const { return n * sizeof(T); }
};

// RAII pinned (page-locked) host allocation for the small host<->device
// transfers of the solver loops: asynchronous copies from and to it are
// ordered on the stream and complete at the next stream synchronisation.
template <typename T> struct PinnedBuf {
  T* p = nullptr;
  size_t n = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    hipError_t e = hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocDefault);
    if (e != hipSuccess) {
      p = nullptr;
      throw Error(ERR_ALLOC, "hipHostMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed: " + hipGetErrorString(e));
    }
    n = count;
  }
  void ensure(size_t count) { if (n < count) alloc(count); }
  void release() { if (p) (void)hipHostFree(p); p = nullptr; n = 0; }
  T& operator[](size_t i) { return p[i]; }
  const T& operator[](size_t i) const { return p[i]; }
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;   // caller-provided or the null stream
  int num_cus = 256;
  size_t lds_per_block = 65536;
};

// Integer helpers.
inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Mixed-radix factorisation used by the FFT plans (radices 8,4,5,3,2,7).
bool factor_smooth(int n, std::vector<int>& radices);
// Smallest n' >= n that factors over {2,3,5,7}.
int next_smooth(int n);

}  // namespace vh
