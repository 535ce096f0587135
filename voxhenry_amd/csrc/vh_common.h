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
  }
  void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
  size_t bytes() const { return n * sizeof(T); }
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
