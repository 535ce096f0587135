// LDS-staged mixed-radix Stockham FFT engine for 64-wide wavefronts.
//
// A workgroup owns T pencils of length n held in LDS; element (p, i) of pencil
// p lives at buf[p*ps + i*is].  Each stage of radix R performs n/R butterflies
// per pencil: butterfly j reads elements j + r*n/R, multiplies by the twiddle
// w_n^{(j mod Ns) * r * n/(Ns R)}, applies an R-point DFT and writes elements
// (j/Ns)*Ns*R + (j mod Ns) + r*Ns (decimation-in-time Stockham autosort).  All
// butterflies of a stage are first gathered into registers, the workgroup
// synchronises, then the results are written back in place: one LDS buffer,
// no ping-pong.  Twiddles come from a per-length table of w_n^k built on the
// host in long double.
#pragma once
#include "vh_common.h"

namespace vh {

constexpr int FFT_MAX_STAGES = 12;
constexpr int FFT_THREADS = 256;

struct FftPlan {
  int n;
  int nstage;
  int radix[FFT_MAX_STAGES];
  const void* tw;       // n entries of w_n^k = exp(-2 pi i k / n), precision of the kernel
};

// ----------------------------------------------------------------------------
// small DFTs.  DIR = -1 forward (exp(-i..)), +1 inverse (unnormalised).
// ----------------------------------------------------------------------------
template <int DIR, typename C> __device__ __forceinline__ C rot90(C a) {
  // multiply by DIR*i :  (-i)*(x+iy) = y - ix ;  (+i)*(x+iy) = -y + ix
  C r;
  if (DIR < 0) { r.x = a.y; r.y = -a.x; } else { r.x = -a.y; r.y = a.x; }
  return r;
}

// cos(2 pi m / P), sin(2 pi m / P) for the odd radices 3, 5, 7 (exact to double).
__host__ __device__ constexpr double kc(int P, int m) {
  return P == 3 ? (m == 0 ? 1.0 : -0.5)
       : P == 5 ? (m == 0 ? 1.0 : (m == 1 || m == 4) ? 0.30901699437494742410229341718281906 : -0.80901699437494742410229341718281906)
       : (m == 0 ? 1.0 : (m == 1 || m == 6) ? 0.62348980185873353052500488400423981
          : (m == 2 || m == 5) ? -0.22252093395631440428890256449679476 : -0.90096886790241912623610231950744505);
}
__host__ __device__ constexpr double ks(int P, int m) {
  return P == 3 ? (m == 0 ? 0.0 : m == 1 ? 0.86602540378443864676372317075293618 : -0.86602540378443864676372317075293618)
       : P == 5 ? (m == 0 ? 0.0 : m == 1 ? 0.95105651629515357211643933337938214 : m == 2 ? 0.58778525229247312916870595463907277
                   : m == 3 ? -0.58778525229247312916870595463907277 : -0.95105651629515357211643933337938214)
       : (m == 0 ? 0.0 : m == 1 ? 0.78183148246802980870844452667405775 : m == 2 ? 0.97492791218182360701813168299393122
          : m == 3 ? 0.43388373911755812047576833284835875 : m == 4 ? -0.43388373911755812047576833284835875
          : m == 5 ? -0.97492791218182360701813168299393122 : -0.78183148246802980870844452667405775);
}

// cos and sin of 2 pi m / P for the one-radix lengths P in {2, 3, 4, 5, 7, 8}
__host__ __device__ constexpr double root_cos(int P, int m) {
  m %= P;
  constexpr double h = 0.70710678118654752440084436210484904;
  return P == 2 ? (m == 0 ? 1.0 : -1.0)
       : P == 4 ? (m == 0 ? 1.0 : m == 2 ? -1.0 : 0.0)
       : P == 8 ? (m == 0 ? 1.0 : m == 4 ? -1.0 : (m == 2 || m == 6) ? 0.0 : (m == 1 || m == 7) ? h : -h)
       : kc(P, m);
}
__host__ __device__ constexpr double root_sin(int P, int m) {
  m %= P;
  constexpr double h = 0.70710678118654752440084436210484904;
  return P == 2 ? 0.0
       : P == 4 ? (m == 1 ? 1.0 : m == 3 ? -1.0 : 0.0)
       : P == 8 ? (m == 0 || m == 4 ? 0.0 : m == 2 ? 1.0 : m == 6 ? -1.0 : (m == 1 || m == 3) ? h : -h)
       : ks(P, m);
}

template <int RAD, int DIR, typename C> struct Dft;

template <int DIR, typename C> struct Dft<2, DIR, C> {
  __device__ __forceinline__ static void run(C* v) {
    C a = v[0], b = v[1];
    v[0] = cadd(a, b); v[1] = csub(a, b);
  }
};

template <int DIR, typename C> struct Dft<4, DIR, C> {
  __device__ __forceinline__ static void run(C* v) {
    C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    C t2 = cadd(v[1], v[3]), t3 = rot90<DIR>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2); v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3); v[3] = csub(t1, t3);
  }
};

template <int DIR, typename C> struct Dft<8, DIR, C> {
  __device__ __forceinline__ static void run(C* v) {
    C e[4] = {v[0], v[2], v[4], v[6]};
    C o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, DIR, C>::run(e);
    Dft<4, DIR, C>::run(o);
    using R = decltype(v[0].x);
    const R h = R(0.70710678118654752440084436210484904);
    C o1, o2, o3;
    // w8^k = exp(DIR * i pi k / 4)
    if (DIR < 0) {
      o1.x = h * (o[1].x + o[1].y); o1.y = h * (o[1].y - o[1].x);
      o3.x = h * (o[3].y - o[3].x); o3.y = -h * (o[3].x + o[3].y);
    } else {
      o1.x = h * (o[1].x - o[1].y); o1.y = h * (o[1].y + o[1].x);
      o3.x = -h * (o[3].x + o[3].y); o3.y = h * (o[3].x - o[3].y);
    }
    o2 = rot90<DIR>(o[2]);
    v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);   v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);   v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);   v[7] = csub(e[3], o3);
  }
};

// Radix 16 as 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2; DFT4 over n1, twiddle
// w16^(n2 k1), DFT4 over n2.
template <int DIR, typename C> struct Dft<16, DIR, C> {
  __device__ __forceinline__ static void run(C* v) {
    using R = decltype(v[0].x);
    const R c1 = R(0.92387953251128675612818318939678829), s1 = R(0.38268343236508977172845998403039887);
    const R h = R(0.70710678118654752440084436210484904);
    C y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      C t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
      Dft<4, DIR, C>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = t[k1];
    }
    // twiddles w16^m, m = n2*k1 in {1,2,3,4,6,9}: cos(m pi/8) + DIR*i*sin(m pi/8)
    auto tw = [&](C a, R c, R s) { C r; const R sd = DIR < 0 ? -s : s; r.x = a.x * c - a.y * sd; r.y = a.x * sd + a.y * c; return r; };
    y[1][1] = tw(y[1][1], c1, s1);
    y[1][2] = tw(y[1][2], h, h);
    y[1][3] = tw(y[1][3], s1, c1);
    y[2][1] = tw(y[2][1], h, h);
    y[2][2] = rot90<DIR>(y[2][2]);
    y[2][3] = tw(y[2][3], -h, h);
    y[3][1] = tw(y[3][1], s1, c1);
    y[3][2] = tw(y[3][2], -h, h);
    y[3][3] = tw(y[3][3], -c1, -s1);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      C t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
      Dft<4, DIR, C>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
  }
};

// Odd radix P in {3,5,7}: X_k = v0 + sum_q cos(th) s_q + DIR*i sum_q sin(th) d_q,
// s_q = v_q + v_{P-q}, d_q = v_q - v_{P-q}, th = 2 pi q k / P.
template <int P, int DIR, typename C> struct DftOdd {
  __device__ __forceinline__ static void run(C* v) {
    using R = decltype(v[0].x);
    constexpr int H = (P - 1) / 2;
    C s[H], d[H];
#pragma unroll
    for (int q = 1; q <= H; ++q) { s[q - 1] = cadd(v[q], v[P - q]); d[q - 1] = csub(v[q], v[P - q]); }
    C x0 = v[0];
#pragma unroll
    for (int q = 0; q < H; ++q) x0 = cadd(x0, s[q]);
    C out[P];
    out[0] = x0;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      C a = v[0], b; b.x = 0; b.y = 0;
#pragma unroll
      for (int q = 1; q <= H; ++q) {
        const R c = R(kc(P, (q * k) % P)), sn = R(ks(P, (q * k) % P));
        a.x += c * s[q - 1].x; a.y += c * s[q - 1].y;
        b.x += sn * d[q - 1].x; b.y += sn * d[q - 1].y;
      }
      // X_k = a + DIR*i*b ; X_{P-k} = a - DIR*i*b
      const C ib = rot90<DIR>(b);
      out[k] = cadd(a, ib);
      out[P - k] = csub(a, ib);
    }
#pragma unroll
    for (int k = 0; k < P; ++k) v[k] = out[k];
  }
};
template <int DIR, typename C> struct Dft<3, DIR, C> { __device__ __forceinline__ static void run(C* v) { DftOdd<3, DIR, C>::run(v); } };
template <int DIR, typename C> struct Dft<5, DIR, C> { __device__ __forceinline__ static void run(C* v) { DftOdd<5, DIR, C>::run(v); } };
template <int DIR, typename C> struct Dft<7, DIR, C> { __device__ __forceinline__ static void run(C* v) { DftOdd<7, DIR, C>::run(v); } };

// ----------------------------------------------------------------------------
// one Stockham stage of radix RAD over T pencils in LDS.
// EMAX bounds the elements per thread (T*n <= EMAX * nthr).
// ----------------------------------------------------------------------------
template <typename R, int DIR, int RAD, int EMAX>
__device__ __forceinline__ void fft_stage(cplx<R>* buf, int n, int T, int ps, int is, int Ns,
                                          const cplx<R>* __restrict__ tw, int tid, int nthr,
                                          bool pencil_fast) {
  using C = cplx<R>;
  constexpr int BMAX = (EMAX + RAD - 1) / RAD;
  const int nb = n / RAD;
  const int total = nb * T;
  const int twstep = n / (Ns * RAD);
  C v[BMAX][RAD];
#pragma unroll
  for (int b = 0; b < BMAX; ++b) {
    const int id = tid + b * nthr;
    if (id < total) {
      int p, j;
      if (pencil_fast) { p = id % T; j = id / T; } else { p = id / nb; j = id % nb; }
      const C* src = buf + p * ps + j * is;
#pragma unroll
      for (int r = 0; r < RAD; ++r) v[b][r] = src[r * nb * is];
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < BMAX; ++b) {
    const int id = tid + b * nthr;
    if (id < total) {
      int p, j;
      if (pencil_fast) { p = id % T; j = id / T; } else { p = id / nb; j = id % nb; }
      const int k = j % Ns;
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < RAD; ++r) {
          C w = tw[k * r * twstep];
          if (DIR > 0) w.y = -w.y;
          v[b][r] = cmul(v[b][r], w);
        }
      }
      Dft<RAD, DIR, C>::run(v[b]);
      const int idxD = (j / Ns) * Ns * RAD + k;
      C* dst = buf + p * ps + idxD * is;
#pragma unroll
      for (int r = 0; r < RAD; ++r) dst[r * Ns * is] = v[b][r];
    }
  }
  __syncthreads();
}

// Full FFT of T pencils in LDS.  Caller must __syncthreads() before (data in
// place) -- the routine ends with a barrier, so results are visible on return.
template <typename R, int DIR, int EMAX>
__device__ void lds_fft(cplx<R>* buf, int T, int ps, int is, const FftPlan& pl, int tid, int nthr,
                        bool pencil_fast) {
  const cplx<R>* tw = (const cplx<R>*)pl.tw;
  int Ns = 1;
  for (int s = 0; s < pl.nstage; ++s) {
    const int rad = pl.radix[s];
    switch (rad) {
      case 8: fft_stage<R, DIR, 8, EMAX>(buf, pl.n, T, ps, is, Ns, tw, tid, nthr, pencil_fast); break;
      case 4: fft_stage<R, DIR, 4, EMAX>(buf, pl.n, T, ps, is, Ns, tw, tid, nthr, pencil_fast); break;
      case 5: fft_stage<R, DIR, 5, EMAX>(buf, pl.n, T, ps, is, Ns, tw, tid, nthr, pencil_fast); break;
      case 3: fft_stage<R, DIR, 3, EMAX>(buf, pl.n, T, ps, is, Ns, tw, tid, nthr, pencil_fast); break;
      case 2: fft_stage<R, DIR, 2, EMAX>(buf, pl.n, T, ps, is, Ns, tw, tid, nthr, pencil_fast); break;
      case 7: fft_stage<R, DIR, 7, EMAX>(buf, pl.n, T, ps, is, Ns, tw, tid, nthr, pencil_fast); break;
      default: break;
    }
    Ns *= rad;
  }
}

// Host side: an FFT length with twiddle tables in both precisions.
struct FftAxis {
  int n = 0;
  std::vector<int> radices;
  DevBuf<double2> tw64;
  DevBuf<float2> tw32;
  void init(int len);
  template <typename R> FftPlan plan() const;
};

}  // namespace vh
