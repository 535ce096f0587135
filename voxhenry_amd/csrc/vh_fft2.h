// Fused-pass FFT engine (Stockham, mixed radix) used by the matvec passes.
//
// A workgroup of NT threads transforms T pencils of length n; each pencil is
// served by tpp = NT/T threads (both powers of two).  Stage s of radix R
// performs n/R butterflies per pencil: butterfly j reads elements j + r*n/R,
// multiplies element r by w_{Ns R}^{(j mod Ns) r} (Ns = product of the earlier
// radices; per-stage table, contiguous per butterfly), applies an R-point DFT
// and writes elements (j div Ns)*Ns*R + (j mod Ns) + r*Ns.  The first stage
// reads its inputs straight from global memory through a loader and the last
// stage writes straight to global memory through a storer, so a two-stage
// transform makes one round trip through LDS and a one-stage transform none.
// Every thread issues all loads of its butterflies before using any of them.
// j div Ns uses a precomputed 32-bit reciprocal (exact for j, Ns < 2^16).
//
// LDS placement is a functor `lx(i)` giving the LDS element offset of pencil
// element i for the calling thread's pencil.
#pragma once
#include "vh_fft.h"

namespace vh {

constexpr int F2_MAX_STAGES = 10;

struct Stage2 {
  int rad;           // radix
  int ns;            // product of the earlier radices
  int nb;            // butterflies per pencil = n / rad
  uint32_t ns_rcp;   // ceil(2^32 / ns) for ns > 1
  int tw_off;        // offset of this stage's twiddles (ns*(rad-1) entries) in the table
};

// Kernel-side view of a plan: the stage table lives in device memory (read
// with uniform loads when a stage starts), so plans cost no registers.
struct Plan2 {
  int n;
  int nstage;
  const Stage2* st;  // nstage entries
  const void* tw;    // concatenated per-stage twiddle tables, precision of the kernel
};

__device__ __forceinline__ int div_ns(int j, const Stage2& S) {
  return S.ns == 1 ? j : (int)__umulhi((uint32_t)j, S.ns_rcp);
}

template <typename R> __device__ __forceinline__ cplx<R> cz() { cplx<R> c; c.x = 0; c.y = 0; return c; }

// Twiddle and DFT of one butterfly held in v (inputs j + r*nb in natural order).
template <typename R, int DIR, int RAD>
__device__ __forceinline__ void butterfly(cplx<R>* v, const cplx<R>* __restrict__ w, bool twiddle) {
  if (twiddle) {
#pragma unroll
    for (int r = 1; r < RAD; ++r) {
      cplx<R> t = w[r - 1];
      if (DIR > 0) t.y = -t.y;
      v[r] = cmul(v[r], t);
    }
  }
  Dft<RAD, DIR, cplx<R>>::run(v);
}

// One stage of radix RAD for the butterflies j = jl + m*tpp of one pencil.
// ld(i) returns input element i; st(o, v) stores output element o.  With
// sync_mid the workgroup synchronises between the reads and the writes (the
// stage works in place in LDS).
template <typename R, int DIR, int RAD, int BM, class LD, class ST>
__device__ __forceinline__ void stage_run(const Stage2& S, const cplx<R>* __restrict__ tws, int jl, int tpp, LD& ld,
                                          ST& st, bool sync_mid) {
  using C = cplx<R>;
  C v[BM][RAD];
#pragma unroll
  for (int m = 0; m < BM; ++m) {
    const int j = jl + m * tpp;
    if (j < S.nb) {
#pragma unroll
      for (int r = 0; r < RAD; ++r) v[m][r] = ld(j + r * S.nb);
    }
  }
  if (sync_mid) __syncthreads();
#pragma unroll
  for (int m = 0; m < BM; ++m) {
    const int j = jl + m * tpp;
    if (j < S.nb) {
      const int q = div_ns(j, S);
      const int k = j - q * S.ns;
      butterfly<R, DIR, RAD>(v[m], tws + k * (RAD - 1), S.ns > 1);
      const int base = q * S.ns * RAD + k;
#pragma unroll
      for (int r = 0; r < RAD; ++r) st(base + r * S.ns, v[m][r]);
    }
  }
}

// Largest power-of-two radix of the plans used with an element bound E.
__host__ __device__ constexpr int rmax_for(int E) { return E >= 16 ? 16 : E >= 8 ? 8 : 4; }

// Runtime radix dispatch; E bounds the elements per thread and pencil
// (ceil(n/tpp)), so a stage of radix R needs ceil(E/R) butterflies per thread.
template <typename R, int DIR, int E, class LD, class ST>
__device__ __forceinline__ void stage_any(const Stage2& S, const cplx<R>* tw, int jl, int tpp, LD& ld, ST& st,
                                          bool sync_mid) {
  const cplx<R>* tws = tw + S.tw_off;
  // plans for this E never use a power-of-two radix above rmax_for(E); leaving
  // those cases out keeps their register demand out of the kernel
  constexpr int RMAX = rmax_for(E);
  switch (S.rad) {
    case 16: if constexpr (RMAX >= 16) stage_run<R, DIR, 16, (E + 15) / 16>(S, tws, jl, tpp, ld, st, sync_mid); break;
    case 8: if constexpr (RMAX >= 8) stage_run<R, DIR, 8, (E + 7) / 8>(S, tws, jl, tpp, ld, st, sync_mid); break;
    case 4: stage_run<R, DIR, 4, (E + 3) / 4>(S, tws, jl, tpp, ld, st, sync_mid); break;
    case 5: stage_run<R, DIR, 5, (E + 4) / 5>(S, tws, jl, tpp, ld, st, sync_mid); break;
    case 3: stage_run<R, DIR, 3, (E + 2) / 3>(S, tws, jl, tpp, ld, st, sync_mid); break;
    case 7: stage_run<R, DIR, 7, (E + 6) / 7>(S, tws, jl, tpp, ld, st, sync_mid); break;
    case 2: stage_run<R, DIR, 2, (E + 1) / 2>(S, tws, jl, tpp, ld, st, sync_mid); break;
    default: break;
  }
}

// Full transform of one pencil per thread group: gld/gst are the global loader
// and storer, lds the workgroup's LDS array, lx the element placement.  Ends
// with the last stage's stores issued (no trailing barrier).
template <typename R, int DIR, int E, class LX, class GLD, class GST>
__device__ __forceinline__ void fft_pencil(const Plan2& P, cplx<R>* lds, const LX& lx, int jl, int tpp, GLD& gld,
                                           GST& gst) {
  const cplx<R>* tw = (const cplx<R>*)P.tw;
  auto L = [&](int i) { return lds[lx(i)]; };
  auto S = [&](int o, cplx<R> v) { lds[lx(o)] = v; };
  if (P.nstage == 1) {
    stage_any<R, DIR, E>(P.st[0], tw, jl, tpp, gld, gst, false);
    return;
  }
  stage_any<R, DIR, E>(P.st[0], tw, jl, tpp, gld, S, false);
  __syncthreads();
  for (int s = 1; s + 1 < P.nstage; ++s) {
    stage_any<R, DIR, E>(P.st[s], tw, jl, tpp, L, S, true);
    __syncthreads();
  }
  stage_any<R, DIR, E>(P.st[P.nstage - 1], tw, jl, tpp, L, gst, false);
}

// Forward part of a register hand-off: the last stage (radix RL) leaves its
// outputs in registers, out[m][r] = element j + r*n/RL of butterfly
// j = jl + m*tpp, in natural order.
template <typename R, int DIR, int E, int RL, class LX, class GLD>
__device__ __forceinline__ void fft_pencil_to_regs(const Plan2& P, cplx<R>* lds, const LX& lx, int jl, int tpp,
                                                   GLD& gld, cplx<R> (&out)[(E + RL - 1) / RL][RL]) {
  using C = cplx<R>;
  constexpr int BL = (E + RL - 1) / RL;
  const C* tw = (const C*)P.tw;
  auto L = [&](int i) { return lds[lx(i)]; };
  auto S = [&](int o, C v) { lds[lx(o)] = v; };
  const Stage2& SL = P.st[P.nstage - 1];
  if (P.nstage > 1) {
    stage_any<R, DIR, E>(P.st[0], tw, jl, tpp, gld, S, false);
    __syncthreads();
    for (int s = 1; s + 1 < P.nstage; ++s) {
      stage_any<R, DIR, E>(P.st[s], tw, jl, tpp, L, S, true);
      __syncthreads();
    }
  }
  const C* tws = tw + SL.tw_off;
#pragma unroll
  for (int m = 0; m < BL; ++m) {
    const int j = jl + m * tpp;
    if (j < SL.nb) {
#pragma unroll
      for (int r = 0; r < RL; ++r) out[m][r] = P.nstage > 1 ? L(j + r * SL.nb) : gld(j + r * SL.nb);
    }
  }
#pragma unroll
  for (int m = 0; m < BL; ++m) {
    const int j = jl + m * tpp;
    // last stage: Ns*RL = n, so j < Ns and j mod Ns = j
    if (j < SL.nb) butterfly<R, DIR, RL>(out[m], tws + j * (RL - 1), SL.ns > 1);
  }
}

// Inverse counterpart: the first stage (radix RL, Ns = 1, no twiddles) takes
// its inputs from registers laid out as fft_pencil_to_regs leaves them.
template <typename R, int DIR, int E, int RL, class LX, class GST>
__device__ __forceinline__ void fft_pencil_from_regs(const Plan2& P, cplx<R>* lds, const LX& lx, int jl, int tpp,
                                                     cplx<R> (&in)[(E + RL - 1) / RL][RL], GST& gst) {
  using C = cplx<R>;
  constexpr int BL = (E + RL - 1) / RL;
  const C* tw = (const C*)P.tw;
  auto L = [&](int i) { return lds[lx(i)]; };
  auto S = [&](int o, C v) { lds[lx(o)] = v; };
  const Stage2& S0 = P.st[0];   // output o = j*RL + r
#pragma unroll
  for (int m = 0; m < BL; ++m) {
    const int j = jl + m * tpp;
    if (j < S0.nb) {
      Dft<RL, DIR, C>::run(in[m]);
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        if (P.nstage == 1) gst(j * RL + r, in[m][r]);
        else S(j * RL + r, in[m][r]);
      }
    }
  }
  if (P.nstage == 1) return;
  __syncthreads();
  for (int s = 1; s + 1 < P.nstage; ++s) {
    stage_any<R, DIR, E>(P.st[s], tw, jl, tpp, L, S, true);
    __syncthreads();
  }
  stage_any<R, DIR, E>(P.st[P.nstage - 1], tw, jl, tpp, L, gst, false);
}

// Host: plans of one transform length for every largest power-of-two radix
// (16, 8, 4) in forward and reversed stage order, both precisions.
struct FftAxis2 {
  struct Entry {
    std::vector<int> radices;
    std::vector<Stage2> st;
    DevBuf<Stage2> dst;   // device copy of st
    DevBuf<double2> tw64;
    DevBuf<float2> tw32;
  };
  int n = 0;
  Entry e[3][2];   // [16, 8, 4][forward order, reversed order]
  void init(int len);
  // Plan whose power-of-two radices are at most rmax (16, 8 or 4).
  template <typename R> Plan2 plan(int rmax, bool reversed) const;
  int last_radix(int rmax) const;   // last radix of the forward-order plan
  const std::vector<Stage2>& stages(int rmax) const;   // forward-order stages (host copy)
};

// Threads per pencil: smallest power of two t <= nt_max with ceil(n/t) <= E.
inline int fft_tpp(int n, int E, int nt_max) {
  int t = 1;
  while (t < nt_max && (n + t - 1) / t > E) t <<= 1;
  return ((n + t - 1) / t > E) ? -1 : t;
}

// ---------------------------------------------------------------------------
// Compile-time plans for power-of-two lengths.  P2Plan<N, RMAX, REV> has the
// radix sequence of the run-time plan built for (N, RMAX) -- the power-of-two
// exponent split into as few stages of at most RMAX as possible, larger
// stages first -- in forward (REV = false) or reversed order, so the
// run-time twiddle tables of that plan serve it unchanged.
// ---------------------------------------------------------------------------
constexpr int clog2(int v) { return v <= 1 ? 0 : 1 + clog2(v / 2); }

template <int N, int RMAX, bool REV> struct P2Plan {
  static constexpr int L = clog2(N), LM = clog2(RMAX);
  static constexpr int S = (L + LM - 1) / LM;
  // log2 radix of stage s of the forward order
  static constexpr int lr_fwd(int s) { return L / S + (s < L % S ? 1 : 0); }
  static constexpr int rad(int s) { return 1 << lr_fwd(REV ? S - 1 - s : s); }
  static constexpr int ns(int s) { return s == 0 ? 1 : ns(s - 1) * rad(s - 1); }
  static constexpr int tw_off(int s) { return s == 0 ? 0 : tw_off(s - 1) + ns(s - 1) * (rad(s - 1) - 1); }
};

// One stage of a compile-time plan on TPP threads per pencil (TPP divides
// N/RAD): butterflies j = jl + m*TPP take elements j + r*N/RAD (in v, or read
// through ld) and write (j div NS)*NS*RAD + (j mod NS) + r*NS through st.
template <int N, int RAD, int TPP> struct SStage {
  static constexpr int NB = N / RAD;
  static_assert(NB % TPP == 0, "threads per pencil must divide the butterflies of every stage");
  static constexpr int BM = NB / TPP;
};

template <typename R, int DIR, int N, int RAD, int NS, int TPP, class ST>
__device__ __forceinline__ void sstage_regs(const cplx<R>* __restrict__ tws, int jl,
                                            cplx<R> (&v)[SStage<N, RAD, TPP>::BM][RAD], ST& st) {
  constexpr int BM = SStage<N, RAD, TPP>::BM;
#pragma unroll
  for (int m = 0; m < BM; ++m) {
    const int j = jl + m * TPP;
    const int k = j & (NS - 1);
    const int q = j >> clog2(NS);
    butterfly<R, DIR, RAD>(v[m], tws + k * (RAD - 1), NS > 1);
#pragma unroll
    for (int r = 0; r < RAD; ++r) st(q * NS * RAD + k + r * NS, v[m][r]);
  }
}

template <typename R, int DIR, int N, int RAD, int NS, int TPP, class LD, class ST>
__device__ __forceinline__ void sstage(const cplx<R>* __restrict__ tws, int jl, LD& ld, ST& st) {
  constexpr int NB = SStage<N, RAD, TPP>::NB, BM = SStage<N, RAD, TPP>::BM;
  cplx<R> v[BM][RAD];
#pragma unroll
  for (int m = 0; m < BM; ++m)
#pragma unroll
    for (int r = 0; r < RAD; ++r) v[m][r] = ld(jl + m * TPP + r * NB);
  sstage_regs<R, DIR, N, RAD, NS, TPP>(tws, jl, v, st);
}

}  // namespace vh
