// Geometry setup: FFT plans, voxel map, incidence matrix, impedance, and the
// drop-in path that accepts the reference's padded tensor spectrum.
#include "vh_geom.h"
#include "vh_fft2.h"
#include <algorithm>
#include <cmath>

namespace vh {

bool factor_smooth(int n, std::vector<int>& radices) {
  radices.clear();
  if (n < 1) return false;
  int m = n;
  for (int r : {8, 4, 5, 3, 2, 7}) {
    while (m % r == 0) { radices.push_back(r); m /= r; }
  }
  if (m != 1 || (int)radices.size() > FFT_MAX_STAGES) return false;
  if (radices.empty()) radices.push_back(1);
  return true;
}

int next_smooth(int n) {
  std::vector<int> r;
  for (int m = std::max(n, 1);; ++m)
    if (m >= 2 && factor_smooth(m, r)) return m;
}

void FftAxis::init(int len) {
  n = len;
  VH_REQUIRE(len >= 2 && factor_smooth(len, radices), "FFT length " + std::to_string(len) + " is not {2,3,5,7}-smooth");
  std::vector<double2> h64(n);
  std::vector<float2> h32(n);
  for (int k = 0; k < n; ++k) {
    // w_n^k = exp(-2 pi i k/n), evaluated in long double then rounded
    const long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
    const long double c = cosl(a), s = sinl(a);
    h64[k].x = (double)c; h64[k].y = (double)s;
    h32[k].x = (float)c; h32[k].y = (float)s;
  }
  tw64.alloc(n);
  tw32.alloc(n);
  VH_HIP(hipMemcpy(tw64.p, h64.data(), n * sizeof(double2), hipMemcpyHostToDevice));
  VH_HIP(hipMemcpy(tw32.p, h32.data(), n * sizeof(float2), hipMemcpyHostToDevice));
}

template <> FftPlan FftAxis::plan<double>() const {
  FftPlan p{};
  p.n = n; p.nstage = (int)radices.size();
  for (int i = 0; i < p.nstage; ++i) p.radix[i] = radices[i];
  p.tw = tw64.p;
  return p;
}
template <> FftPlan FftAxis::plan<float>() const {
  FftPlan p{};
  p.n = n; p.nstage = (int)radices.size();
  for (int i = 0; i < p.nstage; ++i) p.radix[i] = radices[i];
  p.tw = tw32.p;
  return p;
}

// Radix sequence of a v2 plan: the power-of-two part split into as few stages
// of at most rmax as possible, as evenly as possible (larger first), followed
// by the radices 7, 5, 3.
static std::vector<int> radices2(int n, int rmax) {
  int m = n, e = 0;
  while (m % 2 == 0) { m /= 2; ++e; }
  std::vector<int> odd;
  for (int r : {7, 5, 3})
    while (m % r == 0) { odd.push_back(r); m /= r; }
  VH_REQUIRE(m == 1, "FFT length " + std::to_string(n) + " is not {2,3,5,7}-smooth");
  const int lm = rmax == 16 ? 4 : rmax == 8 ? 3 : 2;
  std::vector<int> rad;
  if (e > 0) {
    const int s = (e + lm - 1) / lm;
    for (int i = 0; i < s; ++i) rad.push_back(1 << (e / s + (i < e % s ? 1 : 0)));
  }
  rad.insert(rad.end(), odd.begin(), odd.end());
  if (rad.empty()) rad.push_back(1);
  return rad;
}

void FftAxis2::init(int len) {
  n = len;
  const int rm[3] = {16, 8, 4};
  for (int a = 0; a < 3; ++a)
    for (int rev = 0; rev < 2; ++rev) {
      Entry& E = e[a][rev];
      E.radices = radices2(len, rm[a]);
      if (rev) std::reverse(E.radices.begin(), E.radices.end());
      VH_REQUIRE((int)E.radices.size() <= F2_MAX_STAGES, "FFT length has too many stages");
      VH_REQUIRE(len >= 2 && E.radices[0] >= 2, "FFT length must be >= 2");
      std::vector<double2> h64;
      std::vector<float2> h32;
      E.st.clear();
      int ns = 1;
      for (int r : E.radices) {
        Stage2 S{};
        S.rad = r; S.ns = ns; S.nb = len / r;
        S.ns_rcp = ns > 1 ? (uint32_t)(((1ULL << 32) + ns - 1) / ns) : 0u;
        S.tw_off = (int)h64.size();
        for (int k = 0; k < ns; ++k)
          for (int q = 1; q < r; ++q) {
            // w_{ns r}^{k q} = exp(-2 pi i k q / (ns r)), in long double, then rounded
            const long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)(k * q) /
                                  (long double)(ns * r);
            double2 w; w.x = (double)cosl(a); w.y = (double)sinl(a);
            float2 f; f.x = (float)cosl(a); f.y = (float)sinl(a);
            h64.push_back(w); h32.push_back(f);
          }
        E.st.push_back(S);
        ns *= r;
      }
      E.dst.alloc(E.st.size());
      VH_HIP(hipMemcpy(E.dst.p, E.st.data(), E.st.size() * sizeof(Stage2), hipMemcpyHostToDevice));
      const size_t nt = std::max<size_t>(h64.size(), 1);
      E.tw64.alloc(nt); E.tw32.alloc(nt);
      if (!h64.empty()) {
        VH_HIP(hipMemcpy(E.tw64.p, h64.data(), h64.size() * sizeof(double2), hipMemcpyHostToDevice));
        VH_HIP(hipMemcpy(E.tw32.p, h32.data(), h32.size() * sizeof(float2), hipMemcpyHostToDevice));
      }
    }
}

static int rmax_slot(int rmax) { return rmax >= 16 ? 0 : rmax >= 8 ? 1 : 2; }

template <typename R> Plan2 FftAxis2::plan(int rmax, bool reversed) const {
  const Entry& E = e[rmax_slot(rmax)][reversed ? 1 : 0];
  Plan2 p{};
  p.n = n;
  p.nstage = (int)E.st.size();
  p.st = E.dst.p;
  if constexpr (sizeof(R) == 8) p.tw = E.tw64.p; else p.tw = E.tw32.p;
  return p;
}
template Plan2 FftAxis2::plan<double>(int, bool) const;
template Plan2 FftAxis2::plan<float>(int, bool) const;

int FftAxis2::last_radix(int rmax) const { return e[rmax_slot(rmax)][0].radices.back(); }
const std::vector<Stage2>& FftAxis2::stages(int rmax) const { return e[rmax_slot(rmax)][0].st; }

void geometry_init(Geometry& G, Context* ctx, int L, int M, int N, double dx, int64_t nv, const int64_t* idx,
                   int pad_mode) {
  VH_REQUIRE(L >= 2 && M >= 2 && N >= 2, "grid dims must be >= 2 (near-field table covers a 2x2x2 block)");
  VH_REQUIRE(dx > 0, "dx must be positive");
  const int64_t V = (int64_t)L * M * N;
  VH_REQUIRE(nv >= 1 && nv <= V, "voxel count out of range");
  VH_REQUIRE(V < (1LL << 31) && 5 * nv < (1LL << 31), "grid too large for 32-bit voxel indexing");
  G.ctx = ctx;
  G.dx = dx;
  G.g.L = L; G.g.M = M; G.g.N = N;
  G.g.nv = (int)nv;
  if (pad_mode == 0) {
    G.g.Px = 2 * L; G.g.Py = 2 * M; G.g.Pz = 2 * N;
  } else {
    G.g.Px = next_smooth(2 * L - 1); G.g.Py = next_smooth(2 * M - 1); G.g.Pz = next_smooth(2 * N - 1);
  }
  G.g.Hx = G.g.Px / 2 + 1; G.g.Hy = G.g.Py / 2 + 1; G.g.Hz = G.g.Pz / 2 + 1;
  G.ax.init(G.g.Px); G.ay.init(G.g.Py); G.az.init(G.g.Pz);
  G.fx.init(G.g.Px); G.fy.init(G.g.Py); G.fz.init(G.g.Pz);
  G.num_curr = 5 * nv;
  G.vox.assign(idx, idx + nv);
  std::vector<int> vmap(V, -1);
  for (int64_t k = 0; k < nv; ++k) {
    VH_REQUIRE(idx[k] >= 0 && idx[k] < V, "voxel index out of range");
    VH_REQUIRE(k == 0 || idx[k] > idx[k - 1], "voxel indices must be strictly increasing (column-major order)");
    vmap[idx[k]] = (int)k;
  }
  G.dense = (nv == V);
  std::vector<int> lines;
  for (int64_t line = 0; line < (int64_t)M * N; ++line) {
    bool any = false;
    for (int i = 0; i < L && !any; ++i) any = vmap[i + (int64_t)L * line] >= 0;
    if (any) lines.push_back((int)line);
  }
  G.nlines = (int)lines.size();
  std::vector<int4> runs(lines.size());
  for (size_t l = 0; l < lines.size(); ++l) {
    const int* vm = vmap.data() + (int64_t)L * lines[l];
    int x0 = 0;
    while (vm[x0] < 0) ++x0;
    int x1 = x0;
    while (x1 < L && vm[x1] >= 0) ++x1;
    bool one = true;
    for (int i = x1; i < L && one; ++i) one = vm[i] < 0;
    runs[l] = one ? make_int4(x0, x1, vm[x0] - x0, 0) : make_int4(-1, 0, 0, 0);
  }
  G.row_bits_on = M <= 256 && lines.size() < (size_t)M * N;
  if (G.row_bits_on) {
    std::vector<unsigned long long> bits((size_t)4 * N, 0ull);
    for (int line : lines) bits[(size_t)4 * (line / M) + (line % M) / 64] |= 1ull << ((line % M) % 64);
    G.rowbits.alloc(bits.size());
    VH_HIP(hipMemcpy(G.rowbits.p, bits.data(), bits.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
  }
  G.runs.alloc(std::max<size_t>(runs.size(), 1));
  if (!runs.empty()) VH_HIP(hipMemcpy(G.runs.p, runs.data(), runs.size() * sizeof(int4), hipMemcpyHostToDevice));
  G.vmap.alloc(V);
  VH_HIP(hipMemcpy(G.vmap.p, vmap.data(), V * sizeof(int), hipMemcpyHostToDevice));
  G.lines.alloc(std::max<size_t>(lines.size(), 1));
  if (!lines.empty()) VH_HIP(hipMemcpy(G.lines.p, lines.data(), lines.size() * sizeof(int), hipMemcpyHostToDevice));
}

static void upload_csr(DevCsr& D, int rows, int cols, const std::vector<int>& ptr, const std::vector<int>& col,
                       const std::vector<double>& val) {
  D.rows = rows; D.cols = cols; D.nnz = (int)col.size();
  D.ptr.alloc(rows + 1);
  VH_HIP(hipMemcpy(D.ptr.p, ptr.data(), (rows + 1) * sizeof(int), hipMemcpyHostToDevice));
  const size_t nz = std::max<size_t>(col.size(), 1);
  D.col.alloc(nz); D.v64.alloc(nz); D.v32.alloc(nz);
  if (!col.empty()) {
    std::vector<float> v32(val.begin(), val.end());
    VH_HIP(hipMemcpy(D.col.p, col.data(), col.size() * sizeof(int), hipMemcpyHostToDevice));
    VH_HIP(hipMemcpy(D.v64.p, val.data(), val.size() * sizeof(double), hipMemcpyHostToDevice));
    VH_HIP(hipMemcpy(D.v32.p, v32.data(), v32.size() * sizeof(float), hipMemcpyHostToDevice));
  }
}

void geometry_set_incidence_csc(Geometry& G, int64_t num_node, int64_t nnz, const int64_t* colptr,
                                const int64_t* rowind, const double* val) {
  const int64_t nc = G.num_curr;
  VH_REQUIRE(num_node >= 0 && num_node < (1LL << 31), "num_node out of range");
  VH_REQUIRE(colptr[0] == 0 && colptr[nc] == nnz, "CSC column pointer inconsistent with nnz");
  // Ar' in CSR is Ar in CSC: rows of Ar' = columns of Ar.
  std::vector<int> tp(nc + 1), tc(nnz);
  std::vector<double> tv(nnz);
  std::vector<int> cnt(num_node + 1, 0);
  for (int64_t j = 0; j <= nc; ++j) tp[j] = (int)colptr[j];
  for (int64_t k = 0; k < nnz; ++k) {
    VH_REQUIRE(rowind[k] >= 0 && rowind[k] < num_node, "CSC row index out of range");
    tc[k] = (int)rowind[k];
    tv[k] = val[k];
    cnt[rowind[k] + 1]++;
  }
  for (int64_t i = 0; i < num_node; ++i) cnt[i + 1] += cnt[i];
  std::vector<int> rp(cnt.begin(), cnt.end()), rc(nnz);
  std::vector<double> rv(nnz);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int64_t j = 0; j < nc; ++j)
    for (int64_t k = colptr[j]; k < colptr[j + 1]; ++k) {
      const int r = (int)rowind[k];
      rc[fill[r]] = (int)j;
      rv[fill[r]] = val[k];
      fill[r]++;
    }
  G.num_node = num_node;
  upload_csr(G.Ar, (int)num_node, (int)nc, rp, rc, rv);
  upload_csr(G.ArT, (int)nc, (int)num_node, tp, tc, tv);
  // Packed rows of Ar' (pass Xi): the currents of one component touch a fixed
  // small number of nodes (2 for Jx, Jy, Jz, up to 4 for J2d, 6 for J3d) with
  // coefficients +-1 or +-1/2 (lse_compute_Ae_matrix.m), so component c keeps
  // k[c] (even) entries per current, entry = node << 3 | code (code 0 none,
  // 1 +1, 2 -1, 3 +1/2, 4 -1/2), and a row is read without the row-pointer
  // chain.  Any other matrix keeps the CSR path.
  auto code_of = [](double v) { return v == 1.0 ? 1 : v == -1.0 ? 2 : v == 0.5 ? 3 : v == -0.5 ? 4 : 0; };
  const int64_t nvx = G.g.nv;
  bool packable = num_node < (1LL << 28) && nc == 5 * nvx;
  PackedInc pinc{};
  int64_t total = 0;
  for (int c = 0; c < 5 && packable; ++c) {
    int kmax = 0;
    for (int64_t v = 0; v < nvx; ++v) {
      const int64_t j = c * nvx + v;
      kmax = std::max<int>(kmax, (int)(colptr[j + 1] - colptr[j]));
    }
    kmax = std::max(2, (kmax + 1) & ~1);
    if (kmax > 6) packable = false;
    pinc.k[c] = kmax;
    pinc.off[c] = (int)total;
    total += (int64_t)kmax * nvx;
  }
  packable = packable && total < (1LL << 31);
  std::vector<int> pk(packable ? std::max<int64_t>(total, 2) : 0, 0);
  for (int c = 0; c < 5 && packable; ++c)
    for (int64_t v = 0; v < nvx && packable; ++v) {
      const int64_t j = c * nvx + v;
      int* e = pk.data() + pinc.off[c] + v * pinc.k[c];
      for (int64_t k = colptr[j]; k < colptr[j + 1]; ++k) {
        const int code = code_of(val[k]);
        if (code == 0) { packable = false; break; }
        e[k - colptr[j]] = (int)(rowind[k] << 3) | code;
      }
    }
  G.atpk.release();
  G.pinc = pinc;
  if (packable) {
    G.atpk.alloc(pk.size());
    VH_HIP(hipMemcpy(G.atpk.p, pk.data(), pk.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  G.arp_h = std::move(rp); G.arc_h = std::move(rc); G.arv_h = std::move(rv);
  G.atp_h = std::move(tp); G.atc_h = std::move(tc); G.atv_h = std::move(tv);
  G.has_inc = true;
}

void geometry_set_impedance(Geometry& G, const double* z) {
  const int nv = G.g.nv;
  std::vector<float2> z32(nv);
  G.zh.resize(nv);
  for (int k = 0; k < nv; ++k) {
    z32[k].x = (float)z[2 * k]; z32[k].y = (float)z[2 * k + 1];
    G.zh[k] = std::complex<double>(z[2 * k], z[2 * k + 1]);
  }
  if (G.z64.n != (size_t)nv) { G.z64.alloc(nv); G.z32.alloc(nv); }
  VH_HIP(hipMemcpy(G.z64.p, z, nv * sizeof(double2), hipMemcpyHostToDevice));
  VH_HIP(hipMemcpy(G.z32.p, z32.data(), nv * sizeof(float2), hipMemcpyHostToDevice));
  G.has_z = true;
}

// Keep one octant of a full padded spectrum (host, column-major [x][y][z][c]).
void geometry_set_spectrum_full(Geometry& G, const double* fN, int64_t P0, int64_t P1, int64_t P2, double sre,
                                double sim) {
  const GridDims& g = G.g;
  VH_REQUIRE(P0 == g.Px && P1 == g.Py && P2 == g.Pz,
             "tensor spectrum dims differ from the geometry's padded grid (create the geometry with pad_mode 0)");
  const size_t O = G.octant();
  const double invP = 1.0 / ((double)g.Px * g.Py * g.Pz);
  std::vector<double2> h(TENSOR_STRIDE * O, double2{0, 0});
  std::vector<float> h32(TENSOR_STRIDE * O, 0.f);
  for (int c = 0; c < 7; ++c)
    for (int kz = 0; kz < g.Hz; ++kz)
      for (int ky = 0; ky < g.Hy; ++ky)
        for (int kx = 0; kx < g.Hx; ++kx) {
          const size_t src = (size_t)kx + (size_t)g.Px * (ky + (size_t)g.Py * (kz + (size_t)g.Pz * c));
          const double re = fN[2 * src], im = fN[2 * src + 1];
          double2 v;
          v.x = (re * sre - im * sim) * invP;
          v.y = (re * sim + im * sre) * invP;
          const size_t dst = (kx + (size_t)g.Hx * (ky + (size_t)g.Hy * kz)) * TENSOR_STRIDE + c;
          h[dst] = v;
          const bool odd = (c == 1 || c == 2 || c == 4);
          h32[dst] = (float)(odd ? v.y : v.x);
        }
  G.t64.alloc(TENSOR_STRIDE * O);
  G.t32.alloc(TENSOR_STRIDE * O);
  VH_HIP(hipMemcpy(G.t64.p, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice));
  VH_HIP(hipMemcpy(G.t32.p, h32.data(), h32.size() * sizeof(float), hipMemcpyHostToDevice));
  G.has_tensor = true;
}

}  // namespace vh
