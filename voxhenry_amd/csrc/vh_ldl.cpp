// Host sparse LDL^T of the SPD Schur matrix S = Ar*diag(A_inv)*Ar'.
//
// Stands in for CHOLMOD 3.0.12 ldlchol (called at lse_sparse_precon_prepare.m:154):
// the factor is exact up to rounding, so the preconditioner apply of
// lse_sparse_precon_multiply.m:84-88 is unchanged whatever ordering is used.
//   * ordering: nested dissection with breadth-first level-set separators
//     (pseudo-peripheral start node, middle level as the vertex separator),
//     which keeps fill low on the grid-like node graphs and the elimination
//     tree shallow for level-scheduled triangular solves on the GPU;
//   * symbolic: elimination tree and column counts by row-subtree traversal;
//   * numeric: up-looking LDL^T (one row of L per step, pattern from the
//     elimination-tree reach), the published algorithm of T. Davis' LDL
//     package (ACM TOMS 31(4), 2005), written here from that description.
#include "vh_ldl.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <queue>
#include <stdexcept>

namespace vh {

namespace {

struct Graph {
  int n;
  const int* p;
  const int* i;
};

// BFS restricted to vertices with mark[v] == stamp; returns levels.
static int bfs_levels(const Graph& g, int start, const std::vector<int>& mark, int stamp, std::vector<int>& level,
                      std::vector<int>& order) {
  order.clear();
  order.push_back(start);
  level[start] = 0;
  size_t head = 0;
  int maxl = 0;
  while (head < order.size()) {
    const int v = order[head++];
    for (int q = g.p[v]; q < g.p[v + 1]; ++q) {
      const int w = g.i[q];
      if (mark[w] == stamp && level[w] < 0) {
        level[w] = level[v] + 1;
        maxl = std::max(maxl, level[w]);
        order.push_back(w);
      }
    }
  }
  return maxl;
}

struct ND {
  const Graph& g;
  std::vector<int> mark, level, order, out, blocks;   // blocks: sizes of the emitted leaf / separator blocks
  int stamp = 0;
  static constexpr int LEAF = 48;

  explicit ND(const Graph& gg) : g(gg), mark(gg.n, -1), level(gg.n, -1) { out.reserve(gg.n); }

  void emit(const std::vector<int>& v) {
    if (v.empty()) return;
    out.insert(out.end(), v.begin(), v.end());
    blocks.push_back((int)v.size());
  }

  void run(std::vector<int> verts) {
    // explicit stack to avoid deep recursion on long chains
    struct Item { std::vector<int> v; int phase; std::vector<int> sep; };
    std::vector<Item> stack;
    stack.push_back({std::move(verts), 0, {}});
    while (!stack.empty()) {
      Item it = std::move(stack.back());
      stack.pop_back();
      if (it.phase == 1) {           // emit the separator after both halves
        emit(it.sep);
        continue;
      }
      std::vector<int>& vs = it.v;
      if ((int)vs.size() <= LEAF) {
        emit(vs);
        continue;
      }
      ++stamp;
      for (int v : vs) { mark[v] = stamp; level[v] = -1; }
      // pseudo-peripheral node: repeat BFS from the farthest node
      int start = vs[0];
      int ecc = bfs_levels(g, start, mark, stamp, level, order);
      if ((int)order.size() < (int)vs.size()) {
        // disconnected: split off this component, handle the rest separately
        std::vector<int> comp(order.begin(), order.end());
        std::vector<int> rest;
        rest.reserve(vs.size() - comp.size());
        for (int v : vs) if (level[v] < 0) rest.push_back(v);
        stack.push_back({std::move(rest), 0, {}});
        stack.push_back({std::move(comp), 0, {}});
        continue;
      }
      for (int rep = 0; rep < 3; ++rep) {
        const int far = order.back();
        for (int v : vs) level[v] = -1;
        const int e2 = bfs_levels(g, far, mark, stamp, level, order);
        if (e2 <= ecc) { break; }
        ecc = e2;
        start = far;
      }
      if (ecc < 2) {
        emit(vs);
        continue;
      }
      // level sizes; separator = the smallest level in the middle band
      std::vector<int> cnt(ecc + 1, 0);
      for (int v : vs) cnt[level[v]]++;
      const int total = (int)vs.size();
      int acc = 0, best = -1;
      long bestscore = -1;
      for (int l = 1; l < ecc; ++l) {
        const int below = acc + cnt[l - 1];
        acc = below;
        const int above = total - below - cnt[l];
        if (below < total / 4 || above < total / 4) continue;
        const long score = (long)cnt[l];
        if (best < 0 || score < bestscore) { best = l; bestscore = score; }
      }
      if (best < 0) {
        // fall back to the level holding the median vertex
        acc = 0;
        for (int l = 0; l <= ecc; ++l) { acc += cnt[l]; if (acc >= total / 2) { best = std::max(1, std::min(l, ecc - 1)); break; } }
      }
      std::vector<int> A, B, S;
      for (int v : vs) {
        if (level[v] < best) A.push_back(v);
        else if (level[v] > best) B.push_back(v);
        else S.push_back(v);
      }
      stack.push_back({{}, 1, std::move(S)});
      stack.push_back({std::move(B), 0, {}});
      stack.push_back({std::move(A), 0, {}});
    }
  }
};

}  // namespace

std::vector<int> nd_ordering(int n, const int* Ap, const int* Ai, std::vector<int>* blocks) {
  Graph g{n, Ap, Ai};
  ND nd(g);
  std::vector<int> all(n);
  std::iota(all.begin(), all.end(), 0);
  nd.run(std::move(all));
  if ((int)nd.out.size() != n) throw std::runtime_error("nested dissection lost vertices");
  if (blocks) *blocks = nd.blocks;
  return nd.out;
}

// Geometric nested dissection for the panel (node) graph of a voxel grid.
// Node i is the face of a voxel normal to axis type[i] at coordinate xyz[i]
// (voxel units, faces normal to the axis at integer positions, the other two
// coordinates at half-integers).  A voxel couples only its own six faces, so
// the faces normal to axis d at the integer plane c separate every node with
// coordinate < c from every node with coordinate > c.  Each step cuts the
// longest extent of the current node set at its median plane.
std::vector<int> nd_ordering_geometric(int n, const int* Ap, const int* Ai, const float* xyz, const int* type,
                                       std::vector<int>* blocks) {
  std::vector<int> out, bl;
  out.reserve(n);
  auto emit = [&](const std::vector<int>& v) {
    if (v.empty()) return;
    out.insert(out.end(), v.begin(), v.end());
    bl.push_back((int)v.size());
  };
  struct Item { std::vector<int> v; int phase; std::vector<int> sep; };
  std::vector<Item> stack;
  {
    std::vector<int> all(n);
    std::iota(all.begin(), all.end(), 0);
    stack.push_back({std::move(all), 0, {}});
  }
  const int LEAF = 32;
  Graph g{n, Ap, Ai};
  std::vector<int> pos;
  while (!stack.empty()) {
    Item it = std::move(stack.back());
    stack.pop_back();
    if (it.phase == 1) { emit(it.sep); continue; }
    std::vector<int>& vs = it.v;
    if ((int)vs.size() <= LEAF) { emit(vs); continue; }
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (int v : vs)
      for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], xyz[3 * v + d]); hi[d] = std::max(hi[d], xyz[3 * v + d]); }
    int axes[3] = {0, 1, 2};
    std::sort(axes, axes + 3, [&](int a, int b) { return hi[a] - lo[a] > hi[b] - lo[b]; });
    bool done = false;
    for (int ai = 0; ai < 3 && !done; ++ai) {
      const int d = axes[ai];
      pos.clear();
      for (int v : vs) if (type[v] == d) pos.push_back((int)std::lround(xyz[3 * v + d]));
      if (pos.size() < 2) continue;
      std::nth_element(pos.begin(), pos.begin() + pos.size() / 2, pos.end());
      const int c = pos[pos.size() / 2];
      std::vector<int> A, B, S;
      for (int v : vs) {
        const float cv = xyz[3 * v + d];
        if (type[v] == d && (int)std::lround(cv) == c) S.push_back(v);
        else if (cv < (float)c) A.push_back(v);
        else B.push_back(v);
      }
      if (A.empty() || B.empty()) continue;
      stack.push_back({{}, 1, std::move(S)});
      stack.push_back({std::move(B), 0, {}});
      stack.push_back({std::move(A), 0, {}});
      done = true;
    }
    if (!done) {
      // no plane splits this set: order it with the graph dissection
      ND nd(g);
      nd.run(vs);
      out.insert(out.end(), nd.out.begin(), nd.out.end());
      bl.insert(bl.end(), nd.blocks.begin(), nd.blocks.end());
    }
  }
  if ((int)out.size() != n) throw std::runtime_error("geometric dissection lost vertices");
  if (blocks) *blocks = bl;
  return out;
}

void ldl_symbolic(int n, const int* Ap, const int* Ai, const std::vector<int>& perm, LdlFactor& F) {
  F.n = n;
  F.perm = perm;
  F.iperm.assign(n, 0);
  for (int k = 0; k < n; ++k) F.iperm[perm[k]] = k;
  F.parent.assign(n, -1);
  std::vector<int> flag(n), lnz(n, 0);
  for (int k = 0; k < n; ++k) {
    flag[k] = k;
    const int kk = perm[k];
    for (int p = Ap[kk]; p < Ap[kk + 1]; ++p) {
      int i = F.iperm[Ai[p]];
      if (i < k) {
        for (; flag[i] != k; i = F.parent[i]) {
          if (F.parent[i] == -1) F.parent[i] = k;
          lnz[i]++;
          flag[i] = k;
        }
      }
    }
  }
  F.Lp.assign(n + 1, 0);
  for (int k = 0; k < n; ++k) F.Lp[k + 1] = F.Lp[k] + lnz[k];
  const int64_t nnz = F.Lp[n];
  F.Li.assign(nnz, 0);
  F.Lx.assign(nnz, 0.0);
  F.D.assign(n, 0.0);
}

void ldl_numeric(int n, const int* Ap, const int* Ai, const double* Ax, LdlFactor& F) {
  std::vector<double> Y(n, 0.0);
  std::vector<int> pattern(n), flag(n), lnz(n, 0);
  for (int k = 0; k < n; ++k) {
    Y[k] = 0.0;
    int top = n;
    flag[k] = k;
    lnz[k] = 0;
    const int kk = F.perm[k];
    for (int p = Ap[kk]; p < Ap[kk + 1]; ++p) {
      int i = F.iperm[Ai[p]];
      if (i <= k) {
        Y[i] += Ax[p];
        int len = 0;
        for (; flag[i] != k; i = F.parent[i]) {
          pattern[len++] = i;
          flag[i] = k;
        }
        while (len > 0) pattern[--top] = pattern[--len];
      }
    }
    double dk = Y[k];
    Y[k] = 0.0;
    for (; top < n; ++top) {
      const int i = pattern[top];
      const double yi = Y[i];
      Y[i] = 0.0;
      const int p2 = F.Lp[i] + lnz[i];
      for (int p = F.Lp[i]; p < p2; ++p) Y[F.Li[p]] -= F.Lx[p] * yi;
      const double lki = yi / F.D[i];
      dk -= lki * yi;
      F.Li[p2] = k;
      F.Lx[p2] = lki;
      lnz[i]++;
    }
    if (!(dk > 0.0) || !std::isfinite(dk))
      throw std::runtime_error("LDL^T: non-positive pivot at step " + std::to_string(k) + " (Schur matrix not SPD)");
    F.D[k] = dk;
  }
}

void ldl_levels(const LdlFactor& F, std::vector<int>& Rp, std::vector<int>& Ri, std::vector<double>& Rx,
                std::vector<int>& lev_f, std::vector<int>& lev_b) {
  const int n = F.n;
  // CSR of L (rows) from its CSC
  Rp.assign(n + 1, 0);
  for (int j = 0; j < n; ++j)
    for (int p = F.Lp[j]; p < F.Lp[j + 1]; ++p) Rp[F.Li[p] + 1]++;
  for (int i = 0; i < n; ++i) Rp[i + 1] += Rp[i];
  Ri.assign(Rp[n], 0);
  Rx.assign(Rp[n], 0.0);
  std::vector<int> fill(Rp.begin(), Rp.end() - 1);
  for (int j = 0; j < n; ++j)
    for (int p = F.Lp[j]; p < F.Lp[j + 1]; ++p) {
      const int i = F.Li[p];
      Ri[fill[i]] = j;
      Rx[fill[i]] = F.Lx[p];
      fill[i]++;
    }
  lev_f.assign(n, 0);
  for (int i = 0; i < n; ++i) {
    int l = 0;
    for (int p = Rp[i]; p < Rp[i + 1]; ++p) l = std::max(l, lev_f[Ri[p]] + 1);
    lev_f[i] = l;
  }
  lev_b.assign(n, 0);
  for (int i = n - 1; i >= 0; --i) {
    int l = 0;
    for (int p = F.Lp[i]; p < F.Lp[i + 1]; ++p) l = std::max(l, lev_b[F.Li[p]] + 1);
    lev_b[i] = l;
  }
}

}  // namespace vh
