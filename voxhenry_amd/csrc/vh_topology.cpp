// Node numbering, incidence matrix and port right-hand sides of a voxel
// geometry (host, linear time).
//
// Restates, in the reference's numbering order:
//   lse_generate_panelIDs.m:31-172   -- faces numbered per axis in the loop
//       orders (x: z,y,x; y: x,z,y; z: y,x,z), a face shared by two adjacent
//       voxels gets one id, x faces first, then y, then z;
//   lse_compute_Ae_matrix.m:29-88    -- Ae(node, current): Jx/Jy/Jz -1 on the
//       lower face and +1 on the upper face; J2d +1/2 on both x faces and -1/2
//       on both y faces; J3d +1/2 on x and y faces and -1 on z faces;
//   lse_assign_exc_grnd_nodes.m:6-94 -- port faces -> node ids;
//   VoxHenry_executer.m:165-192      -- rhs = -sum of the excitation rows of Ae,
//       port rows removed (Ar), rhs truncated to the reduced size.
#include "vh_geom.h"
#include <algorithm>

namespace vh {

namespace {
// One axis of the face numbering: visit voxels with `axis` the fastest index.
void number_axis(const std::vector<int>& occ_id, int L, int M, int N, int axis, std::vector<int64_t>& lo,
                 std::vector<int64_t>& hi, int64_t& count) {
  const int dims[3] = {L, M, N};
  // loop order slowest -> fastest
  int ord[3];
  if (axis == 0) { ord[0] = 2; ord[1] = 1; ord[2] = 0; }
  else if (axis == 1) { ord[0] = 0; ord[1] = 2; ord[2] = 1; }
  else { ord[0] = 1; ord[1] = 0; ord[2] = 2; }
  int64_t ind = 1;
  int c[3];
  for (c[ord[0]] = 0; c[ord[0]] < dims[ord[0]]; ++c[ord[0]])
    for (c[ord[1]] = 0; c[ord[1]] < dims[ord[1]]; ++c[ord[1]])
      for (c[ord[2]] = 0; c[ord[2]] < dims[ord[2]]; ++c[ord[2]]) {
        const int64_t lin = c[0] + (int64_t)L * (c[1] + (int64_t)M * c[2]);
        const int v = occ_id[lin];
        if (v < 0) continue;
        int nb[3] = {c[0], c[1], c[2]};
        nb[axis] += 1;
        bool next = nb[axis] < dims[axis];
        if (next) next = occ_id[nb[0] + (int64_t)L * (nb[1] + (int64_t)M * nb[2])] >= 0;
        lo[v] = ind;
        ++ind;
        hi[v] = ind;
        if (!next) ++ind;
      }
  int64_t mx = 0;
  for (int64_t v : hi) mx = std::max(mx, v);
  count = mx;
}
}  // namespace

void geometry_build_topology(Geometry& G, const std::vector<std::vector<int64_t>>& Pfaces,
                             const std::vector<std::vector<int64_t>>& Nfaces, const std::vector<int64_t>& well,
                             std::vector<double>& rhs_out) {
  const int L = G.g.L, M = G.g.M, N = G.g.N;
  const int nv = G.g.nv;
  const int64_t V = (int64_t)L * M * N;
  std::vector<int> occ(V, -1);
  for (int v = 0; v < nv; ++v) occ[G.vox[v]] = v;
  std::vector<int64_t> lo[3], hi[3];
  int64_t cnt[3];
  for (int d = 0; d < 3; ++d) {
    lo[d].assign(nv, 0);
    hi[d].assign(nv, 0);
    number_axis(occ, L, M, N, d, lo[d], hi[d], cnt[d]);
  }
  // 0-based node ids: x faces, then y (+cnt[0]), then z (+cnt[0]+cnt[1])
  const int64_t off[3] = {-1, cnt[0] - 1, cnt[0] + cnt[1] - 1};
  const int64_t num_nodes = cnt[0] + cnt[1] + cnt[2];
  auto face = [&](int v, int side) -> int64_t {   // side 0..5 = -x,+x,-y,+y,-z,+z
    const int d = side / 2;
    return (side % 2 == 0 ? lo[d][v] : hi[d][v]) + off[d];
  };
  // columns of Ae: (node, value) pairs per current unknown
  const int64_t nc = 5 * (int64_t)nv;
  std::vector<int64_t> colptr(nc + 1, 0);
  for (int64_t j = 0; j < nc; ++j) {
    const int comp = (int)(j / nv);
    colptr[j + 1] = colptr[j] + (comp < 3 ? 2 : comp == 3 ? 4 : 6);
  }
  std::vector<int64_t> row(colptr[nc]);
  std::vector<double> val(colptr[nc]);
  for (int v = 0; v < nv; ++v) {
    for (int d = 0; d < 3; ++d) {
      const int64_t j = (int64_t)d * nv + v, p = colptr[j];
      row[p] = face(v, 2 * d); val[p] = -1.0;
      row[p + 1] = face(v, 2 * d + 1); val[p + 1] = 1.0;
    }
    {
      const int64_t p = colptr[3 * (int64_t)nv + v];
      row[p] = face(v, 0); val[p] = 0.5;
      row[p + 1] = face(v, 1); val[p + 1] = 0.5;
      row[p + 2] = face(v, 2); val[p + 2] = -0.5;
      row[p + 3] = face(v, 3); val[p + 3] = -0.5;
    }
    {
      const int64_t p = colptr[4 * (int64_t)nv + v];
      row[p] = face(v, 0); val[p] = 0.5;
      row[p + 1] = face(v, 1); val[p + 1] = 0.5;
      row[p + 2] = face(v, 2); val[p + 2] = 0.5;
      row[p + 3] = face(v, 3); val[p + 3] = 0.5;
      row[p + 4] = face(v, 4); val[p + 4] = -1.0;
      row[p + 5] = face(v, 5); val[p + 5] = -1.0;
    }
  }
  // port nodes
  auto lookup = [&](const int64_t* r) -> int64_t {   // (x, y, z, side) 1-based
    VH_REQUIRE(r[0] >= 1 && r[0] <= L && r[1] >= 1 && r[1] <= M && r[2] >= 1 && r[2] <= N && r[3] >= 1 && r[3] <= 6,
               "port node outside the grid");
    const int64_t lin = (r[0] - 1) + (int64_t)L * ((r[1] - 1) + (int64_t)M * (r[2] - 1));
    const int v = occ[lin];
    VH_REQUIRE(v >= 0, "port node on an empty voxel");
    return face(v, (int)r[3] - 1);
  };
  const int np = (int)Pfaces.size();
  std::vector<char> drop(num_nodes, 0);
  // rows of Ae as CSR to sum excitation rows
  std::vector<int64_t> rp(num_nodes + 1, 0);
  for (int64_t p = 0; p < colptr[nc]; ++p) rp[row[p] + 1]++;
  for (int64_t i = 0; i < num_nodes; ++i) rp[i + 1] += rp[i];
  std::vector<int64_t> rc(colptr[nc]);
  std::vector<double> rv(colptr[nc]);
  {
    std::vector<int64_t> fill(rp.begin(), rp.end() - 1);
    for (int64_t j = 0; j < nc; ++j)
      for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) { rc[fill[row[p]]] = j; rv[fill[row[p]]] = val[p]; fill[row[p]]++; }
  }
  std::vector<double> rhs_full((size_t)np * nc, 0.0);
  for (int p = 0; p < np; ++p) {
    const auto& P = Pfaces[p];
    for (size_t k = 0; k + 3 < P.size(); k += 4) {
      const int64_t node = lookup(&P[k]);
      drop[node] = 1;
      for (int64_t q = rp[node]; q < rp[node + 1]; ++q) rhs_full[(size_t)p * nc + rc[q]] -= rv[q];
    }
    const auto& Nn = Nfaces[p];
    for (size_t k = 0; k + 3 < Nn.size(); k += 4) drop[lookup(&Nn[k])] = 1;
  }
  for (size_t k = 0; k + 3 < well.size(); k += 4) drop[lookup(&well[k])] = 1;
  // renumber the kept nodes, build Ar in CSC
  std::vector<int64_t> newid(num_nodes, -1);
  int64_t nn = 0;
  for (int64_t i = 0; i < num_nodes; ++i) if (!drop[i]) newid[i] = nn++;
  std::vector<int64_t> cp(nc + 1, 0), ri;
  std::vector<double> vv;
  ri.reserve(row.size());
  vv.reserve(row.size());
  for (int64_t j = 0; j < nc; ++j) {
    // rows within a column in increasing order (CSC canonical form)
    std::vector<std::pair<int64_t, double>> e;
    for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p)
      if (newid[row[p]] >= 0) e.push_back({newid[row[p]], val[p]});
    std::sort(e.begin(), e.end());
    for (auto& t : e) { ri.push_back(t.first); vv.push_back(t.second); }
    cp[j + 1] = (int64_t)ri.size();
  }
  geometry_set_incidence_csc(G, nn, (int64_t)ri.size(), cp.data(), ri.data(), vv.data());
  G.num_nodes_full = num_nodes;
  // rhs rows of length n = nc + nn (node part zero), port-major
  const int64_t n = nc + nn;
  rhs_out.assign((size_t)np * n, 0.0);
  for (int p = 0; p < np; ++p)
    std::copy(rhs_full.begin() + (size_t)p * nc, rhs_full.begin() + (size_t)(p + 1) * nc, rhs_out.begin() + (size_t)p * n);
}

}  // namespace vh
