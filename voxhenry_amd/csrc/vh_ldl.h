// Host sparse LDL^T (ordering, symbolic, numeric) for the Schur matrix.
#pragma once
#include <cstdint>
#include <vector>

namespace vh {

struct LdlFactor {
  int n = 0;
  std::vector<int> perm, iperm;   // C = P S P': C(i,j) = S(perm[i], perm[j])
  std::vector<int> parent;        // elimination tree of C
  std::vector<int> Lp, Li;        // strictly lower L, CSC (rows increasing)
  std::vector<double> Lx;
  std::vector<double> D;
};

// Nested-dissection ordering of the graph of a symmetric pattern (CSR/CSC,
// diagonal entries allowed).  Returns perm (new -> old).
// blocks (optional) receives the sizes of the consecutive blocks of the
// ordering (leaves and separators, in elimination order).
std::vector<int> nd_ordering(int n, const int* Ap, const int* Ai, std::vector<int>* blocks = nullptr);
// Nested dissection by panel planes for node graphs of voxel grids: xyz (3n)
// are face-centre coordinates in voxel units, type[i] the axis the face is
// normal to.
std::vector<int> nd_ordering_geometric(int n, const int* Ap, const int* Ai, const float* xyz, const int* type,
                                       std::vector<int>* blocks = nullptr);
// Symbolic factorisation of the full symmetric pattern under perm.
void ldl_symbolic(int n, const int* Ap, const int* Ai, const std::vector<int>& perm, LdlFactor& F);
// Numeric factorisation (same pattern as the symbolic call).
void ldl_numeric(int n, const int* Ap, const int* Ai, const double* Ax, LdlFactor& F);
// CSR copy of L and forward / backward level numbers of every row.
void ldl_levels(const LdlFactor& F, std::vector<int>& Rp, std::vector<int>& Ri, std::vector<double>& Rx,
                std::vector<int>& lev_f, std::vector<int>& lev_b);

}  // namespace vh
