// ordering.hpp -- fill-reducing orderings for the multilevel ILDL.
#pragma once
#include <vector>

namespace jdhip {

// Approximate minimum degree ordering of the symmetric pattern given as a
// 0-based CSR of both triangles (the diagonal may be present).  Returns
// order[new] = old.
std::vector<int> amd_order(int n, const int *ptr, const int *idx);

// Reverse Cuthill-McKee ordering of the same kind of pattern (George & Liu;
// one BFS per connected component from a pseudo-peripheral node).  Returns
// order[new] = old.
std::vector<int> rcm_order(int n, const int *ptr, const int *idx);

}  // namespace jdhip
