// matching.hpp -- symmetric weighted matching, scaling and pair-preserving
// ordering of one level (see matching.cpp for the reference routines each
// step restates).
#pragma once
#include <vector>

#include "ildl_factor.hpp"

namespace jdhip {

// SYMSCALE: the power-of-two symmetric scaling vector d (A -> D A D).
std::vector<double> symscale(const SymCsr &A);

struct MatchResult {
    std::vector<int> match;     // row i matched to column match[i]
    std::vector<double> scale;  // power-of-two symmetric scaling from the duals
    bool structurally_singular = false;
};
// Maximum-product perfect matching of |A| on its epsilon graph, with the
// symmetric scaling derived from the dual variables.
MatchResult mwm_symmetric(const SymCsr &A);

// Break the cycles of the matching into 1x1 and 2x2 blocks; order = the
// singletons (ascending), then the 2x2 blocks, two consecutive entries each.
// Returns the number of singletons.
int symwm_pairs(const SymCsr &A, const std::vector<int> &match, std::vector<int> &order);

// Approximate minimum degree on the graph with each 2x2 block merged into
// one node, expanded with the two rows of a block adjacent; pair_start[k] = 1
// where position k starts a 2x2 block.
std::vector<int> pair_amd_order(const SymCsr &A, const std::vector<int> &wm_order, int nsingle,
                                std::vector<char> &pair_start);

}  // namespace jdhip
