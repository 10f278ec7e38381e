// ildl_factor.hpp -- host multilevel inverse-based incomplete LDL^T.
//
// Structure of the reference's multilevel ILDL (DSYMAMGfactor,
// src/ilupack/symAMGsetup.c:121-1647, with the Crout kernel DSYMpiluc,
// src/ilupack/sympiluc.F:169-6998): per level the symmetric matrix is scaled,
// reordered and factored by a left-looking (Crout) LDL^T in which
//   * a pivot is deferred to the next level when the running estimate of
//     ||L^{-1}|| (look-ahead, both signs tried, sympiluc.F:1480-1510)
//     exceeds condest;
//   * an entry L(l,k) is dropped when |d_k L(l,k)| max_m |L(m,k)| <=
//     droptol / maxcondest (sympiluc.F:2191-2239);
// the deferred block's Schur complement (with dropping) is the next level,
// and a small final level is factored densely.  A level whose reduction is
// poor (more than 75 % deferred) is refactored with condest doubled
// (symAMGsetup.c:783-809).
#pragma once
#include <cstdint>
#include <vector>

namespace jdhip {

// Symmetric matrix, full pattern (both triangles), 0-based CSR.
struct SymCsr {
    int n = 0;
    std::vector<int> ptr, idx;
    std::vector<double> val;
};

template <typename T>
struct IldlLevel {
    int n = 0, nB = 0;
    std::vector<int> perm;  // perm[i]: level-input index of position i ([B; C])
    std::vector<T> scal;    // scaling by level-input index (powers of two)
    // strict lower L_B (rows/cols = B positions), CSR
    std::vector<int> lb_ptr, lb_idx;
    std::vector<T> lb_val;
    // coupling block (rows = C positions 0..nC-1, cols = B positions), CSR:
    // F^T, the rows of the level's scaled, permuted input matrix between C
    // and B (coupling 0, the reference's structure), or L_CB, the factor's
    // own rows of the deferred pivots (coupling 1)
    std::vector<int> lc_ptr, lc_idx;
    std::vector<T> lc_val;
    std::vector<T> d;  // diagonal of the block diagonal D of B (SPD-converted: |D|)
    // 2x2 pivots: dpair[b] = the other row of b's 2x2 block (adjacent, b +- 1)
    // or -1 for a 1x1 pivot; doff[b] = the block's off-diagonal entry
    std::vector<T> doff;
    std::vector<int> dpair;
};

template <typename T>
struct IldlMultilevel {
    std::vector<IldlLevel<T>> levels;
    // 0: every level stores F^T = A_CB and the apply solves with B twice
    //    (forward and backward block substitution of DSYMAMGsol,
    //    src/ilupack/symAMGsol.c:110-631; the F block is extracted by
    //    DSYMAMGextract, symAMGextract.c:112);
    // 1: every level stores L_CB and the apply is the block LDL^T
    //    substitution [L_B 0; L_CB I] [D 0; 0 S] [L_B^T L_CB^T; 0 I]
    //    (one solve with L_B and one with L_B^T per level)
    int coupling = 0;
    int m = 0;                 // size of the dense final level (0: none)
    std::vector<double> dinv;  // m x m inverse of the (converted) final block
    long nnz_factor = 0;       // stored entries (L, D, dense)
    int nneg = 0;              // negative pivots before SPD conversion
};

struct IldlParams {
    double droptol = 1e-3;
    double condest = 5.0;
    double mem = 0.0;        // elbow: max stored entries / nnz(A_upper) (<=0: none)
    int dense_max = 1500;    // final level solved densely when n <= dense_max
    int max_levels = 40;
    bool spd = false;        // convert pivots to |d| (ISEARCH <= 1)
    int ordering = 2;        // 0 natural, 1 reverse Cuthill-McKee, 2 approximate minimum degree
    int coupling = 0;        // coupling block stored per level (IldlMultilevel::coupling)
    bool matching = true;    // symmetric weighted matching + compressed AMD (ordering 2)
    int threads = 1;         // host threads for the Schur complements
};

struct IldlStats {
    int ierr = 0;  // 0 ok, -3 elbow space exceeded, >0 zero pivot in dense level
    int nlev = 0;
    long nnz = 0;
    int nneg = 0;
    double condest_used = 0;
    std::vector<int> level_n, level_nB;
};

template <typename T>
IldlStats ildl_factor(const SymCsr &A, const IldlParams &prm, IldlMultilevel<T> &out,
                      long nnz_upper_ref);

// Build a SymCsr from a 1-based upper (or lower) triangle CSR with a
// constant `shift` subtracted from the diagonal.
SymCsr sym_from_upper(int n, const int *ia, const int *ja, const double *a, double shift);

}  // namespace jdhip
