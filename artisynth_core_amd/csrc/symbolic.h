// Host symbolic analysis for the sparse symmetric (indefinite) KKT solver:
// fill-reducing ordering, elimination tree, column counts, relaxed
// supernodes and their row structures.  Everything here is host C++; the
// numeric factorization uploads the resulting structure to the GPU once.
//
// Input convention follows the matrices that maspack.solvers.KKTSolver hands
// to a DirectSolver (KKTSolver.java:171-288): 0-based CSR of the UPPER
// triangle (column >= row), diagonal always stored.  Entries below the
// diagonal are ignored, as an upper-view symmetric solver does
// (cudssBridge.cc:220-233 selects CUDSS_MVIEW_UPPER for symmetric types).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace kkt {

struct SymOptions {
  int ordering = 1;        // 0 = natural order, 1 = nested dissection
  int leaf_size = 16;      // compressed vertices per nested-dissection leaf
  // relaxed supernode amalgamation (column limits and zero fractions): a child
  // merges into its parent when the merged supernode has <= nrelax0 columns, or
  // <= nrelax1 columns and a zero fraction below zrelax0, or <= nrelax2 and below
  // zrelax1, or any width below zrelax2.  Wider merges than CHOLMOD's defaults:
  // fewer, larger fronts and a shallower tree (fewer level launches per solve)
  // for a few percent more fill
  int nrelax0 = 8, nrelax1 = 96, nrelax2 = 256;
  double zrelax0 = 0.8, zrelax1 = 0.3, zrelax2 = 0.1;
  int max_super = 2048;    // widest supernode (the solve kernels' limit); wider
                           // ones are split into a chain
};

struct Symbolic {
  int n = 0;
  std::vector<int> perm;       // perm[new] = old
  std::vector<int> iperm;      // iperm[old] = new
  std::vector<int> parent;     // column elimination tree of the permuted matrix
  std::vector<int> colcount;   // exact column counts of L (diagonal included)
  int nsuper = 0;
  std::vector<int> sup_ptr;    // nsuper+1: first column of each supernode
  std::vector<int> col2sup;    // n
  std::vector<int> sparent;    // nsuper: assembly-tree parent, -1 at roots
  std::vector<int64_t> rows_ptr;  // nsuper+1
  std::vector<int> rows;       // below-diagonal rows of each supernode (sorted)
  std::vector<int> level;      // height in the assembly tree (leaves = 0)
  int nlevels = 0;
  int ndelayed = 0;            // rows ordered after all their neighbours
  int64_t nnzL = 0;            // stored entries of L incl. amalgamation zeros
  int64_t nnzL_exact = 0;      // sum of exact column counts
  double flops = 0;            // sum_s k^3/3 + r k^2 + r^2 k
};

// delay (may be null): rows that must be eliminated after every row they
// couple to (zero-diagonal constraint rows of a KKT system).
int symbolic_analyze(int n, const int* rowOffs, const int* colIdxs,
                     const unsigned char* delay, const SymOptions& opt,
                     Symbolic& S, std::string& err);

// Validates an upper CSR pattern; returns 0 or a negative code with err set.
int validate_pattern(int n, int nnz, const int* rowOffs, const int* colIdxs,
                     std::string& err);

}  // namespace kkt
