/* CPU oracle for the sparse KKT direct-solve path.  TEST INFRASTRUCTURE ONLY:
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
 * as the checker; the product path never calls it.
 *
 * What it restates (reference file:line):
 *  - symmetric SpMV from upper-CSR storage, A x = U x + U^T x - diag(U) x
 *    (cudssBridge.cc:1665-1703 applyA, cudssBridgeKernels.cu:9-24 axmy);
 *  - symmetric-indefinite LDL^T with Bunch-Kaufman 1x1/2x2 pivots and static
 *    pivot perturbation, as configured for PARDISO mtype=-2 (pardisoMkl.cc:
 *    89-110 iparm[9]=8 perturbation 1e-8, iparm[20]=1 Bunch-Kaufman), pivoting
 *    restricted to supernode panels (the PARDISO supernode-BK scheme);
 *  - inertia / perturbed-pivot counts (pardisoMkl.cc:500-509, iparm[13,21,22]);
 *  - BiCGStab with the stale factor as preconditioner
 *    (cudssBridge.cc:1705-2006: x0 = M^-1 b, r_hat = r0, early exit on ||s||,
 *    breakdown threshold 1e-30, returns iterations or a negative code);
 *  - symbolic structure (elimination tree, column counts) by explicit
 *    column-by-column symbolic elimination, an algorithm independent of the
 *    product's row-subtree counting.
 */
#ifndef KKT_ORACLE_H
#define KKT_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_factor orc_factor;

/* y = A x, A symmetric given by its upper CSR (entries below the diagonal ignored). */
void orc_spmv_upper(int n, const int* rowOffs, const int* colIdxs, const double* vals,
                    const double* x, double* y);

/* Symbolic elimination of the permuted pattern: parent[j], colcount[j] (incl. diag).
 * perm[new] = old.  Returns 0 on success. */
int orc_symbolic(int n, const int* rowOffs, const int* colIdxs, const int* perm,
                 int* parent, int* colcount);

/* Supernode row structures for a given partition (sup_ptr, nsuper+1 entries):
 * rows_ptr (nsuper+1), rows (caller buffer of rows_cap ints). Returns total rows or <0. */
long orc_supernode_rows(int n, const int* rowOffs, const int* colIdxs, const int* perm,
                        int nsuper, const int* sup_ptr, long* rows_ptr, int* rows, long rows_cap);

/* Multifrontal LDL^T (FP64) on the given ordering and supernode partition, with
 * Bunch-Kaufman pivoting restricted to panels of `nb` columns inside each
 * supernode and pivots smaller than pert*max|a_ij| perturbed, the maximum taken
 * over the pivot's connected component (its assembly tree).  NULL on failure. */
orc_factor* orc_factor_create(int n, const int* rowOffs, const int* colIdxs, const double* vals,
                              const int* perm, int nsuper, const int* sup_ptr, int nb,
                              double pert);
void orc_factor_free(orc_factor* f);
/* x = A^-1 b using the factor (no refinement). */
void orc_factor_solve(const orc_factor* f, const double* b, double* x);
/* info: [0]=n_pos [1]=n_neg [2]=n_perturbed [3]=n_2x2 [4]=nnzL(stored) */
void orc_factor_info(const orc_factor* f, long* info);
/* BiCGStab preconditioned by `f` on the matrix (rowOffs, colIdxs, vals);
 * returns iterations (>=1) or a negative status (-12 breakdown, -13 no convergence). */
int orc_bicgstab(const orc_factor* f, int n, const int* rowOffs, const int* colIdxs,
                 const double* vals, const double* b, double* x, double tolRel, int maxIter);
/* Fixed-point refinement x <- x + M^-1 (b - A x) until ||b-Ax||/||b|| <= tol;
 * returns the number of correction steps or -1 if not reached in maxSteps. */
int orc_refine(const orc_factor* f, int n, const int* rowOffs, const int* colIdxs,
               const double* vals, const double* b, double* x, double tol, int maxSteps);

#ifdef __cplusplus
}
#endif
#endif
