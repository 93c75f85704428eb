/*
 * kkt_mi355x.h -- C ABI of the MI355X sparse KKT direct solver.
 *
 * A drop-in native backend for ArtiSynth's maspack.solvers.DirectSolver
 * contract (DirectSolver.java:26-254).  Each entry point below replaces one
 * native method of the reference's cuDSS bridge (CuDssSolver.java:87-144,
 * implemented in CuDssJNI.cc and cudssBridge.cc); the citation is given per
 * function.  Plain pointers and sizes only; indices are 0-based; symmetric
 * matrices are passed as the CSR of their UPPER triangle, exactly as
 * KKTSolver builds them (KKTSolver.java:171-288, 386-454).
 *
 * Status codes: 0 = OK, negative = error (kkt_last_error() has the text).
 * The first thirteen codes keep the meaning of CUDSS_BRIDGE_* in
 * cudssBridge.h:39-52; codes below -13 are new.
 */
#ifndef KKT_MI355X_H
#define KKT_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

/* matrix types, same values as CUDSS_BRIDGE_MT_* (cudssBridge.h:35-37) */
#define KKT_MT_GENERAL 0
#define KKT_MT_SYMMETRIC 1
#define KKT_MT_SPD 2

#define KKT_OK 0
#define KKT_ERR_INIT -1
#define KKT_ERR_ALLOC -2
#define KKT_ERR_COPY -3
#define KKT_ERR_PATTERN -4
#define KKT_ERR_ANALYSIS -5
#define KKT_ERR_FACTORIZATION -6
#define KKT_ERR_SOLVE -7
#define KKT_ERR_STATE -8
#define KKT_ERR_UNKNOWN_MTYPE -9
#define KKT_ERR_SPMV -10
#define KKT_ERR_REDUCTION -11
#define KKT_ERR_ITER_BREAKDOWN -12
#define KKT_ERR_ITER_NO_CONVERGE -13
#define KKT_ERR_ARGUMENT -14
#define KKT_ERR_DEVICE -15
#define KKT_ERR_REFINE_STALL -16

/* factor precision */
#define KKT_PREC_FP64 0        /* FP64 fronts, FP64 solves */
#define KKT_PREC_FP32_REFINE 1 /* FP32 fronts on MFMA + FP64 iterative refinement */

typedef struct kkt_solver kkt_solver;

/* Statistics of the current symbolic analysis / factor. */
typedef struct kkt_info {
  int n;
  int nsuper;
  int nlevels;
  int ndelayed;          /* zero-diagonal rows ordered after their neighbours */
  long long nnz;         /* upper-triangle entries of A */
  long long nnzL;        /* stored entries of L (with amalgamation zeros) */
  long long nnzL_exact;  /* exact column-count sum */
  double factor_flops;   /* sum over supernodes of k^3/3 + r k^2 + r^2 k */
  long long front_bytes; /* HBM bytes held by the fronts */
  int npos, nneg;        /* inertia of D (PARDISO iparm[21], iparm[22]) */
  int nperturbed;        /* perturbed pivots (PARDISO iparm[13]) */
  int n2x2;              /* 2x2 Bunch-Kaufman pivots */
  int precision;         /* KKT_PREC_* of the current factor */
  int refine_steps;      /* refinement steps of the last solve */
  int fp64_fallbacks;    /* FP64 refactors forced by stalled refinement */
  double last_rel_residual; /* ||b - A x|| / ||b|| of the last refined solve */
  int pivoted_blocks;    /* 64-column diagonal blocks that needed Bunch-Kaufman
                            interchanges or 2x2 pivots (the rest took 1x1 pivots
                            in order) */
  long long device_bytes;        /* device memory the handle holds now */
  long long factored_bytes;      /* ... once factored in the requested precision
                                    (structure, pattern, fronts, solve vectors) */
  long long fp64_fallback_bytes; /* what an FP64 refactor after stalled FP32
                                    refinement would add (0 if already held) */
  int precond_half;      /* 1: the iterative solve's preconditioner reads the
                            half-precision copy of this FP32 factor's solve
                            arrays (kkt_set_iterative_half) */
} kkt_info;

/* CuDssSolver.doInit (CuDssJNI.cc:11-23): creates a handle (NULL on failure). */
kkt_solver* kkt_create(void);
/* CuDssSolver.doDispose (CuDssJNI.cc:718-724; cudssBridge.cc:2008-2023).
 * Releases all device memory; safe on NULL. */
void kkt_destroy(kkt_solver* h);

/* CuDssSolver.doSetPattern (CuDssJNI.cc:25-45; cudssBridge.cc:291-382).
 * Copies the pattern; drops any previous analysis and factor. */
int kkt_set_pattern(kkt_solver* h, int n, int nnz, const int* rowOffs, const int* colIdxs,
                    int mtype);
/* CuDssSolver.doAnalyze (CuDssJNI.cc:47-53; cudssBridge.cc:384-407).
 * Pattern-only analysis; rows whose diagonal turns out to be zero at the
 * first factor are re-ordered then (once per pattern). */
int kkt_analyze(kkt_solver* h);
/* PardisoSolver.analyze(vals,...) (PardisoSolver.java:1378-1432, phase 11 with
 * values, pardisoMkl.cc:440-511): analysis that also sees the values, so
 * zero-diagonal (constraint) rows are ordered after their neighbours at once. */
int kkt_analyze_values(kkt_solver* h, const double* vals);
/* CuDssSolver.doFactor (CuDssJNI.cc:55-70; cudssBridge.cc:409-455): host
 * values (nnz doubles) -> numeric LDL^T on the GPU. */
int kkt_factor(kkt_solver* h, const double* vals);
/* Same with the values already resident in device memory (no H2D copy). */
int kkt_factor_device(kkt_solver* h, const double* d_vals);

/* Device-side value assembly (the step before the factor).  The handle keeps
 * one device CSR value buffer per analyzed pattern (nnz doubles, the array
 * kkt_factor copies into); these calls edit it in place and
 * kkt_factor_device_values factors it.  All require a previous analyze.
 * Slots are 0-based positions in the CSR value array; negative slots are
 * skipped.  Contributions to one slot are summed in a fixed order (no atomics),
 * so results are bitwise reproducible.  Slot lists are planned once and cached
 * by content (up to 16 per handle). */
/* CuDssSolver.doClearDeviceValues (CuDssJNI.cc:66-71; cudssBridge.h:88): zero. */
int kkt_clear_device_values(kkt_solver* h);
/* CuDssSolver.doGetDeviceValues (CuDssJNI.cc:73-84; cudssBridge.h:135): copy
 * the buffer to vals (nnz doubles). */
int kkt_get_device_values(kkt_solver* h, double* vals);
/* CuDssSolver.doAddDeviceValues (CuDssJNI.cc:86-104; cudssBridge.h:89):
 * buf[slots[i]] += scale * vals[i], i < nvals. */
int kkt_add_device_values(kkt_solver* h, const int* slots, const double* vals, int nvals,
                          double scale);
/* CuDssSolver.doAddScaledDiagonal3DeviceValues (CuDssJNI.cc:106-129):
 * buf[diagSlots[3i+k]] += scale * masses[i], i < nblocks, k < 3. */
int kkt_add_scaled_diagonal3_device_values(kkt_solver* h, const int* diagSlots,
                                           const double* masses, int nblocks, double scale);
/* CuDssSolver.doAddScaledBlock3DeviceValues (CuDssJNI.cc:131-162):
 * buf[blockSlots[9i+j]] += scale * blockScales[i] * blockVals[9i+j]. */
int kkt_add_scaled_block3_device_values(kkt_solver* h, const int* blockSlots,
                                        const double* blockVals, const double* blockScales,
                                        int nblocks, double scale);
/* CuDssSolver.doAddMaterialStiffness3DeviceValues (CuDssJNI.cc:164-202;
 * FemUtilities.addMaterialStiffness): per block i the 3x3 matrix
 * K = dv_i (B(gi)^T D_i B(gj) + (gi^T sigma_i gj) I), B the 6x3 Voigt
 * strain-displacement matrix (order xx,yy,zz,xy,yz,xz), D_i 6x6 row-major,
 * sigma_i = [s00,s11,s22,s01,s12,s02]; buf[blockSlots[9i+j]] += scale * K[j]
 * (row-major). */
int kkt_add_material_stiffness3_device_values(kkt_solver* h, const int* blockSlots,
                                              const double* gis, const double* gjs,
                                              const double* Ds, const double* sigmas,
                                              const double* dvs, int nblocks, double scale);
/* Element-batched forms.  Per element e: elemNodeCounts[e] nodes; node pairs
 * p in [elemPairOffsets[e], elemPairOffsets[e+1]) with element-local node
 * indices pairNodeIdxs[2p], pairNodeIdxs[2p+1]; integration points in
 * [elemIpOffsets[e], elemIpOffsets[e+1]), each with elemNodeCounts[e] gradients
 * starting at the element's gradient offset.  Each element node q (global index
 * elemNodeOffsets[e] + local) has a target dimension nodeDims[q] (3 free, 6
 * frame-attached) and an 18-double row-major (dim x 3) transform T_q; the pair
 * block K_ij (summed over integration points, in order) is added as
 * T_i K_ij T_j^T at blockSlots[36p + r * dim_j + c] (row-major
 * dim_i x dim_j; the rest of the 36 are unused).  Negative slots are skipped. */
/* CuDssSolver.doAddMaterialStiffness3ElementDeviceValues (CuDssJNI.cc:203-311;
 * cudssBridgeKernels.cu:381-449): K_ij = the material stiffness block above with
 * the integration point's D (6x6), sigma (6) and dv, and gradients g_i, g_j. */
int kkt_add_material_stiffness3_element_device_values(
    kkt_solver* h, const int* elemNodeCounts, const int* elemNodeOffsets, const int* elemPairOffsets,
    const int* elemIpOffsets, const int* elemGradOffsets, const int* pairNodeIdxs, const int* blockSlots,
    const int* nodeDims, const double* nodeTransforms, const double* grads, const double* Ds,
    const double* sigmas, const double* dvs, int nelems, double scale);
/* CuDssSolver.doAddLinearElasticStiffness3ElementDeviceValues (CuDssJNI.cc:312-391;
 * cudssBridgeKernels.cu:451-512): isotropic linear elasticity from
 * elemParams[2e..2e+1] = (E, nu), sigma = 0, no node reduction: 9 slots per
 * pair at blockSlots[9p + 3r + c]. */
int kkt_add_linear_elastic_stiffness3_element_device_values(
    kkt_solver* h, const int* elemNodeCounts, const int* elemPairOffsets, const int* elemIpOffsets,
    const int* elemGradOffsets, const int* pairNodeIdxs, const int* blockSlots, const double* elemParams,
    const double* grads, const double* dvs, int nelems, double scale);
/* CuDssSolver.doAddLinearElasticStiffness3ElementGeometryDeviceValues
 * (CuDssJNI.cc:392-514; cudssBridgeKernels.cu:531-640): as the linear elastic
 * form, with each integration point's Jacobian J = sum_n x_n (dN_n/dxi)^T from
 * elemNodePositions (3 per element node) and naturalGrads, spatial gradients
 * J^-T dN/dxi and dv = det(J) ipWeights[ip]; node reduction as above. */
int kkt_add_linear_elastic_stiffness3_element_geometry_device_values(
    kkt_solver* h, const int* elemNodeCounts, const int* elemNodeOffsets, const int* elemPairOffsets,
    const int* elemIpOffsets, const int* elemNaturalGradOffsets, const int* pairNodeIdxs,
    const int* blockSlots, const int* nodeDims, const double* nodeTransforms, const double* elemParams,
    const double* elemNodePositions, const double* naturalGrads, const double* ipWeights, int nelems,
    double scale);
/* CuDssSolver.doAddDilationalStiffness3ElementDeviceValues (CuDssJNI.cc:515-628;
 * cudssBridgeKernels.cu:658-717): K_ij = sum_{a,b < np} c_{i,a} Rinv(a,b) c_{j,b}^T
 * with np = elemPressureCounts[e] (elements with np <= 0 add nothing), the
 * 3-vectors c_{i,a} = constraints[3 (elemConstraintOffsets[e] + i np + a)] and
 * Rinv row-major at rinvs[elemRinvOffsets[e]]; node reduction as above. */
int kkt_add_dilational_stiffness3_element_device_values(
    kkt_solver* h, const int* elemNodeCounts, const int* elemNodeOffsets, const int* elemPressureCounts,
    const int* elemPairOffsets, const int* elemConstraintOffsets, const int* elemRinvOffsets,
    const int* pairNodeIdxs, const int* blockSlots, const int* nodeDims, const double* nodeTransforms,
    const double* constraints, const double* rinvs, int nelems, double scale);
/* CuDssSolver.doFactorDeviceValues (CuDssJNI.cc:629-652; cudssBridge.h:130):
 * factor the device value buffer (no host copy). */
int kkt_factor_device_values(kkt_solver* h);
/* CuDssSolver.doSolve (CuDssJNI.cc:640-651; cudssBridge.cc:1376-1417):
 * x = A^-1 b (host arrays of n doubles), plus the configured refinement. */
int kkt_solve(kkt_solver* h, const double* b, double* x);
/* Same on device arrays. */
int kkt_solve_device(kkt_solver* h, const double* d_b, double* d_x);
/* No counterpart in the reference: kkt_solve_device for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Allocates nothing: the call's working set
 * (triangular-solve vectors, refinement residual and correction, reduction
 * partials) lives in a static 64 MiB device region per GPU, held from the
 * library's load until the process ends; a system whose working set is larger
 * (about 16 (4n + sum r + sum m) bytes with FP64 fronts) is rejected with
 * KKT_ERR_ARGUMENT.  Launches are eager, and a stalled FP32 refinement returns
 * KKT_ERR_REFINE_STALL rather than refactoring in FP64.  Every call on a GPU
 * shares that GPU's region and leaves its contents behind: calls that overlap on
 * one GPU (from several threads or handles) race. */
int kkt_solve_device_noalloc(kkt_solver* h, const double* d_b, double* d_x);
/* CuDssSolver.doSolveMulti (cudssBridge.cc:1454-1523; CuDssSolver.java:814-826):
 * column-major n x nrhs.  The right-hand sides are solved eight at a time in one
 * launch sequence (each front's factor data read once per eight), with the
 * configured refinement per column; each column's result equals its kkt_solve
 * result. */
int kkt_solve_multi(kkt_solver* h, int nrhs, const double* B, double* X);
/* Same on device arrays (B and X must not overlap). */
int kkt_solve_multi_device(kkt_solver* h, int nrhs, const double* d_B, double* d_X);
/* CuDssSolver.doMultiply (CuDssJNI.cc:654-675; cudssBridge.cc:1419-1452;
 * CuDssSolver.java:443-465): y = A x with the matrix in the device value buffer
 * (the last factor's or iterative solve's values, or device assembly).  Needs a
 * previous analyze. */
int kkt_multiply(kkt_solver* h, const double* x, double* y);
/* CuDssSolver.doIterativeSolve (CuDssJNI.cc:700-716; cudssBridge.cc:1705-2006):
 * BiCGStab on the NEW values `vals`, preconditioned by the existing (stale)
 * factor.  Returns iterations (>= 1) or a negative status. */
int kkt_iterative_solve(kkt_solver* h, const double* vals, const double* b, double* x,
                        double tolRel, int maxIter);
/* Same on device arrays. */
int kkt_iterative_solve_device(kkt_solver* h, const double* d_vals, const double* d_b,
                               double* d_x, double tolRel, int maxIter);
/* Group iterative solves (no counterpart in the reference, whose bridges each
 * drive one handle on its own stream, cudssBridge.cc:162-179): the
 * kkt_iterative_solve_device of nh distinct factored handles in lockstep, each
 * step of a BiCGStab iteration one launch for all handles still iterating, so
 * many small systems share every launch.  Handle i solves with d_vals[i], d_b[i]
 * into d_x[i]; status[i] receives what kkt_iterative_solve_device would return
 * for it (iterations >= 1, or a negative status; the handle's last error is set
 * as by that call), and its solution is bitwise the one that call gives.
 * Handles whose current factors differ in precision (FP64, FP32, FP32 with the
 * half-precision preconditioner copy) iterate as separate lockstep groups, one
 * after the other.  No handle may be in use elsewhere during the call.  The group object owns
 * a stream and staging buffers; one group object is driven by one thread at a
 * time.  Returns KKT_OK, or an error code (then every handle not yet finished
 * has that status). */
#define KKT_GROUP_MAX 128
typedef struct kkt_group kkt_group;
kkt_group* kkt_group_create(void);
void kkt_group_destroy(kkt_group* g);
int kkt_group_iterative_solve_device(kkt_group* g, int nh, kkt_solver* const* hs,
                                     const double* const* d_vals, const double* const* d_b,
                                     double* const* d_x, double tolRel, int maxIter, int* status);
/* The same on host arrays: vals[i], b[i] are copied to handle i's own device
 * buffers (as kkt_iterative_solve does) and x[i] is written on return for every
 * handle whose status is positive; bitwise what kkt_iterative_solve gives. */
int kkt_group_iterative_solve(kkt_group* g, int nh, kkt_solver* const* hs, const double* const* vals,
                              const double* const* b, double* const* x, double tolRel, int maxIter,
                              int* status);
const char* kkt_group_last_error(kkt_group* g);
/* Frees the factor's device memory (fronts, packed solve arrays, workspaces,
 * solve vectors) and keeps the analysis, its plan and the pattern resident:
 * the handle is back in the analyzed state, and the next factor allocates again
 * and needs no new analysis (no counterpart in the reference; for callers that
 * hold many systems under a device-memory budget, as the serving loop does when
 * it evicts an idle instance). */
int kkt_release_factor(kkt_solver* h);
/* CuDssSolver.doSetIrSteps (cudssBridge.cc setIterativeRefinementSteps):
 * refinement steps after each FP64 solve (default 0; FP32 factors always
 * refine to the refinement tolerance). */
int kkt_set_ir_steps(kkt_solver* h, int nsteps);
/* CuDssSolver.doGetLastError (cudssBridge.h:187-189): NULL after success. */
const char* kkt_last_error(kkt_solver* h);
/* CuDssSolver.doGetVersion. */
const char* kkt_version(void);
/* CuDssSolver.doSetTimingEnabled (cudssBridge.cc:79-91): per-phase stderr lines. */
void kkt_set_timing_enabled(int on);

/* ---- extensions of this backend (no counterpart in the cuDSS bridge) ---- */
/* KKT_PREC_FP64 or KKT_PREC_FP32_REFINE; applies from the next factor. */
int kkt_set_precision(kkt_solver* h, int prec);
/* relative-residual target of the FP64 refinement of FP32 factors (default 1e-10) */
int kkt_set_refine_tol(kkt_solver* h, double tol, int max_steps);
/* 1 (default): the factor and triangular-solve launch sequences are captured once
 * per analysis and precision into HIP graphs and replayed; 0: launched eagerly. */
int kkt_set_graphs(kkt_solver* h, int on);
/* ordering: 0 natural, 1 nested dissection (default) */
int kkt_set_ordering(kkt_solver* h, int ordering);
/* Widest supernode the analysis builds (64..2048 columns, default 2048; wider
 * separators become chains), applied from the next analysis.  No counterpart in
 * the reference.  Wider supernodes mean fewer tree levels, so fewer launches per
 * triangular solve, for more fill and factor work from the relaxed amalgamation:
 * on a 210k-DOF tet box 2048 gives 19 levels and 793 GFLOP per factor, 1024
 * gives 27 levels and 705 GFLOP.  Solve-heavy callers (stale-factor iterations)
 * gain from the default, direct-solve-heavy ones may prefer 1024. */
int kkt_set_max_supernode(kkt_solver* h, int cols);
/* 1: an FP32 factor also keeps a half-precision copy of its packed
 * triangular-solve arrays (2 bytes an entry, half the FP32 arrays' size), and the
 * stale-factor iterative solve (kkt_iterative_solve*, the BiCGStab
 * preconditioner M^-1) reads that copy, with products and sums in FP32; direct
 * solves and their refinement keep the FP32 arrays.  A factor with an entry
 * beyond the half range (|v| >= 65504) keeps to the FP32 arrays (kkt_info
 * precond_half = 0).  0 (default): the preconditioner reads the FP32 arrays (on
 * the 88k-DOF beam sequence the copy makes each BiCGStab iteration 3% faster
 * but needs 15% more iterations).  Applies from the next factor.  No counterpart in the reference: the preconditioner only
 * steers the iteration, whose residual test is FP64, so the solution meets the
 * same tolerance. */
int kkt_set_iterative_half(kkt_solver* h, int on);
int kkt_get_info(kkt_solver* h, kkt_info* out);
/* Copies of the host symbolic structure (for structure-parity checks):
 * perm (n), sup_ptr (nsuper+1), sparent (nsuper), rows_ptr (nsuper+1, int64),
 * rows (rows_ptr[nsuper]).  Any pointer may be NULL. */
int kkt_get_symbolic(kkt_solver* h, int* perm, int* sup_ptr, int* sparent, long long* rows_ptr,
                     int* rows);
/* Reads back the device-resident structure arrays and compares them with the
 * host arrays byte for byte; returns 0 when identical, 1 when not. */
int kkt_verify_device_structure(kkt_solver* h);
/* Kernel timing (HIP events on the solver's stream) accumulated since the
 * last reset, per kernel class: 0 assemble, 1 panel-diag (+ inverse rows),
 * 2 panel-trsm (+ the solve-form post-passes), 3 panel-update (MFMA),
 * 4 forward solve (one count per k_fwd_level launch), 5 backward solve (one per
 * k_bwd_cols launch), 6 spmv, 7 vector ops (including the D^-1 pass).
 * enable: 0 off, -1 all classes, else a bit mask (1 << class).  Factors and
 * solves replay their captured HIP graphs unless one of their classes is timed
 * (then they launch eagerly), so time outside a measured region.
 * ms[c], launches[c]; `update_flops` = algorithmic flops of class-3 launches. */
void kkt_timing_reset(kkt_solver* h, int enable);
/* Solve-shape autotuner (no counterpart in the reference; the analysis picks a
 * default).  The forward-rows and backward-columns kernels give each row or
 * column of a front 2^lg lanes, the smallest power of two with
 * epl * 2^lg >= its length (16 by default).  The candidates are the nine
 * uniform (forward, backward) pairs from {8, 16, 32} and a per-level pick from
 * timing each tree level's launches (HIP events, best of reps), which also
 * chooses per level between one forward launch (each row block of a large front
 * gathers the front's right-hand side itself) and two (a gather launch, then the
 * rows; the same sums in the same order, so bitwise the same); every candidate
 * is timed over 4 reps whole triangular solves as solves run (captured graph
 * or eager), on the current factor, and the fastest is kept.  The captured
 * solve graphs are dropped.  Needs a factor.  A shape changes only the lane mapping, so
 * solutions differ at rounding level (the summation order of a row or column);
 * they stay bitwise reproducible for a given shape. */
int kkt_autotune_solve(kkt_solver* h, int reps);
/* The current per-level shapes (kkt_info.nlevels entries each; NULL skips). */
int kkt_get_solve_shapes(kkt_solver* h, int* fwd_elems, int* bwd_elems);
/* The forward form per tree level (kkt_info.nlevels entries): 1 = one launch
 * (the default), 2 = a gather launch, then the rows (autotuner's pick). */
int kkt_get_forward_forms(kkt_solver* h, int* launches);
/* Sets the forward form per tree level (1 or 2, as above), e.g. to pin a tuned
 * plan; the solution is bitwise the same either way. */
int kkt_set_forward_forms(kkt_solver* h, const int* launches);
/* Row groups per forward row block of the one-launch form, per tree level
 * (kkt_info.nlevels entries; 1, 2, 4 or 8): the row blocks of a large front
 * each gather the front's right-hand side once for that many groups of rows.
 * The default leaves about 1024 blocks per level; the autotuner times 1-8 per
 * level.  The solution is bitwise the same for any choice (no counterpart in
 * the reference). */
int kkt_get_forward_row_groups(kkt_solver* h, int* groups);
int kkt_set_forward_row_groups(kkt_solver* h, const int* groups);
int kkt_timing_get(kkt_solver* h, double* ms, long long* launches, double* update_flops);
/* Kernel durations without per-launch events (no counterpart in the
 * reference): one forward sweep (what 0, the k_fwd_level launches of a
 * triangular solve), one backward sweep (1, the k_bwd_cols launches) or one
 * residual SpMV (2) on the current factor, captured into a graph and replayed
 * `reps` times between two HIP events on the handle's stream.  Returns the
 * average duration of one launch and the launches per replay. */
int kkt_time_kernels(kkt_solver* h, int what, int reps, double* ms_per_launch, int* launches_per_rep);

/* ---- batched systems ----
 * Many independent KKT systems of uneven size (the contact / rigid-body
 * mixes of a scene) factored and solved together.  No counterpart in the
 * reference, whose cuDSS bridge keeps one handle and stream per system
 * (cudssBridge.cc:157-218, CuDssJNI.cc:11-23); each system's arrays have the
 * DirectSolver array layout of kkt_set_pattern (KKTSolver.java:171-454).
 * The systems become one block-diagonal matrix with one assembly tree per
 * system, so each factor and each triangular solve is one captured launch
 * sequence covering every system.  Pivot perturbation thresholds and inertia
 * are per system, and refinement stops per system on that system's own
 * residual, so each result is what the system gets alone (within the
 * refinement tolerance; bitwise when the systems' refinement decisions match). */
typedef struct kkt_batch kkt_batch;
typedef struct kkt_batch_sysinfo {
  int n;                    /* rows of the system */
  int npos, nneg;           /* inertia of D */
  int nperturbed;           /* perturbed pivots */
  int n2x2;                 /* 2x2 Bunch-Kaufman pivots */
  int refine_steps;         /* refinement steps of the last solve */
  double last_rel_residual; /* ||b - A x|| / ||b|| of the last refined solve (0 without refinement) */
} kkt_batch_sysinfo;
kkt_batch* kkt_batch_create(void);
void kkt_batch_destroy(kkt_batch* b);
/* nsys systems; system q has sysN[q] rows.  rowOffs holds each system's own
 * 0-based upper-CSR row offsets one system after another (sysN[q] + 1 entries
 * each, starting at 0), colIdxs each system's local column indices one system
 * after another.  Values, right-hand sides and solutions are passed the same
 * way: concatenated in system order. */
int kkt_batch_set_systems(kkt_batch* b, int nsys, const int* sysN, const int* rowOffs, const int* colIdxs,
                          int mtype);
/* Analysis of every system (vals: the concatenated values, so zero-diagonal
 * constraint rows are ordered after their neighbours; NULL: pattern only). */
int kkt_batch_analyze(kkt_batch* b, const double* vals);
int kkt_batch_factor(kkt_batch* b, const double* vals);
int kkt_batch_factor_device(kkt_batch* b, const double* d_vals);
int kkt_batch_solve(kkt_batch* b, const double* B, double* X);
int kkt_batch_solve_device(kkt_batch* b, const double* d_B, double* d_X);
int kkt_batch_set_precision(kkt_batch* b, int prec);
int kkt_batch_set_refine_tol(kkt_batch* b, double tol, int max_steps);
int kkt_batch_set_ir_steps(kkt_batch* b, int nsteps);
/* picks the batch's solve shapes per tree level by timing, as kkt_autotune_solve
 * (after a factor; the batch's levels hold the fronts of every system) */
int kkt_batch_autotune_solve(kkt_batch* b, int reps);
/* statistics of the whole batch (as kkt_get_info) and of one system */
int kkt_batch_get_info(kkt_batch* b, kkt_info* out);
int kkt_batch_system_info(kkt_batch* b, int sys, kkt_batch_sysinfo* out);
const char* kkt_batch_last_error(kkt_batch* b);

/* ---- serving loop ----
 * Step-solve requests from many concurrent simulation instances (no
 * counterpart in the reference, where every simulation drives its own solver).
 * Each registered instance gets its own handle (own HIP stream) on its first
 * request; its analysis and factor stay resident in HBM between requests.
 * Admission by an HBM budget: after an instance's analysis, what its handle
 * will hold once factored (kkt_info.factored_bytes) is reserved before the
 * factor allocates, evicting idle resident instances least recently used
 * first; an instance larger than the whole budget fails its requests with
 * KKT_ERR_ALLOC.  Requests enter one FIFO queue served by `nworkers` threads,
 * each taking the oldest request whose instance is idle (continuous batching:
 * arrivals start as soon as a worker and memory are free, instances' device
 * work overlaps on their streams).  An instance's requests run in submission
 * order, one at a time, each a KKT step with the iteration-limit part of
 * KKTSolver.factorAndSolve's policy (KKTSolver.java:930-1014): BiCGStab with the
 * resident factor while it converges within max_iter iterations, else a factor
 * and a refined solve (no timing-based early refactor: a request's time in a
 * shared server includes other instances' work). */
typedef struct kkt_server kkt_server;
typedef struct kkt_serve_config {
  long long budget_bytes; /* device memory the resident instances may hold (default 200 GiB) */
  int nworkers;           /* worker threads, 1..64 (default 12) */
  int precision;          /* KKT_PREC_* of every instance (default FP32 + refinement) */
  int tol_exp;            /* iterative solves to 10^-tol_exp relative residual (default 10) */
  int max_iter;           /* BiCGStab iterations before a refactor (default 30) */
  double refine_tol;      /* refinement target of FP32 factors (default 1e-10) */
  int group_runners;      /* 0..8 threads solving iterative steps in groups (default 2); 0: on the workers */
  int group_max;          /* requests per group solve, 1..KKT_GROUP_MAX (default 64) */
} kkt_serve_config;
typedef struct kkt_serve_result {
  int status;               /* KKT_OK or a negative KKT_ERR_* code */
  int kind;                 /* 1: iterative solve with the resident factor, 2: factor + solve */
  int iterations;           /* BiCGStab iterations (kind 1) */
  double rel_residual;      /* the handle's last refinement residual */
  double submit_ms, start_ms, done_ms; /* since the server's creation */
  char message[160];        /* error text when status != 0 */
} kkt_serve_result;
typedef struct kkt_serve_stats {
  long long submitted, completed, errors, direct, iterative;
  long long admissions, evictions;
  long long queue_depth, queue_max;      /* requests waiting now / at most */
  long long budget_bytes, bytes_resident, bytes_peak;
  int instances_resident;
  double service_ms;                     /* sum of the requests' solve times */
  double elapsed_ms;                     /* since the server's creation */
  long long group_calls, group_requests; /* group solves run / requests they took */
  long long releases;                    /* idle instances' factors released (analysis kept);
                                            evictions counts instances destroyed */
} kkt_serve_stats;
/* The defaults listed above. */
void kkt_serve_default_config(kkt_serve_config* cfg);
/* Starts the worker threads; NULL on a bad configuration. */
kkt_server* kkt_serve_create(const kkt_serve_config* cfg);
/* Finishes every submitted request, stops the workers and frees every handle.
 * Requests not yet waited for are dropped with the server. */
void kkt_serve_destroy(kkt_server* s);
/* An instance's pattern (0-based upper CSR, as kkt_set_pattern) and the values
 * its analysis orders by (zero-diagonal rows after their neighbours); copied. */
int kkt_serve_register(kkt_server* s, int iid, int n, int nnz, const int* rowOffs, const int* colIdxs,
                       const double* vals0, int mtype);
/* Queues one step of instance iid: the step's values (nnz) and right-hand side
 * (n) are copied at submission; x (n doubles) is written when the request
 * completes and must stay valid until kkt_serve_wait returns.  Returns the
 * request id (> 0) or a negative status. */
long long kkt_serve_submit(kkt_server* s, int iid, const double* vals, const double* b, double* x);
/* The same with device arrays (used in place: they must hold their data when
 * submitted and stay valid until kkt_serve_wait returns). */
long long kkt_serve_submit_device(kkt_server* s, int iid, const double* d_vals, const double* d_b, double* d_x);
/* Blocks until the request completes; fills out (may be NULL) and returns its
 * status.  Each request is waited for once. */
int kkt_serve_wait(kkt_server* s, long long req, kkt_serve_result* out);
int kkt_serve_get_stats(kkt_server* s, kkt_serve_stats* out);
const char* kkt_serve_last_error(kkt_server* s);

/* Host-only symbolic analysis (no GPU needed): structure statistics and the
 * same arrays kkt_get_symbolic returns, for pattern checks on the CPU. */
typedef struct kkt_symbolic kkt_symbolic;
kkt_symbolic* kkt_symbolic_create(int n, int nnz, const int* rowOffs, const int* colIdxs,
                                  const unsigned char* delay, int ordering);
int kkt_symbolic_info(kkt_symbolic* s, kkt_info* out);
int kkt_symbolic_get(kkt_symbolic* s, int* perm, int* parent, int* colcount, int* sup_ptr,
                     int* sparent, long long* rows_ptr, int* rows, int* level);
void kkt_symbolic_destroy(kkt_symbolic* s);

#ifdef __cplusplus
}
#endif
#endif
