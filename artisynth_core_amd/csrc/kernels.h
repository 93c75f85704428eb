// Device kernels of the supernodal LDL^T solver (gfx950) and their launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace kkt {

constexpr int NB = 64;    // panel width (pivot columns per panel step)
constexpr int TILE = 64;  // assembly / Schur-update tile edge
constexpr int UPD_GROUP = 4;  // panels whose Schur updates are applied together (K = 256)
// fronts with k <= FWD_SMALL_K pivot columns and m <= FWD_SMALL_M rows run their
// whole forward step (gather + rows) in one workgroup
constexpr int FWD_SMALL_K = 32;
constexpr int FWD_SMALL_M = 256;
// widest supernode the solve kernels take (LDS vectors of KMAX entries); the
// symbolic analysis splits wider ones, and the solve plan checks it
constexpr int KMAX = 2048;
// rows of one forward row block of a large front at most (its row groups share
// one gather of the front's right-hand side; the contribution rows' entries are
// staged in LDS)
constexpr int FWD_GROUP_ROWS = 1024;
// right-hand sides one multi-RHS solve pass carries (vectors stored one after
// another, each with its own leading dimension)
constexpr int SOLVE_NRHS = 8;

// Device copy of the supernodal structure (all arrays device pointers).
struct DevSym {
  const int* sn_first;       // first permuted column
  const int* sn_k;           // pivot columns
  const int* sn_m;           // front size k + r
  const long long* sn_foff;  // front offset (elements), m*m column-major
  const long long* sn_roff;  // offset into rows[] of the r below-diagonal rows
  const long long* sn_woff;  // panel workspace offset (m * NB elements)
  const long long* sn_uoff;  // solve update-vector offset (r elements)
  const int* ch_ptr;         // children of s: ch_list[ch_ptr[s] .. ch_ptr[s+1])
  const int* ch_list;
  const int* rel;            // for child c: parent-front position of each CB row,
                             // stored at rel[sn_roff[c] ...]
  const int* rows;           // global permuted row index of each below-diagonal row
  int* lperm;                // per permuted column: original local index in supernode
  int* ilperm;               // inverse of lperm (local index -> pivot position)
  signed char* piv;          // 1 = 1x1, 2 = first of 2x2, 0 = second of 2x2
  // forward gather map: front position p of s receives u[gsrc[e]] for
  // e in [gptr[sn_goff[s] + p], gptr[sn_goff[s] + p + 1]), children in ascending order
  const long long* sn_goff;
  const int* gptr;
  const int* gsrc;
  // backward vector: supernode s owns m entries at sn voff, [z of its pivots (pivot
  // order); -x of its below-diagonal rows].  Solve position g is a below-diagonal
  // row of the supernodes whose entries are push_pos[push_ptr[g] .. push_ptr[g+1])
  // (ascending); wpos[q] is the entry of pivot position q
  const int* push_ptr;
  const int* push_pos;
  const int* wpos;
  // connected component (assembly tree) of each supernode, 0..nroots-1
  const int* sn_root;
  // offset of each supernode's block in the packed solve arrays
  const long long* sn_poff;
};

// Work item of the solve kernels: one workgroup's share of a supernode, with the
// supernode's structure inlined so a kernel issues its data loads right away
struct alignas(16) SolveItem {
  int s;     // supernode
  int blk;   // block of positions / rows / columns of the supernode
  int lg;    // 2^lg lanes per row or column (rows / columns kernels)
  int kind;  // forward gather: 1 = gather block of a large front, 0 = whole small front
  int k, m, first;
  int pad;   // forward row blocks (kind 2): row groups per block
  long long foff;  // front offset (elements)
  long long uoff;  // update-vector offset
  long long roff;  // offset of the below-diagonal rows in rows[]
  long long goff;  // offset of the front positions in the gather map
  long long voff;  // offset of the supernode's backward vector entries
  long long poff;  // offset of the supernode's block in the packed solve arrays
  long long soff;  // offset of the supernode's update-slot block (nch x m)
  long long sdst;  // where its contribution rows go: its column of the parent's
                   // slot block (the parent's soff + j m, j = its child index)
  int nch;         // number of children
  int pad1, pad2, pad3;
};

// Packed solve arrays (written after each factor from the solve form of the
// fronts, supernodes in level-major order, k(k-1)/2 + k r entries each, so one
// tree level's data is one contiguous range):
//   forward  row i of supernode s: G(i, 0 .. min(i, k)) -- X(i, :) without the
//            unit diagonal for i < k, Z(i - k, :) below -- at poff + prow_off(i, k);
//   backward column c < k: G(c+1 .. m, c) at poff + pcol_off(c, m).
// After both, the forward update slots: supernode s with nch children has an
// nch x m block at soff, child j's update entry for parent position p at
// soff + j m + p (the child writes it, the parent reads the column of p); an
// entry no child row maps to stays zero (the block is zeroed when allocated and
// those entries are never written), so the parent adds all nch entries of a
// position in ascending child order with no index loads.
__host__ __device__ inline long long prow_off(int i, int k) {
  return i < k ? (long long)i * (i - 1) / 2 : (long long)k * (k - 1) / 2 + (long long)(i - k) * k;
}
__host__ __device__ inline long long pcol_off(int c, int m) {
  return (long long)c * (m - 1) - (long long)c * (c - 1) / 2;
}

struct FactorCounters {  // device ints
  int npos, nneg, npert, n2x2;
  int npivoted;  // diagonal blocks that needed the full pivoting loop
  int half_overflow;  // an FP32 solve entry too large for the half-precision copy
};

// ---- group launches: one launch covers the same step of many handles ----
// The workgroups of a group launch are dealt to its segments (one per handle)
// by the prefix counts pre[0..nseg] (pre[0] = 0, pre[nseg] = the grid); a
// segment's blocks run the same kernel body, with the same block indices and
// arguments, as that handle's own launch of the step.
constexpr int SEG_MAX = 256;  // segments (handles) per group launch
enum SegVecOp : int {
  SEG_GATHER = 1,  // tout[j] = a[idx[j]] in the front precision             n idx a tout
  SEG_GATHER_P,    // BiCGStab p = r + beta (p - omega v), gathered to tout    n idx sc a=r b=v d=p tout
  SEG_SCATTER,     // d[idx[j]] = tin[j]                                      n idx tin d
  SEG_DSOLVE,      // tout[wpos[g]] = E0 y + E1 y_partner                     n tin=E0 tin2=E1 piv wpos tin3=y tout
  SEG_EXPAND,      // d[e] = a[idx[e]] (full symmetric copy of new values)    n idx a d
  SEG_FILL,        // d[i] = v                                                n v d
  SEG_COPY,        // d[i] = a[i]                                             n a d
  SEG_BICG_S,      // s = r - alpha v, |s|^2 partials                         n sc a=r b=v d=s part1
  SEG_BICG_XR,     // x, r updates, |r|^2 and (r_hat, r) partials             n sc a=y b=z c=s e=t d=x f=r g=rh part1 part2
  SEG_BICG_XA,     // x += alpha y                                            n sc a=y d=x
  SEG_BICG_INIT,   // the first iteration's scalars                           d=sc
  SEG_COLLECT,     // d[0..15] = a[0..15] (one block)                         a d
};
struct SegVec {
  int op, pad;
  long long n;
  const int* idx;
  const signed char* piv;
  const int* wpos;
  const double* sc;
  const double *a, *b, *c, *e, *g;
  double *d, *f, *part1, *part2;
  const void *tin, *tin2, *tin3;
  void* tout;
  double v;
};
enum SegRedOp : int {
  SEG_RED2 = 1,  // k_reduce2: p1 p2 np out1 out2 sc
  SEG_DOTP,      // k_dot_partial over RED_BLOCKS blocks: n a=p1 b=p2 partial=out1
  SEG_REDF,      // k_reduce_final (sum): partial=p1 np out1
};
struct SegRed {
  int op, np, n, pad;
  const double *p1, *p2;
  double *out1, *out2, *sc;
};
struct SegSpmv {
  int n, mode;
  const int *rowptr, *col;
  const double *val, *x, *b, *a1;
  double *y, *part1, *part2;
};
struct SegFwd {
  DevSym d;
  const SolveItem* items;
  const void* F;
  const void* bp;
  void *fp, *y, *u;
};
struct SegBwd {
  DevSym d;
  const SolveItem* items;
  const void* F;
  void *wv, *xp;
};
int red_blocks();
template <typename T>
void launch_seg_vec(const int* pre, int nseg, const SegVec* a, int nblk, hipStream_t st);
void launch_seg_red(const int* pre, int nseg, const SegRed* a, int nblk, hipStream_t st);
void launch_seg_spmv(const int* pre, int nseg, const SegSpmv* a, int nblk, hipStream_t st);
template <typename T, typename S>
void launch_seg_fwd(const int* pre, int nseg, const SegFwd* a, int nblk, hipStream_t st);
template <typename T, typename S>
void launch_seg_bwd(const int* pre, int nseg, const SegBwd* a, int nblk, hipStream_t st);

template <typename T>
void launch_assemble(const int4* items, int nitems, const DevSym& d, T* F, hipStream_t st);
template <typename T>
void launch_scatter(const int* aidx, const long long* apos, int cnt, const double* vals, T* F,
                    hipStream_t st);
template <typename T>
void launch_panel_diag(const int* items, int nitems, int p, const DevSym& d, T* F, T* E0, T* E1,
                       double pert, const double* amax, FactorCounters* ctr, int* rctr, hipStream_t st);
// W of panel p goes to the supernode's W-buffer slot `slot` (m * NB each)
template <typename T>
void launch_panel_trsm(const int2* items, int nitems, int p, int slot, const DevSym& d, T* F, T* W,
                       hipStream_t st);
// C -= sum of W_p L_p^T over the panels pfirst.. of the items (see kernels.hip)
template <typename T>
void launch_panel_update(const int4* items, int nitems, int pfirst, const DevSym& d, T* F,
                         const T* W, hipStream_t st);

template <typename T>
void launch_panel_xrow(const int2* items, int nitems, int p, const DevSym& d, T* F, hipStream_t st);

// after the factor: Z^T = (L21 X)^T into the upper-right block, then the lower
// part becomes G = [X; Z] (transposed copies of the upper part)
template <typename T>
void launch_zt(const int4* items, int nitems, const DevSym& d, T* F, hipStream_t st);
template <typename T>
void launch_gform(const int4* items, int nitems, const DevSym& d, T* F, hipStream_t st);
// copies the solve form of every front into the packed solve arrays (items:
// (supernode, block of 4 front columns))
template <typename T>
void launch_pack_solve(const int2* items, int nitems, const DevSym& d, const T* F, T* PF, T* PB, hipStream_t st);
// half-precision copy of n FP32 packed entries (ctr->half_overflow set when one
// does not fit)
void launch_pack_half(long long n, const float* in, _Float16* out, FactorCounters* ctr, hipStream_t st);
// level-scheduled solves (vectors in the front precision T) on the packed
// arrays: PF forward rows, PB backward columns, SL the forward update slots;
// the packed entries in S (T, or _Float16 for T = float)
template <typename T, typename S>
void launch_fwd_level(const SolveItem* items, int nitems, const DevSym& d, const S* PF, const T* bp,
                      T* fp, T* y, T* SL, hipStream_t st);
template <typename T>
void launch_dsolve(int n, const T* E0, const T* E1, const signed char* piv, const int* wpos,
                   const T* y, T* wv, hipStream_t st);
template <typename T, typename S>
void launch_bwd_cols(const SolveItem* items, int nitems, const DevSym& d, const S* PB, T* wv,
                     T* xp, hipStream_t st);

// the same for nr <= SOLVE_NRHS right-hand sides at once (vector j of a set at
// j * ld; ldn, ldu, ldw: n-, update- and backward-vector sets)
template <typename T>
void launch_fwd_gather_m(const SolveItem* items, int nitems, const DevSym& d, const T* F, int nr, const T* bp,
                         T* fp, T* y, T* u, long long ldn, long long ldu, hipStream_t st);
template <typename T>
void launch_fwd_rows_m(const SolveItem* items, int nitems, const DevSym& d, const T* F, int nr, const T* fp,
                       T* y, T* u, long long ldn, long long ldu, hipStream_t st);
template <typename T>
void launch_dsolve_m(int n, int nr, const T* E0, const T* E1, const signed char* piv, const int* wpos,
                     const T* y, T* wv, long long ldn, long long ldw, hipStream_t st);
template <typename T>
void launch_bwd_cols_m(const SolveItem* items, int nitems, const DevSym& d, const T* F, int nr, T* wv, T* xp,
                       long long ldn, long long ldw, hipStream_t st);
template <typename TI, typename TO>
void launch_gather_perm_m(int n, int nr, const int* perm, const TI* in, long long ldi, TO* out, long long ldo,
                          hipStream_t st);
template <typename TI, typename TO>
void launch_scatter_perm_m(int n, int nr, const int* perm, const TI* in, long long ldi, TO* out, long long ldo,
                           hipStream_t st);

// permutation / precision conversion: out[j] = (TO) in[perm[j]]  (gather)
template <typename TI, typename TO>
void launch_gather_perm(int n, const int* perm, const TI* in, TO* out, hipStream_t st);
// out[perm[j]] = (TO) in[j]  (scatter)
template <typename TI, typename TO>
void launch_scatter_perm(int n, const int* perm, const TI* in, TO* out, hipStream_t st);
// gather in pivot order with BiCGStab's p = r + beta (p - omega v) fused in (p updated in place)
template <typename TO>
void launch_gather_bicg_p(int n, const int* perm, const double* sc, const double* r, const double* v, double* p,
                          TO* out, hipStream_t st);

// sparse symmetric matrix (full CSR copy) in FP64
void launch_expand_vals(long long nnzf, const int* map, const double* vals, double* valsf,
                        hipStream_t st);
// y = A x (mode 0) or y = b - A x (mode 1); with part1 / part2, per-block partial
// sums of a1 .* y and y .* y (spmv_blocks(n) of each) for launch_reduce2
void launch_spmv(int n, const int* rowptr, const int* col, const double* val, const double* x,
                 const double* b, double* y, int mode, hipStream_t st, const double* a1 = nullptr,
                 double* part1 = nullptr, double* part2 = nullptr);
int spmv_blocks(int n);
int vec_blocks(int n);
// out1 / out2 = the sums of np partials (fixed order); with sc, the next BiCGStab
// iteration's scalars are formed from the BiCGStab slots, out2 being the next rho
void launch_reduce2(const double* p1, const double* p2, int np, double* out1, double* out2, hipStream_t st,
                    double* sc = nullptr);
// Y_j = B_j - A X_j and out[j] = (A_j, B_j) for nr <= SOLVE_NRHS columns (column j at
// j * ld), each column's value equal to launch_spmv (mode 1) / launch_dot's;
// partial holds dot_m_partials() doubles
void launch_spmv_resid_m(int n, int nr, const int* rowptr, const int* col, const double* val, const double* X,
                         const double* B, double* Y, long long ld, hipStream_t st);
int dot_m_partials();
void launch_dot_m(int n, int nr, const double* A, const double* B, long long ld, double* partial, double* out,
                  hipStream_t st);
void launch_absmax(long long nnz, const double* v, double* partial, double* out, hipStream_t st);
// out[slot] = sum a.*b  (deterministic two-stage)
void launch_dot(int n, const double* a, const double* b, double* partial, double* out,
                hipStream_t st);
// generic fused vector op: y = alpha*x + beta*y (+ gamma*z), scalars from device when ptrs given
void launch_axpby(int n, double alpha, const double* x, double beta, double* y, hipStream_t st);
void launch_axpbypcz(int n, double a, const double* x, double b, const double* y, double c,
                     double* z, hipStream_t st);


// BiCGStab updates with device-resident scalars (see kernels.hip for the slots)
// device-side value assembly: contributions into scratch c, then the ordered
// per-slot scatter vals[uslot[u]] += scale * sum_{e in segment u} c[order[e]]
void launch_contrib_block3(int nblocks, const double* bv, const double* bs, double* c, hipStream_t st);
void launch_contrib_diag3(int nblocks, const double* m, double* c, hipStream_t st);
void launch_contrib_matstiff3(int nblocks, const double* gi, const double* gj, const double* D,
                              const double* sg, const double* dv, double* c, hipStream_t st);
// element-batched forms: contributions at c[36 p + r * coldim + q] of node pair p
// (c[9 p + ...] for the linear elastic form without node reduction)
void launch_contrib_matstiff3_elem(int nelems, const int* nodeCnt, const int* nodeOff, const int* pairOff,
                                   const int* ipOff, const int* gradOff, const int* pairNodes,
                                   const int* nodeDims, const double* T, const double* grads,
                                   const double* Ds, const double* sigmas, const double* dvs, double* c,
                                   hipStream_t st);
void launch_contrib_linelastic3_elem(int nelems, const int* nodeCnt, const int* pairOff, const int* ipOff,
                                     const int* gradOff, const int* pairNodes, const double* params,
                                     const double* grads, const double* dvs, double* c, hipStream_t st);
void launch_contrib_linelastic3_geom_elem(int nelems, const int* nodeCnt, const int* nodeOff,
                                          const int* pairOff, const int* ipOff, const int* ngOff,
                                          const int* pairNodes, const int* nodeDims, const double* T,
                                          const double* params, const double* pos, const double* ngrads,
                                          const double* w, double* c, hipStream_t st);
void launch_contrib_dilational3_elem(int nelems, const int* nodeOff, const int* presCnt, const int* pairOff,
                                     const int* consOff, const int* rinvOff, const int* pairNodes,
                                     const int* nodeDims, const double* T, const double* cons,
                                     const double* rinv, double* c, hipStream_t st);
void launch_seg_scatter(int nuniq, const int* useg, const int* uslot, const int* order,
                        const double* c, double scale, double* vals, hipStream_t st);

// s = r - alpha v, with per-block partials of |s|^2 (vec_blocks(n))
void launch_bicg_s(int n, const double* sc, const double* r, const double* v, double* s_, double* part,
                   hipStream_t st);
// x, r updates, with per-block partials of |r|^2 and (r_hat, r)
void launch_bicg_xr(int n, const double* sc, const double* y, const double* z, const double* s_,
                    const double* t, double* x, double* r, const double* rh, double* part1, double* part2,
                    hipStream_t st);
void launch_bicg_xa(int n, const double* sc, const double* y, double* x, hipStream_t st);
void launch_bicg_init(double* sc, hipStream_t st);
void launch_fill(long long n, double v, double* x, hipStream_t st);
// zeroes the totals and the nr per-component counters (four per component:
// positive, negative, perturbed pivots, 2x2 pivots)
void launch_zero_counters(FactorCounters* c, int* rctr, int nr, hipStream_t st);
// out[g] = sum of v_i^2 over the ranges of segment g: chunks [ca[q], cb[q]) for
// q in [cptr[g], cptr[g+1]), summed in a fixed order; partial holds nchunks doubles
void launch_seg_sumsq(int nchunks, const long long* ca, const long long* cb, const double* v, double* partial,
                      int nseg, const int* cptr, double* out, hipStream_t st);
// x_i += d_i where active[row_seg[i]] != 0
void launch_axpy_masked(long long n, const int* row_seg, const unsigned char* active, const double* d, double* x,
                        hipStream_t st);
// amax[r] = max |v| over the value ranges of component r: chunks [ca[q], cb[q])
// for q in [cptr[r], cptr[r+1]); partial holds nchunks doubles
void launch_absmax_roots(int nchunks, const long long* ca, const long long* cb, const double* v,
                         double* partial, int nroots, const int* cptr, double* amax, hipStream_t st);

}  // namespace kkt
