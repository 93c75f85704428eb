/* CPU oracle for the sparse KKT direct-solve path -- TEST INFRASTRUCTURE.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * as the checker.  See kkt_oracle.h for the reference citations per function.
 *
 * Parity pins (tests/test_oracle_golden.py): PardisoSolverTest.java:407-543
 * (5x5 indefinite, inertia 2-/3+, 0 perturbed pivots, three value sets),
 * CuDssSolverTest.java:55-107,343-349,870-947 (SPD5, symmetric indefinite 3x3,
 * 5x5 KKT), PardisoSolverTest.java:608-688 (testMatrix.mat residual and hybrid
 * solve), dense LAPACK solves and Sylvester inertia on seeded systems. */
#include "kkt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

void orc_spmv_upper(int n, const int* rowOffs, const int* colIdxs, const double* vals,
                    const double* x, double* y) {
  for (int i = 0; i < n; i++) y[i] = 0.0;
  for (int i = 0; i < n; i++) {
    double acc = 0.0;
    for (int p = rowOffs[i]; p < rowOffs[i + 1]; p++) {
      int j = colIdxs[p];
      if (j < i) continue;
      double v = vals[p];
      acc += v * x[j];
      if (j != i) y[j] += v * x[i];
    }
    y[i] += acc;
  }
}

/* ---------------------------------------------------------------------------
 * Symbolic elimination: struct(L(:,j)) = {i>j: A(i,j)!=0} U children's structs
 * minus j.  Column structures kept as sorted int arrays.
 * ------------------------------------------------------------------------- */
static int cmp_int(const void* a, const void* b) {
  int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

/* full symmetric adjacency of the permuted pattern (lower part: rows > col) */
typedef struct {
  int n;
  long* ptr;
  int* idx;
} lower_pat;

static int build_lower(int n, const int* rowOffs, const int* colIdxs, const int* perm,
                       lower_pat* L) {
  int* iperm = (int*)malloc(sizeof(int) * n);
  if (!iperm) return -1;
  for (int k = 0; k < n; k++) iperm[perm[k]] = k;
  long* cnt = (long*)calloc(n + 1, sizeof(long));
  for (int i = 0; i < n; i++)
    for (int p = rowOffs[i]; p < rowOffs[i + 1]; p++) {
      int j = colIdxs[p];
      if (j <= i) continue;
      int a = iperm[i], b = iperm[j];
      int c = a < b ? a : b;
      cnt[c + 1]++;
    }
  for (int i = 0; i < n; i++) cnt[i + 1] += cnt[i];
  int* idx = (int*)malloc(sizeof(int) * (cnt[n] + 1));
  long* pos = (long*)malloc(sizeof(long) * n);
  for (int i = 0; i < n; i++) pos[i] = cnt[i];
  for (int i = 0; i < n; i++)
    for (int p = rowOffs[i]; p < rowOffs[i + 1]; p++) {
      int j = colIdxs[p];
      if (j <= i) continue;
      int a = iperm[i], b = iperm[j];
      int c = a < b ? a : b, r = a < b ? b : a;
      idx[pos[c]++] = r;
    }
  for (int c = 0; c < n; c++) qsort(idx + cnt[c], cnt[c + 1] - cnt[c], sizeof(int), cmp_int);
  free(pos);
  free(iperm);
  L->n = n;
  L->ptr = cnt;
  L->idx = idx;
  return 0;
}

int orc_symbolic(int n, const int* rowOffs, const int* colIdxs, const int* perm, int* parent,
                 int* colcount) {
  lower_pat A;
  if (build_lower(n, rowOffs, colIdxs, perm, &A)) return -1;
  int** cs = (int**)calloc(n, sizeof(int*));
  int* cl = (int*)calloc(n, sizeof(int));
  int* mark = (int*)malloc(sizeof(int) * n);
  int* buf = (int*)malloc(sizeof(int) * n);
  /* children lists built incrementally */
  int* head = (int*)malloc(sizeof(int) * n);
  int* next = (int*)malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++) { mark[i] = -1; head[i] = -1; next[i] = -1; }
  for (int j = 0; j < n; j++) {
    int m = 0;
    for (long p = A.ptr[j]; p < A.ptr[j + 1]; p++) {
      int r = A.idx[p];
      if (mark[r] != j) { mark[r] = j; buf[m++] = r; }
    }
    for (int c = head[j]; c >= 0; c = next[c]) {
      for (int q = 0; q < cl[c]; q++) {
        int r = cs[c][q];
        if (r > j && mark[r] != j) { mark[r] = j; buf[m++] = r; }
      }
      free(cs[c]);
      cs[c] = NULL;
    }
    qsort(buf, m, sizeof(int), cmp_int);
    cs[j] = (int*)malloc(sizeof(int) * (m + 1));
    memcpy(cs[j], buf, sizeof(int) * m);
    cl[j] = m;
    colcount[j] = m + 1;
    parent[j] = m > 0 ? buf[0] : -1;
    if (m > 0) { next[j] = head[buf[0]]; head[buf[0]] = j; }
  }
  for (int j = 0; j < n; j++) free(cs[j]);
  free(cs); free(cl); free(mark); free(buf); free(head); free(next);
  free(A.ptr); free(A.idx);
  return 0;
}

long orc_supernode_rows(int n, const int* rowOffs, const int* colIdxs, const int* perm,
                        int nsuper, const int* sup_ptr, long* rows_ptr, int* rows, long rows_cap) {
  lower_pat A;
  if (build_lower(n, rowOffs, colIdxs, perm, &A)) return -1;
  int* c2s = (int*)malloc(sizeof(int) * n);
  for (int s = 0; s < nsuper; s++)
    for (int j = sup_ptr[s]; j < sup_ptr[s + 1]; j++) c2s[j] = s;
  int* mark = (int*)malloc(sizeof(int) * n);
  int* buf = (int*)malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++) mark[i] = -1;
  /* children: a supernode's parent is the supernode holding its first row */
  int* head = (int*)malloc(sizeof(int) * nsuper);
  int* next = (int*)malloc(sizeof(int) * nsuper);
  for (int s = 0; s < nsuper; s++) { head[s] = -1; next[s] = -1; }
  long tot = 0;
  rows_ptr[0] = 0;
  for (int s = 0; s < nsuper; s++) {
    int f = sup_ptr[s], l = sup_ptr[s + 1] - 1, m = 0;
    for (int c = f; c <= l; c++)
      for (long p = A.ptr[c]; p < A.ptr[c + 1]; p++) {
        int r = A.idx[p];
        if (r > l && mark[r] != s) { mark[r] = s; buf[m++] = r; }
      }
    for (int ch = head[s]; ch >= 0; ch = next[ch])
      for (long q = rows_ptr[ch]; q < rows_ptr[ch + 1]; q++) {
        int r = rows[q];
        if (r > l && mark[r] != s) { mark[r] = s; buf[m++] = r; }
      }
    qsort(buf, m, sizeof(int), cmp_int);
    if (tot + m > rows_cap) { tot = -2; break; }
    memcpy(rows + tot, buf, sizeof(int) * m);
    tot += m;
    rows_ptr[s + 1] = tot;
    if (m > 0) {
      int p = c2s[buf[0]];
      next[s] = head[p];
      head[p] = s;
    }
  }
  free(c2s); free(mark); free(buf); free(head); free(next);
  free(A.ptr); free(A.idx);
  return tot;
}

/* ---------------------------------------------------------------------------
 * Multifrontal numeric factorization.
 * ------------------------------------------------------------------------- */
struct orc_factor {
  int n, ns, nb;
  int* perm;       /* perm[new] = old */
  int* sup_ptr;    /* ns+1 */
  long* rows_ptr;  /* ns+1 */
  int* rows;       /* row structures */
  int* sparent;
  long* foff;      /* front offsets (m*m each) */
  double* F;       /* fronts, column-major m x m, lower triangle used */
  int* lperm;      /* per column: local pivot permutation (n entries, by supernode) */
  int* piv;        /* per column: 1 = 1x1, 2 = first of 2x2, 0 = second of 2x2 */
  long npos, nneg, npert, n2x2, nnzL;
};

static inline double* FE(double* F, int m, int i, int j) { return F + (long)j * m + i; }

/* Bunch-Kaufman on the w x w diagonal block at (j0,j0) of an m x m front,
 * restricted to the block.  Applies swaps to: the block, rows of columns
 * [0,j0), the panel columns below the block.  Records local permutation. */
static void bk_panel(double* F, int m, int j0, int w, double tau, int* lp, int* piv,
                     long* npos, long* nneg, long* npert, long* n2x2) {
  const double alpha = (1.0 + sqrt(17.0)) / 8.0;
  int k = 0;
  while (k < w) {
    int kstep = 1, kp = k;
    int gk = j0 + k;
    double absakk = fabs(*FE(F, m, gk, gk));
    int imax = -1;
    double colmax = 0.0;
    for (int i = k + 1; i < w; i++) {
      double v = fabs(*FE(F, m, j0 + i, gk));
      if (v > colmax) { colmax = v; imax = i; }
    }
    int perturb = 0;
    if ((absakk > colmax ? absakk : colmax) < tau) {
      perturb = 1;
    } else if (absakk >= alpha * colmax) {
      kp = k;
    } else {
      double rowmax = 0.0;
      for (int j = k; j < imax; j++) {
        double v = fabs(*FE(F, m, j0 + imax, j0 + j));
        if (v > rowmax) rowmax = v;
      }
      for (int i = imax + 1; i < w; i++) {
        double v = fabs(*FE(F, m, j0 + i, j0 + imax));
        if (v > rowmax) rowmax = v;
      }
      if (absakk * rowmax >= alpha * colmax * colmax) kp = k;
      else if (fabs(*FE(F, m, j0 + imax, j0 + imax)) >= alpha * rowmax) kp = imax;
      else { kp = imax; kstep = 2; }
    }
    int kk = k + kstep - 1;
    if (kp != kk) {
      int a = j0 + kk, b = j0 + kp; /* a < b */
      double t;
      /* diagonal */
      t = *FE(F, m, a, a); *FE(F, m, a, a) = *FE(F, m, b, b); *FE(F, m, b, b) = t;
      /* between */
      for (int j = a + 1; j < b; j++) {
        t = *FE(F, m, j, a); *FE(F, m, j, a) = *FE(F, m, b, j); *FE(F, m, b, j) = t;
      }
      /* below b within the block and below the block (whole remaining front rows) */
      for (int i = b + 1; i < m; i++) {
        t = *FE(F, m, i, a); *FE(F, m, i, a) = *FE(F, m, i, b); *FE(F, m, i, b) = t;
      }
      /* columns left of a: rows a and b (includes earlier panels and [j0, gk)) */
      for (int j = 0; j < a; j++) {
        t = *FE(F, m, a, j); *FE(F, m, a, j) = *FE(F, m, b, j); *FE(F, m, b, j) = t;
      }
      int ti = lp[a]; lp[a] = lp[b]; lp[b] = ti;
    }
    if (kstep == 1) {
      double d = *FE(F, m, gk, gk);
      if (perturb || fabs(d) < tau) {
        d = (d >= 0.0) ? tau : -tau;
        *FE(F, m, gk, gk) = d;
        (*npert)++;
      }
      if (d > 0) (*npos)++; else (*nneg)++;
      piv[gk] = 1;
      double dinv = 1.0 / d;
      /* rank-1 update of the remaining block (cols gk+1 .. j0+w-1); rows below
       * the block are handled by the panel TRSM/update afterwards */
      for (int j = k + 1; j < w; j++) {
        double lj = *FE(F, m, j0 + j, gk) * dinv;
        for (int i = j; i < w; i++)
          *FE(F, m, j0 + i, j0 + j) -= *FE(F, m, j0 + i, gk) * lj;
      }
      for (int i = k + 1; i < w; i++) *FE(F, m, j0 + i, gk) *= dinv;
    } else {
      int g1 = gk, g2 = gk + 1;
      double d11 = *FE(F, m, g1, g1), d21 = *FE(F, m, g2, g1), d22 = *FE(F, m, g2, g2);
      double det = d11 * d22 - d21 * d21;
      double i11 = d22 / det, i21 = -d21 / det, i22 = d11 / det;
      if (det < 0) { (*npos)++; (*nneg)++; }
      else if (d11 + d22 > 0) (*npos) += 2;
      else (*nneg) += 2;
      (*n2x2)++;
      piv[g1] = 2;
      piv[g2] = 0;
      for (int j = k + 2; j < w; j++) {
        double a1 = *FE(F, m, j0 + j, g1), a2 = *FE(F, m, j0 + j, g2);
        double l1 = a1 * i11 + a2 * i21, l2 = a1 * i21 + a2 * i22;
        for (int i = j; i < w; i++)
          *FE(F, m, j0 + i, j0 + j) -= *FE(F, m, j0 + i, g1) * l1 + *FE(F, m, j0 + i, g2) * l2;
      }
      for (int i = k + 2; i < w; i++) {
        double a1 = *FE(F, m, j0 + i, g1), a2 = *FE(F, m, j0 + i, g2);
        *FE(F, m, j0 + i, g1) = a1 * i11 + a2 * i21;
        *FE(F, m, j0 + i, g2) = a1 * i21 + a2 * i22;
      }
    }
    k += kstep;
  }
}

/* Partial factorization of the first k columns of an m x m front. */
static void front_factor(double* F, int m, int k, int nb, double tau, int* lp, int* piv,
                         long* npos, long* nneg, long* npert, long* n2x2, double* W) {
  for (int j0 = 0; j0 < k; j0 += nb) {
    int w = (k - j0 < nb) ? (k - j0) : nb;
    bk_panel(F, m, j0, w, tau, lp, piv, npos, nneg, npert, n2x2);
    int r0 = j0 + w, mr = m - r0;
    if (mr <= 0) continue;
    /* TRSM: rows r0..m-1 of the panel: x L11^T = a ; W = x ; L = x D^-1 */
    for (int i = r0; i < m; i++) {
      for (int t = 0; t < w; t++) {
        double x = *FE(F, m, i, j0 + t);
        for (int l = 0; l < t; l++) {
          if (piv[j0 + l] == 2 && l + 1 == t) continue; /* D off-diagonal slot */
          x -= W[(long)l * mr + (i - r0)] * *FE(F, m, j0 + t, j0 + l);
        }
        W[(long)t * mr + (i - r0)] = x;
      }
      for (int t = 0; t < w;) {
        int g = j0 + t;
        if (piv[g] == 1) {
          *FE(F, m, i, g) = W[(long)t * mr + (i - r0)] / *FE(F, m, g, g);
          t += 1;
        } else {
          double d11 = *FE(F, m, g, g), d21 = *FE(F, m, g + 1, g), d22 = *FE(F, m, g + 1, g + 1);
          double det = d11 * d22 - d21 * d21;
          double a1 = W[(long)t * mr + (i - r0)], a2 = W[(long)(t + 1) * mr + (i - r0)];
          *FE(F, m, i, g) = (a1 * d22 - a2 * d21) / det;
          *FE(F, m, i, g + 1) = (a2 * d11 - a1 * d21) / det;
          t += 2;
        }
      }
    }
    /* trailing update: A(i,j) -= sum_t W(i,t) L(j,t), i >= j >= r0 */
    for (int j = r0; j < m; j++) {
      double* col = FE(F, m, 0, j);
      for (int t = 0; t < w; t++) {
        double lj = *FE(F, m, j, j0 + t);
        const double* wc = W + (long)t * mr - r0;
        for (int i = j; i < m; i++) col[i] -= wc[i] * lj;
      }
    }
  }
}

static int find_row(const int* rows, long lo, long hi, int g) {
  long a = lo, b = hi;
  while (a < b) {
    long mid = (a + b) / 2;
    if (rows[mid] < g) a = mid + 1; else b = mid;
  }
  return (a < hi && rows[a] == g) ? (int)(a - lo) : -1;
}

orc_factor* orc_factor_create(int n, const int* rowOffs, const int* colIdxs, const double* vals,
                              const int* perm, int nsuper, const int* sup_ptr, int nb,
                              double pert) {
  orc_factor* f = (orc_factor*)calloc(1, sizeof(orc_factor));
  if (!f) return NULL;
  f->n = n; f->ns = nsuper; f->nb = nb > 0 ? nb : 64;
  f->perm = (int*)malloc(sizeof(int) * n);
  memcpy(f->perm, perm, sizeof(int) * n);
  f->sup_ptr = (int*)malloc(sizeof(int) * (nsuper + 1));
  memcpy(f->sup_ptr, sup_ptr, sizeof(int) * (nsuper + 1));
  /* row structures: upper bound on total rows = nnz(L) bounded by n*n; grow */
  long cap = 16L * n + 1024;
  f->rows_ptr = (long*)malloc(sizeof(long) * (nsuper + 1));
  for (;;) {
    f->rows = (int*)malloc(sizeof(int) * cap);
    long t = orc_supernode_rows(n, rowOffs, colIdxs, perm, nsuper, sup_ptr, f->rows_ptr, f->rows, cap);
    if (t >= 0) break;
    free(f->rows);
    if (t != -2) { orc_factor_free(f); return NULL; }
    cap *= 4;
  }
  int* iperm = (int*)malloc(sizeof(int) * n);
  for (int k = 0; k < n; k++) iperm[perm[k]] = k;
  int* c2s = (int*)malloc(sizeof(int) * n);
  for (int s = 0; s < nsuper; s++)
    for (int j = sup_ptr[s]; j < sup_ptr[s + 1]; j++) c2s[j] = s;
  f->sparent = (int*)malloc(sizeof(int) * nsuper);
  f->foff = (long*)malloc(sizeof(long) * (nsuper + 1));
  f->foff[0] = 0;
  int maxm = 0;
  for (int s = 0; s < nsuper; s++) {
    long r = f->rows_ptr[s + 1] - f->rows_ptr[s];
    int k = sup_ptr[s + 1] - sup_ptr[s];
    int m = k + (int)r;
    if (m > maxm) maxm = m;
    f->foff[s + 1] = f->foff[s] + (long)m * m;
    f->sparent[s] = r > 0 ? c2s[f->rows[f->rows_ptr[s]]] : -1;
    f->nnzL += (long)k * (k + 1) / 2 + (long)k * r;
  }
  f->F = (double*)calloc(f->foff[nsuper] > 0 ? f->foff[nsuper] : 1, sizeof(double));
  f->lperm = (int*)malloc(sizeof(int) * n);
  f->piv = (int*)malloc(sizeof(int) * n);
  if (!f->F) { free(iperm); free(c2s); orc_factor_free(f); return NULL; }
  /* perturbation threshold pert * max|a_ij| per connected component: the
   * assembly tree of the supernode (PARDISO's iparm[9] threshold, pardisoMkl.cc:
   * 89-110, taken per independent block so that a block-diagonal batch factors
   * each system as it would alone; one component = the whole-matrix maximum) */
  int* sroot = (int*)malloc(sizeof(int) * (nsuper > 0 ? nsuper : 1));
  for (int s = nsuper - 1; s >= 0; s--) sroot[s] = f->sparent[s] < 0 ? s : sroot[f->sparent[s]];
  double* amax = (double*)calloc(nsuper > 0 ? nsuper : 1, sizeof(double));
  for (int i = 0; i < n; i++)
    for (int p = rowOffs[i]; p < rowOffs[i + 1]; p++) {
      int j = colIdxs[p];
      if (j < i) continue;
      int c = iperm[i] < iperm[j] ? iperm[i] : iperm[j];
      int r = sroot[c2s[c]];
      if (fabs(vals[p]) > amax[r]) amax[r] = fabs(vals[p]);
    }
  /* scatter A into fronts */
  for (int i = 0; i < n; i++)
    for (int p = rowOffs[i]; p < rowOffs[i + 1]; p++) {
      int j = colIdxs[p];
      if (j < i) continue;
      int a = iperm[i], b = iperm[j];
      int c = a < b ? a : b, r = a < b ? b : a;
      int s = c2s[c];
      int fs = sup_ptr[s], k = sup_ptr[s + 1] - fs;
      int m = k + (int)(f->rows_ptr[s + 1] - f->rows_ptr[s]);
      int rp = (r < fs + k) ? r - fs : k + find_row(f->rows, f->rows_ptr[s], f->rows_ptr[s + 1], r);
      f->F[f->foff[s] + (long)(c - fs) * m + rp] += vals[p];
    }
  double* W = (double*)malloc(sizeof(double) * (size_t)maxm * f->nb + 1);
  int* rel = (int*)malloc(sizeof(int) * (maxm + 1));
  for (int s = 0; s < nsuper; s++) {
    int fs = sup_ptr[s], k = sup_ptr[s + 1] - fs;
    int m = k + (int)(f->rows_ptr[s + 1] - f->rows_ptr[s]);
    double* Fs = f->F + f->foff[s];
    for (int t = 0; t < k; t++) f->lperm[fs + t] = t;
    front_factor(Fs, m, k, f->nb, pert * amax[sroot[s]], f->lperm + fs, f->piv + fs, &f->npos, &f->nneg,
                 &f->npert, &f->n2x2, W);
    /* extend-add CB into parent */
    int p = f->sparent[s];
    if (p >= 0 && m > k) {
      int fp = sup_ptr[p], kp = sup_ptr[p + 1] - fp;
      int mp = kp + (int)(f->rows_ptr[p + 1] - f->rows_ptr[p]);
      double* Fp = f->F + f->foff[p];
      for (int i = 0; i < m - k; i++) {
        int g = f->rows[f->rows_ptr[s] + i];
        rel[i] = (g < fp + kp) ? g - fp : kp + find_row(f->rows, f->rows_ptr[p], f->rows_ptr[p + 1], g);
      }
      for (int j = 0; j < m - k; j++)
        for (int i = j; i < m - k; i++)
          Fp[(long)rel[j] * mp + rel[i]] += Fs[(long)(k + j) * m + (k + i)];
    }
  }
  free(W); free(rel); free(iperm); free(c2s); free(sroot); free(amax);
  return f;
}

void orc_factor_free(orc_factor* f) {
  if (!f) return;
  free(f->perm); free(f->sup_ptr); free(f->rows_ptr); free(f->rows); free(f->sparent);
  free(f->foff); free(f->F); free(f->lperm); free(f->piv);
  free(f);
}

void orc_factor_info(const orc_factor* f, long* info) {
  info[0] = f->npos; info[1] = f->nneg; info[2] = f->npert; info[3] = f->n2x2; info[4] = f->nnzL;
}

void orc_factor_solve(const orc_factor* f, const double* b, double* x) {
  int n = f->n, ns = f->ns;
  double* bp = (double*)malloc(sizeof(double) * n);
  double* z = (double*)malloc(sizeof(double) * n);
  double* xp = (double*)malloc(sizeof(double) * n);
  for (int k = 0; k < n; k++) bp[k] = b[f->perm[k]];
  /* update vectors per supernode */
  double** u = (double**)calloc(ns, sizeof(double*));
  int maxm = 0;
  for (int s = 0; s < ns; s++) {
    int k = f->sup_ptr[s + 1] - f->sup_ptr[s];
    int m = k + (int)(f->rows_ptr[s + 1] - f->rows_ptr[s]);
    if (m > maxm) maxm = m;
  }
  double* g = (double*)malloc(sizeof(double) * (maxm + 1));
  double* fv = (double*)malloc(sizeof(double) * (maxm + 1));
  /* children lists */
  int* head = (int*)malloc(sizeof(int) * ns);
  int* next = (int*)malloc(sizeof(int) * ns);
  for (int s = 0; s < ns; s++) head[s] = -1;
  for (int s = ns - 1; s >= 0; s--) {
    int p = f->sparent[s];
    if (p >= 0) { next[s] = head[p]; head[p] = s; }
  }
  for (int s = 0; s < ns; s++) {
    int fs = f->sup_ptr[s], k = f->sup_ptr[s + 1] - fs;
    int r = (int)(f->rows_ptr[s + 1] - f->rows_ptr[s]), m = k + r;
    const double* F = f->F + f->foff[s];
    for (int t = 0; t < m; t++) fv[t] = (t < k) ? bp[fs + t] : 0.0;
    for (int c = head[s]; c >= 0; c = next[c]) {
      int kc = f->sup_ptr[c + 1] - f->sup_ptr[c];
      int rc = (int)(f->rows_ptr[c + 1] - f->rows_ptr[c]);
      (void)kc;
      for (int i = 0; i < rc; i++) {
        int gr = f->rows[f->rows_ptr[c] + i];
        int pos = (gr < fs + k) ? gr - fs : k + find_row(f->rows, f->rows_ptr[s], f->rows_ptr[s + 1], gr);
        fv[pos] += u[c][i];
      }
      free(u[c]);
      u[c] = NULL;
    }
    for (int t = 0; t < k; t++) g[t] = fv[f->lperm[fs + t]];
    for (int t = k; t < m; t++) g[t] = fv[t];
    for (int t = 0; t < k; t++) {
      double gt = g[t];
      int i0 = t + 1;
      if (f->piv[fs + t] == 2) i0 = t + 2;
      for (int i = i0; i < m; i++) g[i] -= F[(long)t * m + i] * gt;
    }
    /* z = D^-1 y */
    for (int t = 0; t < k;) {
      if (f->piv[fs + t] == 1) { z[fs + t] = g[t] / F[(long)t * m + t]; t++; }
      else {
        double d11 = F[(long)t * m + t], d21 = F[(long)t * m + t + 1], d22 = F[(long)(t + 1) * m + t + 1];
        double det = d11 * d22 - d21 * d21;
        z[fs + t] = (g[t] * d22 - g[t + 1] * d21) / det;
        z[fs + t + 1] = (g[t + 1] * d11 - g[t] * d21) / det;
        t += 2;
      }
    }
    if (r > 0) {
      u[s] = (double*)malloc(sizeof(double) * r);
      for (int i = 0; i < r; i++) u[s][i] = g[k + i];
    }
  }
  for (int s = ns - 1; s >= 0; s--) {
    int fs = f->sup_ptr[s], k = f->sup_ptr[s + 1] - fs;
    int r = (int)(f->rows_ptr[s + 1] - f->rows_ptr[s]), m = k + r;
    const double* F = f->F + f->foff[s];
    for (int t = 0; t < k; t++) {
      double v = z[fs + t];
      for (int i = 0; i < r; i++) v -= F[(long)t * m + k + i] * xp[f->rows[f->rows_ptr[s] + i]];
      g[t] = v;
    }
    for (int t = k - 1; t >= 0; t--) {
      double v = g[t];
      int i0 = t + 1;
      if (f->piv[fs + t] == 2) i0 = t + 2;
      for (int i = i0; i < k; i++) v -= F[(long)t * m + i] * g[i];
      g[t] = v;
    }
    for (int t = 0; t < k; t++) xp[fs + f->lperm[fs + t]] = g[t];
  }
  for (int k = 0; k < n; k++) x[f->perm[k]] = xp[k];
  for (int s = 0; s < ns; s++) free(u[s]);
  free(u); free(g); free(fv); free(head); free(next); free(bp); free(z); free(xp);
}

/* ---------------------------------------------------------------------------
 * BiCGStab (cudssBridge.cc:1705-2006) and refinement.
 * ------------------------------------------------------------------------- */
static double dotv(int n, const double* a, const double* b) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

int orc_bicgstab(const orc_factor* f, int n, const int* rowOffs, const int* colIdxs,
                 const double* vals, const double* b, double* x, double tolRel, int maxIter) {
  double* r = (double*)malloc(sizeof(double) * n * 8);
  double *rh = r + n, *p = r + 2 * n, *v = r + 3 * n, *s = r + 4 * n, *t = r + 5 * n,
         *y = r + 6 * n, *zz = r + 7 * n;
  orc_factor_solve(f, b, x);
  orc_spmv_upper(n, rowOffs, colIdxs, vals, x, r);
  for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
  double bn = sqrt(dotv(n, b, b)), rn = sqrt(dotv(n, r, r));
  if (bn == 0.0) bn = 1.0;
  int ret;
  if (rn / bn <= tolRel) { free(r); return 1; }
  memcpy(rh, r, sizeof(double) * n);
  double rho_prev = 1.0, alpha = 1.0, omega = 1.0;
  const double eps = 1e-30;
  memset(p, 0, sizeof(double) * n);
  memset(v, 0, sizeof(double) * n);
  int it;
  ret = 0;
  for (it = 1; it <= maxIter; it++) {
    double rho = dotv(n, rh, r);
    if (fabs(rho) < eps) { ret = -12; break; }
    double beta = (rho / rho_prev) * (alpha / omega);
    for (int i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    orc_factor_solve(f, p, y);
    orc_spmv_upper(n, rowOffs, colIdxs, vals, y, v);
    double rhv = dotv(n, rh, v);
    if (fabs(rhv) < eps) { ret = -12; break; }
    alpha = rho / rhv;
    for (int i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
    double sn = sqrt(dotv(n, s, s));
    if (sn / bn <= tolRel) {
      for (int i = 0; i < n; i++) x[i] += alpha * y[i];
      ret = it;
      break;
    }
    orc_factor_solve(f, s, zz);
    orc_spmv_upper(n, rowOffs, colIdxs, vals, zz, t);
    double ts = dotv(n, t, s), tt = dotv(n, t, t);
    if (tt < eps) { ret = -12; break; }
    omega = ts / tt;
    for (int i = 0; i < n; i++) x[i] += alpha * y[i] + omega * zz[i];
    for (int i = 0; i < n; i++) r[i] = s[i] - omega * t[i];
    rn = sqrt(dotv(n, r, r));
    if (rn / bn <= tolRel) { ret = it; break; }
    if (fabs(omega) < eps) { ret = -12; break; }
    rho_prev = rho;
  }
  if (it > maxIter) ret = -13;
  free(r);
  return ret;
}

int orc_refine(const orc_factor* f, int n, const int* rowOffs, const int* colIdxs,
               const double* vals, const double* b, double* x, double tol, int maxSteps) {
  double* r = (double*)malloc(sizeof(double) * n * 2);
  double* d = r + n;
  double bn = sqrt(dotv(n, b, b));
  if (bn == 0.0) bn = 1.0;
  int steps = 0;
  for (;;) {
    orc_spmv_upper(n, rowOffs, colIdxs, vals, x, r);
    for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
    if (sqrt(dotv(n, r, r)) / bn <= tol) break;
    if (steps >= maxSteps) { steps = -1; break; }
    orc_factor_solve(f, r, d);
    for (int i = 0; i < n; i++) x[i] += d[i];
    steps++;
  }
  free(r);
  return steps;
}
