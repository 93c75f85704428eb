// gfx950 kernels of the supernodal multifrontal LDL^T solver.
//
// Fronts: one dense m x m column-major array per supernode (lower triangle
// used), resident in HBM.  The numeric factorization runs level by level over
// the assembly tree; inside a level, all fronts advance together one 64-column
// panel step at a time:
//   assemble   : zero + extend-add of the children's contribution blocks
//                (tile-owned gather, deterministic) + scatter of A's entries
//   panel_diag : Bunch-Kaufman LDL^T of the 64x64 diagonal block (FP64 in LDS)
//   panel_trsm : L21 = A21 L11^-T D^-1 for the rows below, W = L21 D
//   panel_update: C -= W L^T on 64x64 tiles with MFMA (f64 or f32)
// After the last level, post-passes form X = L11^-1 (xrow), Z^T = (L21 X)^T
// (zt, MFMA) and the lower copy G = [X; Z] (gform).  Solves then run one launch
// per level and sweep (forward bottom-up, backward top-down) with the update
// vectors of the multifrontal scheme.  No floating-point atomics are used
// anywhere: results are bitwise reproducible run to run.
#include <algorithm>
#include "kernels.h"
#include "wave.h"

namespace kkt {

__device__ inline int lower_bound_i(const int* a, int n, int key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// ------------------------------------------------------------------------
// MFMA 16x16x4 wrappers (f64 / f32).  Lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; the 16x16 result sits in lane l, element v, at column l%16 and
// row Mfma<T>::row(l/16, v): 4*(l/16)+v for f32, l/16+4*v for f64 (gfx950).
// ------------------------------------------------------------------------
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 op(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // row of accumulator element v held by a lane in row group lk = lane/16
  static __device__ inline int row(int lk, int v) { return lk + 4 * v; }
};
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 op(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lk, int v) { return 4 * lk + v; }
};

// Work item of workgroup b of n: the hardware deals workgroups to the eight
// XCDs round robin (b % 8), so XCD x gets the contiguous item range [x q +
// min(x, n % 8), ...) with q = n / 8, and neighbouring tiles, which share operand
// panels, share that XCD's L2.  A bijection of [0, n) whatever the placement.
__device__ inline int xcd_item(int b, int n) {
  const int x = b & 7, i = b >> 3, q = n >> 3, rmd = n & 7;
  return x * q + min(x, rmd) + i;
}

// ------------------------------------------------------------------------
// assembly: one work item = one 64x64 lower tile of a parent front
// ------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_assemble(const int4* __restrict__ items, DevSym d,
                                                  T* __restrict__ F) {
  __shared__ T tile[TILE * (TILE + 1)];  // tile[c*65 + r]
  const int4 it = items[xcd_item(blockIdx.x, gridDim.x)];
  const int s = it.x, I = it.y, J = it.z;
  const int m = d.sn_m[s];
  const int r0 = I * TILE, c0 = J * TILE;
  const int nr = min(TILE, m - r0), nc = min(TILE, m - c0);
  for (int e = threadIdx.x; e < TILE * (TILE + 1); e += 256) tile[e] = (T)0;
  __syncthreads();
  for (int q = d.ch_ptr[s]; q < d.ch_ptr[s + 1]; q++) {
    const int c = d.ch_list[q];
    const int kc = d.sn_k[c], mc = d.sn_m[c], rc = mc - kc;
    const int* rel = d.rel + d.sn_roff[c];
    const int ia = lower_bound_i(rel, rc, r0), ib = lower_bound_i(rel, rc, r0 + nr);
    const int ja = lower_bound_i(rel, rc, c0), jb = lower_bound_i(rel, rc, c0 + nc);
    if (ia < ib && ja < jb) {
      const T* CB = F + d.sn_foff[c] + (size_t)kc * mc + kc;
      for (int j = ja + (threadIdx.x >> 6); j < jb; j += 4) {
        const int cj = rel[j] - c0;
        const T* col = CB + (size_t)j * mc;
        for (int i = max(ia, j) + (threadIdx.x & 63); i < ib; i += 64) {
          tile[cj * (TILE + 1) + (rel[i] - r0)] += col[i];
        }
      }
    }
    __syncthreads();
  }
  T* Fs = F + d.sn_foff[s];
  for (int e = threadIdx.x; e < TILE * TILE; e += 256) {
    const int rr = e & (TILE - 1), cc = e / TILE;
    if (rr < nr && cc < nc && r0 + rr >= c0 + cc)
      Fs[(size_t)(c0 + cc) * m + r0 + rr] = tile[cc * (TILE + 1) + rr];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_scatter(const int* __restrict__ aidx,
                                                 const long long* __restrict__ apos, int cnt,
                                                 const double* __restrict__ vals,
                                                 T* __restrict__ F) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e < cnt) F[apos[e]] += (T)vals[aidx[e]];
}

// 1/d from the hardware reciprocal and two Newton steps (within an ulp; the pivots
// are bounded away from 0 and from overflow by the static perturbation)
__device__ inline double rcp_newton(double d) {
  double y = __builtin_amdgcn_rcp(d);
  double e = fma(-d, y, 1.0);
  y = fma(y, e, y);
  e = fma(-d, y, 1.0);
  return fma(y, e, y);
}

// ------------------------------------------------------------------------
// panel diagonal block: Bunch-Kaufman restricted to the block, FP64 in LDS
//
// Four waves per front.  A block whose every column passes the no-interchange
// test is factored slab by slab (one wave pivots 16 columns in registers, all
// four apply the rank-16 update).  Otherwise the full pivoting loop runs: every
// wave does the pivot search on the same LDS data (DPP reductions, identical
// decisions), wave 0 applies interchanges, and each wave applies the rank-1/2
// update to its 16 columns of every trailing row; pivot columns stay unscaled
// until the loop ends, so a step needs one barrier (three with an interchange).
// The inverse X = L11^-1 is then formed in 16x16 blocks: diagonal blocks by
// substitution, off-diagonal blocks as X(I,J) = -X(I,I) sum_{J<=Q<I} L(I,Q) X(Q,J)
// with f64 MFMA.
// ------------------------------------------------------------------------
template <typename T>
// three workgroups per CU (VGPRs <= 168 per lane, 44 KB of LDS each)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_panel_diag(const int* __restrict__ items, int p, DevSym d,
                                                    T* __restrict__ F, T* __restrict__ E0,
                                                    T* __restrict__ E1, double pert,
                                                    const double* __restrict__ amax,
                                                    FactorCounters* __restrict__ ctr,
                                                    int* __restrict__ rctr) {
  __shared__ double A[64 * 65];     // A[row*65 + col]: lower part L/D, upper part X
  __shared__ double dv[64][3];      // 1x1: {1/d, d}; 2x2 (first column): D^-1 {i11, i21, i22}
  __shared__ double Sb[64 * 16];    // fast path: Sb[q*64 + t] = unscaled a(t, c0 + q) of the slab
  __shared__ int sfail;
  __shared__ int lp[64];
  __shared__ signed char pt[64];
  const int s = items[blockIdx.x];
  const int k = d.sn_k[s], m = d.sn_m[s], first = d.sn_first[s];
  const int j0 = p * NB, w = min(NB, k - j0);
  T* Fs = F + d.sn_foff[s];
  const int tid = threadIdx.x, t = tid & 63, wv = tid >> 6;
  auto load_block = [&]() {
    // addresses clamped into the block so the 16 loads issue back to back
    T ld[16];
    const int tc = min(t, w - 1);
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int jc = min(wv + 4 * q, w - 1);
      ld[q] = Fs[(size_t)(j0 + jc) * m + j0 + tc];
    }
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int j = wv + 4 * q;
      A[t * 65 + j] = (j < w && t >= j && t < w) ? (double)ld[q] : 0.0;
    }
    if (wv == 0) {
      lp[t] = t;
      pt[t] = 1;
    }
  };
  load_block();
  if (tid == 0) sfail = 0;
  __syncthreads();
  const double alpha = (1.0 + sqrt(17.0)) / 8.0;
  // perturbation threshold: pert * max|a_ij| of the front's connected component
  // (its assembly tree), so a system's pivots do not depend on other systems
  // factored in the same batch
  const int root = d.sn_root[s];
  const double tau = pert * amax[root];
  int npos = 0, nneg = 0, npert = 0, n2 = 0;
  // Fast path: Bunch-Kaufman takes the 1x1 pivot without interchange whenever
  // |a_kk| >= alpha * colmax (or perturbs a pivot below tau).  Wave 0 factors the
  // block in 16-column slabs held in registers (lane t: row t of the slab).  Of a
  // pivot column's entries at the slab's later rows, the one the next column
  // needs comes as a scalar (readlane); the others go through LDS, read back
  // while the pivot test (two ballots) and the reciprocal run.  A failed test
  // only clears ok (no branch in the step).  Then all four waves apply the
  // slab's rank-16 update to the rest of the block.  A slab with a failed column
  // sends the whole block back to the full pivoting loop below, from the
  // original data (the columns before the first failure are the same as that
  // loop's, so the failure is found on the values it would test).  On success
  // the lower part holds L with D on the diagonal (what the pivoting loop and its
  // scaling pass produce).
  bool fast = true;
  for (int c0 = 0; c0 < w; c0 += 16) {
    const int c1 = min(c0 + 16, w);
    if (wv == 0) {
      double r[16];
#pragma unroll
      for (int q = 0; q < 16; q++) r[q] = A[t * 65 + c0 + q];
      bool ok = true;
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int c = c0 + q;
        if (c < c1) {
          const double rq = r[q];  // column c, not yet scaled
          const double akk = readlane_d(rq, c);
          const double bn = (q < 15) ? readlane_d(rq, c0 + q + 1) : 0.0;  // a(c + 1, c)
          Sb[q * 64 + t] = rq;
          __builtin_amdgcn_wave_barrier();
          const double absakk = fabs(akk);
          const bool below = t > c && t < w;
          const double av = fabs(rq);
          // colmax < tau and |a_kk| < tau: perturbed 1x1 pivot; otherwise the
          // test |a_kk| >= alpha colmax must hold in every row below
          const unsigned long long big = __builtin_amdgcn_ballot_w64(below && av >= tau);
          const unsigned long long bad = __builtin_amdgcn_ballot_w64(below && absakk < alpha * av);
          const bool pz = (absakk < tau) & (big == 0ULL);
          ok = ok & (pz | (bad == 0ULL));
          const bool pert = pz | (absakk < tau);
          const double dd = pert ? ((akk >= 0.0) ? tau : -tau) : akk;
          npert += pert ? 1 : 0;
          npos += (dd > 0) ? 1 : 0;
          nneg += (dd > 0) ? 0 : 1;
          const double dinv = rcp_newton(dd);
          // rows on or above c take lt = 0 and keep their values (a(c0 + j, c)
          // is finite), so the update needs no select
          const double lt = below ? rq * dinv : 0.0;
          if (q < 15) r[q + 1] = fma(-lt, bn, r[q + 1]);
#pragma unroll
          for (int j = q + 2; j < 16; j++) r[j] = fma(-lt, Sb[q * 64 + c0 + j], r[j]);
          r[q] = below ? lt : (t == c ? dd : rq);
          if (t == 0) {
            dv[c][0] = dinv;
            dv[c][1] = dd;
          }
        }
      }
      if (!ok) {
        if (t == 0) sfail = 1;
      } else {
#pragma unroll
        for (int q = 0; q < 16; q++) A[t * 65 + c0 + q] = r[q];
      }
    }
    __syncthreads();
    if (uni(sfail)) {
      fast = false;
      break;
    }
    // A(t, j) -= sum_{q in slab} L(t, q) a(j, q) on the trailing block (rows and
    // columns from c1; a full slab whenever rows remain below it): its 16x16
    // tiles on and below the diagonal, one per wave at a time, K = 16 as four
    // f64 MFMA steps.  Rows and columns past w hold zeros and are never stored.
    {
      typedef Mfma<double>::v4 v4;
      const int nt = (w - c1 + 15) >> 4;  // tiles per side, at most 3
      const int lr = t & 15, lk = t >> 4;
      for (int tile = wv; tile < nt * (nt + 1) / 2; tile += 4) {
        int bi = 0, bj = tile;  // tile -> (bi, bj), bj <= bi, row by row
        while (bj > bi) {
          bj -= bi + 1;
          bi++;
        }
        const int R0 = c1 + 16 * bi, C0 = c1 + 16 * bj;
        v4 acc;
#pragma unroll
        for (int v = 0; v < 4; v++) acc[v] = A[(R0 + Mfma<double>::row(lk, v)) * 65 + C0 + lr];
#pragma unroll
        for (int kk = 0; kk < 4; kk++)
          acc = Mfma<double>::op(-A[(R0 + lr) * 65 + c0 + 4 * kk + lk],
                                 Sb[(4 * kk + lk) * 64 + C0 + lr], acc);
#pragma unroll
        for (int v = 0; v < 4; v++) A[(R0 + Mfma<double>::row(lk, v)) * 65 + C0 + lr] = acc[v];
      }
    }
    __syncthreads();
  }
  if (!fast) {
    __syncthreads();  // every wave has finished reading the block before it is reloaded
    load_block();
    npos = nneg = npert = 0;
    __syncthreads();
  }
  const int q0 = wv * 16;  // first of this wave's 16 update columns
  int kc = fast ? w : 0;
  while (kc < w) {
    const double absakk = uni(fabs(A[kc * 65 + kc]));
    double cv = (t > kc && t < w) ? fabs(A[t * 65 + kc]) : -1.0;
    int ci = t;
    wave_argmax_u(cv, ci);
    const double colmax = cv > 0.0 ? cv : 0.0;
    const int imax = cv > 0.0 ? ci : -1;
    int kstep = 1, kp = kc;
    bool perturb = false;
    if (fmax(absakk, colmax) < tau) {
      perturb = true;
    } else if (absakk >= alpha * colmax) {
      kp = kc;
    } else {
      double rv = 0.0;
      if (t >= kc && t < w && t != imax)
        rv = (t < imax) ? fabs(A[imax * 65 + t]) : fabs(A[t * 65 + imax]);
      const double rowmax = wave_max_u(rv);
      if (absakk * rowmax >= alpha * colmax * colmax) kp = kc;
      else if (uni(fabs(A[imax * 65 + imax])) >= alpha * rowmax) kp = imax;
      else { kp = imax; kstep = 2; }
    }
    const int kk = kc + kstep - 1;
    if (kp != kk) {
      // symmetric interchange of kk and kp inside the block (lower storage); the
      // same rows of earlier panels are permuted once at the end (lp)
      __syncthreads();  // every wave is done reading column kc and row imax
      if (wv == 0) {
        const int a = kk, b = kp;
        if (t == 0) {
          double tmp = A[a * 65 + a]; A[a * 65 + a] = A[b * 65 + b]; A[b * 65 + b] = tmp;
          int ti = lp[a]; lp[a] = lp[b]; lp[b] = ti;
        }
        if (t > a && t < b) { double tmp = A[t * 65 + a]; A[t * 65 + a] = A[b * 65 + t]; A[b * 65 + t] = tmp; }
        if (t > b && t < w) { double tmp = A[t * 65 + a]; A[t * 65 + a] = A[t * 65 + b]; A[t * 65 + b] = tmp; }
        if (t < a) { double tmp = A[a * 65 + t]; A[a * 65 + t] = A[b * 65 + t]; A[b * 65 + t] = tmp; }
      }
      __syncthreads();
    }
    if (kstep == 1) {
      double dd = uni(A[kc * 65 + kc]);
      const bool pz = perturb || fabs(dd) < tau;
      if (pz) {
        dd = (dd >= 0.0) ? tau : -tau;
        npert++;
      }
      if (dd > 0) npos++; else nneg++;
      const double dinv = 1.0 / dd;
      // trailing rows: A(t, j) -= (a(t) / d) a(j) for j > kc; entries above the
      // diagonal are scratch while the pivot loop runs
      if (q0 + 16 > kc + 1 && t > kc && t < w) {
        const double lt = A[t * 65 + kc] * dinv;
        double* row = A + t * 65 + q0;
        double rr[16], cc[16];
#pragma unroll
        for (int q = 0; q < 16; q++) {
          rr[q] = row[q];
          cc[q] = A[(q0 + q) * 65 + kc];
        }
        if (q0 > kc) {
#pragma unroll
          for (int q = 0; q < 16; q++) row[q] = fma(-lt, cc[q], rr[q]);
        } else {
#pragma unroll
          for (int q = 0; q < 16; q++) row[q] = (q0 + q > kc) ? fma(-lt, cc[q], rr[q]) : rr[q];
        }
      }
      if (tid == 0) {
        pt[kc] = 1;
        dv[kc][0] = dinv;
        dv[kc][1] = dd;
      }
    } else {
      const int g1 = kc, g2 = kc + 1;
      const double d11 = uni(A[g1 * 65 + g1]), d21 = uni(A[g2 * 65 + g1]), d22 = uni(A[g2 * 65 + g2]);
      const double det = d11 * d22 - d21 * d21;
      const double rdet = 1.0 / det;
      const double i11 = d22 * rdet, i21 = -d21 * rdet, i22 = d11 * rdet;
      if (det < 0) { npos++; nneg++; }
      else if (d11 + d22 > 0) npos += 2;
      else nneg += 2;
      n2++;
      // A(t, j) -= [a1(t) a2(t)] D^-1 [a1(j) a2(j)]^T for j > g2
      if (q0 + 16 > g2 + 1 && t > g2 && t < w) {
        const double a1 = A[t * 65 + g1], a2 = A[t * 65 + g2];
        const double l1 = a1 * i11 + a2 * i21, l2 = a1 * i21 + a2 * i22;
        double* row = A + t * 65 + q0;
        double rr[16], x1[16], x2[16];
#pragma unroll
        for (int q = 0; q < 16; q++) {
          rr[q] = row[q];
          x1[q] = A[(q0 + q) * 65 + g1];
          x2[q] = A[(q0 + q) * 65 + g2];
        }
        if (q0 > g2) {
#pragma unroll
          for (int q = 0; q < 16; q++) row[q] = fma(-l2, x2[q], fma(-l1, x1[q], rr[q]));
        } else {
#pragma unroll
          for (int q = 0; q < 16; q++)
            row[q] = (q0 + q > g2) ? fma(-l2, x2[q], fma(-l1, x1[q], rr[q])) : rr[q];
        }
      }
      if (tid == 0) {
        pt[g1] = 2;
        pt[g2] = 0;
        dv[g1][0] = i11;
        dv[g1][1] = i21;
        dv[g1][2] = i22;
      }
    }
    __syncthreads();
    kc += kstep;
  }
  // scale the pivot columns: L(t,c) = a(t,c) / d_c, or [a1 a2] D^-1 for a 2x2 pivot
  // (the wave of a pair's first column writes both columns)
  for (int c = wv; c < w && !fast; c += 4) {
    const int pc = pt[c];
    if (pc == 1) {
      if (t > c && t < w) A[t * 65 + c] *= dv[c][0];
      if (t == c) A[c * 65 + c] = dv[c][1];
    } else if (pc == 2) {
      if (t > c + 1 && t < w) {
        const double a1 = A[t * 65 + c], a2 = A[t * 65 + c + 1];
        A[t * 65 + c] = a1 * dv[c][0] + a2 * dv[c][1];
        A[t * 65 + c + 1] = a1 * dv[c][1] + a2 * dv[c][2];
      }
    }
  }
  // rows j0..j0+w of the earlier panels' columns follow the interchanges (new row
  // t = old row lp[t]); each wave moves its own columns, loading before storing
  if (j0 > 0 && !fast) {
    int moved = 0;
    for (int q = 0; q < w; q++) moved |= (lp[q] != q);
    if (uni(moved)) {
      const int src = (t < w) ? lp[t] : t;
      for (int c = 4 * wv; c < j0; c += 16) {
        T v0 = (T)0, v1 = (T)0, v2 = (T)0, v3 = (T)0;
        if (t < w) {
          v0 = Fs[(size_t)c * m + j0 + src];
          v1 = Fs[(size_t)(c + 1) * m + j0 + src];
          v2 = Fs[(size_t)(c + 2) * m + j0 + src];
          v3 = Fs[(size_t)(c + 3) * m + j0 + src];
        }
        __builtin_amdgcn_wave_barrier();
        if (t < w) {
          Fs[(size_t)c * m + j0 + t] = v0;
          Fs[(size_t)(c + 1) * m + j0 + t] = v1;
          Fs[(size_t)(c + 2) * m + j0 + t] = v2;
          Fs[(size_t)(c + 3) * m + j0 + t] = v3;
        }
      }
    }
  }
  __syncthreads();
  // X = L11^-1, unit lower.  The slot (c+1,c) of a 2x2 pivot holds d21 and counts
  // as 0 in L11: it is zeroed while X is formed and restored afterwards.
  // X(i,l), l < i, is stored at A[l*65 + i].
  const bool slot = (wv == 0) && t < w && pt[t] == 2;
  double d21 = 0.0;
  if (slot) {
    d21 = A[(t + 1) * 65 + t];
    A[(t + 1) * 65 + t] = 0.0;
  }
  __syncthreads();
  // diagonal 16x16 blocks: wave b inverts block b, lane c < 16 forms column c by
  // substitution, X(r,c) = -sum_{l<r} L(r,l) X(l,c), with the column kept in Sb
  // as X(l,c) = 0 (l < c), 1 (l = c), so every lane runs the same loop
  if (t < 16) {
    const int b0 = wv * 16, c = t;
    double x[16];  // column c of X(b, b), in registers
#pragma unroll
    for (int l = 0; l < 16; l++) x[l] = (l == c) ? 1.0 : 0.0;
#pragma unroll
    for (int r = 1; r < 16; r++) {
      // row r of L, the same in every lane.  Its offset is tied to x[r - 2], so
      // its loads issue while row r - 1 is computed and no earlier (issued all
      // at once, the 120 loads would not fit in registers)
      int off = (b0 + r) * 65 + b0;
      if (r >= 2) __asm__ __volatile__("" : "+v"(off) : "v"(x[r - 2]));
      const double* lrow = A + off;
      double a0 = 0.0, a1 = 0.0;  // two chains halve the dependent FMAs
#pragma unroll
      for (int l = 0; l < r; l++) {
        if (l & 1) a1 = fma(lrow[l], x[l], a1);
        else a0 = fma(lrow[l], x[l], a0);
      }
      x[r] = (r > c) ? -(a0 + a1) : x[r];
    }
#pragma unroll
    for (int r = 1; r < 16; r++)
      if (r > c) A[(b0 + c) * 65 + b0 + r] = x[r];
  }
  __syncthreads();
  // off-diagonal 16x16 blocks with f64 MFMA, block row bi after block row bi-1;
  // wave bj forms T = sum_{bj<=bl<bi} L(bi,bl) X(bl,bj), then X(bi,bj) = -X(bi,bi) T.
  // T's accumulator element v is exactly the B operand of the second product's
  // k-step v, so T stays in registers.
  {
    typedef Mfma<double>::v4 v4;
    const int lr = t & 15, lk = t >> 4;
    for (int bi = 1; bi < 4; bi++) {
      if (wv < bi) {
        const int bj = wv;
        v4 acc = (v4)(0.0);
        for (int bl = bj; bl < bi; bl++) {
#pragma unroll
          for (int kk = 0; kk < 4; kk++) {
            const int R = 16 * bi + lr, C = 16 * bl + 4 * kk + lk;
            const int XC = 16 * bj + lr;  // X(C, XC)
            const double b = (C > XC) ? A[XC * 65 + C] : (C == XC ? 1.0 : 0.0);
            acc = Mfma<double>::op(A[R * 65 + C], b, acc);
          }
        }
        v4 res = (v4)(0.0);
#pragma unroll
        for (int kk = 0; kk < 4; kk++) {
          const int XR = 16 * bi + lr, XC = 16 * bi + 4 * kk + lk;  // X(XR, XC)
          const double a = (XR > XC) ? A[XC * 65 + XR] : (XR == XC ? 1.0 : 0.0);
          res = Mfma<double>::op(a, acc[kk], res);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) A[(16 * bj + lr) * 65 + 16 * bi + Mfma<double>::row(lk, v)] = -res[v];
      }
      __syncthreads();
    }
  }
  if (slot) A[(t + 1) * 65 + t] = d21;
  __syncthreads();
  // lower: L and D; upper (row t, column j > t): X(j, t)
  for (int j = wv; j < w; j += 4)
    if (t < w) Fs[(size_t)(j0 + j) * m + j0 + t] = (T)A[t * 65 + j];
  if (wv == 0 && t < w) {
    d.lperm[first + j0 + t] = j0 + lp[t];
    d.ilperm[first + j0 + lp[t]] = j0 + t;
    d.piv[first + j0 + t] = pt[t];
    // D^-1 for the solves: z_t = E0_t y_t + E1_t y_partner (partner t+1 / t-1
    // for the first / second column of a 2x2 pivot)
    const int pc = pt[t];
    const double e0 = (pc == 1) ? dv[t][0] : (pc == 2 ? dv[t][0] : dv[t - 1][2]);
    const double e1 = (pc == 1) ? 0.0 : (pc == 2 ? dv[t][1] : dv[t - 1][1]);
    E0[first + j0 + t] = (T)e0;
    E1[first + j0 + t] = (T)e1;
  }
  if (tid == 0) {
    atomicAdd(&ctr->npos, npos);
    atomicAdd(&ctr->nneg, nneg);
    atomicAdd(&ctr->npert, npert);
    atomicAdd(&ctr->n2x2, n2);
    if (!fast) atomicAdd(&ctr->npivoted, 1);
    int* rc = rctr + 4 * root;
    atomicAdd(rc, npos);
    atomicAdd(rc + 1, nneg);
    atomicAdd(rc + 2, npert);
    atomicAdd(rc + 3, n2);
  }
}

// ------------------------------------------------------------------------
// panel TRSM as a GEMM with the stored inverse: W = A21 L11^-T, L21 = W D^-1
// item = (supernode, 64-row block below the diagonal block)
// ------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_panel_trsm(const int2* __restrict__ items, int p,
                                                    int slot, DevSym d, T* __restrict__ F,
                                                    T* __restrict__ Wb) {
  constexpr int LD = 65;
  __shared__ T As[64 * LD];  // As[r*LD + l]: row r, permuted panel column l
  __shared__ T Li[64 * LD];  // Li[q*LD + l] = Linv[q][l]
  __shared__ T Dg[64], Dsub[64];
  __shared__ int sig[64];
  __shared__ signed char pv[64];
  typedef typename Mfma<T>::v4 v4;
  const int2 it = items[xcd_item(blockIdx.x, gridDim.x)];
  const int s = it.x, rb = it.y;
  const int k = d.sn_k[s], m = d.sn_m[s], first = d.sn_first[s];
  const int j0 = p * NB, w = min(NB, k - j0), r0 = j0 + w;
  const int rbase = r0 + rb * 64, nr = min(64, m - rbase);
  T* Fs = F + d.sn_foff[s];
  if (threadIdx.x < 64) {
    const int q = threadIdx.x;
    sig[q] = 0; pv[q] = 1; Dg[q] = (T)1; Dsub[q] = (T)0;
    if (q < w) {
      sig[q] = d.lperm[first + j0 + q] - j0;
      pv[q] = d.piv[first + j0 + q];
      Dg[q] = Fs[(size_t)(j0 + q) * m + j0 + q];
      if (q + 1 < w) Dsub[q] = Fs[(size_t)(j0 + q) * m + j0 + q + 1];
    }
  }
  __syncthreads();
  {
    // this thread's 16 elements (row r, columns c00 + 4q): all loads (addresses
    // clamped into the block) issue before the LDS stores
    const int r = threadIdx.x & 63, c00 = threadIdx.x >> 6;
    const int rl = min(r, w - 1), rn = min(r, nr - 1);
    T lvv[16], avv[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int cl = min(c00 + 4 * q, w - 1);
      // Linv[c][r] for r < c lives at front (row j0+r, col j0+c)
      lvv[q] = Fs[(size_t)(j0 + cl) * m + j0 + rl];
      avv[q] = Fs[(size_t)(j0 + sig[cl]) * m + rbase + rn];
    }
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int c = c00 + 4 * q;
      Li[c * LD + r] = (c < w && r < w) ? ((r < c) ? lvv[q] : (r == c ? (T)1 : (T)0)) : (T)0;
      As[r * LD + c] = (r < nr && c < w) ? avv[q] : (T)0;
    }
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  v4 acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++) acc[bi][bj] = (v4)((T)0);
  const int kmax = (w + 3) & ~3;
  for (int kk = 0; kk < kmax; kk += 4) {
    const T a0 = As[(wr + lr) * LD + kk + lk];
    const T a1 = As[(wr + 16 + lr) * LD + kk + lk];
    const T b0 = Li[(wc + lr) * LD + kk + lk];
    const T b1 = Li[(wc + 16 + lr) * LD + kk + lk];
    acc[0][0] = Mfma<T>::op(a0, b0, acc[0][0]);
    acc[0][1] = Mfma<T>::op(a0, b1, acc[0][1]);
    acc[1][0] = Mfma<T>::op(a1, b0, acc[1][0]);
    acc[1][1] = Mfma<T>::op(a1, b1, acc[1][1]);
  }
  __syncthreads();
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int v = 0; v < 4; v++)
        As[(wr + bi * 16 + Mfma<T>::row(lk, v)) * LD + wc + bj * 16 + lr] = acc[bi][bj][v];
  __syncthreads();
  T* Ws = Wb + d.sn_woff[s] + (size_t)slot * m * NB;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int q = e >> 6, r = e & 63;
    if (r >= nr || q >= w) continue;
    const T xq = As[r * LD + q];
    T l;
    const int pq = pv[q];
    if (pq == 1) {
      l = xq / Dg[q];
    } else if (pq == 2) {
      const T det = Dg[q] * Dg[q + 1] - Dsub[q] * Dsub[q];
      l = (xq * Dg[q + 1] - As[r * LD + q + 1] * Dsub[q]) / det;
    } else {
      const T det = Dg[q - 1] * Dg[q] - Dsub[q - 1] * Dsub[q - 1];
      l = (xq * Dg[q - 1] - As[r * LD + q - 1] * Dsub[q - 1]) / det;
    }
    Ws[(size_t)q * m + rbase + r] = xq;
    Fs[(size_t)(j0 + q) * m + rbase + r] = l;
  }
}

// ------------------------------------------------------------------------
// Schur update with MFMA, C -= sum over up to UPD_GROUP panels of W_p L_p^T:
// the panel W = L21 D of each panel (its W-buffer slot) times that panel's
// columns of L at the tile's column rows.  C is read and written once per group
// of panels (K = 64 per panel, accumulated in registers).  item =
// (supernode, row tile I, column tile J, np | rs << 3): the region starts at
// front row / column rs; J = -1 is the look-ahead form, whose columns are the
// next panel's pivot columns [rs, rs + min(64, k - rs)).  Only the lower part
// (row >= column) is stored.
// ------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_panel_update(const int4* __restrict__ items, int pfirst,
                                                      DevSym d, T* __restrict__ F,
                                                      const T* __restrict__ Wb) {
  constexpr int LD = 65;
  __shared__ T Ws[64 * LD];  // Ws[r*LD + t]
  __shared__ T Ls[64 * LD];
  typedef typename Mfma<T>::v4 v4;
  const int4 it = items[xcd_item(blockIdx.x, gridDim.x)];
  const int s = it.x, I = it.y, J = it.z, np = it.w & 7, rs = it.w >> 3;
  const int k = d.sn_k[s], m = d.sn_m[s];
  const bool look = J < 0;
  const int rI = rs + I * TILE, rJ = look ? rs : rs + J * TILE;
  const int nI = min(TILE, m - rI), nJ = look ? min(TILE, k - rs) : min(TILE, m - rJ);
  const T* Wsrc = Wb + d.sn_woff[s];
  T* Fs = F + d.sn_foff[s];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  v4 acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = wr + bi * 16 + Mfma<T>::row(lk, v), col = wc + bj * 16 + lr;
        acc[bi][bj][v] = (row < nI && col < nJ) ? Fs[(size_t)(rJ + col) * m + rI + row] : (T)0;
      }
  for (int pp = 0; pp < np; pp++) {
    const int j0 = (pfirst + pp) * NB, w = min(NB, k - j0);
    const T* Wp = Wsrc + (size_t)pp * m * NB;
    if (pp > 0) __syncthreads();  // the last panel's tiles are consumed
    {
      // row r, panel columns t00 + 4q: loads first (clamped addresses), then LDS
      const int r = threadIdx.x & 63, t00 = threadIdx.x >> 6;
      const int rw = min(r, nI - 1), rl = min(r, nJ - 1);
      T wv[16], lv[16];
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int tc = min(t00 + 4 * q, w - 1);
        wv[q] = Wp[(size_t)tc * m + rI + rw];
        lv[q] = Fs[(size_t)(j0 + tc) * m + rJ + rl];
      }
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int t = t00 + 4 * q;
        Ws[r * LD + t] = (r < nI && t < w) ? wv[q] : (T)0;
        Ls[r * LD + t] = (r < nJ && t < w) ? lv[q] : (T)0;
      }
    }
    __syncthreads();
    const int kmax = (w + 3) & ~3;
    for (int kk = 0; kk < kmax; kk += 4) {
      const T a0 = -Ws[(wr + lr) * LD + kk + lk];
      const T a1 = -Ws[(wr + 16 + lr) * LD + kk + lk];
      const T b0 = Ls[(wc + lr) * LD + kk + lk];
      const T b1 = Ls[(wc + 16 + lr) * LD + kk + lk];
      acc[0][0] = Mfma<T>::op(a0, b0, acc[0][0]);
      acc[0][1] = Mfma<T>::op(a0, b1, acc[0][1]);
      acc[1][0] = Mfma<T>::op(a1, b0, acc[1][0]);
      acc[1][1] = Mfma<T>::op(a1, b1, acc[1][1]);
    }
  }
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = wr + bi * 16 + Mfma<T>::row(lk, v), col = wc + bj * 16 + lr;
        if (row < nI && col < nJ && rI + row >= rJ + col)
          Fs[(size_t)(rJ + col) * m + rI + row] = acc[bi][bj][v];
      }
}

// ------------------------------------------------------------------------
// Inverse of a supernode's whole unit-lower pivot block, X = L11^-1, built one
// 64-row block per panel step: X(P,CB) = -Linv_P * sum_{Q=CB}^{P-1} L(P,Q) X(Q,CB).
// X(i,l), l < i, is kept above the diagonal at front position (row l, col i);
// the 64x64 diagonal blocks of X are the per-panel inverses from k_panel_diag.
// item = (supernode, column block cb < p)
// ------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_panel_xrow(const int2* __restrict__ items, int p,
                                                    DevSym d, T* __restrict__ F) {
  constexpr int LD = 65;
  __shared__ T As[64 * LD];  // A operand, As[r*LD + kk]
  __shared__ T Bn[64 * LD];  // B operand stored n-major, Bn[n*LD + kk]
  typedef typename Mfma<T>::v4 v4;
  const int2 it = items[xcd_item(blockIdx.x, gridDim.x)];
  const int s = it.x, cb = it.y;
  const int k = d.sn_k[s], m = d.sn_m[s];
  const int j0 = p * NB, w = min(NB, k - j0), c0 = cb * 64;
  T* Fs = F + d.sn_foff[s];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  v4 acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++) acc[bi][bj] = (v4)((T)0);
  for (int q = cb; q < p; q++) {
    const int q0 = q * 64;
    {
      // row rr, columns cc00 + 4i: loads first (clamped), then LDS
      const int rr = threadIdx.x & 63, cc00 = threadIdx.x >> 6, rw = min(rr, w - 1);
      T av[16], xw[16];
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const int cc = cc00 + 4 * i;
        av[i] = Fs[(size_t)(q0 + cc) * m + j0 + rw];  // L(j0+rr, q0+cc): column q0+cc
        xw[i] = Fs[(size_t)(q0 + cc) * m + c0 + rr];  // X(q0+cc, c0+rr)
      }
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const int cc = cc00 + 4 * i;
        As[rr * LD + cc] = (rr < w) ? av[i] : (T)0;
        // X(q0+cc, c0+rr) as Bn[n=rr][kk=cc]
        Bn[rr * LD + cc] = (q == cb) ? ((cc > rr) ? xw[i] : (cc == rr ? (T)1 : (T)0)) : xw[i];
      }
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < 64; kk += 4) {
      const T a0 = As[(wr + lr) * LD + kk + lk], a1 = As[(wr + 16 + lr) * LD + kk + lk];
      const T b0 = Bn[(wc + lr) * LD + kk + lk], b1 = Bn[(wc + 16 + lr) * LD + kk + lk];
      acc[0][0] = Mfma<T>::op(a0, b0, acc[0][0]);
      acc[0][1] = Mfma<T>::op(a0, b1, acc[0][1]);
      acc[1][0] = Mfma<T>::op(a1, b0, acc[1][0]);
      acc[1][1] = Mfma<T>::op(a1, b1, acc[1][1]);
    }
    __syncthreads();
  }
  // R (64 x 64, rows of P, columns of CB) -> Bn[n][kk=r]; Linv_P -> As
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int v = 0; v < 4; v++)
        Bn[(wc + bj * 16 + lr) * LD + wr + bi * 16 + Mfma<T>::row(lk, v)] = acc[bi][bj][v];
  {
    // Linv_P(rr, cc) at front (row j0+cc, col j0+rr); column cc fixed per thread
    const int cc = threadIdx.x & 63, rr00 = threadIdx.x >> 6, cw = min(cc, w - 1);
    T lv[16];
#pragma unroll
    for (int i = 0; i < 16; i++) lv[i] = Fs[(size_t)(j0 + min(rr00 + 4 * i, w - 1)) * m + j0 + cw];
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int rr = rr00 + 4 * i;
      const T v = (rr < w && cc < w) ? ((cc < rr) ? lv[i] : (cc == rr ? (T)1 : (T)0)) : (T)0;
      As[rr * LD + cc] = -v;
    }
  }
  __syncthreads();
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++) acc[bi][bj] = (v4)((T)0);
#pragma unroll 4
  for (int kk = 0; kk < 64; kk += 4) {
    const T a0 = As[(wr + lr) * LD + kk + lk], a1 = As[(wr + 16 + lr) * LD + kk + lk];
    const T b0 = Bn[(wc + lr) * LD + kk + lk], b1 = Bn[(wc + 16 + lr) * LD + kk + lk];
    acc[0][0] = Mfma<T>::op(a0, b0, acc[0][0]);
    acc[0][1] = Mfma<T>::op(a0, b1, acc[0][1]);
    acc[1][0] = Mfma<T>::op(a1, b0, acc[1][0]);
    acc[1][1] = Mfma<T>::op(a1, b1, acc[1][1]);
  }
  // X(j0+r, c0+n) at front (row c0+n, col j0+r)
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = wr + bi * 16 + Mfma<T>::row(lk, v), n = wc + bj * 16 + lr;
        if (r < w) Fs[(size_t)(j0 + r) * m + c0 + n] = acc[bi][bj][v];
      }
}

// ------------------------------------------------------------------------
// Solve form of the factor.  After the last panel step two post-passes turn
// every front into
//   lower part (rows > column):  G = [X; Z],  X = L11^-1, Z = L21 X
//   upper part (row < column):   G^T (X^T in the pivot block, Z^T to its right)
// with D^-1 kept in two vectors (E0, E1).  Then per supernode
//   forward : gather f, then [y; u] = [f; f_cb] + [X - I; -Z] f
//             (one wave per front row, reading column i of the upper copy)
//   z = D^-1 y for all pivots at once
//   backward: x = z + G^T [z; -x_rows]
//             (one wave per pivot column, reading column c of the lower copy)
// so every product streams contiguous columns, and a level costs one launch
// forward (two where a separate gather launch is faster) and one backward.
// ------------------------------------------------------------------------

// Z^T(c, j) = sum_{i >= c} X(i, c) L21(j, i) into front (row c, column k + j);
// item = (supernode, 64-column block of Z, 64-row block of L21)
template <typename T>
__global__ __launch_bounds__(256) void k_zt(const int4* __restrict__ items, DevSym d,
                                            T* __restrict__ F) {
  constexpr int LD = 65;
  __shared__ T As[64 * LD];  // As[c*LD + i] = X(i0 + i, c0 + c)
  __shared__ T Bs[64 * LD];  // Bs[j*LD + i] = L21(jr0 + j, i0 + i)
  typedef typename Mfma<T>::v4 v4;
  const int4 it = items[xcd_item(blockIdx.x, gridDim.x)];
  const int s = it.x, cb = it.y, jb = it.z;
  const int k = d.sn_k[s], m = d.sn_m[s], r = m - k;
  const int c0 = cb * 64, jr0 = jb * 64;
  const int nc = min(64, k - c0), nj = min(64, r - jr0);
  T* Fs = F + d.sn_foff[s];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  v4 acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++) acc[bi][bj] = (v4)((T)0);
  for (int i0 = c0; i0 < k; i0 += 64) {
    const int ni = min(64, k - i0);
    {
      // column offset cc fixed per thread, rows ii00 + 4q: loads first (clamped)
      const int cc = threadIdx.x & 63, ii00 = threadIdx.x >> 6;
      const int cx = min(cc, nc - 1), cl = min(cc, nj - 1);
      T xw[16], lw[16];
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int gi = i0 + min(ii00 + 4 * q, ni - 1);
        xw[q] = Fs[(size_t)gi * m + c0 + cx];
        lw[q] = Fs[(size_t)gi * m + k + jr0 + cl];
      }
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int ii = ii00 + 4 * q, gi = i0 + ii, gc = c0 + cc;
        T xv = (T)0, lv = (T)0;
        if (ii < ni) {
          if (cc < nc) xv = (gi > gc) ? xw[q] : (gi == gc ? (T)1 : (T)0);
          if (cc < nj) lv = lw[q];
        }
        As[cc * LD + ii] = xv;
        Bs[cc * LD + ii] = lv;
      }
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < 64; kk += 4) {
      const T a0 = As[(wr + lr) * LD + kk + lk], a1 = As[(wr + 16 + lr) * LD + kk + lk];
      const T b0 = Bs[(wc + lr) * LD + kk + lk], b1 = Bs[(wc + 16 + lr) * LD + kk + lk];
      acc[0][0] = Mfma<T>::op(a0, b0, acc[0][0]);
      acc[0][1] = Mfma<T>::op(a0, b1, acc[0][1]);
      acc[1][0] = Mfma<T>::op(a1, b0, acc[1][0]);
      acc[1][1] = Mfma<T>::op(a1, b1, acc[1][1]);
    }
    __syncthreads();
  }
  // C(c, j) -> As[j*LD + c], then columns k + jr0 + j are written contiguously in c
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int v = 0; v < 4; v++)
        As[(wc + bj * 16 + lr) * LD + wr + bi * 16 + Mfma<T>::row(lk, v)] = acc[bi][bj][v];
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int jj = e >> 6, cc = e & 63;
    if (jj < nj && cc < nc) Fs[(size_t)(k + jr0 + jj) * m + c0 + cc] = As[jj * LD + cc];
  }
}

// lower part <- transpose of the upper part, one 64x64 tile per item:
//   {s, ib, cb, 0}: pivot block, lower tile (ib, cb), ib >= cb, strictly below the diagonal
//   {s, cb, jb, 1}: Z tile, rows k + 64 jb.., columns 64 cb..
template <typename T>
__global__ __launch_bounds__(256) void k_gform(const int4* __restrict__ items, DevSym d,
                                               T* __restrict__ F) {
  __shared__ T tile[64 * 65];  // tile[x*65 + y]: source column offset x, source row offset y
  const int4 it = items[xcd_item(blockIdx.x, gridDim.x)];
  const int s = it.x;
  const int k = d.sn_k[s], m = d.sn_m[s];
  T* Fs = F + d.sn_foff[s];
  int srow0, scol0, nrow, ncol;  // source tile in the upper part: rows, columns
  if (it.w == 0) {
    srow0 = it.z * 64; scol0 = it.y * 64;
    nrow = min(64, k - srow0); ncol = min(64, k - scol0);
  } else {
    srow0 = it.y * 64; scol0 = k + it.z * 64;
    nrow = min(64, k - srow0); ncol = min(64, m - scol0);
  }
  {
    // source row y fixed per thread, columns x00 + 4q: loads first (clamped)
    const int y = threadIdx.x & 63, x00 = threadIdx.x >> 6, yc = min(y, nrow - 1);
    T v[16];
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = Fs[(size_t)(scol0 + min(x00 + 4 * q, ncol - 1)) * m + srow0 + yc];
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int x = x00 + 4 * q;
      if (x < ncol && y < nrow) tile[x * 65 + y] = v[q];
    }
  }
  __syncthreads();
  // destination: row scol0 + x, column srow0 + y
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int y = e >> 6, x = e & 63;
    if (x < ncol && y < nrow && scol0 + x > srow0 + y)
      Fs[(size_t)(srow0 + y) * m + scol0 + x] = tile[x * 65 + y];
  }
}

// 16-byte vectors of T.  Global loads need only element (dword) alignment on
// this target, so any contiguous run of a front column or vector loads as
// 4 floats / 2 doubles per instruction.
template <typename T>
struct V16;
template <>
struct V16<float> {
  typedef float v __attribute__((ext_vector_type(4)));
};
template <>
struct V16<double> {
  typedef double v __attribute__((ext_vector_type(2)));
};
template <>
struct V16<_Float16> {
  typedef _Float16 v __attribute__((ext_vector_type(8)));
};
template <typename T>
__device__ inline typename V16<T>::v ldv(const T* p) {
  typename V16<T>::v r;
  __builtin_memcpy(&r, p, 16);
  return r;
}
// W consecutive entries of a vector of T: the operand that pairs with one
// 16-byte chunk of the factor's storage type S (W = 16 / sizeof(S); one 16-byte
// load when S is T, two for half-precision storage of an FP32 solve)
template <typename T, int W>
struct VecW {
  T e[W];
};
template <typename T, int W>
__device__ inline VecW<T, W> ldw(const T* p) {
  VecW<T, W> r;
  __builtin_memcpy(&r, p, W * sizeof(T));
  return r;
}
template <typename T>
__device__ inline T vdot(typename V16<T>::v a, typename V16<T>::v b, T acc) {
#pragma unroll
  for (int w = 0; w < (int)(16 / sizeof(T)); w++) acc = fma(a[w], b[w], acc);
  return acc;
}
// Partial sum_{e < len} a[e] b[e] of lane sl of a segment of seg lanes, a in the
// storage type S of the factor (T, or half precision for an FP32 solve; products
// and sums in T).  The lane takes the 16-byte chunks of a (W = 16 / sizeof(S)
// entries) sl, sl + seg, ... in groups of G whose loads all issue before the
// first use; the last chunk may run up to W-1 elements past len (the callers'
// arrays are followed by mapped memory), and every element at or past len is
// dropped by a select, so whatever it holds never reaches the sum.
template <typename T, typename S = T, int G = 4>
__device__ inline T seg_dot(const S* __restrict__ a, const T* __restrict__ b, int len, int sl,
                            int seg) {
  constexpr int W = 16 / sizeof(S);
  typedef typename V16<S>::v vs;
  const int nc = (len + W - 1) / W;
  T acc = (T)0;
  for (int q0 = sl; q0 < nc; q0 += G * seg) {
    vs x[G];
    VecW<T, W> y[G];
#pragma unroll
    for (int u = 0; u < G; u++) {
      const int q = q0 + u * seg, qc = q < nc ? q : 0;
      x[u] = ldv(a + qc * W);
      y[u] = ldw<T, W>(b + qc * W);
    }
#pragma unroll
    for (int u = 0; u < G; u++)
#pragma unroll
      for (int w = 0; w < W; w++) {
        const T t = fma((T)x[u][w], y[u].e[w], acc);
        acc = ((q0 + u * seg) * W + w < len) ? t : acc;
      }
  }
  return acc;
}

template <typename T>
struct VI16;  // the index vector that goes with V16<T>
template <>
struct VI16<float> {
  typedef int v __attribute__((ext_vector_type(4)));
};
template <>
struct VI16<double> {
  typedef int v __attribute__((ext_vector_type(2)));
};
// a + the children's update entries at one front position: sl points at the
// position's entry of child 0 in the supernode's slot block, child j's is
// nch_m = m entries further on each; up to eight loads issue before the first
// add (nch is uniform over the workgroup), added in ascending child order
template <typename T>
__device__ inline T slot_sum(const T* __restrict__ sl, int nch, int m, T a) {
  for (int j0 = 0; j0 < nch; j0 += 8) {
    T v[8];
#pragma unroll
    for (int q = 0; q < 8; q++)
      if (j0 + q < nch) v[q] = sl[(size_t)(j0 + q) * m];
#pragma unroll
    for (int q = 0; q < 8; q++)
      if (j0 + q < nch) a += v[q];
  }
  return a;
}

// the same for the four positions q0 + 256 j (j < 4) below qend: a[j] += the
// children's entries at position q0 + 256 j, in ascending child order; every
// load of a pass (four positions x four children) issues before the first add
template <typename T>
__device__ inline void slot_sum4(const T* __restrict__ sb, int nch, int m, int q0, int qend, T a[4]) {
  for (int c0 = 0; c0 < nch; c0 += 4) {
    T v[4][4];
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const int q = q0 + 256 * j;
        v[j][c] = (q < qend && c0 + c < nch) ? sb[q + (size_t)(c0 + c) * m] : (T)0;
      }
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int c = 0; c < 4; c++)
        if (c0 + c < nch) a[j] += v[j][c];
  }
}

// forward, a whole small front (k <= FWD_SMALL_K, m <= FWD_SMALL_M) in one
// workgroup: one thread per position p gathers f_p = b_p + the children's update
// entries at p (summed in ascending order) into LDS in pivot order, then one
// thread per row: y_i = f_i + G(i,:) f (i < k), and for i >= k the update entry
// f_i - G(i,:) f goes to the supernode's slot in its parent's block, G(i,:) read
// from the row's packed forward entries (contiguous per thread)
static_assert(FWD_SMALL_M == 256, "one thread per row of a small front");
template <typename T, typename S>
__device__ inline void fwd_small_front(const SolveItem& it, const DevSym& d, const S* __restrict__ F,
                                       const T* __restrict__ bp, T* __restrict__ y, T* __restrict__ u) {
  __shared__ T g[256];
  __shared__ __align__(16) T fs[FWD_SMALL_K + 4];  // chunk loads may run 3 past k
  const int k = it.k, m = it.m, first = it.first;
  const int p = threadIdx.x;
  // row p of G from the lower copy (column l, row p): consecutive threads read
  // consecutive entries (a column past the row's length re-reads column 0 and is
  // dropped by a select).  The first HOIST loads issue before the gather, the
  // rest after it: half of them, so the forward kernel's registers stay at 64
  // (FP32) and eight waves fit per SIMD.
  constexpr int HOIST = FWD_SMALL_K / 2;
  const int len = (p < m) ? min(p, k) : 0;
  const S* rowp = F + it.poff + (p < m ? prow_off(p, k) : 0);
  T v[FWD_SMALL_K];
#pragma unroll
  for (int l = 0; l < HOIST; l++) v[l] = (T)rowp[l < len ? l : 0];
  // destination of a contribution row's update entry in the parent's block
  const int dr = (p >= k && p < m) ? d.rel[it.roff + p - k] : 0;
  T acc = (T)0;
  int ip = 0;
  if (p < m) {
    if (p < k) {
      acc = bp[first + p];
      ip = d.ilperm[first + p];
    }
    acc = slot_sum(u + it.soff + p, it.nch, m, acc);
  }
  g[threadIdx.x] = acc;
  __syncthreads();
  if (p < k) fs[ip] = acc;
  __syncthreads();
  if (p < m) {
#pragma unroll
    for (int l = HOIST; l < FWD_SMALL_K; l++) v[l] = (T)rowp[l < len ? l : 0];
    T acc = (T)0;
#pragma unroll
    for (int l = 0; l < FWD_SMALL_K; l++) {
      const T t = fma(v[l], fs[l], acc);
      acc = (l < len) ? t : acc;
    }
    if (p < k) y[first + p] = fs[p] + acc;
    else u[it.sdst + dr] = g[threadIdx.x] - acc;
  }
}

// sum over aligned segments of 2^lg lanes (every lane of a segment gets the sum)
template <typename T>
__device__ inline T seg_sum(T v, int lg) {
  for (int off = (1 << lg) >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// forward gather block of a large front (two-launch form, first launch): one
// thread per position p, f_p = b_p + the children's update entries at p (summed
// in ascending order); fp = f in pivot order (p < k); on a contribution row
// (p >= k) the sum goes to the row's slot in the parent's block, where the rows
// launch reads it and leaves the row's update entry
template <typename T>
__device__ inline void fwd_gather_block(const SolveItem& it, const DevSym& d, const T* __restrict__ bp,
                                        T* __restrict__ fp, T* __restrict__ u) {
  const int k = it.k, m = it.m, first = it.first;
  const int p = it.blk * 256 + threadIdx.x;
  if (p >= m) return;
  const int dr = p >= k ? d.rel[it.roff + p - k] : 0;
  T acc = (T)0;
  int ip = 0;
  if (p < k) {
    acc = bp[first + p];
    ip = d.ilperm[first + p];
  }
  acc = slot_sum(u + it.soff + p, it.nch, m, acc);
  if (p < k) fp[first + ip] = acc;
  else u[it.sdst + dr] = acc;
}

// forward rows of a large front (two-launch form, second launch): products from
// the upper copy (column i, rows l < min(i, k), contiguous), y_i = f_i + G(i,:) f
// (i < k) or u_{i-k} -= G(i,:) f (i >= k), f = fp from the gather launch
template <typename T, typename S>
__device__ inline void fwd_rows_block(const SolveItem& it, const DevSym& d, const S* __restrict__ F,
                                      const T* __restrict__ fp, T* __restrict__ y, T* __restrict__ u) {
  const int lg = it.lg;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int seg = 1 << lg, sl = lane & (seg - 1);
  const int i = (it.blk * 4 + wv) * (64 >> lg) + (lane >> lg);
  const int k = it.k, m = it.m, first = it.first;
  const S* col = F + it.poff + (i < m ? prow_off(i, k) : 0);
  const T* f = fp + first;
  const bool own = sl == 0 && i < m;
  T base = (T)0;
  T* dst = nullptr;  // the contribution row's slot in the parent's block
  if (own && i >= k) dst = u + it.sdst + d.rel[it.roff + i - k];
  if (own) base = (i < k) ? f[i] : *dst;
  const int len = (i < m) ? min(i, k) : 0;
  const T acc = seg_sum(seg_dot<T, S>(col, f, len, sl, seg), lg);
  if (own) {
    if (i < k) y[first + i] = base + acc;
    else *dst = base - acc;
  }
}

// forward, one or two launches per tree level, by item kind:
//   0  a whole small front (fwd_small_front);
//   1  a gather block of a large front (fwd_gather_block), then in the next
//   3  launch its row blocks (fwd_rows_block);
//   2  a row block of a large front that gathers for itself (one launch per
//      level): products from the upper copy as in kind 3,
//        y_i = f_i + G(i,:) f (i < k),   u_i = g_i - G(i,:) f (i >= k),
//      with every such item gathering the front's pivot part f = b + the
//      children's update entries into LDS in pivot order, and g_i (the children's
//      entries on its own contribution rows) for its rows.
// The update entry u_i of a contribution row goes to the supernode's slot for
// that row in its parent's slot block.  Row blocks: a segment of 2^lg lanes per
// row, 64 >> lg rows per wave (item = (s, block of 4 * (64 >> lg) rows, lg)).
// Every sum adds the children's entries in ascending order and every row
// product runs in the seg_dot order, so both forms give bitwise the same y and
// update entries.
template <typename T, typename S>
__device__ inline void fwd_level_body(int bx, const SolveItem* __restrict__ items, const DevSym& d,
                                      const S* __restrict__ F, const T* __restrict__ bp, T* __restrict__ fp,
                                      T* __restrict__ y, T* __restrict__ u) {
  __shared__ __align__(16) T fs[KMAX + 8];  // chunk loads may run 7 past k
  const SolveItem it = items[bx];
  if (it.kind == 0) {
    fwd_small_front<T, S>(it, d, F, bp, y, u);
    return;
  }
  if (it.kind == 1) {
    fwd_gather_block<T>(it, d, bp, fp, u);
    return;
  }
  if (it.kind == 3) {
    fwd_rows_block<T, S>(it, d, F, fp, y, u);
    return;
  }
  // kind 2: nrg = it.pad row groups of rpg = 4 * (64 >> lg) rows each (rows
  // row0 .. row0 + nrg * rpg), one gather for all of them
  const int lg = it.lg;
  const int nrg = it.pad > 0 ? it.pad : 1;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int seg = 1 << lg, sl = lane & (seg - 1);
  const int rpg = 4 * (64 >> lg);
  const int row0 = it.blk * nrg * rpg;
  const int myrow = wv * (64 >> lg) + (lane >> lg);  // row within a group
  const int k = it.k, m = it.m, first = it.first;
  // the first PF chunks of this lane's part of its row in group 0 (the seg_dot
  // order: chunks sl, sl + seg, ..., 16 bytes of S each), loaded before the
  // gather, then GC chunks per loop pass; the chunks are added in ascending
  // order whatever the grouping.  GC = 2 (against 4) keeps the FP32 kernel at
  // 64 registers: the same per-level time on one large system and 7 % less time
  // when a group launch holds many systems' levels (more waves resident).
  constexpr int W = 16 / sizeof(S), GC = 2, PF = 3;
  typedef typename V16<S>::v vt;
  typedef VecW<T, W> vw;
  vt x0[PF];
  auto prefetch = [&](int rg) {
    const int i = row0 + rg * rpg + myrow;
    const S* col = F + it.poff + (i < m ? prow_off(i, k) : 0);
    const int len = (i < m) ? min(i, k) : 0;
    const int nc = (len + W - 1) / W;
#pragma unroll
    for (int q = 0; q < PF; q++) {
      const int c = sl + q * seg;
      x0[q] = ldv(col + (c < nc ? c : 0) * W);
    }
  };
  prefetch(0);
  // gather: f in pivot order (positions p < k) into fs, and the children's
  // entries of the item's contribution rows (k <= r < m) into gb, from the
  // supernode's slot block
  __shared__ T gb[FWD_GROUP_ROWS];
  {
    const T* sb = u + it.soff;
    // four positions per thread and pass, their right-hand side, permutation and
    // slot loads in flight together
    for (int q0 = threadIdx.x; q0 < k; q0 += 1024) {
      T a[4];
      int ip[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int q = q0 + 256 * j;
        a[j] = q < k ? bp[first + q] : (T)0;
        ip[j] = q < k ? d.ilperm[first + q] : 0;
      }
      slot_sum4(sb, it.nch, m, q0, k, a);
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (q0 + 256 * j < k) fs[ip[j]] = a[j];
    }
    const int rlo = max(row0, k), rhi = min(row0 + nrg * rpg, m);
    for (int q0 = rlo + threadIdx.x; q0 < rhi; q0 += 1024) {
      T a[4] = {(T)0, (T)0, (T)0, (T)0};
      slot_sum4(sb, it.nch, m, q0, rhi, a);
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (q0 + 256 * j < rhi) gb[q0 + 256 * j - row0] = a[j];
    }
  }
  __syncthreads();
  for (int rg = 0; rg < nrg; rg++) {
    const int i = row0 + rg * rpg + myrow;
    const bool own = sl == 0 && i < m;
    const S* col = F + it.poff + (i < m ? prow_off(i, k) : 0);
    const int len = (i < m) ? min(i, k) : 0;
    const int nc = (len + W - 1) / W;
    // a contribution row's slot in the parent's block (loaded before the product)
    const int dr = (own && i >= k) ? d.rel[it.roff + i - k] : 0;
    vt xc[PF];
#pragma unroll
    for (int q = 0; q < PF; q++) xc[q] = x0[q];
    if (rg + 1 < nrg) prefetch(rg + 1);  // the next group's first chunks in flight
    // seg_dot(col, fs, len) with its first PF chunks from xc (the same products
    // in the same order; a chunk past the row adds nothing)
    T acc = (T)0;
#pragma unroll
    for (int q = 0; q < PF; q++) {
      const int c = sl + q * seg;
      const vw yv = ldw<T, W>(fs + (c < nc ? c : 0) * W);
#pragma unroll
      for (int w = 0; w < W; w++) {
        const T t = fma((T)xc[q][w], yv.e[w], acc);
        acc = (c * W + w < len) ? t : acc;
      }
    }
    for (int q0 = sl + PF * seg; q0 < nc; q0 += GC * seg) {
      vt x[GC];
      vw yv[GC];
#pragma unroll
      for (int q = 0; q < GC; q++) {
        const int c = q0 + q * seg, cc = c < nc ? c : 0;
        x[q] = ldv(col + cc * W);
        yv[q] = ldw<T, W>(fs + cc * W);
      }
#pragma unroll
      for (int q = 0; q < GC; q++)
#pragma unroll
        for (int w = 0; w < W; w++) {
          const T t = fma((T)x[q][w], yv[q].e[w], acc);
          acc = ((q0 + q * seg) * W + w < len) ? t : acc;
        }
    }
    acc = seg_sum(acc, lg);
    if (own) {
      if (i < k) y[first + i] = fs[i] + acc;
      else u[it.sdst + dr] = gb[i - row0] - acc;
    }
  }
}
template <typename T, typename S>
__global__ __launch_bounds__(256) void k_fwd_level(const SolveItem* __restrict__ items, DevSym d,
                                                   const S* __restrict__ F,
                                                   const T* __restrict__ bp, T* __restrict__ fp,
                                                   T* __restrict__ y, T* __restrict__ u) {
  fwd_level_body<T, S>(blockIdx.x, items, d, F, bp, fp, y, u);
}

// packed solve arrays from the solve form of the fronts: front column i gives
// forward row i (its rows 0 .. min(i, k), the upper copy) and, for i < k,
// backward column i (its rows i+1 .. m, the lower copy G); one wave per column,
// item = (supernode, block of 4 columns)
template <typename T>
__global__ __launch_bounds__(256) void k_pack_solve(const int2* __restrict__ items, DevSym d,
                                                    const T* __restrict__ F, T* __restrict__ PF,
                                                    T* __restrict__ PB) {
  const int2 it = items[blockIdx.x];
  const int s = it.x, i = it.y * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int k = d.sn_k[s], m = d.sn_m[s];
  if (i >= m) return;
  const T* colp = F + d.sn_foff[s] + (size_t)i * m;
  const long long po = d.sn_poff[s];
  const int len = min(i, k);
  T* fr = PF + po + prow_off(i, k);
  for (int t = lane; t < len; t += 64) fr[t] = colp[t];
  if (i < k) {
    T* bc = PB + po + pcol_off(i, m);
    for (int t = lane; t < m - 1 - i; t += 64) bc[t] = colp[i + 1 + t];
  }
}

// half-precision copy of the FP32 packed solve arrays (the iterative solve's
// preconditioner reads it); ctr->half_overflow is set when an entry is too large
// for half precision, so the copy is not used for that factor
__global__ void k_pack_half(long long n, const float* __restrict__ in, _Float16* __restrict__ out,
                            FactorCounters* __restrict__ ctr) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  const float v = in[e];
  out[e] = (_Float16)v;
  if (!(fabsf(v) < 65504.0f)) ctr->half_overflow = 1;
}

// z = D^-1 y over all pivots: z_g = E0_g y_g + E1_g y_partner, stored at the
// pivot's entry of the backward vector
template <typename T>
__device__ inline void dsolve_body(int bx, int n, const T* __restrict__ E0, const T* __restrict__ E1,
                                   const signed char* __restrict__ piv, const int* __restrict__ wpos,
                                   const T* __restrict__ y, T* __restrict__ wv) {
  const int g = bx * 256 + threadIdx.x;
  if (g >= n) return;
  const int pc = piv[g];
  const int pg = (pc == 2) ? g + 1 : (pc == 0 ? g - 1 : g);
  wv[wpos[g]] = E0[g] * y[g] + E1[g] * y[pg];
}
template <typename T>
__global__ __launch_bounds__(256) void k_dsolve(int n, const T* __restrict__ E0,
                                                const T* __restrict__ E1,
                                                const signed char* __restrict__ piv,
                                                const int* __restrict__ wpos,
                                                const T* __restrict__ y, T* __restrict__ wv) {
  dsolve_body<T>(blockIdx.x, n, E0, E1, piv, wpos, y, wv);
}

// backward from the lower copy (column c, rows below the diagonal, contiguous):
// x_c = z_c + sum_{c<i<k} X(i,c) z_i - sum_j Z(j,c) x[rows_j] = z_c + G(:,c) . w[c+1..m)
// with w the supernode's backward vector [z; -x[rows]].  The -x entries were put
// there by the ancestors' backward launches: every computed x_g is also written,
// negated, to all entries that name g (the push map), so a column is one
// contiguous dot product, like a forward row.
// A segment of 2^lg lanes per column (lg <= 8: a long column takes two or four
// waves); item = (s, block of 256 >> lg columns, lg)
template <typename T, typename S>
__device__ inline void bwd_cols_body(int bx, const SolveItem* __restrict__ items, const DevSym& d,
                                     const S* __restrict__ F, T* wv, T* __restrict__ xp) {
  __shared__ T part[4];
  const SolveItem it = items[bx];
  const int lg = it.lg;  // 2^lg lanes per column, up to the whole workgroup (lg = 8)
  const int seg = 1 << lg, sl = threadIdx.x & (seg - 1);
  const int c = it.blk * (256 >> lg) + (threadIdx.x >> lg);
  const int k = it.k, m = it.m, first = it.first;
  const bool ok = c < k;
  // every lane of the segment loads these (the push below uses g and x); they are
  // used only after the product, so their loads stay in flight meanwhile
  const int cq = ok ? c : 0;
  const int lp = d.lperm[first + cq];
  const T* w = wv + it.voff;
  const T zc = w[cq];
  const S* col = F + it.poff + pcol_off(cq, m);
  // every lane of a segment ends with the same sum (xor butterfly; the wave sums
  // of a column spanning 2 or 4 waves are combined in order)
  T acc = seg_sum(seg_dot<T, S>(col, w + cq + 1, ok ? m - c - 1 : 0, sl, seg), lg < 6 ? lg : 6);
  if (lg > 6) {
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    const int w0 = (threadIdx.x >> lg) << (lg - 6);  // first wave of the column
    acc = part[w0];
    for (int q = 1; q < (seg >> 6); q++) acc += part[w0 + q];
  }
  if (ok) {
    const int g = first + lp;
    const T x = zc + acc;
    if (sl == 0) xp[g] = x;
    const int e1 = d.push_ptr[g + 1];
    for (int e = d.push_ptr[g] + sl; e < e1; e += seg) wv[d.push_pos[e]] = -x;
  }
}
template <typename T, typename S>
__global__ __launch_bounds__(256) void k_bwd_cols(const SolveItem* __restrict__ items, DevSym d,
                                                  const S* __restrict__ F, T* wv,
                                                  T* __restrict__ xp) {
  bwd_cols_body<T, S>(blockIdx.x, items, d, F, wv, xp);
}

// ------------------------------------------------------------------------
// Multi-RHS triangular solves (LCP block solves, KKTSolver.java:1537-1588):
// the same work items, launches and per-lane accumulation order as the
// single-RHS kernels above, for nr <= SOLVE_NRHS right-hand sides at once, so
// a front's factor data is read once per pass and each result equals the
// single-RHS one.  Vector j of a set starts at j * ld (ldn: n-vectors, ldu:
// update vectors, ldw: backward vectors; each with the spare entries the chunk
// loads may run into).
// ------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_gather_m(const SolveItem* __restrict__ items, DevSym d,
                                                      const T* __restrict__ F, int nr,
                                                      const T* __restrict__ bp, T* __restrict__ fp,
                                                      T* __restrict__ y, T* __restrict__ u, long long ldn,
                                                      long long ldu) {
  __shared__ T g[SOLVE_NRHS][256];
  __shared__ __align__(16) T fs[SOLVE_NRHS][FWD_SMALL_K + 4];
  const SolveItem it = items[blockIdx.x];
  const int k = it.k, m = it.m, first = it.first;
  const int p = it.blk * 256 + threadIdx.x;
  T acc[SOLVE_NRHS];
#pragma unroll
  for (int j = 0; j < SOLVE_NRHS; j++) acc[j] = (T)0;
  int ip = 0;
  if (p < m) {
    const int* gp = d.gptr + it.goff + p;
    const int e0 = gp[0], e1 = gp[1];
    if (p < k) {
#pragma unroll
      for (int j = 0; j < SOLVE_NRHS; j++)
        if (j < nr) acc[j] = bp[j * ldn + first + p];
      ip = d.ilperm[first + p];
    }
    // children in ascending order (the single-RHS order), the index shared by
    // every right-hand side
    for (int e = e0; e < e1; e++) {
      const int src = d.gsrc[e];
#pragma unroll
      for (int j = 0; j < SOLVE_NRHS; j++)
        if (j < nr) acc[j] += u[j * ldu + src];
    }
  }
  if (it.kind == 1) {
#pragma unroll
    for (int j = 0; j < SOLVE_NRHS; j++)
      if (j < nr) {
        if (p < k) fp[j * ldn + first + ip] = acc[j];
        else if (p < m) u[j * ldu + it.uoff + p - k] = acc[j];
      }
    return;
  }
#pragma unroll
  for (int j = 0; j < SOLVE_NRHS; j++) g[j][threadIdx.x] = acc[j];
  __syncthreads();
  if (p < k) {
#pragma unroll
    for (int j = 0; j < SOLVE_NRHS; j++) fs[j][ip] = acc[j];
  }
  __syncthreads();
  if (p < m) {
    const int len = min(p, k);
    const T* rowp = F + it.foff + p;
    T v[FWD_SMALL_K];
#pragma unroll
    for (int l = 0; l < FWD_SMALL_K; l++) v[l] = rowp[(size_t)(l < len ? l : 0) * m];
#pragma unroll
    for (int j = 0; j < SOLVE_NRHS; j++) {
      if (j >= nr) break;
      T a = (T)0;
#pragma unroll
      for (int l = 0; l < FWD_SMALL_K; l++) {
        const T t = fma(v[l], fs[j][l], a);
        a = (l < len) ? t : a;
      }
      if (p < k) y[j * ldn + first + p] = fs[j][p] + a;
      else u[j * ldu + it.uoff + p - k] = g[j][threadIdx.x] - a;
    }
  }
}

// Lane sl's partial sums sum_{e < len} a[e] b_j[e] for nr vectors b_j = b + j ld,
// over the lane's chunks in the order of seg_dot (groups of G chunks, G * the
// number of vectors <= 16 chunk registers; the loads of a group issue before its
// products)
template <typename T>
__device__ inline void seg_dot_m(const T* __restrict__ a, const T* __restrict__ b, long long ld, int nr,
                                 int len, int sl, int seg, T* acc) {
  constexpr int W = 16 / sizeof(T), G = (SOLVE_NRHS >= 8) ? 1 : 16 / (2 * SOLVE_NRHS);
  typedef typename V16<T>::v vt;
  const int nc = (len + W - 1) / W;
#pragma unroll
  for (int j = 0; j < SOLVE_NRHS; j++) acc[j] = (T)0;
  for (int q0 = sl; q0 < nc; q0 += G * seg) {
    vt x[G], y[G][SOLVE_NRHS];
#pragma unroll
    for (int u = 0; u < G; u++) {
      const int q = q0 + u * seg, qc = q < nc ? q : 0;
      x[u] = ldv(a + qc * W);
#pragma unroll
      for (int j = 0; j < SOLVE_NRHS; j++) y[u][j] = ldv(b + (j < nr ? j : 0) * ld + qc * W);
    }
#pragma unroll
    for (int u = 0; u < G; u++)
#pragma unroll
      for (int j = 0; j < SOLVE_NRHS; j++)
#pragma unroll
        for (int w = 0; w < W; w++) {
          const T t = fma(x[u][w], y[u][j][w], acc[j]);
          acc[j] = ((q0 + u * seg) * W + w < len) ? t : acc[j];
        }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_fwd_rows_m(const SolveItem* __restrict__ items, DevSym d,
                                                    const T* __restrict__ F, int nr,
                                                    const T* __restrict__ fp, T* __restrict__ y,
                                                    T* __restrict__ u, long long ldn, long long ldu) {
  const SolveItem it = items[blockIdx.x];
  const int lg = it.lg;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int seg = 1 << lg, sl = lane & (seg - 1);
  const int i = (it.blk * 4 + wv) * (64 >> lg) + (lane >> lg);
  const int k = it.k, m = it.m, first = it.first;
  const T* col = F + it.foff + (size_t)(i < m ? i : 0) * m;
  const bool own = sl == 0 && i < m;
  T base[SOLVE_NRHS];
#pragma unroll
  for (int j = 0; j < SOLVE_NRHS; j++) {
    base[j] = (T)0;
    if (own && j < nr) base[j] = (i < k) ? fp[j * ldn + first + i] : u[j * ldu + it.uoff + i - k];
  }
  const int len = (i < m) ? min(i, k) : 0;
  T acc[SOLVE_NRHS];
  seg_dot_m<T>(col, fp + first, ldn, nr, len, sl, seg, acc);
#pragma unroll
  for (int j = 0; j < SOLVE_NRHS; j++) {
    const T a = seg_sum(acc[j], lg);
    if (own && j < nr) {
      if (i < k) y[j * ldn + first + i] = base[j] + a;
      else u[j * ldu + it.uoff + i - k] = base[j] - a;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_dsolve_m(int n, int nr, const T* __restrict__ E0,
                                                  const T* __restrict__ E1,
                                                  const signed char* __restrict__ piv,
                                                  const int* __restrict__ wpos, const T* __restrict__ y,
                                                  T* __restrict__ wv, long long ldn, long long ldw) {
  const int g = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  if (g >= n || j >= nr) return;
  const int pc = piv[g];
  const int pg = (pc == 2) ? g + 1 : (pc == 0 ? g - 1 : g);
  wv[j * ldw + wpos[g]] = E0[g] * y[j * ldn + g] + E1[g] * y[j * ldn + pg];
}

template <typename T>
__global__ __launch_bounds__(256) void k_bwd_cols_m(const SolveItem* __restrict__ items, DevSym d,
                                                    const T* __restrict__ F, int nr, T* wv,
                                                    T* __restrict__ xp, long long ldn, long long ldw) {
  __shared__ T part[SOLVE_NRHS][4];
  const SolveItem it = items[blockIdx.x];
  const int lg = it.lg;
  const int seg = 1 << lg, sl = threadIdx.x & (seg - 1);
  const int c = it.blk * (256 >> lg) + (threadIdx.x >> lg);
  const int k = it.k, m = it.m, first = it.first;
  const bool ok = c < k;
  const int cq = ok ? c : 0;
  const int lp = d.lperm[first + cq];
  const T* w = wv + it.voff;
  T zc[SOLVE_NRHS];
#pragma unroll
  for (int j = 0; j < SOLVE_NRHS; j++) zc[j] = w[(j < nr ? j : 0) * ldw + cq];
  const T* col = F + it.foff + (size_t)cq * m + cq + 1;
  T acc[SOLVE_NRHS];
  seg_dot_m<T>(col, w + cq + 1, ldw, nr, ok ? m - c - 1 : 0, sl, seg, acc);
#pragma unroll
  for (int j = 0; j < SOLVE_NRHS; j++) acc[j] = seg_sum(acc[j], lg < 6 ? lg : 6);
  if (lg > 6) {
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
      for (int j = 0; j < SOLVE_NRHS; j++) part[j][threadIdx.x >> 6] = acc[j];
    }
    __syncthreads();
    const int w0 = (threadIdx.x >> lg) << (lg - 6);
#pragma unroll
    for (int j = 0; j < SOLVE_NRHS; j++) {
      acc[j] = part[j][w0];
      for (int q = 1; q < (seg >> 6); q++) acc[j] += part[j][w0 + q];
    }
  }
  if (ok) {
    const int g = first + lp;
    const int e0 = d.push_ptr[g], e1 = d.push_ptr[g + 1];
#pragma unroll
    for (int j = 0; j < SOLVE_NRHS; j++) {
      if (j >= nr) break;
      const T x = zc[j] + acc[j];
      if (sl == 0) xp[j * ldn + g] = x;
      for (int e = e0 + sl; e < e1; e += seg) wv[j * ldw + d.push_pos[e]] = -x;
    }
  }
}

template <typename TI, typename TO>
__global__ void k_gather_perm_m(int n, int nr, const int* __restrict__ perm, const TI* __restrict__ in,
                                long long ldi, TO* __restrict__ out, long long ldo) {
  const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  if (i < n && j < nr) out[j * ldo + i] = (TO)in[j * ldi + perm[i]];
}
template <typename TI, typename TO>
__global__ void k_scatter_perm_m(int n, int nr, const int* __restrict__ perm, const TI* __restrict__ in,
                                 long long ldi, TO* __restrict__ out, long long ldo) {
  const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  if (i < n && j < nr) out[j * ldo + perm[i]] = (TO)in[j * ldi + i];
}

// ------------------------------------------------------------------------
// Device-side value assembly (CuDssSolver.*DeviceValues, cudssBridgeKernels.cu:57-380).
// Each call computes its contributions into a scratch array, then one thread
// per distinct CSR slot adds them in a fixed order (the plan lists the
// contributions of every slot, ascending by contribution index): no atomics,
// bitwise reproducible.
// ------------------------------------------------------------------------
__global__ void k_contrib_block3(int nblocks, const double* __restrict__ blockVals,
                                 const double* __restrict__ blockScales, double* __restrict__ c) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e < 9 * nblocks) c[e] = blockScales[e / 9] * blockVals[e];
}
__global__ void k_contrib_diag3(int nblocks, const double* __restrict__ masses, double* __restrict__ c) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e < 3 * nblocks) c[e] = masses[e / 3];
}
// K = dv (B_i^T D B_j + (g_i^T sigma g_j) I), with B the 6x3 Voigt strain-
// displacement matrix of gradient g (strain order xx, yy, zz, xy, yz, xz,
// engineering shears), D 6x6 row-major, sigma = [s00, s11, s22, s01, s12, s02]
// (FemUtilities.addMaterialStiffness); K row-major 3x3
__device__ void matstiff3(const double* gi, const double* gj, const double* D, const double* sg,
                          double dv, double* K) {
  // Voigt rows of B(g) for displacement component k: bv(g, k)[r]
  auto bv = [](const double* g, int k, double* out) {
    for (int r = 0; r < 6; r++) out[r] = 0.0;
    out[k] = g[k];
    if (k == 0) { out[3] = g[1]; out[5] = g[2]; }
    else if (k == 1) { out[3] = g[0]; out[4] = g[2]; }
    else { out[4] = g[1]; out[5] = g[0]; }
  };
  double bi[3][6], bj[3][6];
  for (int k = 0; k < 3; k++) {
    bv(gi, k, bi[k]);
    bv(gj, k, bj[k]);
  }
  const double S[3][3] = {{sg[0], sg[3], sg[5]}, {sg[3], sg[1], sg[4]}, {sg[5], sg[4], sg[2]}};
  double geo = 0.0;
  for (int a = 0; a < 3; a++)
    for (int q = 0; q < 3; q++) geo += gi[a] * S[a][q] * gj[q];
  for (int r = 0; r < 3; r++)
    for (int q = 0; q < 3; q++) {
      double acc = 0.0;
      for (int u = 0; u < 6; u++) {
        double db = 0.0;
        for (int v = 0; v < 6; v++) db += D[6 * u + v] * bj[q][v];
        acc += bi[r][u] * db;
      }
      K[3 * r + q] = dv * (acc + (r == q ? geo : 0.0));
    }
}
__global__ void k_contrib_matstiff3(int nblocks, const double* __restrict__ gis,
                                    const double* __restrict__ gjs, const double* __restrict__ Ds,
                                    const double* __restrict__ sigmas,
                                    const double* __restrict__ dvs, double* __restrict__ c) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= nblocks) return;
  const double gi[3] = {gis[3 * b], gis[3 * b + 1], gis[3 * b + 2]};
  const double gj[3] = {gjs[3 * b], gjs[3 * b + 1], gjs[3 * b + 2]};
  matstiff3(gi, gj, Ds + 36 * (size_t)b, sigmas + 6 * (size_t)b, dvs[b], c + 9 * (size_t)b);
}

// ---- element-batched forms (cudssBridgeKernels.cu:336-739) ----
// One wave per element, one lane per node pair (looping); the integration
// points of a pair are summed in order, then the pair's 3x3 block K is reduced
// to its target block T_i K T_j^T (T_i row-major dim x 3 in an 18-double slot:
// I for a free node, the 6x3 attachment block of a frame-attached one) and
// stored at c[36 p + r * coldim + q].  Contributions are written only where the
// slot list has entries, which the ordered scatter then adds.
__device__ void reduce3_store(const double* K, const double* Ti, int rdim, const double* Tj,
                              int cdim, double* c) {
  for (int r = 0; r < rdim; r++) {
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;  // (T_i K)(r, :)
    for (int k = 0; k < 3; k++) {
      t0 += Ti[3 * r + k] * K[3 * k];
      t1 += Ti[3 * r + k] * K[3 * k + 1];
      t2 += Ti[3 * r + k] * K[3 * k + 2];
    }
    for (int q = 0; q < cdim; q++) c[r * cdim + q] = t0 * Tj[3 * q] + t1 * Tj[3 * q + 1] + t2 * Tj[3 * q + 2];
  }
}
// D of isotropic linear elasticity from (E, nu), Voigt order with engineering
// shears (cudssBridgeKernels.cu:451-468)
__device__ void linear_elastic_D(double E, double nu, double* D) {
  for (int q = 0; q < 36; q++) D[q] = 0.0;
  const double a = E / (1.0 + nu);
  const double dia = (1.0 - nu) / (1.0 - 2.0 * nu) * a, off = nu / (1.0 - 2.0 * nu) * a;
  for (int r = 0; r < 3; r++)
    for (int q = 0; q < 3; q++) D[6 * r + q] = (r == q) ? dia : off;
  D[21] = D[28] = D[35] = 0.5 * a;
}
__global__ __launch_bounds__(64) void k_contrib_matstiff3_elem(
    const int* __restrict__ nodeCnt, const int* __restrict__ nodeOff, const int* __restrict__ pairOff,
    const int* __restrict__ ipOff, const int* __restrict__ gradOff, const int* __restrict__ pairNodes,
    const int* __restrict__ nodeDims, const double* __restrict__ T, const double* __restrict__ grads,
    const double* __restrict__ Ds, const double* __restrict__ sigmas, const double* __restrict__ dvs,
    double* __restrict__ c) {
  const int e = blockIdx.x;
  const int nn = nodeCnt[e], n0 = nodeOff[e], p0 = pairOff[e], np = pairOff[e + 1] - p0;
  const int i0 = ipOff[e], ni = ipOff[e + 1] - i0, g0 = gradOff[e];
  for (int pr = threadIdx.x; pr < np; pr += 64) {
    const int pidx = p0 + pr, i = pairNodes[2 * pidx], j = pairNodes[2 * pidx + 1];
    double K[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Kq[9];
    for (int ip = 0; ip < ni; ip++) {
      const double* gb = grads + 3 * (size_t)(g0 + ip * nn);
      matstiff3(gb + 3 * i, gb + 3 * j, Ds + 36 * (size_t)(i0 + ip), sigmas + 6 * (size_t)(i0 + ip),
                dvs[i0 + ip], Kq);
      for (int q = 0; q < 9; q++) K[q] += Kq[q];
    }
    reduce3_store(K, T + 18 * (size_t)(n0 + i), nodeDims[n0 + i], T + 18 * (size_t)(n0 + j),
                  nodeDims[n0 + j], c + 36 * (size_t)pidx);
  }
}
__global__ __launch_bounds__(64) void k_contrib_linelastic3_elem(
    const int* __restrict__ nodeCnt, const int* __restrict__ pairOff, const int* __restrict__ ipOff,
    const int* __restrict__ gradOff, const int* __restrict__ pairNodes,
    const double* __restrict__ params, const double* __restrict__ grads,
    const double* __restrict__ dvs, double* __restrict__ c) {
  const int e = blockIdx.x;
  const int nn = nodeCnt[e], p0 = pairOff[e], np = pairOff[e + 1] - p0;
  const int i0 = ipOff[e], ni = ipOff[e + 1] - i0, g0 = gradOff[e];
  double D[36];
  linear_elastic_D(params[2 * e], params[2 * e + 1], D);
  const double sg[6] = {0, 0, 0, 0, 0, 0};
  for (int pr = threadIdx.x; pr < np; pr += 64) {
    const int pidx = p0 + pr, i = pairNodes[2 * pidx], j = pairNodes[2 * pidx + 1];
    double K[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Kq[9];
    for (int ip = 0; ip < ni; ip++) {
      const double* gb = grads + 3 * (size_t)(g0 + ip * nn);
      matstiff3(gb + 3 * i, gb + 3 * j, D, sg, dvs[i0 + ip], Kq);
      for (int q = 0; q < 9; q++) K[q] += Kq[q];
    }
    for (int q = 0; q < 9; q++) c[9 * (size_t)pidx + q] = K[q];
  }
}
// spatial gradients from the Jacobian J = sum_n x_n (dN_n/dxi)^T of each
// integration point: g = J^-T dN/dxi, dv = det(J) w (cudssBridgeKernels.cu:531-640)
__global__ __launch_bounds__(64) void k_contrib_linelastic3_geom_elem(
    const int* __restrict__ nodeCnt, const int* __restrict__ nodeOff, const int* __restrict__ pairOff,
    const int* __restrict__ ipOff, const int* __restrict__ ngOff, const int* __restrict__ pairNodes,
    const int* __restrict__ nodeDims, const double* __restrict__ T, const double* __restrict__ params,
    const double* __restrict__ pos, const double* __restrict__ ngrads, const double* __restrict__ w,
    double* __restrict__ c) {
  const int e = blockIdx.x;
  const int nn = nodeCnt[e], n0 = nodeOff[e], p0 = pairOff[e], np = pairOff[e + 1] - p0;
  const int i0 = ipOff[e], ni = ipOff[e + 1] - i0, g0 = ngOff[e];
  double D[36];
  linear_elastic_D(params[2 * e], params[2 * e + 1], D);
  const double sg[6] = {0, 0, 0, 0, 0, 0};
  for (int pr = threadIdx.x; pr < np; pr += 64) {
    const int pidx = p0 + pr, i = pairNodes[2 * pidx], j = pairNodes[2 * pidx + 1];
    double K[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Kq[9];
    for (int ip = 0; ip < ni; ip++) {
      const double* gb = ngrads + 3 * (size_t)(g0 + ip * nn);
      double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int n = 0; n < nn; n++) {
        const double* x = pos + 3 * (size_t)(n0 + n);
        for (int r = 0; r < 3; r++)
          for (int q = 0; q < 3; q++) J[3 * r + q] += x[r] * gb[3 * n + q];
      }
      // inverse by cofactors
      const double c00 = J[4] * J[8] - J[5] * J[7], c01 = J[5] * J[6] - J[3] * J[8],
                   c02 = J[3] * J[7] - J[4] * J[6];
      const double det = J[0] * c00 + J[1] * c01 + J[2] * c02, id = 1.0 / det;
      const double Ji[9] = {c00 * id, (J[2] * J[7] - J[1] * J[8]) * id, (J[1] * J[5] - J[2] * J[4]) * id,
                            c01 * id, (J[0] * J[8] - J[2] * J[6]) * id, (J[2] * J[3] - J[0] * J[5]) * id,
                            c02 * id, (J[1] * J[6] - J[0] * J[7]) * id, (J[0] * J[4] - J[1] * J[3]) * id};
      double gi[3], gj[3];
      for (int q = 0; q < 3; q++) {
        gi[q] = Ji[q] * gb[3 * i] + Ji[3 + q] * gb[3 * i + 1] + Ji[6 + q] * gb[3 * i + 2];
        gj[q] = Ji[q] * gb[3 * j] + Ji[3 + q] * gb[3 * j + 1] + Ji[6 + q] * gb[3 * j + 2];
      }
      matstiff3(gi, gj, D, sg, det * w[i0 + ip], Kq);
      for (int q = 0; q < 9; q++) K[q] += Kq[q];
    }
    reduce3_store(K, T + 18 * (size_t)(n0 + i), nodeDims[n0 + i], T + 18 * (size_t)(n0 + j),
                  nodeDims[n0 + j], c + 36 * (size_t)pidx);
  }
}
// K_ij = sum_{a,b} c_{i,a} Rinv(a,b) c_{j,b}^T over the element's np pressure
// degrees of freedom (c_{i,a}: 3-vector); elements with np <= 0 contribute zeros
// (cudssBridgeKernels.cu:658-717)
__global__ __launch_bounds__(64) void k_contrib_dilational3_elem(
    const int* __restrict__ nodeOff, const int* __restrict__ presCnt, const int* __restrict__ pairOff,
    const int* __restrict__ consOff, const int* __restrict__ rinvOff, const int* __restrict__ pairNodes,
    const int* __restrict__ nodeDims, const double* __restrict__ T, const double* __restrict__ cons,
    const double* __restrict__ rinv, double* __restrict__ c) {
  const int e = blockIdx.x;
  const int npr = presCnt[e], n0 = nodeOff[e], p0 = pairOff[e], np = pairOff[e + 1] - p0;
  const int c0 = consOff[e], r0 = rinvOff[e];
  for (int pr = threadIdx.x; pr < np; pr += 64) {
    const int pidx = p0 + pr, i = pairNodes[2 * pidx], j = pairNodes[2 * pidx + 1];
    double K[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < npr; a++) {
      const double* ci = cons + 3 * (size_t)(c0 + i * npr + a);
      for (int b = 0; b < npr; b++) {
        const double rv = rinv[r0 + a * npr + b];
        const double* cj = cons + 3 * (size_t)(c0 + j * npr + b);
        for (int r = 0; r < 3; r++)
          for (int q = 0; q < 3; q++) K[3 * r + q] += ci[r] * rv * cj[q];
      }
    }
    reduce3_store(K, T + 18 * (size_t)(n0 + i), nodeDims[n0 + i], T + 18 * (size_t)(n0 + j),
                  nodeDims[n0 + j], c + 36 * (size_t)pidx);
  }
}
__global__ void k_seg_scatter(int nuniq, const int* __restrict__ useg, const int* __restrict__ uslot,
                              const int* __restrict__ order, const double* __restrict__ c,
                              double scale, double* __restrict__ vals) {
  const int u = blockIdx.x * 256 + threadIdx.x;
  if (u >= nuniq) return;
  double acc = 0.0;
  for (int e = useg[u]; e < useg[u + 1]; e++) acc += c[order[e]];
  vals[uslot[u]] += scale * acc;
}

// ------------------------------------------------------------------------
// vectors and SpMV
// ------------------------------------------------------------------------
// Kernel bodies take their block index bx as an argument: the kernels below
// pass blockIdx.x, and the group kernels (many handles' launches of one step in
// one launch, further down) pass the block's index within its handle's share, so
// a handle's blocks compute the same values either way.
template <typename TI, typename TO>
__device__ inline void gather_perm_body(int bx, int n, const int* __restrict__ perm, const TI* __restrict__ in,
                                        TO* __restrict__ out) {
  const int j = bx * 256 + threadIdx.x;
  if (j < n) out[j] = (TO)in[perm[j]];
}
template <typename TI, typename TO>
__global__ void k_gather_perm(int n, const int* __restrict__ perm, const TI* __restrict__ in,
                              TO* __restrict__ out) {
  gather_perm_body<TI, TO>(blockIdx.x, n, perm, in, out);
}
template <typename TI, typename TO>
__device__ inline void scatter_perm_body(int bx, int n, const int* __restrict__ perm, const TI* __restrict__ in,
                                         TO* __restrict__ out) {
  const int j = bx * 256 + threadIdx.x;
  if (j < n) out[perm[j]] = (TO)in[j];
}
template <typename TI, typename TO>
__global__ void k_scatter_perm(int n, const int* __restrict__ perm, const TI* __restrict__ in,
                               TO* __restrict__ out) {
  scatter_perm_body<TI, TO>(blockIdx.x, n, perm, in, out);
}
// the gather of a BiCGStab preconditioner application with the search-direction
// update of k_bicg_p fused in: for i = perm[j], p_i = r_i + beta (p_i - omega v_i),
// written back to p and, in the front precision, to out[j]
template <typename TO>
__device__ inline void gather_bicg_p_body(int bx, int n, const int* __restrict__ perm,
                                          const double* __restrict__ sc, const double* __restrict__ r,
                                          const double* __restrict__ v, double* __restrict__ p,
                                          TO* __restrict__ out) {
  const int j = bx * 256 + threadIdx.x;
  if (j >= n) return;
  const int i = perm[j];
  const double beta = (sc[0] / sc[1]) * (sc[2] / sc[3]);
  const double pi = r[i] + beta * (p[i] - sc[3] * v[i]);
  p[i] = pi;
  out[j] = (TO)pi;
}
template <typename TO>
__global__ void k_gather_bicg_p(int n, const int* __restrict__ perm, const double* __restrict__ sc,
                                const double* __restrict__ r, const double* __restrict__ v,
                                double* __restrict__ p, TO* __restrict__ out) {
  gather_bicg_p_body<TO>(blockIdx.x, n, perm, sc, r, v, p, out);
}

__device__ inline void expand_vals_body(int bx, long long nnzf, const int* __restrict__ map,
                                        const double* __restrict__ vals, double* __restrict__ valsf) {
  const long long e = (long long)bx * 256 + threadIdx.x;
  if (e < nnzf) valsf[e] = vals[map[e]];
}
__global__ void k_expand_vals(long long nnzf, const int* __restrict__ map,
                              const double* __restrict__ vals, double* __restrict__ valsf) {
  expand_vals_body(blockIdx.x, nnzf, map, vals, valsf);
}

// one wave per row
// y = A x (mode 0) or b - A x (mode 1): SPMV_LANES lanes per row, SPMV_ROWS rows
// per workgroup.  Lane sl takes the row's entries sl, sl + 16, sl + 32, ...
// (its q-th at rowptr[row] + sl + 16 q), eight per round: every column index and
// value load of a round issues first, then the round's eight x loads, so a row of
// up to 128 entries costs two dependent round trips.  Entries with q even are
// added into one accumulator and q odd into another, then the two, then the
// lanes of the row (xor butterfly 8, 4, 2, 1); k_spmv_resid_m adds in this same
// order.  Optional fused dot products of the result (BiCGStab): part1[block] =
// sum of a1_i y_i, part2[block] = sum of y_i^2 over the block's rows, in a fixed
// order.
constexpr int SPMV_LANES = 16;
constexpr int SPMV_ROWS = 256 / SPMV_LANES;
__device__ inline void spmv_body(int bx, int n, const int* __restrict__ rowptr, const int* __restrict__ col,
                                 const double* __restrict__ val, const double* __restrict__ x,
                                 const double* __restrict__ b, double* __restrict__ y, int mode,
                                 const double* __restrict__ a1, double* __restrict__ part1,
                                 double* __restrict__ part2) {
  __shared__ double sm[2][4];
  const int row = bx * SPMV_ROWS + threadIdx.x / SPMV_LANES;
  const int sl = threadIdx.x & (SPMV_LANES - 1);
  double acc = 0.0;
  if (row < n) {
    const int e1 = rowptr[row + 1];
    double a0 = 0.0, b0 = 0.0;
    for (int e = rowptr[row] + sl; e < e1; e += 8 * SPMV_LANES) {
      int c[8];
      double v[8], xv[8];
      bool ok[8];
#pragma unroll
      for (int q = 0; q < 8; q++) {
        ok[q] = e + q * SPMV_LANES < e1;
        c[q] = 0;
        v[q] = 0.0;
        if (ok[q]) {
          c[q] = col[e + q * SPMV_LANES];
          v[q] = val[e + q * SPMV_LANES];
        }
      }
#pragma unroll
      for (int q = 0; q < 8; q++) xv[q] = ok[q] ? x[c[q]] : 0.0;
#pragma unroll
      for (int q = 0; q < 8; q += 2) {
        a0 = ok[q] ? fma(v[q], xv[q], a0) : a0;
        b0 = ok[q + 1] ? fma(v[q + 1], xv[q + 1], b0) : b0;
      }
    }
    acc = a0 + b0;
  }
  for (int off = SPMV_LANES / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  double d1 = 0.0, d2 = 0.0;
  if (row < n && sl == 0) {
    const double yv = mode ? (b[row] - acc) : acc;
    y[row] = yv;
    if (a1) d1 = a1[row] * yv;
    d2 = yv * yv;
  }
  if (part1 || part2) {
    d1 = wave_sum(d1);
    d2 = wave_sum(d2);
    if ((threadIdx.x & 63) == 0) {
      sm[0][threadIdx.x >> 6] = d1;
      sm[1][threadIdx.x >> 6] = d2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      if (part1) part1[bx] = (sm[0][0] + sm[0][1]) + (sm[0][2] + sm[0][3]);
      if (part2) part2[bx] = (sm[1][0] + sm[1][1]) + (sm[1][2] + sm[1][3]);
    }
  }
}
__global__ __launch_bounds__(256) void k_spmv(int n, const int* __restrict__ rowptr,
                                              const int* __restrict__ col,
                                              const double* __restrict__ val,
                                              const double* __restrict__ x,
                                              const double* __restrict__ b,
                                              double* __restrict__ y, int mode,
                                              const double* __restrict__ a1,
                                              double* __restrict__ part1,
                                              double* __restrict__ part2) {
  spmv_body(blockIdx.x, n, rowptr, col, val, x, b, y, mode, a1, part1, part2);
}

constexpr int RED_BLOCKS = 512;

__device__ inline void dot_partial_body(int bx, int n, const double* __restrict__ a,
                                        const double* __restrict__ b, double* __restrict__ partial) {
  __shared__ double sm[4];
  double acc = 0.0;
  for (int i = bx * 256 + threadIdx.x; i < n; i += RED_BLOCKS * 256) acc += a[i] * b[i];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[bx] = (sm[0] + sm[1]) + (sm[2] + sm[3]);
}
__global__ __launch_bounds__(256) void k_dot_partial(int n, const double* __restrict__ a,
                                                     const double* __restrict__ b,
                                                     double* __restrict__ partial) {
  dot_partial_body(blockIdx.x, n, a, b, partial);
}
__device__ inline void reduce_final_body(const double* __restrict__ partial, int np, double* __restrict__ out,
                                         int is_max);
__global__ __launch_bounds__(256) void k_reduce_final(const double* __restrict__ partial, int np,
                                                      double* __restrict__ out, int is_max) {
  reduce_final_body(partial, np, out, is_max);
}
__device__ inline void reduce_final_body(const double* __restrict__ partial, int np, double* __restrict__ out,
                                         int is_max) {
  __shared__ double sm[4];
  double acc = 0.0;
  // thread t takes entries t, t + 256, ... in turn; loads issue 8 at a time
  for (int i0 = threadIdx.x; i0 < np; i0 += 256 * 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = (i0 + 256 * u < np) ? partial[i0 + 256 * u] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (i0 + 256 * u < np) acc = is_max ? fmax(acc, v[u]) : acc + v[u];
  }
  if (is_max) acc = wave_max(acc); else acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0)
    *out = is_max ? fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3])) : (sm[0] + sm[1]) + (sm[2] + sm[3]);
}
// Residuals and dot products of nr <= SOLVE_NRHS columns (column j at j * ld) in
// one launch each, per column in the order of k_spmv (mode 1) and of
// k_dot_partial + k_reduce_final, so each column's value equals the single one
__global__ __launch_bounds__(256) void k_spmv_resid_m(int n, int nr, const int* __restrict__ rowptr,
                                                      const int* __restrict__ col,
                                                      const double* __restrict__ val,
                                                      const double* __restrict__ X,
                                                      const double* __restrict__ Bv, double* __restrict__ Y,
                                                      long long ld) {
  const int row = blockIdx.x * SPMV_ROWS + threadIdx.x / SPMV_LANES;
  const int sl = threadIdx.x & (SPMV_LANES - 1);
  double acc[SOLVE_NRHS];
#pragma unroll
  for (int j = 0; j < SOLVE_NRHS; j++) acc[j] = 0.0;
  if (row < n) {
    const int e1 = rowptr[row + 1];
    double a0[SOLVE_NRHS], b0[SOLVE_NRHS];
#pragma unroll
    for (int j = 0; j < SOLVE_NRHS; j++) a0[j] = b0[j] = 0.0;
    // k_spmv's order: the lane's q-th entry into a0 (q even) or b0 (q odd)
    for (int e = rowptr[row] + sl; e < e1; e += 2 * SPMV_LANES) {
      const bool ok1 = e + SPMV_LANES < e1;
      const double v0 = val[e], v1 = ok1 ? val[e + SPMV_LANES] : 0.0;
      const int c0 = col[e], c1 = ok1 ? col[e + SPMV_LANES] : 0;
#pragma unroll
      for (int j = 0; j < SOLVE_NRHS; j++) {
        if (j >= nr) break;
        const double* x = X + j * ld;
        const double x0 = x[c0], x1 = ok1 ? x[c1] : 0.0;
        a0[j] = fma(v0, x0, a0[j]);
        b0[j] = ok1 ? fma(v1, x1, b0[j]) : b0[j];
      }
    }
#pragma unroll
    for (int j = 0; j < SOLVE_NRHS; j++) acc[j] = a0[j] + b0[j];
  }
#pragma unroll
  for (int j = 0; j < SOLVE_NRHS; j++)
    for (int off = SPMV_LANES / 2; off > 0; off >>= 1) acc[j] += __shfl_xor(acc[j], off, 64);
  if (row < n && sl == 0) {
#pragma unroll
    for (int j = 0; j < SOLVE_NRHS; j++)
      if (j < nr) Y[j * ld + row] = Bv[j * ld + row] - acc[j];
  }
}
__global__ __launch_bounds__(256) void k_dot_partial_m(int n, const double* __restrict__ A,
                                                       const double* __restrict__ Bv, long long ld,
                                                       double* __restrict__ partial) {
  __shared__ double sm[4];
  const double* a = A + blockIdx.y * ld;
  const double* b = Bv + blockIdx.y * ld;
  double acc = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += RED_BLOCKS * 256) acc += a[i] * b[i];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.y * RED_BLOCKS + blockIdx.x] = (sm[0] + sm[1]) + (sm[2] + sm[3]);
}
__global__ __launch_bounds__(256) void k_reduce_final_m(const double* __restrict__ partial,
                                                        double* __restrict__ out) {
  __shared__ double sm[4];
  const double* p = partial + blockIdx.x * RED_BLOCKS;
  double acc = 0.0;
  for (int i = threadIdx.x; i < RED_BLOCKS; i += 256) acc = acc + p[i];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = (sm[0] + sm[1]) + (sm[2] + sm[3]);
}
// One workgroup: out1 = sum of p1[0, np), out2 = sum of p2[0, np) in a fixed
// order (thread t adds entries t, t + 256, ... in turn, then the wave sums, then
// the four waves in order).  A thread's loads issue RED_AHEAD at a time before its
// adds, so the partials cost a few memory round trips instead of one per entry.
// With sc, thread 0 then starts the next BiCGStab iteration's scalars (slots of
// the BiCGStab kernels below): alpha = rho / (r_hat, v), omega = (t, s) / (t, t),
// rho_prev = rho, rho = out2's value (the caller passes out2 = sc + 10).
constexpr int RED_AHEAD = 8;
__device__ inline void reduce2_body(const double* __restrict__ p1, const double* __restrict__ p2, int np,
                                    double* __restrict__ out1, double* __restrict__ out2,
                                    double* __restrict__ sc);
__global__ __launch_bounds__(256) void k_reduce2(const double* __restrict__ p1,
                                                 const double* __restrict__ p2, int np,
                                                 double* __restrict__ out1,
                                                 double* __restrict__ out2, double* __restrict__ sc) {
  reduce2_body(p1, p2, np, out1, out2, sc);
}
__device__ inline void reduce2_body(const double* __restrict__ p1, const double* __restrict__ p2, int np,
                                    double* __restrict__ out1, double* __restrict__ out2,
                                    double* __restrict__ sc) {
  __shared__ double sm[2][4];
  double a = 0.0, c = 0.0;
  for (int i0 = threadIdx.x; i0 < np; i0 += 256 * RED_AHEAD) {
    double va[RED_AHEAD], vc[RED_AHEAD];
#pragma unroll
    for (int u = 0; u < RED_AHEAD; u++) {
      const int i = i0 + 256 * u;
      va[u] = (i < np) ? p1[i] : 0.0;
      vc[u] = (p2 && i < np) ? p2[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < RED_AHEAD; u++) {
      if (i0 + 256 * u < np) {
        a += va[u];
        c += vc[u];
      }
    }
  }
  a = wave_sum(a);
  c = wave_sum(c);
  if ((threadIdx.x & 63) == 0) {
    sm[0][threadIdx.x >> 6] = a;
    sm[1][threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double s1 = (sm[0][0] + sm[0][1]) + (sm[0][2] + sm[0][3]);
    const double s2 = (sm[1][0] + sm[1][1]) + (sm[1][2] + sm[1][3]);
    *out1 = s1;
    if (out2) *out2 = s2;
    if (sc) {
      const double rho = sc[0];
      sc[2] = rho / sc[4];
      sc[3] = sc[6] / sc[7];
      sc[1] = rho;
      sc[0] = s2;
    }
  }
}
// one block's sum of v over its 256 threads into part[bx] (fixed order)
__device__ inline void block_sum_to(double v, double* part, int bx) {
  __shared__ double sm[4];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) part[bx] = (sm[0] + sm[1]) + (sm[2] + sm[3]);
}
__global__ __launch_bounds__(256) void k_absmax_partial(long long n, const double* __restrict__ v,
                                                        double* __restrict__ partial) {
  __shared__ double sm[4];
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)RED_BLOCKS * 256)
    acc = fmax(acc, fabs(v[i]));
  acc = wave_max(acc);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
}

__global__ void k_axpby(int n, double alpha, const double* __restrict__ x, double beta,
                        double* __restrict__ y) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = alpha * x[i] + beta * y[i];
}
__global__ void k_axpbypcz(int n, double a, const double* __restrict__ x, double b,
                           const double* __restrict__ y, double c, double* __restrict__ z) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) z[i] = a * x[i] + b * y[i] + c * z[i];
}

// ------------------------------------------------------------------------
// BiCGStab vector updates with device-resident scalars (no host round trips):
// slots: 0 rho, 1 rho_prev, 2 alpha, 3 omega, 4 (r_hat,v), 5 |s|^2, 6 (t,s), 7 (t,t),
// 8 |r|^2, 10 the next rho (r_hat, r)
// ------------------------------------------------------------------------
// s = r - alpha v with alpha = rho / (r_hat, v); part[block] = sum of s_i^2
__device__ inline void bicg_s_body(int bx, int n, const double* __restrict__ sc, const double* __restrict__ r,
                                   const double* __restrict__ v, double* __restrict__ s,
                                   double* __restrict__ part) {
  const int i = bx * 256 + threadIdx.x;
  double d = 0.0;
  if (i < n) {
    const double alpha = sc[0] / sc[4];
    const double si = r[i] - alpha * v[i];
    s[i] = si;
    d = si * si;
  }
  block_sum_to(d, part, bx);
}
__global__ __launch_bounds__(256) void k_bicg_s(int n, const double* __restrict__ sc,
                                                const double* __restrict__ r,
                                                const double* __restrict__ v, double* __restrict__ s,
                                                double* __restrict__ part) {
  bicg_s_body(blockIdx.x, n, sc, r, v, s, part);
}
// x += alpha y + omega z ; r = s - omega t, omega = (t,s)/(t,t);
// part1[block] = sum of r_i^2, part2[block] = sum of r_hat_i r_i (the next rho)
__device__ inline void bicg_xr_body(int bx, int n, const double* __restrict__ sc, const double* __restrict__ y,
                                    const double* __restrict__ z, const double* __restrict__ s,
                                    const double* __restrict__ t, double* __restrict__ x, double* __restrict__ r,
                                    const double* __restrict__ rh, double* __restrict__ part1,
                                    double* __restrict__ part2) {
  const int i = bx * 256 + threadIdx.x;
  double d1 = 0.0, d2 = 0.0;
  if (i < n) {
    const double alpha = sc[0] / sc[4], omega = sc[6] / sc[7];
    x[i] = x[i] + alpha * y[i] + omega * z[i];
    const double ri = s[i] - omega * t[i];
    r[i] = ri;
    d1 = ri * ri;
    d2 = rh[i] * ri;
  }
  block_sum_to(d1, part1, bx);
  __syncthreads();  // block_sum_to's LDS is reused
  block_sum_to(d2, part2, bx);
}
__global__ __launch_bounds__(256) void k_bicg_xr(int n, const double* __restrict__ sc,
                                                 const double* __restrict__ y,
                                                 const double* __restrict__ z,
                                                 const double* __restrict__ s,
                                                 const double* __restrict__ t, double* __restrict__ x,
                                                 double* __restrict__ r, const double* __restrict__ rh,
                                                 double* __restrict__ part1, double* __restrict__ part2) {
  bicg_xr_body(blockIdx.x, n, sc, y, z, s, t, x, r, rh, part1, part2);
}
// x += alpha y (early exit on small |s|)
__device__ inline void bicg_xa_body(int bx, int n, const double* __restrict__ sc, const double* __restrict__ y,
                                    double* __restrict__ x) {
  const int i = bx * 256 + threadIdx.x;
  if (i >= n) return;
  x[i] += (sc[0] / sc[4]) * y[i];
}
__global__ void k_bicg_xa(int n, const double* __restrict__ sc, const double* __restrict__ y,
                          double* __restrict__ x) {
  bicg_xa_body(blockIdx.x, n, sc, y, x);
}
// the first iteration's scalars: rho = |r0|^2 (r_hat = r0), rho_prev = alpha =
// omega = 1 (later iterations' come from k_reduce2 with sc)
__device__ inline void bicg_init_body(int bx, double* __restrict__ sc) {
  if (threadIdx.x != 0 || bx != 0) return;
  sc[0] = sc[8];
  sc[1] = 1.0;
  sc[2] = 1.0;
  sc[3] = 1.0;
}
__global__ void k_bicg_init(double* __restrict__ sc) { bicg_init_body(blockIdx.x, sc); }
__device__ inline void fill_body(int bx, long long n, double v, double* __restrict__ x) {
  const long long i = (long long)bx * 256 + threadIdx.x;
  if (i < n) x[i] = v;
}
__global__ void k_fill(long long n, double v, double* __restrict__ x) { fill_body(blockIdx.x, n, v, x); }
__global__ void k_zero_counters(FactorCounters* c, int* __restrict__ rctr, int nr) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i == 0) { c->npos = 0; c->nneg = 0; c->npert = 0; c->n2x2 = 0; c->npivoted = 0; c->half_overflow = 0; }
  if (i < nr) rctr[i] = 0;
}
// max |v| over each chunk [ca[q], cb[q]) of the value array (one workgroup each)
__global__ __launch_bounds__(256) void k_absmax_chunks(const long long* __restrict__ ca,
                                                       const long long* __restrict__ cb,
                                                       const double* __restrict__ v,
                                                       double* __restrict__ partial) {
  __shared__ double sm[4];
  const long long a = ca[blockIdx.x], b = cb[blockIdx.x];
  double acc = 0.0;
  for (long long i = a + threadIdx.x; i < b; i += 256) acc = fmax(acc, fabs(v[i]));
  acc = wave_max(acc);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
}
// sum of squares of v over each chunk [ca[q], cb[q]) (one workgroup each, fixed
// order: thread-strided partial sums, wave sums, then the four waves in order)
__global__ __launch_bounds__(256) void k_sumsq_chunks(const long long* __restrict__ ca,
                                                      const long long* __restrict__ cb,
                                                      const double* __restrict__ v,
                                                      double* __restrict__ partial) {
  __shared__ double sm[4];
  const long long a = ca[blockIdx.x], b = cb[blockIdx.x];
  double acc = 0.0;
  for (long long i = a + threadIdx.x; i < b; i += 256) acc += v[i] * v[i];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (sm[0] + sm[1]) + (sm[2] + sm[3]);
}
// per segment g: the sum of its chunks' partial sums in chunk order
__global__ void k_seg_sum(int nseg, const int* __restrict__ cptr, const double* __restrict__ partial,
                          double* __restrict__ out) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= nseg) return;
  double acc = 0.0;
  for (int q = cptr[g]; q < cptr[g + 1]; q++) acc += partial[q];
  out[g] = acc;
}
// x_i += d_i for the rows whose segment is still active
__global__ void k_axpy_masked(long long n, const int* __restrict__ row_seg,
                              const unsigned char* __restrict__ active, const double* __restrict__ d,
                              double* __restrict__ x) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n && active[row_seg[i]]) x[i] += d[i];
}
// per component r: max of its chunks' maxima, chunks [cptr[r], cptr[r+1])
__global__ void k_absmax_roots(int nroots, const int* __restrict__ cptr, const double* __restrict__ partial,
                               double* __restrict__ amax) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= nroots) return;
  double acc = 0.0;
  for (int q = cptr[r]; q < cptr[r + 1]; q++) acc = fmax(acc, partial[q]);
  amax[r] = acc;
}

// ------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------
static inline unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }

template <typename T>
void launch_assemble(const int4* items, int nitems, const DevSym& d, T* F, hipStream_t st) {
  if (nitems > 0) hipLaunchKernelGGL(k_assemble<T>, dim3(nitems), dim3(256), 0, st, items, d, F);
}
template <typename T>
void launch_scatter(const int* aidx, const long long* apos, int cnt, const double* vals, T* F,
                    hipStream_t st) {
  if (cnt > 0)
    hipLaunchKernelGGL(k_scatter<T>, dim3(nblk(cnt, 256)), dim3(256), 0, st, aidx, apos, cnt, vals, F);
}
template <typename T>
void launch_panel_diag(const int* items, int nitems, int p, const DevSym& d, T* F, T* E0, T* E1,
                       double pert, const double* amax, FactorCounters* ctr, int* rctr, hipStream_t st) {
  if (nitems > 0)
    hipLaunchKernelGGL(k_panel_diag<T>, dim3(nitems), dim3(256), 0, st, items, p, d, F, E0, E1, pert,
                       amax, ctr, rctr);
}
template <typename T>
void launch_panel_trsm(const int2* items, int nitems, int p, int slot, const DevSym& d, T* F, T* W,
                       hipStream_t st) {
  if (nitems > 0)
    hipLaunchKernelGGL(k_panel_trsm<T>, dim3(nitems), dim3(256), 0, st, items, p, slot, d, F, W);
}
template <typename T>
void launch_panel_update(const int4* items, int nitems, int pfirst, const DevSym& d, T* F,
                         const T* W, hipStream_t st) {
  if (nitems > 0)
    hipLaunchKernelGGL(k_panel_update<T>, dim3(nitems), dim3(256), 0, st, items, pfirst, d, F, W);
}
template <typename T>
void launch_panel_xrow(const int2* items, int nitems, int p, const DevSym& d, T* F, hipStream_t st) {
  if (nitems > 0) hipLaunchKernelGGL(k_panel_xrow<T>, dim3(nitems), dim3(256), 0, st, items, p, d, F);
}
template <typename T>
void launch_zt(const int4* items, int nitems, const DevSym& d, T* F, hipStream_t st) {
  if (nitems > 0) hipLaunchKernelGGL(k_zt<T>, dim3(nitems), dim3(256), 0, st, items, d, F);
}
template <typename T>
void launch_gform(const int4* items, int nitems, const DevSym& d, T* F, hipStream_t st) {
  if (nitems > 0) hipLaunchKernelGGL(k_gform<T>, dim3(nitems), dim3(256), 0, st, items, d, F);
}
template <typename T>
void launch_pack_solve(const int2* items, int nitems, const DevSym& d, const T* F, T* PF, T* PB, hipStream_t st) {
  if (nitems > 0) hipLaunchKernelGGL(k_pack_solve<T>, dim3(nitems), dim3(256), 0, st, items, d, F, PF, PB);
}
template <typename T, typename S>
void launch_fwd_level(const SolveItem* items, int nitems, const DevSym& d, const S* F, const T* bp, T* fp,
                      T* y, T* u, hipStream_t st) {
  if (nitems > 0)
    hipLaunchKernelGGL((k_fwd_level<T, S>), dim3(nitems), dim3(256), 0, st, items, d, F, bp, fp, y, u);
}
template <typename T>
void launch_dsolve(int n, const T* E0, const T* E1, const signed char* piv, const int* wpos,
                   const T* y, T* wv, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_dsolve<T>, dim3(nblk(n, 256)), dim3(256), 0, st, n, E0, E1, piv, wpos, y, wv);
}
template <typename T, typename S>
void launch_bwd_cols(const SolveItem* items, int nitems, const DevSym& d, const S* F, T* wv, T* xp,
                     hipStream_t st) {
  if (nitems > 0) hipLaunchKernelGGL((k_bwd_cols<T, S>), dim3(nitems), dim3(256), 0, st, items, d, F, wv, xp);
}
template <typename T>
void launch_fwd_gather_m(const SolveItem* items, int nitems, const DevSym& d, const T* F, int nr, const T* bp,
                         T* fp, T* y, T* u, long long ldn, long long ldu, hipStream_t st) {
  if (nitems > 0)
    hipLaunchKernelGGL(k_fwd_gather_m<T>, dim3(nitems), dim3(256), 0, st, items, d, F, nr, bp, fp, y, u, ldn,
                       ldu);
}
template <typename T>
void launch_fwd_rows_m(const SolveItem* items, int nitems, const DevSym& d, const T* F, int nr, const T* fp,
                       T* y, T* u, long long ldn, long long ldu, hipStream_t st) {
  if (nitems > 0)
    hipLaunchKernelGGL(k_fwd_rows_m<T>, dim3(nitems), dim3(256), 0, st, items, d, F, nr, fp, y, u, ldn, ldu);
}
template <typename T>
void launch_dsolve_m(int n, int nr, const T* E0, const T* E1, const signed char* piv, const int* wpos,
                     const T* y, T* wv, long long ldn, long long ldw, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_dsolve_m<T>, dim3(nblk(n, 256), nr), dim3(256), 0, st, n, nr, E0, E1, piv, wpos, y,
                       wv, ldn, ldw);
}
template <typename T>
void launch_bwd_cols_m(const SolveItem* items, int nitems, const DevSym& d, const T* F, int nr, T* wv, T* xp,
                       long long ldn, long long ldw, hipStream_t st) {
  if (nitems > 0)
    hipLaunchKernelGGL(k_bwd_cols_m<T>, dim3(nitems), dim3(256), 0, st, items, d, F, nr, wv, xp, ldn, ldw);
}
template <typename TI, typename TO>
void launch_gather_perm_m(int n, int nr, const int* perm, const TI* in, long long ldi, TO* out, long long ldo,
                          hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL((k_gather_perm_m<TI, TO>), dim3(nblk(n, 256), nr), dim3(256), 0, st, n, nr, perm, in,
                       ldi, out, ldo);
}
template <typename TI, typename TO>
void launch_scatter_perm_m(int n, int nr, const int* perm, const TI* in, long long ldi, TO* out, long long ldo,
                           hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL((k_scatter_perm_m<TI, TO>), dim3(nblk(n, 256), nr), dim3(256), 0, st, n, nr, perm, in,
                       ldi, out, ldo);
}
template <typename TI, typename TO>
void launch_gather_perm(int n, const int* perm, const TI* in, TO* out, hipStream_t st) {
  hipLaunchKernelGGL((k_gather_perm<TI, TO>), dim3(nblk(n, 256)), dim3(256), 0, st, n, perm, in, out);
}
template <typename TI, typename TO>
void launch_scatter_perm(int n, const int* perm, const TI* in, TO* out, hipStream_t st) {
  hipLaunchKernelGGL((k_scatter_perm<TI, TO>), dim3(nblk(n, 256)), dim3(256), 0, st, n, perm, in, out);
}
template <typename TO>
void launch_gather_bicg_p(int n, const int* perm, const double* sc, const double* r, const double* v, double* p,
                          TO* out, hipStream_t st) {
  hipLaunchKernelGGL((k_gather_bicg_p<TO>), dim3(nblk(n, 256)), dim3(256), 0, st, n, perm, sc, r, v, p, out);
}
void launch_expand_vals(long long nnzf, const int* map, const double* vals, double* valsf,
                        hipStream_t st) {
  if (nnzf > 0)
    hipLaunchKernelGGL(k_expand_vals, dim3(nblk(nnzf, 256)), dim3(256), 0, st, nnzf, map, vals, valsf);
}
void launch_spmv(int n, const int* rowptr, const int* col, const double* val, const double* x,
                 const double* b, double* y, int mode, hipStream_t st, const double* a1, double* part1,
                 double* part2) {
  hipLaunchKernelGGL(k_spmv, dim3(nblk(n, SPMV_ROWS)), dim3(256), 0, st, n, rowptr, col, val, x, b, y, mode,
                     a1, part1, part2);
}
int spmv_blocks(int n) { return (int)nblk(n, SPMV_ROWS); }
void launch_reduce2(const double* p1, const double* p2, int np, double* out1, double* out2, hipStream_t st,
                    double* sc) {
  hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(256), 0, st, p1, p2, np, out1, out2, sc);
}
void launch_absmax(long long nnz, const double* v, double* partial, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_absmax_partial, dim3(RED_BLOCKS), dim3(256), 0, st, nnz, v, partial);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, (const double*)partial, RED_BLOCKS, out, 1);
}
void launch_dot(int n, const double* a, const double* b, double* partial, double* out,
                hipStream_t st) {
  hipLaunchKernelGGL(k_dot_partial, dim3(RED_BLOCKS), dim3(256), 0, st, n, a, b, partial);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, (const double*)partial, RED_BLOCKS, out, 0);
}
void launch_spmv_resid_m(int n, int nr, const int* rowptr, const int* col, const double* val, const double* X,
                         const double* B, double* Y, long long ld, hipStream_t st) {
  if (n > 0 && nr > 0)
    hipLaunchKernelGGL(k_spmv_resid_m, dim3(nblk(n, SPMV_ROWS)), dim3(256), 0, st, n, nr, rowptr, col, val, X, B,
                       Y, ld);
}
int dot_m_partials() { return RED_BLOCKS * SOLVE_NRHS; }
void launch_dot_m(int n, int nr, const double* A, const double* B, long long ld, double* partial, double* out,
                  hipStream_t st) {
  if (nr <= 0) return;
  hipLaunchKernelGGL(k_dot_partial_m, dim3(RED_BLOCKS, nr), dim3(256), 0, st, n, A, B, ld, partial);
  hipLaunchKernelGGL(k_reduce_final_m, dim3(nr), dim3(256), 0, st, (const double*)partial, out);
}
void launch_axpby(int n, double alpha, const double* x, double beta, double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_axpby, dim3(nblk(n, 256)), dim3(256), 0, st, n, alpha, x, beta, y);
}
void launch_axpbypcz(int n, double a, const double* x, double b, const double* y, double c,
                     double* z, hipStream_t st) {
  hipLaunchKernelGGL(k_axpbypcz, dim3(nblk(n, 256)), dim3(256), 0, st, n, a, x, b, y, c, z);
}

void launch_contrib_block3(int nblocks, const double* bv, const double* bs, double* c, hipStream_t st) {
  if (nblocks > 0)
    hipLaunchKernelGGL(k_contrib_block3, dim3(nblk(9LL * nblocks, 256)), dim3(256), 0, st, nblocks, bv, bs, c);
}
void launch_contrib_diag3(int nblocks, const double* m, double* c, hipStream_t st) {
  if (nblocks > 0)
    hipLaunchKernelGGL(k_contrib_diag3, dim3(nblk(3LL * nblocks, 256)), dim3(256), 0, st, nblocks, m, c);
}
void launch_contrib_matstiff3(int nblocks, const double* gi, const double* gj, const double* D,
                              const double* sg, const double* dv, double* c, hipStream_t st) {
  if (nblocks > 0)
    hipLaunchKernelGGL(k_contrib_matstiff3, dim3(nblk(nblocks, 256)), dim3(256), 0, st, nblocks, gi, gj,
                       D, sg, dv, c);
}
void launch_contrib_matstiff3_elem(int nelems, const int* nodeCnt, const int* nodeOff, const int* pairOff,
                                   const int* ipOff, const int* gradOff, const int* pairNodes,
                                   const int* nodeDims, const double* T, const double* grads,
                                   const double* Ds, const double* sigmas, const double* dvs, double* c,
                                   hipStream_t st) {
  if (nelems > 0)
    hipLaunchKernelGGL(k_contrib_matstiff3_elem, dim3(nelems), dim3(64), 0, st, nodeCnt, nodeOff, pairOff,
                       ipOff, gradOff, pairNodes, nodeDims, T, grads, Ds, sigmas, dvs, c);
}
void launch_contrib_linelastic3_elem(int nelems, const int* nodeCnt, const int* pairOff, const int* ipOff,
                                     const int* gradOff, const int* pairNodes, const double* params,
                                     const double* grads, const double* dvs, double* c, hipStream_t st) {
  if (nelems > 0)
    hipLaunchKernelGGL(k_contrib_linelastic3_elem, dim3(nelems), dim3(64), 0, st, nodeCnt, pairOff, ipOff,
                       gradOff, pairNodes, params, grads, dvs, c);
}
void launch_contrib_linelastic3_geom_elem(int nelems, const int* nodeCnt, const int* nodeOff,
                                          const int* pairOff, const int* ipOff, const int* ngOff,
                                          const int* pairNodes, const int* nodeDims, const double* T,
                                          const double* params, const double* pos, const double* ngrads,
                                          const double* w, double* c, hipStream_t st) {
  if (nelems > 0)
    hipLaunchKernelGGL(k_contrib_linelastic3_geom_elem, dim3(nelems), dim3(64), 0, st, nodeCnt, nodeOff,
                       pairOff, ipOff, ngOff, pairNodes, nodeDims, T, params, pos, ngrads, w, c);
}
void launch_contrib_dilational3_elem(int nelems, const int* nodeOff, const int* presCnt, const int* pairOff,
                                     const int* consOff, const int* rinvOff, const int* pairNodes,
                                     const int* nodeDims, const double* T, const double* cons,
                                     const double* rinv, double* c, hipStream_t st) {
  if (nelems > 0)
    hipLaunchKernelGGL(k_contrib_dilational3_elem, dim3(nelems), dim3(64), 0, st, nodeOff, presCnt, pairOff,
                       consOff, rinvOff, pairNodes, nodeDims, T, cons, rinv, c);
}
void launch_seg_scatter(int nuniq, const int* useg, const int* uslot, const int* order,
                        const double* c, double scale, double* vals, hipStream_t st) {
  if (nuniq > 0)
    hipLaunchKernelGGL(k_seg_scatter, dim3(nblk(nuniq, 256)), dim3(256), 0, st, nuniq, useg, uslot, order,
                       c, scale, vals);
}
void launch_bicg_s(int n, const double* sc, const double* r, const double* v, double* s_, double* part,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_bicg_s, dim3(nblk(n, 256)), dim3(256), 0, st, n, sc, r, v, s_, part);
}
void launch_bicg_xr(int n, const double* sc, const double* y, const double* z, const double* s_,
                    const double* t, double* x, double* r, const double* rh, double* part1, double* part2,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_bicg_xr, dim3(nblk(n, 256)), dim3(256), 0, st, n, sc, y, z, s_, t, x, r, rh, part1,
                     part2);
}
int vec_blocks(int n) { return (int)nblk(n, 256); }
void launch_bicg_xa(int n, const double* sc, const double* y, double* x, hipStream_t st) {
  hipLaunchKernelGGL(k_bicg_xa, dim3(nblk(n, 256)), dim3(256), 0, st, n, sc, y, x);
}
void launch_bicg_init(double* sc, hipStream_t st) {
  hipLaunchKernelGGL(k_bicg_init, dim3(1), dim3(64), 0, st, sc);
}
void launch_fill(long long n, double v, double* x, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_fill, dim3(nblk(n, 256)), dim3(256), 0, st, n, v, x);
}
void launch_pack_half(long long n, const float* in, _Float16* out, FactorCounters* ctr, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_pack_half, dim3(nblk(n, 256)), dim3(256), 0, st, n, in, out, ctr);
}
void launch_zero_counters(FactorCounters* c, int* rctr, int nr, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_counters, dim3(nblk(std::max(nr, 1), 256)), dim3(256), 0, st, c, rctr, nr);
}
void launch_seg_sumsq(int nchunks, const long long* ca, const long long* cb, const double* v, double* partial,
                      int nseg, const int* cptr, double* out, hipStream_t st) {
  if (nchunks > 0) hipLaunchKernelGGL(k_sumsq_chunks, dim3(nchunks), dim3(256), 0, st, ca, cb, v, partial);
  if (nseg > 0)
    hipLaunchKernelGGL(k_seg_sum, dim3(nblk(nseg, 256)), dim3(256), 0, st, nseg, cptr, (const double*)partial,
                       out);
}
void launch_axpy_masked(long long n, const int* row_seg, const unsigned char* active, const double* d, double* x,
                        hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_axpy_masked, dim3(nblk(n, 256)), dim3(256), 0, st, n, row_seg, active, d, x);
}
void launch_absmax_roots(int nchunks, const long long* ca, const long long* cb, const double* v,
                         double* partial, int nroots, const int* cptr, double* amax, hipStream_t st) {
  if (nchunks > 0) hipLaunchKernelGGL(k_absmax_chunks, dim3(nchunks), dim3(256), 0, st, ca, cb, v, partial);
  if (nroots > 0)
    hipLaunchKernelGGL(k_absmax_roots, dim3(nblk(nroots, 256)), dim3(256), 0, st, nroots, cptr,
                       (const double*)partial, amax);
}

#define KKT_INST(T)                                                                              \
  template void launch_assemble<T>(const int4*, int, const DevSym&, T*, hipStream_t);            \
  template void launch_scatter<T>(const int*, const long long*, int, const double*, T*,          \
                                  hipStream_t);                                                  \
  template void launch_panel_diag<T>(const int*, int, int, const DevSym&, T*, T*, T*, double,   \
                                     const double*, FactorCounters*, int*, hipStream_t);         \
  template void launch_panel_trsm<T>(const int2*, int, int, int, const DevSym&, T*, T*, hipStream_t); \
  template void launch_panel_update<T>(const int4*, int, int, const DevSym&, T*, const T*,       \
                                       hipStream_t);                                             \
  template void launch_panel_xrow<T>(const int2*, int, int, const DevSym&, T*, hipStream_t);     \
  template void launch_zt<T>(const int4*, int, const DevSym&, T*, hipStream_t);                \
  template void launch_gform<T>(const int4*, int, const DevSym&, T*, hipStream_t);             \
  template void launch_pack_solve<T>(const int2*, int, const DevSym&, const T*, T*, T*, hipStream_t); \
  template void launch_fwd_level<T, T>(const SolveItem*, int, const DevSym&, const T*, const T*, T*, T*, \
                                       T*, hipStream_t);                                         \
  template void launch_dsolve<T>(int, const T*, const T*, const signed char*, const int*,        \
                                 const T*, T*, hipStream_t);                                     \
  template void launch_bwd_cols<T, T>(const SolveItem*, int, const DevSym&, const T*, T*, T*,       \
                                      hipStream_t);
KKT_INST(double)
KKT_INST(float)
template void launch_fwd_level<float, _Float16>(const SolveItem*, int, const DevSym&, const _Float16*, const float*,
                                                float*, float*, float*, hipStream_t);
template void launch_bwd_cols<float, _Float16>(const SolveItem*, int, const DevSym&, const _Float16*, float*,
                                               float*, hipStream_t);
template void launch_gather_perm<double, double>(int, const int*, const double*, double*, hipStream_t);
template void launch_gather_perm<double, float>(int, const int*, const double*, float*, hipStream_t);
template void launch_gather_bicg_p<double>(int, const int*, const double*, const double*, const double*, double*,
                                           double*, hipStream_t);
template void launch_gather_bicg_p<float>(int, const int*, const double*, const double*, const double*, double*,
                                          float*, hipStream_t);
template void launch_scatter_perm<double, double>(int, const int*, const double*, double*, hipStream_t);
template void launch_scatter_perm<float, double>(int, const int*, const float*, double*, hipStream_t);
#define KKT_INST_M(T)                                                                                        \
  template void launch_fwd_gather_m<T>(const SolveItem*, int, const DevSym&, const T*, int, const T*, T*, T*, \
                                       T*, long long, long long, hipStream_t);                               \
  template void launch_fwd_rows_m<T>(const SolveItem*, int, const DevSym&, const T*, int, const T*, T*, T*,   \
                                     long long, long long, hipStream_t);                                     \
  template void launch_dsolve_m<T>(int, int, const T*, const T*, const signed char*, const int*, const T*,   \
                                   T*, long long, long long, hipStream_t);                                   \
  template void launch_bwd_cols_m<T>(const SolveItem*, int, const DevSym&, const T*, int, T*, T*, long long,  \
                                     long long, hipStream_t);                                                \
  template void launch_gather_perm_m<double, T>(int, int, const int*, const double*, long long, T*,          \
                                                long long, hipStream_t);                                     \
  template void launch_scatter_perm_m<T, double>(int, int, const int*, const T*, long long, double*,         \
                                                 long long, hipStream_t);
KKT_INST_M(float)
KKT_INST_M(double)

// ------------------------------------------------------------------------
// Group launches: the same step of many handles' solves in one launch.  Block
// b belongs to segment h with pre[h] <= b < pre[h + 1] (the prefix counts go
// through LDS, then a binary search) and runs the step's body with block index
// b - pre[h] and the segment's arguments, i.e. exactly what that handle's own
// launch of the step would run in that block.
// ------------------------------------------------------------------------
// A pointer read from a segment's argument table is generic to the compiler
// (flat loads and stores, which wait on both the vector-memory and the LDS
// counters and hold their addresses in more registers).  Rebuilt from its
// integer value as a global-address-space pointer, it gets global loads and
// stores, as a kernel's own pointer arguments do (a direct cast from the generic
// pointer is folded away).
template <typename P>
__device__ __forceinline__ P* gbl(P* p) {
  typedef __attribute__((address_space(1))) P gp;
  return (P*)(gp*)(unsigned long long)p;
}
// the symbolic arrays the solve bodies read, from a segment's table
__device__ __forceinline__ DevSym gbl_sym(const DevSym& s) {
  DevSym d = s;
  d.rel = gbl(s.rel);
  d.ilperm = gbl(s.ilperm);
  d.lperm = gbl(s.lperm);
  d.gptr = gbl(s.gptr);
  d.gsrc = gbl(s.gsrc);
  d.push_ptr = gbl(s.push_ptr);
  d.push_pos = gbl(s.push_pos);
  return d;
}

__device__ inline int seg_find(const int* __restrict__ pre, int nseg, int& bx) {
  __shared__ int sp[SEG_MAX + 1];
  for (int i = threadIdx.x; i <= nseg; i += 256) sp[i] = pre[i];
  __syncthreads();
  const int b = blockIdx.x;
  int lo = 0, hi = nseg;  // sp[lo] <= b < sp[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (sp[mid] <= b) lo = mid;
    else hi = mid;
  }
  // every lane found the same segment: say so, so that the segment's arguments
  // load into scalar registers
  const int h = __builtin_amdgcn_readfirstlane(lo);
  bx = __builtin_amdgcn_readfirstlane(b - sp[lo]);
  return h;
}

template <typename T>
__global__ __launch_bounds__(256) void k_seg_vec(const int* __restrict__ pre, int nseg,
                                                 const SegVec* __restrict__ args) {
  int bx;
  const int h = seg_find(pre, nseg, bx);
  const SegVec& a = args[h];
  const int n = (int)a.n;
  const int* idx = gbl(a.idx);
  const double* sc = gbl(a.sc);
  const double *va = gbl(a.a), *vb = gbl(a.b), *vc = gbl(a.c), *ve = gbl(a.e), *vg = gbl(a.g);
  double *vd = gbl(a.d), *vf = gbl(a.f), *p1 = gbl(a.part1), *p2 = gbl(a.part2);
  const T *tin = gbl((const T*)a.tin), *tin2 = gbl((const T*)a.tin2), *tin3 = gbl((const T*)a.tin3);
  T* tout = gbl((T*)a.tout);
  switch (a.op) {
    case SEG_GATHER: gather_perm_body<double, T>(bx, n, idx, va, tout); break;
    case SEG_GATHER_P: gather_bicg_p_body<T>(bx, n, idx, sc, va, vb, vd, tout); break;
    case SEG_SCATTER: scatter_perm_body<T, double>(bx, n, idx, tin, vd); break;
    case SEG_DSOLVE: dsolve_body<T>(bx, n, tin, tin2, gbl(a.piv), gbl(a.wpos), tin3, tout); break;
    case SEG_EXPAND: expand_vals_body(bx, a.n, idx, va, vd); break;
    case SEG_FILL: fill_body(bx, a.n, a.v, vd); break;
    case SEG_COPY: {
      const long long i = (long long)bx * 256 + threadIdx.x;
      if (i < a.n) vd[i] = va[i];
    } break;
    case SEG_BICG_S: bicg_s_body(bx, n, sc, va, vb, vd, p1); break;
    case SEG_BICG_XR: bicg_xr_body(bx, n, sc, va, vb, vc, ve, vd, vf, vg, p1, p2); break;
    case SEG_BICG_XA: bicg_xa_body(bx, n, sc, va, vd); break;
    case SEG_BICG_INIT: bicg_init_body(bx, vd); break;
    case SEG_COLLECT:
      if (bx == 0 && threadIdx.x < 16) vd[threadIdx.x] = va[threadIdx.x];
      break;
    default: break;
  }
}

__global__ __launch_bounds__(256) void k_seg_red(const int* __restrict__ pre, int nseg,
                                                 const SegRed* __restrict__ args) {
  int bx;
  const int h = seg_find(pre, nseg, bx);
  const SegRed& a = args[h];
  const double *p1 = gbl(a.p1), *p2 = gbl(a.p2);
  double *o1 = gbl(a.out1), *o2 = gbl(a.out2), *sc = gbl(a.sc);
  switch (a.op) {
    case SEG_RED2: reduce2_body(p1, p2, a.np, o1, o2, sc); break;
    case SEG_DOTP: dot_partial_body(bx, a.n, p1, p2, o1); break;
    case SEG_REDF: reduce_final_body(p1, a.np, o1, 0); break;
    default: break;
  }
}

__global__ __launch_bounds__(256) void k_seg_spmv(const int* __restrict__ pre, int nseg,
                                                  const SegSpmv* __restrict__ args) {
  int bx;
  const int h = seg_find(pre, nseg, bx);
  const SegSpmv& a = args[h];
  spmv_body(bx, a.n, gbl(a.rowptr), gbl(a.col), gbl(a.val), gbl(a.x), gbl(a.b), gbl(a.y), a.mode, gbl(a.a1),
            gbl(a.part1), gbl(a.part2));
}

template <typename T, typename S>
__global__ __launch_bounds__(256) void k_seg_fwd(const int* __restrict__ pre, int nseg,
                                                 const SegFwd* __restrict__ args) {
  int bx;
  const int h = seg_find(pre, nseg, bx);
  const SegFwd& a = args[h];
  const DevSym d = gbl_sym(a.d);
  fwd_level_body<T, S>(bx, gbl(a.items), d, gbl((const S*)a.F), gbl((const T*)a.bp), gbl((T*)a.fp),
                       gbl((T*)a.y), gbl((T*)a.u));
}

template <typename T, typename S>
__global__ __launch_bounds__(256) void k_seg_bwd(const int* __restrict__ pre, int nseg,
                                                 const SegBwd* __restrict__ args) {
  int bx;
  const int h = seg_find(pre, nseg, bx);
  const SegBwd& a = args[h];
  const DevSym d = gbl_sym(a.d);
  bwd_cols_body<T, S>(bx, gbl(a.items), d, gbl((const S*)a.F), gbl((T*)a.wv), gbl((T*)a.xp));
}

int red_blocks() { return RED_BLOCKS; }
template <typename T>
void launch_seg_vec(const int* pre, int nseg, const SegVec* a, int nblk, hipStream_t st) {
  if (nblk > 0) hipLaunchKernelGGL(k_seg_vec<T>, dim3(nblk), dim3(256), 0, st, pre, nseg, a);
}
void launch_seg_red(const int* pre, int nseg, const SegRed* a, int nblk, hipStream_t st) {
  if (nblk > 0) hipLaunchKernelGGL(k_seg_red, dim3(nblk), dim3(256), 0, st, pre, nseg, a);
}
void launch_seg_spmv(const int* pre, int nseg, const SegSpmv* a, int nblk, hipStream_t st) {
  if (nblk > 0) hipLaunchKernelGGL(k_seg_spmv, dim3(nblk), dim3(256), 0, st, pre, nseg, a);
}
template <typename T, typename S>
void launch_seg_fwd(const int* pre, int nseg, const SegFwd* a, int nblk, hipStream_t st) {
  if (nblk > 0) hipLaunchKernelGGL((k_seg_fwd<T, S>), dim3(nblk), dim3(256), 0, st, pre, nseg, a);
}
template <typename T, typename S>
void launch_seg_bwd(const int* pre, int nseg, const SegBwd* a, int nblk, hipStream_t st) {
  if (nblk > 0) hipLaunchKernelGGL((k_seg_bwd<T, S>), dim3(nblk), dim3(256), 0, st, pre, nseg, a);
}
template void launch_seg_vec<float>(const int*, int, const SegVec*, int, hipStream_t);
template void launch_seg_vec<double>(const int*, int, const SegVec*, int, hipStream_t);
template void launch_seg_fwd<double, double>(const int*, int, const SegFwd*, int, hipStream_t);
template void launch_seg_fwd<float, float>(const int*, int, const SegFwd*, int, hipStream_t);
template void launch_seg_fwd<float, _Float16>(const int*, int, const SegFwd*, int, hipStream_t);
template void launch_seg_bwd<double, double>(const int*, int, const SegBwd*, int, hipStream_t);
template void launch_seg_bwd<float, float>(const int*, int, const SegBwd*, int, hipStream_t);
template void launch_seg_bwd<float, _Float16>(const int*, int, const SegBwd*, int, hipStream_t);

}  // namespace kkt
