// Host symbolic analysis: ordering, elimination tree, column counts,
// supernodes.  See symbolic.h for the input convention.
#include "symbolic.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace kkt {

namespace {

struct Graph {
  int n = 0;
  std::vector<int64_t> ptr;
  std::vector<int> adj;
  int deg(int v) const { return (int)(ptr[v + 1] - ptr[v]); }
};

// Symmetric adjacency (no self loops, deduplicated) of the upper pattern.
void build_graph(int n, const int* rowOffs, const int* colIdxs, Graph& g) {
  g.n = n;
  std::vector<int64_t> cnt(n + 1, 0);
  for (int i = 0; i < n; i++)
    for (int p = rowOffs[i]; p < rowOffs[i + 1]; p++) {
      int j = colIdxs[p];
      if (j > i) { cnt[i]++; cnt[j]++; }
    }
  g.ptr.assign(n + 1, 0);
  for (int i = 0; i < n; i++) g.ptr[i + 1] = g.ptr[i] + cnt[i];
  g.adj.resize(g.ptr[n]);
  std::vector<int64_t> pos(g.ptr.begin(), g.ptr.end() - 1);
  for (int i = 0; i < n; i++)
    for (int p = rowOffs[i]; p < rowOffs[i + 1]; p++) {
      int j = colIdxs[p];
      if (j > i) { g.adj[pos[i]++] = j; g.adj[pos[j]++] = i; }
    }
  // sort + dedup each list, then compact
  std::vector<int64_t> nptr(n + 1, 0);
  int64_t w = 0;
  for (int i = 0; i < n; i++) {
    int64_t b = g.ptr[i], e = g.ptr[i + 1];
    std::sort(g.adj.begin() + b, g.adj.begin() + e);
    nptr[i] = w;
    int last = -1;
    for (int64_t p = b; p < e; p++) {
      if (g.adj[p] != last) { g.adj[w++] = g.adj[p]; last = g.adj[p]; }
    }
  }
  nptr[n] = w;
  g.adj.resize(w);
  g.ptr.swap(nptr);
}

inline uint64_t mix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}

// ---------------------------------------------------------------------------
// Nested dissection on the compressed (quotient) graph.
// ---------------------------------------------------------------------------
class NestedDissection {
 public:
  NestedDissection(const Graph& g, const std::vector<int>& wgt, int leaf)
      : g_(g), w_(wgt), leaf_(std::max(leaf, 2)) {
    tag_.assign(g.n, -1);
    dist_.assign(g.n, -1);
    mark_.assign(g.n, -1);
    side_.assign(g.n, 0);
  }

  void run(std::vector<int>& order) {
    std::vector<int> all(g_.n);
    std::iota(all.begin(), all.end(), 0);
    order.clear();
    order.reserve(g_.n);
    // explicit work stack: entries are either "dissect V" or "emit V"
    struct Item { std::vector<int> v; bool emit; };
    std::vector<Item> stack;
    stack.push_back({std::move(all), false});
    while (!stack.empty()) {
      Item it = std::move(stack.back());
      stack.pop_back();
      if (it.emit) { order.insert(order.end(), it.v.begin(), it.v.end()); continue; }
      std::vector<int>& V = it.v;
      if ((int)V.size() <= leaf_) { leaf_order(V, order); continue; }
      int st = ++stamp_;
      for (int v : V) tag_[v] = st;
      // connected components
      std::vector<std::vector<int>> comps;
      components(V, st, comps);
      if (comps.size() > 1) {
        // process components in order: push in reverse
        for (int c = (int)comps.size() - 1; c >= 0; c--)
          stack.push_back({std::move(comps[c]), false});
        continue;
      }
      std::vector<int> A, B, S;
      if (!bisect(V, st, A, B, S)) { leaf_order(V, order); continue; }
      // order: A, B, then separator S
      stack.push_back({std::move(S), true});
      stack.push_back({std::move(B), false});
      stack.push_back({std::move(A), false});
    }
  }

 private:
  const Graph& g_;
  const std::vector<int>& w_;
  int leaf_;
  int stamp_ = 0;
  std::vector<int> tag_, dist_, mark_;
  std::vector<char> side_;

  void components(const std::vector<int>& V, int st, std::vector<std::vector<int>>& comps) {
    int cst = ++stamp_;
    for (int s : V) {
      if (mark_[s] == cst) continue;
      std::vector<int> comp;
      comp.push_back(s);
      mark_[s] = cst;
      for (size_t h = 0; h < comp.size(); h++) {
        int v = comp[h];
        for (int64_t p = g_.ptr[v]; p < g_.ptr[v + 1]; p++) {
          int u = g_.adj[p];
          if (tag_[u] == st && mark_[u] != cst) { mark_[u] = cst; comp.push_back(u); }
        }
      }
      comps.push_back(std::move(comp));
    }
  }

  // BFS inside the tagged subgraph; returns levels (vertices grouped by distance)
  void bfs(int root, int st, std::vector<int>& ordr, std::vector<int>& lvl_ptr) {
    int bst = ++stamp_;
    ordr.clear();
    lvl_ptr.clear();
    ordr.push_back(root);
    mark_[root] = bst;
    dist_[root] = 0;
    lvl_ptr.push_back(0);
    size_t h = 0;
    int cur = 0;
    while (h < ordr.size()) {
      int v = ordr[h];
      if (dist_[v] != cur) { lvl_ptr.push_back((int)h); cur = dist_[v]; }
      h++;
      for (int64_t p = g_.ptr[v]; p < g_.ptr[v + 1]; p++) {
        int u = g_.adj[p];
        if (tag_[u] == st && mark_[u] != bst) {
          mark_[u] = bst;
          dist_[u] = dist_[v] + 1;
          ordr.push_back(u);
        }
      }
    }
    lvl_ptr.push_back((int)ordr.size());
  }

  int pseudo_peripheral(int start, int st, std::vector<int>& ordr, std::vector<int>& lp) {
    int root = start;
    bfs(root, st, ordr, lp);
    int ecc = (int)lp.size() - 2;
    for (int it = 0; it < 8; it++) {
      // min-degree vertex of the last level
      int best = -1, bd = 1 << 30;
      for (int q = lp[lp.size() - 2]; q < lp.back(); q++) {
        int v = ordr[q];
        if (g_.deg(v) < bd) { bd = g_.deg(v); best = v; }
      }
      std::vector<int> o2, l2;
      bfs(best, st, o2, l2);
      int e2 = (int)l2.size() - 2;
      if (e2 <= ecc) break;
      ecc = e2;
      root = best;
      ordr.swap(o2);
      lp.swap(l2);
    }
    return root;
  }

  // Candidate separator from a BFS level structure: cut between level l-1
  // and level l, separator = minimum vertex cover of the cut edges.
  // Returns separator weight (or large if no valid cut), filling A/B/S.
  int64_t level_cut(const std::vector<int>& ordr, const std::vector<int>& lp, int st,
                    int64_t total, std::vector<int>& A, std::vector<int>& B,
                    std::vector<int>& S) {
    int nl = (int)lp.size() - 1;
    if (nl < 3) return INT64_MAX;
    std::vector<int64_t> lw(nl, 0);
    for (int l = 0; l < nl; l++)
      for (int q = lp[l]; q < lp[l + 1]; q++) lw[l] += w_[ordr[q]];
    // choose level l (1..nl-1) as the "B-side boundary"
    int64_t acc = 0;
    int bestl = -1;
    double bestscore = 1e300;
    for (int l = 1; l < nl; l++) {
      acc += lw[l - 1];  // weight of levels < l
      int64_t a = acc, b = total - acc;
      double lo = 0.30 * (double)total;
      if ((double)a < lo || (double)(b - lw[l]) < lo * 0.5) continue;
      double sepw = (double)std::min(lw[l], lw[l - 1]);
      double imb = std::fabs((double)a - (double)(b)) / (double)total;
      double score = sepw * (1.0 + 0.5 * imb);
      if (score < bestscore) { bestscore = score; bestl = l; }
    }
    if (bestl < 0) {
      // fall back to the median level
      acc = 0;
      for (int l = 1; l < nl; l++) {
        acc += lw[l - 1];
        if (2 * acc >= total) { bestl = l; break; }
      }
      if (bestl < 0 || bestl >= nl) bestl = nl / 2;
    }
    int L = bestl;
    // sides: levels < L -> 0 (A), >= L -> 1 (B)
    for (int l = 0; l < nl; l++)
      for (int q = lp[l]; q < lp[l + 1]; q++) side_[ordr[q]] = (l < L) ? 0 : 1;
    // bipartite cut graph between level L-1 (left) and level L (right)
    std::vector<int> left, right;
    for (int q = lp[L - 1]; q < lp[L]; q++) left.push_back(ordr[q]);
    for (int q = lp[L]; q < lp[L + 1]; q++) right.push_back(ordr[q]);
    std::vector<int> cover;
    min_vertex_cover(left, right, st, cover);
    int cst = ++stamp_;
    for (int v : cover) mark_[v] = cst;
    A.clear(); B.clear(); S.clear();
    int64_t sw = 0;
    for (int q = 0; q < (int)ordr.size(); q++) {
      int v = ordr[q];
      if (mark_[v] == cst) { S.push_back(v); sw += w_[v]; }
      else if (side_[v] == 0) A.push_back(v);
      else B.push_back(v);
    }
    if (A.empty() || B.empty()) return INT64_MAX;
    return sw;
  }

  // Koenig: minimum vertex cover of the bipartite graph formed by the edges
  // between `left` and `right` (restricted to tagged vertices).
  void min_vertex_cover(const std::vector<int>& left, const std::vector<int>& right, int st,
                        std::vector<int>& cover) {
    int nlft = (int)left.size(), nrgt = (int)right.size();
    int lst = ++stamp_;
    // local ids: left i -> i, right j -> j (stored in dist_ temporarily via mark)
    std::vector<int> rid_of;  // map vertex -> right index via dist_ reuse is unsafe; use hash
    // use mark_ as membership stamp for right and dist_ as index
    for (int j = 0; j < nrgt; j++) { mark_[right[j]] = lst; dist_[right[j]] = j; }
    std::vector<int64_t> eptr(nlft + 1, 0);
    std::vector<int> eadj;
    for (int i = 0; i < nlft; i++) {
      int v = left[i];
      for (int64_t p = g_.ptr[v]; p < g_.ptr[v + 1]; p++) {
        int u = g_.adj[p];
        if (tag_[u] == st && mark_[u] == lst) eadj.push_back(dist_[u]);
      }
      eptr[i + 1] = (int64_t)eadj.size();
    }
    std::vector<int> matchL(nlft, -1), matchR(nrgt, -1);
    // greedy init
    for (int i = 0; i < nlft; i++)
      for (int64_t p = eptr[i]; p < eptr[i + 1]; p++)
        if (matchR[eadj[p]] < 0) { matchR[eadj[p]] = i; matchL[i] = eadj[p]; break; }
    // augmenting paths (iterative DFS)
    std::vector<int> visR(nrgt, -1);
    std::vector<int> stkL, stkP;
    std::vector<int64_t> itp(nlft);
    std::vector<int> viaR(nlft, -1);
    for (int s = 0; s < nlft; s++) {
      if (matchL[s] >= 0 || eptr[s] == eptr[s + 1]) continue;
      stkL.clear();
      stkL.push_back(s);
      itp[s] = eptr[s];
      bool found = false;
      while (!stkL.empty() && !found) {
        int i = stkL.back();
        if (itp[i] >= eptr[i + 1]) { stkL.pop_back(); continue; }
        int j = eadj[itp[i]++];
        if (visR[j] == s) continue;
        visR[j] = s;
        if (matchR[j] < 0) {
          // augment along the stack
          int jj = j;
          for (int d = (int)stkL.size() - 1; d >= 0; d--) {
            int li = stkL[d];
            int prev = matchL[li];
            matchL[li] = jj;
            matchR[jj] = li;
            jj = prev;
          }
          found = true;
        } else {
          int nl2 = matchR[j];
          itp[nl2] = eptr[nl2];
          stkL.push_back(nl2);
        }
      }
    }
    // alternating reachability from unmatched left vertices
    std::vector<char> zL(nlft, 0), zR(nrgt, 0);
    std::vector<int> q;
    for (int i = 0; i < nlft; i++)
      if (matchL[i] < 0) { zL[i] = 1; q.push_back(i); }
    for (size_t h = 0; h < q.size(); h++) {
      int i = q[h];
      for (int64_t p = eptr[i]; p < eptr[i + 1]; p++) {
        int j = eadj[p];
        if (matchL[i] == j) continue;
        if (!zR[j]) {
          zR[j] = 1;
          int i2 = matchR[j];
          if (i2 >= 0 && !zL[i2]) { zL[i2] = 1; q.push_back(i2); }
        }
      }
    }
    cover.clear();
    for (int i = 0; i < nlft; i++)
      if (!zL[i] && eptr[i + 1] > eptr[i]) cover.push_back(left[i]);
    for (int j = 0; j < nrgt; j++)
      if (zR[j]) cover.push_back(right[j]);
  }

  bool bisect(const std::vector<int>& V, int st, std::vector<int>& A, std::vector<int>& B,
              std::vector<int>& S) {
    int64_t total = 0;
    for (int v : V) total += w_[v];
    std::vector<int> ordr, lp;
    int u = pseudo_peripheral(V[0], st, ordr, lp);
    (void)u;
    std::vector<int> A1, B1, S1;
    int64_t s1 = level_cut(ordr, lp, st, total, A1, B1, S1);
    // second candidate: BFS from the far end
    int far = ordr.back();
    {
      int best = far, bd = 1 << 30;
      for (int q = lp[lp.size() - 2]; q < lp.back(); q++)
        if (g_.deg(ordr[q]) < bd) { bd = g_.deg(ordr[q]); best = ordr[q]; }
      far = best;
    }
    std::vector<int> o2, l2;
    bfs(far, st, o2, l2);
    std::vector<int> A2, B2, S2;
    int64_t s2 = level_cut(o2, l2, st, total, A2, B2, S2);
    if (s1 == INT64_MAX && s2 == INT64_MAX) return false;
    if (s2 < s1) { A.swap(A2); B.swap(B2); S.swap(S2); }
    else { A.swap(A1); B.swap(B1); S.swap(S1); }
    return true;
  }

  // minimum degree on a small vertex set (<= 64) using bit masks
  void leaf_order(const std::vector<int>& V, std::vector<int>& order) {
    int m = (int)V.size();
    if (m <= 2 || m > 64) { order.insert(order.end(), V.begin(), V.end()); return; }
    int lst = ++stamp_;
    for (int i = 0; i < m; i++) { mark_[V[i]] = lst; dist_[V[i]] = i; }
    std::vector<uint64_t> nb(m, 0);
    std::vector<int64_t> ext(m, 0);
    for (int i = 0; i < m; i++) {
      int v = V[i];
      for (int64_t p = g_.ptr[v]; p < g_.ptr[v + 1]; p++) {
        int u = g_.adj[p];
        if (mark_[u] == lst) nb[i] |= (1ULL << dist_[u]);
        else ext[i] += w_[u];
      }
    }
    uint64_t alive = (m == 64) ? ~0ULL : ((1ULL << m) - 1);
    for (int step = 0; step < m; step++) {
      int best = -1;
      int64_t bd = INT64_MAX;
      for (int i = 0; i < m; i++) {
        if (!((alive >> i) & 1ULL)) continue;
        uint64_t msk = nb[i] & alive;
        int64_t d = ext[i];
        while (msk) { int b = __builtin_ctzll(msk); d += w_[V[b]]; msk &= msk - 1; }
        if (d < bd) { bd = d; best = i; }
      }
      alive &= ~(1ULL << best);
      uint64_t msk = nb[best] & alive;
      uint64_t m2 = msk;
      while (m2) {
        int b = __builtin_ctzll(m2);
        nb[b] |= msk & ~(1ULL << b);
        ext[b] = std::max(ext[b], ext[best]);
        m2 &= m2 - 1;
      }
      order.push_back(V[best]);
    }
  }
};

}  // namespace

int validate_pattern(int n, int nnz, const int* rowOffs, const int* colIdxs, std::string& err) {
  if (n <= 0) { err = "matrix size must be positive"; return -1; }
  if (!rowOffs || !colIdxs) { err = "null pattern array"; return -1; }
  if (rowOffs[0] != 0) { err = "rowOffs[0] must be 0 (0-based CSR)"; return -1; }
  for (int i = 0; i < n; i++)
    if (rowOffs[i + 1] < rowOffs[i]) { err = "rowOffs not monotone"; return -1; }
  if (rowOffs[n] != nnz) { err = "rowOffs[n] != nnz"; return -1; }
  for (int64_t p = 0; p < nnz; p++)
    if (colIdxs[p] < 0 || colIdxs[p] >= n) { err = "column index out of range"; return -1; }
  // duplicate check per row (small rows: sort a copy)
  std::vector<int> tmp;
  for (int i = 0; i < n; i++) {
    tmp.assign(colIdxs + rowOffs[i], colIdxs + rowOffs[i + 1]);
    std::sort(tmp.begin(), tmp.end());
    for (size_t q = 1; q < tmp.size(); q++)
      if (tmp[q] == tmp[q - 1]) { err = "duplicate entry in a row"; return -1; }
  }
  return 0;
}

int symbolic_analyze(int n, const int* rowOffs, const int* colIdxs, const unsigned char* delay,
                     const SymOptions& opt, Symbolic& S, std::string& err) {
  S = Symbolic();
  S.n = n;
  Graph g;
  build_graph(n, rowOffs, colIdxs, g);

  std::vector<char> isdel(n, 0);
  if (delay)
    for (int i = 0; i < n; i++) isdel[i] = delay[i] ? 1 : 0;

  // ---------------- ordering ----------------
  std::vector<int> order;  // order[k] = original vertex eliminated k-th
  order.reserve(n);
  if (opt.ordering == 0) {
    for (int i = 0; i < n; i++)
      if (!isdel[i]) order.push_back(i);
  } else {
    // supervariables: identical closed adjacency and same delay flag
    std::vector<uint64_t> h(n);
    for (int v = 0; v < n; v++) {
      uint64_t s = mix64((uint64_t)v) + (isdel[v] ? 0x5555ULL : 0);
      for (int64_t p = g.ptr[v]; p < g.ptr[v + 1]; p++) s += mix64((uint64_t)g.adj[p]);
      h[v] = s;
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int a, int b) {
      if (h[a] != h[b]) return h[a] < h[b];
      return a < b;
    });
    std::vector<int> sv(n, -1);  // supervariable id of each vertex
    std::vector<int> rep;        // representative vertex of each supervariable
    auto closed_equal = [&](int a, int b) {
      if (g.deg(a) != g.deg(b) || isdel[a] != isdel[b]) return false;
      // closed lists: adj(a) + a  vs adj(b) + b
      std::vector<int> la(g.adj.begin() + g.ptr[a], g.adj.begin() + g.ptr[a + 1]);
      std::vector<int> lb(g.adj.begin() + g.ptr[b], g.adj.begin() + g.ptr[b + 1]);
      la.push_back(a);
      lb.push_back(b);
      std::sort(la.begin(), la.end());
      std::sort(lb.begin(), lb.end());
      return la == lb;
    };
    for (int q = 0; q < n;) {
      int r = q;
      while (r < n && h[idx[r]] == h[idx[q]]) r++;
      // group [q, r) share a hash: cluster exact matches
      for (int a = q; a < r; a++) {
        int va = idx[a];
        if (sv[va] >= 0) continue;
        sv[va] = (int)rep.size();
        rep.push_back(va);
        for (int b = a + 1; b < r; b++) {
          int vb = idx[b];
          if (sv[vb] < 0 && closed_equal(va, vb)) sv[vb] = sv[va];
        }
      }
      q = r;
    }
    int nsv = (int)rep.size();
    std::vector<std::vector<int>> members(nsv);
    for (int v = 0; v < n; v++) members[sv[v]].push_back(v);
    // compressed graph on non-delayed supervariables
    std::vector<int> cid(nsv, -1);
    std::vector<int> cmap;
    for (int s = 0; s < nsv; s++)
      if (!isdel[rep[s]]) { cid[s] = (int)cmap.size(); cmap.push_back(s); }
    int nc = (int)cmap.size();
    Graph cg;
    cg.n = nc;
    std::vector<std::vector<int>> cadj(nc);
    std::vector<int> stamp(nc, -1);
    for (int c = 0; c < nc; c++) {
      int v = rep[cmap[c]];
      stamp[c] = c;
      for (int64_t p = g.ptr[v]; p < g.ptr[v + 1]; p++) {
        int u = g.adj[p];
        int cu = cid[sv[u]];
        if (cu >= 0 && stamp[cu] != c) { stamp[cu] = c; cadj[c].push_back(cu); }
      }
    }
    // cliques induced by delayed rows (the pattern of S + G^T G)
    for (int d = 0; d < n; d++) {
      if (!isdel[d]) continue;
      std::vector<int> nbs;
      for (int64_t p = g.ptr[d]; p < g.ptr[d + 1]; p++) {
        int cu = cid[sv[g.adj[p]]];
        if (cu >= 0) nbs.push_back(cu);
      }
      std::sort(nbs.begin(), nbs.end());
      nbs.erase(std::unique(nbs.begin(), nbs.end()), nbs.end());
      if (nbs.size() > 64) continue;
      for (size_t a = 0; a < nbs.size(); a++)
        for (size_t b = 0; b < nbs.size(); b++)
          if (a != b) cadj[nbs[a]].push_back(nbs[b]);
    }
    cg.ptr.assign(nc + 1, 0);
    for (int c = 0; c < nc; c++) {
      std::sort(cadj[c].begin(), cadj[c].end());
      cadj[c].erase(std::unique(cadj[c].begin(), cadj[c].end()), cadj[c].end());
      cg.ptr[c + 1] = cg.ptr[c] + (int64_t)cadj[c].size();
    }
    cg.adj.reserve(cg.ptr[nc]);
    for (int c = 0; c < nc; c++) cg.adj.insert(cg.adj.end(), cadj[c].begin(), cadj[c].end());
    std::vector<int> wgt(nc);
    for (int c = 0; c < nc; c++) wgt[c] = (int)members[cmap[c]].size();
    std::vector<int> corder;
    if (nc > 0) {
      NestedDissection nd(cg, wgt, opt.leaf_size);
      nd.run(corder);
    }
    for (int c : corder)
      for (int v : members[cmap[c]]) order.push_back(v);
  }
  // insert delayed rows right after their last-ordered neighbour
  {
    std::vector<int> pos(n, -1);
    for (int k = 0; k < (int)order.size(); k++) pos[order[k]] = k;
    std::vector<std::pair<int, int>> dl;  // (after position, vertex)
    for (int d = 0; d < n; d++) {
      if (!isdel[d]) continue;
      int last = -1;
      for (int64_t p = g.ptr[d]; p < g.ptr[d + 1]; p++) {
        int u = g.adj[p];
        if (!isdel[u]) last = std::max(last, pos[u]);
      }
      if (last < 0) last = (int)order.size() - 1;
      dl.push_back({last, d});
      S.ndelayed++;
    }
    std::sort(dl.begin(), dl.end());
    std::vector<int> full;
    full.reserve(n);
    size_t q = 0;
    while (q < dl.size() && dl[q].first < 0) full.push_back(dl[q++].second);
    for (int k = 0; k < (int)order.size(); k++) {
      full.push_back(order[k]);
      while (q < dl.size() && dl[q].first == k) full.push_back(dl[q++].second);
    }
    while (q < dl.size()) full.push_back(dl[q++].second);
    order.swap(full);
  }
  if ((int)order.size() != n) { err = "internal: ordering size mismatch"; return -5; }

  // ---------------- elimination tree + postorder ----------------
  std::vector<int> iperm(n);
  for (int k = 0; k < n; k++) iperm[order[k]] = k;
  auto etree = [&](const std::vector<int>& pm, const std::vector<int>& ipm, std::vector<int>& par) {
    par.assign(n, -1);
    std::vector<int> anc(n, -1);
    for (int j = 0; j < n; j++) {
      int v = pm[j];
      for (int64_t p = g.ptr[v]; p < g.ptr[v + 1]; p++) {
        int i = ipm[g.adj[p]];
        if (i >= j) continue;
        int r = i;
        while (anc[r] != -1 && anc[r] != j) { int t = anc[r]; anc[r] = j; r = t; }
        if (anc[r] == -1) { anc[r] = j; par[r] = j; }
      }
    }
  };
  std::vector<int> par;
  etree(order, iperm, par);
  // postorder (children visited in increasing order)
  std::vector<int> head(n, -1), next(n, -1);
  for (int j = n - 1; j >= 0; j--)
    if (par[j] >= 0) { next[j] = head[par[j]]; head[par[j]] = j; }
  std::vector<int> post;
  post.reserve(n);
  {
    std::vector<int> stk;
    for (int r = 0; r < n; r++) {
      if (par[r] != -1) continue;
      stk.push_back(r);
      while (!stk.empty()) {
        int v = stk.back();
        if (head[v] >= 0) {
          int c = head[v];
          head[v] = next[c];
          stk.push_back(c);
        } else {
          stk.pop_back();
          post.push_back(v);
        }
      }
    }
  }
  S.perm.resize(n);
  for (int k = 0; k < n; k++) S.perm[k] = order[post[k]];
  S.iperm.resize(n);
  for (int k = 0; k < n; k++) S.iperm[S.perm[k]] = k;
  etree(S.perm, S.iperm, S.parent);

  // ---------------- column counts (row-subtree traversal) ----------------
  S.colcount.assign(n, 1);
  {
    std::vector<int> mark(n, -1);
    for (int i = 0; i < n; i++) {
      mark[i] = i;
      int v = S.perm[i];
      for (int64_t p = g.ptr[v]; p < g.ptr[v + 1]; p++) {
        int k = S.iperm[g.adj[p]];
        if (k >= i) continue;
        while (mark[k] != i) { S.colcount[k]++; mark[k] = i; k = S.parent[k]; }
      }
    }
  }
  S.nnzL_exact = 0;
  for (int j = 0; j < n; j++) S.nnzL_exact += S.colcount[j];

  // ---------------- fundamental (maximal) supernodes ----------------
  std::vector<int> fsup;  // first column of each
  fsup.push_back(0);
  for (int j = 1; j < n; j++) {
    bool join = (S.parent[j - 1] == j) && (S.colcount[j - 1] == S.colcount[j] + 1);
    if (!join) fsup.push_back(j);
  }
  int nf = (int)fsup.size();
  fsup.push_back(n);
  // relaxed amalgamation
  std::vector<int> first(nf), kk(nf);
  std::vector<int64_t> rr(nf), stored(nf), zeros(nf);
  std::vector<int> fpar(nf, -1);
  std::vector<int> c2f(n);
  for (int s = 0; s < nf; s++)
    for (int j = fsup[s]; j < fsup[s + 1]; j++) c2f[j] = s;
  for (int s = 0; s < nf; s++) {
    first[s] = fsup[s];
    kk[s] = fsup[s + 1] - fsup[s];
    rr[s] = S.colcount[fsup[s]] - kk[s];
    stored[s] = (int64_t)kk[s] * (kk[s] + 1) / 2 + (int64_t)kk[s] * rr[s];
    zeros[s] = 0;
    int lastc = fsup[s + 1] - 1;
    fpar[s] = S.parent[lastc] >= 0 ? c2f[S.parent[lastc]] : -1;
  }
  std::vector<int> alive(nf, 1);
  std::vector<int> endcol_owner(n + 1, -1);  // supernode whose last column is j
  for (int s = 0; s < nf; s++) endcol_owner[fsup[s + 1] - 1] = s;
  // merge map
  std::vector<int> mergedInto(nf, -1);
  for (int p = 0; p < nf; p++) {
    while (true) {
      int fc = first[p];
      if (fc == 0) break;
      int c = endcol_owner[fc - 1];
      if (c < 0 || !alive[c]) break;
      // c must be a child of p (supernodal tree)
      int cp = fpar[c];
      while (cp >= 0 && mergedInto[cp] >= 0) cp = mergedInto[cp];
      if (cp != p) break;
      int64_t k = kk[c] + kk[p];
      int64_t newst = k * (k + 1) / 2 + k * rr[p];
      int64_t z = newst - (stored[c] - zeros[c]) - (stored[p] - zeros[p]);
      double zf = newst > 0 ? (double)z / (double)newst : 0.0;
      bool ok = (k <= opt.nrelax0) || (k <= opt.nrelax1 && zf < opt.zrelax0) ||
                (k <= opt.nrelax2 && zf < opt.zrelax1) || (zf < opt.zrelax2);
      if (!ok) break;
      alive[c] = 0;
      mergedInto[c] = p;
      endcol_owner[fsup[c + 1] - 1] = -1;
      first[p] = first[c];
      kk[p] = (int)k;
      stored[p] = newst;
      zeros[p] = z;
    }
  }
  S.sup_ptr.clear();
  for (int s = 0; s < nf; s++) {
    if (!alive[s]) continue;
    // split supernodes wider than max_super into a chain of chunks
    int f0 = first[s], f1 = fsup[s + 1];
    for (int c = f0; c < f1; c += std::max(1, opt.max_super)) S.sup_ptr.push_back(c);
  }
  S.nsuper = (int)S.sup_ptr.size();
  S.sup_ptr.push_back(n);
  // supernodes must tile [0,n) in order
  for (int s = 0; s < S.nsuper; s++)
    if (S.sup_ptr[s] >= S.sup_ptr[s + 1]) { err = "internal: supernode partition"; return -5; }
  S.col2sup.resize(n);
  for (int s = 0; s < S.nsuper; s++)
    for (int j = S.sup_ptr[s]; j < S.sup_ptr[s + 1]; j++) S.col2sup[j] = s;

  // ---------------- supernode row structures ----------------
  int ns = S.nsuper;
  S.rows_ptr.assign(ns + 1, 0);
  S.rows.clear();
  S.sparent.assign(ns, -1);
  std::vector<int> shead(ns, -1), snext(ns, -1);
  std::vector<int> mark(n, -1);
  std::vector<int> tmp;
  for (int s = 0; s < ns; s++) {
    int f = S.sup_ptr[s], l = S.sup_ptr[s + 1] - 1;
    tmp.clear();
    for (int c = f; c <= l; c++) {
      int v = S.perm[c];
      for (int64_t p = g.ptr[v]; p < g.ptr[v + 1]; p++) {
        int i = S.iperm[g.adj[p]];
        if (i > l && mark[i] != s) { mark[i] = s; tmp.push_back(i); }
      }
    }
    for (int ch = shead[s]; ch >= 0; ch = snext[ch]) {
      for (int64_t q = S.rows_ptr[ch]; q < S.rows_ptr[ch + 1]; q++) {
        int i = S.rows[q];
        if (i > l && mark[i] != s) { mark[i] = s; tmp.push_back(i); }
      }
    }
    std::sort(tmp.begin(), tmp.end());
    S.rows.insert(S.rows.end(), tmp.begin(), tmp.end());
    S.rows_ptr[s + 1] = (int64_t)S.rows.size();
    if (!tmp.empty()) {
      int p = S.col2sup[tmp[0]];
      S.sparent[s] = p;
      if (p <= s) { err = "internal: assembly tree order"; return -5; }
      // prepend to children list of p (children are listed in decreasing order;
      // consumers re-sort as needed)
      snext[s] = shead[p];
      shead[p] = s;
    }
  }
  // levels and statistics
  S.level.assign(ns, 0);
  S.nnzL = 0;
  S.flops = 0;
  for (int s = 0; s < ns; s++) {
    int p = S.sparent[s];
    if (p >= 0) S.level[p] = std::max(S.level[p], S.level[s] + 1);
    double k = S.sup_ptr[s + 1] - S.sup_ptr[s];
    double r = (double)(S.rows_ptr[s + 1] - S.rows_ptr[s]);
    S.nnzL += (int64_t)(k * (k + 1) / 2 + k * r);
    S.flops += k * k * k / 3.0 + r * k * k + r * r * k;
  }
  S.nlevels = 0;
  for (int s = 0; s < ns; s++) S.nlevels = std::max(S.nlevels, S.level[s] + 1);
  return 0;
}

}  // namespace kkt
