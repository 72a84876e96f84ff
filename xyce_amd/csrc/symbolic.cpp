// Sparse LU symbolic analysis for the GPU refactor.
//
// Replaces KLU's analyse step as Xyce calls it (AmesosSolver::doSolve first
// call -> Amesos_Klu SymbolicFactorization, src/LinearAlgebraServicesPKG/
// N_LAS_AmesosSolver.C:335): a maximum transversal gives a zero-free diagonal
// (KLU's BTF step), then a symmetric fill-reducing order is chosen on the
// pattern of A+A^T.  The order is chosen for the GPU: multiple-minimum-degree
// elimination of independent sets (small systems) and nested dissection with
// level-set separators (large systems) both give shallow elimination trees, so
// the numeric refactor runs level by level.  Rows/columns denser than
// 10*sqrt(n) are ordered last, as AMD does with dense rows.
//
// Pivots are static (the diagonal of the permuted matrix); the numeric kernels
// check every pivot and xg_lu_numeric re-analyses with the values when one
// fails, which stands in for KLU's re-pivoting at every factorisation
// (repivot_, N_LAS_AmesosSolver.C:87,318).
#include <algorithm>
#include <cmath>
#include <deque>
#include <numeric>
#include <stdexcept>

#include "xg_internal.h"

namespace xg {

namespace {

// Maximum transversal: match[col] = row with A(row,col) != 0, preferring the
// diagonal, then (if values are given) the largest entry of the column.
std::vector<int> max_transversal(int n, const int* rowptr, const int* colind, const double* values) {
  // column-wise pattern
  std::vector<int> cptr(n + 1, 0), crow, cpos;
  for (int r = 0; r < n; ++r) for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) cptr[colind[k] + 1]++;
  for (int c = 0; c < n; ++c) cptr[c + 1] += cptr[c];
  crow.resize(cptr[n]); cpos.resize(cptr[n]);
  {
    std::vector<int> f(cptr.begin(), cptr.end() - 1);
    for (int r = 0; r < n; ++r)
      for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) { int c = colind[k]; crow[f[c]] = r; cpos[f[c]++] = k; }
  }
  // order candidate rows of each column by |value| (descending)
  if (values) {
    for (int c = 0; c < n; ++c) {
      std::vector<std::pair<double, int>> v;
      for (int k = cptr[c]; k < cptr[c + 1]; ++k) v.push_back({-std::fabs(values[cpos[k]]), crow[k]});
      std::stable_sort(v.begin(), v.end());
      for (int k = cptr[c]; k < cptr[c + 1]; ++k) crow[k] = v[k - cptr[c]].second;
    }
  }
  std::vector<int> match_row(n, -1), match_col(n, -1);
  auto nonzero = [&](int r, int c) {
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k)
      if (colind[k] == c) return !values || values[k] != 0.0;
    return false;
  };
  for (int c = 0; c < n; ++c)
    if (nonzero(c, c)) { match_col[c] = c; match_row[c] = c; }
  // When values are given, a diagonal that is tiny relative to its column is
  // released so the augmenting search can pick a better row.
  if (values) {
    for (int c = 0; c < n; ++c) {
      if (match_col[c] != c) continue;
      double dmax = 0, dg = 0;
      for (int k = cptr[c]; k < cptr[c + 1]; ++k) {
        double a = std::fabs(values[cpos[k]]);
        dmax = std::max(dmax, a);
      }
      for (int k = rowptr[c]; k < rowptr[c + 1]; ++k) if (colind[k] == c) dg = std::fabs(values[k]);
      if (dg < 1e-3 * dmax) { match_col[c] = -1; match_row[c] = -1; }
    }
  }
  // augmenting paths (iterative DFS over columns; rows_on_path[i] is the row
  // tried from column stk[i]).  Each column entered first looks for a free
  // row among its own entries (MC21's look-ahead), so short paths win over
  // deep ones: a voltage-source branch pairs with its node in two hops
  // instead of shifting a chain of diagonal pivots across the circuit.
  std::vector<int> visited(n, -1), rows_on_path;
  std::vector<std::pair<int, int>> stk;
  auto free_row = [&](int c) {
    for (int k = cptr[c]; k < cptr[c + 1]; ++k)
      if (match_row[crow[k]] < 0) return crow[k];
    return -1;
  };
  for (int c0 = 0; c0 < n; ++c0) {
    if (match_col[c0] >= 0) continue;
    stk.assign(1, {c0, cptr[c0]});
    rows_on_path.clear();
    bool found = false;
    {
      const int r = free_row(c0);
      if (r >= 0) { rows_on_path.push_back(r); found = true; }
    }
    while (!found && !stk.empty()) {
      const int c = stk.back().first, k = stk.back().second;
      if (k >= cptr[c + 1]) {
        stk.pop_back();
        if (!rows_on_path.empty()) rows_on_path.pop_back();
        continue;
      }
      stk.back().second = k + 1;
      const int r = crow[k];
      if (visited[r] == c0) continue;
      visited[r] = c0;
      rows_on_path.push_back(r);
      if (match_row[r] < 0) { found = true; break; }
      const int c2 = match_row[r];
      stk.push_back({c2, cptr[c2]});
      const int r2 = free_row(c2);
      if (r2 >= 0) { rows_on_path.push_back(r2); found = true; break; }
    }
    if (!found) throw std::runtime_error("structurally singular matrix");
    for (size_t i = 0; i < stk.size(); ++i) {
      const int c = stk[i].first, r = rows_on_path[i];
      match_col[c] = r; match_row[r] = c;
    }
  }
  return match_col;
}

using Graph = std::vector<std::vector<int>>;

// Multiple minimum degree on an explicit elimination graph: each round
// eliminates an independent set of minimum-degree nodes, so a round is one
// level of the elimination tree.
void mmd_order(const Graph& g0, const std::vector<int>& nodes, std::vector<int>& order) {
  const int m = (int)nodes.size();
  if (m == 0) return;
  std::vector<int> loc(g0.size(), -1);
  for (int i = 0; i < m; ++i) loc[nodes[i]] = i;
  Graph g(m);
  for (int i = 0; i < m; ++i) {
    for (int a : g0[nodes[i]]) if (loc[a] >= 0 && a != nodes[i]) g[i].push_back(loc[a]);
    std::sort(g[i].begin(), g[i].end());
    g[i].erase(std::unique(g[i].begin(), g[i].end()), g[i].end());
  }
  std::vector<char> gone(m, 0), mark(m, 0);
  int left = m;
  while (left > 0) {
    int dmin = m + 1;
    for (int i = 0; i < m; ++i) if (!gone[i]) dmin = std::min(dmin, (int)g[i].size());
    std::vector<int> sel;
    std::fill(mark.begin(), mark.end(), 0);
    for (int i = 0; i < m; ++i) {
      if (gone[i] || (int)g[i].size() != dmin || mark[i]) continue;
      sel.push_back(i);
      mark[i] = 1;
      for (int a : g[i]) mark[a] = 1;
    }
    for (int v : sel) {
      order.push_back(nodes[v]);
      gone[v] = 1; --left;
      const std::vector<int> nb = g[v];
      for (int a : nb) {
        auto& ga = g[a];
        ga.erase(std::remove(ga.begin(), ga.end(), v), ga.end());
        std::vector<int> merged;
        merged.reserve(ga.size() + nb.size());
        std::set_union(ga.begin(), ga.end(), nb.begin(), nb.end(), std::back_inserter(merged));
        merged.erase(std::remove(merged.begin(), merged.end(), a), merged.end());
        ga.swap(merged);
      }
      g[v].clear();
    }
  }
}

// Nested dissection with BFS level-set separators (George's automatic
// nested dissection) down to leaves ordered by MMD.
// Approximate minimum degree ordering on the quotient graph, after
// Amestoy, Davis and Duff, "An approximate minimum degree ordering
// algorithm", SIAM J. Matrix Anal. Appl. 17(4), 1996 (the ordering KLU uses,
// here without supervariable detection).  Eliminated pivots become elements
// standing for the cliques they created; a variable's degree is the
// approximate external degree of that paper; elements covered by the new
// one are absorbed.
void amd_order(const Graph& g0, const std::vector<int>& nodes, std::vector<int>& order) {
  const int m = (int)nodes.size();
  if (m == 0) return;
  std::vector<int> loc(g0.size(), -1);
  for (int i = 0; i < m; ++i) loc[nodes[i]] = i;
  std::vector<std::vector<int>> A(m), E(m), L(m);
  for (int i = 0; i < m; ++i)
    for (int a : g0[nodes[i]])
      if (loc[a] >= 0 && loc[a] != i) A[i].push_back(loc[a]);
  for (auto& v : A) { std::sort(v.begin(), v.end()); v.erase(std::unique(v.begin(), v.end()), v.end()); }
  enum { VAR = 0, ELEM = 1, DEAD = 2 };
  std::vector<char> state(m, VAR);
  std::vector<int> deg(m), head(m + 1, -1), nxt(m, -1), prv(m, -1);
  auto bucket_add = [&](int i, int d) {
    deg[i] = d; prv[i] = -1; nxt[i] = head[d];
    if (head[d] >= 0) prv[head[d]] = i;
    head[d] = i;
  };
  auto bucket_del = [&](int i) {
    if (prv[i] >= 0) nxt[prv[i]] = nxt[i]; else head[deg[i]] = nxt[i];
    if (nxt[i] >= 0) prv[nxt[i]] = prv[i];
  };
  for (int i = 0; i < m; ++i) bucket_add(i, (int)A[i].size());
  std::vector<int> mark(m, -1), wflag(m, -1), w(m, 0), Lp;
  int mindeg = 0, nleft = m;
  for (int step = 0; step < m; ++step) {
    while (head[mindeg] < 0) ++mindeg;
    const int p = head[mindeg];
    bucket_del(p);
    order.push_back(nodes[p]);
    --nleft;
    // the new element: p's variables and those of its elements (absorbed)
    Lp.clear();
    mark[p] = step;
    for (int j : A[p])
      if (state[j] == VAR && mark[j] != step) { mark[j] = step; Lp.push_back(j); }
    for (int e : E[p]) {
      if (state[e] != ELEM) continue;
      for (int j : L[e])
        if (state[j] == VAR && mark[j] != step) { mark[j] = step; Lp.push_back(j); }
      state[e] = DEAD;
      std::vector<int>().swap(L[e]);
    }
    state[p] = ELEM;
    std::vector<int>().swap(A[p]);
    std::vector<int>().swap(E[p]);
    L[p] = Lp;
    // |Le \ Lp| for the elements next to the new element's variables
    for (int i : Lp)
      for (int e : E[i]) {
        if (state[e] != ELEM || e == p) continue;
        if (wflag[e] != step) { wflag[e] = step; w[e] = (int)L[e].size(); }
        --w[e];
      }
    const int lp = (int)Lp.size();
    for (int i : Lp) {
      // prune variables now reached through the new element
      int k = 0;
      for (int j : A[i])
        if (state[j] == VAR && mark[j] != step) A[i][k++] = j;
      A[i].resize(k);
      int ext = 0;
      k = 0;
      for (int e : E[i]) {
        if (state[e] != ELEM || e == p) continue;
        if (wflag[e] == step && w[e] == 0) { state[e] = DEAD; std::vector<int>().swap(L[e]); continue; }  // e inside Lp
        E[i][k++] = e;
        ext += wflag[e] == step ? w[e] : (int)L[e].size();
      }
      E[i].resize(k);
      E[i].push_back(p);
      long long d = (long long)A[i].size() + (lp - 1) + ext;
      d = std::min<long long>(d, (long long)deg[i] + lp - 1);
      d = std::min<long long>(d, nleft - 1);
      if (d < 0) d = 0;
      bucket_del(i);
      bucket_add(i, (int)d);
      if (d < mindeg) mindeg = (int)d;
    }
  }
}

// Height of the elimination tree of the symmetric pattern g restricted to
// `order` (eliminated in that sequence; Liu's algorithm with path compression).
int etree_height(const Graph& g, const std::vector<int>& order) {
  const int m = (int)order.size();
  std::vector<int> pos(g.size(), -1);
  for (int i = 0; i < m; ++i) pos[order[i]] = i;
  std::vector<int> parent(m, -1), anc(m, -1);
  for (int i = 0; i < m; ++i) {
    for (int a : g[order[i]]) {
      int k = pos[a];
      if (k < 0 || k >= i) continue;
      while (k != -1 && k < i) {
        const int nx = anc[k];
        anc[k] = i;
        if (nx == -1) { parent[k] = i; break; }
        k = nx;
      }
    }
  }
  std::vector<int> h(m, 0);
  int best = 0;
  for (int i = 0; i < m; ++i) {
    if (parent[i] >= 0) h[parent[i]] = std::max(h[parent[i]], h[i] + 1);
    best = std::max(best, h[i]);
  }
  return best + 1;
}

// Nested dissection by breadth-first level structures (George & Liu): from a
// pseudo-peripheral node, the level of smallest size near the half-way count
// separates the levels before it from the levels after it; the two parts are
// ordered recursively, then the separator.  Parts of at most `leaf` nodes get
// approximate minimum degree.  Gives a shallow elimination tree (about
// log2(n / leaf) separator levels) where minimum degree leaves long chains.
void nd_order(const Graph& g, std::vector<int> nodes, std::vector<int>& order, int leaf, std::vector<int>& tag,
              int& stamp, std::vector<int>& lvl) {
  if ((int)nodes.size() <= leaf) { amd_order(g, nodes, order); return; }
  const int id = ++stamp;
  for (int v : nodes) tag[v] = id;
  auto bfs = [&](int root, std::vector<int>& seq, std::vector<int>& lptr) {
    seq.clear(); lptr.assign(1, 0);
    seq.push_back(root);
    lvl[root] = 0;
    tag[root] = -id;
    size_t head = 0;
    while (head < seq.size()) {
      const size_t end = seq.size();
      for (; head < end; ++head)
        for (int a : g[seq[head]])
          if (tag[a] == id) { tag[a] = -id; lvl[a] = lvl[seq[head]] + 1; seq.push_back(a); }
      lptr.push_back((int)end);
    }
    if (lptr.back() != (int)seq.size()) lptr.push_back((int)seq.size());
    for (int v : seq) tag[v] = id;
  };
  std::vector<int> seq, lptr;
  int root = nodes[0];
  for (int it = 0; it < 3; ++it) {      // pseudo-peripheral node: deepest level, smallest degree
    bfs(root, seq, lptr);
    const int last = lptr[lptr.size() - 2];
    int best = seq[last];
    for (int q = last; q < (int)seq.size(); ++q)
      if (g[seq[q]].size() < g[best].size()) best = seq[q];
    if (best == root) break;
    root = best;
  }
  bfs(root, seq, lptr);
  if (seq.size() < nodes.size()) {
    // disconnected: the component and the rest are independent
    std::vector<int> rest;
    for (int v : seq) tag[v] = -id;
    for (int v : nodes) if (tag[v] == id) rest.push_back(v);
    nd_order(g, seq, order, leaf, tag, stamp, lvl);
    nd_order(g, rest, order, leaf, tag, stamp, lvl);
    return;
  }
  const int nl = (int)lptr.size() - 1;
  if (nl < 3) { amd_order(g, nodes, order); return; }
  const int half = (int)seq.size() / 2;
  int mid = 1;
  while (mid < nl - 1 && lptr[mid + 1] <= half) ++mid;
  // smallest level within a window around the middle that keeps both sides
  // at least a third of the nodes
  int sep = mid;
  for (int L = 1; L < nl - 1; ++L) {
    if (lptr[L] < (int)seq.size() / 3 || (int)seq.size() - lptr[L + 1] < (int)seq.size() / 3) continue;
    if (lptr[L + 1] - lptr[L] < lptr[sep + 1] - lptr[sep]) sep = L;
  }
  std::vector<int> A(seq.begin(), seq.begin() + lptr[sep]), B(seq.begin() + lptr[sep + 1], seq.end()), S;
  // separator nodes with no neighbour beyond the separator join the first part
  for (int q = lptr[sep]; q < lptr[sep + 1]; ++q) {
    const int v = seq[q];
    bool beyond = false;
    for (int a : g[v]) if (tag[a] == id && lvl[a] == sep + 1) { beyond = true; break; }
    (beyond ? S : A).push_back(v);
  }
  nd_order(g, std::move(A), order, leaf, tag, stamp, lvl);
  nd_order(g, std::move(B), order, leaf, tag, stamp, lvl);
  amd_order(g, S, order);
}

}  // namespace

int lu_analyse(int n, const int* rowptr, const int* colind, const double* values, LuSymbolic& s, bool want_schedule,
               bool keep_tree) {
  s = LuSymbolic();
  s.n = n; s.nnz_a = rowptr[n];
  std::vector<int> match = max_transversal(n, rowptr, colind, values);  // A'(c,:) = A(match[c],:)
  // A' pattern (row c of A' = row match[c] of A), symmetrised
  Graph g(n);
  for (int c = 0; c < n; ++c) {
    const int r = match[c];
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
      const int j = colind[k];
      if (j != c) { g[c].push_back(j); g[j].push_back(c); }
    }
  }
  for (auto& v : g) { std::sort(v.begin(), v.end()); v.erase(std::unique(v.begin(), v.end()), v.end()); }
  // dense nodes last
  const int dense_thresh = std::max(16, (int)(10.0 * std::sqrt((double)n)));
  std::vector<int> sparse_nodes, dense_nodes;
  for (int v = 0; v < n; ++v) ((int)g[v].size() > dense_thresh ? dense_nodes : sparse_nodes).push_back(v);
  std::vector<int> order;
  order.reserve(n);
  // small systems (the batched engine's classes): multiple minimum degree,
  // whose independent sets keep the level schedule shallow; larger: AMD
  if ((int)sparse_nodes.size() <= 256) mmd_order(g, sparse_nodes, order);
  else {
    amd_order(g, sparse_nodes, order);
    // a large system whose minimum-degree elimination tree is a long chain
    // (gate arrays wired with locality) is dissected instead: the GPU
    // factorisation proceeds one tree level after another
    if ((int)sparse_nodes.size() > 65535) {
      const int h_amd = etree_height(g, order);
      if (h_amd > 2000) {
        std::vector<int> nd, tag(n, 0), lvl(n, 0);
        int stamp = 0;
        nd_order(g, sparse_nodes, nd, 4096, tag, stamp, lvl);
        if ((int)nd.size() == (int)sparse_nodes.size() && etree_height(g, nd) * 4 < h_amd) order.swap(nd);
      }
    }
  }
  order.insert(order.end(), dense_nodes.begin(), dense_nodes.end());
  s.dense_tail = (int)dense_nodes.size();
  // B(i,j) = A'(order[i], order[j]) = A(match[order[i]], order[j])
  s.perm_row.resize(n); s.perm_col.resize(n);
  std::vector<int> pinv(n);
  for (int i = 0; i < n; ++i) { s.perm_col[i] = order[i]; s.perm_row[i] = match[order[i]]; pinv[order[i]] = i; }
  std::vector<int> rinv(n);
  for (int i = 0; i < n; ++i) rinv[s.perm_row[i]] = i;
  // symmetric pattern of B, lower part per row (for the etree / fill)
  std::vector<std::vector<int>> blow(n);
  for (int r = 0; r < n; ++r) {
    const int i = rinv[r];
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
      const int j = pinv[colind[k]];
      if (j < i) blow[i].push_back(j);
      else if (j > i) blow[j].push_back(i);
    }
  }
  // elimination tree (Liu) and row structure of L via ereach
  std::vector<int> parent(n, -1), ancestor(n, -1);
  for (int i = 0; i < n; ++i) {
    for (int j : blow[i]) {
      int k = j;
      while (k != -1 && k < i) {
        int next = ancestor[k];
        ancestor[k] = i;
        if (next == -1) { parent[k] = i; break; }
        k = next;
      }
    }
  }
  std::vector<std::vector<int>> lrow(n);
  {
    std::vector<int> mark(n, -1);
    for (int i = 0; i < n; ++i) {
      mark[i] = i;
      for (int j : blow[i]) {
        int k = j;
        while (k != -1 && k < i && mark[k] != i) { lrow[i].push_back(k); mark[k] = i; k = parent[k]; }
      }
      std::sort(lrow[i].begin(), lrow[i].end());
    }
  }
  blow.clear(); blow.shrink_to_fit();
  // U pattern = transpose of L pattern: urow[k] = { i : k in lrow[i] }
  std::vector<std::vector<int>> urow(n);
  for (int i = 0; i < n; ++i) for (int k : lrow[i]) urow[k].push_back(i);
  // row-major L+U storage
  s.lu_rowptr.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) s.lu_rowptr[i + 1] = s.lu_rowptr[i] + (int)lrow[i].size() + 1 + (int)urow[i].size();
  s.nlu = s.lu_rowptr[n];
  s.lu_colind.resize(s.nlu); s.lu_diag.resize(n);
  for (int i = 0; i < n; ++i) {
    int p = s.lu_rowptr[i];
    for (int k : lrow[i]) s.lu_colind[p++] = k;
    s.lu_diag[i] = p; s.lu_colind[p++] = i;
    for (int j : urow[i]) s.lu_colind[p++] = j;
  }
  auto pos = [&](int i, int j) -> int {
    const int* b = s.lu_colind.data() + s.lu_rowptr[i];
    const int* e = s.lu_colind.data() + s.lu_rowptr[i + 1];
    const int* f = std::lower_bound(b, e, j);
    if (f == e || *f != j) return -1;
    return (int)(f - s.lu_colind.data());
  };
  s.a2lu.resize(s.nnz_a);
  for (int r = 0; r < n; ++r)
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
      const int p = pos(rinv[r], pinv[colind[k]]);
      if (p < 0) throw std::runtime_error("A entry missing from LU pattern");
      s.a2lu[k] = p;
    }
  // flops: sum over pivots of 2*|L(:,k)|*|U(k,:)| + |L(:,k)|
  {
    double fl = 0;
    for (int k = 0; k < n; ++k) fl += 2.0 * (double)urow[k].size() * (double)urow[k].size() + (double)urow[k].size();
    s.flops = fl;
  }
  if (keep_tree) {
    s.etree = parent;
    s.lcol = urow;
    s.rinv = rinv;
    s.pinv = pinv;
  }
  if (!want_schedule) return XG_OK;

  // ---- entry-level factor schedule: target (i,j) += -L(i,k)*U(k,j), k ascending
  std::vector<int> level(s.nlu, 0);
  std::vector<std::vector<int>> tops(s.nlu);   // not stored per target to save memory: build CSR directly
  // count ops per target
  std::vector<int> opcnt(s.nlu, 0);
  for (int i = 0; i < n; ++i)
    for (int k : lrow[i])
      for (int j : urow[k]) { int p = pos(i, j); if (p >= 0) opcnt[p]++; }
  tops.clear();
  std::vector<int> op_ptr(s.nlu + 1, 0);
  for (int p = 0; p < s.nlu; ++p) op_ptr[p + 1] = op_ptr[p] + opcnt[p];
  std::vector<int> ops((size_t)2 * op_ptr[s.nlu]);
  {
    std::vector<int> fill(op_ptr.begin(), op_ptr.end() - 1);
    for (int i = 0; i < n; ++i)
      for (int k : lrow[i]) {
        const int pik = pos(i, k);
        for (int j : urow[k]) {
          const int p = pos(i, j);
          if (p < 0) continue;
          const int f = fill[p]++;
          ops[2 * f] = pik; ops[2 * f + 1] = pos(k, j);
        }
      }
    // the diagonal U(k,k)... ops with j==k do not occur (urow excludes k)
  }
  // also U(k,j) for j == i with i > k: covered (urow[k] contains i)
  // levels in row-major order
  std::vector<int> diag_of(s.nlu, -1);
  int maxlvl = 0;
  for (int i = 0; i < n; ++i) {
    for (int p = s.lu_rowptr[i]; p < s.lu_rowptr[i + 1]; ++p) {
      const int j = s.lu_colind[p];
      int lv = 0;
      for (int o = op_ptr[p]; o < op_ptr[p + 1]; ++o) lv = std::max(lv, std::max(level[ops[2 * o]], level[ops[2 * o + 1]]) + 1);
      if (j < i) { diag_of[p] = s.lu_diag[j]; lv = std::max(lv, level[s.lu_diag[j]] + 1); }
      level[p] = lv;
      maxlvl = std::max(maxlvl, lv);
    }
  }
  // bucket targets by level (targets with no ops and no division still get
  // scheduled at level 0: they are plain copies of A)
  std::vector<int> lcnt(maxlvl + 2, 0);
  for (int p = 0; p < s.nlu; ++p) lcnt[level[p] + 1]++;
  s.f_lvl_ptr.assign(maxlvl + 2, 0);
  for (int l = 0; l <= maxlvl; ++l) s.f_lvl_ptr[l + 1] = s.f_lvl_ptr[l] + lcnt[l + 1];
  s.f_tgt.resize(s.nlu);
  {
    std::vector<int> fill(s.f_lvl_ptr.begin(), s.f_lvl_ptr.end() - 1);
    for (int p = 0; p < s.nlu; ++p) s.f_tgt[fill[level[p]]++] = p;
  }
  s.f_op_ptr.assign(s.nlu + 1, 0);
  s.f_div.resize(s.nlu);
  s.f_ops.reserve(ops.size());
  for (int t = 0; t < s.nlu; ++t) {
    const int p = s.f_tgt[t];
    for (int o = op_ptr[p]; o < op_ptr[p + 1]; ++o) { s.f_ops.push_back(ops[2 * o]); s.f_ops.push_back(ops[2 * o + 1]); }
    s.f_op_ptr[t + 1] = (int)(s.f_ops.size() / 2);
    s.f_div[t] = diag_of[p];
  }
  // ---- solve schedules
  std::vector<int> flev(n, 0), blev(n, 0);
  int fmax = 0, bmax = 0;
  for (int i = 0; i < n; ++i) {
    int lv = 0;
    for (int k : lrow[i]) lv = std::max(lv, flev[k] + 1);
    flev[i] = lv; fmax = std::max(fmax, lv);
  }
  for (int i = n - 1; i >= 0; --i) {
    int lv = 0;
    for (int j : urow[i]) lv = std::max(lv, blev[j] + 1);
    blev[i] = lv; bmax = std::max(bmax, lv);
  }
  auto bucket = [&](const std::vector<int>& lev, int mx, std::vector<int>& ptr, std::vector<int>& rows) {
    ptr.assign(mx + 2, 0);
    for (int i = 0; i < n; ++i) ptr[lev[i] + 1]++;
    for (int l = 0; l <= mx; ++l) ptr[l + 1] += ptr[l];
    rows.resize(n);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < n; ++i) rows[fill[lev[i]]++] = i;
  };
  bucket(flev, fmax, s.fs_lvl_ptr, s.fs_rows);
  bucket(blev, bmax, s.bs_lvl_ptr, s.bs_rows);
  return XG_OK;
}

}  // namespace xg
