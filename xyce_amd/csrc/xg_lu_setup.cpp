// Symbolic analysis of the fixed-pattern sparse LU used for the Newton
// solve of each independent sub-circuit.
//
// The reference solves J dx = -F with KLU (N_LAS_AmesosSolver.C:231-470):
// a fill-reducing ordering, then partial pivoting with a preference for the
// diagonal (a diagonal entry is kept while |a_kk| >= 0.001 max_i |a_ik|).
// Here the pivot sequence is frozen at setup from the structure alone:
//
//  1. matching: each row is assigned a column so that the diagonal is
//     structurally nonzero.  A voltage source's branch row takes its
//     positive (else negative) node's column and that node's row takes the
//     branch column (the +-1 entries of N_DEV_Vsrc.C:1023-1036); the other
//     rows keep their own column when they can, and augmenting paths
//     settle the rest.
//  2. ordering: nested dissection of the symmetrised graph (BFS level-set
//     separators from a pseudo-peripheral node), with rows of very high
//     degree (supply rails) last.  It keeps the elimination tree shallow,
//     so the factorisation has few dependent levels.
//  3. fill: the row structure of L from the elimination tree (the filled
//     pattern is made structurally symmetric); U is its transpose.
//  4. tasks: every value of L and U is one task (a Crout dot product,
//     terms in ascending order), grouped into levels of independent tasks;
//     the two triangular solves likewise, one task per row.
//
// The device checks every solve's backward error (nonzero finite pivots,
// ||J dx - rhs|| <= 1e-12 (||J|| ||dx|| + ||rhs||)) and refactors a
// component that fails with the dense partial-pivoting kernel.
//
// Everything here uses component-local indices only, so a sub-circuit gets
// the same factorisation alone and inside a batch.
#include <algorithm>
#include <atomic>
#include <functional>
#include <list>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <iterator>
#include <queue>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <string>
#include <new>
#include <stdexcept>
#include <system_error>
#include <thread>
#include <vector>

#include "xg_internal.h"
#include "xg_setup.h"

namespace xg {

namespace {

// Per-component limits of the static factorisation: strictly-lower fill
// entries and multiply-adds per factorisation.
constexpr int64_t kLuMaxFill = 4 << 20;
constexpr int64_t kLuMaxOps = 64 << 20;
// Past those, the component is factored by dense fronts (multifrontal);
// the fill they may hold (strictly-lower entries of L).
constexpr int64_t kFrontMaxFill = (int64_t)1 << 28;
// Element steps the minimum-degree order of a front component may take
// (bounds its setup time to seconds).
constexpr int64_t kFrontMdWork = (int64_t)4 << 30;

using Graph = std::vector<std::vector<int32_t>>;

// Breadth-first levels of the vertices with in[v] == tag, from root.
int bfs_levels(const Graph &g, const std::vector<int32_t> &in, int32_t tag,
               int32_t root, std::vector<int32_t> &lev,
               std::vector<int32_t> &order) {
  order.clear();
  order.push_back(root);
  lev[root] = 0;
  int depth = 0;
  for (size_t h = 0; h < order.size(); ++h) {
    const int32_t v = order[h];
    for (int32_t u : g[v])
      if (in[u] == tag && lev[u] < 0) {
        lev[u] = lev[v] + 1;
        depth = std::max(depth, lev[u]);
        order.push_back(u);
      }
  }
  return depth + 1;
}

struct Dissector {
  const Graph &g;
  std::vector<int32_t> in;     // current tag of each vertex (-1: ordered)
  std::vector<int32_t> lev;
  std::vector<int32_t> &out;   // elimination order
  int32_t next_tag = 1;
  Dissector(const Graph &g_, std::vector<int32_t> &o)
      : g(g_), in(g_.size(), -1), lev(g_.size(), -1), out(o) {}

  // Orders the vertex set `set` (all tagged `tag`).
  void run(std::vector<int32_t> set, int32_t tag) {
    if (set.size() <= 8) {
      for (int32_t v : set) {
        out.push_back(v);
        in[v] = -1;
      }
      return;
    }
    // connected pieces are ordered one after the other
    std::vector<int32_t> order;
    for (int32_t v : set) lev[v] = -1;
    std::vector<std::vector<int32_t>> pieces;
    for (int32_t v : set)
      if (lev[v] < 0) {
        bfs_levels(g, in, tag, v, lev, order);
        pieces.push_back(order);
      }
    if (pieces.size() > 1) {
      for (auto &p : pieces) {
        const int32_t t = next_tag++;
        for (int32_t v : p) in[v] = t;
        run(p, t);
      }
      return;
    }
    // pseudo-peripheral root: the last vertex of a BFS, twice
    int32_t root = set[0];
    int depth = 0;
    for (int it = 0; it < 2; ++it) {
      for (int32_t v : set) lev[v] = -1;
      depth = bfs_levels(g, in, tag, root, lev, order);
      root = order.back();
    }
    for (int32_t v : set) lev[v] = -1;
    depth = bfs_levels(g, in, tag, root, lev, order);
    if (depth < 3) {
      for (int32_t v : set) {
        out.push_back(v);
        in[v] = -1;
      }
      return;
    }
    std::vector<int32_t> cnt(depth, 0);
    for (int32_t v : set) cnt[lev[v]]++;
    // the smallest level in the middle third is the separator
    int lo = std::max(1, depth / 3), hi = std::min(depth - 2, (2 * depth) / 3);
    if (hi < lo) lo = hi = depth / 2;
    int sep = lo;
    for (int l = lo; l <= hi; ++l)
      if (cnt[l] < cnt[sep]) sep = l;
    // sides: 0 below the separator level, 1 above, 2 on it.  A level
    // vertex with no neighbour on one side joins the other (leaves hanging
    // off the level, such as a device's input and branch nodes, need not
    // separate anything); a move never joins the two sides, as a vertex
    // moves only when it has no neighbour on the side it leaves open.
    for (int32_t v : set) lev[v] = lev[v] < sep ? 0 : lev[v] > sep ? 1 : 2;
    for (int32_t v : set) {
      if (lev[v] != 2) continue;
      bool below = false, above = false;
      for (int32_t u : g[v])
        if (in[u] == tag) {
          below |= lev[u] == 0;
          above |= lev[u] == 1;
        }
      if (!above)
        lev[v] = 0;
      else if (!below)
        lev[v] = 1;
    }
    std::vector<int32_t> a, b, s;
    for (int32_t v : set) (lev[v] == 0 ? a : lev[v] == 1 ? b : s).push_back(v);
    const int32_t ta = next_tag++, tb = next_tag++;
    for (int32_t v : a) in[v] = ta;
    for (int32_t v : b) in[v] = tb;
    for (int32_t v : s) in[v] = -2;  // kept out of both halves
    run(a, ta);
    run(b, tb);
    for (int32_t v : s) {
      out.push_back(v);
      in[v] = -1;
    }
  }
};

// Row -> column matching with a structurally nonzero diagonal.
// Returns false if the pattern is structurally singular.
bool match_rows(const Graph &rows, int32_t n, const std::vector<int32_t> &pair_r,
                const std::vector<int32_t> &pair_c, std::vector<int32_t> &col_of_row,
                std::vector<int32_t> &row_of_col) {
  col_of_row.assign(n, -1);
  row_of_col.assign(n, -1);
  auto has = [&](int32_t r, int32_t c) {
    return std::binary_search(rows[r].begin(), rows[r].end(), c);
  };
  for (size_t k = 0; k < pair_r.size(); ++k) {
    const int32_t br = pair_r[k], nd = pair_c[k];
    if (col_of_row[br] >= 0 || col_of_row[nd] >= 0 || row_of_col[nd] >= 0 ||
        row_of_col[br] >= 0 || !has(br, nd) || !has(nd, br))
      continue;
    col_of_row[br] = nd;
    row_of_col[nd] = br;
    col_of_row[nd] = br;
    row_of_col[br] = nd;
  }
  for (int32_t r = 0; r < n; ++r)
    if (col_of_row[r] < 0 && row_of_col[r] < 0 && has(r, r)) {
      col_of_row[r] = r;
      row_of_col[r] = r;
    }
  std::vector<int32_t> seen(n, -1);
  std::vector<std::pair<int32_t, size_t>> stack;
  std::vector<int32_t> path_col;
  for (int32_t r0 = 0; r0 < n; ++r0) {
    if (col_of_row[r0] >= 0) continue;
    // depth-first augmenting path from row r0
    stack.assign(1, {r0, 0});
    path_col.clear();
    bool found = false;
    while (!stack.empty() && !found) {
      auto &top = stack.back();
      const int32_t r = top.first;
      if (top.second >= rows[r].size()) {
        stack.pop_back();
        if (!path_col.empty()) path_col.pop_back();
        continue;
      }
      const int32_t c = rows[r][top.second++];
      if (seen[c] == r0) continue;
      seen[c] = r0;
      path_col.push_back(c);
      if (row_of_col[c] < 0) {
        found = true;
      } else {
        stack.push_back({row_of_col[c], 0});
      }
    }
    if (!found) return false;
    // flip the path: stack rows pair with path columns
    for (size_t k = 0; k < stack.size(); ++k) {
      const int32_t r = stack[k].first, c = path_col[k];
      col_of_row[r] = c;
      row_of_col[c] = r;
    }
  }
  return true;
}

// Minimum-degree order of the symmetric graph g (sorted, duplicate-free
// neighbour lists): repeatedly eliminate a vertex of least current degree
// (smallest index on ties), joining its neighbours into a clique of the
// explicit elimination graph.  The vertices in `last` (supply rails, whose
// rows are dense anyway) stay out of the elimination graph and follow in
// the given order.  False when that graph would exceed `budget` (directed)
// edges or the list merges `work` element steps (bounding the setup time).
bool min_degree_order(const Graph &g, const std::vector<int32_t> &last,
                      std::vector<int32_t> &order, int64_t budget, int64_t work) {
  const int32_t n = (int32_t)g.size();
  std::vector<char> gone(n, 0);
  for (int32_t v : last) gone[v] = 1;
  Graph e(n);
  int64_t edges = 0;
  for (int32_t v = 0; v < n; ++v) {
    if (gone[v]) continue;
    for (int32_t u : g[v])
      if (!gone[u]) e[v].push_back(u);
    edges += (int64_t)e[v].size();
  }
  using Item = std::pair<int32_t, int32_t>;  // (degree, vertex)
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
  for (int32_t v = 0; v < n; ++v)
    if (!gone[v]) pq.push({(int32_t)e[v].size(), v});
  order.clear();
  order.reserve(n);
  std::vector<int32_t> merged;
  while (!pq.empty()) {
    const Item it = pq.top();
    pq.pop();
    const int32_t v = it.second;
    if (gone[v] || it.first != (int32_t)e[v].size()) continue;  // stale entry
    gone[v] = 1;
    order.push_back(v);
    std::vector<int32_t> nb;
    nb.swap(e[v]);
    edges -= (int64_t)nb.size();
    for (int32_t a : nb) {
      // a's neighbours become (theirs and v's) minus a and v
      merged.clear();
      std::set_union(e[a].begin(), e[a].end(), nb.begin(), nb.end(),
                     std::back_inserter(merged));
      merged.erase(std::remove_if(merged.begin(), merged.end(),
                                  [&](int32_t x) { return x == a || x == v; }),
                   merged.end());
      edges += (int64_t)merged.size() - (int64_t)e[a].size();
      work -= (int64_t)merged.size();
      e[a].swap(merged);
      if (edges > budget || work < 0) return false;
      pq.push({(int32_t)e[a].size(), a});
    }
  }
  for (int32_t v : last) order.push_back(v);
  return true;
}

// Strictly-lower fill and elimination depth (the longest chain of pivots
// that depend on each other through L) of an order, by the row-subtree
// walk of the factorisation below; false when the fill exceeds the budget.
// With `flops` given, also the factorisation's floating-point operations
// (per pivot j with c_j entries below it: c_j divisions and 2 c_j^2 for
// the update), and the fill may reach `cap`.
bool order_cost(const Graph &sym, const std::vector<int32_t> &order, int64_t &fill,
                int32_t &depth, int64_t *flops = nullptr, int64_t cap = kLuMaxFill) {
  const int32_t n = (int32_t)sym.size();
  std::vector<int32_t> pinv(n), parent(n, -1), mark(n, -1), lev(n, 0);
  std::vector<int64_t> cc(flops ? n : 0, 0);
  for (int32_t i = 0; i < n; ++i) pinv[order[i]] = i;
  fill = 0;
  depth = 0;
  for (int32_t i = 0; i < n; ++i) {
    mark[i] = i;
    for (int32_t u : sym[order[i]]) {
      int32_t j = pinv[u];
      if (j >= i) continue;
      while (mark[j] != i) {
        ++fill;
        if (flops) cc[j]++;
        lev[i] = std::max(lev[i], lev[j] + 1);
        mark[j] = i;
        if (parent[j] < 0) {
          parent[j] = i;
          break;
        }
        j = parent[j];
      }
    }
    depth = std::max(depth, lev[i]);
    if (fill > cap) return false;
  }
  if (flops) {
    *flops = 0;
    for (int64_t c : cc) *flops += c + 2 * c * c;
  }
  return true;
}

// Components up to this size also get a minimum-degree order (kept when it
// factors in fewer dependent steps than the nested dissection).
constexpr int32_t kMinDegreeMaxN = 100000;

// The fronts' cost of one more dependent elimination level, in
// floating-point operations: a level is about three launches (factor,
// forward, backward; ~20 us), while the fronts factor at ~50 GFLOP/s.  An
// order with a long elimination chain (the minimum-degree order of a
// ladder is the chain itself) loses to a shallower one with more work.
constexpr int64_t kFrontLevelFlops = 1000000;

// One component's static factorisation, with task indices local to the
// component (lu_analyse appends the components in order).  The component's
// entries of prow / qcol / qpos / lu_pos (disjoint from every other
// component's) are written in place.
// Dense fronts of one component (local positions and front ids).
struct FrontLocal {
  std::vector<int32_t> m, k, height;   // per front
  std::vector<int64_t> fo;             // value offsets (n_fronts + 1)
  std::vector<int64_t> ro;             // row-list offsets (n_fronts + 1)
  std::vector<int32_t> rows;           // front rows, ascending positions
  std::vector<int32_t> cbmap;          // at ro[s] + t, t >= k: row's index in the parent
  std::vector<int32_t> ch_ptr, ch;     // children, ascending
  std::vector<int32_t> ent_k;          // CSR entries of the component
  std::vector<int64_t> ent_fpos;       // their value positions
  int64_t flops = 0;
  int32_t max_m = 0;
};

struct CompLu {
  int rc = XG_OK;
  std::string err;
  bool ok = false, in_lds = true, fronts = false;
  FrontLocal fr;
  int64_t fn = 0, factor_ops = 0;
  int nlev[3] = {0, 0, 0};
  std::vector<int32_t> t_dst, t_div, ops;
  std::vector<int32_t> t_end;                    // per task: end of its ops (pairs)
  std::vector<int32_t> lv_end, lv_mid, lv_long;  // per level, task indices
};

// Supernodal multifrontal plan from the elimination tree `parent` and the
// strictly-lower column structures `ucol` (ascending) of a component in
// pivot positions.  Supernodes: column j joins j - 1 when j is the parent
// of j - 1 and L(:, j-1) = {j} + L(:, j) (other children of j then hang off
// the supernode: their contribution rows lie in its rows).  Front s covers
// columns [f, l) with rows R = [f, l) + L(:, l-1) (ascending); its dense
// m x m block (row-major) receives the matrix entries whose smaller
// position falls in [f, l), and the contribution blocks of its children.
// Entry e is (ei[e], ej[e]) in positions.
bool build_fronts(int32_t n, const std::vector<int32_t> &parent,
                  const std::vector<std::vector<int32_t>> &ucol,
                  const std::vector<int32_t> &ei, const std::vector<int32_t> &ej,
                  FrontLocal &F, std::string &err) {
  std::vector<int32_t> nchild(n, 0);
  for (int32_t j = 0; j < n; ++j)
    if (parent[j] >= 0) nchild[parent[j]]++;
  std::vector<int32_t> first;  // supernode starts
  std::vector<int32_t> sn_of(n);
  for (int32_t j = 0; j < n; ++j) {
    const bool join = j > 0 && parent[j - 1] == j &&
                      ucol[j - 1].size() == ucol[j].size() + 1;
    if (!join) first.push_back(j);
    sn_of[j] = (int32_t)first.size() - 1;
  }
  const int32_t ns = (int32_t)first.size();
  first.push_back(n);
  F.m.resize(ns);
  F.k.resize(ns);
  F.height.assign(ns, 0);
  F.fo.assign(ns + 1, 0);
  F.ro.assign(ns + 1, 0);
  std::vector<int32_t> spar(ns, -1);
  for (int32_t s = 0; s < ns; ++s) {
    const int32_t f = first[s], l = first[s + 1];
    const int32_t k = l - f, m = k + (int32_t)ucol[l - 1].size();
    F.k[s] = k;
    F.m[s] = m;
    F.max_m = std::max(F.max_m, m);
    F.fo[s + 1] = F.fo[s] + (int64_t)m * m;
    F.ro[s + 1] = F.ro[s] + m;
    if (parent[l - 1] >= 0) spar[s] = sn_of[parent[l - 1]];
    // flops of the partial factorisation: per pivot a column scaling and a
    // rank-1 update of the trailing block
    for (int64_t j = 0; j < k; ++j) F.flops += (m - j - 1) + 2 * (m - j - 1) * (m - j - 1);
  }
  F.rows.resize(F.ro[ns]);
  F.cbmap.assign(F.ro[ns], -1);
  for (int32_t s = 0; s < ns; ++s) {
    int32_t *r = F.rows.data() + F.ro[s];
    const int32_t f = first[s], l = first[s + 1];
    for (int32_t j = f; j < l; ++j) *r++ = j;
    for (int32_t i : ucol[l - 1]) *r++ = i;
  }
  auto local = [&](int32_t s, int32_t pos) -> int32_t {
    const int32_t *b = F.rows.data() + F.ro[s], *e = b + F.m[s];
    const int32_t *it = std::lower_bound(b, e, pos);
    return (it != e && *it == pos) ? (int32_t)(it - b) : -1;
  };
  F.ch_ptr.assign(ns + 1, 0);
  for (int32_t s = 0; s < ns; ++s)
    if (spar[s] >= 0) F.ch_ptr[spar[s] + 1]++;
  for (int32_t s = 0; s < ns; ++s) F.ch_ptr[s + 1] += F.ch_ptr[s];
  F.ch.resize(F.ch_ptr[ns]);
  {
    std::vector<int32_t> fillp(F.ch_ptr.begin(), F.ch_ptr.end() - 1);
    for (int32_t s = 0; s < ns; ++s)
      if (spar[s] >= 0) F.ch[fillp[spar[s]]++] = s;
  }
  for (int32_t s = 0; s < ns; ++s) {
    const int32_t p = spar[s];
    if (p < 0) continue;
    F.height[p] = std::max(F.height[p], F.height[s] + 1);
    for (int32_t t = F.k[s]; t < F.m[s]; ++t) {
      const int32_t q = local(p, F.rows[F.ro[s] + t]);
      if (q < 0) {
        err = "sparse LU: a contribution row is missing from its parent front";
        return false;
      }
      F.cbmap[F.ro[s] + t] = q;
    }
  }
  if (std::getenv("XG_LU_DEBUG")) {
    int32_t hmax = 0;
    for (int32_t h : F.height) hmax = std::max(hmax, h);
    std::vector<int64_t> cnt(hmax + 1, 0), kk(hmax + 1, 0), mm(hmax + 1, 0);
    for (int32_t s = 0; s < ns; ++s) {
      cnt[F.height[s]]++;
      kk[F.height[s]] += F.k[s];
      mm[F.height[s]] = std::max<int64_t>(mm[F.height[s]], F.m[s]);
    }
    for (int32_t h = 0; h <= hmax; h += std::max(1, hmax / 40))
      std::fprintf(stderr, "  level %d: fronts %lld, columns %lld, max m %lld\n", h,
                   (long long)cnt[h], (long long)kk[h], (long long)mm[h]);
  }
  F.ent_fpos.resize(ei.size());
  for (size_t e = 0; e < ei.size(); ++e) {
    const int32_t i = ei[e], j = ej[e];
    const int32_t s = sn_of[std::min(i, j)];
    const int32_t li = local(s, i), lj = local(s, j);
    if (li < 0 || lj < 0) {
      err = "sparse LU: a matrix entry falls outside its front";
      return false;
    }
    F.ent_fpos[e] = F.fo[s] + (int64_t)li * F.m[s] + lj;
  }
  return true;
}

void analyse_component_uncached(const Plan &plan, LuPlan &L, int32_t c,
                                const std::vector<std::pair<int32_t, int32_t>> &pairs,
                                CompLu &out) {
  const int32_t G = plan.n_nodes;
  const int32_t beg = plan.comp_start[c], n = plan.comp_start[c + 1] - beg;
  const int32_t *nodes = plan.comp_nodes.data() + beg;
  auto loc = [&](int32_t g) { return plan.comp_local[g]; };
  // dense fronts by choice: always (XG_LU_FRONTS), or (XG_LU_AUTO) when the
  // task lists cannot fit in LDS -- their global-memory form is slower than
  // the fronts.  With `fill` strictly-lower entries the filled pattern has
  // n + 2 fill values, the task lists at least n + fill tasks and 2 fill
  // solve terms, a lower bound of their LDS bytes.
  auto fronts_for = [&](int64_t fill) {
    if (plan.lu_mode == XG_LU_FRONTS) return true;
    if (plan.lu_mode != XG_LU_AUTO) return false;
    return sparse_lu_lds_bytes(n, n + 2 * fill, n + fill, 2 * fill, 0) > sparse_lu_lds_budget();
  };
  // local structure of the component's rows
  Graph rows(n);
  for (int32_t r = 0; r < n; ++r) {
    const int32_t g = nodes[r];
    for (int32_t k = plan.rowptr[g]; k < plan.rowptr[g + 1]; ++k) {
      const int32_t col = plan.colidx[k];
      if (col < G) rows[r].push_back(loc(col));
    }
    std::sort(rows[r].begin(), rows[r].end());
    rows[r].erase(std::unique(rows[r].begin(), rows[r].end()), rows[r].end());
  }
  std::vector<int32_t> pr, pc;
  for (auto &p : pairs) {
    pr.push_back(loc(p.first));
    pc.push_back(loc(p.second));
  }
  std::vector<int32_t> col_of_row, row_of_col;
  if (!match_rows(rows, n, pr, pc, col_of_row, row_of_col)) {
    for (int32_t i = 0; i < n; ++i) {
      L.prow[beg + i] = nodes[i];
      L.qcol[beg + i] = nodes[i];
      L.qpos[nodes[i]] = i;
    }
    return;
  }
  // symmetrised graph of the matched matrix M(c, j) = A(row_of_col[c], j)
  Graph sym(n);
  for (int32_t cc = 0; cc < n; ++cc)
    for (int32_t j : rows[row_of_col[cc]])
      if (j != cc) {
        sym[cc].push_back(j);
        sym[j].push_back(cc);
      }
  size_t edges = 0;
  for (auto &a : sym) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
    edges += a.size();
  }
  // ordering: nested dissection; rows of very high degree last
  const double avg = n ? (double)edges / n : 0.0;
  const size_t hub = (size_t)std::max(32.0, std::max(8.0 * avg, 4.0 * std::sqrt((double)n)));
  std::vector<int32_t> order, hubs;
  order.reserve(n);
  {
    Dissector d(sym, order);
    std::vector<int32_t> rest;
    for (int32_t v = 0; v < n; ++v) (sym[v].size() > hub ? hubs : rest).push_back(v);
    for (int32_t v : rest) d.in[v] = 0;
    d.next_tag = 1;
    d.run(rest, 0);
    std::stable_sort(hubs.begin(), hubs.end(), [&](int32_t a, int32_t b) {
      return sym[a].size() < sym[b].size();
    });
    for (int32_t v : hubs) order.push_back(v);
  }
  // minimum degree instead where its elimination chain is shorter (random
  // logic: a quarter of the fill and a third of the depth; rings and
  // ladders keep the dissection's logarithmic depth)
  if (n <= kMinDegreeMaxN) {
    std::vector<int32_t> md;
    if (min_degree_order(sym, hubs, md, 2 * kLuMaxFill, kLuMaxOps)) {
      int64_t f_nd = 0, f_md = 0;
      int32_t d_nd = 0, d_md = 0;
      const bool ok_nd = order_cost(sym, order, f_nd, d_nd);
      const bool ok_md = order_cost(sym, md, f_md, d_md);
      if (ok_md && (!ok_nd || d_md < d_nd || (d_md == d_nd && f_md < f_nd))) order.swap(md);
    }
  }
  // A component beyond the task lists' budgets is factored by dense
  // fronts, whose cost is their floating-point work: the minimum-degree
  // order with the fronts' budgets replaces the dissection when it needs
  // fewer operations.
  {
    int64_t f0 = 0, w0 = 0;
    int32_t d0 = 0;
    const bool fits = order_cost(sym, order, f0, d0, &w0, kFrontMaxFill);
    if (!fits || f0 > kLuMaxFill || (w0 - f0) / 2 > kLuMaxOps || fronts_for(f0)) {
      std::vector<int32_t> md;
      int64_t f1 = 0, w1 = 0;
      int32_t d1 = 0;
      // (its explicit elimination graph grows too large on big meshes)
      if (n <= kMinDegreeMaxN &&
          min_degree_order(sym, hubs, md, 2 * kFrontMaxFill, kFrontMdWork) &&
          order_cost(sym, md, f1, d1, &w1, kFrontMaxFill) &&
          (!fits || w1 + kFrontLevelFlops * d1 < w0 + kFrontLevelFlops * d0))
        order.swap(md);
      if (std::getenv("XG_LU_DEBUG"))
        std::fprintf(stderr, "lu: n %d nd fill %lld flops %.3g depth %d | md fill %lld flops %.3g depth %d\n",
                     n, (long long)f0, (double)w0, d0, (long long)f1, (double)w1, d1);
    }
  }
  std::vector<int32_t> pinv(n);
  for (int32_t i = 0; i < n; ++i) pinv[order[i]] = i;
  for (int32_t i = 0; i < n; ++i) {
    L.qcol[beg + i] = nodes[order[i]];
    L.qpos[nodes[order[i]]] = i;
    L.prow[beg + i] = nodes[row_of_col[order[i]]];
  }
  // elimination tree and row structure of L (Liu's row-subtree walk),
  // abandoned when the fill or the factorisation work exceeds a budget
  // (the component then has no static factorisation)
  std::vector<std::vector<int32_t>> lrow(n), ucol(n);
  std::vector<int32_t> parent(n, -1), mark(n, -1);
  int64_t fill = 0;
  bool too_big = false;
  for (int32_t i = 0; i < n && !too_big; ++i) {
    mark[i] = i;
    for (int32_t u : sym[order[i]]) {
      int32_t j = pinv[u];
      if (j >= i) continue;
      while (mark[j] != i) {
        lrow[i].push_back(j);
        mark[j] = i;
        if (parent[j] < 0) {
          parent[j] = i;
          break;
        }
        j = parent[j];
      }
    }
    std::sort(lrow[i].begin(), lrow[i].end());
    for (int32_t j : lrow[i]) ucol[j].push_back(i);
    fill += (int64_t)lrow[i].size();
    too_big = fill > kFrontMaxFill;
  }
  if (too_big) {  // beyond the fronts' budget too: no factorisation
    out.in_lds = false;
    return;
  }
  {
    int64_t work = 0;  // multiply-adds: sum over pivots of (column count)^2
    for (int32_t k = 0; k < n; ++k) work += (int64_t)ucol[k].size() * (int64_t)ucol[k].size();
    if (fill > kLuMaxFill || work > kLuMaxOps || fronts_for(fill)) {
      // too large for one workgroup's task lists: dense fronts
      std::vector<int32_t> rpos(n);
      for (int32_t i = 0; i < n; ++i) rpos[row_of_col[order[i]]] = i;
      std::vector<int32_t> ei, ej;
      for (int32_t r = 0; r < n; ++r) {
        const int32_t g = nodes[r];
        for (int32_t k = plan.rowptr[g]; k < plan.rowptr[g + 1]; ++k) {
          const int32_t col = plan.colidx[k];
          if (col >= G) continue;
          out.fr.ent_k.push_back(k);
          ei.push_back(rpos[r]);
          ej.push_back(pinv[loc(col)]);
        }
      }
      if (!build_fronts(n, parent, ucol, ei, ej, out.fr, out.err)) {
        out.rc = XG_E_ARG;
        return;
      }
      if (front_form(out.fr.max_m) == 0) {  // no panel holds its largest front
        out.fr = FrontLocal();
        out.in_lds = false;
        return;
      }
      out.fronts = true;
      out.in_lds = false;
      return;
    }
  }
  // filled row-major pattern: row i = lrow[i], i, ucol[i] (all sorted)
  std::vector<int64_t> rp(n + 1, 0);
  for (int32_t i = 0; i < n; ++i)
    rp[i + 1] = rp[i] + (int64_t)lrow[i].size() + 1 + (int64_t)ucol[i].size();
  const int64_t fn = rp[n];
  if (fn > INT32_MAX) {
    out.rc = XG_E_ARG;
    out.err = "sparse LU: filled pattern of a component exceeds 2^31 entries";
    return;
  }
  auto pos = [&](int32_t i, int32_t j) -> int32_t {
    if (j < i) {
      auto it = std::lower_bound(lrow[i].begin(), lrow[i].end(), j);
      return (int32_t)(rp[i] + (it - lrow[i].begin()));
    }
    if (j == i) return (int32_t)(rp[i] + (int64_t)lrow[i].size());
    auto it = std::lower_bound(ucol[i].begin(), ucol[i].end(), j);
    return (int32_t)(rp[i] + (int64_t)lrow[i].size() + 1 + (it - ucol[i].begin()));
  };
  // U(m, j) exactly where L(j, m) is: uof[lo[j] + y] is the position of
  // U(lrow[j][y], j) (ucol[m] lists its rows in ascending order)
  std::vector<int64_t> lo(n + 1, 0);
  for (int32_t i = 0; i < n; ++i) lo[i + 1] = lo[i] + (int64_t)lrow[i].size();
  std::vector<int32_t> uof(lo[n]);
  {
    std::vector<int32_t> next(n, 0);
    for (int32_t j = 0; j < n; ++j)
      for (size_t y = 0; y < lrow[j].size(); ++y) {
        const int32_t m = lrow[j][y];
        uof[lo[j] + y] = (int32_t)(rp[m] + (int64_t)lrow[m].size() + 1 + next[m]++);
      }
  }
  // CSR entry -> position: entry (g, col) of row g lands at
  // (position of g's equation, position of col's unknown)
  std::vector<int32_t> rpos(n);
  for (int32_t i = 0; i < n; ++i) rpos[row_of_col[order[i]]] = i;
  for (int32_t r = 0; r < n; ++r) {
    const int32_t g = nodes[r];
    for (int32_t k = plan.rowptr[g]; k < plan.rowptr[g + 1]; ++k) {
      const int32_t col = plan.colidx[k];
      if (col < G) L.lu_pos[k] = pos(rpos[r], pinv[loc(col)]);
    }
  }
  out.fn = fn;
  out.ok = true;
  // ---- factor tasks: values in the order diag, column of L, row of U.
  // Tasks are kept flat: task k's (a, b) operand pairs are
  // ops[2 * beg[k], 2 * beg[k + 1]).
  std::vector<int32_t> lev(fn, -1);
  struct Tasks {
    std::vector<int32_t> dst, div, lev, ops;
    std::vector<int64_t> beg{0};
    void open(int32_t d, int32_t v, int32_t l) {
      dst.push_back(d);
      div.push_back(v);
      lev.push_back(l);
    }
    void close() { beg.push_back((int64_t)ops.size() / 2); }
    void drop() {  // the open task (no operands)
      dst.pop_back();
      div.pop_back();
      lev.pop_back();
    }
    size_t size() const { return dst.size(); }
  };
  Tasks tasks, bwd;
  // terms L(i, m) U(m, j) over m in lrow[i] and lrow[j] (m < min(i, j)),
  // ascending m.  The pivot p's row is scattered into at_p (at_p[m] = the
  // index of m in lrow[p]), so each task walks only its other row.
  std::vector<int32_t> at_p(n, -1);
  int32_t lv = 0;
  auto term = [&](int32_t pa, int32_t pb) {
    tasks.ops.push_back(pa);
    tasks.ops.push_back(pb);
    lv = std::max(lv, std::max(lev[pa], lev[pb]) + 1);
  };
  for (int32_t p = 0; p < n; ++p) {
    const auto &rowp = lrow[p];
    for (size_t x = 0; x < rowp.size(); ++x) at_p[rowp[x]] = (int32_t)x;
    const int32_t dp = (int32_t)(rp[p] + (int64_t)rowp.size());
    lv = 0;
    for (size_t x = 0; x < rowp.size(); ++x)
      term((int32_t)(rp[p] + (int64_t)x), uof[lo[p] + x]);
    tasks.open(dp, -2, lv);
    tasks.close();
    lev[dp] = lv;
    for (int32_t i : ucol[p]) {  // L(i, p), i > p
      const int32_t d = pos(i, p);
      lv = lev[dp] + 1;
      const auto &a = lrow[i];
      for (size_t x = 0; x < a.size() && a[x] < p; ++x) {
        const int32_t y = at_p[a[x]];
        if (y >= 0) term((int32_t)(rp[i] + (int64_t)x), uof[lo[p] + y]);
      }
      tasks.open(d, dp, lv);
      tasks.close();
      lev[d] = lv;
    }
    for (size_t z = 0; z < ucol[p].size(); ++z) {  // U(p, j), j > p
      const int32_t j = ucol[p][z];
      const int64_t b0 = (int64_t)tasks.ops.size();
      lv = 0;
      const auto &b = lrow[j];
      for (size_t y = 0; y < b.size() && b[y] < p; ++y) {
        const int32_t x = at_p[b[y]];
        if (x >= 0) term((int32_t)(rp[p] + (int64_t)x), uof[lo[j] + y]);
      }
      if ((int64_t)tasks.ops.size() == b0) continue;  // the scattered value is final
      const int32_t d = dp + 1 + (int32_t)z;
      tasks.open(d, -1, lv);
      tasks.close();
      lev[d] = lv;
    }
    for (int32_t m : rowp) at_p[m] = -1;
  }
  // ---- forward solve (L y = b, unit diagonal) and backward (U x = y).
  // A forward row needs its L entries and the earlier y it multiplies, not
  // the whole factorisation: it joins the factor phase at the level after
  // both, so the forward solve adds no levels of its own.
  {
    std::vector<int32_t> lf(n, -1), lb(n, -1);
    for (int32_t i = 0; i < n; ++i) {
      if (lrow[i].empty()) continue;
      int32_t l = 0;
      for (size_t x = 0; x < lrow[i].size(); ++x) {
        const int32_t k = lrow[i][x];
        const int32_t pl = (int32_t)(rp[i] + (int64_t)x);
        tasks.ops.push_back(pl);
        tasks.ops.push_back(k);
        l = std::max(l, std::max(lev[pl], lf[k]) + 1);
      }
      tasks.open(i, kLuFwdTask, l);
      tasks.close();
      lf[i] = l;
    }
    for (int32_t i = n - 1; i >= 0; --i) {
      const int32_t di = (int32_t)(rp[i] + (int64_t)lrow[i].size());
      int32_t l = 0;
      for (size_t z = 0; z < ucol[i].size(); ++z) {
        const int32_t j = ucol[i][z];
        bwd.ops.push_back(di + 1 + (int32_t)z);
        bwd.ops.push_back(j);
        l = std::max(l, lb[j] + 1);
      }
      bwd.open(i, di, l);
      bwd.close();
      lb[i] = l;
    }
  }
  // ---- the three phases, tasks bucketed by level (short tasks first, then
  // mid, then long within a level; stable within each class)
  auto emit = [&](Tasks &ts, int ph) {
    const int32_t nt = (int32_t)ts.size();
    int nlev = 0;
    for (int32_t l : ts.lev) nlev = std::max(nlev, l + 1);
    auto cls = [&](int32_t k) {
      const int64_t m = ts.beg[k + 1] - ts.beg[k];
      return m <= kLuShortOps ? 0 : m <= kLuLongOps ? 1 : 2;
    };
    // counting sort by (level, class), stable in task order
    std::vector<int64_t> cnt((size_t)3 * nlev + 1, 0);
    for (int32_t k = 0; k < nt; ++k) cnt[(size_t)3 * ts.lev[k] + cls(k) + 1]++;
    for (size_t b = 1; b < cnt.size(); ++b) cnt[b] += cnt[b - 1];
    std::vector<int32_t> sorted(nt);
    {
      std::vector<int64_t> at(cnt.begin(), cnt.end() - 1);
      for (int32_t k = 0; k < nt; ++k) sorted[at[(size_t)3 * ts.lev[k] + cls(k)]++] = k;
    }
    out.t_dst.reserve(out.t_dst.size() + nt);
    out.t_div.reserve(out.t_div.size() + nt);
    out.t_end.reserve(out.t_end.size() + nt);
    out.ops.reserve(out.ops.size() + ts.ops.size());
    const int32_t base = (int32_t)out.t_dst.size();
    for (int l = 0; l < nlev; ++l) {
      for (int64_t q = cnt[(size_t)3 * l]; q < cnt[(size_t)3 * l + 3]; ++q) {
        const int32_t k = sorted[q];
        out.t_dst.push_back(ts.dst[k]);
        out.t_div.push_back(ts.div[k]);
        out.ops.insert(out.ops.end(), ts.ops.begin() + 2 * ts.beg[k],
                       ts.ops.begin() + 2 * ts.beg[k + 1]);
        out.t_end.push_back((int32_t)(out.ops.size() / 2));
        if (ph == 0) out.factor_ops += ts.beg[k + 1] - ts.beg[k];
      }
      out.lv_mid.push_back(base + (int32_t)cnt[(size_t)3 * l + 1]);
      out.lv_long.push_back(base + (int32_t)cnt[(size_t)3 * l + 2]);
      out.lv_end.push_back(base + (int32_t)cnt[(size_t)3 * l + 3]);
    }
    out.nlev[ph] = nlev;
    ts = Tasks();
  };
  Tasks none;
  emit(tasks, 0);
  emit(none, 1);
  emit(bwd, 2);
  out.in_lds = sparse_lu_lds_bytes(n, fn, (int64_t)out.t_dst.size(),
                                   (int64_t)out.ops.size() / 2,
                                   out.nlev[0] + out.nlev[1] + out.nlev[2]) <=
               sparse_lu_lds_budget();
  if (out.ops.size() / 2 > (size_t)INT32_MAX) {
    out.rc = XG_E_ARG;
    out.err = "sparse LU: task lists exceed 2^31 entries";
  }
}

// ---- analysed components by topology.  A component's analysis depends
// only on its local structure (rows in CSR order with component-local
// columns, source pairings, the solve form), so a component whose
// structure was analysed before -- the same sub-circuit in an earlier
// batch or in this one -- takes the stored result, mapped onto its own
// nodes and CSR entries.  Keyed by a 64-bit hash of that structure, the
// whole structure compared on a match; least recently used entries leave
// beyond the byte limit.  No counterpart in the reference, which analyses
// one circuit once.
struct CacheEntry {
  std::vector<int32_t> key;
  std::vector<int32_t> qcol, prow;  // local unknown / equation per position
  std::vector<int32_t> lu_pos;      // per entry in key order (-1: none)
  CompLu res;                       // fronts: ent_k are entry ordinals
  size_t bytes = 0;
};

struct PlanCache {
  std::mutex mu;
  std::list<std::shared_ptr<CacheEntry>> lru;  // most recent first
  std::unordered_multimap<uint64_t, std::list<std::shared_ptr<CacheEntry>>::iterator> map;
  size_t bytes = 0, limit = (size_t)256 << 20;
  int64_t hits = 0, misses = 0;
};

PlanCache &plan_cache() {
  static PlanCache c;
  return c;
}

uint64_t fnv64(const std::vector<int32_t> &v) {
  uint64_t h = 1469598103934665603ull;
  for (int32_t x : v) {
    h ^= (uint32_t)x;
    h *= 1099511628211ull;
  }
  return h;
}

template <class T>
size_t vbytes(const std::vector<T> &v) { return v.size() * sizeof(T); }

size_t entry_bytes(const CacheEntry &e) {
  const CompLu &r = e.res;
  const FrontLocal &f = r.fr;
  return sizeof(CacheEntry) + vbytes(e.key) + vbytes(e.qcol) + vbytes(e.prow) +
         vbytes(e.lu_pos) + vbytes(r.t_dst) + vbytes(r.t_div) + vbytes(r.ops) + vbytes(r.t_end) +
         vbytes(r.lv_end) + vbytes(r.lv_mid) + vbytes(r.lv_long) + vbytes(f.m) + vbytes(f.k) +
         vbytes(f.height) + vbytes(f.fo) + vbytes(f.ro) + vbytes(f.rows) + vbytes(f.cbmap) +
         vbytes(f.ch_ptr) + vbytes(f.ch) + vbytes(f.ent_k) + vbytes(f.ent_fpos);
}

// The structure key of component c, and its CSR entries in key order.
void component_key(const Plan &plan, int32_t c,
                   const std::vector<std::pair<int32_t, int32_t>> &pairs,
                   std::vector<int32_t> &key, std::vector<int32_t> &ents) {
  const int32_t G = plan.n_nodes;
  const int32_t beg = plan.comp_start[c], n = plan.comp_start[c + 1] - beg;
  const int32_t *nodes = plan.comp_nodes.data() + beg;
  key.assign({1, n, plan.lu_mode, (int32_t)pairs.size()});
  ents.clear();
  for (auto &p : pairs) {
    key.push_back(plan.comp_local[p.first]);
    key.push_back(plan.comp_local[p.second]);
  }
  for (int32_t r = 0; r < n; ++r) {
    const int32_t g = nodes[r];
    const size_t at = key.size();
    key.push_back(0);
    for (int32_t k = plan.rowptr[g]; k < plan.rowptr[g + 1]; ++k) {
      const int32_t col = plan.colidx[k];
      if (col >= G) continue;
      key.push_back(plan.comp_local[col]);
      ents.push_back(k);
    }
    key[at] = (int32_t)(key.size() - at - 1);
  }
}

void analyse_component(const Plan &plan, LuPlan &L, int32_t c,
                       const std::vector<std::pair<int32_t, int32_t>> &pairs, CompLu &out) {
  PlanCache &pc = plan_cache();
  {
    std::lock_guard<std::mutex> g(pc.mu);
    if (pc.limit == 0) {
      pc.misses++;
      analyse_component_uncached(plan, L, c, pairs, out);
      return;
    }
  }
  const int32_t beg = plan.comp_start[c], n = plan.comp_start[c + 1] - beg;
  const int32_t *nodes = plan.comp_nodes.data() + beg;
  std::vector<int32_t> key, ents;
  component_key(plan, c, pairs, key, ents);
  const uint64_t h = fnv64(key);
  std::shared_ptr<CacheEntry> hit;
  {
    std::lock_guard<std::mutex> g(pc.mu);
    auto range = pc.map.equal_range(h);
    for (auto it = range.first; it != range.second; ++it)
      if ((*it->second)->key == key) {
        hit = *it->second;
        pc.lru.splice(pc.lru.begin(), pc.lru, it->second);  // most recent
        break;
      }
    if (hit) pc.hits++;
    else pc.misses++;
  }
  if (hit) {
    for (int32_t i = 0; i < n; ++i) {
      L.qcol[beg + i] = nodes[hit->qcol[i]];
      L.qpos[nodes[hit->qcol[i]]] = i;
      L.prow[beg + i] = nodes[hit->prow[i]];
    }
    for (size_t e = 0; e < ents.size(); ++e) L.lu_pos[ents[e]] = hit->lu_pos[e];
    out = hit->res;
    for (auto &k : out.fr.ent_k) k = ents[k];
    return;
  }
  analyse_component_uncached(plan, L, c, pairs, out);
  if (out.rc != XG_OK) return;
  auto e = std::make_shared<CacheEntry>();
  e->qcol.resize(n);
  e->prow.resize(n);
  for (int32_t i = 0; i < n; ++i) {
    e->qcol[i] = plan.comp_local[L.qcol[beg + i]];
    e->prow[i] = plan.comp_local[L.prow[beg + i]];
  }
  e->lu_pos.resize(ents.size());
  for (size_t q = 0; q < ents.size(); ++q) e->lu_pos[q] = L.lu_pos[ents[q]];
  e->res = out;
  // the fronts list the entries in this same order: store their ordinals
  if (e->res.fr.ent_k.size() != ents.size() && !e->res.fr.ent_k.empty()) return;
  for (size_t q = 0; q < e->res.fr.ent_k.size(); ++q) e->res.fr.ent_k[q] = (int32_t)q;
  e->key = std::move(key);
  e->bytes = entry_bytes(*e);
  std::lock_guard<std::mutex> g(pc.mu);
  if (e->bytes > pc.limit / 4) return;  // too large to keep
  pc.lru.push_front(e);
  pc.map.emplace(h, pc.lru.begin());
  pc.bytes += e->bytes;
  while (pc.bytes > pc.limit && !pc.lru.empty()) {
    auto last = std::prev(pc.lru.end());
    const uint64_t hk = fnv64((*last)->key);
    auto range = pc.map.equal_range(hk);
    for (auto it = range.first; it != range.second; ++it)
      if (it->second == last) {
        pc.map.erase(it);
        break;
      }
    pc.bytes -= (*last)->bytes;
    pc.lru.erase(last);
  }
}

}  // namespace

int lu_analyse(Plan &plan, std::string &err) {
  const int32_t n_all = plan.n_nodes, G = plan.n_nodes;
  const int32_t nc = (int32_t)plan.comp_start.size() - 1;
  LuPlan &L = plan.lu;
  L = LuPlan{};
  L.prow.assign(n_all, 0);
  L.qcol.assign(n_all, 0);
  L.qpos.assign(n_all + 1, -1);
  L.lu_pos.assign(plan.colidx.size(), -1);
  L.voff.assign(nc + 1, 0);
  L.ok.assign(nc, 0);
  L.in_lds.assign(nc, 1);
  L.ent_off.assign(1, 0);
  L.ent_rs.assign((size_t)n_all + nc, 0);
  L.lv_comp.assign((size_t)3 * nc + 1, 0);
  L.lv_ptr.assign(1, 0);
  L.t_op.assign(1, 0);
  // voltage-source pairings (branch row, node column), global ids
  std::vector<std::pair<int32_t, int32_t>> vpairs;
  for (int32_t s = 0; s < plan.lin_slots; ++s) {
    if (plan.lin_type[s] != LIN_V) continue;
    const size_t S = (size_t)plan.lin_slots;
    const int32_t p = plan.lin_lid[s], q = plan.lin_lid[S + s],
                  br = plan.lin_lid[2 * S + s];
    const int32_t nd = p != G ? p : q;
    if (nd != G && br != G) vpairs.push_back({br, nd});
  }
  std::vector<std::vector<std::pair<int32_t, int32_t>>> comp_pairs(nc);
  std::vector<int32_t> comp_of(n_all, -1);
  for (int32_t c = 0; c < nc; ++c)
    for (int32_t j = plan.comp_start[c]; j < plan.comp_start[c + 1]; ++j)
      comp_of[plan.comp_nodes[j]] = c;
  for (auto &pr : vpairs) comp_pairs[comp_of[pr.first]].push_back(pr);

  // components analysed on a pool of host threads (largest first), then
  // appended in component order, so the plan does not depend on the pool
  std::vector<CompLu> res(nc);
  std::vector<int32_t> by_size(nc);
  for (int32_t c = 0; c < nc; ++c) by_size[c] = c;
  std::stable_sort(by_size.begin(), by_size.end(), [&](int32_t a, int32_t b) {
    return plan.comp_start[a + 1] - plan.comp_start[a] >
           plan.comp_start[b + 1] - plan.comp_start[b];
  });
  // components of one structure: the first of each (largest first) is
  // analysed in a first pass, the others in a second pass, where they take
  // its stored analysis -- a repeated sub-circuit of a batch is analysed
  // once
  std::vector<int32_t> first, again;
  {
    std::vector<int32_t> key, ents;
    std::unordered_map<uint64_t, int32_t> seen;
    for (int32_t c : by_size) {
      component_key(plan, c, comp_pairs[c], key, ents);
      (seen.emplace(fnv64(key), c).second ? first : again).push_back(c);
    }
  }
  // an allocation failure inside a component's analysis is that
  // component's result (XG_E_NOMEM), never an exception leaving a thread
  auto run_pass = [&](const std::vector<int32_t> &list) {
    std::atomic<int32_t> next{0};
    const int32_t cnt = (int32_t)list.size();
    auto work = [&]() {
      for (int32_t k; (k = next.fetch_add(1)) < cnt;) {
        const int32_t c = list[k];
        try {
          analyse_component(plan, L, c, comp_pairs[c], res[c]);
        } catch (const std::bad_alloc &) {
          res[c] = CompLu();
          res[c].rc = XG_E_NOMEM;
          res[c].err = "sparse LU analysis: out of host memory";
        } catch (const std::exception &e) {
          res[c] = CompLu();
          res[c].rc = XG_E_ARG;
          res[c].err = std::string("sparse LU analysis: ") + e.what();
        }
      }
    };
    const int32_t n_threads = std::max(
        1, std::min<int32_t>({cnt, 16, (int32_t)std::thread::hardware_concurrency()}));
    std::vector<std::thread> pool;
    pool.reserve(n_threads);
    try {
      for (int32_t t = 1; t < n_threads; ++t) pool.emplace_back(work);
    } catch (const std::exception &) {
      // no more threads: the ones started and this one share the work
    }
    work();  // catches inside, so the pool is always joined below
    for (auto &th : pool) th.join();
  };
  run_pass(first);
  run_pass(again);
  L.front_comp.assign(nc, 0);
  L.fr_fo.assign(1, 0);
  L.fr_ro.assign(1, 0);
  L.fr_ch_ptr.assign(1, 0);
  std::vector<int32_t> fr_height;
  for (int32_t c = 0; c < nc; ++c) {
    CompLu &r = res[c];
    if (r.rc != XG_OK) {
      err = r.err;
      return r.rc;
    }
    if (r.fronts) {
      // dense fronts: appended with global front ids, positions and values
      FrontLocal &F = r.fr;
      const int32_t s0 = (int32_t)L.fr_m.size(), ns = (int32_t)F.m.size();
      const int64_t v0 = L.fr_fo.back(), r0 = L.fr_ro.back();
      const int32_t beg = plan.comp_start[c];
      L.front_comp[c] = 1;
      L.n_front_comp++;
      // no task lists: the LDS pass of the task kernel (launched always)
      // hands the component to the front kernels
      L.in_lds[c] = 1;
      for (int32_t s = 0; s < ns; ++s) {
        L.fr_comp.push_back(c);
        L.fr_m.push_back(F.m[s]);
        L.fr_k.push_back(F.k[s]);
        L.fr_fo.push_back(v0 + F.fo[s + 1]);
        L.fr_ro.push_back(r0 + F.ro[s + 1]);
        L.fr_ch_ptr.push_back(L.fr_ch_ptr.back() + F.ch_ptr[s + 1] - F.ch_ptr[s]);
        fr_height.push_back(F.height[s]);
      }
      for (int32_t v : F.rows) L.fr_rows.push_back(beg + v);
      L.fr_cbmap.insert(L.fr_cbmap.end(), F.cbmap.begin(), F.cbmap.end());
      for (int32_t ch : F.ch) L.fr_ch.push_back(s0 + ch);
      for (size_t e = 0; e < F.ent_k.size(); ++e) {
        L.fr_ent_k.push_back(F.ent_k[e]);
        L.fr_ent_fpos.push_back(v0 + F.ent_fpos[e]);
      }
      L.fr_flops += F.flops;
      L.fr_max_m = std::max(L.fr_max_m, F.max_m);
      L.voff[c + 1] = L.voff[c];
      for (int ph = 0; ph < 3; ++ph)
        L.lv_comp[(size_t)3 * c + ph + 1] = (int32_t)L.lv_ptr.size() - 1;
      r = CompLu();
      continue;
    }
    if (!r.ok) {
      L.n_static_fail++;
      for (int ph = 0; ph < 3; ++ph)
        L.lv_comp[(size_t)3 * c + ph + 1] = (int32_t)L.lv_ptr.size() - 1;
      L.voff[c + 1] = L.voff[c];
      L.in_lds[c] = r.in_lds ? 1 : 0;
      continue;
    }
    L.voff[c + 1] = L.voff[c] + r.fn;
    L.fnnz += r.fn;
    L.ok[c] = 1;
    L.in_lds[c] = r.in_lds ? 1 : 0;
    const int64_t tb = (int64_t)L.t_dst.size(), ob = (int64_t)L.ops.size() / 2;
    if (tb + (int64_t)r.t_dst.size() > INT32_MAX ||
        ob + (int64_t)r.ops.size() / 2 > INT32_MAX) {
      err = "sparse LU: task lists exceed 2^31 entries";
      return XG_E_ARG;
    }
    L.t_dst.insert(L.t_dst.end(), r.t_dst.begin(), r.t_dst.end());
    L.t_div.insert(L.t_div.end(), r.t_div.begin(), r.t_div.end());
    L.ops.insert(L.ops.end(), r.ops.begin(), r.ops.end());
    for (int32_t e : r.t_end) L.t_op.push_back((int32_t)(ob + e));
    size_t lvl = 0;
    for (int ph = 0; ph < 3; ++ph) {
      for (int k = 0; k < r.nlev[ph]; ++k, ++lvl) {
        L.lv_long.push_back((int32_t)(tb + r.lv_long[lvl]));
        L.lv_mid.push_back((int32_t)(tb + r.lv_mid[lvl]));
        L.lv_ptr.push_back((int32_t)(tb + r.lv_end[lvl]));
      }
      L.lv_comp[(size_t)3 * c + ph + 1] = (int32_t)L.lv_ptr.size() - 1;
    }
    L.factor_ops += r.factor_ops;
    L.max_factor_levels = std::max(L.max_factor_levels, r.nlev[0]);
    L.max_solve_levels = std::max(L.max_solve_levels, r.nlev[1] + r.nlev[2]);
    r = CompLu();  // release the component's lists
  }
  // front levels: heights from the leaves, processed bottom-up
  {
    int32_t nl = 0;
    for (int32_t h : fr_height) nl = std::max(nl, h + 1);
    L.fr_lv_ptr.assign(nl + 1, 0);
    for (int32_t h : fr_height) L.fr_lv_ptr[h + 1]++;
    for (int32_t l = 0; l < nl; ++l) L.fr_lv_ptr[l + 1] += L.fr_lv_ptr[l];
    L.fr_lv.resize(fr_height.size());
    L.fr_lv_maxm.assign(nl, 0);
    std::vector<int32_t> at(L.fr_lv_ptr.begin(), L.fr_lv_ptr.end() - 1);
    for (int32_t s = 0; s < (int32_t)fr_height.size(); ++s) {
      const int32_t h = fr_height[s];
      L.fr_lv[at[h]++] = s;
      L.fr_lv_maxm[h] = std::max(L.fr_lv_maxm[h], L.fr_m[s]);
    }
  }
  // entry lists (every component, with or without a static factorisation)
  for (int32_t c = 0; c < nc; ++c) {
    const int32_t beg = plan.comp_start[c], n = plan.comp_start[c + 1] - beg;
    const int32_t eb = L.ent_off.back();
    int32_t *rs = L.ent_rs.data() + beg + c;
    for (int32_t i = 0; i < n; ++i) {
      rs[i] = (int32_t)L.ent_k.size() - eb;
      const int32_t g = L.prow[beg + i];
      for (int32_t k = plan.rowptr[g]; k < plan.rowptr[g + 1]; ++k) {
        const int32_t col = plan.colidx[k];
        if (col >= G) continue;
        L.ent_k.push_back(k);
        L.ent_pos.push_back(L.lu_pos[k]);
        L.ent_col.push_back(L.qpos[col]);
      }
    }
    rs[n] = (int32_t)L.ent_k.size() - eb;
    if (L.ent_k.size() > (size_t)INT32_MAX) {
      err = "sparse LU: entry lists exceed 2^31 entries";
      return XG_E_ARG;
    }
    L.ent_off.push_back((int32_t)L.ent_k.size());
  }
  return XG_OK;
}

void lu_cache_stats(int64_t out[3]) {
  PlanCache &pc = plan_cache();
  std::lock_guard<std::mutex> g(pc.mu);
  out[0] = pc.hits;
  out[1] = pc.misses;
  out[2] = (int64_t)pc.bytes;
}

void lu_cache_set_limit(int64_t bytes) {
  PlanCache &pc = plan_cache();
  std::lock_guard<std::mutex> g(pc.mu);
  pc.limit = bytes > 0 ? (size_t)bytes : 0;
  pc.lru.clear();
  pc.map.clear();
  pc.bytes = 0;
  pc.hits = pc.misses = 0;
}

}  // namespace xg
