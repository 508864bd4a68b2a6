// Supernodal symbolic analysis for the multifrontal LDL^T (host).
//
// Replaces the symbolic half of MUMPS' analysis (JOB = 1; src/factor/mumps_factorization.cpp:188-283,
// estimates read at :262-283) with an independent implementation:
//   1. elimination tree of the permuted pattern (Liu's algorithm with path halving);
//   2. postorder, composed into the ordering so that every subtree is numbered contiguously;
//   3. column counts of L by the row-subtree skeleton method (Gilbert, Ng and Peyton, 1994);
//   4. fundamental supernodes, then relaxed amalgamation along each node's last child;
//   5. front row structures bottom-up (own columns, children's contribution rows, original
//      entries), extend-add maps, storage offsets and the level schedule (height in the tree).
#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <condition_variable>
#include <cstdint>
#include <numeric>
#include <stdexcept>
#include <thread>
#include <vector>

#include "analysis.hpp"

namespace hsg {

namespace {

// body(i0, i1) over nt contiguous ranges of [0, n) on as many threads (range 0 on the caller).
template <class F>
void over_ranges(int32_t n, F&& body) {
  const int nt = n < 32768 ? 1 : std::min(16, (int)std::max(1u, std::thread::hardware_concurrency()));
  auto range = [&](int t) {
    body(static_cast<int32_t>(int64_t{n} * t / nt), static_cast<int32_t>(int64_t{n} * (t + 1) / nt));
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back([&, t] { range(t); });
  range(0);
  for (auto& th : pool) th.join();
}

// Lower-triangular pattern (rows > col) of P A P^T in CSC form, new numbering. Columns are
// independent: counted, then filled and sorted over column ranges on several threads.
void permuted_lower(int32_t n, const Graph& g, const std::vector<int32_t>& order,
                    const std::vector<int32_t>& iorder, std::vector<int64_t>& cp, std::vector<int32_t>& ri) {
  cp.assign(static_cast<size_t>(n) + 1, 0);
  over_ranges(n, [&](int32_t j0, int32_t j1) {
    for (int32_t jn = j0; jn < j1; ++jn) {
      const int32_t j = order[jn];
      int64_t c = 0;
      for (int64_t p = g.ptr[j]; p < g.ptr[j + 1]; ++p) c += iorder[g.adj[p]] > jn;
      cp[jn + 1] = c;
    }
  });
  for (int32_t j = 0; j < n; ++j) cp[j + 1] += cp[j];
  ri.resize(static_cast<size_t>(cp[n]));
  over_ranges(n, [&](int32_t j0, int32_t j1) {
    for (int32_t jn = j0; jn < j1; ++jn) {
      const int32_t j = order[jn];
      int64_t q = cp[jn];
      for (int64_t p = g.ptr[j]; p < g.ptr[j + 1]; ++p) {
        const int32_t in = iorder[g.adj[p]];
        if (in > jn) ri[q++] = in;
      }
      std::sort(ri.begin() + cp[jn], ri.begin() + cp[jn + 1]);
    }
  });
}

// Elimination tree of the pattern in the new numbering (Liu's algorithm with path halving): row
// i's entries left of the diagonal are the neighbours of order[i] numbered before i. The algorithm
// gives the elimination tree whatever order a row's columns are visited in, so no sorted lower
// pattern is needed.
// Liu's algorithm with path compression: column i (new numbering) walks up from each lower
// neighbour j to the root of j's current tree and makes i its parent. With `ranges` (closed below:
// see nested_dissection), the columns of each range are processed first, the ranges on several
// threads (a range's walks stay inside it), then the other columns in increasing order; the
// parents are those of the plain column-by-column pass, since a range's walks touch no column a
// later walk from outside it would see differently. A walk leaving its range (ranges not closed
// below) makes the whole pass run again column by column.
std::vector<int32_t> etree_from_graph(int32_t n, const Graph& g, const std::vector<int32_t>& order,
                                      const std::vector<int32_t>& iorder,
                                      const std::vector<std::pair<int32_t, int32_t>>& ranges) {
  std::vector<int32_t> parent(n, -1), anc(n, -1);
  auto column = [&](int32_t i, int32_t lo) {
    const int32_t v = order[i];
    bool ok = true;
    for (int64_t p = g.ptr[v]; p < g.ptr[v + 1]; ++p) {
      int32_t j = iorder[g.adj[p]];
      if (j < lo) ok = false;
      while (j != -1 && j < i) {
        const int32_t next = anc[j];
        anc[j] = i;
        if (next == -1) {
          parent[j] = i;
          break;
        }
        j = next;
      }
    }
    return ok;
  };
  std::vector<char> in_range;
  if (!ranges.empty()) {
    in_range.assign(n, 0);
    for (const auto& r : ranges) {
      if (r.first < 0 || r.second > n || r.first > r.second) throw std::runtime_error("etree: bad range");
      for (int32_t i = r.first; i < r.second; ++i) {
        if (in_range[i]) throw std::runtime_error("etree: overlapping ranges");
        in_range[i] = 1;
      }
    }
    std::atomic<size_t> next{0};
    std::atomic<bool> closed{true};
    auto worker = [&]() {
      for (size_t k; (k = next.fetch_add(1)) < ranges.size();)
        for (int32_t i = ranges[k].first; i < ranges[k].second; ++i)
          if (!column(i, ranges[k].first)) closed = false;
    };
    const int nt = std::min<int>((int)ranges.size(),
                                 std::min(16, (int)std::max(1u, std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    if (!closed) {
      in_range.clear();
      std::fill(parent.begin(), parent.end(), -1);
      std::fill(anc.begin(), anc.end(), -1);
    }
  }
  for (int32_t i = 0; i < n; ++i)
    if (in_range.empty() || !in_range[i]) column(i, 0);
  return parent;
}

std::vector<int32_t> postorder(int32_t n, const std::vector<int32_t>& parent) {
  std::vector<int32_t> head(n, -1), next(n, -1);
  for (int32_t j = n - 1; j >= 0; --j) {
    if (parent[j] < 0) continue;
    next[j] = head[parent[j]];
    head[parent[j]] = j;
  }
  std::vector<int32_t> post;
  post.reserve(n);
  std::vector<int32_t> stack;
  for (int32_t r = 0; r < n; ++r) {
    if (parent[r] != -1) continue;
    stack.push_back(r);
    while (!stack.empty()) {
      const int32_t v = stack.back();
      const int32_t c = head[v];
      if (c == -1) {
        post.push_back(v);
        stack.pop_back();
      } else {
        head[v] = next[c];
        stack.push_back(c);
      }
    }
  }
  return post;
}

// Column counts (incl. diagonal) of L for a postordered etree (post = identity here) by the
// row-subtree skeleton method.
std::vector<int64_t> column_counts(int32_t n, const std::vector<int32_t>& parent, const std::vector<int64_t>& cp,
                                   const std::vector<int32_t>& ri) {
  std::vector<int32_t> first(n, -1), maxfirst(n, -1), prevleaf(n, -1), anc(n);
  std::vector<int64_t> delta(n, 0);
  // first[j]: smallest postorder number in the subtree of j (identity postorder)
  for (int32_t k = 0; k < n; ++k) {
    int32_t j = k;
    delta[j] = first[j] == -1 ? 1 : 0;
    for (; j != -1 && first[j] == -1; j = parent[j]) first[j] = k;
  }
  std::iota(anc.begin(), anc.end(), 0);
  for (int32_t j = 0; j < n; ++j) {
    if (parent[j] != -1) --delta[parent[j]];
    for (int64_t p = cp[j]; p < cp[j + 1]; ++p) {
      const int32_t i = ri[p];  // i > j: entry (i, j) of the lower pattern
      if (first[j] <= maxfirst[i]) continue;  // j is not a leaf of row subtree i
      maxfirst[i] = first[j];
      const int32_t jprev = prevleaf[i];
      prevleaf[i] = j;
      ++delta[j];
      if (jprev != -1) {
        int32_t q = jprev;
        while (q != anc[q]) q = anc[q];
        for (int32_t s = jprev; s != q;) {
          const int32_t sp = anc[s];
          anc[s] = q;
          s = sp;
        }
        --delta[q];
      }
    }
    if (parent[j] != -1) anc[j] = parent[j];
  }
  for (int32_t j = 0; j < n; ++j)
    if (parent[j] != -1) delta[parent[j]] += delta[j];
  return delta;
}

}  // namespace

Symbolic analyze_pattern(int32_t n, const int64_t* rowptr, const int32_t* colind, const OrderingOptions& oopt,
                         const AmalgamationOptions& aopt, const int32_t* given_order) {
  if (n <= 0) throw std::runtime_error("analyze_pattern: empty matrix");
  Graph g = graph_from_csr(n, rowptr, colind);
  // ranges of the dissection for the elimination tree on several threads (large problems only)
  std::vector<std::pair<int32_t, int32_t>> ranges;
  std::vector<int32_t> order = given_order ? std::vector<int32_t>(given_order, given_order + n)
                               : nested_dissection(g, oopt, n >= 65536 ? &ranges : nullptr, n / 32);
  std::vector<int32_t> iorder(n, -1);
  for (int32_t k = 0; k < n; ++k) {
    if (order[k] < 0 || order[k] >= n || iorder[order[k]] != -1)
      throw std::runtime_error("analyze_pattern: invalid ordering");
    iorder[order[k]] = k;
  }
  std::vector<int32_t> parent = etree_from_graph(n, g, order, iorder, ranges);
  // postorder and compose; a postorder is a topological order of the tree, so the tree of the
  // postordered pattern is the same tree relabelled (no second elimination-tree pass)
  {
    const std::vector<int32_t> post = postorder(n, parent);
    std::vector<int32_t> ipost(n), order2(n), parent2(n);
    for (int32_t k = 0; k < n; ++k) ipost[post[k]] = k;
    for (int32_t k = 0; k < n; ++k) {
      order2[k] = order[post[k]];
      const int32_t pa = parent[post[k]];
      parent2[k] = pa < 0 ? -1 : ipost[pa];
    }
    order.swap(order2);
    parent.swap(parent2);
    for (int32_t k = 0; k < n; ++k) iorder[order[k]] = k;
  }
  std::vector<int64_t> cp;
  std::vector<int32_t> ri;
  permuted_lower(n, g, order, iorder, cp, ri);
  const std::vector<int64_t> cc = column_counts(n, parent, cp, ri);
  // number of children per column
  std::vector<int32_t> nchild(n, 0);
  for (int32_t j = 0; j < n; ++j)
    if (parent[j] >= 0) ++nchild[parent[j]];
  // fundamental supernodes
  std::vector<int32_t> fs_first;
  for (int32_t j = 0; j < n; ++j) {
    const bool merge = j > 0 && parent[j - 1] == j && nchild[j] == 1 && cc[j - 1] == cc[j] + 1;
    if (!merge) fs_first.push_back(j);
  }
  const int32_t nfs = static_cast<int32_t>(fs_first.size());
  fs_first.push_back(n);
  std::vector<int32_t> col_fs(n);
  for (int32_t s = 0; s < nfs; ++s)
    for (int32_t j = fs_first[s]; j < fs_first[s + 1]; ++j) col_fs[j] = s;
  // supernode tree and true entries
  std::vector<int32_t> fs_parent(nfs, -1);
  std::vector<int64_t> fs_m(nfs), fs_entries(nfs, 0);
  std::vector<int32_t> fs_ncol(nfs);
  for (int32_t s = 0; s < nfs; ++s) {
    const int32_t last = fs_first[s + 1] - 1;
    fs_parent[s] = parent[last] >= 0 ? col_fs[parent[last]] : -1;
    fs_ncol[s] = fs_first[s + 1] - fs_first[s];
    fs_m[s] = cc[fs_first[s]];
    for (int32_t j = fs_first[s]; j < fs_first[s + 1]; ++j) fs_entries[s] += cc[j];
  }
  // relaxed amalgamation: supernode s absorbs its last child c (c ends right before s starts)
  // while the merged front stays within the relaxation rule. Merged nodes are tracked by
  // union-find on the fundamental supernodes; columns stay contiguous.
  std::vector<int32_t> rep(nfs);
  std::iota(rep.begin(), rep.end(), 0);
  std::vector<int32_t> m_first(fs_first.begin(), fs_first.end() - 1), m_ncol = fs_ncol;
  std::vector<int64_t> m_m = fs_m, m_entries = fs_entries;
  std::vector<char> absorbed(nfs, 0);
  for (int32_t s = 0; s < nfs; ++s) {
    for (;;) {
      const int32_t f = m_first[s];
      if (f == 0) break;
      const int32_t c = rep[col_fs[f - 1]];
      // c must be a child of s in the amalgamated tree
      int32_t pc = fs_parent[c];
      if (pc < 0) break;
      while (rep[pc] != pc) pc = rep[pc];
      if (pc != s) break;
      const int64_t ncol = m_ncol[c] + m_ncol[s];
      const int64_t mm = m_ncol[c] + m_m[s];
      const double dense = static_cast<double>(ncol) * static_cast<double>(mm) -
                           0.5 * static_cast<double>(ncol) * static_cast<double>(ncol - 1);
      const double zeros = dense - static_cast<double>(m_entries[c] + m_entries[s]);
      bool ok = false;
      for (int r = 0; r < 3 && !ok; ++r)
        ok = ncol <= aopt.ncol[r] && zeros <= aopt.zero_frac[r] * dense;
      if (!ok) break;
      // merge c into s
      rep[c] = s;
      absorbed[c] = 1;
      m_first[s] = m_first[c];
      m_ncol[s] = static_cast<int32_t>(ncol);
      m_m[s] = mm;
      m_entries[s] = m_entries[c] + m_entries[s];  // true entries; the zeros are re-derived
    }
  }
  auto find = [&](int32_t x) {
    while (rep[x] != x) x = rep[x];
    return x;
  };
  Symbolic S;
  S.n = n;
  S.order = order;
  S.iorder = iorder;
  std::vector<int32_t> sn_of_fs(nfs, -1);
  for (int32_t s = 0; s < nfs; ++s)
    if (!absorbed[s]) {
      sn_of_fs[s] = S.nsn++;
      S.sn_first.push_back(m_first[s]);
      S.sn_ncol.push_back(m_ncol[s]);
    }
  // supernodes are in increasing first-column order because fundamental ones are and every kept
  // node keeps its position relative to the others (it only absorbs columns right before it)
  S.col_sn.assign(n, -1);
  for (int32_t s = 0; s < S.nsn; ++s)
    for (int32_t j = S.sn_first[s]; j < S.sn_first[s] + S.sn_ncol[s]; ++j) S.col_sn[j] = s;
  for (int32_t j = 0; j < n; ++j)
    if (S.col_sn[j] < 0) throw std::runtime_error("analyze_pattern: column without supernode");
  S.sn_parent.assign(S.nsn, -1);
  for (int32_t s = 0; s < S.nsn; ++s) {
    const int32_t last = S.sn_first[s] + S.sn_ncol[s] - 1;
    S.sn_parent[s] = parent[last] >= 0 ? S.col_sn[parent[last]] : -1;
    if (S.sn_parent[s] >= 0 && S.sn_parent[s] <= s) throw std::runtime_error("analyze_pattern: tree not postordered");
  }
  (void)find;
  // children lists
  S.child_ptr.assign(static_cast<size_t>(S.nsn) + 1, 0);
  for (int32_t s = 0; s < S.nsn; ++s)
    if (S.sn_parent[s] >= 0) ++S.child_ptr[S.sn_parent[s] + 1];
  for (int32_t s = 0; s < S.nsn; ++s) S.child_ptr[s + 1] += S.child_ptr[s];
  S.child_list.resize(S.child_ptr[S.nsn]);
  {
    std::vector<int32_t> fill(S.child_ptr.begin(), S.child_ptr.end() - 1);
    for (int32_t s = 0; s < S.nsn; ++s)
      if (S.sn_parent[s] >= 0) S.child_list[fill[S.sn_parent[s]]++] = s;
  }
  // levels (height): children first
  S.level.assign(S.nsn, 0);
  for (int32_t s = 0; s < S.nsn; ++s)
    for (int32_t q = S.child_ptr[s]; q < S.child_ptr[s + 1]; ++q)
      S.level[s] = std::max(S.level[s], S.level[S.child_list[q]] + 1);
  S.nlevels = 0;
  for (int32_t s = 0; s < S.nsn; ++s) S.nlevels = std::max(S.nlevels, S.level[s] + 1);
  S.level_ptr.assign(static_cast<size_t>(S.nlevels) + 1, 0);
  for (int32_t s = 0; s < S.nsn; ++s) ++S.level_ptr[S.level[s] + 1];
  for (int32_t l = 0; l < S.nlevels; ++l) S.level_ptr[l + 1] += S.level_ptr[l];
  S.level_sn.resize(S.nsn);
  {
    std::vector<int32_t> fill(S.level_ptr.begin(), S.level_ptr.end() - 1);
    for (int32_t s = 0; s < S.nsn; ++s) S.level_sn[fill[S.level[s]]++] = s;
  }
  // row structures, level by level (a front's children lie on lower levels): the fronts of a level
  // with many of them on up to 16 threads, each thread with its own marks; the lists are then
  // concatenated in front order, so the result does not depend on the thread count
  S.sn_rowoff.assign(static_cast<size_t>(S.nsn) + 1, 0);
  S.sn_m.assign(S.nsn, 0);
  {
    std::vector<std::vector<int32_t>> rl(S.nsn);
    const int nthr = n < 8192 ? 1 : std::min(16, (int)std::max(1u, std::thread::hardware_concurrency()));
    std::vector<std::vector<int32_t>> marks(nthr);
    std::vector<char> bad(nthr, 0);
    auto fronts = [&](int t, int32_t k0, int32_t k1) {
      std::vector<int32_t>& mark = marks[t];
      if (mark.empty()) mark.assign(n, -1);
      for (int32_t k = k0; k < k1; ++k) {
        const int32_t s = S.level_sn[k];
        const int32_t f = S.sn_first[s], l = f + S.sn_ncol[s];
        std::vector<int32_t>& tmp = rl[s];
        for (int32_t j = f; j < l; ++j) {
          mark[j] = s;
          tmp.push_back(j);
        }
        for (int32_t j = f; j < l; ++j)
          for (int64_t p = cp[j]; p < cp[j + 1]; ++p) {
            const int32_t i = ri[p];
            if (mark[i] != s) {
              mark[i] = s;
              tmp.push_back(i);
            }
          }
        for (int32_t q = S.child_ptr[s]; q < S.child_ptr[s + 1]; ++q) {
          const std::vector<int32_t>& cr = rl[S.child_list[q]];
          for (size_t p = static_cast<size_t>(S.sn_ncol[S.child_list[q]]); p < cr.size(); ++p) {
            const int32_t i = cr[p];
            if (mark[i] != s) {
              mark[i] = s;
              tmp.push_back(i);
            }
          }
        }
        std::sort(tmp.begin() + (l - f), tmp.end());
        for (size_t q = static_cast<size_t>(l - f); q < tmp.size(); ++q)
          if (tmp[q] < l) bad[t] = 1;
      }
    };
    // one pool of threads for all levels: the fronts of a level are handed out one by one by an
    // atomic counter, and the threads meet at a barrier between levels (a front's children lie on
    // lower levels)
    std::vector<std::atomic<int32_t>> next(S.nlevels);
    for (int32_t lv = 0; lv < S.nlevels; ++lv) next[lv].store(S.level_ptr[lv], std::memory_order_relaxed);
    std::mutex bmu;
    std::condition_variable bcv;
    int arrived = 0, generation = 0;
    auto barrier = [&]() {
      std::unique_lock<std::mutex> lk(bmu);
      const int gen = generation;
      if (++arrived == nthr) {
        arrived = 0;
        ++generation;
        bcv.notify_all();
      } else {
        bcv.wait(lk, [&] { return generation != gen; });
      }
    };
    auto run = [&](int t) {
      for (int32_t lv = 0; lv < S.nlevels; ++lv) {
        const int32_t k1 = S.level_ptr[lv + 1];
        for (int32_t k; (k = next[lv].fetch_add(1, std::memory_order_relaxed)) < k1;) fronts(t, k, k + 1);
        if (nthr > 1) barrier();
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nthr; ++t) pool.emplace_back(run, t);
    run(0);
    for (auto& th : pool) th.join();
    for (char b : bad)
      if (b) throw std::runtime_error("analyze_pattern: row above the supernode's columns");
    for (int32_t s = 0; s < S.nsn; ++s) {
      S.sn_rowoff[s + 1] = S.sn_rowoff[s] + static_cast<int64_t>(rl[s].size());
      S.sn_m[s] = static_cast<int32_t>(rl[s].size());
    }
    S.rows.resize(static_cast<size_t>(S.sn_rowoff[S.nsn]));
    for (int32_t s = 0; s < S.nsn; ++s) std::copy(rl[s].begin(), rl[s].end(), S.rows.begin() + S.sn_rowoff[s]);
  }
  // extend-add maps: local index of each child CB row in the parent's front
  // (fronts are independent: over ranges of fronts on several threads, errors flagged per range)
  S.relmap.assign(S.rows.size(), -1);
  {
    std::vector<char> bad_root(S.nsn, 0), bad_row(S.nsn, 0);
    over_ranges(S.nsn, [&](int32_t c0, int32_t c1) {
      for (int32_t c = c0; c < c1; ++c) {
        const int32_t p = S.sn_parent[c];
        if (p < 0) {
          if (S.sn_m[c] != S.sn_ncol[c]) bad_root[c] = 1;
          continue;
        }
        const int32_t* prow = &S.rows[S.sn_rowoff[p]];
        const int32_t pm = S.sn_m[p];
        int32_t q = 0;
        for (int64_t k = S.sn_rowoff[c] + S.sn_ncol[c]; k < S.sn_rowoff[c + 1]; ++k) {
          const int32_t i = S.rows[k];
          while (q < pm && prow[q] < i) ++q;
          if (q >= pm || prow[q] != i) {
            bad_row[c] = 1;
            break;
          }
          S.relmap[k] = q;
        }
      }
    });
    for (int32_t c = 0; c < S.nsn; ++c) {
      if (bad_root[c]) throw std::runtime_error("analyze_pattern: root with contribution rows");
      if (bad_row[c]) throw std::runtime_error("analyze_pattern: child row missing in parent");
    }
  }
  // offsets, statistics
  S.panel_off.assign(static_cast<size_t>(S.nsn) + 1, 0);
  for (int32_t s = 0; s < S.nsn; ++s) {
    const int64_t m = S.sn_m[s], nc = S.sn_ncol[s], cb = m - nc;
    S.panel_off[s + 1] = S.panel_off[s] + m * nc;
    S.cb_sum += cb * cb;
    S.factor_entries += static_cast<double>(m * nc) - 0.5 * static_cast<double>(nc * (nc - 1));
    const double pp = static_cast<double>(nc), c = static_cast<double>(cb);
    S.flops += pp * pp * pp / 3.0 + pp * pp * c + pp * c * c;
    S.max_front = std::max<int32_t>(S.max_front, static_cast<int32_t>(m));
    S.max_ncol = std::max<int32_t>(S.max_ncol, static_cast<int32_t>(nc));
  }
  plan_contribution_blocks(S);
  return S;
}

void plan_contribution_blocks(Symbolic& S) {
  // A front's contribution block is written while its level is factorized and read while its
  // parent's level is (levels run one after the other): it is live over the levels
  // [level(s), level(parent)]. Offsets are assigned level by level over one address range: at the
  // start of level l the blocks whose parent level was l - 1 are released (coalescing free
  // neighbours; a free range that reaches the top lowers the top), then the level's blocks are
  // placed, largest first, each in the smallest free range that holds it (best fit, found by
  // length in a sorted map), else at the top. Blocks whose lifetimes overlap never share storage.
  const int32_t nsn = S.nsn;
  S.cb_off.assign(static_cast<size_t>(nsn) + 1, 0);
  std::vector<std::vector<int32_t>> ending(static_cast<size_t>(std::max(1, S.nlevels)));
  std::map<int64_t, int64_t> by_off;            // free ranges: offset -> length
  std::multimap<int64_t, int64_t> by_len;       // the same ranges: length -> offset
  int64_t top = 0;
  auto drop_len = [&](int64_t len, int64_t off) {
    auto r = by_len.equal_range(len);
    for (auto it = r.first; it != r.second; ++it)
      if (it->second == off) {
        by_len.erase(it);
        return;
      }
  };
  auto release = [&](int64_t off, int64_t len) {
    if (len == 0) return;
    auto nx = by_off.lower_bound(off);
    if (nx != by_off.end() && off + len == nx->first) {  // merge with the range after
      len += nx->second;
      drop_len(nx->second, nx->first);
      nx = by_off.erase(nx);
    }
    if (nx != by_off.begin()) {  // merge with the range before
      auto pv = std::prev(nx);
      if (pv->first + pv->second == off) {
        off = pv->first;
        len += pv->second;
        drop_len(pv->second, pv->first);
        by_off.erase(pv);
      }
    }
    if (off + len == top) {
      top = off;
      return;
    }
    by_off.emplace(off, len);
    by_len.emplace(len, off);
  };
  auto place = [&](int64_t len) -> int64_t {
    auto it = by_len.lower_bound(len);
    if (it == by_len.end()) {
      const int64_t off = top;
      top += len;
      return off;
    }
    const int64_t flen = it->first, off = it->second;
    by_len.erase(it);
    by_off.erase(off);
    if (flen > len) {
      by_off.emplace(off + len, flen - len);
      by_len.emplace(flen - len, off + len);
    }
    return off;
  };
  int64_t peak = 0;
  for (int32_t l = 0; l < S.nlevels; ++l) {
    if (l > 0)
      for (int32_t s : ending[l - 1]) {
        const int64_t cb = S.sn_m[s] - S.sn_ncol[s];
        release(S.cb_off[s], cb * cb);
      }
    std::vector<int32_t> here(S.level_sn.begin() + S.level_ptr[l], S.level_sn.begin() + S.level_ptr[l + 1]);
    std::stable_sort(here.begin(), here.end(), [&](int32_t a, int32_t b) {
      return S.sn_m[a] - S.sn_ncol[a] > S.sn_m[b] - S.sn_ncol[b];
    });
    for (int32_t s : here) {
      const int64_t cb = S.sn_m[s] - S.sn_ncol[s];
      S.cb_off[s] = cb > 0 ? place(cb * cb) : 0;
      const int32_t end = S.sn_parent[s] >= 0 ? S.level[S.sn_parent[s]] : l;
      ending[static_cast<size_t>(end)].push_back(s);
    }
    peak = std::max(peak, top);
  }
  S.cb_off[nsn] = peak;
  S.cb_total = peak;
}

std::vector<int64_t> assembly_map(const Symbolic& S, int32_t n, const int64_t* rowptr, const int32_t* colind) {
  std::vector<int64_t> dst(static_cast<size_t>(rowptr[n]), -1);
  // rows are independent: contiguous row ranges on up to 16 threads (each entry is written once)
  auto rows = [&](int32_t i0, int32_t i1, bool& bad) {
    for (int32_t i = i0; i < i1; ++i) {
      const int32_t in = S.iorder[i];
      for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
        const int32_t jn = S.iorder[colind[p]];
        if (in < jn) continue;  // mirrored copy above the diagonal
        const int32_t s = S.col_sn[jn];
        const int32_t* r = &S.rows[S.sn_rowoff[s]];
        const int32_t m = S.sn_m[s];
        const int32_t* it = std::lower_bound(r, r + m, in);
        if (it == r + m || *it != in) {
          bad = true;
          return;
        }
        const int64_t lr = it - r, lc = jn - S.sn_first[s];
        dst[p] = S.panel_off[s] + lc * m + lr;
      }
    }
  };
  const int nt = n < 8192 ? 1 : std::min(16, (int)std::max(1u, std::thread::hardware_concurrency()));
  std::vector<char> bad(nt, 0);
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t)
    pool.emplace_back([&, t] {
      bool b = false;
      rows((int32_t)((int64_t)n * t / nt), (int32_t)((int64_t)n * (t + 1) / nt), b);
      bad[t] = b;
    });
  {
    bool b = false;
    rows(0, (int32_t)((int64_t)n / nt), b);
    bad[0] = b;
  }
  for (auto& th : pool) th.join();
  for (char b : bad)
    if (b) throw std::runtime_error("assembly_map: entry outside the front");
  return dst;
}

}  // namespace hsg
