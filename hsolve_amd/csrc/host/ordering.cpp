// Nested-dissection fill-reducing ordering (host).
//
// Plays the role of hsolve's own nested dissection (src/factor/ordering.cpp, contract
// include/hsolve/factor.hpp:197-246, 598-633: multilevel separators, refinement, minimum-degree
// leaves), written independently: every subgraph larger than a leaf is bisected by a multilevel
// edge partition (heavy-edge matching coarsening, greedy graph growing on the coarsest graph,
// Fiduccia-Mattheyses boundary refinement while uncoarsening); the cut edges' boundary on the
// smaller side becomes the vertex separator, which is then made minimal. Separators are numbered
// last; leaves get a minimum-degree order computed on bit-set elimination graphs. The result is
// deterministic (a fixed-seed generator drives the matching order).
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <exception>
#include <mutex>
#include <numeric>
#include <queue>
#include <stdexcept>
#include <thread>
#include <utility>
#include <vector>

#include "analysis.hpp"

namespace hsg {

namespace {

// Threads for a loop over `items` independent items: one below `min_items`, else up to 16.
int worker_count(int64_t items, int64_t min_items) {
  if (items < min_items) return 1;
  return std::min(16, static_cast<int>(std::max(1u, std::thread::hardware_concurrency())));
}

// body(t, i0, i1) for the nt contiguous ranges [i0, i1) of [0, n), range 0 on the calling thread.
template <class F>
void parallel_ranges(int32_t n, int nt, F&& body) {
  auto range = [&](int t) {
    body(t, static_cast<int32_t>(int64_t{n} * t / nt), static_cast<int32_t>(int64_t{n} * (t + 1) / nt));
  };
  if (nt <= 1) {
    range(0);
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back([&, t] { range(t); });
  range(0);
  for (auto& th : pool) th.join();
}

}  // namespace

Graph graph_from_csr(int32_t n, const int64_t* rowptr, const int32_t* colind) {
  // Count both directions, then sort/unique each row.
  std::vector<int64_t> deg(static_cast<size_t>(n) + 1, 0);
  for (int32_t i = 0; i < n; ++i)
    for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
      const int32_t j = colind[p];
      if (j == i) continue;
      if (j < 0 || j >= n) throw std::runtime_error("pattern column index out of range");
      ++deg[i + 1];
      ++deg[j + 1];
    }
  for (int32_t i = 0; i < n; ++i) deg[i + 1] += deg[i];
  std::vector<int32_t> tmp(static_cast<size_t>(deg[n]));
  std::vector<int64_t> fill(deg.begin(), deg.end() - 1);
  for (int32_t i = 0; i < n; ++i)
    for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
      const int32_t j = colind[p];
      if (j == i) continue;
      tmp[fill[i]++] = j;
      tmp[fill[j]++] = i;
    }
  // rows sorted and deduplicated over row ranges on several threads, concatenated in row order
  Graph g;
  g.n = n;
  g.ptr.assign(static_cast<size_t>(n) + 1, 0);
  const int nt = worker_count(n, 32768);
  std::vector<int64_t> kept(static_cast<size_t>(n) + 1, 0);
  parallel_ranges(n, nt, [&](int, int32_t i0, int32_t i1) {
    for (int32_t i = i0; i < i1; ++i) {
      auto b = tmp.begin() + deg[i], e = tmp.begin() + deg[i + 1];
      std::sort(b, e);
      kept[i + 1] = std::unique(b, e) - b;
    }
  });
  for (int32_t i = 0; i < n; ++i) g.ptr[i + 1] = g.ptr[i] + kept[i + 1];
  g.adj.resize(static_cast<size_t>(g.ptr[n]));
  parallel_ranges(n, nt, [&](int, int32_t i0, int32_t i1) {
    for (int32_t i = i0; i < i1; ++i)
      std::copy(tmp.begin() + deg[i], tmp.begin() + deg[i] + kept[i + 1], g.adj.begin() + g.ptr[i]);
  });
  return g;
}

namespace {

// Small deterministic generator (splitmix64) for tie-breaking orders.
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
};

// Weighted graph of one multilevel stage.
struct WGraph {
  int32_t n = 0;
  std::vector<int64_t> ptr;
  std::vector<int32_t> adj;
  std::vector<int32_t> ew;  // edge weights
  std::vector<int32_t> vw;  // vertex weights
  int64_t total_vw = 0;
};

// Heavy-edge matching: returns coarse graph and the fine->coarse map.
bool coarsen(const WGraph& g, WGraph& c, std::vector<int32_t>& cmap, Rng& rng) {
  const int32_t n = g.n;
  std::vector<int32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  for (int32_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[rng.next() % (uint64_t)(i + 1)]);
  std::vector<int32_t> match(n, -1);
  const int32_t max_vw = static_cast<int32_t>(std::max<int64_t>(1, g.total_vw / 20));
  for (int32_t v : perm) {
    if (match[v] >= 0) continue;
    int32_t best = -1, bw = -1;
    for (int64_t p = g.ptr[v]; p < g.ptr[v + 1]; ++p) {
      const int32_t u = g.adj[p];
      if (match[u] >= 0 || u == v) continue;
      if (g.vw[u] + g.vw[v] > max_vw) continue;
      if (g.ew[p] > bw) {
        bw = g.ew[p];
        best = u;
      }
    }
    if (best >= 0) {
      match[v] = best;
      match[best] = v;
    } else {
      match[v] = v;
    }
  }
  cmap.assign(n, -1);
  int32_t nc = 0;
  for (int32_t v = 0; v < n; ++v) {
    if (cmap[v] >= 0) continue;
    cmap[v] = nc;
    if (match[v] != v) cmap[match[v]] = nc;
    ++nc;
  }
  if (nc > static_cast<int32_t>(0.92 * n)) return false;  // too little reduction
  c.n = nc;
  c.vw.assign(nc, 0);
  for (int32_t v = 0; v < n; ++v) c.vw[cmap[v]] += g.vw[v];
  c.total_vw = g.total_vw;
  // Visit coarse vertices in order; the fine members are v and match[v].
  std::vector<int32_t> first(nc, -1), second(nc, -1);
  for (int32_t v = 0; v < n; ++v) {
    const int32_t cv = cmap[v];
    if (first[cv] < 0) first[cv] = v;
    else second[cv] = v;
  }
  // coarse rows are independent: built over ranges of coarse vertices, each range into vectors of
  // its own thread (moved out at the end, so no two threads write near the same cache line), then
  // concatenated in order
  const int nt = std::min(8, worker_count(nc, 24576));
  std::vector<std::vector<int32_t>> radj(nt), rew(nt);
  std::vector<int64_t> rdeg(static_cast<size_t>(nc) + 1, 0);
  parallel_ranges(nc, nt, [&](int t, int32_t b0, int32_t e0) {
    std::vector<int32_t> A, W;
    A.reserve(g.adj.size() / 2 / static_cast<size_t>(nt) + 16);
    W.reserve(A.capacity());
    std::vector<int32_t> slot(nc, -1);  // position of coarse neighbour in the current row
    for (int32_t cv = b0; cv < e0; ++cv) {
      const size_t row_start = A.size();
      for (int k = 0; k < 2; ++k) {
        const int32_t v = k == 0 ? first[cv] : second[cv];
        if (v < 0) continue;
        for (int64_t p = g.ptr[v]; p < g.ptr[v + 1]; ++p) {
          const int32_t cu = cmap[g.adj[p]];
          if (cu == cv) continue;
          if (slot[cu] >= 0) {
            W[slot[cu]] += g.ew[p];
          } else {
            slot[cu] = static_cast<int32_t>(A.size());
            A.push_back(cu);
            W.push_back(g.ew[p]);
          }
        }
      }
      for (size_t q = row_start; q < A.size(); ++q) slot[A[q]] = -1;
      rdeg[cv + 1] = static_cast<int64_t>(A.size() - row_start);
    }
    radj[t] = std::move(A);
    rew[t] = std::move(W);
  });
  c.ptr.assign(static_cast<size_t>(nc) + 1, 0);
  for (int32_t cv = 0; cv < nc; ++cv) c.ptr[cv + 1] = c.ptr[cv] + rdeg[cv + 1];
  c.adj.resize(static_cast<size_t>(c.ptr[nc]));
  c.ew.resize(static_cast<size_t>(c.ptr[nc]));
  for (int t = 0; t < nt; ++t) {
    const int32_t b0 = static_cast<int32_t>(int64_t{nc} * t / nt);
    std::copy(radj[t].begin(), radj[t].end(), c.adj.begin() + c.ptr[b0]);
    std::copy(rew[t].begin(), rew[t].end(), c.ew.begin() + c.ptr[b0]);
  }
  return true;
}

int64_t cut_of(const WGraph& g, const std::vector<int8_t>& part) {
  int64_t cut = 0;
  for (int32_t v = 0; v < g.n; ++v)
    for (int64_t p = g.ptr[v]; p < g.ptr[v + 1]; ++p)
      if (part[v] != part[g.adj[p]]) cut += g.ew[p];
  return cut / 2;
}

// Fiduccia-Mattheyses style refinement of a 2-way edge partition with a balance bound; several
// passes, each moving vertices by best gain with rollback to the best prefix.
constexpr size_t kFmStall = 100;
void fm_refine(const WGraph& g, std::vector<int8_t>& part, double max_frac, int passes) {
  // Gains and external weights are computed once and then kept exact by the same update a move
  // makes (also when a pass rolls moves back), so a pass costs its moves, not a sweep of the
  // graph; a pass starts from the boundary vertices (external weight > 0), kept in a list that
  // only grows (entries whose external weight fell to 0 are skipped). The result is the one a
  // full recount per pass gives: the heap order depends on (gain, vertex) values only.
  const int32_t n = g.n;
  const int64_t maxw = static_cast<int64_t>(max_frac * static_cast<double>(g.total_vw)) + 1;
  std::vector<int64_t> pw(2, 0);
  for (int32_t v = 0; v < n; ++v) pw[part[v]] += g.vw[v];
  std::vector<int32_t> gain(n), ext(n);
  std::vector<char> moved(n, 0), listed(n, 0);
  std::vector<int32_t> boundary;
  for (int32_t v = 0; v < n; ++v) {
    int32_t e = 0, i = 0;
    for (int64_t p = g.ptr[v]; p < g.ptr[v + 1]; ++p) (part[g.adj[p]] != part[v] ? e : i) += g.ew[p];
    gain[v] = e - i;
    ext[v] = e;
    if (e > 0) {
      boundary.push_back(v);
      listed[v] = 1;
    }
  }
  // moves v to the other side and updates its neighbours' gains and external weights
  auto flip = [&](int32_t v) {
    const int to = 1 - part[v];
    pw[part[v]] -= g.vw[v];
    pw[to] += g.vw[v];
    part[v] = static_cast<int8_t>(to);
    gain[v] = -gain[v];
    int32_t e = 0;
    for (int64_t p = g.ptr[v]; p < g.ptr[v + 1]; ++p) {
      const int32_t u = g.adj[p];
      const int32_t w = g.ew[p];
      if (part[u] == to) {
        gain[u] -= 2 * w;
        ext[u] -= w;
      } else {
        gain[u] += 2 * w;
        ext[u] += w;
        e += w;
        if (!listed[u]) {
          listed[u] = 1;
          boundary.push_back(u);
        }
      }
    }
    ext[v] = e;
    if (e > 0 && !listed[v]) {
      listed[v] = 1;
      boundary.push_back(v);
    }
  };
  for (int pass = 0; pass < passes; ++pass) {
    using Item = std::pair<int32_t, int32_t>;  // (gain, -vertex) max-heap
    std::priority_queue<Item> heap[2];
    for (int32_t v : boundary)
      if (ext[v] > 0) heap[part[v]].push({gain[v], -v});
    std::vector<int32_t> seq;
    int64_t cur = 0, best = 0;
    size_t best_len = 0;
    const size_t limit = static_cast<size_t>(std::max<int32_t>(50, n / 10));
    // a pass also ends after kFmStall moves in a row that did not improve on the best cut: on
    // large graphs the moves beyond that are almost always rolled back
    while (seq.size() < limit && seq.size() - best_len <= kFmStall) {
      // pick the side to move from: the heavier side first if it has candidates
      int from = pw[0] >= pw[1] ? 0 : 1;
      int32_t v = -1;
      for (int attempt = 0; attempt < 2 && v < 0; ++attempt) {
        auto& h = heap[from];
        while (!h.empty()) {
          const auto [gv, nv] = h.top();
          const int32_t cand = -nv;
          if (moved[cand] || part[cand] != from || gv != gain[cand]) {
            h.pop();
            continue;
          }
          if (pw[1 - from] + g.vw[cand] > maxw) break;
          v = cand;
          h.pop();
          break;
        }
        if (v < 0) from = 1 - from;
      }
      if (v < 0) break;
      cur -= gain[v];
      flip(v);
      moved[v] = 1;
      seq.push_back(v);
      for (int64_t p = g.ptr[v]; p < g.ptr[v + 1]; ++p) {
        const int32_t u = g.adj[p];
        if (!moved[u]) heap[part[u]].push({gain[u], -u});
      }
      if (cur < best) {
        best = cur;
        best_len = seq.size();
      }
    }
    // roll back beyond the best prefix (the inverse moves, in reverse order)
    for (size_t k = seq.size(); k > best_len; --k) flip(seq[k - 1]);
    for (int32_t v : seq) moved[v] = 0;
    if (best_len == 0) break;
  }
}

// Greedy graph growing from seed: BFS adding vertices to part 0 until half the weight.
void grow_from(const WGraph& g, int32_t seed, std::vector<int8_t>& part) {
  part.assign(g.n, 1);
  int64_t w = 0;
  const int64_t half = g.total_vw / 2;
  std::vector<int32_t> queue;
  queue.reserve(g.n);
  std::vector<char> seen(g.n, 0);
  queue.push_back(seed);
  seen[seed] = 1;
  size_t head = 0;
  while (w < half) {
    if (head == queue.size()) {  // disconnected: take any unseen vertex
      int32_t v = -1;
      for (int32_t u = 0; u < g.n; ++u)
        if (!seen[u]) {
          v = u;
          break;
        }
      if (v < 0) break;
      seen[v] = 1;
      queue.push_back(v);
    }
    const int32_t v = queue[head++];
    part[v] = 0;
    w += g.vw[v];
    for (int64_t p = g.ptr[v]; p < g.ptr[v + 1]; ++p) {
      const int32_t u = g.adj[p];
      if (!seen[u]) {
        seen[u] = 1;
        queue.push_back(u);
      }
    }
  }
}

int32_t bfs_farthest(const WGraph& g, int32_t s) {
  std::vector<int32_t> dist(g.n, -1);
  std::vector<int32_t> q{s};
  dist[s] = 0;
  int32_t last = s;
  for (size_t h = 0; h < q.size(); ++h) {
    const int32_t v = q[h];
    last = v;
    for (int64_t p = g.ptr[v]; p < g.ptr[v + 1]; ++p)
      if (dist[g.adj[p]] < 0) {
        dist[g.adj[p]] = dist[v] + 1;
        q.push_back(g.adj[p]);
      }
  }
  return last;
}

// Multilevel 2-way edge partition of g; part[v] in {0,1}.
void multilevel_bisect(const WGraph& g0, std::vector<int8_t>& part, Rng& rng) {
  // coarse graphs by level; level 0 is g0 itself (not copied)
  std::vector<WGraph> coarse;
  std::vector<std::vector<int32_t>> maps;
  auto level = [&](size_t l) -> const WGraph& { return l == 0 ? g0 : coarse[l - 1]; };
  while (level(coarse.size()).n > 160) {
    WGraph c;
    std::vector<int32_t> cmap;
    if (!coarsen(level(coarse.size()), c, cmap, rng)) break;
    coarse.push_back(std::move(c));
    maps.push_back(std::move(cmap));
  }
  const WGraph& gc = level(coarse.size());
  // initial partitions: grown from a few seeds (peripheral and random), keep the best cut
  std::vector<int8_t> best;
  int64_t best_cut = -1;
  std::vector<int32_t> seeds;
  seeds.push_back(bfs_farthest(gc, 0));
  seeds.push_back(bfs_farthest(gc, seeds[0]));
  for (int t = 0; t < 4 && gc.n > 0; ++t) seeds.push_back(static_cast<int32_t>(rng.next() % (uint64_t)gc.n));
  for (int32_t s : seeds) {
    std::vector<int8_t> p;
    grow_from(gc, s, p);
    fm_refine(gc, p, 0.55, 6);
    const int64_t c = cut_of(gc, p);
    if (best_cut < 0 || c < best_cut) {
      best_cut = c;
      best = p;
    }
  }
  part = std::move(best);
  for (size_t l = coarse.size(); l > 0; --l) {
    const WGraph& fine = level(l - 1);
    const std::vector<int32_t>& cmap = maps[l - 1];
    std::vector<int8_t> fp(fine.n);
    for (int32_t v = 0; v < fine.n; ++v) fp[v] = part[cmap[v]];
    fm_refine(fine, fp, 0.55, 4);
    part = std::move(fp);
  }
}

// Minimum-degree order of a small subgraph by explicit elimination on bit sets.
void min_degree_leaf(const Graph& g, const std::vector<int32_t>& verts, const std::vector<int32_t>& local,
                     std::vector<int32_t>& out) {
  const int32_t L = static_cast<int32_t>(verts.size());
  const int32_t W = (L + 63) / 64;
  std::vector<uint64_t> bits(static_cast<size_t>(L) * W, 0);
  for (int32_t a = 0; a < L; ++a) {
    const int32_t v = verts[a];
    for (int64_t p = g.ptr[v]; p < g.ptr[v + 1]; ++p) {
      const int32_t b = local[g.adj[p]];
      if (b >= 0 && b != a) bits[static_cast<size_t>(a) * W + b / 64] |= 1ull << (b % 64);
    }
  }
  std::vector<char> gone(L, 0);
  for (int32_t step = 0; step < L; ++step) {
    int32_t best = -1, bd = 1 << 30;
    for (int32_t a = 0; a < L; ++a) {
      if (gone[a]) continue;
      int32_t d = 0;
      for (int32_t w = 0; w < W; ++w) d += __builtin_popcountll(bits[static_cast<size_t>(a) * W + w]);
      if (d < bd) {
        bd = d;
        best = a;
      }
    }
    gone[best] = 1;
    out.push_back(verts[best]);
    uint64_t* rb = &bits[static_cast<size_t>(best) * W];
    for (int32_t b = 0; b < L; ++b) {
      if (!(rb[b / 64] >> (b % 64) & 1ull)) continue;
      uint64_t* r = &bits[static_cast<size_t>(b) * W];
      for (int32_t w = 0; w < W; ++w) r[w] |= rb[w];
      r[b / 64] &= ~(1ull << (b % 64));
      r[best / 64] &= ~(1ull << (best % 64));
    }
    for (int32_t w = 0; w < W; ++w) rb[w] = 0;
  }
}

}  // namespace

std::vector<int32_t> nested_dissection(const Graph& g, const OrderingOptions& opt,
                                       std::vector<std::pair<int32_t, int32_t>>* ranges, int32_t range_max) {
  // Subgraphs are independent once their separator is numbered, so they are dissected by a pool
  // of threads. Every task owns the number range [lo, hi) of its vertices and a generator seeded
  // from its position in the dissection tree, so the result does not depend on which thread
  // runs which task or when.
  const int32_t n = g.n;
  std::vector<int32_t> perm(n, -1);
  // subgraph id of each unnumbered vertex; read for neighbours that other tasks may relabel at the
  // same time (a neighbour outside the task never carries the task's id, before or after)
  std::vector<std::atomic<int32_t>> label(n);
  for (auto& l : label) l.store(0, std::memory_order_relaxed);
  struct Task {
    std::vector<int32_t> verts;
    int32_t lo = 0, hi = 0;  // numbers [lo, hi) are assigned to these vertices
    int32_t id = 0;
    uint64_t seed = 0;
    bool ranged = false;     // the task or an ancestor has its range recorded
  };
  auto child_seed = [](uint64_t parent, uint64_t k) {
    Rng r(parent ^ (0x9E3779B97F4A7C15ull * (k + 1)));
    return r.next();
  };
  std::atomic<int32_t> next_id{1};
  std::mutex mu, range_mu;
  if (ranges) ranges->clear();
  std::condition_variable cv;
  std::vector<Task> stack;
  int busy = 0;
  std::exception_ptr failure;
  {
    Task t;
    t.verts.resize(n);
    std::iota(t.verts.begin(), t.verts.end(), 0);
    t.lo = 0;
    t.hi = n;
    t.id = 0;
    t.seed = 0x5EED2026ull;
    stack.push_back(std::move(t));
  }
  // one task: a leaf gets its minimum-degree order, a component split or a bisection pushes
  // the parts as new tasks; local is the calling thread's scratch map (all -1 between tasks)
  auto process = [&](Task& t, std::vector<int32_t>& local, std::vector<Task>& out) {
    const int32_t nv = static_cast<int32_t>(t.verts.size());
    if (nv == 0) return;
    if (ranges && !t.ranged && nv <= range_max) {
      // the first subgraph on this branch small enough: its numbers [lo, hi) go to no separator
      // above it, and its vertices' neighbours outside it are such separators (numbered later)
      std::lock_guard<std::mutex> lk(range_mu);
      ranges->emplace_back(t.lo, t.hi);
      t.ranged = true;
    }
    for (int32_t v : t.verts) label[v].store(t.id, std::memory_order_relaxed);
    if (nv <= opt.leaf_size) {
      std::vector<int32_t> ord;
      ord.reserve(nv);
      if (opt.min_degree_leaves) {
        for (int32_t a = 0; a < nv; ++a) local[t.verts[a]] = a;
        min_degree_leaf(g, t.verts, local, ord);
        for (int32_t v : t.verts) local[v] = -1;
      } else {
        ord = t.verts;
      }
      for (int32_t a = 0; a < nv; ++a) perm[t.lo + a] = ord[a];
      return;
    }
    // local weighted graph of the subgraph (large ones: degrees, then rows, over vertex ranges on
    // several threads; the result is the same)
    for (int32_t a = 0; a < nv; ++a) local[t.verts[a]] = a;
    WGraph wg;
    wg.n = nv;
    wg.ptr.assign(static_cast<size_t>(nv) + 1, 0);
    {
      auto inside = [&](int32_t u) { return label[u].load(std::memory_order_relaxed) == t.id; };
      const int nt = worker_count(nv, 32768);
      parallel_ranges(nv, nt, [&](int, int32_t a0, int32_t a1) {
        for (int32_t a = a0; a < a1; ++a) {
          const int32_t v = t.verts[a];
          int64_t d = 0;
          for (int64_t p = g.ptr[v]; p < g.ptr[v + 1]; ++p) d += inside(g.adj[p]);
          wg.ptr[a + 1] = d;
        }
      });
      for (int32_t a = 0; a < nv; ++a) wg.ptr[a + 1] += wg.ptr[a];
      wg.adj.resize(static_cast<size_t>(wg.ptr[nv]));
      parallel_ranges(nv, nt, [&](int, int32_t a0, int32_t a1) {
        for (int32_t a = a0; a < a1; ++a) {
          const int32_t v = t.verts[a];
          int64_t q = wg.ptr[a];
          for (int64_t p = g.ptr[v]; p < g.ptr[v + 1]; ++p) {
            const int32_t u = g.adj[p];
            if (inside(u)) wg.adj[q++] = local[u];
          }
        }
      });
    }
    wg.ew.assign(wg.adj.size(), 1);
    wg.vw.assign(nv, 1);
    wg.total_vw = nv;
    for (int32_t v : t.verts) local[v] = -1;
    // connected components: dissect each separately (no separator needed between them)
    std::vector<int32_t> comp(nv, -1);
    int32_t ncomp = 0;
    {
      std::vector<int32_t> q;
      for (int32_t s0 = 0; s0 < nv; ++s0) {
        if (comp[s0] >= 0) continue;
        q.assign(1, s0);
        comp[s0] = ncomp;
        for (size_t h = 0; h < q.size(); ++h)
          for (int64_t p = wg.ptr[q[h]]; p < wg.ptr[q[h] + 1]; ++p)
            if (comp[wg.adj[p]] < 0) {
              comp[wg.adj[p]] = ncomp;
              q.push_back(wg.adj[p]);
            }
        ++ncomp;
      }
    }
    if (ncomp > 1) {
      std::vector<std::vector<int32_t>> parts(ncomp);
      for (int32_t a = 0; a < nv; ++a) parts[comp[a]].push_back(t.verts[a]);
      int32_t lo = t.lo;
      for (int32_t k = 0; k < ncomp; ++k) {
        Task c;
        c.lo = lo;
        c.hi = lo + static_cast<int32_t>(parts[k].size());
        lo = c.hi;
        c.verts = std::move(parts[k]);
        c.id = next_id.fetch_add(1);
        c.seed = child_seed(t.seed, static_cast<uint64_t>(k));
        c.ranged = t.ranged;
        out.push_back(std::move(c));
      }
      return;
    }
    Rng rng(t.seed);
    std::vector<int8_t> part;
    multilevel_bisect(wg, part, rng);
    // vertex separator: boundary vertices of the side with fewer of them
    std::vector<int32_t> bnd[2];
    for (int32_t a = 0; a < nv; ++a) {
      for (int64_t p = wg.ptr[a]; p < wg.ptr[a + 1]; ++p)
        if (part[wg.adj[p]] != part[a]) {
          bnd[part[a]].push_back(a);
          break;
        }
    }
    const int side = bnd[0].size() <= bnd[1].size() ? 0 : 1;
    std::vector<int8_t> where(nv);  // 0, 1 or 2 (separator)
    for (int32_t a = 0; a < nv; ++a) where[a] = part[a];
    for (int32_t a : bnd[side]) where[a] = 2;
    // minimal separator: a separator vertex with no neighbour on the other side joins its side
    for (int32_t a : bnd[side]) {
      bool touches_other = false;
      for (int64_t p = wg.ptr[a]; p < wg.ptr[a + 1] && !touches_other; ++p)
        touches_other = where[wg.adj[p]] == 1 - side;
      if (!touches_other) where[a] = static_cast<int8_t>(side);
    }
    // and those touching only the other side (never the kept side) may move there
    for (int32_t a : bnd[side]) {
      if (where[a] != 2) continue;
      bool touches_side = false;
      for (int64_t p = wg.ptr[a]; p < wg.ptr[a + 1] && !touches_side; ++p)
        touches_side = where[wg.adj[p]] == side;
      if (!touches_side) where[a] = static_cast<int8_t>(1 - side);
    }
    std::vector<int32_t> p0, p1, sep;
    for (int32_t a = 0; a < nv; ++a) {
      const int32_t v = t.verts[a];
      (where[a] == 0 ? p0 : where[a] == 1 ? p1 : sep).push_back(v);
    }
    if (p0.empty() || p1.empty()) {  // degenerate: order as a leaf
      for (int32_t a = 0; a < nv; ++a) perm[t.lo + a] = t.verts[a];
      return;
    }
    const int32_t ns = static_cast<int32_t>(sep.size());
    for (int32_t a = 0; a < ns; ++a) perm[t.hi - ns + a] = sep[a];
    for (int32_t v : sep) label[v].store(-1, std::memory_order_relaxed);
    Task c0, c1;
    c0.lo = t.lo;
    c0.hi = t.lo + static_cast<int32_t>(p0.size());
    c1.lo = c0.hi;
    c1.hi = t.hi - ns;
    c0.verts = std::move(p0);
    c1.verts = std::move(p1);
    c0.id = next_id.fetch_add(1);
    c1.id = next_id.fetch_add(1);
    c0.seed = child_seed(t.seed, 0);
    c1.seed = child_seed(t.seed, 1);
    c0.ranged = c1.ranged = t.ranged;
    out.push_back(std::move(c0));
    out.push_back(std::move(c1));
  };
  auto worker = [&]() {
    std::vector<int32_t> local(n, -1);
    for (;;) {
      Task t;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return !stack.empty() || busy == 0; });
        if (stack.empty()) return;
        t = std::move(stack.back());
        stack.pop_back();
        ++busy;
      }
      std::vector<Task> out;
      try {
        if (!failure) process(t, local, out);
      } catch (...) {
        std::lock_guard<std::mutex> lk(mu);
        if (!failure) failure = std::current_exception();
        out.clear();
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        for (Task& c : out) stack.push_back(std::move(c));
        --busy;
      }
      cv.notify_all();
    }
  };
  int nthreads = opt.threads > 0 ? opt.threads : std::min(16, (int)std::max(1u, std::thread::hardware_concurrency()));
  if (n < 4096) nthreads = 1;
  std::vector<std::thread> pool;
  for (int i = 1; i < nthreads; ++i) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  if (failure) std::rethrow_exception(failure);
  for (int32_t i = 0; i < n; ++i)
    if (perm[i] < 0) throw std::runtime_error("nested_dissection: incomplete ordering");
  return perm;
}

}  // namespace hsg
