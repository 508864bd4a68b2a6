// Host-side symbolic analysis of a symmetric sparsity pattern for the multifrontal LDL^T:
// nested-dissection ordering, elimination tree, column counts, relaxed supernodes, front row
// structures, extend-add maps, the assembly map of the original entries and the level schedule.
//
// The reference keeps this step inside MUMPS (JOB = 1, src/factor/mumps_factorization.cpp:188-283)
// with hsolve's own nested dissection handed over as PERM_IN (include/hsolve/factor.hpp:197-246,
// 598-633). Here everything up to the numeric values is computed on the host; the numeric
// factorization runs on the GPU.
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace hsg {

struct OrderingOptions {
  int leaf_size = 64;        // subgraphs at most this large are not dissected further
  int min_degree_leaves = 1; // 1: minimum-degree order inside leaves, 0: natural order
  int threads = 0;           // worker threads of the dissection (0: min(16, hardware threads))
};

struct AmalgamationOptions {
  // Relaxed supernodes (merge a child into its parent when the merged front is small or the
  // explicit zeros it adds are a small fraction): up to ncol[i] merged columns with at most
  // zero_frac[i] explicit zeros.
  int ncol[3] = {8, 32, 96};
  double zero_frac[3] = {0.8, 0.25, 0.08};
};

/// Symmetric pattern in compressed form: for every vertex the sorted list of its neighbours
/// (both triangles, no diagonal).
struct Graph {
  int32_t n = 0;
  std::vector<int64_t> ptr;
  std::vector<int32_t> adj;
};

/// Builds the adjacency graph of a square pattern given as CSR (rowptr, colind); the pattern is
/// symmetrised and the diagonal dropped.
Graph graph_from_csr(int32_t n, const int64_t* rowptr, const int32_t* colind);

/// Nested-dissection ordering: perm[new] = old. With `ranges`, also the ranges [lo, hi) of new
/// numbers held by the subgraphs of the dissection that first have at most `range_max` vertices
/// (disjoint; their separators above lie in no range). Each range is closed below: a vertex
/// numbered in [lo, hi) has all of its lower-numbered neighbours inside [lo, hi).
std::vector<int32_t> nested_dissection(const Graph& g, const OrderingOptions& opt,
                                       std::vector<std::pair<int32_t, int32_t>>* ranges = nullptr,
                                       int32_t range_max = 0);

struct Symbolic {
  int32_t n = 0;
  // order[new] = old index of the input; iorder[old] = new. Supernodal (postordered) numbering.
  std::vector<int32_t> order, iorder;
  int32_t nsn = 0;
  std::vector<int32_t> sn_first, sn_ncol, sn_m, sn_parent;  // per supernode
  std::vector<int64_t> sn_rowoff;   // offset of the front's row list in rows[] (size nsn + 1)
  std::vector<int32_t> rows;        // front rows, new numbering, ascending; first ncol = own columns
  std::vector<int32_t> relmap;      // parallel to rows: for CB rows, the local index in the parent front
  std::vector<int32_t> child_ptr, child_list;  // children of each supernode (ascending)
  std::vector<int64_t> panel_off;   // m x ncol panel of L per supernode (size nsn + 1)
  // cb x cb contribution block per supernode: offsets into one array of cb_total entries shared by
  // blocks whose lifetimes (own level to parent level) do not overlap (plan_contribution_blocks);
  // cb_off[nsn] = cb_total
  std::vector<int64_t> cb_off;
  int64_t cb_total = 0;             // entries of the contribution-block array (the plan's peak)
  int64_t cb_sum = 0;               // sum of all blocks' entries (what one block per front would take)
  std::vector<int32_t> level;       // height in the assembly tree (leaves 0)
  int32_t nlevels = 0;
  std::vector<int32_t> level_ptr, level_sn;  // supernodes grouped by level, ascending
  std::vector<int32_t> col_sn;      // supernode of each (new) column
  // statistics
  double factor_entries = 0;        // lower-trapezoid entries of all panels
  double flops = 0;                 // sum over fronts of p^3/3 + p^2 c + p c^2
  int32_t max_front = 0, max_ncol = 0;
};

/// Full symbolic analysis of the pattern (CSR, any symmetric pattern incl. diagonal).
Symbolic analyze_pattern(int32_t n, const int64_t* rowptr, const int32_t* colind,
                         const OrderingOptions& oopt, const AmalgamationOptions& aopt,
                         const int32_t* given_order /* may be null: new -> old */);

/// Offsets of the contribution blocks (cb_off, cb_total) from the level schedule: blocks live from
/// their front's level to their parent's level and share storage otherwise.
void plan_contribution_blocks(Symbolic& s);

/// Assembly destinations of the original entries: for every entry e of the CSR pattern, the panel
/// index (sn's panel_off + local_col * m + local_row) of K(i, j) when, in the new numbering, row
/// i lies on or below column j; -1 for the mirrored copy above the diagonal.
std::vector<int64_t> assembly_map(const Symbolic& s, int32_t n, const int64_t* rowptr,
                                  const int32_t* colind);

}  // namespace hsg
