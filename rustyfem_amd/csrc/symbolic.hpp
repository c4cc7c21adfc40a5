// Host-side symbolic phase of the MI355X solver: nested-dissection ordering,
// elimination tree with supernodes (fronts), front row patterns and the
// lower-triangle pattern.  Integer work, restating src/sparse.rs:78-275.
#pragma once
#include <cstdint>
#include <vector>

namespace rf {

struct Front {
  std::vector<int64_t> nodes;     // mesh nodes in labelling order
  std::vector<int64_t> children;  // at most two, in dissection order
  int64_t start = 0, k = 0;       // first equation and pivot count
  std::vector<int64_t> rows;      // later equations of the front, sorted
  std::vector<int64_t> row_nodes; // nodes of `rows`, sorted by first equation
  double work = 0.0;              // flops of the subtree (src/sparse.rs:226-228)
  int64_t parent = -1;
};

struct Symbolic {
  std::vector<Front> fronts;  // postorder (children before parents)
  std::vector<int64_t> roots;
  std::vector<int64_t> first_dof;
  int64_t n = 0;

  double work() const;
  int64_t factor_size() const;
  // lower-triangle CSC pattern in internal numbering (src/sparse.rs:256-275)
  void pattern(const std::vector<int64_t>& adj_ptr, const std::vector<int64_t>& adj_idx,
               const std::vector<int64_t>& ndof, std::vector<int64_t>& colptr,
               std::vector<int64_t>& rowidx) const;
};

// coords: nn x 3; adjacency as CSR over nodes; ndof per node.
// Restates src/sparse.rs:167-232 (analyse_with_leaf).
Symbolic analyse(const double* coords, int64_t nn, const std::vector<int64_t>& adj_ptr,
                 const std::vector<int64_t>& adj_idx, const std::vector<int64_t>& ndof, int64_t leaf_max);

}  // namespace rf
