// Nested-dissection symbolic analysis (restates src/sparse.rs:78-275).
//
// Node sets are sorted by a coordinate with the restated unstable sort of
// Rust's standard library (rust_sort.hpp), as src/sparse.rs:108 sorts them:
// its order among equal coordinates decides the labels inside a supernode.
// The sorts by first equation (src/sparse.rs:224, 260, 263) have distinct
// keys; they use the same sort.
#include "symbolic.hpp"

#include "rust_sort.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <limits>

namespace rf {

namespace {

class Dissector {
 public:
  Dissector(const double* xyz, const std::vector<int64_t>& ap, const std::vector<int64_t>& ai, int64_t leaf,
            std::vector<Front>& out)
      : x_(xyz), ap_(ap), ai_(ai), leaf_(leaf), tree_(out), side_(ap.size() - 1, 2) {}

  // src/sparse.rs:78-157; returns the tree index, -1 for an empty set
  int64_t run(std::vector<int64_t> set) {
    if (set.empty()) return -1;
    if (static_cast<int64_t>(set.size()) <= leaf_) return push(std::move(set), {});
    const size_t n = set.size();
    std::array<double, 3> lo{}, hi{};
    lo.fill(std::numeric_limits<double>::infinity());
    hi.fill(-std::numeric_limits<double>::infinity());
    for (int64_t v : set)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::fmin(lo[a], x_[3 * v + a]);
        hi[a] = std::fmax(hi[a], x_[3 * v + a]);
      }
    std::array<int, 3> axes{0, 1, 2};
    std::stable_sort(axes.begin(), axes.end(), [&](int a, int b) { return (hi[b] - lo[b]) < (hi[a] - lo[a]); });
    bool have = false, best_left = true;
    int best_ax = 0;
    double best_score = 0.0, best_t = 0.0;
    for (int ax : axes) {
      if (!(hi[ax] > lo[ax])) continue;
      rust_sort_unstable(set.data(), set.size(), [&](int64_t a, int64_t b) { return x_[3 * a + ax] < x_[3 * b + ax]; });
      for (size_t q : {n / 2, 2 * n / 5, 3 * n / 5}) {
        const double t = x_[3 * set[q] + ax];
        const size_t nl = std::partition_point(set.begin(), set.end(), [&](int64_t v) { return x_[3 * v + ax] < t; }) -
                          set.begin();
        if (nl == 0 || nl == n) continue;
        for (size_t i = 0; i < n; ++i) side_[set[i]] = i >= nl ? 1 : 0;
        size_t sep_l = 0, sep_r = 0;
        for (size_t i = 0; i < nl; ++i) sep_l += touches(set[i], 1);
        for (size_t i = nl; i < n; ++i) sep_r += touches(set[i], 0);
        const bool in_left = !(sep_r < sep_l);
        const size_t s = in_left ? sep_l : sep_r;
        const double score = static_cast<double>(s) / (static_cast<double>(std::min(nl, n - nl)) / static_cast<double>(n));
        if (!have || score < best_score) {
          have = true;
          best_score = score;
          best_ax = ax;
          best_t = t;
          best_left = in_left;
        }
      }
    }
    std::vector<int64_t> left, right;
    left.reserve(n);
    right.reserve(n);
    if (have) {
      for (int64_t v : set) (x_[3 * v + best_ax] < best_t ? left : right).push_back(v);
    } else {  // all nodes at one point: split on index
      left.assign(set.begin(), set.begin() + n / 2);
      right.assign(set.begin() + n / 2, set.end());
    }
    std::vector<int64_t>().swap(set);
    const bool in_left = have ? best_left : true;
    for (int64_t v : left) side_[v] = 0;
    for (int64_t v : right) side_[v] = 1;
    std::vector<int64_t> sep, a, b;
    if (in_left) {
      for (int64_t v : left) (touches(v, 1) ? sep : a).push_back(v);
      b = right;
    } else {
      for (int64_t v : right) (touches(v, 0) ? sep : b).push_back(v);
      a = left;
    }
    for (int64_t v : left) side_[v] = 2;
    for (int64_t v : right) side_[v] = 2;
    std::vector<int64_t>().swap(left);
    std::vector<int64_t>().swap(right);
    std::vector<int64_t> ch;
    const int64_t ca = run(std::move(a));
    const int64_t cb = run(std::move(b));
    if (ca >= 0) ch.push_back(ca);
    if (cb >= 0) ch.push_back(cb);
    return push(std::move(sep), std::move(ch));
  }

 private:
  size_t touches(int64_t v, uint8_t s) const {
    for (int64_t p = ap_[v]; p < ap_[v + 1]; ++p)
      if (side_[ai_[p]] == s) return 1;
    return 0;
  }
  int64_t push(std::vector<int64_t> nodes, std::vector<int64_t> children) {
    Front f;
    f.nodes = std::move(nodes);
    f.children = std::move(children);
    tree_.push_back(std::move(f));
    return static_cast<int64_t>(tree_.size()) - 1;
  }

  const double* x_;
  const std::vector<int64_t>& ap_;
  const std::vector<int64_t>& ai_;
  int64_t leaf_;
  std::vector<Front>& tree_;
  std::vector<uint8_t> side_;
};

}  // namespace

Symbolic analyse(const double* coords, int64_t nn, const std::vector<int64_t>& adj_ptr,
                 const std::vector<int64_t>& adj_idx, const std::vector<int64_t>& ndof, int64_t leaf_max) {
  Symbolic s;
  leaf_max = std::max<int64_t>(leaf_max, 1);
  std::vector<int64_t> all(nn);
  for (int64_t i = 0; i < nn; ++i) all[i] = i;
  if (nn > 0) {
    Dissector d(coords, adj_ptr, adj_idx, leaf_max, s.fronts);
    s.roots.push_back(d.run(std::move(all)));
  }
  // Children are pushed before their parent, child 0's subtree first, so the
  // construction order is the postorder of src/sparse.rs:175-187.
  auto& T = s.fronts;
  s.first_dof.assign(nn, 0);
  int64_t next = 0;
  for (size_t t = 0; t < T.size(); ++t) {
    T[t].start = next;
    for (int64_t v : T[t].nodes) {
      s.first_dof[v] = next;
      next += ndof[v];
    }
    T[t].k = next - T[t].start;
    for (int64_t c : T[t].children) T[c].parent = static_cast<int64_t>(t);
  }
  s.n = next;
  // front pattern at node level (src/sparse.rs:201-230)
  std::vector<int64_t> stamp(nn, -1);
  for (size_t t = 0; t < T.size(); ++t) {
    const int64_t end = T[t].start + T[t].k;
    std::vector<int64_t> list;
    auto take = [&](int64_t w) {
      if (s.first_dof[w] >= end && stamp[w] != static_cast<int64_t>(t)) {
        stamp[w] = static_cast<int64_t>(t);
        list.push_back(w);
      }
    };
    for (int64_t v : T[t].nodes)
      for (int64_t p = adj_ptr[v]; p < adj_ptr[v + 1]; ++p) take(adj_idx[p]);
    for (int64_t c : T[t].children)
      for (int64_t w : T[c].row_nodes) take(w);
    rust_sort_unstable(list.data(), list.size(), [&](int64_t a, int64_t b) { return s.first_dof[a] < s.first_dof[b]; });
    for (int64_t w : list)
      for (int64_t d = 0; d < ndof[w]; ++d) T[t].rows.push_back(s.first_dof[w] + d);
    const double k = static_cast<double>(T[t].k), u = static_cast<double>(T[t].rows.size());
    const double own = k * k * k / 3.0 + u * k * k + u * u * k;
    double cs = 0.0;
    for (int64_t c : T[t].children) cs += T[c].work;
    T[t].work = own + cs;
    T[t].row_nodes = std::move(list);
  }
  return s;
}

double Symbolic::work() const {
  double w = 0.0;
  for (int64_t r : roots) w += fronts[r].work;
  return w;
}

int64_t Symbolic::factor_size() const {
  int64_t tot = 0;
  for (const auto& f : fronts) tot += f.k * std::max<int64_t>(f.k - 1, 0) / 2 + f.k * static_cast<int64_t>(f.rows.size());
  return tot;
}

void Symbolic::pattern(const std::vector<int64_t>& adj_ptr, const std::vector<int64_t>& adj_idx,
                       const std::vector<int64_t>& ndof, std::vector<int64_t>& colptr,
                       std::vector<int64_t>& rowidx) const {
  const int64_t nn = static_cast<int64_t>(first_dof.size());
  std::vector<int64_t> order(nn);
  for (int64_t i = 0; i < nn; ++i) order[i] = i;
  rust_sort_unstable(order.data(), order.size(), [&](int64_t a, int64_t b) { return first_dof[a] < first_dof[b]; });
  colptr.assign(n + 1, 0);
  rowidx.clear();
  std::vector<int64_t> nb;
  for (int64_t v : order) {
    nb.assign(adj_idx.begin() + adj_ptr[v], adj_idx.begin() + adj_ptr[v + 1]);
    nb.push_back(v);
    rust_sort_unstable(nb.data(), nb.size(), [&](int64_t a, int64_t b) { return first_dof[a] < first_dof[b]; });
    const int64_t f = first_dof[v];
    for (int64_t c = f; c < f + ndof[v]; ++c) {
      for (int64_t w : nb) {
        const int64_t fw = first_dof[w];
        for (int64_t r = std::max(fw, c); r < fw + ndof[w]; ++r) rowidx.push_back(r);
      }
      colptr[c + 1] = static_cast<int64_t>(rowidx.size());
    }
  }
}

}  // namespace rf
