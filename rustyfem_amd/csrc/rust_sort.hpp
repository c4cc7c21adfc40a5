// The unstable sort of Rust's standard library (core::slice::sort::unstable,
// "ipnsort", Rust 1.81 and later) for slices of 8-byte integers, restated so
// that sorting node lists by a coordinate leaves ties in the order
// src/sparse.rs:108 (`set.sort_unstable_by(...)`) leaves them.  Which nodes
// tie is decided by the coordinates; where they end up decides the labels
// inside a supernode of the nested-dissection ordering.
//
// Steps (for 8-byte Copy elements: sorting-network small sort, branchless
// cyclic Lomuto partition):
//   - len <= 20: stable insertion sort;
//   - a leading non-descending or strictly descending run over the whole
//     slice: kept as is, or reversed;
//   - otherwise quicksort, at most 2 floor(log2(len | 1)) imbalanced
//     partitions before heapsort takes over; slices of <= 32 go to the small
//     sort; the pivot is the median of three samples at 0, 4(len/8) and
//     7(len/8) (a recursive median of medians from 64 elements on); a pivot
//     that is not above the previous (ancestor) pivot splits off all
//     elements equal to it.
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <utility>

namespace rf {
namespace rsort {

template <typename Less>
void insertion_from(int64_t* v, size_t len, size_t sorted, Less& lt) {
  for (size_t i = sorted; i < len; ++i) {
    if (!lt(v[i], v[i - 1])) continue;
    const int64_t x = v[i];
    size_t j = i;
    for (;;) {
      v[j] = v[j - 1];
      --j;
      if (j == 0 || !lt(x, v[j - 1])) break;
    }
    v[j] = x;
  }
}

// comparator pairs of the 9- and 13-input sorting networks, applied in order
constexpr uint8_t kNet9[25][2] = {{0, 3}, {1, 7}, {2, 5}, {4, 8}, {0, 7}, {2, 4}, {3, 8}, {5, 6}, {0, 2},
                                  {1, 3}, {4, 5}, {7, 8}, {1, 4}, {3, 6}, {5, 7}, {0, 1}, {2, 4}, {3, 5},
                                  {6, 8}, {2, 3}, {4, 5}, {6, 7}, {1, 2}, {3, 4}, {5, 6}};
constexpr uint8_t kNet13[45][2] = {{0, 12}, {1, 10}, {2, 9},  {3, 7},   {5, 11}, {6, 8},   {1, 6},  {2, 3},  {4, 11},
                                   {7, 9},  {8, 10}, {0, 4},  {1, 2},   {3, 6},  {7, 8},   {9, 10}, {11, 12}, {4, 6},
                                   {5, 9},  {8, 11}, {10, 12}, {0, 5},  {3, 8},  {4, 7},   {6, 11}, {9, 10}, {0, 1},
                                   {2, 5},  {6, 9},  {7, 8},  {10, 11}, {1, 3},  {2, 4},   {5, 6},  {9, 10}, {1, 2},
                                   {3, 4},  {5, 7},  {6, 8},  {2, 3},   {4, 5},  {6, 7},   {8, 9},  {3, 4},  {5, 6}};

template <size_t N, typename Less>
void network(int64_t* v, const uint8_t (&net)[N][2], Less& lt) {
  for (size_t i = 0; i < N; ++i) {
    const int a = net[i][0], b = net[i][1];
    if (lt(v[b], v[a])) std::swap(v[a], v[b]);
  }
}

// merges the sorted halves [0, len/2) and [len/2, len) from both ends at once
template <typename Less>
void merge_both_ends(const int64_t* v, size_t len, int64_t* out, Less& lt) {
  const ptrdiff_t h = static_cast<ptrdiff_t>(len / 2);
  ptrdiff_t lo_l = 0, lo_r = h, lo_o = 0;
  ptrdiff_t hi_l = h - 1, hi_r = static_cast<ptrdiff_t>(len) - 1, hi_o = static_cast<ptrdiff_t>(len) - 1;
  for (ptrdiff_t s = 0; s < h; ++s) {
    const bool from_l = !lt(v[lo_r], v[lo_l]);  // ties: left first
    out[lo_o++] = from_l ? v[lo_l] : v[lo_r];
    lo_l += from_l;
    lo_r += !from_l;
    const bool back_l = lt(v[hi_r], v[hi_l]);  // ties: right last
    out[hi_o--] = back_l ? v[hi_l] : v[hi_r];
    hi_l -= back_l;
    hi_r -= !back_l;
  }
  if (len & 1) out[lo_o] = lo_l <= hi_l ? v[lo_l] : v[lo_r];
}

template <typename Less>
void small_sort(int64_t* v, size_t len, Less& lt) {
  if (len < 2) return;
  auto presort = [&](int64_t* r, size_t rl) {
    size_t done = 1;
    if (rl >= 13) {
      network(r, kNet13, lt);
      done = 13;
    } else if (rl >= 9) {
      network(r, kNet9, lt);
      done = 9;
    }
    insertion_from(r, rl, done, lt);
  };
  if (len < 18) {
    presort(v, len);
    return;
  }
  const size_t h = len / 2;
  presort(v, h);
  presort(v + h, len - h);
  int64_t tmp[32];
  merge_both_ends(v, len, tmp, lt);
  std::memcpy(v, tmp, len * sizeof(int64_t));
}

template <typename Less>
void heap_sort(int64_t* v, size_t len, Less& lt) {
  auto sift = [&](size_t node, size_t n) {
    for (size_t c = 2 * node + 1; c < n; c = 2 * node + 1) {
      if (c + 1 < n && lt(v[c], v[c + 1])) ++c;
      if (!lt(v[node], v[c])) break;
      std::swap(v[node], v[c]);
      node = c;
    }
  };
  for (size_t i = len + len / 2; i-- > 0;) {
    if (i >= len) {
      sift(i - len, len);
    } else {
      std::swap(v[0], v[i]);
      sift(0, i);
    }
  }
}

template <typename Less>
size_t med3(const int64_t* v, size_t a, size_t b, size_t c, Less& lt) {
  const bool ab = lt(v[a], v[b]), ac = lt(v[a], v[c]);
  if (ab != ac) return a;
  return (lt(v[b], v[c]) != ab) ? c : b;
}

template <typename Less>
size_t med3_rec(const int64_t* v, size_t a, size_t b, size_t c, size_t n, Less& lt) {
  if (n >= 8) {
    const size_t s = n / 8;
    a = med3_rec(v, a, a + 4 * s, a + 7 * s, s, lt);
    b = med3_rec(v, b, b + 4 * s, b + 7 * s, s, lt);
    c = med3_rec(v, c, c + 4 * s, c + 7 * s, s, lt);
  }
  return med3(v, a, b, c, lt);
}

// Moves the pivot v[p] to the front, then sends every element x of the rest
// with below(x, pivot) to the front part, in the cyclic Lomuto order: the
// rest's first element waits in a register and is placed last.  The pivot
// ends at the returned index.
template <typename Below>
size_t partition_at(int64_t* v, size_t len, size_t p, Below below) {
  std::swap(v[0], v[p]);
  const int64_t piv = v[0];
  int64_t* w = v + 1;
  const size_t m = len - 1;
  size_t cnt = 0;
  if (m > 0) {
    const int64_t held = w[0];
    size_t hole = 0;
    for (size_t i = 1; i < m; ++i) {
      const int64_t x = w[i];
      w[hole] = w[cnt];
      w[cnt] = x;
      hole = i;
      cnt += below(x, piv);
    }
    w[hole] = w[cnt];
    w[cnt] = held;
    cnt += below(held, piv);
  }
  std::swap(v[0], v[cnt]);
  return cnt;
}

template <typename Less>
void quick(int64_t* v, size_t len, bool has_anc, int64_t anc, uint32_t limit, Less& lt) {
  for (;;) {
    if (len <= 32) {
      small_sort(v, len, lt);
      return;
    }
    if (limit == 0) {
      heap_sort(v, len, lt);
      return;
    }
    --limit;
    const size_t s = len / 8;
    const size_t p = len < 64 ? med3(v, 0, 4 * s, 7 * s, lt) : med3_rec(v, 0, 4 * s, 7 * s, s, lt);
    if (has_anc && !lt(anc, v[p])) {
      const size_t le = partition_at(v, len, p, [&](int64_t x, int64_t q) { return !lt(q, x); });
      v += le + 1;
      len -= le + 1;
      has_anc = false;
      continue;
    }
    const size_t below = partition_at(v, len, p, [&](int64_t x, int64_t q) { return lt(x, q); });
    quick(v, below, has_anc, anc, limit, lt);
    anc = v[below];
    has_anc = true;
    v += below + 1;
    len -= below + 1;
  }
}

}  // namespace rsort

// slice::sort_unstable_by with the strict order lt(a, b) on int64_t items
template <typename Less>
void rust_sort_unstable(int64_t* v, size_t len, Less lt) {
  if (len < 2) return;
  if (len <= 20) {
    rsort::insertion_from(v, len, 1, lt);
    return;
  }
  const bool desc = lt(v[1], v[0]);
  size_t run = 2;
  while (run < len && (desc ? lt(v[run], v[run - 1]) : !lt(v[run], v[run - 1]))) ++run;
  if (run == len) {
    if (desc)
      for (size_t i = 0, j = len - 1; i < j; ++i, --j) std::swap(v[i], v[j]);
    return;
  }
  uint32_t lg = 0;
  for (size_t x = len | 1; x > 1; x >>= 1) ++lg;
  rsort::quick(v, len, false, 0, 2 * lg, lt);
}

}  // namespace rf
