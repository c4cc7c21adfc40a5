/* Test infrastructure (the oracle): a restatement of the unstable sort of
 * Rust's standard library, which src/sparse.rs:108 calls to order the nodes
 * of a nested-dissection set by one coordinate:
 *
 *   set.sort_unstable_by(|&a, &b| coords[a][ax].partial_cmp(&coords[b][ax]).unwrap())
 *
 * The sort is not stable, so the order of nodes with equal coordinates -- and
 * with it the labels inside every supernode -- is whatever this algorithm
 * leaves.  The reference has no dependencies (Cargo.lock), so the algorithm is
 * the one of its toolchain's `core` library: ipnsort, the unstable sort of
 * core::slice::sort::unstable since Rust 1.81 (the reference needs >= 1.82:
 * src/sparse.rs:123 calls Option::is_none_or).  Its published steps, restated
 * for slices of usize (8-byte Copy + Freeze elements, which select the
 * sorting-network small sort and the branchless cyclic Lomuto partition):
 *
 *   sort:        len < 2: done; len <= 20: insertion sort (stable);
 *                otherwise ipnsort.
 *   ipnsort:     the leading run (non-descending, or strictly descending) is
 *                measured; a run over the whole slice is kept (reversed when
 *                descending); otherwise quicksort with a limit of
 *                2 * floor(log2(len | 1)) imbalanced partitions.
 *   quicksort:   len <= 32: small sort; limit 0: heapsort; pivot = median of
 *                three at 0, 4*(len/8), 7*(len/8) (recursive pseudo-median of
 *                nine-way samples from len 64 on); a pivot not above the
 *                ancestor pivot splits off the elements equal to it;
 *                otherwise partition, recurse into the left side, loop on the
 *                right with the pivot as the new ancestor.
 *   partition:   pivot swapped to the front, branchless Lomuto with a cyclic
 *                permutation over the rest (the first element held out in a
 *                temporary and processed last), pivot swapped into place.
 *   small sort:  len < 18: sort13/sort9 network on the prefix, insertion sort
 *                of the rest; otherwise the same on each half and a
 *                bidirectional merge.
 *
 * Pinned by tests/test_oracle_rsort.py: sortedness and permutation on inputs
 * full of ties, the stable insertion sort at len <= 20, presorted and strictly
 * descending inputs, the two networks as sorting networks (0-1 principle over
 * all 2^9 and 2^13 inputs), and the product's ordering equal to this one.
 */
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include "ora.h"
#include "ora_util.h"

typedef struct {
  const double *dkey; /* key of item v: dkey[v * stride + off] (NULL: use ikey) */
  size_t stride, off;
  const int64_t *ikey; /* key of item v: ikey[v] (NULL with dkey NULL: v itself) */
} rs_ctx;

static int rs_less(const rs_ctx *c, int64_t a, int64_t b) {
  if (c->dkey) return c->dkey[(size_t)a * c->stride + c->off] < c->dkey[(size_t)b * c->stride + c->off];
  if (c->ikey) return c->ikey[a] < c->ikey[b];
  return a < b;
}

/* insertion_sort_shift_left: v[0..offset) is sorted; insert the rest one by one */
static void rs_insertion(int64_t *v, size_t len, size_t offset, const rs_ctx *c) {
  for (size_t i = offset; i < len; i++) {
    if (!rs_less(c, v[i], v[i - 1])) continue;
    int64_t tmp = v[i];
    size_t j = i;
    do {
      v[j] = v[j - 1];
      j--;
    } while (j > 0 && rs_less(c, tmp, v[j - 1]));
    v[j] = tmp;
  }
}

static void rs_swap_if_less(int64_t *v, size_t a, size_t b, const rs_ctx *c) {
  if (rs_less(c, v[b], v[a])) {
    int64_t t = v[a];
    v[a] = v[b];
    v[b] = t;
  }
}

/* optimal 9- and 13-input networks (25 and 45 comparators), in their order */
static const unsigned char NET9[25][2] = {{0, 3}, {1, 7}, {2, 5}, {4, 8}, {0, 7}, {2, 4}, {3, 8}, {5, 6}, {0, 2},
                                          {1, 3}, {4, 5}, {7, 8}, {1, 4}, {3, 6}, {5, 7}, {0, 1}, {2, 4}, {3, 5},
                                          {6, 8}, {2, 3}, {4, 5}, {6, 7}, {1, 2}, {3, 4}, {5, 6}};
static const unsigned char NET13[45][2] = {
    {0, 12}, {1, 10}, {2, 9},  {3, 7},  {5, 11}, {6, 8},  {1, 6},  {2, 3},  {4, 11}, {7, 9},  {8, 10}, {0, 4},
    {1, 2},  {3, 6},  {7, 8},  {9, 10}, {11, 12}, {4, 6},  {5, 9},  {8, 11}, {10, 12}, {0, 5},  {3, 8},  {4, 7},
    {6, 11}, {9, 10}, {0, 1},  {2, 5},  {6, 9},  {7, 8},  {10, 11}, {1, 3},  {2, 4},  {5, 6},  {9, 10}, {1, 2},
    {3, 4},  {5, 7},  {6, 8},  {2, 3},  {4, 5},  {6, 7},  {8, 9},  {3, 4},  {5, 6}};

void ora_rsort_net9(int64_t *v, const int64_t *ikey) {
  rs_ctx c = {NULL, 0, 0, ikey};
  for (int i = 0; i < 25; i++) rs_swap_if_less(v, NET9[i][0], NET9[i][1], &c);
}

void ora_rsort_net13(int64_t *v, const int64_t *ikey) {
  rs_ctx c = {NULL, 0, 0, ikey};
  for (int i = 0; i < 45; i++) rs_swap_if_less(v, NET13[i][0], NET13[i][1], &c);
}

static void rs_net(int64_t *v, int which, const rs_ctx *c) {
  if (which == 13)
    for (int i = 0; i < 45; i++) rs_swap_if_less(v, NET13[i][0], NET13[i][1], c);
  else
    for (int i = 0; i < 25; i++) rs_swap_if_less(v, NET9[i][0], NET9[i][1], c);
}

/* bidirectional merge of the sorted halves v[0..len/2) and v[len/2..len) into dst */
static void rs_bidir_merge(const int64_t *v, size_t len, int64_t *dst, const rs_ctx *c) {
  const ptrdiff_t half = (ptrdiff_t)(len / 2);
  ptrdiff_t l = 0, r = half, d = 0;
  ptrdiff_t lr = half - 1, rr = (ptrdiff_t)len - 1, dr = (ptrdiff_t)len - 1;
  for (ptrdiff_t i = 0; i < half; i++) {
    /* front: take the left element unless the right one is smaller */
    const int take_l = !rs_less(c, v[r], v[l]);
    dst[d++] = take_l ? v[l] : v[r];
    r += !take_l;
    l += take_l;
    /* back: take the right element unless it is smaller than the left one */
    const int take_lr = rs_less(c, v[rr], v[lr]);
    dst[dr--] = take_lr ? v[lr] : v[rr];
    rr -= !take_lr;
    lr -= take_lr;
  }
  if (len % 2 != 0) {
    const int left_nonempty = l < lr + 1;
    dst[d] = left_nonempty ? v[l] : v[r];
  }
}

static void rs_small_sort(int64_t *v, size_t len, const rs_ctx *c) {
  if (len < 2) return;
  const size_t half = len / 2;
  const int no_merge = len < 18;
  int64_t *region = v;
  size_t rlen = no_merge ? len : half;
  for (;;) {
    size_t pre = 1;
    if (rlen >= 13) {
      rs_net(region, 13, c);
      pre = 13;
    } else if (rlen >= 9) {
      rs_net(region, 9, c);
      pre = 9;
    }
    rs_insertion(region, rlen, pre, c);
    if (no_merge) return;
    if (region != v) break;
    region = v + half;
    rlen = len - half;
  }
  int64_t scratch[32];
  rs_bidir_merge(v, len, scratch, c);
  memcpy(v, scratch, len * sizeof(int64_t));
}

static void rs_sift_down(int64_t *v, size_t len, size_t node, const rs_ctx *c) {
  for (;;) {
    size_t child = 2 * node + 1;
    if (child >= len) break;
    if (child + 1 < len) child += (size_t)rs_less(c, v[child], v[child + 1]);
    if (!rs_less(c, v[node], v[child])) break;
    int64_t t = v[node];
    v[node] = v[child];
    v[child] = t;
    node = child;
  }
}

static void rs_heapsort(int64_t *v, size_t len, const rs_ctx *c) {
  for (size_t i = len + len / 2; i-- > 0;) {
    size_t sift, hl;
    if (i >= len) {
      sift = i - len;
      hl = len;
    } else {
      int64_t t = v[0];
      v[0] = v[i];
      v[i] = t;
      sift = 0;
      hl = i;
    }
    if (sift < hl) rs_sift_down(v, hl, sift, c);
  }
}

static size_t rs_median3(const int64_t *v, size_t a, size_t b, size_t cc, const rs_ctx *c) {
  const int x = rs_less(c, v[a], v[b]);
  const int y = rs_less(c, v[a], v[cc]);
  if (x == y) {
    const int z = rs_less(c, v[b], v[cc]);
    return (z ^ x) ? cc : b;
  }
  return a;
}

static size_t rs_median3_rec(const int64_t *v, size_t a, size_t b, size_t cc, size_t n, const rs_ctx *c) {
  if (n * 8 >= 64) {
    const size_t n8 = n / 8;
    a = rs_median3_rec(v, a, a + n8 * 4, a + n8 * 7, n8, c);
    b = rs_median3_rec(v, b, b + n8 * 4, b + n8 * 7, n8, c);
    cc = rs_median3_rec(v, cc, cc + n8 * 4, cc + n8 * 7, n8, c);
  }
  return rs_median3(v, a, b, cc, c);
}

static size_t rs_choose_pivot(const int64_t *v, size_t len, const rs_ctx *c) {
  const size_t d8 = len / 8;
  const size_t a = 0, b = d8 * 4, cc = d8 * 7;
  if (len < 64) return rs_median3(v, a, b, cc, c);
  return rs_median3_rec(v, a, b, cc, d8, c);
}

/* branchless cyclic Lomuto over v[0..len) against the pivot item p;
 * returns the number of items x with lt(x, p) (mode 0: x < p, mode 1: x <= p) */
static size_t rs_lomuto_cyclic(int64_t *v, size_t len, int64_t p, int mode, const rs_ctx *c) {
  if (len == 0) return 0;
  int64_t held = v[0];
  size_t num_lt = 0;
  int64_t *gap = &v[0];
  for (size_t i = 1; i <= len; i++) {
    int64_t *rp = i < len ? &v[i] : &held;
    const int lt = mode == 0 ? rs_less(c, *rp, p) : !rs_less(c, p, *rp);
    int64_t *left = &v[num_lt];
    *gap = *left;
    *left = *rp;
    gap = rp;
    num_lt += (size_t)lt;
  }
  return num_lt;
}

static size_t rs_partition(int64_t *v, size_t len, size_t pivot, int mode, const rs_ctx *c) {
  int64_t t = v[0];
  v[0] = v[pivot];
  v[pivot] = t;
  const size_t num_lt = rs_lomuto_cyclic(v + 1, len - 1, v[0], mode, c);
  t = v[0];
  v[0] = v[num_lt];
  v[num_lt] = t;
  return num_lt;
}

static void rs_quicksort(int64_t *v, size_t len, const int64_t *ancestor, uint32_t limit, const rs_ctx *c) {
  int has_anc = ancestor != NULL;
  int64_t anc = has_anc ? *ancestor : 0;
  for (;;) {
    if (len <= 32) {
      rs_small_sort(v, len, c);
      return;
    }
    if (limit == 0) {
      rs_heapsort(v, len, c);
      return;
    }
    limit--;
    const size_t piv = rs_choose_pivot(v, len, c);
    if (has_anc && !rs_less(c, anc, v[piv])) {
      /* the pivot equals the ancestor pivot: split off everything <= it */
      const size_t num_le = rs_partition(v, len, piv, 1, c);
      v += num_le + 1;
      len -= num_le + 1;
      has_anc = 0;
      continue;
    }
    const size_t num_lt = rs_partition(v, len, piv, 0, c);
    const int64_t pv = v[num_lt];
    if (has_anc)
      rs_quicksort(v, num_lt, &anc, limit, c);
    else
      rs_quicksort(v, num_lt, NULL, limit, c);
    v += num_lt + 1;
    len -= num_lt + 1;
    anc = pv;
    has_anc = 1;
  }
}

static void rs_sort(int64_t *v, size_t len, const rs_ctx *c) {
  if (len < 2) return;
  if (len <= 20) {
    rs_insertion(v, len, 1, c);
    return;
  }
  /* the leading run */
  size_t run = 2;
  const int desc = rs_less(c, v[1], v[0]);
  if (desc)
    while (run < len && rs_less(c, v[run], v[run - 1])) run++;
  else
    while (run < len && !rs_less(c, v[run], v[run - 1])) run++;
  if (run == len) {
    if (desc)
      for (size_t i = 0, j = len - 1; i < j; i++, j--) {
        int64_t t = v[i];
        v[i] = v[j];
        v[j] = t;
      }
    return;
  }
  uint32_t lg = 0;
  for (size_t x = len | 1; x > 1; x >>= 1) lg++;
  rs_quicksort(v, len, NULL, 2 * lg, c);
}

void ora_rsort_by_dkey(int64_t *items, size_t n, const double *key, size_t key_stride, size_t key_off) {
  rs_ctx c = {key, key_stride, key_off, NULL};
  rs_sort(items, n, &c);
}

void ora_rsort_by_ikey(int64_t *items, size_t n, const int64_t *key) {
  rs_ctx c = {NULL, 0, 0, key};
  rs_sort(items, n, &c);
}
