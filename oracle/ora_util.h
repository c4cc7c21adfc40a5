/* Small helpers for the C oracle: growable arrays, stable sorts, threads.
 * Test infrastructure only (see ora.h). */
#ifndef ORA_UTIL_H
#define ORA_UTIL_H
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

typedef struct { int64_t *a; size_t n, cap; } vi;
typedef struct { double *a; size_t n, cap; } vd;

static inline void *ora_xmalloc(size_t bytes) {
  void *p = malloc(bytes ? bytes : 1);
  if (!p) { fprintf(stderr, "oracle: out of memory (%zu bytes)\n", bytes); abort(); }
  return p;
}
static inline void *ora_xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz ? sz : 1);
  if (!p) { fprintf(stderr, "oracle: out of memory (%zu x %zu bytes)\n", n, sz); abort(); }
  return p;
}
static inline void *ora_xrealloc(void *p, size_t bytes) {
  void *q = realloc(p, bytes ? bytes : 1);
  if (!q) { fprintf(stderr, "oracle: out of memory (%zu bytes)\n", bytes); abort(); }
  return q;
}
static inline void vi_push(vi *v, int64_t x) {
  if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 8; v->a = (int64_t *)ora_xrealloc(v->a, v->cap * sizeof(int64_t)); }
  v->a[v->n++] = x;
}
static inline void vd_push(vd *v, double x) {
  if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 8; v->a = (double *)ora_xrealloc(v->a, v->cap * sizeof(double)); }
  v->a[v->n++] = x;
}
static inline void vi_free(vi *v) { free(v->a); v->a = NULL; v->n = v->cap = 0; }
static inline void vd_free(vd *v) { free(v->a); v->a = NULL; v->n = v->cap = 0; }

/* stable merge sort of int64 items by a double key (key[item]); ties keep the current order */
void ora_sort_by_dkey(int64_t *items, size_t n, const double *key, size_t key_stride, size_t key_off);
/* stable merge sort of int64 items by an int64 key array */
void ora_sort_by_ikey(int64_t *items, size_t n, const int64_t *key);
/* sort int64 ascending */
void ora_sort_i64(int64_t *a, size_t n);
/* Rust's unstable sort (ipnsort) restated, ora_rsort.c: the tie order of
 * slice::sort_unstable_by on usize items */
void ora_rsort_by_dkey(int64_t *items, size_t n, const double *key, size_t key_stride, size_t key_off);
void ora_rsort_by_ikey(int64_t *items, size_t n, const int64_t *key);
size_t ora_dedup_i64(int64_t *a, size_t n);
/* binary search of x in sorted a[0..n); returns index or -1 */
static inline int64_t ora_bsearch(const int64_t *a, size_t n, int64_t x) {
  size_t lo = 0, hi = n;
  while (lo < hi) {
    size_t mid = (lo + hi) / 2;
    if (a[mid] < x) lo = mid + 1; else if (a[mid] > x) hi = mid; else return (int64_t)mid;
  }
  return -1;
}

/* run f(arg_i) for i in [0, n) on up to n threads (pthread), joining all */
typedef void (*ora_task_fn)(void *arg);
void ora_run_parallel(ora_task_fn f, void **args, int n);
int ora_threads(void);

/* src/par.rs:46-77: contiguous row blocks of roughly equal cost; bounds has capacity t+1; returns count+1 */
size_t ora_cost_bounds(size_t nrows, int threads, int cost_linear, size_t *bounds);
#endif
