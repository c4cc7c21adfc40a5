/* Helpers for the C oracle.  Test infrastructure only (see ora.h). */
#include "ora_util.h"
#include <pthread.h>
#include <unistd.h>

static int g_threads = 0;

void ora_set_threads(int t) { g_threads = t; }

int ora_threads(void) {
  if (g_threads > 0) return g_threads;
  long n = sysconf(_SC_NPROCESSORS_ONLN);
  return n > 0 ? (int)n : 1;
}

typedef struct { const double *key; size_t stride, off; } dkey_ctx;

static void merge_d(int64_t *a, int64_t *tmp, size_t n, const dkey_ctx *c) {
  if (n < 2) return;
  if (n <= 16) { /* stable insertion sort */
    for (size_t i = 1; i < n; i++) {
      int64_t x = a[i];
      double kx = c->key[(size_t)x * c->stride + c->off];
      size_t j = i;
      while (j > 0 && c->key[(size_t)a[j - 1] * c->stride + c->off] > kx) { a[j] = a[j - 1]; j--; }
      a[j] = x;
    }
    return;
  }
  size_t h = n / 2;
  merge_d(a, tmp, h, c);
  merge_d(a + h, tmp, n - h, c);
  size_t i = 0, j = h, k = 0;
  while (i < h && j < n) {
    double ki = c->key[(size_t)a[i] * c->stride + c->off], kj = c->key[(size_t)a[j] * c->stride + c->off];
    if (kj < ki) tmp[k++] = a[j++]; else tmp[k++] = a[i++];
  }
  while (i < h) tmp[k++] = a[i++];
  while (j < n) tmp[k++] = a[j++];
  memcpy(a, tmp, n * sizeof(int64_t));
}

void ora_sort_by_dkey(int64_t *items, size_t n, const double *key, size_t key_stride, size_t key_off) {
  int64_t *tmp = (int64_t *)ora_xmalloc(n * sizeof(int64_t));
  dkey_ctx c = {key, key_stride, key_off};
  merge_d(items, tmp, n, &c);
  free(tmp);
}

static void merge_i(int64_t *a, int64_t *tmp, size_t n, const int64_t *key) {
  if (n < 2) return;
  if (n <= 16) {
    for (size_t i = 1; i < n; i++) {
      int64_t x = a[i];
      int64_t kx = key ? key[x] : x;
      size_t j = i;
      while (j > 0 && (key ? key[a[j - 1]] : a[j - 1]) > kx) { a[j] = a[j - 1]; j--; }
      a[j] = x;
    }
    return;
  }
  size_t h = n / 2;
  merge_i(a, tmp, h, key);
  merge_i(a + h, tmp, n - h, key);
  size_t i = 0, j = h, k = 0;
  while (i < h && j < n) {
    int64_t ki = key ? key[a[i]] : a[i], kj = key ? key[a[j]] : a[j];
    if (kj < ki) tmp[k++] = a[j++]; else tmp[k++] = a[i++];
  }
  while (i < h) tmp[k++] = a[i++];
  while (j < n) tmp[k++] = a[j++];
  memcpy(a, tmp, n * sizeof(int64_t));
}

void ora_sort_by_ikey(int64_t *items, size_t n, const int64_t *key) {
  int64_t *tmp = (int64_t *)ora_xmalloc(n * sizeof(int64_t));
  merge_i(items, tmp, n, key);
  free(tmp);
}

void ora_sort_i64(int64_t *a, size_t n) { ora_sort_by_ikey(a, n, NULL); }

size_t ora_dedup_i64(int64_t *a, size_t n) {
  if (n == 0) return 0;
  size_t w = 1;
  for (size_t i = 1; i < n; i++)
    if (a[i] != a[w - 1]) a[w++] = a[i];
  return w;
}

typedef struct { ora_task_fn f; void *arg; } thr_arg;
static void *thr_main(void *p) {
  thr_arg *t = (thr_arg *)p;
  t->f(t->arg);
  return NULL;
}

void ora_run_parallel(ora_task_fn f, void **args, int n) {
  if (n <= 1) {
    if (n == 1) f(args[0]);
    return;
  }
  pthread_t *th = (pthread_t *)ora_xmalloc(sizeof(pthread_t) * (size_t)n);
  thr_arg *ta = (thr_arg *)ora_xmalloc(sizeof(thr_arg) * (size_t)n);
  int *ok = (int *)ora_xcalloc((size_t)n, sizeof(int));
  for (int i = 1; i < n; i++) {
    ta[i].f = f;
    ta[i].arg = args[i];
    ok[i] = pthread_create(&th[i], NULL, thr_main, &ta[i]) == 0;
    if (!ok[i]) f(args[i]);
  }
  f(args[0]);
  for (int i = 1; i < n; i++)
    if (ok[i]) pthread_join(th[i], NULL);
  free(th);
  free(ta);
  free(ok);
}

/* src/par.rs:51-66: bounds of t contiguous blocks with cost(r) = 1 (cost_linear 0) or r+1 (1) */
size_t ora_cost_bounds(size_t nrows, int threads, int cost_linear, size_t *bounds) {
  size_t t = (size_t)threads < nrows ? (size_t)threads : nrows;
  if (t < 1) t = 1;
  size_t nb = 0;
  bounds[nb++] = 0;
  if (t == 1) {
    bounds[nb++] = nrows;
    return nb;
  }
  double total = 0.0;
  for (size_t r = 0; r < nrows; r++) total += cost_linear ? (double)(r + 1) : 1.0;
  double acc = 0.0;
  for (size_t r = 0; r < nrows; r++) {
    acc += cost_linear ? (double)(r + 1) : 1.0;
    if (acc >= total * (double)nb / (double)t && nb < t) bounds[nb++] = r + 1;
  }
  bounds[nb++] = nrows;
  /* dedup */
  size_t w = 1;
  for (size_t i = 1; i < nb; i++)
    if (bounds[i] != bounds[w - 1]) bounds[w++] = bounds[i];
  return w;
}
