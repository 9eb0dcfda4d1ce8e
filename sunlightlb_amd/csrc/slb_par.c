/*
 * slb_par.c -- the O(N) host passes over the node map (object lookups,
 * body-force fills, counts) split over host threads.
 *
 * A pass over [0, n) is cut into one contiguous range per thread; each
 * thread writes only its own range and its own slot of any partial result,
 * and the caller combines the slots in range order, so every result is the
 * one the sequential loop gives (integer counts and sums of integer-valued
 * doubles below 2^53 are exact in any order).  Threads are started per pass:
 * a pass worth splitting is tens of milliseconds, thread start-up is tens of
 * microseconds.  SLB_HOST_THREADS sets the count (default: the cores this
 * process may run on, at most 16).
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <sched.h>
#include <stdlib.h>
#include <unistd.h>
#include "slb_internal.h"

int slb_host_threads(void) {
  static int nt = 0;
  if (nt > 0) return nt;
  int n = 0;
  const char *e = getenv("SLB_HOST_THREADS");
  if (e && atoi(e) > 0) n = atoi(e);
  if (n == 0) {
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof set, &set) == 0) n = CPU_COUNT(&set);
    if (n <= 0) n = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (n > 16) n = 16;
  }
  if (n < 1) n = 1;
  if (n > SLB_PAR_MAX) n = SLB_PAR_MAX;
  nt = n;
  return nt;
}

typedef struct {
  slb_range_fn fn;
  void *ctx;
  long lo, hi;
  int tid;
} par_job;

static void *par_run(void *arg) {
  par_job *j = (par_job *)arg;
  j->fn(j->ctx, j->lo, j->hi, j->tid);
  return NULL;
}

int slb_par_for(long n, slb_range_fn fn, void *ctx) {
  int nt = slb_host_threads();
  /* small passes are not worth a thread */
  if (n < (1L << 18) || nt == 1) {
    fn(ctx, 0, n, 0);
    return 1;
  }
  if ((long)nt > n) nt = (int)n;
  par_job job[SLB_PAR_MAX];
  pthread_t th[SLB_PAR_MAX];
  int started[SLB_PAR_MAX] = {0};
  const long chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; t++) {
    job[t].fn = fn; job[t].ctx = ctx; job[t].tid = t;
    job[t].lo = chunk * t < n ? chunk * t : n;
    job[t].hi = chunk * (t + 1) < n ? chunk * (t + 1) : n;
  }
  /* ranges 1 .. nt-1 on new threads, range 0 on this one; a thread that
   * cannot be started runs its range here */
  for (int t = 1; t < nt; t++) started[t] = pthread_create(&th[t], NULL, par_run, &job[t]) == 0;
  par_run(&job[0]);
  for (int t = 1; t < nt; t++) {
    if (started[t]) pthread_join(th[t], NULL);
    else par_run(&job[t]);
  }
  return nt;
}

/* ------------------------------------------------------------------ */
/* Host mirrors of per-node arrays.  Large ones are anonymous mappings
 * aligned to 2 MiB with transparent huge pages requested: their first touch
 * (an O(N) host pass after newlb) then faults 2 MiB at a time instead of
 * 4 KiB, which also lets the threads of a pass fault in parallel.  The
 * memory is zero-filled, as calloc's.  A registry maps each such array to
 * its mapping for slb_big_free; small arrays come from calloc. */
#include <string.h>
#include <sys/mman.h>

#define BIG_MIN (2L << 20)
#define BIG_ALIGN (2UL << 20)

typedef struct big_rec {
  void *p, *raw;
  size_t len;
  struct big_rec *next;
} big_rec;

static big_rec *big_head = NULL;
static pthread_mutex_t big_lock = PTHREAD_MUTEX_INITIALIZER;

void *slb_big_calloc(long n, size_t size) {
  const size_t bytes = (size_t)(n > 0 ? n : 1) * size;
  if ((long)bytes < BIG_MIN) return calloc((size_t)(n > 0 ? n : 1), size);
  const size_t len = bytes + BIG_ALIGN;
  void *raw = mmap(NULL, len, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
  if (raw == MAP_FAILED) return NULL;
  void *p = (void *)(((uintptr_t)raw + BIG_ALIGN - 1) & ~(uintptr_t)(BIG_ALIGN - 1));
  madvise(p, bytes, MADV_HUGEPAGE);   /* a hint: without it the mapping still works */
  big_rec *r = (big_rec *)malloc(sizeof *r);
  if (r == NULL) { munmap(raw, len); return NULL; }
  r->p = p; r->raw = raw; r->len = len;
  pthread_mutex_lock(&big_lock);
  r->next = big_head;
  big_head = r;
  pthread_mutex_unlock(&big_lock);
  return p;
}

void slb_big_free(void *p) {
  if (p == NULL) return;
  pthread_mutex_lock(&big_lock);
  big_rec **q = &big_head;
  while (*q != NULL && (*q)->p != p) q = &(*q)->next;
  big_rec *r = *q;
  if (r != NULL) *q = r->next;
  pthread_mutex_unlock(&big_lock);
  if (r == NULL) { free(p); return; }
  munmap(r->raw, r->len);
  free(r);
}
