/*
 * oracle_pool.c -- pthread worker pool for the CPU oracle (test
 * infrastructure; see gviz_oracle.h).
 *
 * Restates the phase structure of the reference's gvizThreadPoolForRange
 * (src/core/gvizThreadPool.c:248-304): the caller plus up to
 * min(threads, chunks-1) helpers pull `grain`-sized chunks of [begin, end)
 * from one atomic counter, and the call returns once every chunk is done.
 * A NULL pool runs the range serially, as the reference does.
 */
#include "gviz_oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <unistd.h>

struct oc_pool {
  int nthreads;
  pthread_t *threads;
  pthread_mutex_t mu;
  pthread_cond_t wake;
  pthread_cond_t done;
  uint64_t generation;
  int helpersWanted;
  int helpersActive;
  int shutdown;
  /* current job */
  oc_range_fn fn;
  void *ctx;
  int64_t begin, end, grain;
  _Atomic int64_t next;
};

static void run_chunks(struct oc_pool *p) {
  for (;;) {
    int64_t s = atomic_fetch_add(&p->next, p->grain);
    if (s >= p->end)
      return;
    int64_t e = s + p->grain;
    if (e > p->end)
      e = p->end;
    p->fn(p->ctx, s, e);
  }
}

static void *worker(void *arg) {
  struct oc_pool *p = arg;
  uint64_t seen = 0;
  pthread_mutex_lock(&p->mu);
  for (;;) {
    while (!p->shutdown && (p->generation == seen || p->helpersWanted == 0))
      pthread_cond_wait(&p->wake, &p->mu);
    if (p->shutdown)
      break;
    seen = p->generation;
    p->helpersWanted--;
    p->helpersActive++;
    pthread_mutex_unlock(&p->mu);
    run_chunks(p);
    pthread_mutex_lock(&p->mu);
    p->helpersActive--;
    if (p->helpersActive == 0)
      pthread_cond_broadcast(&p->done);
  }
  pthread_mutex_unlock(&p->mu);
  return NULL;
}

oc_pool *oc_pool_create(int threads) {
  if (threads <= 0) {
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    threads = n > 0 ? (int)n : 1;
  }
  struct oc_pool *p = calloc(1, sizeof(*p));
  if (!p)
    return NULL;
  p->nthreads = threads;
  pthread_mutex_init(&p->mu, NULL);
  pthread_cond_init(&p->wake, NULL);
  pthread_cond_init(&p->done, NULL);
  p->threads = calloc((size_t)threads, sizeof(pthread_t));
  for (int i = 0; i < threads; i++)
    pthread_create(&p->threads[i], NULL, worker, p);
  return p;
}

void oc_pool_destroy(oc_pool *p) {
  if (!p)
    return;
  pthread_mutex_lock(&p->mu);
  p->shutdown = 1;
  pthread_cond_broadcast(&p->wake);
  pthread_mutex_unlock(&p->mu);
  for (int i = 0; i < p->nthreads; i++)
    pthread_join(p->threads[i], NULL);
  pthread_mutex_destroy(&p->mu);
  pthread_cond_destroy(&p->wake);
  pthread_cond_destroy(&p->done);
  free(p->threads);
  free(p);
}

int oc_pool_threads(const oc_pool *p) { return p ? p->nthreads : 1; }

void oc_pool_for_range(oc_pool *p, int64_t begin, int64_t end, int64_t grain,
                       oc_range_fn fn, void *ctx) {
  if (end <= begin)
    return;
  if (grain < 1)
    grain = 1;
  int64_t chunks = (end - begin + grain - 1) / grain;
  if (!p || chunks <= 1) {
    fn(ctx, begin, end);
    return;
  }
  int64_t helpers = chunks - 1 < p->nthreads ? chunks - 1 : p->nthreads;
  pthread_mutex_lock(&p->mu);
  p->fn = fn;
  p->ctx = ctx;
  p->begin = begin;
  p->end = end;
  p->grain = grain;
  atomic_store(&p->next, begin);
  p->helpersWanted = (int)helpers;
  p->generation++;
  pthread_cond_broadcast(&p->wake);
  pthread_mutex_unlock(&p->mu);

  run_chunks(p);

  pthread_mutex_lock(&p->mu);
  /* helpers that never woke up are no longer wanted */
  p->helpersWanted = 0;
  while (p->helpersActive > 0)
    pthread_cond_wait(&p->done, &p->mu);
  pthread_mutex_unlock(&p->mu);
}
