/*
 * vs_pool.c -- fixed thread pool with a shared index counter (see vs_pool.h).
 */
#include "vs_pool.h"

#include <pthread.h>
#include <stdlib.h>

struct VSPool {
  int nthreads;
  pthread_t* th;
  pthread_mutex_t mu;
  pthread_cond_t start, done;
  unsigned gen;        /* bumped for every run */
  int quit;
  int running;         /* workers still inside the current run */
  void (*fn)(void*, int);
  void* ctx;
  int n;
  int next;            /* next index to hand out (atomic) */
};

static void drain(VSPool* p) {
  for (;;) {
    const int i = __atomic_fetch_add(&p->next, 1, __ATOMIC_RELAXED);
    if (i >= p->n) return;
    p->fn(p->ctx, i);
  }
}

static void* worker(void* arg) {
  VSPool* p = (VSPool*)arg;
  unsigned seen = 0;
  pthread_mutex_lock(&p->mu);
  for (;;) {
    while (!p->quit && p->gen == seen) pthread_cond_wait(&p->start, &p->mu);
    if (p->quit) break;
    seen = p->gen;
    pthread_mutex_unlock(&p->mu);
    drain(p);
    pthread_mutex_lock(&p->mu);
    if (--p->running == 0) pthread_cond_signal(&p->done);
  }
  pthread_mutex_unlock(&p->mu);
  return NULL;
}

VSPool* vs_pool_create(int nthreads) {
  if (nthreads < 1) nthreads = 1;
  VSPool* p = (VSPool*)calloc(1, sizeof(VSPool));
  if (!p) return NULL;
  p->nthreads = 1;
  pthread_mutex_init(&p->mu, NULL);
  pthread_cond_init(&p->start, NULL);
  pthread_cond_init(&p->done, NULL);
  if (nthreads > 1) {
    p->th = (pthread_t*)calloc((size_t)nthreads - 1, sizeof(pthread_t));
    if (!p->th) {
      vs_pool_destroy(p);
      return NULL;
    }
    for (int k = 0; k < nthreads - 1; k++) {
      if (pthread_create(&p->th[k], NULL, worker, p) != 0) break;
      p->nthreads++;
    }
  }
  return p;
}

int vs_pool_threads(const VSPool* p) { return p ? p->nthreads : 1; }

void vs_pool_destroy(VSPool* p) {
  if (!p) return;
  pthread_mutex_lock(&p->mu);
  p->quit = 1;
  pthread_cond_broadcast(&p->start);
  pthread_mutex_unlock(&p->mu);
  for (int k = 0; k < p->nthreads - 1; k++) pthread_join(p->th[k], NULL);
  free(p->th);
  pthread_cond_destroy(&p->start);
  pthread_cond_destroy(&p->done);
  pthread_mutex_destroy(&p->mu);
  free(p);
}

void vs_pool_run(VSPool* p, int n, void (*fn)(void*, int), void* ctx) {
  if (n <= 0) return;
  if (!p || p->nthreads == 1 || n == 1) {
    for (int i = 0; i < n; i++) fn(ctx, i);
    return;
  }
  pthread_mutex_lock(&p->mu);
  p->fn = fn;
  p->ctx = ctx;
  p->n = n;
  __atomic_store_n(&p->next, 0, __ATOMIC_RELAXED);
  p->running = p->nthreads - 1;
  p->gen++;
  pthread_cond_broadcast(&p->start);
  pthread_mutex_unlock(&p->mu);
  drain(p);
  pthread_mutex_lock(&p->mu);
  while (p->running > 0) pthread_cond_wait(&p->done, &p->mu);
  pthread_mutex_unlock(&p->mu);
}
