/*
 * vs_pool.c -- fixed thread pool with a shared index counter (see vs_pool.h).
 *
 * A run of n items is published by bumping the generation in `ticket`
 * (which also carries how many workers take part); the first
 * min(n, threads) - 1 workers then take indices from `next` until it passes
 * n, and the run returns once each of them has checked out (`running`
 * reaches zero).  The other workers sit the run out, asleep: a few items
 * (one frame's staging copy) are not slowed by threads that have nothing to
 * do.  After a run, a worker that took part spins on `ticket` for a short
 * while before it sleeps on the condition variable, and the caller spins on
 * `running` the same way -- as an OpenMP runtime's threads wait actively
 * before they block -- so back-to-back runs (one vsMotionDetection call
 * after another) do not pay a sleep and a wake-up per thread.  A worker reads
 * the generation and its part in the run from one atomic word, so a worker
 * late for a run it sits out cannot mistake the next run's participant count
 * for its own.  Every sleeper re-checks its condition under the mutex that
 * the waker holds while it changes the state and signals, so no wake-up is
 * lost.
 */
#include "vs_pool.h"

#include <pthread.h>
#include <stdlib.h>
#include <time.h>

#if defined(__x86_64__) || defined(__i386__)
#include <immintrin.h>
#define VS_RELAX() _mm_pause()
#else
#define VS_RELAX() ((void)0)
#endif

/* how long a waiting thread spins before it sleeps */
#define VS_POOL_SPIN_NS 200000LL

struct VSPool {
  int nthreads;
  pthread_t* th;
  pthread_mutex_t mu;
  pthread_cond_t start, done;
  unsigned long long ticket;   /* generation << 16 | workers taking part (atomic; written under mu) */
  int quit;            /* atomic; written under mu */
  int running;         /* workers still inside the current run (atomic) */
  void (*fn)(void*, int);
  void* ctx;
  int n;
  int next;            /* next index to hand out (atomic) */
};

static long long mono_ns(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (long long)ts.tv_sec * 1000000000LL + ts.tv_nsec;
}

static void drain(VSPool* p) {
  for (;;) {
    const int i = __atomic_fetch_add(&p->next, 1, __ATOMIC_RELAXED);
    if (i >= p->n) return;
    p->fn(p->ctx, i);
  }
}

/* the ticket after `seen` (another generation), or any value once quit is
   set; spins first when `spin` */
static unsigned long long wait_run(VSPool* p, unsigned long long seen, int spin) {
  const long long until = mono_ns() + (spin ? VS_POOL_SPIN_NS : 0);
  for (int k = 0;; k++) {
    const unsigned long long t = __atomic_load_n(&p->ticket, __ATOMIC_ACQUIRE);
    if (t != seen || __atomic_load_n(&p->quit, __ATOMIC_ACQUIRE)) return t;
    if ((k & 63) == 63 && mono_ns() > until) break;
    VS_RELAX();
  }
  pthread_mutex_lock(&p->mu);
  while (!p->quit && __atomic_load_n(&p->ticket, __ATOMIC_ACQUIRE) == seen) pthread_cond_wait(&p->start, &p->mu);
  const unsigned long long t = __atomic_load_n(&p->ticket, __ATOMIC_ACQUIRE);
  pthread_mutex_unlock(&p->mu);
  return t;
}

typedef struct {
  VSPool* pool;
  int index;   /* 0 .. nthreads - 2 */
} Worker;

static void* worker(void* arg) {
  VSPool* p = ((Worker*)arg)->pool;
  const int index = ((Worker*)arg)->index;
  free(arg);
  unsigned long long seen = 0;
  int spin = 1;
  for (;;) {
    seen = wait_run(p, seen, spin);
    if (__atomic_load_n(&p->quit, __ATOMIC_ACQUIRE)) break;
    /* a worker not needed for this run sleeps until the next one: spinning
       beside the few that work (often on their cores' sibling threads)
       slows them down */
    spin = index < (int)(seen & 0xFFFF);
    if (!spin) continue;
    drain(p);
    if (__atomic_sub_fetch(&p->running, 1, __ATOMIC_ACQ_REL) == 0) {
      pthread_mutex_lock(&p->mu);
      pthread_cond_signal(&p->done);
      pthread_mutex_unlock(&p->mu);
    }
  }
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
      Worker* w = (Worker*)malloc(sizeof(Worker));
      if (!w) break;
      w->pool = p;
      w->index = k;
      if (pthread_create(&p->th[k], NULL, worker, w) != 0) {
        free(w);
        break;
      }
      p->nthreads++;
    }
  }
  return p;
}

int vs_pool_threads(const VSPool* p) { return p ? p->nthreads : 1; }

void vs_pool_destroy(VSPool* p) {
  if (!p) return;
  pthread_mutex_lock(&p->mu);
  __atomic_store_n(&p->quit, 1, __ATOMIC_RELEASE);
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
  /* every worker of the previous run has checked out, so none reads these
     until the generation below is published */
  const int active = (n < p->nthreads ? n : p->nthreads) - 1;
  p->fn = fn;
  p->ctx = ctx;
  p->n = n;
  __atomic_store_n(&p->next, 0, __ATOMIC_RELAXED);
  __atomic_store_n(&p->running, active, __ATOMIC_RELAXED);
  pthread_mutex_lock(&p->mu);
  const unsigned long long gen = (__atomic_load_n(&p->ticket, __ATOMIC_RELAXED) >> 16) + 1;
  __atomic_store_n(&p->ticket, (gen << 16) | (unsigned long long)active, __ATOMIC_RELEASE);
  pthread_cond_broadcast(&p->start);
  pthread_mutex_unlock(&p->mu);
  drain(p);
  const long long until = mono_ns() + VS_POOL_SPIN_NS;
  for (int k = 0; __atomic_load_n(&p->running, __ATOMIC_ACQUIRE) > 0; k++) {
    if ((k & 63) == 63 && mono_ns() > until) {
      pthread_mutex_lock(&p->mu);
      while (__atomic_load_n(&p->running, __ATOMIC_ACQUIRE) > 0) pthread_cond_wait(&p->done, &p->mu);
      pthread_mutex_unlock(&p->mu);
      break;
    }
    VS_RELAX();
  }
}
