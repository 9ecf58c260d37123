/*
 * release_stress.c -- the bulk release of LocalMotions arrays
 * (vid.stab_amd/csrc/host/vs_release.c) from several caller threads at once:
 * arrays of random lengths (below and above the asynchronous threshold) of
 * vectors filled by vs_vector_append_dup, released with vsLocalMotionsArrayDel
 * and reused at once.  Counting vs_malloc / vs_zalloc / vs_realloc / vs_free
 * hooks check that after vsLocalMotionsReleaseWait every block handed out has
 * come back exactly once, and every released vector is empty on return.
 *
 * Usage: release_stress <caller threads> <rounds> <seed>; exit status 0 on
 * success.
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "motiondetect.h"
#include "vidstab_amd.h"

static long long g_alloc, g_free;

static void* count_malloc(size_t n) {
  __atomic_fetch_add(&g_alloc, 1, __ATOMIC_RELAXED);
  return malloc(n);
}
static void* count_zalloc(size_t n) {
  __atomic_fetch_add(&g_alloc, 1, __ATOMIC_RELAXED);
  return calloc(1, n);
}
static void* count_realloc(void* p, size_t n) {
  if (!p) __atomic_fetch_add(&g_alloc, 1, __ATOMIC_RELAXED);
  return realloc(p, n);
}
static void count_free(void* p) {
  if (p) __atomic_fetch_add(&g_free, 1, __ATOMIC_RELAXED);
  free(p);
}

typedef struct {
  int rounds;
  unsigned seed;
  int bad;
} Caller;

static unsigned next(unsigned* s) {
  *s = *s * 1103515245u + 12345u;
  return *s >> 8;
}

static void* caller(void* p) {
  Caller* c = (Caller*)p;
  LocalMotions v[64];
  for (int r = 0; r < c->rounds; r++) {
    const int n = 1 + (int)(next(&c->seed) % 64);
    for (int i = 0; i < n; i++) {
      vs_vector_init(&v[i], 0);
      const int m = (int)(next(&c->seed) % 40);
      LocalMotion lm;
      memset(&lm, 0, sizeof(lm));
      for (int k = 0; k < m; k++) {
        lm.match = k;
        vs_vector_append_dup(&v[i], &lm, sizeof(lm));
      }
    }
    vsLocalMotionsArrayDel(v, n);
    for (int i = 0; i < n; i++)
      if (v[i].data || v[i].nelems || v[i].buffersize) c->bad++;
    if (next(&c->seed) % 16 == 0) vsLocalMotionsReleaseWait();
  }
  return NULL;
}

int main(int argc, char** argv) {
  const int threads = argc > 1 ? atoi(argv[1]) : 4;
  const int rounds = argc > 2 ? atoi(argv[2]) : 1000;
  const unsigned seed = argc > 3 ? (unsigned)atoi(argv[3]) : 1;
  if (threads < 1 || threads > 32) return 2;
  vs_malloc = count_malloc;
  vs_zalloc = count_zalloc;
  vs_realloc = count_realloc;
  vs_free = count_free;
  pthread_t th[32];
  Caller cs[32];
  for (int t = 0; t < threads; t++) {
    cs[t].rounds = rounds;
    cs[t].seed = seed * 7919u + (unsigned)t;
    cs[t].bad = 0;
    pthread_create(&th[t], NULL, caller, &cs[t]);
  }
  int bad = 0;
  for (int t = 0; t < threads; t++) {
    pthread_join(th[t], NULL);
    bad += cs[t].bad;
  }
  vsLocalMotionsReleaseWait();
  const long long a = __atomic_load_n(&g_alloc, __ATOMIC_RELAXED), f = __atomic_load_n(&g_free, __ATOMIC_RELAXED);
  if (bad || a != f) {
    fprintf(stderr, "release_stress: %d vectors not emptied, %lld blocks allocated, %lld freed\n", bad, a, f);
    return 1;
  }
  printf("release_stress: %lld blocks allocated and freed\n", a);
  return 0;
}
