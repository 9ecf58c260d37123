/*
 * pool_stress.c -- the detect pass's host thread pool (vid.stab_amd/csrc/host/
 * vs_pool.c) under many back-to-back runs of random sizes, including runs
 * smaller than the pool (where most workers sit the run out), runs separated
 * by pauses longer than the spin window (workers asleep), and several pools
 * at once.  Every run must execute each index exactly once and the caller
 * must see every item's write when the run returns.
 *
 * Usage: pool_stress <threads> <runs> <seed>; exit status 0 on success.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "vs_pool.h"

#define MAXN 64

typedef struct {
  int hits[MAXN];
  unsigned stamp[MAXN];
  unsigned run;
} Ctx;

static void item(void* p, int i) {
  Ctx* c = (Ctx*)p;
  __atomic_fetch_add(&c->hits[i], 1, __ATOMIC_RELAXED);
  c->stamp[i] = c->run;   /* plain write: the run's return must publish it */
}

static unsigned rnd(unsigned* s) {
  *s = *s * 1103515245u + 12345u;
  return (*s >> 8) & 0xFFFFFF;
}

int main(int argc, char** argv) {
  if (argc != 4) return 2;
  const int threads = atoi(argv[1]), runs = atoi(argv[2]);
  unsigned seed = (unsigned)atoi(argv[3]);
  VSPool* pools[2] = {vs_pool_create(threads), vs_pool_create(threads > 2 ? threads - 1 : 2)};
  if (!pools[0] || !pools[1]) return 3;
  Ctx c;
  memset(&c, 0, sizeof(c));
  for (int r = 0; r < runs; r++) {
    VSPool* p = pools[rnd(&seed) % 2];
    const unsigned kind = rnd(&seed) % 8;
    const int n = kind == 0 ? 1 + (int)(rnd(&seed) % MAXN) : 1 + (int)(rnd(&seed) % 6);
    c.run = (unsigned)r + 1;
    memset(c.hits, 0, sizeof(c.hits));
    vs_pool_run(p, n, item, &c);
    for (int i = 0; i < MAXN; i++) {
      const int want = i < n ? 1 : 0;
      if (c.hits[i] != want || (want && c.stamp[i] != c.run)) {
        fprintf(stderr, "run %d (n %d): item %d executed %d times, stamp %u\n", r, n, i, c.hits[i], c.stamp[i]);
        return 1;
      }
    }
    if (kind == 7) {   /* past the spin window: the workers go to sleep */
      struct timespec ts = {0, 300000};
      nanosleep(&ts, NULL);
    }
  }
  vs_pool_destroy(pools[0]);
  vs_pool_destroy(pools[1]);
  printf("pool ok: %d runs\n", runs);
  return 0;
}
