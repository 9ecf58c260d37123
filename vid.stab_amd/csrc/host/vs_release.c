/*
 * vs_release.c -- bulk release of LocalMotions arrays on a library thread.
 *
 * A LocalMotions vector holds one vs_malloc block per element, so releasing
 * a batch's results is one vs_free per LocalMotion: about 92,000 for the 300
 * frames of a 1080p clip, 1.7-2 ms on one host thread.  Spreading them over
 * threads does not help: the blocks come from the detector's worker threads'
 * allocator arenas and the freeing threads only contend for them.  So
 * vsLocalMotionsArrayDel empties the caller's vectors at once (they can be
 * reused on return) and hands their elements to one release thread, which
 * returns them through vs_free while the caller carries on -- typically with
 * its next detect call, whose time is mostly spent waiting for the GPU.
 *
 * At most kMaxPending releases are queued or running at a time: a caller
 * that outruns the release thread waits in vsLocalMotionsArrayDel.
 * vsLocalMotionsReleaseWait returns once every queued release is done; when
 * the library is unloaded (or the process exits) the queue is drained and
 * the release thread ends and is joined.  Arrays of fewer than
 * kAsyncMin vectors, or any when the thread cannot be started, are released
 * in the calling thread.  The hooks (vs_free) must stay the same while
 * releases are pending.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "vidstab_amd.h"
#include "vsvector.h"

enum { kMaxPending = 4, kAsyncMin = 16 };

typedef struct Release {
  struct Release* next;
  int n;
  LocalMotions v[];   /* the caller's vectors as they were handed over */
} Release;

static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static pthread_cond_t g_queued = PTHREAD_COND_INITIALIZER;   /* a release was queued */
static pthread_cond_t g_done = PTHREAD_COND_INITIALIZER;     /* a release finished */
static Release* g_head = NULL;
static Release* g_tail = NULL;
static int g_pending = 0;   /* releases queued or running */
static int g_thread = 0;    /* 0: not started, 1: running, -1: could not start or ended */
static int g_quit = 0;      /* the thread ends once the queue is empty */
static pthread_t g_tid;

static void* release_main(void* arg) {
  (void)arg;
  pthread_mutex_lock(&g_mu);
  for (;;) {
    while (!g_head && !g_quit) pthread_cond_wait(&g_queued, &g_mu);
    if (!g_head) break;
    Release* r = g_head;
    g_head = r->next;
    if (!g_head) g_tail = NULL;
    pthread_mutex_unlock(&g_mu);
    for (int i = 0; i < r->n; i++) vs_vector_del(&r->v[i]);
    free(r);
    pthread_mutex_lock(&g_mu);
    g_pending--;
    pthread_cond_broadcast(&g_done);
  }
  pthread_mutex_unlock(&g_mu);
  return NULL;
}

/* with g_mu held: 1 when the release thread runs and takes new releases */
static int thread_running(void) {
  if (g_thread == 0 && !g_quit) g_thread = pthread_create(&g_tid, NULL, release_main, NULL) == 0 ? 1 : -1;
  return g_thread == 1 && !g_quit;
}

VS_API void vsLocalMotionsArrayDel(LocalMotions* motions, int n) {
  if (!motions || n <= 0) return;
  Release* r = NULL;
  if (n >= kAsyncMin) r = (Release*)malloc(sizeof(Release) + sizeof(LocalMotions) * (size_t)n);
  if (r) {
    pthread_mutex_lock(&g_mu);
    const int async = thread_running();
    if (async) {
      while (g_pending >= kMaxPending) pthread_cond_wait(&g_done, &g_mu);
      r->next = NULL;
      r->n = n;
      memcpy(r->v, motions, sizeof(LocalMotions) * (size_t)n);
      if (g_tail)
        g_tail->next = r;
      else
        g_head = r;
      g_tail = r;
      g_pending++;
      pthread_cond_signal(&g_queued);
    }
    pthread_mutex_unlock(&g_mu);
    if (async) {
      /* emptied as vs_vector_del leaves a vector */
      for (int i = 0; i < n; i++) vs_vector_init(&motions[i], 0);
      return;
    }
    free(r);
  }
  for (int i = 0; i < n; i++) vs_vector_del(&motions[i]);
}

VS_API void vsLocalMotionsReleaseWait(void) {
  pthread_mutex_lock(&g_mu);
  while (g_pending > 0) pthread_cond_wait(&g_done, &g_mu);
  pthread_mutex_unlock(&g_mu);
}

/* before the library goes away (process exit or unload): the queue is
   drained, then the thread ends and is joined; later releases run in the
   calling thread */
__attribute__((destructor)) static void release_drain(void) {
  pthread_mutex_lock(&g_mu);
  const int running = g_thread == 1;
  g_quit = 1;
  pthread_cond_broadcast(&g_queued);
  pthread_mutex_unlock(&g_mu);
  if (running) pthread_join(g_tid, NULL);
  pthread_mutex_lock(&g_mu);
  g_thread = -1;
  pthread_mutex_unlock(&g_mu);
}
