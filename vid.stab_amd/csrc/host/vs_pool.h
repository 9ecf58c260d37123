/*
 * vs_pool.h -- a small fixed thread pool for the per-frame host steps of a
 * batch (the double-precision transform, LocalMotions construction and the
 * match cleanup).  It stands where the reference's OpenMP loops over fields
 * stand (src/motiondetect.c:752-807): frames of a batch are independent.
 * Not exported.
 */
#ifndef VS_POOL_H
#define VS_POOL_H

typedef struct VSPool VSPool;

/* nthreads >= 1 (the caller counts as one); NULL on failure */
VSPool* vs_pool_create(int nthreads);
void vs_pool_destroy(VSPool* p);
/* threads that run a batch of work, the caller included */
int vs_pool_threads(const VSPool* p);
/* runs fn(ctx, i) for every i < n across the pool and returns when all are done */
void vs_pool_run(VSPool* p, int n, void (*fn)(void* ctx, int i), void* ctx);

#endif /* VS_POOL_H */
