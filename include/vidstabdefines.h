/*
 * vidstabdefines.h -- replaceable allocation/log hooks and the runtime status
 * codes of the vid.stab ABI.
 *
 * These are *variables*, not macros, exactly as in the reference
 * (src/vidstabdefines.h:88-118, defined in src/libvidstab.c:59-75): a caller
 * such as ffmpeg assigns its own allocator/logger and may even redefine the
 * values of VS_OK / VS_ERROR at run time.  Library code therefore compares
 * against the variables, never against literal 0 / -1.
 */
#ifndef VIDSTABDEFINES_H_
#define VIDSTABDEFINES_H_

#include <stddef.h>
#include <stdlib.h>
#include "vidstab_api.h"

#ifdef __GNUC__
#define likely(x)       __builtin_expect(!!(x), 1)
#define unlikely(x)     __builtin_expect(!!(x), 0)
#else
#define likely(x)       (x)
#define unlikely(x)     (x)
#endif

#define VS_MAX(a, b)    (((a) > (b)) ? (a) : (b))
#define VS_MIN(a, b)    (((a) < (b)) ? (a) : (b))
/* clamp x into [a, b] */
#define VS_CLAMP(x, a, b)  VS_MIN(VS_MAX((a), (x)), (b))

#define VS_DEBUG 2

/* single-channel pixel access, with and without a range check */
#define PIXEL(img, linesize, x, y, w, h, def) \
  (((x) < 0 || (y) < 0 || (x) >= (w) || (y) >= (h)) ? (def) : img[(x) + (y) * (linesize)])
#define PIX(img, linesize, x, y) (img[(x) + (y) * (linesize)])
/* N-channel (packed) pixel access */
#define PIXELN(img, linesize, x, y, w, h, N, channel, def)                     \
  (((x) < 0 || (y) < 0 || (x) >= (w) || (y) >= (h)                             \
    || (channel) < 0 || (channel) >= (N))                                      \
   ? (def) : img[(x)*(N) + (y) * (linesize) + (channel)])
#define PIXN(img, linesize, x, y, N, channel) \
  (img[(x)*(N) + (y) * (linesize) + (channel)])

/* hook types -- reference src/vidstabdefines.h:86-92 */
typedef void* (*vs_malloc_t) (size_t size);
typedef void* (*vs_realloc_t) (void* ptr, size_t size);
typedef void  (*vs_free_t) (void* ptr);
typedef void* (*vs_zalloc_t) (size_t size);
typedef int   (*vs_log_t) (int type, const char* tag, const char* format, ...);
typedef char* (*vs_strdup_t) (const char* s);

extern VS_API vs_log_t     vs_log;
extern VS_API int          vs_log_level;
extern VS_API vs_malloc_t  vs_malloc;
extern VS_API vs_realloc_t vs_realloc;
extern VS_API vs_free_t    vs_free;
extern VS_API vs_zalloc_t  vs_zalloc;
extern VS_API vs_strdup_t  vs_strdup;

extern VS_API int VS_ERROR_TYPE;
extern VS_API int VS_WARN_TYPE;
extern VS_API int VS_INFO_TYPE;
extern VS_API int VS_MSG_TYPE;

extern VS_API int VS_ERROR;
extern VS_API int VS_OK;

#define vs_log_error(...) vs_log(VS_ERROR_TYPE, __VA_ARGS__)
#define vs_log_warn(...)  vs_log(VS_WARN_TYPE,  __VA_ARGS__)
#define vs_log_info(...)  vs_log(VS_INFO_TYPE,  __VA_ARGS__)
#define vs_log_msg(...)   vs_log(VS_MSG_TYPE,   __VA_ARGS__)

#endif /* VIDSTABDEFINES_H_ */
