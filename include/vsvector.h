/*
 * vsvector.h -- the growable pointer vector that carries LocalMotions.
 *
 * Layout (16 bytes) and ownership rules follow the reference
 * (src/vsvector.h:35-40): every element is a separately allocated block
 * obtained through vs_malloc, and vs_vector_del frees them with vs_free.
 * The caller of vsMotionDetection owns the returned elements.
 */
#ifndef VSVECTOR_H
#define VSVECTOR_H

#include <stddef.h>
#include <stdio.h>
#include <limits.h>
#include "vidstab_api.h"

typedef struct vsvector_ VSVector;
struct vsvector_ {
  void** data;
  int    buffersize;
  int    nelems;
};

/* src/vsvector.c:36-48 */
VS_API int vs_vector_init(VSVector* V, int buffersize);
/* frees the pointer array only; src/vsvector.c:50-57 */
VS_API int vs_vector_fini(VSVector* V);
/* frees elements and array; src/vsvector.c:59-62 */
VS_API int vs_vector_del(VSVector* V);
/* frees the elements, keeps the array; src/vsvector.c:64-75 */
VS_API int vs_vector_zero(VSVector* V);
VS_API int vs_vector_size(const VSVector* V);
/* src/vsvector.c:82-94 */
VS_API int vs_vector_append(VSVector* V, void* data);
/* appends a vs_malloc'd copy; src/vsvector.c:96-106 */
VS_API int vs_vector_append_dup(VSVector* V, void* data, int data_size);
#define VS_VECTOR_MAX_SIZE ((int)(INT_MAX / sizeof(void*)))
VS_API int vs_vector_set(VSVector* V, int pos, void* data, void** old);
VS_API int vs_vector_set_dup(VSVector* V, int pos, void* data, int data_size, void** old);
/* NULL for an out-of-range position; src/vsvector.c:108-114 */
VS_API void* vs_vector_get(const VSVector* V, int pos);
/* shares (does not copy) the selected elements; src/vsvector.c:180-193 */
VS_API VSVector vs_vector_filter(const VSVector* V, short (*pred)(void*, void*), void* param);
/* moves the element pointers of both into a new vector; src/vsvector.c:195-203 */
VS_API VSVector vs_vector_concat(const VSVector* V1, const VSVector* V2);

/* the double array of match statistics (localmotionsGetMatch); of the
   reference's VSArray helpers (src/vsvector.h:190-222) the detect path needs
   only creation and release */
typedef struct vsarray_ VSArray;
struct vsarray_ {
  double* dat;
  int len;
};
VS_API VSArray vs_array_new(int len);
VS_API void vs_array_free(VSArray a);

#endif /* VSVECTOR_H */
