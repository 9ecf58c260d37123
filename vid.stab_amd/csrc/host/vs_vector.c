/*
 * vs_vector.c -- the LocalMotions container of the detect ABI (interface:
 * include/vsvector.h, which keeps the reference's src/vsvector.h:35-40 layout).
 *
 * A VSVector is a caller-visible {data, buffersize, nelems} triple of element
 * pointers.  Ownership is the contract: every element this library creates
 * is one vs_malloc block, so a caller releases a result with vs_vector_del,
 * which returns each element and then the pointer array to vs_free (ffmpeg
 * installs its own allocator there).  Capacity grows by doubling through
 * vs_realloc; vs_vector_concat and vs_vector_filter share element pointers
 * instead of copying the elements.
 */
#include <string.h>

#include "vidstabdefines.h"
#include "vsvector.h"

/* capacity for at least `need` pointers (need <= VS_VECTOR_MAX_SIZE) */
static int reserve(VSVector* V, int need) {
  if (need <= V->buffersize && V->data) return VS_OK;
  if (need > VS_VECTOR_MAX_SIZE) {
    vs_log_error("vs_vector", "cannot hold %i elements\n", need);
    return VS_ERROR;
  }
  int cap = V->buffersize > 0 ? V->buffersize : 4;
  while (cap < need) cap = cap > VS_VECTOR_MAX_SIZE / 2 ? need : 2 * cap;
  void** grown = (void**)vs_realloc(V->data, sizeof(void*) * (size_t)cap);
  if (!grown) {
    vs_log_error("vs_vector", "no memory for %i element pointers\n", cap);
    return VS_ERROR;
  }
  V->data = grown;
  V->buffersize = cap;
  return VS_OK;
}

/* a vs_malloc'd copy of `size` bytes */
static void* dup_block(const void* src, int size) {
  void* p = vs_malloc((size_t)(size > 0 ? size : 1));
  if (p && size > 0) memcpy(p, src, (size_t)size);
  return p;
}

VS_API int vs_vector_init(VSVector* V, int buffersize) {
  V->nelems = 0;
  V->buffersize = 0;
  V->data = NULL;
  if (buffersize <= 0) return VS_OK;
  V->data = (void**)vs_zalloc(sizeof(void*) * (size_t)buffersize);
  if (!V->data) return VS_ERROR;
  V->buffersize = buffersize;
  return VS_OK;
}

VS_API int vs_vector_fini(VSVector* V) {
  if (V->data) vs_free(V->data);
  return vs_vector_init(V, 0);
}

VS_API int vs_vector_zero(VSVector* V) {
  for (int i = 0; i < V->nelems; ++i)
    if (V->data[i]) vs_free(V->data[i]);
  V->nelems = 0;
  return VS_OK;
}

VS_API int vs_vector_del(VSVector* V) {
  vs_vector_zero(V);
  return vs_vector_fini(V);
}

VS_API int vs_vector_size(const VSVector* V) { return V->nelems; }

VS_API void* vs_vector_get(const VSVector* V, int pos) {
  return pos >= 0 && pos < V->nelems ? V->data[pos] : NULL;
}

VS_API int vs_vector_append(VSVector* V, void* data) {
  if (reserve(V, V->nelems + 1) != VS_OK) return VS_ERROR;
  V->data[V->nelems++] = data;
  return VS_OK;
}

VS_API int vs_vector_append_dup(VSVector* V, void* data, int data_size) {
  void* copy = dup_block(data, data_size);
  if (!copy) return VS_ERROR;
  if (vs_vector_append(V, copy) == VS_OK) return VS_OK;
  vs_free(copy);
  return VS_ERROR;
}

/* element pos becomes data; slots between the old end and pos read NULL */
VS_API int vs_vector_set(VSVector* V, int pos, void* data, void** old) {
  if (old) *old = NULL;
  if (pos < 0 || pos >= VS_VECTOR_MAX_SIZE) {
    vs_log_error("vs_vector", "cannot set element %i\n", pos);
    return VS_ERROR;
  }
  if (reserve(V, pos + 1) != VS_OK) return VS_ERROR;
  while (V->nelems <= pos) V->data[V->nelems++] = NULL;
  if (old) *old = V->data[pos];
  V->data[pos] = data;
  return VS_OK;
}

VS_API int vs_vector_set_dup(VSVector* V, int pos, void* data, int data_size, void** old) {
  if (old) *old = NULL;
  void* copy = dup_block(data, data_size);
  if (!copy) return VS_ERROR;
  if (vs_vector_set(V, pos, copy, old) == VS_OK) return VS_OK;
  vs_free(copy);
  return VS_ERROR;
}

VS_API VSVector vs_vector_filter(const VSVector* V, short (*pred)(void*, void*), void* param) {
  VSVector out;
  vs_vector_init(&out, V->nelems);
  for (int i = 0; i < V->nelems; ++i)
    if (pred(param, V->data[i])) vs_vector_append(&out, V->data[i]);
  return out;
}

VS_API VSVector vs_vector_concat(const VSVector* V1, const VSVector* V2) {
  VSVector out;
  const int n = V1->nelems + V2->nelems;
  if (vs_vector_init(&out, n) != VS_OK) return out;
  for (int i = 0; i < V1->nelems; ++i) out.data[out.nelems++] = V1->data[i];
  for (int i = 0; i < V2->nelems; ++i) out.data[out.nelems++] = V2->data[i];
  return out;
}

/* the double array the match statistics use (localmotionsGetMatch) */
VS_API VSArray vs_array_new(int len) {
  VSArray a;
  a.dat = (double*)vs_zalloc(sizeof(double) * (size_t)(len > 0 ? len : 1));
  a.len = a.dat ? len : 0;
  return a;
}

VS_API void vs_array_free(VSArray a) {
  if (a.dat) vs_free(a.dat);
}
