/*
 * vs_vector.c -- VSVector / VSArray of the vid.stab ABI (reference:
 * src/vsvector.c:36-275).  The detect pass returns its LocalMotions in a
 * VSVector whose elements are separately allocated through vs_malloc, so the
 * caller's vs_vector_del (through vs_free) releases them.
 */
#include <assert.h>
#include <string.h>

#include "vidstabdefines.h"
#include "vsvector.h"

static int vec_resize(VSVector* V, int newsize) {
  if (newsize < 1) newsize = 1;
  void** d = (void**)vs_realloc(V->data, (size_t)newsize * sizeof(void*));
  if (!d) {
    vs_log_error("VS_Vector", "out of memory!");
    return VS_ERROR;
  }
  V->data = d;
  V->buffersize = newsize;
  if (V->nelems > V->buffersize) V->nelems = V->buffersize;
  return VS_OK;
}

VS_API int vs_vector_init(VSVector* V, int buffersize) {
  assert(V);
  if (buffersize > 0) {
    V->data = (void**)vs_zalloc(sizeof(void*) * (size_t)buffersize);
    if (!V->data) return VS_ERROR;
  } else {
    V->data = 0;
    buffersize = 0;
  }
  V->buffersize = buffersize;
  V->nelems = 0;
  return VS_OK;
}

VS_API int vs_vector_fini(VSVector* V) {
  assert(V);
  if (V->data) vs_free(V->data);
  V->data = 0;
  V->buffersize = 0;
  V->nelems = 0;
  return VS_OK;
}

VS_API int vs_vector_zero(VSVector* V) {
  assert(V);
  for (int i = 0; i < V->nelems; i++)
    if (V->data[i]) vs_free(V->data[i]);
  V->nelems = 0;
  return VS_OK;
}

VS_API int vs_vector_del(VSVector* V) {
  vs_vector_zero(V);
  return vs_vector_fini(V);
}

VS_API int vs_vector_size(const VSVector* V) {
  assert(V);
  return V->nelems;
}

VS_API int vs_vector_append(VSVector* V, void* data) {
  assert(V && data);
  if (!V->data || V->buffersize < 1)
    if (vs_vector_init(V, 4) != VS_OK) return VS_ERROR;
  if (V->nelems >= V->buffersize)
    if (vec_resize(V, V->buffersize * 2) != VS_OK) return VS_ERROR;
  V->data[V->nelems++] = data;
  return VS_OK;
}

VS_API int vs_vector_append_dup(VSVector* V, void* data, int data_size) {
  assert(V && data);
  void* d = vs_malloc((size_t)data_size);
  if (!d) return VS_ERROR;
  memcpy(d, data, (size_t)data_size);
  if (vs_vector_append(V, d) != VS_OK) {
    vs_free(d);
    return VS_ERROR;
  }
  return VS_OK;
}

VS_API void* vs_vector_get(const VSVector* V, int pos) {
  assert(V && V->data);
  if (pos < 0 || pos >= V->nelems) return 0;
  return V->data[pos];
}

VS_API int vs_vector_set(VSVector* V, int pos, void* data, void** old) {
  assert(V && data);
  if (old) *old = 0;
  if (pos < 0 || pos >= VS_VECTOR_MAX_SIZE) {
    vs_log_error("VS_Vector", "position %i out of range [0,%i)!", pos, VS_VECTOR_MAX_SIZE);
    return VS_ERROR;
  }
  if (!V->data || V->buffersize < 1)
    if (vs_vector_init(V, 4) != VS_OK) return VS_ERROR;
  if (V->buffersize <= pos) {
    int nsize = V->buffersize;
    while (nsize <= pos && nsize <= VS_VECTOR_MAX_SIZE / 2) nsize *= 2;
    if (nsize <= pos) nsize = pos + 1;
    if (vec_resize(V, nsize) != VS_OK) return VS_ERROR;
  }
  if (pos >= V->nelems) {
    for (int i = V->nelems; i < pos + 1; i++) V->data[i] = 0;
    V->nelems = pos + 1;
  }
  if (old) *old = V->data[pos];
  V->data[pos] = data;
  return VS_OK;
}

VS_API int vs_vector_set_dup(VSVector* V, int pos, void* data, int data_size, void** old) {
  if (old) *old = 0;
  void* d = vs_malloc((size_t)data_size);
  if (!d) return VS_ERROR;
  memcpy(d, data, (size_t)data_size);
  int r = vs_vector_set(V, pos, d, old);
  if (r != VS_OK) vs_free(d);
  return r;
}

VS_API VSVector vs_vector_filter(const VSVector* V, short (*pred)(void*, void*), void* param) {
  VSVector result;
  assert(V);
  vs_vector_init(&result, V->nelems);
  for (int i = 0; i < V->nelems; i++)
    if (pred(param, V->data[i])) vs_vector_append(&result, V->data[i]);
  return result;
}

VS_API VSVector vs_vector_concat(const VSVector* V1, const VSVector* V2) {
  VSVector result;
  assert(V1 && V2);
  vs_vector_init(&result, V1->nelems + V2->nelems);
  if (V1->nelems) memcpy(result.data, V1->data, sizeof(void*) * (size_t)V1->nelems);
  if (V2->nelems) memcpy(result.data + V1->nelems, V2->data, sizeof(void*) * (size_t)V2->nelems);
  result.nelems = V1->nelems + V2->nelems;
  return result;
}

VS_API VSArray vs_array_new(int len) {
  VSArray a;
  a.dat = (double*)vs_zalloc(sizeof(double) * (size_t)len);
  a.len = len;
  return a;
}

VS_API VSArray vs_array(double vals[], int len) {
  VSArray a = vs_array_new(len);
  memcpy(a.dat, vals, sizeof(double) * (size_t)len);
  return a;
}

VS_API VSArray* vs_array_plus(VSArray* c, VSArray a, VSArray b) {
  assert(a.len == b.len);
  if (c->len == 0) *c = vs_array_new(a.len);
  for (int i = 0; i < a.len; i++) c->dat[i] = a.dat[i] + b.dat[i];
  return c;
}

VS_API VSArray* vs_array_scale(VSArray* c, VSArray a, double f) {
  if (c->len == 0) *c = vs_array_new(a.len);
  for (int i = 0; i < a.len; i++) c->dat[i] = a.dat[i] * f;
  return c;
}

VS_API VSArray vs_array_copy(VSArray a) {
  VSArray c = vs_array_new(a.len);
  memcpy(c.dat, a.dat, sizeof(double) * (size_t)a.len);
  return c;
}

VS_API void vs_array_zero(VSArray* a) { memset(a->dat, 0, sizeof(double) * (size_t)a->len); }

VS_API void vs_array_swap(VSArray* a, VSArray* b) {
  VSArray t = *a;
  *a = *b;
  *b = t;
}

VS_API void vs_array_free(VSArray a) { vs_free(a.dat); }

VS_API void vs_array_print(VSArray a, FILE* f) {
  for (int i = 0; i < a.len; i++) fprintf(f, "%g ", a.dat[i]);
}
