/*
 * vs_glue.c -- the replaceable allocation/log hooks and runtime status codes
 * of the vid.stab ABI (reference: src/libvidstab.c:35-75).
 *
 * They are data symbols so that a caller can swap the allocator and logger,
 * or remap VS_OK / VS_ERROR, at run time.  The defaults are the C library's
 * allocator and a stderr logger filtered by vs_log_level.
 */
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "vidstabdefines.h"

static void* vs_default_zalloc(size_t size) {
  void* p = malloc(size);
  if (p) memset(p, 0, size);
  return p;
}

static int vs_default_log(int type, const char* tag, const char* format, ...) {
  if (vs_log_level < type) return 0;
  const char* kind = type == VS_ERROR_TYPE ? "Error: "
                   : type == VS_WARN_TYPE  ? "Warn:  "
                   : type == VS_INFO_TYPE  ? "Info:  "
                   : type == VS_MSG_TYPE   ? "Msg:   " : "Unknown";
  fprintf(stderr, "%s (%s):", kind, tag ? tag : "");
  va_list ap;
  va_start(ap, format);
  vfprintf(stderr, format, ap);
  va_end(ap);
  fputc('\n', stderr);
  return 0;
}

vs_malloc_t  vs_malloc  = malloc;
vs_realloc_t vs_realloc = realloc;
vs_free_t    vs_free    = free;
vs_zalloc_t  vs_zalloc  = vs_default_zalloc;
vs_strdup_t  vs_strdup  = strdup;
vs_log_t     vs_log     = vs_default_log;

int VS_ERROR_TYPE = 0;
int VS_WARN_TYPE  = 1;
int VS_INFO_TYPE  = 2;
int VS_MSG_TYPE   = 3;

int VS_ERROR = -1;
int VS_OK    = 0;

int vs_log_level = 4;
