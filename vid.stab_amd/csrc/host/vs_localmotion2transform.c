/*
 * vs_localmotion2transform.c -- the second pass's per-frame global transform
 * for a whole clip on the GPU: vsMotionsToTransform
 * (src/localmotion2transform.c:213-283) of every frame's LocalMotions, the
 * frames batched into one launch (vsg_motions_to_transform).
 *
 * The reference runs vsMotionsToTransform frame by frame inside
 * vsLocalmotions2Transforms (:40-118).  This entry point covers that loop for
 * the similarity model (VSTransformConfig.fov 0) without the clip-wide lens
 * estimate: its output for frame f is what vsMotionsToTransform returns for
 * motions[f] with fov 0 and the given smoothZoom, up to the last-bit
 * differences of the device's sin and cos.
 */
#include <stdlib.h>
#include <string.h>

#include "localmotion2transform.h"
#include "vidstabdefines.h"
#include "vsg.h"

static int env_device(void) {
  const char* v = getenv("VIDSTAB_AMD_DEVICE");
  return v ? atoi(v) : 0;
}

VS_API int vsMotionsToTransformBatch(const VSFrameInfo* fi, const LocalMotions* motions, int n, int smoothZoom,
                                     VSTransform* out) {
  if (!fi || n < 0 || (n > 0 && (!motions || !out))) return VS_ERROR;
  if (n == 0) return VS_OK;
  long long total = 0;
  int* count = (int*)malloc(sizeof(int) * (size_t)n);
  if (!count) return VS_ERROR;
  for (int f = 0; f < n; f++) {
    count[f] = vs_vector_size(&motions[f]);
    total += count[f];
  }
  const size_t nm = total > 0 ? (size_t)total : 1;
  int16_t* lm = (int16_t*)malloc(sizeof(int16_t) * 4 * nm);
  double* match = (double*)malloc(sizeof(double) * nm);
  int rc = VS_ERROR;
  if (lm && match) {
    size_t k = 0;
    for (int f = 0; f < n; f++)
      for (int i = 0; i < count[f]; i++, k++) {
        const LocalMotion* m = (const LocalMotion*)vs_vector_get(&motions[f], i);
        lm[4 * k] = m->f.x;
        lm[4 * k + 1] = m->f.y;
        lm[4 * k + 2] = m->v.x;
        lm[4 * k + 3] = m->v.y;
        match[k] = m->match;
      }
    if (vsg_motions_to_transform(env_device(), fi->width, fi->height, lm, match, count, n, smoothZoom, out, NULL,
                                 NULL) == 0)
      rc = VS_OK;
    else
      vs_log_error("vidstab_amd", "vsMotionsToTransformBatch: %s\n", vsg_last_error());
  }
  free(lm);
  free(match);
  free(count);
  return rc;
}
