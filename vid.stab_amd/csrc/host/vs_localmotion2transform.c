/*
 * vs_localmotion2transform.c -- the second pass's LocalMotions -> global
 * transforms step on the GPU (reference: src/localmotion2transform.c:40-118,
 * 213-283).
 *
 *   vsLocalmotions2Transforms  the reference's entry point, taking the
 *                              transform pass's VSTransformData: the
 *                              clip-wide lens estimate (vs_lensdistortion.c),
 *                              then every frame's transform in one launch --
 *                              through the estimated lens when it is
 *                              determined and large enough to matter
 *                              (vsg_lens_transforms), else vsMotionsToTransform
 *                              with the configured camera model
 *                              (vsg_motions_to_transform_fov) -- or the simple
 *                              per-frame transform when simpleMotionCalculation
 *                              is set
 *   vsMotionsToTransformBatch  vsMotionsToTransform of every frame of a clip
 *                              with the similarity model, for callers without
 *                              a VSTransformData
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lensdistortion.h"
#include "localmotion2transform.h"
#include "transform.h"
#include "transformtype_operations.h"
#include "vidstabdefines.h"
#include "vs_detect_internal.h"
#include "vsg.h"

static int env_device(void) { return vsd_device(); }

/* src/localmotion2transform.c:40-118 */
VS_API int vsLocalmotions2Transforms(VSTransformData* td, const VSManyLocalMotions* motions, VSTransformations* trans) {
  if (!td || !motions || !trans) return VS_ERROR;
  const char* mod = td->conf.modName ? td->conf.modName : "vidstab_amd";
  if (trans->len != 0 || trans->ts != NULL) {
    vs_log_error(mod, "vsLocalmotions2Transforms: the transformations must be empty\n");
    return VS_ERROR;
  }
  const int len = vs_vector_size(motions);
  trans->ts = (VSTransform*)vs_malloc(sizeof(VSTransform) * (size_t)(len > 0 ? len : 1));
  if (!trans->ts) return VS_ERROR;
  FILE* f = td->conf.storeTransforms ? fopen("global_motions.trf", "w") : NULL;
  int16_t* lm = NULL;
  double* match = NULL;
  int* count = NULL;
  double* residual = NULL;
  int* rounds = NULL;
  VSGLens* gpu = NULL;
  int rc = VS_ERROR;
  const double focal = focal_from_fov(td->conf.fov, td->fiSrc.width);
  if (len == 0) {
    rc = VS_OK;
  } else if (td->conf.simpleMotionCalculation != 0) {
    for (int i = 0; i < len; i++) trans->ts[i] = vsSimpleMotionsToTransform(td->fiSrc, mod, VSMLMGet(motions, i));
    rc = VS_OK;
  } else if (vsd_flatten_motions(motions, &lm, &match, &count) == VS_OK &&
             (residual = (double*)malloc(sizeof(double) * (size_t)len)) != NULL &&
             (rounds = (int*)malloc(sizeof(int) * (size_t)len)) != NULL) {
    /* the lens strength of the whole clip: one number pooled over every
       frame pair, used only when determined and above |k| = 0.01 */
    int useLens = 0;
    VSLensEstimate le;
    memset(&le, 0, sizeof(le));
    VSLensEstimateConfig lcfg = vsLensEstimateGetDefaultConfig();
    lcfg.f = focal;
    int gpuOk = 1;
    if (td->conf.estimateLensDistortion) {
      int failed = 1;
      if (vsg_lens_create(&gpu, env_device(), td->fiSrc.width, td->fiSrc.height, lm, match, count, len) == 0)
        le = vsd_lens_estimate(gpu, len, count, &lcfg, &failed);
      if (failed) {
        vs_log_error(mod, "lens distortion estimate on the GPU failed: %s\n", vsg_last_error());
        gpuOk = 0;
      }
      useLens = gpuOk && le.determined && fabs(le.k) > 0.01;
      /* an explicit conf.lensK wins (vsTransformSetLensK, src/transform.c:166-168) */
      if (fabs(td->conf.lensK) <= 0.01 && useLens) td->lensK = le.k;
      if (gpuOk)
        vs_log_info(mod, "Lens distortion: k=%.4f +- %.4f from %i motions (%i dropped), %s\n", le.k,
                    le.uncertainty, le.used, le.rejected,
                    useLens ? "correcting transforms" : (le.determined ? "too small to matter" : "not determined"));
    }
    if (!gpuOk) {
      /* rc stays VS_ERROR */
    } else if (useLens) {
      int* status = rounds;   /* per-frame status of the lens fits */
      if (vsg_lens_transforms(gpu, le.k, focal, lcfg.gaussNewtonSteps, lcfg.rejectOutliers, lcfg.outlierStddevs,
                              trans->ts, residual, status) == 0) {
        rc = VS_OK;
        if (f)
          for (int i = 0; i < len; i++)
            fprintf(f, "0 %f %f %f %f %i\n#\t\t\t\t\t %f lens\n", trans->ts[i].x, trans->ts[i].y,
                    trans->ts[i].alpha, trans->ts[i].zoom, trans->ts[i].extra, residual[i]);
      } else {
        vs_log_error(mod, "lens transforms on the GPU failed: %s\n", vsg_last_error());
      }
    } else if (vsg_motions_to_transform_fov(env_device(), td->fiSrc.width, td->fiSrc.height, lm, match, count, len,
                                            td->conf.smoothZoom, focal, trans->ts, residual, rounds) == 0) {
      rc = VS_OK;
      if (f)
        for (int i = 0; i < len; i++) {
          const VSTransform* t = &trans->ts[i];
          if (count[i] == 0)
            fprintf(f, "0 0 0 0 0 %i\n# no fields\n", t->extra);
          else
            fprintf(f, "0 %f %f %f %f %i\n#\t\t\t\t\t %f %i\n", t->x, t->y, t->alpha, t->zoom, t->extra,
                    residual[i], rounds[i]);
        }
    } else {
      vs_log_error(mod, "vsLocalmotions2Transforms: %s\n", vsg_last_error());
    }
  }
  if (rc == VS_OK) {
    trans->len = len;
  } else {
    vs_free(trans->ts);
    trans->ts = NULL;
  }
  vsg_lens_destroy(gpu);
  free(lm);
  free(match);
  free(count);
  free(residual);
  free(rounds);
  if (f) fclose(f);
  return rc;
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
