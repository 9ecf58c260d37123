/*
 * localmotion2transform.h -- the LocalMotions->transform steps: the one inside
 * the detect pass (coarse result -> offset of the fine search) and the
 * second pass's, on the GPU: vsLocalmotions2Transforms (the reference's
 * entry point, with the lens estimate and the rotational camera model) and
 * vsMotionsToTransformBatch (its per-frame fit for callers without a
 * VSTransformData).
 */
#ifndef __LOCALMOTION2TRANSFORM_H
#define __LOCALMOTION2TRANSFORM_H

#include "transformtype.h"
#include "frameinfo.h"
#include "serialize.h"
#include "transform.h"
#include "vidstab_api.h"

/* every frame's global transform of a clip's LocalMotions into trans (empty
   on entry; trans->ts is vs_malloc'd): with the transform config's
   simpleMotionCalculation, vsSimpleMotionsToTransform per frame; otherwise,
   with estimateLensDistortion, first the clip's lens strength k (the GPU
   estimate, vsEstimateLensDistortion), written to td->lensK unless
   conf.lensK is set; then, when k is determined and |k| > 0.01, every
   frame's fit through that lens (vsLensMotionsToTransform), else
   vsMotionsToTransform with conf.fov's camera model.  With storeTransforms
   the transforms also go to global_motions.trf.  Transforms through the lens
   agree with the reference within the estimate's tolerance; the others match
   it to the last bits the device's sin and cos leave open.
   src/localmotion2transform.h:38-40, src/localmotion2transform.c:40-118 */
VS_API int vsLocalmotions2Transforms(VSTransformData* td, const VSManyLocalMotions* motions,
                                     VSTransformations* trans);

/* rotation of one field about (center_x, center_y) implied by its motion;
   0 for fields within 2*size (L1) of the centre.
   src/localmotion2transform.c:336-349 */
VS_API double vsCalcAngle(const LocalMotion* lm, int center_x, int center_y);

/* trimmed-mean translation plus trimmed-mean rotation (n >= 6) of a set of
   LocalMotions, compensated for the off-centre rotation centre.
   src/localmotion2transform.h, src/localmotion2transform.c:352-402 */
VS_API VSTransform vsSimpleMotionsToTransform(VSFrameInfo fi, const char* modName,
                                              const LocalMotions* motions);

/* vsMotionsToTransform (src/localmotion2transform.c:213-283; the reference
   declares it at src/localmotion2transform.h:65) for n frames at once on the
   GPU, with the transform config's fov 0 (similarity model) and no lens
   estimate: out[f] is the global transform of motions[f] -- the mean motion
   refined by gradient descent on the fields' squared residuals, with badly
   matched fields and then outliers disabled (up to three descents);
   extra = 1 for a frame without motions or with a residual above 100; the
   zoom is kept only when smoothZoom != 0.  Matches the reference to the last
   bits the device's sin and cos leave open.  Returns VS_OK or VS_ERROR. */
VS_API int vsMotionsToTransformBatch(const VSFrameInfo* fi, const LocalMotions* motions, int n, int smoothZoom,
                                     VSTransform* out);

#endif /* __LOCALMOTION2TRANSFORM_H */
