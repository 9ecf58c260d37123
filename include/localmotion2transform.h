/*
 * localmotion2transform.h -- the LocalMotions->transform steps: the one inside
 * the detect pass (coarse result -> offset of the fine search) and the
 * second pass's per-frame fit, batched over a clip on the GPU.
 *
 * vsLocalmotions2Transforms itself (it takes the transform pass's
 * VSTransformData, with the lens estimate and the fov model) stays outside
 * this library; vsMotionsToTransformBatch covers its per-frame loop for the
 * similarity model.
 */
#ifndef __LOCALMOTION2TRANSFORM_H
#define __LOCALMOTION2TRANSFORM_H

#include "transformtype.h"
#include "frameinfo.h"
#include "vidstab_api.h"

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
