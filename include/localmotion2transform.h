/*
 * localmotion2transform.h -- the one LocalMotions->transform step that sits
 * inside the detect pass (coarse result -> offset of the fine search).
 *
 * The second-pass functions of the reference header (vsLocalmotions2Transforms,
 * vsMotionsToTransform, ...) are outside this library's scope.
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

#endif /* __LOCALMOTION2TRANSFORM_H */
