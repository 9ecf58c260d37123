/*
 * transformtype_operations.h -- the transform/statistics helpers that the
 * detect pass uses between its GPU stages (all host double arithmetic).
 *
 * Only the on-path subset of the reference's src/transformtype_operations.h
 * is provided; each declaration cites the reference function it replaces.
 */
#ifndef __TRANSFORMTYPE_OPERATIONS_H
#define __TRANSFORMTYPE_OPERATIONS_H

#include "transformtype.h"
#include "vidstabdefines.h"
#include "vsvector.h"
#include "frameinfo.h"
#include "vidstab_api.h"

/* element access; src/transformtype_operations.h:35-36 */
#define LMGet(localmotions,index) \
    ((LocalMotion*)vs_vector_get(localmotions,index))

/* src/transformtype.c:38-62 */
VS_API VSTransform null_transform(void);
VS_API VSTransform new_transform(double x, double y, double alpha,
                                 double zoom, double barrel, double rshutter, int extra);

/* Same layout as src/transformtype_operations.h:56-68 (the rotational-model
   members are carried so the struct matches; the host steps of the detect
   pass use the similarity model, f == 0). */
typedef struct _preparedtransform {
  const VSTransform* t;
  double zcos_a;
  double zsin_a;
  double c_x;
  double c_y;
  double f;
  double z;
  double r[9];
} PreparedTransform;

/* src/transformtype.c:158-160 (similarity model only) */
VS_API PreparedTransform prepare_transform(const VSTransform* t, const VSFrameInfo* fi);
/* focal length in px of a horizontal field of view in degrees, 0 for
   fov <= 0 or >= 180; src/transformtype.c:123-127 */
VS_API double focal_from_fov(double fovDeg, int width);
/* truncating integer version; src/transformtype.c:186-191 */
VS_API Vec transform_vec(const PreparedTransform* t, const Vec* v);
/* src/transformtype.c:193-209 */
VS_API void transform_vec_double(double* x, double* y, const PreparedTransform* t, const Vec* v);

/* src/transformtype.c:211-222 */
VS_API Vec sub_vec(Vec v1, Vec v2);
VS_API Vec add_vec(Vec v1, Vec v2);
VS_API Vec field_to_vec(Field f);

/* ascending comparators; src/transformtype.c:240-261 */
VS_API int cmp_double(const void* t1, const void* t2);
VS_API int cmp_int(const void* t1, const void* t2);

/* 20%-trimmed mean; sorts ds.  src/transformtype.c:445-459 */
VS_API double cleanmean(double* ds, int len, double* minimum, double* maximum);

/* src/transformtype.c:464-468, 490-495, 513-534, 536-542 */
VS_API LocalMotion null_localmotion(void);
VS_API LocalMotion sub_localmotion(const LocalMotion* lm1, const LocalMotion* lm2);
VS_API LocalMotion cleanmean_localmotions(const LocalMotions* localmotions);
VS_API VSArray localmotionsGetMatch(const LocalMotions* localmotions);

#endif /* __TRANSFORMTYPE_OPERATIONS_H */
