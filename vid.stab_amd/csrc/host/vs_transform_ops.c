/*
 * vs_transform_ops.c -- the host statistics and transform helpers used
 * between the GPU stages of the detect pass (reference:
 * src/transformtype.c:38-62, 158-222, 240-261, 445-542;
 * src/localmotion2transform.c:336-402).
 *
 * Bit-exactness: the reference ships built with -O3 -msse2 -ffast-math, and
 * its compiled code evaluates some double expressions in a different order
 * than its C source.  This file is compiled with -ffp-contract=off and no
 * fast-math, and writes each expression in the order the compiled reference
 * executes:
 *   cleanmean            two-lane sum (even / odd offsets), lanes added, then
 *                        the odd tail element
 *   transform_vec_double (zcos*rx + zsin*ry) + (t.x + c_x), and
 *                        ((c_x - x)*zsin + ry*zcos) + (t.y + c_y)
 *   vsSimpleMotionsToTransform
 *                        t.x = ((cos a - 1)*px - sin a*py) + mean.x,
 *                        t.y = (mean.y + sin a*px) + (cos a - 1)*py
 *   prepare_transform    z = 1 + zoom*0.01 (the division became a multiply)
 * Transcendentals are glibc's (sincos, atan2), as in the reference.
 */
#define _GNU_SOURCE 1
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "localmotion2transform.h"
#include "transformtype_operations.h"
#include "vidstabdefines.h"

VS_API VSTransform new_transform(double x, double y, double alpha, double zoom, double barrel,
                                 double rshutter, int extra) {
  VSTransform t;
  t.x = x;
  t.y = y;
  while (alpha < -M_PI) alpha += 2.0 * M_PI;
  while (alpha > M_PI) alpha -= 2.0 * M_PI;
  t.alpha = alpha;
  t.zoom = zoom;
  t.barrel = barrel;
  t.rshutter = rshutter;
  t.extra = extra;
  return t;
}

VS_API VSTransform null_transform(void) { return new_transform(0, 0, 0, 0, 0, 0, 0); }

/* src/transformtype.c:123-127, in its compiled order: (width * 0.5) /
   tan(fov * (pi / 360)) */
VS_API double focal_from_fov(double fovDeg, int width) {
  if (fovDeg <= 0.0 || fovDeg >= 180.0 || width <= 0) return 0.0;
  const double halfDegree = M_PI / 360.0;
  return ((double)width * 0.5) / tan(fovDeg * halfDegree);
}

VS_API PreparedTransform prepare_transform(const VSTransform* t, const VSFrameInfo* fi) {
  PreparedTransform pt;
  memset(&pt, 0, sizeof(pt));
  double sn, cs;
  pt.t = t;
  double z = 1.0 + t->zoom * 0.01;
  sincos(t->alpha, &sn, &cs);
  pt.zcos_a = cs * z;
  pt.zsin_a = z * sn;
  pt.c_x = (double)(fi->width / 2);
  pt.c_y = (double)(fi->height / 2);
  pt.z = z;
  pt.f = 0.0;
  return pt;
}

VS_API void transform_vec_double(double* x, double* y, const PreparedTransform* pt, const Vec* v) {
  const double vx = (double)v->x, vy = (double)v->y;
  const double rx = vx - pt->c_x, ry = vy - pt->c_y;
  *x = (rx * pt->zcos_a + pt->zsin_a * ry) + (pt->t->x + pt->c_x);
  *y = ((pt->c_x - vx) * pt->zsin_a + ry * pt->zcos_a) + (pt->t->y + pt->c_y);
}

VS_API Vec transform_vec(const PreparedTransform* pt, const Vec* v) {
  double x, y;
  transform_vec_double(&x, &y, pt, v);
  Vec r;
  r.x = (int16_t)(int)x;   /* truncation, low 16 bits */
  r.y = (int16_t)(int)y;
  return r;
}

VS_API Vec sub_vec(Vec a, Vec b) {
  Vec r = {(int16_t)(a.x - b.x), (int16_t)(a.y - b.y)};
  return r;
}

VS_API Vec add_vec(Vec a, Vec b) {
  Vec r = {(int16_t)(a.x + b.x), (int16_t)(a.y + b.y)};
  return r;
}

VS_API Vec field_to_vec(Field f) {
  Vec r = {f.x, f.y};
  return r;
}

VS_API int cmp_double(const void* a, const void* b) {
  double x = *(const double*)a, y = *(const double*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

VS_API int cmp_int(const void* a, const void* b) {
  int x = *(const int*)a, y = *(const int*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

VS_API double cleanmean(double* ds, int len, double* minimum, double* maximum) {
  const int cut = len / 5;
  qsort(ds, (size_t)len, sizeof(double), cmp_double);
  const int count = len - cut - cut;
  double sum = 0.0;
  if (count == 1) {
    sum = 0.0 + ds[cut];
  } else if (count > 1) {
    double lane0 = 0.0, lane1 = 0.0;
    const double* p = ds + cut;
    int i = 0;
    for (; i + 1 < count; i += 2) {
      lane0 += p[i];
      lane1 += p[i + 1];
    }
    sum = lane1 + lane0;
    if (i < count) sum += p[i];
  }
  if (minimum) *minimum = ds[cut];
  if (maximum) *maximum = ds[len - cut - 1];
  return sum / ((double)len - ((double)cut + (double)cut));
}

VS_API LocalMotion null_localmotion(void) {
  LocalMotion lm;
  memset(&lm, 0, sizeof(lm));
  return lm;
}

VS_API LocalMotion sub_localmotion(const LocalMotion* a, const LocalMotion* b) {
  LocalMotion r = *a;
  r.v.x = (int16_t)(r.v.x - b->v.x);
  r.v.y = (int16_t)(r.v.y - b->v.y);
  return r;
}

VS_API LocalMotion cleanmean_localmotions(const LocalMotions* lms) {
  const int len = vs_vector_size(lms), cut = len / 5;
  LocalMotion m = null_localmotion();
  int* xs = (int*)vs_malloc(sizeof(int) * (size_t)(len > 0 ? len : 1));
  int* ys = (int*)vs_malloc(sizeof(int) * (size_t)(len > 0 ? len : 1));
  for (int i = 0; i < len; i++) {
    xs[i] = LMGet(lms, i)->v.x;
    ys[i] = LMGet(lms, i)->v.y;
  }
  qsort(xs, (size_t)len, sizeof(int), cmp_int);
  qsort(ys, (size_t)len, sizeof(int), cmp_int);
  int16_t sx = 0, sy = 0;   /* the reference accumulates in the int16 field */
  for (int i = cut; i < len - cut; i++) sx = (int16_t)(sx + xs[i]);
  for (int i = cut; i < len - cut; i++) sy = (int16_t)(sy + ys[i]);
  const double d = (double)len - ((double)cut + (double)cut);
  m.v.x = (int16_t)(int)((double)sx / d);
  m.v.y = (int16_t)(int)((double)sy / d);
  vs_free(xs);
  vs_free(ys);
  return m;
}

VS_API VSArray localmotionsGetMatch(const LocalMotions* lms) {
  VSArray m = vs_array_new(vs_vector_size(lms));
  for (int i = 0; i < m.len; i++) m.dat[i] = LMGet(lms, i)->match;
  return m;
}

VS_API double vsCalcAngle(const LocalMotion* lm, int center_x, int center_y) {
  const int dx = lm->f.x - center_x, dy = lm->f.y - center_y;
  if (abs(dx) + abs(dy) < lm->f.size * 2) return 0;
  const double a1 = atan2((double)dy, (double)dx);
  const double a2 = atan2((double)(dy + lm->v.y), (double)(dx + lm->v.x));
  const double diff = a2 - a1;
  return (diff > M_PI) ? diff - 2 * M_PI : ((diff < -M_PI) ? diff + 2 * M_PI : diff);
}

VS_API VSTransform vsSimpleMotionsToTransform(VSFrameInfo fi, const char* modName,
                                              const LocalMotions* motions) {
  VSTransform t = null_transform();
  if (!motions) return t;
  const int n = vs_vector_size(motions);
  if (n < 1) return t;
  int cx = 0, cy = 0;
  for (int i = 0; i < n; i++) {
    cx += LMGet(motions, i)->f.x;
    cy += LMGet(motions, i)->f.y;
  }
  cx /= n;
  cy /= n;
  const LocalMotion mean = cleanmean_localmotions(motions);
  double c1 = 0.0, s = 0.0;   /* cos(alpha) - 1, sin(alpha) */
  if (n >= 6) {
    double* angles = (double*)vs_malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; i++) {
      LocalMotion m = sub_localmotion(LMGet(motions, i), &mean);
      angles[i] = vsCalcAngle(&m, cx, cy);
    }
    double mn, mx;
    t.alpha = -cleanmean(angles, n, &mn, &mx);
    vs_free(angles);
    if (mx - mn > 1.0) {
      t.alpha = 0;
      vs_log_info(modName, "too large variation in angle(%f)\n", mx - mn);
    } else {
      double sn, cs;
      sincos(t.alpha, &sn, &cs);
      c1 = cs - 1.0;
      s = sn;
    }
  }
  const double px = (double)(cx - fi.width / 2), py = (double)(cy - fi.height / 2);
  t.x = (c1 * px - s * py) + (double)mean.v.x;
  t.y = ((double)mean.v.y + s * px) + c1 * py;
  return t;
}
