/*
 * transformtype.h -- measurement field, vector, local motion and transform
 * types of the vid.stab ABI.
 *
 * Layouts are fixed by the reference (src/transformtype.h:35-71) because the
 * caller reads LocalMotion elements directly and the .trf writer serializes
 * them field by field.  LocalMotion is 32 bytes: v@0, f@4, 6 padding bytes,
 * contrast@16, match@24.  Compare LocalMotions fieldwise, never with memcmp.
 */
#ifndef __TRANSFORMTYPE_H
#define __TRANSFORMTYPE_H

#include <stdio.h>
#include <stdint.h>
#include "vsvector.h"
#include "vidstab_api.h"

/* global frame-to-frame transform, 56 bytes */
typedef struct _transform {
  double x;
  double y;
  double alpha;     /* rotation about the frame centre, radians */
  double zoom;      /* percent */
  double barrel;
  double rshutter;
  int extra;        /* -1 ignore, 0 normal, 1 scene cut */
} VSTransform;

/* a square measurement field: centre and edge length, 6 bytes */
typedef struct _field {
  int16_t x;
  int16_t y;
  int16_t size;
} Field;

/* integer displacement, 4 bytes */
typedef struct _vec {
  int16_t x;
  int16_t y;
} Vec;

/* one field's measured displacement, 32 bytes */
typedef struct _localmotion {
  Vec v;
  Field f;
  double contrast;  /* selection contrast of the field */
  double match;     /* mean absolute difference at the best shift */
} LocalMotion;

typedef VSVector LocalMotions;

#endif /* __TRANSFORMTYPE_H */
