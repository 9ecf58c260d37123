/*
 * transform.h -- the types of vid.stab's transform pass that the second-pass
 * entry point of this library reads: vsLocalmotions2Transforms takes the
 * transform pass's VSTransformData, which ffmpeg's vidstabtransform filter
 * fills with vsTransformDataInit (the reference library's) and hands over.
 *
 * Layouts are those of the reference's src/transform.h:44-178 and
 * src/lensmap.h:52-110 in a release build (no TESTING members), so one
 * VSTransformData can pass between the two libraries.  This library reads
 * fiSrc and conf and writes lensK (what vsTransformSetLensK sets,
 * src/transform.c:166-168); everything else belongs to the transform pass,
 * which this library does not implement.
 */
#ifndef __TRANSFORM_H
#define __TRANSFORM_H

#include <stdint.h>

#include "frameinfo.h"
#include "transformtype.h"
#include "vidstab_api.h"

/* one transform per frame; src/transform.h:44-49 */
typedef struct _vstransformations {
  VSTransform* ts;     /* vs_malloc'd array of len transforms */
  int current;         /* index of the transform in use */
  int len;
  short warned_end;
} VSTransformations;

/* src/transform.h:60, 65-66 */
typedef enum { VS_Zero, VS_Linear, VS_BiLinear, VS_BiCubic, VS_NBInterPolTypes } VSInterpolType;
typedef enum { VSKeepBorder = 0, VSCropBorder } VSBorderType;
typedef enum { VSOptimalL1 = 0, VSGaussian, VSAvg } VSCamPathAlgo;

/* how the output is corrected for the estimated lens; src/lensmap.h:52-56 */
typedef enum { VSLensCorrectOff = 0, VSLensCorrectWobble, VSLensCorrectFull } VSLensCorrectMode;

/* per-pixel sampler of the transform pass; src/transform.h:83-85 */
typedef void (*vsInterpolateFun)(uint8_t* rv, int32_t x, int32_t y, const uint8_t* img, int linesize, int width,
                                 int height, uint8_t def);

/* one plane's lens lookup tables (transform pass state); src/lensmap.h:81-110 */
typedef struct _vslensplanemap {
  int active;
  VSLensCorrectMode mode;
  double k;
  double cdx, cdy;
  double csx, csy;
  double tMaxU, tMaxD;
  double tDomD;
  double invRho2;
  double f;
  int sxShift;
  int syShift;
  int32_t idxScaleU;
  int32_t idxScaleD;
  int32_t* gU;
  int32_t* gD;
} VSLensPlaneMap;

/* src/transform.h:85-147 */
typedef struct _VSTransformConfig {
  int relative;
  int smoothing;
  VSBorderType crop;
  int invert;
  double zoom;
  int optZoom;
  double zoomSpeed;
  VSInterpolType interpolType;
  int maxShift;
  double maxAngle;
  const char* modName;
  int verbose;
  int simpleMotionCalculation;   /* 1: vsSimpleMotionsToTransform per frame */
  int storeTransforms;           /* 1: also write global_motions.trf */
  int smoothZoom;                /* 0: the fitted zoom is dropped */
  VSCamPathAlgo camPathAlgo;
  int estimateLensDistortion;    /* 1: estimate one lens strength k for the clip */
  VSLensCorrectMode lensCorrection;
  double lensK;                  /* user's k; 0.0: use the estimate */
  double fov;                    /* horizontal field of view in degrees; 0: similarity model */
} VSTransformConfig;

/* src/transform.h:149-178 (release build) */
typedef struct _VSTransformData {
  VSFrameInfo fiSrc;
  VSFrameInfo fiDest;
  VSFrame src;
  VSFrame destbuf;
  VSFrame dest;
  short srcMalloced;
  vsInterpolateFun interpolate;
  VSTransformConfig conf;
  VSLensCorrectMode lensMode;
  int lensActive;
  double lensK;
  double lensMapK;
  VSLensPlaneMap lensMaps[3];
  int initialized;
} VSTransformData;

#endif /* __TRANSFORM_H */
