/*
 * lensdistortion.h -- the clip-wide radial lens estimate of the second pass
 * (reference: src/lensdistortion.h; single-parameter division model
 * U(x) = x / (1 + k r^2), r in units of half the frame diagonal, barrel for
 * k < 0).  This library estimates k from a clip's LocalMotions on the GPU
 * (vsg_lens_* in vsg.h): every evaluation of the profile objective fits the
 * camera motion of all frames through the lens at once.
 */
#ifndef __LENSDISTORTION_H
#define __LENSDISTORTION_H

#include "frameinfo.h"
#include "serialize.h"
#include "transformtype.h"
#include "vidstab_api.h"

/* src/lensdistortion.h:39-44 */
typedef struct _vslensdistortion {
  double k;
  double cx;
  double cy;
  double rho;
} VSLensDistortion;

/* search settings; src/lensdistortion.h:117-133, defaults src/lensdistortion.c:367-387 */
typedef struct _vslensestimateconfig {
  double kMin;             /* -0.6 */
  double kMax;             /* 0.2 */
  double tolerance;        /* 1e-6, absolute on k */
  int maxIterations;       /* 100 */
  int gaussNewtonSteps;    /* 3 */
  double maxUncertainty;   /* 0.02 */
  int rejectOutliers;      /* 1 */
  double outlierStddevs;   /* 2.5 */
  int outlierPasses;       /* 3 */
  double f;                /* focal length in px for the rotational model, 0: similarity */
} VSLensEstimateConfig;

/* src/lensdistortion.h:136-144 */
typedef struct _vslensestimate {
  double k;
  double residual;      /* RMS image-space residual per correspondence at the minimum, px */
  double curvature;     /* d2E/dk2 there */
  double uncertainty;   /* standard error of k */
  int iterations;       /* objective evaluations */
  int rejected;         /* correspondences masked as outliers */
  int used;             /* correspondences of the final fit */
  int determined;       /* 0: k not identifiable (flat objective, or pinned to the bracket) */
} VSLensEstimate;

/* src/lensdistortion.h:146, src/lensdistortion.c:367-387 */
VS_API VSLensEstimateConfig vsLensEstimateGetDefaultConfig(void);

/* the lens strength of a clip from its LocalMotions (fields -> field +
   motion), on the GPU; src/lensdistortion.h:169-175, src/lensdistortion.c:641-693.
   Agrees with the reference to within the search's tolerance: the per-frame
   fits sum their points in another order. */
VS_API VSLensEstimate vsEstimateLensDistortion(const VSFrameInfo* fi, const VSManyLocalMotions* motions,
                                               const VSLensEstimateConfig* cfg);

#endif /* __LENSDISTORTION_H */
