/*
 * vs_lensdistortion.c -- the clip-wide lens estimate of the second pass
 * (reference: src/lensdistortion.c:367-387, 455-693) over the GPU's
 * per-frame fits (vsg_lens_* in vsg.h).
 *
 * The estimate minimises E(k), the mean squared image-space residual of
 * every frame's camera motion refitted through a lens of strength k, by
 * Brent's method over [kMin, kMax]; between searches it masks outlier
 * correspondences (median / MAD of each frame's residuals at the current k)
 * and searches again; finally it takes E's curvature at the minimum by a
 * central difference, and the standard error of k from it.  Every E(k) is one
 * GPU launch over all frames (one workgroup per frame, vsg_lens_fit); the
 * host only adds the frames' r^2 n in frame order, as the reference's loop
 * does, and steps the search.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "lensdistortion.h"
#include "transformtype_operations.h"
#include "vidstabdefines.h"
#include "vs_detect_internal.h"
#include "vsg.h"

/* src/lensdistortion.c:393, 401, 404 */
#define LENS_PENALTY 1e12
#define LENS_CURVATURE_FLOOR 1e-9
#define LENS_UNDETERMINED_SIGMA 1e6

VS_API VSLensEstimateConfig vsLensEstimateGetDefaultConfig(void) {
  VSLensEstimateConfig c;
  c.kMin = -0.6;
  c.kMax = 0.2;
  c.tolerance = 1e-6;
  c.maxIterations = 100;
  c.gaussNewtonSteps = 3;
  c.maxUncertainty = 0.02;
  c.rejectOutliers = 1;
  c.outlierStddevs = 2.5;
  c.outlierPasses = 3;
  c.f = 0.0;
  return c;
}

/* the profile objective E(k) over the clip, on the GPU */
typedef struct {
  VSGLens* gpu;
  int nframes;
  const int* count;
  int gn;
  double f;
  int useMask;
  int evals;
  int failed;   /* a GPU call failed: the estimate is abandoned */
  double* r;
  int* status;
} Objective;

/* lensObjective, src/lensdistortion.c:455-478 */
static double objective(Objective* o, double k) {
  o->evals++;
  if (o->failed) return LENS_PENALTY;
  if (vsg_lens_fit(o->gpu, k, o->f, o->gn, o->useMask, o->r, o->status) != 0) {
    o->failed = 1;
    return LENS_PENALTY;
  }
  double sum2 = 0;
  long long total = 0;
  for (int i = 0; i < o->nframes; i++) {
    if (o->count[i] < 3) continue;   /* cannot pin a similarity down */
    if (o->status[i] != 0) return LENS_PENALTY;
    sum2 += o->r[i] * o->r[i] * o->count[i];
    total += o->count[i];
  }
  if (total < 1) return LENS_PENALTY;
  sum2 /= (double)total;
  if (!(sum2 == sum2) || sum2 > LENS_PENALTY) return LENS_PENALTY;
  return sum2;
}

/* Brent's minimisation of E on [a, b] (golden section with parabolic steps
   where the parabola behaves), as src/lensdistortion.c:480-526 steps it */
static double brent(double a, double b, double tol, int maxIter, Objective* o, int* usedIter) {
  const double gold = 0.3819660112501051;   /* (3 - sqrt(5)) / 2 */
  double x = a + gold * (b - a), w = x, v = x;
  double fx = objective(o, x), fw = fx, fv = fx;
  double e = 0, d = 0;
  int it;
  for (it = 0; it < maxIter; it++) {
    const double m = 0.5 * (a + b);
    const double tol1 = tol * fabs(x) + 1e-10, tol2 = 2.0 * tol1;
    if (fabs(x - m) <= tol2 - 0.5 * (b - a)) break;
    double p = 0, q = 0, r = 0;
    if (fabs(e) > tol1) {
      r = (x - w) * (fx - fv);
      q = (x - v) * (fx - fw);
      p = (x - v) * q - (x - w) * r;
      q = 2.0 * (q - r);
      if (q > 0)
        p = -p;
      else
        q = -q;
      r = e;
      e = d;
    }
    double u;
    if (fabs(p) < fabs(0.5 * q * r) && p > q * (a - x) && p < q * (b - x)) {
      d = p / q;   /* the parabola's step */
      u = x + d;
      if (u - a < tol2 || b - u < tol2) d = (x < m) ? tol1 : -tol1;
    } else {
      e = (x < m) ? b - x : a - x;   /* golden section */
      d = gold * e;
    }
    u = (fabs(d) >= tol1) ? x + d : x + ((d > 0) ? tol1 : -tol1);
    const double fu = objective(o, u);
    if (fu <= fx) {
      if (u < x)
        b = x;
      else
        a = x;
      v = w;
      fv = fw;
      w = x;
      fw = fx;
      x = u;
      fx = fu;
    } else {
      if (u < x)
        a = u;
      else
        b = u;
      if (fu <= fw || w == x) {
        v = w;
        fv = fw;
        w = u;
        fw = fu;
      } else if (fu <= fv || v == x || v == w) {
        v = u;
        fv = fu;
      }
    }
  }
  if (usedIter) *usedIter = it;
  return x;
}

/* vsEstimateLensDistortionFromMatches (src/lensdistortion.c:528-639) over
   the clip held by `gpu`; *failed is set when a GPU call failed */
VSLensEstimate vsd_lens_estimate(VSGLens* gpu, int nframes, const int* count, const VSLensEstimateConfig* cfg,
                                 int* failed) {
  VSLensEstimate est;
  memset(&est, 0, sizeof(est));
  *failed = 0;
  long long nPoints = 0;
  for (int i = 0; i < nframes; i++) nPoints += count[i];
  if (nPoints < 3) return est;
  est.used = (int)nPoints;
  Objective o;
  memset(&o, 0, sizeof(o));
  o.gpu = gpu;
  o.nframes = nframes;
  o.count = count;
  o.gn = cfg->gaussNewtonSteps;
  o.f = cfg->f;
  o.r = (double*)malloc(sizeof(double) * (size_t)nframes);
  o.status = (int*)malloc(sizeof(int) * (size_t)nframes);
  int* cut = (int*)malloc(sizeof(int) * (size_t)nframes);
  if (!o.r || !o.status || !cut) {
    free(o.r);
    free(o.status);
    free(cut);
    *failed = 1;
    return est;
  }
  int iters = 0;
  est.k = brent(cfg->kMin, cfg->kMax, cfg->tolerance, cfg->maxIterations, &o, &iters);
  /* reject and refit: the mask stays fixed through each search and changes
     only between searches, at the current k */
  if (cfg->rejectOutliers && cfg->outlierPasses > 1 && !o.failed) {
    if (vsg_lens_reset_mask(gpu) != 0) o.failed = 1;
    for (int pass = 1; pass < cfg->outlierPasses && !o.failed; pass++) {
      if (vsg_lens_mask(gpu, est.k, cfg->gaussNewtonSteps, cfg->outlierStddevs, cut) != 0) {
        o.failed = 1;
        break;
      }
      long long c = 0;
      for (int i = 0; i < nframes; i++) c += cut[i];
      if (c == 0) break;   /* nothing new to remove */
      o.useMask = 1;
      est.k = brent(cfg->kMin, cfg->kMax, cfg->tolerance, cfg->maxIterations, &o, &iters);
    }
    o.useMask = 1;   /* the reported residual describes the final fit */
    long long rejected = 0;
    if (!o.failed && vsg_lens_rejected(gpu, &rejected) != 0) o.failed = 1;
    est.rejected = (int)rejected;
    est.used = (int)(nPoints - rejected);
  }
  /* curvature of the profile at the minimum, by central difference */
  const double h = 1e-3;
  const double e0 = objective(&o, est.k), ep = objective(&o, est.k + h), em = objective(&o, est.k - h);
  if (ep >= LENS_PENALTY || em >= LENS_PENALTY)
    est.curvature = 0;
  else
    est.curvature = (ep - 2.0 * e0 + em) / (h * h);
  est.residual = sqrt(e0 >= LENS_PENALTY ? 0 : e0);
  est.iterations = o.evals;
  if (est.curvature > LENS_CURVATURE_FLOOR && est.used > 0)
    est.uncertainty = est.residual / sqrt((double)est.used * est.curvature);
  else
    est.uncertainty = LENS_UNDETERMINED_SIGMA;
  /* a minimum on the bracket's edge was never bracketed */
  const double edge = 1e-6 * (cfg->kMax - cfg->kMin) + cfg->tolerance;
  const int pinned = (est.k - cfg->kMin < edge) || (cfg->kMax - est.k < edge);
  est.determined = !pinned && (est.uncertainty < cfg->maxUncertainty);
  *failed = o.failed;
  if (o.failed) est.determined = 0;
  free(o.r);
  free(o.status);
  free(cut);
  return est;
}

/* a clip's LocalMotions as the GPU takes them: 4 int16 per motion (field x,
   y, motion x, y), the match values, the count per frame; NULL on failure */
int vsd_flatten_motions(const VSManyLocalMotions* motions, int16_t** lm, double** match, int** count) {
  const int n = vs_vector_size(motions);
  long long total = 0;
  *lm = NULL;
  *match = NULL;
  *count = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  if (!*count) return VS_ERROR;
  for (int f = 0; f < n; f++) {
    (*count)[f] = vs_vector_size(VSMLMGet(motions, f));
    total += (*count)[f];
  }
  const size_t nm = total > 0 ? (size_t)total : 1;
  *lm = (int16_t*)malloc(sizeof(int16_t) * 4 * nm);
  *match = (double*)malloc(sizeof(double) * nm);
  if (!*lm || !*match) {
    free(*lm);
    free(*match);
    free(*count);
    *lm = NULL;
    *match = NULL;
    *count = NULL;
    return VS_ERROR;
  }
  size_t k = 0;
  for (int f = 0; f < n; f++) {
    const LocalMotions* lms = VSMLMGet(motions, f);
    for (int i = 0; i < (*count)[f]; i++, k++) {
      const LocalMotion* m = LMGet(lms, i);
      (*lm)[4 * k] = m->f.x;
      (*lm)[4 * k + 1] = m->f.y;
      (*lm)[4 * k + 2] = m->v.x;
      (*lm)[4 * k + 3] = m->v.y;
      (*match)[k] = m->match;
    }
  }
  return VS_OK;
}

int vsd_device(void) {
  const char* v = getenv("VIDSTAB_AMD_DEVICE");
  return v ? atoi(v) : 0;
}

/* src/lensdistortion.c:641-693 */
VS_API VSLensEstimate vsEstimateLensDistortion(const VSFrameInfo* fi, const VSManyLocalMotions* motions,
                                               const VSLensEstimateConfig* cfg) {
  VSLensEstimate est;
  memset(&est, 0, sizeof(est));
  if (!fi || !motions) return est;
  const int n = vs_vector_size(motions);
  if (n < 1) return est;
  VSLensEstimateConfig defcfg = vsLensEstimateGetDefaultConfig();
  if (!cfg) cfg = &defcfg;
  int16_t* lm;
  double* match;
  int* count;
  if (vsd_flatten_motions(motions, &lm, &match, &count) != VS_OK) return est;
  long long total = 0;
  for (int f = 0; f < n; f++) total += count[f];
  if (total >= 1) {
    VSGLens* gpu = NULL;
    int failed = 1;
    if (vsg_lens_create(&gpu, vsd_device(), fi->width, fi->height, lm, match, count, n) == 0)
      est = vsd_lens_estimate(gpu, n, count, cfg, &failed);
    if (failed) vs_log_error("vidstab_amd", "lens distortion estimate on the GPU failed: %s\n", vsg_last_error());
    vsg_lens_destroy(gpu);
  }
  free(lm);
  free(match);
  free(count);
  return est;
}
