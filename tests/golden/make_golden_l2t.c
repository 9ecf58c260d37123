/*
 * make_golden_l2t.c -- fixtures of the second pass's LocalMotions -> transforms
 * step, written by the REFERENCE (oracle/_ref/libvsref.so, built from
 * /root/reference/src with its release flags).  Run only where
 * /root/reference exists (make_golden.sh l2t).
 *
 * Each case is a seeded clip of synthetic LocalMotions: a grid of fields
 * whose motions follow a camera move (similarity, or rotation about the
 * optical centre for a given field of view) seen through a radial lens
 * (division model, strength k), rounded to whole pixels, with a fraction of
 * outlier fields (random motions, worse match values).  For every case the
 * file records the motions themselves, then the reference's
 *   - vsLocalmotions2Transforms with the case's transform config (fov,
 *     estimateLensDistortion, simpleMotionCalculation, smoothZoom): every
 *     frame's transform (exact bits) and the lens strength it left in the
 *     transform data (td->lensK);
 *   - vsEstimateLensDistortion with the configuration vsLocalmotions2Transforms
 *     uses (defaults, f from the fov): k, residual, curvature, uncertainty,
 *     evaluations, determined, used, rejected;
 *   - vsMotionsToTransform of every frame with the case's fov (no lens).
 * Match values are multiples of 1/8 so the JSON numbers are exact.
 */
#define _GNU_SOURCE 1
#include <inttypes.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lensdistortion.h"
#include "transformtype_operations.h"
#include "localmotion2transform.h"
#include "transform.h"

static uint64_t dbits(double d) {
  uint64_t u;
  memcpy(&u, &d, 8);
  return u;
}

static uint64_t rng_next(uint64_t* s) {   /* splitmix64 */
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double urand(uint64_t* s) { return (double)(rng_next(s) >> 11) * (1.0 / 9007199254740992.0); }
static double grand(uint64_t* s) {
  double u = urand(s), v = urand(s);
  if (u < 1e-300) u = 1e-300;
  return sqrt(-2.0 * log(u)) * cos(6.283185307179586 * v);
}

typedef struct {
  const char* name;
  int width, height, frames, spacing;
  double k;          /* lens strength of the synthetic footage */
  double fovDeg;     /* camera model of the footage and of the fit (0: similarity) */
  double outliers;   /* fraction of outlier fields */
  int estimate;      /* VSTransformConfig.estimateLensDistortion */
  int simple;        /* VSTransformConfig.simpleMotionCalculation */
  int smoothZoom;
  uint64_t seed;
} L2TCase;

/* forward camera map on a centred, undistorted point: rotation about the
   optical centre at focal length f (yaw = tx/f, pitch = ty/f, roll alpha,
   zoom z), or the similarity for f = 0 */
static void camera(double f, double tx, double ty, double alpha, double z, double ux, double uy, double* wx,
                   double* wy) {
  if (f <= 0) {
    *wx = z * (cos(alpha) * ux + sin(alpha) * uy) + tx;
    *wy = z * (-sin(alpha) * ux + cos(alpha) * uy) + ty;
    return;
  }
  const double yaw = tx / f, pitch = ty / f;
  /* R = Ry(yaw) Rx(pitch) Rz(alpha) applied to (ux, uy, f) */
  double X = cos(alpha) * ux + sin(alpha) * uy, Y = -sin(alpha) * ux + cos(alpha) * uy, Z = f;
  double Y2 = cos(pitch) * Y - sin(pitch) * Z, Z2 = sin(pitch) * Y + cos(pitch) * Z;
  double X3 = cos(yaw) * X + sin(yaw) * Z2, Z3 = -sin(yaw) * X + cos(yaw) * Z2;
  *wx = z * f * X3 / Z3;
  *wy = z * f * Y2 / Z3;
}

static void synth_frame(const L2TCase* cs, double f, uint64_t* s, double tx, double ty, double alpha, double zoom,
                        LocalMotions* out) {
  const double cx = cs->width / 2.0, cy = cs->height / 2.0;
  const double rho = 0.5 * sqrt((double)cs->width * cs->width + (double)cs->height * cs->height);
  vs_vector_init(out, 0);
  for (int gy = cs->spacing; gy < cs->height - cs->spacing / 2; gy += cs->spacing)
    for (int gx = cs->spacing; gx < cs->width - cs->spacing / 2; gx += cs->spacing) {
      const double dx = gx - cx, dy = gy - cy;
      const double r2 = (dx * dx + dy * dy) / (rho * rho);
      const double ux = dx / (1.0 + cs->k * r2), uy = dy / (1.0 + cs->k * r2);
      double wx, wy;
      camera(f, tx, ty, alpha, 1.0 + zoom / 100.0, ux, uy, &wx, &wy);
      const double t = (wx * wx + wy * wy) / (rho * rho);
      const double disc = 1.0 - 4.0 * cs->k * t;
      const double g = disc > 0 ? 2.0 / (1.0 + sqrt(disc)) : 1.0;
      const double qx = cx + wx * g, qy = cy + wy * g;
      LocalMotion lm;
      memset(&lm, 0, sizeof(lm));
      lm.f.x = (short)gx;
      lm.f.y = (short)gy;
      lm.f.size = 32;
      lm.contrast = 0.5;
      if (urand(s) < cs->outliers) {
        lm.v.x = (short)((int)(rng_next(s) % 41) - 20);
        lm.v.y = (short)((int)(rng_next(s) % 41) - 20);
        lm.match = (double)(80 + (int)(rng_next(s) % 400)) / 8.0;
      } else {
        lm.v.x = (short)lround(qx - gx + 0.35 * grand(s));
        lm.v.y = (short)lround(qy - gy + 0.35 * grand(s));
        lm.match = (double)(8 + (int)(rng_next(s) % 120)) / 8.0;
      }
      vs_vector_append_dup(out, &lm, sizeof(lm));
    }
}

static void run_case(FILE* out, const L2TCase* cs, int first) {
  uint64_t s = cs->seed;
  VSFrameInfo fi;
  vsFrameInfoInit(&fi, cs->width, cs->height, PF_YUV420P);
  const double f = focal_from_fov(cs->fovDeg, cs->width);
  VSManyLocalMotions mlms;
  vs_vector_init(&mlms, cs->frames);
  double tx = 0, ty = 0, alpha = 0, zoom = 0;
  for (int t = 0; t < cs->frames; t++) {
    tx = 0.8 * tx + 3.0 * grand(&s);
    ty = 0.8 * ty + 3.0 * grand(&s);
    alpha = 0.8 * alpha + 0.004 * grand(&s);
    zoom = 0.8 * zoom + 0.3 * grand(&s);
    LocalMotions lms;
    if (t == 3) {
      vs_vector_init(&lms, 0);   /* a frame without motions */
    } else {
      synth_frame(cs, f, &s, tx, ty, alpha, zoom, &lms);
    }
    vs_vector_append_dup(&mlms, &lms, sizeof(LocalMotions));
  }
  fprintf(out, "%s{\"name\":\"%s\",\"width\":%d,\"height\":%d,\"fov\":%.17g,\"k_true\":%.17g,"
          "\"estimate\":%d,\"simple\":%d,\"smoothZoom\":%d,\"frames\":[",
          first ? "" : ",\n", cs->name, cs->width, cs->height, cs->fovDeg, cs->k, cs->estimate, cs->simple,
          cs->smoothZoom);
  for (int t = 0; t < cs->frames; t++) {
    const LocalMotions* lms = VSMLMGet(&mlms, t);
    fprintf(out, "%s[", t ? "," : "");
    for (int i = 0; i < vs_vector_size(lms); i++) {
      const LocalMotion* m = LMGet(lms, i);
      fprintf(out, "%s%d,%d,%d,%d,%.17g", i ? "," : "", m->f.x, m->f.y, m->v.x, m->v.y, m->match);
    }
    fprintf(out, "]");
  }
  fprintf(out, "]");

  /* vsLocalmotions2Transforms with the case's configuration */
  VSTransformData td;
  memset(&td, 0, sizeof(td));
  VSTransformConfig conf = vsTransformGetDefaultConfig("golden");
  conf.fov = cs->fovDeg;
  conf.estimateLensDistortion = cs->estimate;
  conf.simpleMotionCalculation = cs->simple;
  conf.smoothZoom = cs->smoothZoom;
  conf.storeTransforms = 0;
  td.conf = conf;
  td.fiSrc = fi;
  td.fiDest = fi;
  VSTransformations trans;
  vsTransformationsInit(&trans);
  if (vsLocalmotions2Transforms(&td, &mlms, &trans) != VS_OK) {
    fprintf(stderr, "%s: vsLocalmotions2Transforms failed\n", cs->name);
    exit(1);
  }
  fprintf(out, ",\"lensK\":\"%016" PRIx64 "\",\"transforms\":[", dbits(td.lensK));
  for (int t = 0; t < trans.len; t++) {
    const VSTransform* tr = &trans.ts[t];
    fprintf(out, "%s[\"%016" PRIx64 "\",\"%016" PRIx64 "\",\"%016" PRIx64 "\",\"%016" PRIx64 "\",\"%016" PRIx64
            "\",%d]",
            t ? "," : "", dbits(tr->x), dbits(tr->y), dbits(tr->alpha), dbits(tr->zoom), dbits(tr->barrel),
            tr->extra);
  }
  fprintf(out, "]");
  vsTransformationsCleanup(&trans);

  /* the estimate alone, as vsLocalmotions2Transforms configures it */
  VSLensEstimateConfig lcfg = vsLensEstimateGetDefaultConfig();
  lcfg.f = f;
  VSLensEstimate le = vsEstimateLensDistortion(&fi, &mlms, &lcfg);
  fprintf(out, ",\"estimate_result\":{\"k\":\"%016" PRIx64 "\",\"residual\":\"%016" PRIx64 "\",\"curvature\":\"%016"
          PRIx64 "\",\"uncertainty\":\"%016" PRIx64 "\",\"iterations\":%d,\"determined\":%d,\"used\":%d,"
          "\"rejected\":%d}",
          dbits(le.k), dbits(le.residual), dbits(le.curvature), dbits(le.uncertainty), le.iterations, le.determined,
          le.used, le.rejected);

  /* vsMotionsToTransform of every frame with the case's fov, no lens */
  fprintf(out, ",\"motions_to_transform\":[");
  for (int t = 0; t < cs->frames; t++) {
    VSTransform tr = vsMotionsToTransform(&td, VSMLMGet(&mlms, t), NULL);
    fprintf(out, "%s[\"%016" PRIx64 "\",\"%016" PRIx64 "\",\"%016" PRIx64 "\",\"%016" PRIx64 "\",%d]", t ? "," : "",
            dbits(tr.x), dbits(tr.y), dbits(tr.alpha), dbits(tr.zoom), tr.extra);
  }
  fprintf(out, "]}");
  for (int t = 0; t < cs->frames; t++) vs_vector_del(VSMLMGet(&mlms, t));
  vs_vector_del(&mlms);
  fprintf(stderr, "l2t case %s: k %.5f +- %.5f (%s), lensK %.5f\n", cs->name, le.k, le.uncertainty,
          le.determined ? "determined" : "not determined", td.lensK);
}

int main(int argc, char** argv) {
  const char* path = argc > 1 ? argv[1] : "l2t_golden.json";
  vs_log_level = 0;
  static const L2TCase cases[] = {
    /* name             w     h    frames spacing k      fov   outl  est simple sz  seed */
    { "fov0_plain",     1280, 720, 24,    64,     0.0,   0.0,  0.08, 0,  0,     1,  101 },
    { "fov65",          1280, 720, 24,    64,     0.0,   65.0, 0.08, 0,  0,     1,  102 },
    { "fov90_noise",    960,  540, 20,    48,     0.0,   90.0, 0.15, 0,  0,     0,  103 },
    { "lens_barrel",    1280, 720, 24,    64,     -0.15, 0.0,  0.08, 1,  0,     1,  104 },
    { "lens_fov70",     1280, 720, 24,    64,     -0.10, 70.0, 0.08, 1,  0,     1,  105 },
    { "lens_none",      1280, 720, 24,    64,     0.0,   0.0,  0.08, 1,  0,     1,  106 },
    { "lens_pincushion",960,  540, 20,    48,     0.08,  0.0,  0.05, 1,  0,     0,  107 },
    { "simple",         1280, 720, 16,    64,     0.0,   0.0,  0.08, 1,  1,     1,  108 },
    /* 1,947 fields a frame, near the 2,048 a frame may hold */
    { "lens_dense_4k",  3840, 2160, 6,    64,     -0.2,  0.0,  0.06, 1,  0,     1,  109 },
  };
  FILE* out = fopen(path, "w");
  if (!out) return 1;
  fprintf(out, "{\"source\":\"reference vsLocalmotions2Transforms / vsEstimateLensDistortion / vsMotionsToTransform "
               "(oracle/_ref/libvsref.so) on synthetic LocalMotions\",\"cases\":[\n");
  for (size_t i = 0; i < sizeof(cases) / sizeof(cases[0]); i++) run_case(out, &cases[i], i == 0);
  fprintf(out, "\n]}\n");
  fclose(out);
  return 0;
}
