/*
 * vsynth.h -- seeded synthetic clip generator for the detect pass.
 *
 * Produces the luma planes of the benchmark/test clips: 4-octave value noise
 * (mean 128, std 40, clamped), warped per frame by a global translation and
 * rotation with bilinear sampling, optional independently moving textured
 * rectangles, optional box motion blur, plus approximately Gaussian noise.
 *
 * Everything is integer arithmetic or IEEE basic double operations (+ - * /
 * and sqrt, all correctly rounded); no libm transcendental is called and the
 * file must be compiled without -ffast-math and with -ffp-contract=off.  The
 * same seed therefore gives byte-identical frames on every x86-64 host.
 */
#ifndef VSYNTH_H
#define VSYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* per-frame global motion: translation in px, rotation in radians */
typedef struct {
  double tx, ty, alpha;
} VSynMotion;

/* one moving rectangle (textured with its own value-noise seed) */
typedef struct {
  int w, h;
  double x0, y0;      /* top-left at frame 0 */
  double vx, vy;      /* px per frame */
  uint64_t seed;
  uint8_t* tex;       /* w x h texture, owned by the clip */
} VSynRect;

typedef struct {
  int width, height;
  int margin;          /* master texture border around the frame */
  uint64_t seed;
  int noise_sigma_q8;  /* Gaussian noise sigma in 1/256 px (512 = sigma 2) */
  int nrects;
  VSynRect rects[8];
  int blur_len_max;    /* >0: box motion blur of 1..blur_len_max px along motion */
  uint8_t* master;     /* (width+2*margin) x (height+2*margin) */
} VSynClip;

/* 4-octave value noise, normalised to mean 128 / std 40, clamped to 0..255 */
void vsyn_value_noise(uint8_t* out, int w, int h, int stride, uint64_t seed);

/* allocate and fill the master texture; returns 0 on success */
int vsyn_clip_init(VSynClip* c, int width, int height, uint64_t seed, int margin);
void vsyn_clip_free(VSynClip* c);

/* motion models.  mode 0: i.i.d. translation U(+-tmax) and rotation
   U(+-amax_deg) about the master; mode 1: random walk with N(0, tstep) and
   N(0, astep_deg) increments (reflected to stay inside the margin). */
void vsyn_motion_path(VSynMotion* out, int nframes, int mode, double tpar, double apar_deg,
                      int margin, uint64_t seed);

/* 3..6 rectangles of 64..512 px with velocity U(+-12) px/frame */
void vsyn_add_rects(VSynClip* c, uint64_t seed, int nmin, int nmax, int smin, int smax, double vmax);

/* render frame t into dst (luma, stride >= width) */
void vsyn_render(const VSynClip* c, const VSynMotion* m, int t, uint8_t* dst, int stride);

/* Clip specifications of the detect workloads:
     1: 640x360 x60, i.i.d. motion U(+-8 px) / U(+-1 deg), step 6, mincontrast 0.3
     2: 1920x1080 x300, random walk N(0,4 px) / N(0,0.5 deg), shakiness 8, step 6
     3: 3840x2160 x240, walk + 3..6 moving rectangles, shakiness 10, step 4
     4: batched clip `clip`: resolution U{360p,720p,1080p,1440p,2160p},
        30..120 frames, shakiness U[4,10], walk motion, step 6
     5: 3840x2160 x240, walk + box motion blur 1..9 px, step 6
   accuracy is 15 everywhere; noise sigma is 2 px. */
typedef struct {
  int config, clip;
  int width, height, nframes;
  int shakiness, accuracy, stepsize;
  double mincontrast;
  uint64_t seed;
} VSynSpec;

int vsyn_spec(int config, int clip, VSynSpec* s);
/* build the clip and its motion path (path must hold s->nframes entries) */
int vsyn_open(const VSynSpec* s, VSynClip* c, VSynMotion* path);

/* a packed frame (bpp 3 or 4 bytes a pixel) from a rendered luma plane, one
   distinct per-pixel map per byte: Y, (7*Y) mod 256, 255 - Y, Y xor 0x5A.
   Every channel moves with the content but has its own min/max. */
void vsyn_pack(const uint8_t* luma, int w, int h, int lstride, int bpp, uint8_t* dst, int dstride);

/* FNV-1a over a plane (row by row), for fixtures */
uint64_t vsyn_hash_plane(const uint8_t* p, int w, int h, int stride);

/* libm-free sine/cosine (range-reduced Taylor series), exposed for tests */
double vsyn_sin(double x);
double vsyn_cos(double x);

#ifdef __cplusplus
}
#endif
#endif /* VSYNTH_H */
