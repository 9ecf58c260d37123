/*
 * vsynth.c -- seeded synthetic clip generator (see vsynth.h).
 *
 * Determinism contract: integer arithmetic plus IEEE basic double operations
 * only, compiled with -ffp-contract=off and without -ffast-math.  No libm
 * transcendental is used (sine/cosine below are a range-reduced Taylor
 * series), so frames are byte-identical on every x86-64 host.
 */
#include "vsynth.h"

#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static const double kPi = 3.14159265358979323846;

/* ---- counter-based random numbers (splitmix64 finaliser) ---------------- */

static inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

typedef struct { uint64_t s; } Rng;

static inline uint64_t rng_next(Rng* r) {
  r->s += 0x9E3779B97F4A7C15ull;
  return mix64(r->s);
}

/* uniform in [0,1) with 53 random bits */
static inline double rng_unit(Rng* r) {
  return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0);
}

static inline double rng_uniform(Rng* r, double lo, double hi) {
  return lo + (hi - lo) * rng_unit(r);
}

/* approximately standard normal: Irwin-Hall sum of 12 uniforms minus 6 */
static double rng_gauss(Rng* r) {
  double s = 0.0;
  for (int i = 0; i < 12; ++i) s += rng_unit(r);
  return s - 6.0;
}

/* ---- libm-free sine / cosine --------------------------------------------- */

static double taylor_sin(double x) {  /* |x| <= pi/2 */
  double x2 = x * x, term = x, sum = x;
  for (int k = 1; k <= 11; ++k) {
    term = -term * x2 / (double)((2 * k) * (2 * k + 1));
    sum += term;
  }
  return sum;
}

double vsyn_sin(double x) {
  /* reduce to [-pi, pi] */
  while (x > kPi) x -= 2.0 * kPi;
  while (x < -kPi) x += 2.0 * kPi;
  /* reflect into [-pi/2, pi/2] */
  if (x > 0.5 * kPi) x = kPi - x;
  else if (x < -0.5 * kPi) x = -kPi - x;
  return taylor_sin(x);
}

double vsyn_cos(double x) { return vsyn_sin(x + 0.5 * kPi); }

/* ---- value noise ---------------------------------------------------------- */

/* 16-bit lattice value for (octave, ix, iy) */
static inline uint32_t lattice(uint64_t seed, int oct, int ix, int iy) {
  uint64_t k = seed ^ ((uint64_t)(uint32_t)oct * 0xD6E8FEB86659FD93ull)
                    ^ ((uint64_t)(uint32_t)ix * 0x9E3779B97F4A7C15ull)
                    ^ ((uint64_t)(uint32_t)iy * 0xC2B2AE3D27D4EB4Full);
  return (uint32_t)(mix64(k) >> 48);
}

/* smoothstep in Q16: s = f^2 (3 - 2f) */
static inline uint32_t smooth_q16(uint32_t f) {
  uint64_t f2 = ((uint64_t)f * f) >> 16;
  return (uint32_t)((f2 * (3u * 65536u - 2u * f)) >> 16);
}

/* sum over 4 octaves (cells 64,32,16,8 px; amplitudes 8,4,2,1) -> 0..15*65535 */
static uint32_t noise_total(uint64_t seed, int x, int y) {
  uint32_t total = 0;
  for (int o = 0; o < 4; ++o) {
    int cell = 64 >> o;
    int ix = x / cell, iy = y / cell;            /* x, y >= 0 */
    uint32_t fx = (uint32_t)(((x - ix * cell) << 16) / cell);
    uint32_t fy = (uint32_t)(((y - iy * cell) << 16) / cell);
    uint32_t sx = smooth_q16(fx), sy = smooth_q16(fy);
    uint32_t v00 = lattice(seed, o, ix, iy), v10 = lattice(seed, o, ix + 1, iy);
    uint32_t v01 = lattice(seed, o, ix, iy + 1), v11 = lattice(seed, o, ix + 1, iy + 1);
    uint64_t top = (uint64_t)v00 * (65536u - sx) + (uint64_t)v10 * sx;  /* Q32 */
    uint64_t bot = (uint64_t)v01 * (65536u - sx) + (uint64_t)v11 * sx;
    uint64_t v = ((top >> 16) * (65536u - sy) + (bot >> 16) * sy) >> 16;
    total += (uint32_t)v << (3 - o);
  }
  return total;
}

void vsyn_value_noise(uint8_t* out, int w, int h, int stride, uint64_t seed) {
  size_t n = (size_t)w * (size_t)h;
  uint32_t* t = (uint32_t*)malloc(n * sizeof(uint32_t));
  if (!t) return;
  int y;
#pragma omp parallel for schedule(static)
  for (y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) t[(size_t)y * w + x] = noise_total(seed, x, y);
  /* exact integer mean and second moment (sequential, order-independent) */
  uint64_t s1 = 0;
  for (size_t i = 0; i < n; ++i) s1 += t[i];
  double mean = (double)s1 / (double)n;
  double var = 0.0;
  for (size_t i = 0; i < n; ++i) {
    double d = (double)t[i] - mean;
    var += d * d;
  }
  var /= (double)n;
  double sd = __builtin_sqrt(var);
  double scale = sd > 0.0 ? 40.0 / sd : 0.0;
#pragma omp parallel for schedule(static)
  for (y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      double v = 128.0 + ((double)t[(size_t)y * w + x] - mean) * scale;
      int q = v <= 0.0 ? 0 : (v >= 255.0 ? 255 : (int)(v + 0.5));
      out[(size_t)y * stride + x] = (uint8_t)q;
    }
  free(t);
}

/* ---- clip ----------------------------------------------------------------- */

int vsyn_clip_init(VSynClip* c, int width, int height, uint64_t seed, int margin) {
  memset(c, 0, sizeof(*c));
  c->width = width;
  c->height = height;
  c->margin = margin;
  c->seed = seed;
  c->noise_sigma_q8 = 512;
  int mw = width + 2 * margin, mh = height + 2 * margin;
  c->master = (uint8_t*)malloc((size_t)mw * mh);
  if (!c->master) return -1;
  vsyn_value_noise(c->master, mw, mh, mw, seed);
  return 0;
}

void vsyn_clip_free(VSynClip* c) {
  free(c->master);
  c->master = NULL;
  for (int i = 0; i < c->nrects; ++i) {
    free(c->rects[i].tex);
    c->rects[i].tex = NULL;
  }
  c->nrects = 0;
}

static double reflect(double v, double lim) {
  /* fold v into [-lim, lim] */
  for (int i = 0; i < 64 && (v > lim || v < -lim); ++i) {
    if (v > lim) v = 2.0 * lim - v;
    if (v < -lim) v = -2.0 * lim - v;
  }
  return v;
}

void vsyn_motion_path(VSynMotion* out, int nframes, int mode, double tpar, double apar_deg,
                      int margin, uint64_t seed) {
  Rng r = { seed ^ 0x5A17C0DEull };
  const double deg = kPi / 180.0;
  double tlim = (double)margin * 0.6, alim = 4.0 * deg;
  double x = 0.0, y = 0.0, a = 0.0;
  for (int t = 0; t < nframes; ++t) {
    if (mode == 0) {
      x = rng_uniform(&r, -tpar, tpar);
      y = rng_uniform(&r, -tpar, tpar);
      a = rng_uniform(&r, -apar_deg, apar_deg) * deg;
    } else if (t > 0) {
      x = reflect(x + tpar * rng_gauss(&r), tlim);
      y = reflect(y + tpar * rng_gauss(&r), tlim);
      a = reflect(a + apar_deg * deg * rng_gauss(&r), alim);
    }
    out[t].tx = x;
    out[t].ty = y;
    out[t].alpha = a;
  }
}

void vsyn_add_rects(VSynClip* c, uint64_t seed, int nmin, int nmax, int smin, int smax, double vmax) {
  Rng r = { seed ^ 0x3EC7A11Eull };
  int n = nmin + (int)(rng_next(&r) % (uint64_t)(nmax - nmin + 1));
  if (n > 8) n = 8;
  c->nrects = n;
  for (int i = 0; i < n; ++i) {
    VSynRect* q = &c->rects[i];
    q->w = smin + (int)(rng_next(&r) % (uint64_t)(smax - smin + 1));
    q->h = smin + (int)(rng_next(&r) % (uint64_t)(smax - smin + 1));
    if (q->w > c->width) q->w = c->width;
    if (q->h > c->height) q->h = c->height;
    q->x0 = rng_uniform(&r, 0.0, (double)(c->width - q->w));
    q->y0 = rng_uniform(&r, 0.0, (double)(c->height - q->h));
    q->vx = rng_uniform(&r, -vmax, vmax);
    q->vy = rng_uniform(&r, -vmax, vmax);
    q->seed = rng_next(&r);
    q->tex = (uint8_t*)malloc((size_t)q->w * q->h);
    if (q->tex) vsyn_value_noise(q->tex, q->w, q->h, q->w, q->seed);
  }
}

/* position on [0, L] of a point moving with constant speed, bouncing at the ends */
static int bounce(double p, int L) {
  if (L <= 0) return 0;
  long period = 2L * L;
  long q = (long)(p < 0.0 ? -p : p);   /* |p|, truncated */
  q %= period;
  if (q > L) q = period - q;
  return (int)q;
}

static inline uint8_t sample_bilinear(const uint8_t* m, int mw, int mh, double u, double v) {
  /* floor for possibly negative coordinates without libm */
  long iu = (long)u; if ((double)iu > u) iu -= 1;
  long iv = (long)v; if ((double)iv > v) iv -= 1;
  int fu = (int)((u - (double)iu) * 256.0);
  int fv = (int)((v - (double)iv) * 256.0);
  long u0 = iu < 0 ? 0 : (iu >= mw ? mw - 1 : iu);
  long u1 = iu + 1 < 0 ? 0 : (iu + 1 >= mw ? mw - 1 : iu + 1);
  long v0 = iv < 0 ? 0 : (iv >= mh ? mh - 1 : iv);
  long v1 = iv + 1 < 0 ? 0 : (iv + 1 >= mh ? mh - 1 : iv + 1);
  uint32_t a = m[v0 * mw + u0], b = m[v0 * mw + u1];
  uint32_t cc = m[v1 * mw + u0], d = m[v1 * mw + u1];
  uint32_t top = a * (256u - fu) + b * fu;
  uint32_t bot = cc * (256u - fu) + d * fu;
  return (uint8_t)((top * (256u - fv) + bot * fv + 32768u) >> 16);
}

void vsyn_render(const VSynClip* c, const VSynMotion* mo, int t, uint8_t* dst, int stride) {
  const int W = c->width, H = c->height, M = c->margin;
  const int mw = W + 2 * M, mh = H + 2 * M;
  const VSynMotion m = mo[t];
  const double ca = vsyn_cos(m.alpha), sa = vsyn_sin(m.alpha);
  const double cx = (double)(W / 2), cy = (double)(H / 2);
  /* motion blur along this frame's motion relative to the previous one */
  int blen = 1;
  double bdx = 0.0, bdy = 0.0;
  if (c->blur_len_max > 1) {
    Rng br = { c->seed ^ (0xB1B1ull + (uint64_t)t * 0x10001ull) };
    blen = 1 + (int)(rng_next(&br) % (uint64_t)c->blur_len_max);
    double dx = t > 0 ? m.tx - mo[t - 1].tx : 1.0;
    double dy = t > 0 ? m.ty - mo[t - 1].ty : 0.0;
    double nrm = __builtin_sqrt(dx * dx + dy * dy);
    if (nrm > 1e-9) { bdx = dx / nrm; bdy = dy / nrm; } else { bdx = 1.0; bdy = 0.0; }
  }
  int y;
#pragma omp parallel for schedule(static)
  for (y = 0; y < H; ++y) {
    uint8_t* row = dst + (size_t)y * stride;
    for (int x = 0; x < W; ++x) {
      /* source point: inverse of (rotate about centre, then translate) */
      double px = (double)x - cx - m.tx, py = (double)y - cy - m.ty;
      double u = ca * px + sa * py + cx + (double)M;
      double v = -sa * px + ca * py + cy + (double)M;
      uint32_t acc;
      if (blen > 1) {
        uint32_t s = 0;
        for (int k = 0; k < blen; ++k) {
          double o = (double)k - 0.5 * (double)(blen - 1);
          s += sample_bilinear(c->master, mw, mh, u + o * bdx, v + o * bdy);
        }
        acc = (s * 2u + (uint32_t)blen) / (2u * (uint32_t)blen);
      } else {
        acc = sample_bilinear(c->master, mw, mh, u, v);
      }
      row[x] = (uint8_t)acc;
    }
  }
  /* moving rectangles, each with its own texture */
  for (int i = 0; i < c->nrects; ++i) {
    const VSynRect* q = &c->rects[i];
    int rx = bounce(q->x0 + q->vx * (double)t, W - q->w);
    int ry = bounce(q->y0 + q->vy * (double)t, H - q->h);
    if (!q->tex) continue;
    for (int yy = 0; yy < q->h; ++yy)
      memcpy(dst + (size_t)(ry + yy) * stride + rx, q->tex + (size_t)yy * q->w, (size_t)q->w);
  }
  /* additive noise: per-pixel counter-based Irwin-Hall(4) scaled to sigma */
  if (c->noise_sigma_q8 > 0) {
    const int64_t sig = c->noise_sigma_q8;
#pragma omp parallel for schedule(static)
    for (y = 0; y < H; ++y) {
      uint8_t* row = dst + (size_t)y * stride;
      for (int x = 0; x < W; ++x) {
        uint64_t rr = mix64(c->seed ^ ((uint64_t)t << 40) ^ ((uint64_t)y << 20) ^ (uint64_t)x
                            ^ 0x6E015Eull);
        int64_t s = (int64_t)(rr & 0xFFFF) + (int64_t)((rr >> 16) & 0xFFFF)
                  + (int64_t)((rr >> 32) & 0xFFFF) + (int64_t)(rr >> 48) - 131070;
        /* Irwin-Hall(4) has std sqrt(1/3); 443 = sqrt(3)*256 */
        int64_t nq8 = (s * 443 * sig) >> 24;
        int64_t val = ((int64_t)row[x] << 8) + nq8 + 128;
        val >>= 8;
        row[x] = (uint8_t)(val < 0 ? 0 : (val > 255 ? 255 : val));
      }
    }
  }
}

void vsyn_pack(const uint8_t* luma, int w, int h, int lstride, int bpp, uint8_t* dst, int dstride) {
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const uint8_t v = luma[(size_t)y * lstride + x];
      uint8_t* d = dst + (size_t)y * dstride + (size_t)x * bpp;
      d[0] = v;
      d[1] = (uint8_t)(7 * v);
      d[2] = (uint8_t)(255 - v);
      if (bpp > 3) d[3] = (uint8_t)(v ^ 0x5A);
    }
}

uint64_t vsyn_hash_plane(const uint8_t* p, int w, int h, int stride) {
  uint64_t hsh = 0xCBF29CE484222325ull;
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      hsh ^= p[(size_t)y * stride + x];
      hsh *= 0x100000001B3ull;
    }
  return hsh;
}

/* ---- workload specifications ---------------------------------------------- */

int vsyn_spec(int config, int clip, VSynSpec* s) {
  memset(s, 0, sizeof(*s));
  s->config = config;
  s->clip = clip;
  s->accuracy = 15;
  s->stepsize = 6;
  s->shakiness = 5;
  s->mincontrast = 0.25;
  s->seed = 1000ull * (uint64_t)config + (uint64_t)clip;
  switch (config) {
    case 1:
      s->width = 640; s->height = 360; s->nframes = 60; s->mincontrast = 0.3;
      break;
    case 2:
      s->width = 1920; s->height = 1080; s->nframes = 300; s->shakiness = 8;
      break;
    case 3:
      s->width = 3840; s->height = 2160; s->nframes = 240; s->shakiness = 10; s->stepsize = 4;
      break;
    case 4: {
      static const int res[5][2] = { {640, 360}, {1280, 720}, {1920, 1080}, {2560, 1440}, {3840, 2160} };
      Rng r = { s->seed ^ 0xBA7C4ull };
      int k = (int)(rng_next(&r) % 5u);
      s->width = res[k][0]; s->height = res[k][1];
      s->nframes = 30 + (int)(rng_next(&r) % 91u);
      s->shakiness = 4 + (int)(rng_next(&r) % 7u);
      break;
    }
    case 5:
      s->width = 3840; s->height = 2160; s->nframes = 240; s->shakiness = 8;
      break;
    default:
      return -1;
  }
  return 0;
}

int vsyn_open(const VSynSpec* s, VSynClip* c, VSynMotion* path) {
  int margin = s->config == 1 ? 32 : (s->height >= 2000 ? 384 : 256);
  if (vsyn_clip_init(c, s->width, s->height, s->seed, margin) != 0) return -1;
  if (s->config == 1)
    vsyn_motion_path(path, s->nframes, 0, 8.0, 1.0, margin, s->seed);
  else
    vsyn_motion_path(path, s->nframes, 1, 4.0, 0.5, margin, s->seed);
  if (s->config == 3) vsyn_add_rects(c, s->seed, 3, 6, 64, 512, 12.0);
  if (s->config == 5) c->blur_len_max = 9;
  return 0;
}
