/*
 * vs_draw.c -- the show = 1 / 2 overlays of the detect pass
 * (reference: src/motiondetect.c:271-286 for the sequence, :816-952 for the
 * primitives).  They are drawn on the host into the caller's frame once the
 * frame's LocalMotions are final, as the reference draws into md->currorig.
 *
 * Every colour is a grey level written to plane 0: the luma of a planar
 * frame, or each of the first three bytes of a packed pixel (a fourth byte,
 * alpha, is never touched).  Runs are clipped to the frame.
 */
#include <math.h>
#include <stdlib.h>

#include "vs_detect_internal.h"

typedef struct {
  uint8_t* img;
  int stride;   /* bytes per row */
  int w, h;     /* pixels */
  int bpp;      /* bytes per pixel of the surface (1: planar luma) */
} Surface;

static void put(const Surface* s, int x, int y, uint8_t c) {
  uint8_t* p = s->img + (size_t)y * s->stride + (size_t)x * s->bpp;
  p[0] = c;
  if (s->bpp >= 3) {
    p[1] = c;
    p[2] = c;
  }
}

/* src/motiondetect.c:868-879 */
static void hline(const Surface* s, int x, int y, int len, uint8_t c) {
  if (y < 0 || y >= s->h || len <= 0) return;
  int x0 = x < 0 ? 0 : x;
  int x1 = x + len > s->w ? s->w : x + len;
  for (int i = x0; i < x1; i++) put(s, i, y, c);
}

/* src/motiondetect.c:885-895 */
static void vline(const Surface* s, int x, int y, int len, uint8_t c) {
  if (x < 0 || x >= s->w || len <= 0) return;
  int y0 = y < 0 ? 0 : y;
  int y1 = y + len > s->h ? s->h : y + len;
  for (int j = y0; j < y1; j++) put(s, x, j, c);
}

/* src/motiondetect.c:901-906: filled, centred at (x, y) */
static void box(const Surface* s, int x, int y, int sx, int sy, uint8_t c) {
  for (int j = 0; j < sy; j++) hline(s, x - sx / 2, y - sy / 2 + j, sx, c);
}

/* src/motiondetect.c:912-922: outline, centred at (x, y) */
static void rectangle(const Surface* s, int x, int y, int sx, int sy, uint8_t c) {
  hline(s, x - sx / 2, y - sy / 2, sx, c);
  hline(s, x - sx / 2, y + sy / 2, sx, c);
  vline(s, x - sx / 2, y - sy / 2, sy, c);
  vline(s, x + sx / 2, y - sy / 2, sy, c);
}

/* src/motiondetect.c:928-952.  The start point is the motion's field position;
   for an axis-parallel line pointing left or up the reference swaps the ends
   by storing the end point into it, which the returned LocalMotion keeps.
   A slanted line is drawn row by row, horlen + 1 pixels starting at
   x = (int)((a.x - horlen/2) + m*dy): the order in which the reference's
   -ffast-math build evaluates a.x + m*dy - horlen/2. */
static void line_from_field(const Surface* s, Field* a, int bx, int by, int thick, uint8_t c) {
  int dx = bx - a->x, dy = by - a->y;
  if (dy == 0) {
    if (dx < 0) {
      a->x = (int16_t)bx;
      a->y = (int16_t)by;
      dx = -dx;
    }
    for (int r = -thick / 2; r <= thick / 2; r++) hline(s, a->x, a->y + r, dx + 1, c);
    return;
  }
  if (dx == 0) {
    if (dy < 0) {
      a->x = (int16_t)bx;
      a->y = (int16_t)by;
      dy = -dy;
    }
    for (int r = -thick / 2; r <= thick / 2; r++) vline(s, a->x + r, a->y, dy + 1, c);
    return;
  }
  const double m = (double)dx / (double)dy;
  const int horlen = (int)((double)thick + fabs(m));
  const double x0 = (double)(a->x - horlen / 2);
  const int n = abs(dy);
  for (int k = 0; k <= n; k++) {
    const int d = dy < 0 ? -k : k;
    hline(s, (int)(x0 + m * (double)d), a->y + d, horlen + 1, c);
  }
}

/* src/motiondetect.c:271-286 */
void vsd_draw_overlays(const VSMotionDetect* md, uint8_t* img, int linesize, LocalMotions* coarse,
                       LocalMotions* fine) {
  if (!img || md->conf.show <= 0) return;
  Surface s;
  s.img = img;
  s.stride = linesize;
  s.w = md->fi.width;
  s.h = md->fi.height;
  s.bpp = md->fi.pFormat > PF_PACKED ? md->fi.bytesPerPixel : 1;
  const int nc = vs_vector_size(coarse), nf = vs_vector_size(fine);
  const int scan = md->fieldscoarse.maxShift;
  if (md->conf.show > 1)
    for (int i = 0; i < nc; i++) {
      const LocalMotion* m = (const LocalMotion*)vs_vector_get(coarse, i);
      rectangle(&s, m->f.x, m->f.y, m->f.size + 2 * scan, m->f.size + 2 * scan, 80);
    }
  for (int i = 0; i < nc; i++) {
    const LocalMotion* m = (const LocalMotion*)vs_vector_get(coarse, i);
    box(&s, m->f.x, m->f.y, m->f.size, m->f.size, 40);
  }
  for (int i = 0; i < nf; i++) {
    const LocalMotion* m = (const LocalMotion*)vs_vector_get(fine, i);
    rectangle(&s, m->f.x, m->f.y, m->f.size, m->f.size, 40);
  }
  for (int pass = 0; pass < 2; pass++) {
    LocalMotions* v = pass ? fine : coarse;
    const int n = pass ? nf : nc;
    const uint8_t color = pass ? 64 : 180;
    for (int i = 0; i < n; i++) {
      LocalMotion* m = (LocalMotion*)vs_vector_get(v, i);
      const int ex = m->f.x + m->v.x, ey = m->f.y + m->v.y;
      box(&s, m->f.x, m->f.y, 5, 5, 0);
      box(&s, ex, ey, 5, 5, 250);
      line_from_field(&s, &m->f, ex, ey, 3, color);
    }
  }
}
