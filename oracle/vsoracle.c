/*
 * vsoracle.c -- CPU restatement of vid.stab's detect pass (vsMotionDetection)
 * for the planar formats and the packed RGB24 / BGR24 / RGBA formats.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU library.  Nothing in the
 * product links this file.
 *
 * Each function cites the reference code it restates.  Where the reference's
 * release build (-O3 -msse2 -ffast-math) evaluates a double expression in a
 * different order than its C source, the order below is the one the compiled
 * reference executes (read from the x86 code of the reference built from its
 * own sources with its own flags); this file is compiled with
 * -ffp-contract=off and without -ffast-math so that order is kept.
 */
#define _GNU_SOURCE 1   /* sincos */
#include "vsoracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_MAX(a, b) ((a) > (b) ? (a) : (b))
#define OR_MIN(a, b) ((a) < (b) ? (a) : (b))

/* ---- geometry: src/motiondetect.c:314-358 (initFields) ------------------- */

static int or_init_fields(OrFields* fs, int w, int h, int accuracy, int size, int maxShift,
                          int stepSize, int keepBorder, int spacing, double ct) {
  fs->fieldSize = size;
  fs->maxShift = maxShift;
  fs->stepSize = stepSize;
  fs->contrastThreshold = ct;
  int rows = OR_MAX(3, (h - maxShift * 2) / (size + spacing) - 1);
  int cols = OR_MAX(3, (w - maxShift * 2) / (size + spacing) - 1);
  fs->fieldNum = rows * cols;
  fs->fieldRows = rows;
  fs->fields = (Field*)malloc(sizeof(Field) * (size_t)fs->fieldNum);
  if (!fs->fields) return 0;
  int border = keepBorder ? size / 2 + maxShift + stepSize : stepSize;
  int step_x = (w - 2 * border) / OR_MAX(cols - 1, 1);
  int step_y = (h - 2 * border) / OR_MAX(rows - 1, 1);
  for (int j = 0; j < rows; j++)
    for (int i = 0; i < cols; i++) {
      Field* f = &fs->fields[j * cols + i];
      f->x = (int16_t)(border + i * step_x);
      f->y = (int16_t)(border + j * step_y);
      f->size = (int16_t)size;
    }
  fs->maxFields = accuracy * fs->fieldNum / 15;
  return 1;
}

/* src/motiondetect.c:83-165 (vsMotionDetectInit), planar formats */
int or_init(OrState* s, int width, int height, int shakiness, int accuracy, int stepSize,
            double contrastThreshold, int virtualTripod) {
  memset(s, 0, sizeof(*s));
  s->width = width;
  s->height = height;
  s->virtualTripod = virtualTripod;
  s->contrastThreshold = contrastThreshold;
  s->shakiness = OR_MIN(10, OR_MAX(1, shakiness));
  s->accuracy = OR_MIN(15, OR_MAX(1, accuracy));
  if (s->accuracy < s->shakiness / 2) s->accuracy = s->shakiness / 2;
  if (s->accuracy > 9 && stepSize > 6) stepSize = 6;
  s->stepSize = stepSize;
  int minDim = OR_MIN(width, height);
  int maxShift = OR_MAX(16, minDim / 7);
  int fieldSize = OR_MAX(16, minDim / 10);
  int fieldSizeFine = OR_MAX(6, minDim / 60);
  fieldSize = (fieldSize / 16 + 1) * 16;          /* always rounds up */
  fieldSizeFine = (fieldSizeFine / 16 + 1) * 16;
  if (!or_init_fields(&s->coarse, width, height, s->accuracy, fieldSize, maxShift, stepSize, 1, 0,
                      contrastThreshold))
    return -1;
  if (!or_init_fields(&s->fine, width, height, s->accuracy, fieldSizeFine, fieldSizeFine, 2, 1,
                      fieldSizeFine, contrastThreshold / 2))
    return -1;
  s->blurSize = or_blur_size(stepSize, width, height);
  s->bpp = 1;
  size_t n = (size_t)width * height;
  s->curr = (uint8_t*)calloc(n, 1);
  s->prev = (uint8_t*)calloc(n, 1);
  s->tmp = (uint8_t*)calloc(n, 1);
  return (s->curr && s->prev && s->tmp) ? 0 : -1;
}

/* src/motiondetect.c:83-165 for PF_RGB24 / PF_BGR24 / PF_RGBA: the same field
   geometry, no blur (:197-201), planes of width*bpp bytes a row */
int or_init_packed(OrState* s, int width, int height, int bpp, int shakiness, int accuracy, int stepSize,
                   double contrastThreshold, int virtualTripod) {
  if (bpp != 3 && bpp != 4) return -1;
  if (or_init(s, width, height, shakiness, accuracy, stepSize, contrastThreshold, virtualTripod) != 0) return -1;
  s->bpp = bpp;
  s->blurSize = 1;
  size_t n = (size_t)width * height * bpp;
  free(s->curr);
  free(s->prev);
  free(s->tmp);
  s->curr = (uint8_t*)calloc(n, 1);
  s->prev = (uint8_t*)calloc(n, 1);
  s->tmp = (uint8_t*)calloc(n, 1);
  return (s->curr && s->prev && s->tmp) ? 0 : -1;
}

void or_free(OrState* s) {
  free(s->coarse.fields);
  free(s->fine.fields);
  free(s->curr);
  free(s->prev);
  free(s->tmp);
  memset(s, 0, sizeof(*s));
}

/* ---- box blur: src/boxblur.c:104-161, 190-332 ------------------------------
   Closed form of the two running-sum passes: each pass is
   floor(sum_{k=-s2..s2} in[clamp(i+k)] / size) along its axis, with an 8-bit
   result between the passes.  Returns 0/1 for "plain copy". */
int or_blur_size(int stepSize, int width, int height) {
  if (stepSize < 2) return 1;
  int size = (stepSize / 2) * 2 + 1;
  size = OR_MIN(OR_MAX(3, size), OR_MIN(height / 2, width / 2));
  return size;
}

void or_boxblur(uint8_t* dst, const uint8_t* src, int w, int h, int sstride, int size) {
  if (size < 2) {
    for (int y = 0; y < h; y++) memcpy(dst + (size_t)y * w, src + (size_t)y * sstride, (size_t)w);
    return;
  }
  int s2 = size / 2;
  uint8_t* hor = (uint8_t*)malloc((size_t)w * h);
  if (!hor) return;
  int y;
#pragma omp parallel for schedule(static)
  for (y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      unsigned acc = 0;
      for (int k = -s2; k <= s2; k++) acc += src[(size_t)y * sstride + OR_MIN(OR_MAX(x + k, 0), w - 1)];
      hor[(size_t)y * w + x] = (uint8_t)(acc / (unsigned)size);
    }
#pragma omp parallel for schedule(static)
  for (y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      unsigned acc = 0;
      for (int k = -s2; k <= s2; k++) acc += hor[(size_t)OR_MIN(OR_MAX(y + k, 0), h - 1) * w + x];
      dst[(size_t)y * w + x] = (uint8_t)(acc / (unsigned)size);
    }
  free(hor);
}

/* ---- contrast: src/motiondetect.c:389-407 --------------------------------- */

void or_minmax(const uint8_t* img, int stride, int x, int y, int size, int* mn, int* mx) {
  int s2 = size / 2, lo = 255, hi = 0;
  for (int j = 0; j < size; j++) {
    const uint8_t* p = img + (size_t)(y - s2 + j) * stride + (x - s2);
    for (int k = 0; k < size; k++) {
      lo = p[k] < lo ? p[k] : lo;
      hi = p[k] > hi ? p[k] : hi;
    }
  }
  *mn = lo;
  *mx = hi;
}

double or_contrast(int mn, int mx) { return (double)(mx - mn) / ((double)(mx + mn) + 0.1); }

/* ---- SAD: src/motiondetect.c:971-996 --------------------------------------
   sum |cur(field) - prev(field + d)|; stops after a row once the sum exceeds
   thr (the caller only needs to know it is not below its current minimum). */
static uint32_t or_sad_thr(const uint8_t* cur, const uint8_t* prev, int stride, int fx, int fy,
                           int size, int dx, int dy, uint32_t thr) {
  int s2 = size / 2;
  const uint8_t* p1 = cur + (size_t)(fy - s2) * stride + (fx - s2);
  const uint8_t* p2 = prev + (size_t)(fy - s2 + dy) * stride + (fx - s2 + dx);
  uint32_t sum = 0;
  for (int j = 0; j < size; j++) {
    for (int k = 0; k < size; k++) sum += (uint32_t)abs((int)p1[k] - (int)p2[k]);
    if (sum > thr) break;
    p1 += stride;
    p2 += stride;
  }
  return sum;
}

uint32_t or_sad(const uint8_t* cur, const uint8_t* prev, int stride, int fx, int fy, int size,
                int dx, int dy) {
  return or_sad_thr(cur, prev, stride, fx, fy, size, dx, dy, UINT_MAX);
}

/* ---- spiral order: src/motiondetect.c:466-518 ----------------------------- */

/* visit order of the outgoing spiral in grid units, simulated step by step */
void or_spiral_walk(int n, int* as, int* bs) {
  int i = 0, j = 0, limit = 1, step = 0, dir = 0, k = 0;
  while (j >= -n && j <= n && i >= -n && i <= n) {
    as[k] = i;
    bs[k] = j;
    k++;
    step++;
    switch (dir) {
      case 0: i++; if (step == limit) { dir = 1; step = 0; } break;
      case 1: j++; if (step == limit) { dir = 2; step = 0; limit++; } break;
      case 2: i--; if (step == limit) { dir = 3; step = 0; } break;
      case 3: j--; if (step == limit) { dir = 0; step = 0; limit++; } break;
    }
  }
}

/* closed form of the same order: ring k = max(|a|,|b|) starts at (2k-1)^2 */
int or_spiral_rank(int a, int b) {
  int k = OR_MAX(abs(a), abs(b));
  if (k == 0) return 0;
  int base = (2 * k - 1) * (2 * k - 1);
  if (a == k && b > -k) return base + (b + k - 1);      /* right column, going down */
  if (b == k) return base + 2 * k - 1 + (k - a);         /* bottom row, going left */
  if (a == -k) return base + 4 * k - 1 + (k - b);        /* left column, going up */
  return base + 6 * k - 1 + (a + k);                     /* top row, going right */
}

/* ---- selection: src/motiondetect.c:665-743 --------------------------------
   glibc's qsort (2.35) is a stable merge sort, so equal contrasts keep their
   index order; this stable merge sort reproduces that. */
typedef struct { double c; int i; } OrCI;

static void or_stable_sort_desc(OrCI* a, int n) {
  if (n < 2) return;
  OrCI* t = (OrCI*)malloc(sizeof(OrCI) * (size_t)n);
  for (int width = 1; width < n; width *= 2) {
    for (int lo = 0; lo < n; lo += 2 * width) {
      int mid = OR_MIN(lo + width, n), hi = OR_MIN(lo + 2 * width, n);
      int p = lo, q = mid, o = lo;
      while (p < mid && q < hi) t[o++] = (a[q].c > a[p].c) ? a[q++] : a[p++];
      while (p < mid) t[o++] = a[p++];
      while (q < hi) t[o++] = a[q++];
    }
    memcpy(a, t, sizeof(OrCI) * (size_t)n);
  }
  free(t);
}

int or_select(const OrFields* fs, const double* contrast, int* idx, double* sel_contrast) {
  int n = fs->fieldNum, cnt = 0;
  OrCI* ci = (OrCI*)malloc(sizeof(OrCI) * (size_t)n);
  OrCI* seg = (OrCI*)malloc(sizeof(OrCI) * (size_t)n);
  for (int i = 0; i < n; i++) {
    ci[i].c = contrast[i] < fs->contrastThreshold ? 0.0 : contrast[i];
    ci[i].i = i;
  }
  memcpy(seg, ci, sizeof(OrCI) * (size_t)n);
  int numsegms = fs->fieldRows + 1;
  int segmlen = n / (fs->fieldRows + 1) + 1;
  for (int s = 0; s < numsegms; s++) {
    int start = segmlen * s, end = OR_MIN(segmlen * (s + 1), n);
    if (end > start) or_stable_sort_desc(seg + start, end - start);
    for (int j = 0; j < fs->maxFields / numsegms; j++) {
      if (start + j >= end) continue;
      if (seg[start + j].c > 0) {
        idx[cnt] = seg[start + j].i;
        sel_contrast[cnt] = ci[seg[start + j].i].c;
        cnt++;
        seg[start + j].c = 0;
      }
    }
  }
  int remaining = fs->maxFields - cnt;
  if (remaining > 0) {
    or_stable_sort_desc(seg, n);
    for (int j = 0; j < remaining && j < n; j++)
      if (seg[j].c > 0) {
        idx[cnt] = seg[j].i;
        sel_contrast[cnt] = seg[j].c;
        cnt++;
      }
  }
  free(ci);
  free(seg);
  return cnt;
}

/* ---- one field: src/motiondetect.c:439-586 (calcFieldTransPlanar) -------- */

int or_field_search(const uint8_t* cur, const uint8_t* prev, int w, int h, int stride,
                    const OrFields* fs, const Field* f, int ox, int oy,
                    int* tx_out, int* ty_out, uint32_t* minerror_out) {
  (void)w; (void)h;
  int stepSize = fs->stepSize, maxShift = fs->maxShift;
  int n = maxShift / stepSize;
  int g = 2 * n + 1;
  int* as = (int*)malloc(sizeof(int) * (size_t)g * g);
  int* bs = (int*)malloc(sizeof(int) * (size_t)g * g);
  or_spiral_walk(n, as, bs);
  uint32_t minerror = UINT_MAX;
  int tx = 0, ty = 0;
  for (int q = 0; q < g * g; q++) {
    int i = as[q] * stepSize, j = bs[q] * stepSize;
    uint32_t e = or_sad_thr(cur, prev, stride, f->x, f->y, f->size, i + ox, j + oy, minerror);
    if (e < minerror) { minerror = e; tx = i; ty = j; }
  }
  free(as);
  free(bs);
  while (stepSize > 1) {
    int txc = tx, tyc = ty;
    int ns = stepSize / 2, r = stepSize - ns;
    for (int i = txc - r; i <= txc + r; i += ns)
      for (int j = tyc - r; j <= tyc + r; j += ns) {
        if (i == txc && j == tyc) continue;
        uint32_t e = or_sad_thr(cur, prev, stride, f->x, f->y, f->size, i + ox, j + oy, minerror);
        if (e < minerror) { minerror = e; tx = i; ty = j; }
      }
    stepSize /= 2;
  }
  *tx_out = tx;
  *ty_out = ty;
  *minerror_out = minerror;
  if (abs(tx) >= maxShift + stepSize - 1 || abs(ty) >= maxShift + stepSize) return 0;
  return 1;
}

/* ---- packed frames: src/motiondetect.c:370-377, 389-407, 591-660, 971-996 -- */

/* mean Michelson contrast of the three colour channels (a fourth byte is left
   out).  The reference's -ffast-math build computes the mean as
   ((c0 + c1) + c2) * (1/3): a multiplication by the rounded reciprocal, not a
   division by 3 (read from its x86 code built here from its own sources). */
double or_contrast_packed(const uint8_t* img, int stride, int bpp, int x, int y, int size) {
  double c[3];
  int s2 = size / 2;
  for (int ch = 0; ch < 3; ch++) {
    int lo = 255, hi = 0;
    for (int j = 0; j < size; j++) {
      const uint8_t* p = img + (size_t)(y - s2 + j) * stride + (size_t)(x - s2) * bpp + ch;
      for (int k = 0; k < size; k++) {
        int v = p[(size_t)k * bpp];
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
      }
    }
    c[ch] = or_contrast(lo, hi);
  }
  return ((c[0] + c[1]) + c[2]) * (1.0 / 3.0);
}

/* SAD over size rows of size*bpp bytes; the column is scaled by bpp, the row
   offset is the byte stride */
static uint32_t or_sad_packed_thr(const uint8_t* cur, const uint8_t* prev, int stride, int bpp, int fx,
                                  int fy, int size, int dx, int dy, uint32_t thr) {
  int s2 = size / 2, rb = size * bpp;
  const uint8_t* p1 = cur + (size_t)(fy - s2) * stride + (size_t)(fx - s2) * bpp;
  const uint8_t* p2 = prev + (size_t)(fy - s2 + dy) * stride + (ptrdiff_t)(fx - s2 + dx) * bpp;
  uint32_t sum = 0;
  for (int j = 0; j < size; j++) {
    for (int k = 0; k < rb; k++) sum += (uint32_t)abs((int)p1[k] - (int)p2[k]);
    if (sum > thr) break;
    p1 += stride;
    p2 += stride;
  }
  return sum;
}

uint32_t or_sad_packed(const uint8_t* cur, const uint8_t* prev, int stride, int bpp, int fx, int fy,
                       int size, int dx, int dy) {
  return or_sad_packed_thr(cur, prev, stride, bpp, fx, fy, size, dx, dy, UINT_MAX);
}

/* calcFieldTransPacked: the unshifted position first, then the raster
   i = -maxShift .. maxShift by stepSize (x outer, y inner, (0,0) skipped),
   then 1-px shifts within stepSize-1 of the best, strict "<" throughout;
   reject at maxShift + stepSize - 1 on either axis */
int or_field_search_packed(const uint8_t* cur, const uint8_t* prev, int stride, int bpp,
                           const OrFields* fs, const Field* f, int ox, int oy,
                           int* tx_out, int* ty_out, uint32_t* minerror_out) {
  int stepSize = fs->stepSize, maxShift = fs->maxShift;
  int tx = 0, ty = 0;
  uint32_t minerror = or_sad_packed_thr(cur, prev, stride, bpp, f->x, f->y, f->size, ox, oy, UINT_MAX);
  for (int i = -maxShift; i <= maxShift; i += stepSize)
    for (int j = -maxShift; j <= maxShift; j += stepSize) {
      if (i == 0 && j == 0) continue;
      uint32_t e = or_sad_packed_thr(cur, prev, stride, bpp, f->x, f->y, f->size, i + ox, j + oy, minerror);
      if (e < minerror) { minerror = e; tx = i; ty = j; }
    }
  if (stepSize > 1) {
    int txc = tx, tyc = ty, r = stepSize - 1;
    for (int i = txc - r; i <= txc + r; i++)
      for (int j = tyc - r; j <= tyc + r; j++) {
        if (i == txc && j == tyc) continue;
        uint32_t e = or_sad_packed_thr(cur, prev, stride, bpp, f->x, f->y, f->size, i + ox, j + oy, minerror);
        if (e < minerror) { minerror = e; tx = i; ty = j; }
      }
  }
  *tx_out = tx;
  *ty_out = ty;
  *minerror_out = minerror;
  if (abs(tx) >= maxShift + stepSize - 1 || abs(ty) >= maxShift + stepSize - 1) return 0;
  return 1;
}

/* ---- statistics: src/transformtype.c:445-459, 513-534 ---------------------- */

static int cmp_d(const void* a, const void* b) {
  double x = *(const double*)a, y = *(const double*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}
static int cmp_i(const void* a, const void* b) {
  int x = *(const int*)a, y = *(const int*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* 20%-trimmed mean.  The reference build sums the kept range in two SSE2
   lanes (even and odd offsets from `cut`, each starting at +0.0), adds the
   lanes, then adds the last element when the kept count is odd. */
double or_cleanmean(double* ds, int len, double* mn, double* mx) {
  int cut = len / 5;
  qsort(ds, (size_t)len, sizeof(double), cmp_d);
  int cnt = len - cut - cut;
  double sum = 0.0;
  if (cnt == 1) {
    sum = 0.0 + ds[cut];
  } else if (cnt > 1) {
    double even = 0.0, odd = 0.0;
    int pairs = cnt / 2;
    for (int p = 0; p < pairs; p++) {
      even += ds[cut + 2 * p];
      odd += ds[cut + 2 * p + 1];
    }
    sum = odd + even;
    if (cnt & 1) sum += ds[cut + cnt - 1];
  }
  if (mn) *mn = ds[cut];
  if (mx) *mx = ds[len - cut - 1];
  return sum / ((double)len - ((double)cut + (double)cut));
}

/* trimmed mean of v (int16 accumulation, truncating division) */
static Vec or_cleanmean_v(const LocalMotion* lms, int n) {
  int cut = n / 5;
  int* xs = (int*)malloc(sizeof(int) * (size_t)n);
  int* ys = (int*)malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; i++) { xs[i] = lms[i].v.x; ys[i] = lms[i].v.y; }
  qsort(xs, (size_t)n, sizeof(int), cmp_i);
  qsort(ys, (size_t)n, sizeof(int), cmp_i);
  int16_t sx = 0, sy = 0;
  for (int i = cut; i < n - cut; i++) sx = (int16_t)(sx + xs[i]);
  for (int i = cut; i < n - cut; i++) sy = (int16_t)(sy + ys[i]);
  double d = (double)n - ((double)cut + (double)cut);
  Vec m;
  m.x = (int16_t)(int)((double)sx / d);
  m.y = (int16_t)(int)((double)sy / d);
  free(xs);
  free(ys);
  return m;
}

/* src/localmotion2transform.c:336-349 */
static double or_calc_angle(const LocalMotion* lm, int cx, int cy) {
  int dx = lm->f.x - cx, dy = lm->f.y - cy;
  if (abs(dx) + abs(dy) < lm->f.size * 2) return 0.0;
  double a1 = atan2((double)dy, (double)dx);
  double a2 = atan2((double)(dy + lm->v.y), (double)(dx + lm->v.x));
  double diff = a2 - a1;
  if (diff > M_PI) return diff - 2.0 * M_PI;
  if (diff < -M_PI) return diff + 2.0 * M_PI;
  return diff;
}

/* src/localmotion2transform.c:352-402, in the compiled reference's order:
   t.x = ((cos a - 1) px - sin a py) + mean.x,  t.y = (mean.y + sin a px) + (cos a - 1) py */
VSTransform or_simple_transform(int w, int h, const LocalMotion* lms, int n) {
  VSTransform t;
  memset(&t, 0, sizeof(t));
  if (n < 1) return t;
  int cx = 0, cy = 0;
  for (int i = 0; i < n; i++) { cx += lms[i].f.x; cy += lms[i].f.y; }
  cx /= n;
  cy /= n;
  Vec mean = or_cleanmean_v(lms, n);
  double c1 = 0.0, s = 0.0, alpha = 0.0;
  if (n >= 6) {
    double* ang = (double*)malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; i++) {
      LocalMotion m = lms[i];
      m.v.x = (int16_t)(m.v.x - mean.x);
      m.v.y = (int16_t)(m.v.y - mean.y);
      ang[i] = or_calc_angle(&m, cx, cy);
    }
    double mn, mx;
    alpha = -or_cleanmean(ang, n, &mn, &mx);
    free(ang);
    if (mx - mn > 1.0) {
      alpha = 0.0;
    } else {
      double sn, cs;
      sincos(alpha, &sn, &cs);
      c1 = cs - 1.0;
      s = sn;
    }
  }
  double px = (double)(cx - w / 2), py = (double)(cy - h / 2);
  t.alpha = alpha;
  t.x = (c1 * px - s * py) + (double)mean.x;
  t.y = ((double)mean.y + s * px) + c1 * py;
  return t;
}

/* prepare_transform + transform_vec (src/transformtype.c:158-209) as compiled:
   x = (zcos rx + zsin ry) + (t.x + c_x),  y = (-rx zsin + ry zcos) + (t.y + c_y),
   then truncation to int16.  zoom is 0, so z = 1 exactly. */
void or_transform_point(const VSTransform* t, int w, int h, int x, int y, int* ox, int* oy) {
  double sn, cs;
  sincos(t->alpha, &sn, &cs);
  double z = 1.0 + t->zoom * 0.01;
  double zcos = cs * z, zsin = z * sn;
  double c_x = (double)(w / 2), c_y = (double)(h / 2);
  double vx = (double)x, vy = (double)y;
  double rx = vx - c_x, ry = vy - c_y;
  double X = (rx * zcos + zsin * ry) + (t->x + c_x);
  double Y = ((c_x - vx) * zsin + ry * zcos) + (t->y + c_y);
  *ox = (int16_t)(int)X;
  *oy = (int16_t)(int)Y;
}

/* ---- overlays: src/motiondetect.c:271-286, 816-952 --------------------------
   Grey levels into plane 0 (every colour byte of a packed pixel, never a
   fourth one), clipped to the frame. */

static void or_run(uint8_t* img, int stride, int w, int h, int bpp, int x, int y, int len, int vertical,
                   uint8_t c) {
  int nch = bpp > 3 ? 3 : bpp;
  if (vertical) {
    if (x < 0 || x >= w) return;
    int y0 = OR_MAX(y, 0), y1 = OR_MIN(y + len, h);
    for (int yy = y0; yy < y1; yy++)
      for (int k = 0; k < nch; k++) img[(size_t)yy * stride + (size_t)x * bpp + k] = c;
  } else {
    if (y < 0 || y >= h) return;
    int x0 = OR_MAX(x, 0), x1 = OR_MIN(x + len, w);
    for (int xx = x0; xx < x1; xx++)
      for (int k = 0; k < nch; k++) img[(size_t)y * stride + (size_t)xx * bpp + k] = c;
  }
}

/* filled box of sx x sy centred at (x, y) */
static void or_fill(uint8_t* img, int stride, int w, int h, int bpp, int x, int y, int sx, int sy, uint8_t c) {
  for (int j = 0; j < sy; j++) or_run(img, stride, w, h, bpp, x - sx / 2, y - sy / 2 + j, sx, 0, c);
}

/* outline: top and bottom rows of sx, left and right columns of sy */
static void or_frame(uint8_t* img, int stride, int w, int h, int bpp, int x, int y, int sx, int sy, uint8_t c) {
  or_run(img, stride, w, h, bpp, x - sx / 2, y - sy / 2, sx, 0, c);
  or_run(img, stride, w, h, bpp, x - sx / 2, y + sy / 2, sx, 0, c);
  or_run(img, stride, w, h, bpp, x - sx / 2, y - sy / 2, sy, 1, c);
  or_run(img, stride, w, h, bpp, x + sx / 2, y - sy / 2, sy, 1, c);
}

/* the reference's drawLine from the field centre (whose x, y it overwrites
   with the end point when it swaps the ends of a leftward or upward
   axis-parallel line) to the end point; the diagonal case steps rows with
   x = (int)((a.x - horlen/2) + m*dy), the order the reference's -ffast-math
   build evaluates it in */
static void or_arrow(uint8_t* img, int stride, int w, int h, int bpp, Field* a, int bx, int by, int thick,
                     uint8_t c) {
  int dx = bx - a->x, dy = by - a->y;
  if (dy == 0) {
    if (dx < 0) { a->x = (int16_t)bx; a->y = (int16_t)by; dx = -dx; }
    for (int r = -thick / 2; r <= thick / 2; r++) or_run(img, stride, w, h, bpp, a->x, a->y + r, dx + 1, 0, c);
  } else if (dx == 0) {
    if (dy < 0) { a->x = (int16_t)bx; a->y = (int16_t)by; dy = -dy; }
    for (int r = -thick / 2; r <= thick / 2; r++) or_run(img, stride, w, h, bpp, a->x + r, a->y, dy + 1, 1, c);
  } else {
    double m = (double)dx / (double)dy;
    int horlen = (int)((double)thick + fabs(m));
    double base = (double)(a->x - horlen / 2);
    for (int k = 0; k <= abs(dy); k++) {
      int step = dy < 0 ? -k : k;
      int x = (int)(base + m * (double)step);
      or_run(img, stride, w, h, bpp, x, a->y + step, horlen + 1, 0, c);
    }
  }
}

static void or_draw(const OrState* s, uint8_t* img, int stride, LocalMotion* co, int nc, LocalMotion* fi, int nf) {
  const int w = s->width, h = s->height, bpp = s->bpp;
  const int scan = s->coarse.maxShift;
  if (s->show > 1)
    for (int i = 0; i < nc; i++)
      or_frame(img, stride, w, h, bpp, co[i].f.x, co[i].f.y, co[i].f.size + 2 * scan, co[i].f.size + 2 * scan, 80);
  for (int i = 0; i < nc; i++) or_fill(img, stride, w, h, bpp, co[i].f.x, co[i].f.y, co[i].f.size, co[i].f.size, 40);
  for (int i = 0; i < nf; i++) or_frame(img, stride, w, h, bpp, fi[i].f.x, fi[i].f.y, fi[i].f.size, fi[i].f.size, 40);
  for (int pass = 0; pass < 2; pass++) {
    LocalMotion* lm = pass ? fi : co;
    int n = pass ? nf : nc;
    uint8_t color = pass ? 64 : 180;
    for (int i = 0; i < n; i++) {
      int ex = lm[i].f.x + lm[i].v.x, ey = lm[i].f.y + lm[i].v.y;
      or_fill(img, stride, w, h, bpp, lm[i].f.x, lm[i].f.y, 5, 5, 0);
      or_fill(img, stride, w, h, bpp, ex, ey, 5, 5, 250);
      or_arrow(img, stride, w, h, bpp, &lm[i].f, ex, ey, 3, color);
    }
  }
}

/* ---- the per-frame driver: src/motiondetect.c:191-305 --------------------- */

typedef struct { LocalMotion lm; int ok; } OrSlot;

static int or_calc_trans_fields(OrState* s, const OrFields* fs, const VSTransform* offset,
                                const double* contrast, LocalMotion* out) {
  int* idx = (int*)malloc(sizeof(int) * (size_t)fs->fieldNum);
  double* sc = (double*)malloc(sizeof(double) * (size_t)fs->fieldNum);
  int nsel = or_select(fs, contrast, idx, sc);
  OrSlot* slots = (OrSlot*)calloc((size_t)(nsel > 0 ? nsel : 1), sizeof(OrSlot));
  int q;
#pragma omp parallel for schedule(dynamic, 1)
  for (q = 0; q < nsel; q++) {
    const Field* f = &fs->fields[idx[q]];
    int ox = 0, oy = 0;
    if (offset) {
      int X, Y;
      or_transform_point(offset, s->width, s->height, f->x, f->y, &X, &Y);
      ox = (int16_t)(X - f->x);
      oy = (int16_t)(Y - f->y);
      int border = f->size / 2 + fs->maxShift + fs->stepSize;
      int x = f->x + ox, y = f->y + oy;
      if (x - border < 0 || x + border >= s->width || y - border < 0 || y + border >= s->height)
        continue;
    }
    int tx, ty;
    uint32_t me;
    int ok = s->bpp == 1
                 ? or_field_search(s->curr, s->prev, s->width, s->height, s->width, fs, f, ox, oy, &tx, &ty, &me)
                 : or_field_search_packed(s->curr, s->prev, s->width * s->bpp, s->bpp, fs, f, ox, oy, &tx, &ty,
                                          &me);
    if (!ok) continue;
    LocalMotion* lm = &slots[q].lm;
    memset(lm, 0, sizeof(*lm));
    lm->f = *f;
    lm->v.x = (int16_t)(tx + ox);
    lm->v.y = (int16_t)(ty + oy);
    lm->match = (double)me / (double)(f->size * f->size);
    lm->contrast = sc[q];
    slots[q].ok = 1;
  }
  int n = 0;
  for (q = 0; q < nsel; q++)
    if (slots[q].ok) out[n++] = slots[q].lm;
  free(slots);
  free(idx);
  free(sc);
  return n;
}

int or_detect(OrState* s, uint8_t* luma, int stride, LocalMotion* out, int maxout) {
  const int W = s->width, H = s->height, bpp = s->bpp;
  or_boxblur(s->curr, luma, W * bpp, H, stride, bpp == 1 ? s->blurSize : 1);
  int tripodLeadIn = s->virtualTripod >= 1 && s->frameNum < s->virtualTripod;
  int total = 0;
  s->nCoarse = s->nFineAll = s->nFineKept = 0;
  if (s->hasSeenOneFrame && !tripodLeadIn) {
    double* cc = (double*)malloc(sizeof(double) * (size_t)s->coarse.fieldNum);
    double* cf = (double*)malloc(sizeof(double) * (size_t)s->fine.fieldNum);
    for (int i = 0; i < s->coarse.fieldNum; i++) {
      int mn, mx;
      const Field* f = &s->coarse.fields[i];
      if (bpp != 1) {
        cc[i] = or_contrast_packed(s->curr, W * bpp, bpp, f->x, f->y, f->size);
        continue;
      }
      or_minmax(s->curr, W, f->x, f->y, f->size, &mn, &mx);
      cc[i] = or_contrast(mn, mx);
    }
    for (int i = 0; i < s->fine.fieldNum; i++) {
      int mn, mx;
      const Field* f = &s->fine.fields[i];
      if (bpp != 1) {
        cf[i] = or_contrast_packed(s->curr, W * bpp, bpp, f->x, f->y, f->size);
        continue;
      }
      or_minmax(s->curr, W, f->x, f->y, f->size, &mn, &mx);
      cf[i] = or_contrast(mn, mx);
    }
    LocalMotion* co = (LocalMotion*)malloc(sizeof(LocalMotion) * (size_t)(s->coarse.fieldNum + 1));
    LocalMotion* fi = (LocalMotion*)malloc(sizeof(LocalMotion) * (size_t)(s->fine.fieldNum + 1));
    int nc = or_calc_trans_fields(s, &s->coarse, NULL, cc, co);
    int nf = 0, nk = 0;
    if (nc >= 1) {
      VSTransform t = or_simple_transform(W, H, co, nc);
      s->lastT = t;
      nf = or_calc_trans_fields(s, &s->fine, &t, cf, fi);
      double* m = (double*)malloc(sizeof(double) * (size_t)nc);
      for (int i = 0; i < nc; i++) m[i] = co[i].match;
      double meanMatch = or_cleanmean(m, nc, NULL, NULL);
      s->lastMeanMatch = meanMatch;
      free(m);
      for (int i = 0; i < nf; i++)
        if (fi[i].match <= meanMatch) fi[nk++] = fi[i];
    }
    if (s->show) or_draw(s, luma, stride, co, nc, fi, nk);
    for (int i = 0; i < nc && total < maxout; i++) out[total++] = co[i];
    for (int i = 0; i < nk && total < maxout; i++) out[total++] = fi[i];
    s->nCoarse = nc;
    s->nFineAll = nf;
    s->nFineKept = nk;
    free(co);
    free(fi);
    free(cc);
    free(cf);
  } else {
    s->hasSeenOneFrame = 1;
  }
  if (s->virtualTripod < 1 || s->frameNum < s->virtualTripod)
    memcpy(s->prev, s->curr, (size_t)W * H * bpp);
  s->frameNum++;
  return total;
}

/* heap-allocated state, for ctypes callers */
OrState* or_new(int width, int height, int shakiness, int accuracy, int stepSize,
                double contrastThreshold, int virtualTripod) {
  OrState* s = (OrState*)calloc(1, sizeof(OrState));
  if (!s) return NULL;
  if (or_init(s, width, height, shakiness, accuracy, stepSize, contrastThreshold, virtualTripod) != 0) {
    or_free(s);
    free(s);
    return NULL;
  }
  return s;
}

OrState* or_new_packed(int width, int height, int bpp, int shakiness, int accuracy, int stepSize,
                       double contrastThreshold, int virtualTripod) {
  OrState* s = (OrState*)calloc(1, sizeof(OrState));
  if (!s) return NULL;
  if (or_init_packed(s, width, height, bpp, shakiness, accuracy, stepSize, contrastThreshold, virtualTripod) != 0) {
    or_free(s);
    free(s);
    return NULL;
  }
  return s;
}

void or_delete(OrState* s) {
  if (!s) return;
  or_free(s);
  free(s);
}

const uint8_t* or_blurred(const OrState* s) { return s->curr; }

/* ---- second pass: vsMotionsToTransform, similarity model (fov 0) ----------------------
   src/localmotion2transform.c:139-283 as the reference's -ffast-math build
   evaluates it (read from its x86 code compiled here):
   - calcTransformQuality (:139-171): prepare_transform_fov with f = 0, i.e.
     z = 1 + zoom*0.01, zcos = cos(alpha)*z, zsin = z*sin(alpha) (one sincos);
     transform_vec_double's order as in or_transform_point; the error sum is
     sequential (transform_vec_double is a call the loop cannot vectorize
     across); the result is (|alpha|*0.2 + |zoom|*0.002) + error/num, the
     divisions by 5 and 500 turned into products with the rounded reciprocals;
   - mean (src/transformtype.c:401-408): two interleaved partial sums
     (even and odd indices, from +0.0) added, then the odd tail; stddev
     (:421-428) sequential;
   - disableFields (:180-211): thresh = sigma*stddevs + mu, strict ">";
   - vsGradientDescent (:288-328): h = +1e-6 on odd sweeps, -1e-6 on even
     ones; the gradient is a true division by h; x2 = x + grad*step over all
     four parameters; step *= 1.2 on success, *= 0.5 otherwise;
   - intMean (:173-177): exact (a sum of integers in double). */
static double or_mean2(const double* ds, int len) {
  double sum = 0.0;
  if (len == 1) {
    sum = 0.0 + ds[0];
  } else if (len > 1) {
    double even = 0.0, odd = 0.0;
    for (int p = 0; p < len / 2; p++) {
      even += ds[2 * p];
      odd += ds[2 * p + 1];
    }
    sum = odd + even;
    if (len & 1) sum += ds[len - 1];
  }
  return sum / (double)len;
}

static double or_stddev(const double* ds, int len, double mu) {
  double sum = 0.0;
  for (int i = 0; i < len; i++) sum += (ds[i] - mu) * (ds[i] - mu);
  return sqrt(sum / (double)len);
}

static int or_disable_fields(double* mask, const double* q, int len, double stddevs) {
  const double mu = or_mean2(q, len);
  const double thresh = or_stddev(q, len, mu) * stddevs + mu;
  int cnt = 0;
  for (int i = 0; i < len; i++)
    if (q[i] > thresh) {
      mask[i] = -1.0;
      cnt++;
    }
  return cnt;
}

static double or_quality(const double p[4], const LocalMotion* lms, int n, double* miss, int w, int h) {
  double sn, cs;
  sincos(p[2], &sn, &cs);
  const double z = p[3] * 0.01 + 1.0;
  const double zcos = cs * z, zsin = z * sn;
  const double c_x = (double)(w / 2), c_y = (double)(h / 2);
  double err = 0.0;
  int num = 1;
  for (int i = 0; i < n; i++) {
    if (miss[i] >= 0) {
      const double vx = (double)lms[i].f.x, vy = (double)lms[i].f.y;
      const double rx = vx - c_x, ry = vy - c_y;
      const double X = (rx * zcos + zsin * ry) + (p[0] + c_x);
      const double Y = ((c_x - vx) * zsin + ry * zcos) + (p[1] + c_y);
      const double dx = (X - vx) - (double)lms[i].v.x, dy = (Y - vy) - (double)lms[i].v.y;
      const double e = dy * dy + dx * dx;
      miss[i] = e;
      err += e;
      num++;
    }
  }
  const double reg = fabs(p[2]) * 0.2 + fabs(p[3]) * 0.002;
  return reg + (n > 0 ? err / (double)num : 0.0);
}

static double or_gradient_descent(double x[4], const LocalMotion* lms, int n, double* miss, int w, int h) {
  double step[4] = {0.2, 0.2, 0.00005, 0.1};
  double v = or_quality(x, lms, n, miss, w, h);
  for (int i = 0; i < 16 * 4 && v > 0.01; i++) {
    const int k = i % 4;
    const double hh = ((i / 4) % 2) ? 1e-6 : -1e-6;
    double x2[4] = {x[0], x[1], x[2], x[3]};
    x2[k] = hh + x2[k];
    double v2 = or_quality(x2, lms, n, miss, w, h);
    double grad[4] = {0.0, 0.0, 0.0, 0.0};
    grad[k] = (v - v2) / hh;
    for (int j = 0; j < 4; j++) x2[j] = x[j] + grad[j] * step[k];
    v2 = or_quality(x2, lms, n, miss, w, h);
    if (v2 < v) {
      for (int j = 0; j < 4; j++) x[j] = x2[j];
      v = v2;
      step[k] *= 1.2;
    } else {
      step[k] *= 0.5;
    }
  }
  return v;
}

VSTransform or_motions_to_transform(int w, int h, const LocalMotion* lms, int n, int smoothZoom, double* residual,
                                    int* rounds) {
  VSTransform t;
  memset(&t, 0, sizeof(t));
  if (residual) *residual = 0.0;
  if (rounds) *rounds = 0;
  if (n <= 0) {
    t.extra = 1;
    return t;
  }
  double sx = 0.0, sy = 0.0;
  for (int i = 0; i < n; i++) {
    sx += (double)lms[i].v.x;
    sy += (double)lms[i].v.y;
  }
  double* miss = (double*)calloc((size_t)n, sizeof(double));
  double* q = (double*)malloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; i++) q[i] = lms[i].match;
  or_disable_fields(miss, q, n, 1.5);
  free(q);
  double params[4] = {sx / (double)n, sy / (double)n, 0.0, 0.0};
  double res = 0.0;
  int k;
  for (k = 0; k < 3; k++) {
    double x[4] = {params[0], params[1], params[2], params[3]};
    res = or_gradient_descent(x, lms, n, miss, w, h);
    for (int j = 0; j < 4; j++) params[j] = x[j];
    if ((k == 0 && res > 0.1) || (k == 1 && res > 20))
      or_disable_fields(miss, miss, n, 1.0);
    else
      break;
  }
  free(miss);
  t.x = params[0];
  t.y = params[1];
  t.alpha = params[2];
  t.zoom = smoothZoom ? params[3] : 0.0;
  t.extra = res > 100 ? 1 : 0;
  if (residual) *residual = res;
  if (rounds) *rounds = k + 1;
  return t;
}
