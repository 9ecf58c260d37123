/*
 * vsoracle.h -- CPU restatement of vid.stab's detect pass.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the GPU library is
 * compared against; it is loaded by tests/, by __graft_entry__.smoke() and by
 * bench.py's CPU-baseline leg, and never by the product library.
 */
#ifndef VSORACLE_H
#define VSORACLE_H

#include <stdint.h>
#include "transformtype.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int fieldSize, maxShift, stepSize;
  int fieldNum, maxFields, fieldRows;
  double contrastThreshold;
  Field* fields;
} OrFields;

typedef struct {
  int width, height;
  int shakiness, accuracy, stepSize, virtualTripod;
  double contrastThreshold;
  OrFields coarse, fine;
  int blurSize;               /* 0/1: plain copy */
  uint8_t *curr, *prev, *tmp; /* width x height, stride width */
  int hasSeenOneFrame, frameNum;
  /* diagnostics of the last detect step */
  int nCoarse, nFineAll, nFineKept;
  VSTransform lastT;
  double lastMeanMatch;
  int bpp;                    /* 1: planar luma; 3 or 4: packed RGB24/BGR24/RGBA plane 0 */
  int show;                   /* 1 / 2: overlays drawn into the frame handed to or_detect */
} OrState;

int  or_init(OrState* s, int width, int height, int shakiness, int accuracy, int stepSize,
             double contrastThreshold, int virtualTripod);
/* packed frames: bpp 3 or 4, curr/prev hold width*bpp bytes a row, unblurred */
int  or_init_packed(OrState* s, int width, int height, int bpp, int shakiness, int accuracy, int stepSize,
                    double contrastThreshold, int virtualTripod);
void or_free(OrState* s);
/* one vsMotionDetection step on a luma plane (packed: plane 0); writes up to
   maxout motions and returns their number (coarse ++ cleaned fine).  With
   s->show > 0 the overlays are drawn into luma, as the reference draws into
   the caller's frame. */
int  or_detect(OrState* s, uint8_t* luma, int stride, LocalMotion* out, int maxout);

/* building blocks, exposed for unit tests */
int      or_blur_size(int stepSize, int width, int height);
void     or_boxblur(uint8_t* dst, const uint8_t* src, int w, int h, int sstride, int size);
void     or_minmax(const uint8_t* img, int stride, int x, int y, int size, int* mn, int* mx);
double   or_contrast(int mn, int mx);
uint32_t or_sad(const uint8_t* cur, const uint8_t* prev, int stride, int fx, int fy, int size,
                int dx, int dy);
int      or_spiral_rank(int a, int b);
void     or_spiral_walk(int n, int* as, int* bs);   /* (2n+1)^2 entries, visit order */
int      or_select(const OrFields* fs, const double* contrast, int* idx, double* sel_contrast);
/* exact search for one field; returns 1 if accepted and fills tx,ty,minerror */
int      or_field_search(const uint8_t* cur, const uint8_t* prev, int w, int h, int stride,
                         const OrFields* fs, const Field* f, int ox, int oy,
                         int* tx, int* ty, uint32_t* minerror);
/* packed building blocks (src/motiondetect.c:370-377, 591-660, 971-996) */
double   or_contrast_packed(const uint8_t* img, int stride, int bpp, int x, int y, int size);
uint32_t or_sad_packed(const uint8_t* cur, const uint8_t* prev, int stride, int bpp, int fx, int fy,
                       int size, int dx, int dy);
int      or_field_search_packed(const uint8_t* cur, const uint8_t* prev, int stride, int bpp,
                                const OrFields* fs, const Field* f, int ox, int oy,
                                int* tx, int* ty, uint32_t* minerror);
double   or_cleanmean(double* ds, int len, double* mn, double* mx);
VSTransform or_simple_transform(int w, int h, const LocalMotion* lms, int n);
void     or_transform_point(const VSTransform* t, int w, int h, int x, int y, int* ox, int* oy);
/* second pass: vsMotionsToTransform of one frame's LocalMotions with the
   similarity model (VSTransformConfig.fov = 0, no lens); residual and the
   number of gradient-descent rounds are optional outputs */
VSTransform or_motions_to_transform(int w, int h, const LocalMotion* lms, int n, int smoothZoom, double* residual,
                                    int* rounds);

OrState*       or_new(int width, int height, int shakiness, int accuracy, int stepSize,
                      double contrastThreshold, int virtualTripod);
OrState*       or_new_packed(int width, int height, int bpp, int shakiness, int accuracy, int stepSize,
                             double contrastThreshold, int virtualTripod);
void           or_delete(OrState* s);
const uint8_t* or_blurred(const OrState* s);

#ifdef __cplusplus
}
#endif
#endif /* VSORACLE_H */
