/*
 * vsg.h -- thin C ABI between the detect library's host C code and its gfx950
 * HIP kernels ("vsg" = vid.stab GPU).  Plain pointers and sizes only.
 *
 * The host side (motiondetect driver) owns every double-precision and
 * ordering decision of the reference (contrast selection, the coarse->fine
 * transform, the match cleanup); this layer owns the device memory and runs
 * the data-parallel integer stages that replace the reference's CPU kernels:
 *
 *   vsg_frames  -- boxblurPlanar (src/boxblur.c:104-332) on a batch of luma
 *                  planes, plus the per-field 8-bit min/max of contrastSubImg
 *                  (src/motiondetect.c:389-407) for every coarse and fine field
 *   vsg_search  -- calcFieldTransPlanar (src/motiondetect.c:439-586) with its
 *                  compareSubImg SAD (src/motiondetect.c:971-996): the spiral
 *                  grid search and the dyadic refinement for a list of fields
 *
 * Every function returns 0 on success and a negative value on failure;
 * vsg_last_error() then describes the failure (thread-local).
 */
#ifndef VSG_H
#define VSG_H

#include <stdint.h>
#include "transformtype.h"
#include "vidstab_api.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct VSGContext VSGContext;

/* one field grid (coarse or fine): VSMotionDetectFields of the reference */
typedef struct {
  int fieldSize;        /* multiple of 16 */
  int maxShift;
  int stepSize;         /* >= 1 */
  int nfields;
  const Field* fields;  /* host array of nfields centres */
} VSGStageDesc;

/* one field search: current frame = batch slot `slot`; previous frame = batch
   slot `prev`, or the stored reference plane when prev < 0; `field` indexes
   the stage's field array; (ox, oy) shifts the search window (fine stage) */
typedef struct {
  int32_t slot;
  int32_t prev;
  int32_t field;
  int16_t ox, oy;
} VSGJob;

/* best shift in px relative to (ox, oy), after refinement, and its SAD */
typedef struct {
  int16_t tx, ty;
  uint32_t minerror;
} VSGResult;

/* Creates the device state for width x height frames on `device`, blur box
   `blurSize` (< 2: no blur), two field grids, and batches of up to maxBatch
   frames.  Replaces the buffers vsMotionDetectInit allocates
   (src/motiondetect.c:104-164). */
VS_API int vsg_create(VSGContext** ctx, int device, int width, int height, int blurSize,
                      const VSGStageDesc* coarse, const VSGStageDesc* fine, int maxBatch);
VS_API void vsg_destroy(VSGContext* ctx);

/* Loads n luma planes into batch slots 0..n-1 (host pointers, or device
   pointers when onDevice != 0; linesize in bytes), blurs them, and returns the
   8-bit (min, max) of every field: minmaxCoarse[n][ncoarse][2],
   minmaxFine[n][nfine][2]. */
VS_API int vsg_frames(VSGContext* ctx, int n, const uint8_t* const* luma, const int* linesize,
                      int onDevice, uint8_t* minmaxCoarse, uint8_t* minmaxFine);

/* Runs njobs field searches of one stage (0 coarse, 1 fine). */
VS_API int vsg_search(VSGContext* ctx, int stage, int njobs, const VSGJob* jobs, VSGResult* out);

/* Makes batch slot `slot` the stored reference plane (the reference's
   vsFrameCopy(&md->prev, &md->curr), src/motiondetect.c:299-303). */
VS_API int vsg_keep_prev(VSGContext* ctx, int slot);

/* Copies a blurred plane to the host (slot < 0: the stored reference). */
VS_API int vsg_read_blurred(VSGContext* ctx, int slot, uint8_t* dst, int dstStride);

VS_API const char* vsg_last_error(void);

/* Optional per-kernel timing (HIP events on the detector's own stream around
   each launch; process-wide counters).  work is the kernel's algorithmic unit
   count: abs-diffs for the searches (candidates x field pixels), bytes for
   the blur (2 x W x H per frame) and the field pass (field bytes). */
enum { VSG_K_BLUR = 0, VSG_K_FIELDS = 1, VSG_K_GRID_COARSE = 2, VSG_K_GRID_FINE = 3,
       VSG_K_REFINE = 4, VSG_K_COUNT = 5 };
typedef struct {
  long long launches[VSG_K_COUNT];
  double ms[VSG_K_COUNT];
  double work[VSG_K_COUNT];
} VSGProfile;
VS_API void vsg_profile_enable(int on);
VS_API void vsg_profile_get(VSGProfile* out);
VS_API void vsg_profile_reset(void);

/* host versions of the device spiral helpers, for tests */
VS_API int vsg_spiral_rank(int a, int b);
VS_API void vsg_spiral_unrank(int rank, int* a, int* b);

#ifdef __cplusplus
}
#endif
#endif /* VSG_H */
