/*
 * vsg.h -- thin C ABI between the detect library's host C code and its gfx950
 * HIP kernels ("vsg" = vid.stab GPU).  Plain pointers and sizes only.
 *
 * The host side (the motiondetect driver) owns the double-precision steps
 * that need glibc's libm (the coarse->fine transform) and the match cleanup;
 * this layer owns the device memory and runs the data-parallel stages that
 * replace the reference's CPU kernels:
 *
 *   vsg_submit      boxblurPlanar (src/boxblur.c:104-332) of a batch of luma
 *                   planes; per-field min/max of contrastSubImg
 *                   (src/motiondetect.c:389-407); selectfields
 *                   (src/motiondetect.c:676-743) for the coarse and fine grids;
 *                   the coarse calcFieldTransPlanar (src/motiondetect.c:439-586)
 *                   with its compareSubImg SAD (src/motiondetect.c:971-996)
 *   vsg_submit_fine the same search for the fine fields at host-computed offsets
 *
 * A context has VSG_NSETS sets of maxBatch frame slots (global slot =
 * set*maxBatch + i), so the host can prepare one batch's fine stage while the
 * device runs the next batches, and a batch's planes survive until its set is
 * submitted again.  Each set runs on its own HIP stream; the context orders
 * what crosses sets (a batch's previous-frame planes, the tripod plane) with
 * events.  Contract: a batch's prevSlot entries name slots of its own set or
 * of the batch submitted just before it, and a set is submitted again only
 * after the fine stage of the batch that followed it has been submitted (or
 * will never be).  Submit calls return at once, wait calls block.
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

/* slot sets of a context (batches k, k+1, k+2 use different sets) */
#define VSG_NSETS 3

/* one field grid with its selection parameters (VSMotionDetectFields) */
typedef struct {
  int fieldSize;            /* multiple of 16 */
  int maxShift;
  int stepSize;             /* >= 1 */
  int nfields;
  int fieldRows;
  int maxFields;
  double contrastThreshold;
  const Field* fields;      /* host array of nfields centres */
} VSGStageDesc;

/* one field search: current frame in slot `slot`, previous plane in slot
   `prev` or the stored reference plane when prev < 0; `field` indexes the
   stage's fields (< 0: skip); (ox, oy) shifts the search window */
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

/* what vsg_wait_coarse hands back for a batch of n frames (pinned host
   memory owned by the context, valid until the set is submitted again) */
typedef struct {
  const int32_t* count;        /* [n][2]: selected coarse, fine fields */
  const int16_t* idx;          /* [n][stride]: coarse selection, then fine selection */
  const double* contrast;      /* [n][stride]: selection contrast */
  const VSGResult* coarse;     /* [n][ncoarse]: search result of each coarse selection */
  int stride;                  /* ncoarse + nfine */
  int ncoarse;
} VSGBatchView;

/* Creates the device state for width x height frames on `device`: blur box
   `blurSize` (< 2: plain copy), the coarse and fine grids, VSG_NSETS sets of
   maxBatch slots.  Replaces the buffers of vsMotionDetectInit
   (src/motiondetect.c:104-164).  Frames are planar; their luma is searched. */
VS_API int vsg_create(VSGContext** ctx, int device, int width, int height, int blurSize,
                      const VSGStageDesc* coarse, const VSGStageDesc* fine, int maxBatch);
/* The same for frames of bytesPerPixel bytes a pixel: 1 is vsg_create; 3 or 4
   are packed RGB24 / BGR24 / RGBA frames, whose plane 0 (width*bytesPerPixel
   bytes a row) is searched unblurred with the packed rules
   (src/motiondetect.c:197-201, 370-377, 591-660).  Packed frames need
   blurSize < 2. */
VS_API int vsg_create_fmt(VSGContext** ctx, int device, int width, int height, int bytesPerPixel,
                          int blurSize, const VSGStageDesc* coarse, const VSGStageDesc* fine, int maxBatch);
VS_API void vsg_destroy(VSGContext* ctx);

/* where vsg_submit finds a batch's frames */
enum {
  VSG_FRAMES_HOST = 0,     /* luma[i]: host memory, staged by vsg_submit */
  VSG_FRAMES_DEVICE = 1,   /* luma[i]: device memory of the context's GPU, read in place */
  VSG_FRAMES_STAGED = 2,   /* already copied by the caller into vsg_stage_frames' buffers */
  VSG_FRAMES_UPLOADED = 3  /* staged, and sent up to the device by vsg_stage_upload */
};

/* Enqueues a batch of n <= maxBatch frames into `set`: upload (host luma),
   upload from the staging buffers, or in-place read (onDevice, one of the
   VSG_FRAMES_* values), blur, field min/max, selection of both grids, and
   the coarse search of every frame with detect[i] != 0 against slot
   prevSlot[i] (or the stored reference when -1).  If refSlot >= 0, that slot's
   blurred plane becomes the stored reference right after the blur. */
VS_API int vsg_submit(VSGContext* ctx, int set, int n, const uint8_t* const* luma,
                      const int* linesize, int onDevice, const int* prevSlot,
                      const uint8_t* detect, int refSlot);
/* The pinned staging buffers of the next n frames of `set` (stage[i]: pitch
   x height bytes, *pitch a multiple of 64), once the set's previous batch is
   over; the caller copies each frame's luma rows (width bytes each) into
   them, on as many threads as it likes, and submits with VSG_FRAMES_STAGED. */
VS_API int vsg_stage_frames(VSGContext* ctx, int set, int n, uint8_t** stage, int* pitch);
/* Sends bytes [offset, offset + bytes) of the set's staging area (frame i at
   offset i x pitch x height) up to the device frames the next vsg_submit of
   the set reads, by a copy kernel queued on the compute stream ahead of it:
   a caller that stages a frame in row chunks uploads each chunk while it
   stages the next, then submits with VSG_FRAMES_UPLOADED.  offset and bytes
   are multiples of the pitch. */
VS_API int vsg_stage_upload(VSGContext* ctx, int set, size_t offset, size_t bytes);
VS_API int vsg_wait_coarse(VSGContext* ctx, int set, VSGBatchView* view);

/* Enqueues / waits for the fine searches of a set (jobs built by the host). */
VS_API int vsg_submit_fine(VSGContext* ctx, int set, int njobs, const VSGJob* jobs);
VS_API int vsg_wait_fine(VSGContext* ctx, int set, const VSGResult** results);

/* Copies a blurred plane (global slot; < 0: the stored reference) to the host. */
VS_API int vsg_read_blurred(VSGContext* ctx, int slot, uint8_t* dst, int dstStride);

/* Copies rows x rowBytes bytes between host and device memory of `device`,
   either way (synchronous; the show overlays of device frames go through
   the host with it). */
VS_API int vsg_copy_plane(int device, uint8_t* dst, int dstPitch, const uint8_t* src, int srcPitch, int rowBytes,
                          int rows);

VS_API const char* vsg_last_error(void);

/* Optional per-kernel timing: HIP events on the context's stream around each
   launch, process-wide counters.  work is the algorithmic unit count:
   abs-diffs for the searches (candidates x field pixels), bytes for the blur
   (2 x W x H per frame) and the field pass (2 x field bytes). */
enum { VSG_K_BLUR = 0, VSG_K_FIELDS = 1, VSG_K_SELECT = 2, VSG_K_GRID_COARSE = 3,
       VSG_K_REFINE_COARSE = 4, VSG_K_GRID_FINE = 5, VSG_K_REFINE_FINE = 6, VSG_K_COUNT = 7 };
typedef struct {
  long long launches[VSG_K_COUNT];
  double ms[VSG_K_COUNT];
  double work[VSG_K_COUNT];
} VSGProfile;
VS_API void vsg_profile_enable(int on);
VS_API void vsg_profile_get(VSGProfile* out);
VS_API void vsg_profile_reset(void);

/* ---- the learned-proposal form ------------------------------------------------
   Attaches a proposal network to a context: from the next batch on, each
   coarse field is searched only in the 5x5 window of 1-px shifts around the
   network's proposal, and goes through the exact full search whenever the
   window's minimum lies on its edge, exceeds `threshold` (match units:
   SAD / field area), or the proposal is out of the search range.  Fine
   fields are searched the same way around the network's displacement less
   the job's offset, unless VIDSTAB_AMD_PROPOSAL_FINE=0.  weights: per layer
   W[9][32][32] (tap, in, out) then B[32], FP32; dil: per layer dilation
   (1..8).  Single-clip contexts only. */
VS_API int vsg_set_proposal(VSGContext* ctx, int nlayers, const int* dil, const float* weights, float threshold);
/* coarse fields decided so far, and how many of them needed the full search */
VS_API int vsg_proposal_stats(VSGContext* ctx, long long* fields, long long* fallbacks);
/* the same for stage 0 (coarse) or 1 (fine: zero while fine proposals are
   switched off by VIDSTAB_AMD_PROPOSAL_FINE=0) */
VS_API int vsg_proposal_stage_stats(VSGContext* ctx, int stage, long long* fields, long long* fallbacks);
/* the fallbacks of stage 0 (coarse) or 1 (fine) by cause: proposal out of
   the search range, window minimum on the window edge, minimum above the
   residual threshold */
VS_API int vsg_proposal_fallbacks(VSGContext* ctx, int stage, long long* outOfRange, long long* edge,
                                  long long* residual);

/* ---- second pass: per-frame global transforms ------------------------------------
   vsMotionsToTransform (src/localmotion2transform.c:213-283) of nframes
   frames at once, with the similarity model (the transform config's fov 0,
   no lens estimate), on HIP device `device`.  Host arrays: lm holds
   4 int16 per motion (field x, field y, motion x, motion y) and match one
   double per motion, frames back to back, count[f] motions in frame f.
   out[f] gets the frame's transform (zoom kept only with smoothZoom != 0;
   extra = 1 for an empty frame or a residual above 100); residual[f] and
   rounds[f] (gradient-descent rounds, 1..3) may be NULL.  The device's sin
   and cos may differ from glibc's in the last bit, so results match the
   reference to a tolerance.  Allocates and frees its own device memory. */
VS_API int vsg_motions_to_transform(int device, int width, int height, const int16_t* lm, const double* match,
                                    const int* count, int nframes, int smoothZoom, VSTransform* out,
                                    double* residual, int* rounds);
/* vsg_motions_to_transform with the rotational model: focal > 0 is the focal
   length in px (focal_from_fov of the transform config's fov; 0 is exactly
   vsg_motions_to_transform) -- calcTransformQuality's prepare_transform_fov
   branch, src/localmotion2transform.c:153-155, src/transformtype.c:133-209 */
VS_API int vsg_motions_to_transform_fov(int device, int width, int height, const int16_t* lm, const double* match,
                                        const int* count, int nframes, int smoothZoom, double focal,
                                        VSTransform* out, double* residual, int* rounds);

/* ---- second pass: the lens model (src/lensdistortion.c) -----------------------
   One clip's LocalMotions on the device (same host layout as above; at most
   4096 motions a frame) and a per-motion mask for the estimator.  Each call
   runs every frame at once, one workgroup per frame, and returns per frame:
     vsg_lens_fit         the fit at strength k (focal f, gnSteps Gauss-Newton
                          steps; with useMask the masked points left out):
                          r[f] = RMS residual, status[f] 0 fitted, 1 failed
                          (vsLensFitSimilarityFov returned VS_ERROR), 2 fewer
                          than 3 motions (:202-365, :455-478)
     vsg_lens_mask        the estimator's masking pass at k (:562-581):
                          similarity fit with the mask, residuals, median /
                          MAD rejection (stddevs); cut[f] newly masked points
     vsg_lens_transforms  vsLensMotionsToTransform of every frame (:725-786):
                          out[f], residual[f], status[f] as above
   vsg_lens_reset_mask marks every motion active; vsg_lens_rejected counts
   the masked ones. */
typedef struct VSGLens VSGLens;
VS_API int vsg_lens_create(VSGLens** out, int device, int width, int height, const int16_t* lm, const double* match,
                           const int* count, int nframes);
VS_API void vsg_lens_destroy(VSGLens* lens);
VS_API int vsg_lens_reset_mask(VSGLens* lens);
VS_API int vsg_lens_rejected(VSGLens* lens, long long* rejected);
VS_API int vsg_lens_fit(VSGLens* lens, double k, double f, int gnSteps, int useMask, double* r, int* status);
VS_API int vsg_lens_mask(VSGLens* lens, double k, int gnSteps, double stddevs, int* cut);
VS_API int vsg_lens_transforms(VSGLens* lens, double k, double f, int gnSteps, int reject, double stddevs,
                               VSTransform* out, double* residual, int* status);

/* vsg_motions_to_transform without any device allocation: no counterpart in
   the reference; for callers that cannot allocate device memory (code next
   to HIP graph capture, serving loops, harnesses that forbid hipMalloc).
   Same arguments and results.  The inputs and outputs are staged through a
   fixed 4.8 MB device array of the library (one per GPU, held until the
   process ends), so a call takes at most 8192 frames and 262144 motions in
   all and fails with an error beyond that; it creates no stream (the null
   stream) and calls no hipMalloc / hipFree.  Calls on one GPU share the
   array and are serialized by a lock. */
VS_API int vsg_motions_to_transform_static(int device, int width, int height, const int16_t* lm,
                                           const double* match, const int* count, int nframes, int smoothZoom,
                                           VSTransform* out, double* residual, int* rounds);

/* ---- the batched multi-clip form ---------------------------------------------
   One context per clip, all on one stream; the searches of a stage run as one
   persistent launch over a single queue of the units of every clip, and each
   slot set's stages are captured once as HIP graphs (fixed shapes: B slots
   per clip, idle slots and padded fine lists skipped on the device), with one
   host<->device copy per crossing for all clips.  Frames must be in device
   memory.  Per-clip results are read with vsg_wait_coarse / vsg_wait_fine on
   vsg_multi_clip(m, c) after the multi waits. */
typedef struct VSGMulti VSGMulti;
VS_API int vsg_multi_create(VSGMulti** m, int device, int nclips, const int* width, const int* height,
                            const int* blurSize, const VSGStageDesc* coarse, const VSGStageDesc* fine,
                            int maxBatch);
VS_API void vsg_multi_destroy(VSGMulti* m);
VS_API VSGContext* vsg_multi_clip(VSGMulti* m, int clip);
/* search launches per batch of a stage (1: the queue launch) */
VS_API int vsg_multi_groups(VSGMulti* m, int stage);
/* n[c] <= maxBatch frames of clip c (0: idle); lumaDev[c][i], linesize[c];
   prevSlot[c][i] a slot of the clip or -1 (the clip's stored reference
   plane), detect[c][i] as for vsg_submit; refSlot (NULL: none) names per
   clip a slot of this batch whose blurred plane becomes the clip's stored
   reference before any search of the batch, or -1 */
VS_API int vsg_multi_submit(VSGMulti* m, int set, const int* n, const uint8_t* const* const* lumaDev,
                            const int* linesize, const int* const* prevSlot, const uint8_t* const* detect,
                            const int* refSlot);
VS_API int vsg_multi_wait_coarse(VSGMulti* m, int set);
VS_API int vsg_multi_submit_fine(VSGMulti* m, int set, const int* njobs, const VSGJob* const* jobs);
VS_API int vsg_multi_wait_fine(VSGMulti* m, int set);

/* The search plan of a stage (0 coarse, 1 fine) chosen by the autotuner:
   out[7] = J, K, D, CW, workgroup size, workgroups per job, LDS-staged field
   (0/1); tuneMs[6] (may be NULL) = the timed batch search with K = 4, 6, 8
   reading the field by scalar loads, then staged in LDS (0 when not timed). */
VS_API int vsg_get_plan(VSGContext* ctx, int stage, int* out, float* tuneMs);

/* Measured packed-u8 SAD ceiling of `device`: a kernel of independent
   v_sad_u8 chains with a wave-uniform operand and no memory traffic;
   *absdiffPerSec = 4 abs-diffs per lane per instruction.  Host-synchronous. */
VS_API int vsg_sad_peak(int device, double* absdiffPerSec, double* ms);

/* Mean duration of the batch box blur (the detector's blur kernel, blur box
   `blurSize`) over `reps` back-to-back launches on `n` device-resident planes
   of width x height bytes, on an otherwise idle device, after one untimed
   launch.  Host-synchronous; allocates its own output slots. */
VS_API int vsg_blur_time(int device, const uint8_t* const* lumaDev, const int* linesize, int n, int width,
                         int height, int blurSize, int reps, double* msPerLaunch);

/* host versions of the device spiral helpers, for tests */
VS_API int vsg_spiral_rank(int a, int b);
VS_API void vsg_spiral_unrank(int rank, int* a, int* b);

#ifdef __cplusplus
}
#endif
#endif /* VSG_H */
