/*
 * vs_detect_internal.h -- host steps of the detect pass shared by the
 * single-clip driver (vs_motiondetect.c) and the multi-clip batched form
 * (vs_multidetect.c).  Not exported.
 */
#ifndef VS_DETECT_INTERNAL_H
#define VS_DETECT_INTERNAL_H

#include "lensdistortion.h"
#include "motiondetect.h"
#include "serialize.h"
#include "vsg.h"

/* vsMotionDetectInit without the GPU context: validation, config clamps, the
   coarse and fine field grids (src/motiondetect.c:83-165, 314-358) */
int vsd_configure(VSMotionDetect* md, const VSMotionDetectConfig* conf, const VSFrameInfo* fi);

/* box size of boxblurPlanar for this detector (src/boxblur.c:109-122) */
int vsd_blur_size(const VSMotionDetect* md);

/* GPU description of one field grid */
VSGStageDesc vsd_stage_desc(const VSMotionDetectFields* fs);

/* Frame i of a batch view: builds its coarse LocalMotions (in selection
   order, rejected fields dropped), computes vsSimpleMotionsToTransform into
   *offset and writes the fine search jobs (slot / prev given) to jobs[];
   jobSel[] gets each job's index in the fine selection.  Returns the number
   of jobs.  Reads md only, so frames may run concurrently. */
int vsd_frame_coarse(const VSMotionDetect* md, const VSGBatchView* v, int i, int fnum, int slot, int prevSlot,
                     LocalMotions* coarse, VSTransform* offset, VSGJob* jobs, int* jobSel);

/* Frame i: fine results of its njobs jobs -> the cleaned fine motions; with
   md->conf.show > 0 and a frame to draw into (plane 0, linesize bytes a
   row, in host memory or, for drawDevice >= 0, on that HIP device; null:
   none), the overlays; then out = coarse ++ fine (the elements of coarse
   move into out).  Reads md only. */
void vsd_frame_fine(const VSMotionDetect* md, const VSGBatchView* v, int i, LocalMotions* coarse,
                    const VSGJob* jobs, const int* jobSel, const VSGResult* res, int njobs, uint8_t* draw,
                    int drawLinesize, int drawDevice, LocalMotions* out);

/* the show overlays of one frame (src/motiondetect.c:271-286, 816-952) into
   plane 0 of the caller's frame; may rewrite the field position of motions
   whose line the reference draws from its end (vs_draw.c) */
void vsd_draw_overlays(const VSMotionDetect* md, uint8_t* img, int linesize, LocalMotions* coarse,
                       LocalMotions* fine);

/* ---- second pass (vs_lensdistortion.c) ---- */

/* the HIP device of this process's detectors (VIDSTAB_AMD_DEVICE, default 0) */
int vsd_device(void);

/* a clip's LocalMotions in the GPU's host layout: 4 int16 per motion (field
   x, y, motion x, y), the match values, the motion count of each frame
   (malloc'd; the caller frees them).  VS_OK or VS_ERROR. */
int vsd_flatten_motions(const VSManyLocalMotions* motions, int16_t** lm, double** match, int** count);

/* the clip-wide lens estimate over the clip held by `gpu`
   (src/lensdistortion.c:528-639); *failed is set when a GPU call failed */
VSLensEstimate vsd_lens_estimate(VSGLens* gpu, int nframes, const int* count, const VSLensEstimateConfig* cfg,
                                 int* failed);

#endif /* VS_DETECT_INTERNAL_H */
