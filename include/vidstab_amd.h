/*
 * vidstab_amd.h -- extensions of the detect API that have no counterpart in
 * the reference.  They batch what the reference does one call at a time; the
 * results are exactly those of n consecutive vsMotionDetection calls.
 *
 * Environment (read by vsMotionDetectInit):
 *   VIDSTAB_AMD_DEVICE     HIP device ordinal of new detectors (default 0)
 *   VIDSTAB_AMD_MAX_BATCH  most frames per device batch (default 128, at most 128;
 *                          a call's frames go through as few batches as that allows,
 *                          all of one size but the last)
 *   VIDSTAB_AMD_SEARCH_LDS 1: field patch in LDS, 0: scalar loads (default: autotuned)
 *   VIDSTAB_AMD_HOST_THREADS  threads for the per-frame host steps of a batch
 *                          (default: numThreads, else 0.8 x cores; at most 16)
 */
#ifndef VIDSTAB_AMD_H
#define VIDSTAB_AMD_H

#include <stdint.h>
#include "motiondetect.h"

#ifdef __cplusplus
extern "C" {
#endif

/* n consecutive detect steps on host frames; motions[i] receives the
   LocalMotions of frames[i].  For callers that hold several frames at once
   (file-based tools); one device batch per VIDSTAB_AMD_MAX_BATCH frames. */
VS_API int vsMotionDetectionBatch(VSMotionDetect* md, LocalMotions* motions, VSFrame* frames, int n);

/* as vsMotionDetectionBatch, but the luma planes are already in device memory
   of the detector's GPU (lumaDev[i], linesize[i] bytes per row).  With
   show > 0 the overlays are drawn into those device frames, as
   vsMotionDetection draws them into a host frame (the frames must then be
   writable). */
VS_API int vsMotionDetectionDevice(VSMotionDetect* md, LocalMotions* motions,
                                   const uint8_t* const* lumaDev, const int* linesize, int n);

/* copies the last blurred plane of the detector (the reference's md->curr
   luma; for packed RGB24 / BGR24 / RGBA frames the unblurred plane 0,
   width * bytesPerPixel bytes a row) into dst; for parity checks */
VS_API int vsMotionDetectGetBlurred(const VSMotionDetect* md, uint8_t* dst, int dstStride);
/* the same for the frame `back` frames before the last one (0: the last);
   the planes of the detector's latest three device batches are kept, older
   ones are overwritten (VS_ERROR) */
VS_API int vsMotionDetectGetBlurredAt(const VSMotionDetect* md, int back, uint8_t* dst, int dstStride);

/* The learned-proposal form: loads a proposal network (the weights file
   written by the project's trainer: magic "VSGP", int32 version 1, int32
   layers, int32 dilations[layers], float32 residual threshold, then per
   layer float32 W[9][32][32] and B[32], little-endian) and switches the
   detector's coarse and fine stages to network proposals refined in a
   +-2 px SAD window, with the exact full search as the fallback for every
   field whose window minimum is on the edge or above the threshold
   (VIDSTAB_AMD_PROPOSAL_FINE=0 keeps the exact fine search).  Everything
   after the searches is unchanged. */
VS_API int vsMotionDetectSetProposal(VSMotionDetect* md, const char* weightsFile);
/* coarse fields decided by the learned form so far, and the fallbacks */
VS_API int vsMotionDetectProposalStats(const VSMotionDetect* md, long long* fields, long long* fallbacks);
/* the same for stage 0 (coarse) or 1 (fine) */
VS_API int vsMotionDetectProposalStageStats(const VSMotionDetect* md, int stage, long long* fields,
                                            long long* fallbacks);

/* The batched multi-clip form: the detect pass of nclips independent clips
   at once, frames already in device memory.  Clip c (config confs[c], frame
   info fis[c], planar; virtualTripod as for vsMotionDetection) has nframes[c] luma planes
   lumaDev[c][0..nframes[c]-1] with linesize[c] bytes per row; motions[c][t]
   receives exactly what the t-th of nframes[c] vsMotionDetection calls on a
   fresh detector for that clip would return; with confs[c].show > 0 the
   overlays are drawn into the clip's device frames, as those calls would
   draw them.  Up to VIDSTAB_AMD_MULTI_BATCH
   (default 64 when the longest clip has 96 frames or more, else 16) frames
   of every clip go through each device batch.  Clips too small for the
   field grid are refused as vsMotionDetectInit refuses them.  The
   device state of a call is kept for the next call over clips of the same
   shapes (frame sizes, settings), so repeated calls pay the set-up once;
   concurrent calls are allowed (one of them uses the kept state). */
VS_API int vsMotionDetectMulti(const VSMotionDetectConfig* confs, const VSFrameInfo* fis, int nclips,
                               const int* nframes, const uint8_t* const* const* lumaDev, const int* linesize,
                               LocalMotions* const* motions);

/* vs_vector_del on motions[0..n-1]: the vectors are empty when it returns
   (and may be reused); the elements of arrays of 16 or more go back through
   vs_free on a library thread while the caller carries on (one vs_free per
   LocalMotion, about 300 per 1080p frame).  At most four such releases are
   pending at a time; vs_free must not change while any is. */
VS_API void vsLocalMotionsArrayDel(LocalMotions* motions, int n);
/* returns once every release vsLocalMotionsArrayDel handed to the library
   thread is done (every element returned to vs_free) */
VS_API void vsLocalMotionsReleaseWait(void);

#ifdef __cplusplus
}
#endif
#endif /* VIDSTAB_AMD_H */
