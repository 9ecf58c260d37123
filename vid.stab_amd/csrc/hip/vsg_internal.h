/*
 * vsg_internal.h -- kernel launch interface between vsg_api.hip (device
 * memory, streams, the C ABI) and vsg_kernels.hip (the gfx950 kernels).
 */
#ifndef VSG_INTERNAL_H
#define VSG_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vsg.h"
#include "vsg_common.h"

/* one source plane of the blur */
struct VSGBlurSrc {
  const uint8_t* p;
  int pitch;
  int pad;
};

/* Two-pass box blur of n planes (src[i] -> dst + i*slotStride, pitch dstPitch).
   size < 2 copies.  tmp (n*dstPitch*H bytes, plane i at i*dstPitch*H) is used
   only when vsg_blur_needs_tmp(W, size): boxes too wide for the one-kernel
   path.  Launches that may overlap need disjoint tmp ranges. */
hipError_t vsg_launch_blur(const VSGBlurSrc* d_src, int n, uint8_t* dst, size_t slotStride,
                           int dstPitch, int W, int H, int size, uint8_t* tmp, hipStream_t st);
bool vsg_blur_needs_tmp(int W, int size);

/* per-field min/max plus the aligned copy of the field ("patch").  Planar
   frames: (min, max) of the luma; packed frames: (min, max) of each of the
   three colour channels, the fourth byte of a 4-byte pixel left out
   (src/motiondetect.c:370-377) */
struct VSGFieldArgs {
  const uint8_t* blurred;
  size_t slotStride;
  int pitch;
  const int16_t* fx;
  const int16_t* fy;
  const int16_t* fsz;
  const uint32_t* patchOff;   /* byte offset of each field's patch in a slot */
  int nfields;
  uint8_t* patches;
  size_t patchSlotStride;
  uint8_t* minmax;            /* [slot][nfields][nmm] */
  const VSGBlurSrc* src;      /* per slot of the launch; p == null: idle slot (skipped) */
  int bpp;                    /* bytes per pixel */
  int nmm;                    /* min/max bytes per field: 2 (planar) or 6 (packed) */
};
hipError_t vsg_launch_fields(const VSGFieldArgs& a, int nslots, hipStream_t st);
/* bytes (a multiple of 16, 16-byte aligned ends) from src to dst by a kernel;
   src may be pinned host memory (its device address) */
hipError_t vsg_launch_upload(const uint8_t* src, uint8_t* dst, size_t bytes, hipStream_t st);

/* one search stage over a job list */
struct VSGSearchArgs {
  const VSGJob* jobs;
  int njobs;
  const int16_t* fx;
  const int16_t* fy;
  const uint8_t* blurred;
  size_t slotStride;
  const uint8_t* ref;
  int pitch;
  const uint8_t* patches;       /* patch of (slot, field): patches + slot*patchSlotStride
                                   + stageOff + field*S*S*bpp */
  size_t patchSlotStride;
  size_t stageOff;
  VSGGrid g;
  int tiles;                    /* workgroups per job in the grid kernel */
  int chunk;                    /* (job, tile) units kept together on one XCD */
  unsigned long long* best;     /* per job: (sad << 32) | spiral rank; ~0 between launches */
  unsigned* done;               /* per job: grid workgroups finished; 0 between launches */
  VSGResult* results;
};

/* field selection of one stage (selectfields, src/motiondetect.c:676-743) */
struct VSGSelectStage {
  int nf;            /* fields of the stage */
  int rows;          /* fieldRows */
  int maxFields;
  int mmOff;         /* first field of the stage inside a slot's min/max row */
  double ct;         /* contrast threshold */
};
struct VSGSelectArgs {
  const uint8_t* minmax;     /* [slot][nfTotal][nmm] */
  int nfTotal;
  int nmm;                   /* 2: one (min, max) per field; 6: one per colour channel */
  const double* ctab;        /* contrast of (min, max): ctab[min*256 + max] */
  VSGSelectStage st[2];
  int slot0;                 /* first slot of the batch */
  const int* prevSlot;       /* [frame] previous plane (slot or -1) */
  const uint8_t* detect;     /* [frame] 0: no measurement this frame */
  int16_t* selIdx;           /* [frame][nfTotal]: coarse picks, then fine picks */
  double* selC;              /* same layout */
  int* count;                /* [frame][2] */
  VSGJob* jobs;              /* [frame][ncoarse]: coarse jobs, field -1 past the count */
};
hipError_t vsg_launch_select(const VSGSelectArgs& a, int nframes, hipStream_t st);

/* fills the grid and J, D, CW, m, U, V, wg for a stage with K candidate rows
   per lane (4, 6 or 8) on frames of bpp bytes per pixel (1: planar luma);
   returns the workgroups per job */
int vsg_plan_grid(VSGGrid* g, int S, int step, int maxShift, int K, int bpp);
/* bytes of one field's patch: S rows of S*bpp bytes */
static inline size_t vsg_patch_bytes(const VSGGrid& g) { return (size_t)g.S * g.S * g.bpp; }
/* modelled instruction count of one job under a plan (for ranking plans) */
double vsg_plan_cost(const VSGGrid& g, int tiles);
/* grid search (argmin into best[]) fused with the refinement, which the
   last workgroup of each job runs (best[] -> results) */
hipError_t vsg_launch_search(const VSGSearchArgs& a, hipStream_t st);

/* the learned form's proposal window search over a job list (either stage):
   pred holds (dx, dy) per (frame of the batch, field), npred fields per
   frame, frame = job slot - slot0; accepted jobs get their result, the
   index of every other job goes to list[stats[1]++] for the exact search
   (in no particular order); stats[0] += jobs, stats[2..4] += fallbacks out
   of range / window minimum on the edge / above the residual threshold.
   A minimum on the window's edge moves the window there, at most `steps`
   windows in all (1: the first edge minimum falls back).  stats[1] must be
   0 at the launch. */
hipError_t vsg_launch_window(const VSGSearchArgs& a, const float* pred, int npred, int slot0, uint32_t maxSad,
                             int* list, unsigned* stats, int steps, hipStream_t st);
/* the exact search (grid + refinement) of the jobs a.jobs[list[i]],
   i < *count (a device count, at most a.njobs), results to
   a.results[list[i]]; list and count both null: every job of a.jobs.
   Each job's rows are split over four waves (latency over throughput). */
hipError_t vsg_launch_fallback(const VSGSearchArgs& a, const int* list, const unsigned* count, hipStream_t st);
/* vsg_launch_search, or the split kernel for a batch too small to fill the
   GPU (same results) */
hipError_t vsg_launch_search_any(const VSGSearchArgs& a, hipStream_t st);

/* one clip's stage inside one XCD queue of a multi-clip queue launch */
struct VSGMultiSeg {
  VSGSearchArgs a;  /* vsg_queue_args of the clip's stage */
  int unitBase;     /* first unit (workgroup-wide tile) of the segment in its queue */
  int jobBase;      /* the segment's first job in this queue */
  int variant;      /* vsg_queue_variant of a.g */
  int pad;
};
/* units per XCD queue, and the units a pull takes (set by the launcher) */
struct VSGQueueTotals {
  int units[VSG_XCDS];
  int queues;   /* 1, or VSG_XCDS: one queue per XCD */
  int chunk;
};
/* the queue kernel's plan variant of a grid (J, D, CW), or -1 if it has none */
int vsg_queue_variant(const VSGGrid& g);
/* a stage's search arguments as the queue kernel runs them: K = 4, the field
   by scalar loads, K candidate rows per lane (4, 6, 8), tiles of wg lanes
   (64, 128 or 256) */
VSGSearchArgs vsg_queue_args(const VSGSearchArgs& a, int wg, int K);
/* one persistent launch over all units of all segments: dSegs holds
   totals.queues tables of nseg segments, each sorted by unitBase; queues:
   two device counters per queue, zero before the launch and left zero
   after it.  split: the segments' units are 64-lane tiles whose prev rows
   the workgroup's waves share (K must be 4; a.tiles counts 64-lane tiles) */
hipError_t vsg_launch_queue(const VSGMultiSeg* dSegs, int nseg, const VSGQueueTotals& totals, int wg, int K,
                            bool split, unsigned* queues, hipStream_t st);

#endif /* VSG_INTERNAL_H */
