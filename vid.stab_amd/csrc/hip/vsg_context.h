/*
 * vsg_context.h -- the device state of one detector (one clip): buffers,
 * the slot sets, their streams and their pinned host twins.  Shared by the single-clip
 * ABI (vsg_api.hip) and the multi-clip batched form (vsg_multi.hip).
 */
#ifndef VSG_CONTEXT_H
#define VSG_CONTEXT_H

#include "vsg_internal.h"

enum { VSG_EV_BLUR0, VSG_EV_BLUR1, VSG_EV_FIELDS1, VSG_EV_SELECT1, VSG_EV_GRID0, VSG_EV_GRID1, VSG_EV_COUNT };
enum { VSG_EVF_0, VSG_EVF_GRID1, VSG_EVF_COUNT };

struct VSGStage {
  int S = 0, maxShift = 0, step = 0, nf = 0, rows = 0, maxFields = 0;
  double ct = 0;
  int16_t* d_fx = nullptr;
  int16_t* d_fy = nullptr;
  VSGGrid g{};
  int tiles = 0;
  size_t patchOff = 0;
  float tuneMs[6] = {0, 0, 0, 0, 0, 0};   /* timed plans K = 4, 6, 8 with scalar, then LDS patch (0: not timed) */
  double candsPerJob = 0;     /* (2n+1)^2 */
  double refinePerJob = 0;    /* refinement candidates the reference visits */
};

struct VSGSet {
  /* device */
  int16_t* d_selIdx = nullptr;
  double* d_selC = nullptr;
  int* d_count = nullptr;
  VSGJob* d_jobsC = nullptr;
  VSGJob* d_jobsF = nullptr;
  unsigned long long* d_best = nullptr;
  unsigned* d_done = nullptr;
  VSGResult* d_resC = nullptr;
  VSGResult* d_resF = nullptr;
  uint8_t* d_desc = nullptr;     /* d_src, d_prev, d_detect in one block */
  VSGBlurSrc* d_src = nullptr;
  int* d_prev = nullptr;
  uint8_t* d_detect = nullptr;
  /* pinned host */
  int16_t* h_selIdx = nullptr;
  double* h_selC = nullptr;
  int* h_count = nullptr;
  VSGResult* h_resC = nullptr;
  VSGJob* h_jobsF = nullptr;
  VSGResult* h_resF = nullptr;
  uint8_t* h_desc = nullptr;     /* h_src, h_prev, h_detect in one block */
  VSGBlurSrc* h_src = nullptr;
  int* h_prev = nullptr;
  uint8_t* h_detect = nullptr;
  uint8_t* h_stage = nullptr;   /* pinned staging of host frames (B x pitch x H), allocated on first use */
  /* device addresses of the pinned buffers above: a small batch's kernels
     read its descriptors and fine jobs and write its selection and results
     there directly, with no copies (zc) */
  int16_t* m_selIdx = nullptr;
  double* m_selC = nullptr;
  int* m_count = nullptr;
  VSGResult* m_resC = nullptr;
  VSGJob* m_jobsF = nullptr;
  VSGResult* m_resF = nullptr;
  uint8_t* m_desc = nullptr;
  uint8_t* m_stage = nullptr;
  bool zc = false;         /* the batch in flight crosses PCIe through the mapped buffers */
  /* state */
  hipStream_t stream = nullptr;   /* all device work of the set runs in order here (by default one
                                     stream shared by the sets: one batch stage at a time) */
  hipEvent_t uploaded = nullptr;  /* the batch's host frames are on the device (upload stream) */
  int n = 0;              /* frames in flight */
  int njobsF = 0;
  int follower = -1;      /* set submitted right after this set's latest batch */
  hipEvent_t done = nullptr, doneF = nullptr;   /* coarse / fine read-back complete */
  hipEvent_t planesReady = nullptr;   /* this and every earlier batch's blurred planes written */
  hipEvent_t searchDone = nullptr;    /* latest search of the set (coarse, then fine) finished */
  hipEvent_t ev[VSG_EV_COUNT] = {};
  hipEvent_t evf[VSG_EVF_COUNT] = {};
  bool prof = false, profF = false;
  bool borrowed = false;   /* the transfer buffers point into a multi-clip arena (not freed here) */
};

struct VSGLearned;   /* vsg_learned.hip */

struct VSGContext {
  int device = 0;
  bool ownStreams = true;
  hipStream_t upStream = nullptr;   /* host-frame uploads of large batches, beside the compute stream */
  int lastSet = -1;       /* set of the latest batch */
  bool oneStream = true;  /* every set runs on the same stream: stream order alone orders the batches */
  int W = 0, H = 0, pitch = 0, rowsAlloc = 0;
  int bpp = 1;            /* bytes per pixel: 1 planar luma, 3 or 4 packed */
  int Wb = 0;             /* bytes per row of a plane: W * bpp */
  int nmm = 2;            /* min/max bytes per field (6: three packed channels) */
  size_t slotStride = 0;
  int blurSize = 0;
  int B = 0;
  VSGStage st[2];
  int nfTotal = 0;
  int16_t* d_allx = nullptr;
  int16_t* d_ally = nullptr;
  int16_t* d_allsz = nullptr;
  uint32_t* d_patchOff = nullptr;
  double* d_ctab = nullptr;
  uint8_t* d_raw = nullptr;
  uint8_t* d_blur = nullptr;
  uint8_t* d_ref = nullptr;
  uint8_t* d_tmp = nullptr;
  uint8_t* d_patches = nullptr;
  size_t patchSlotStride = 0;
  uint8_t* d_minmax = nullptr;
  VSGSet set[VSG_NSETS];
  VSGLearned* learned = nullptr;   /* proposal network of the learned form, or none */
};

/* creation / destruction shared by both ABIs (vsg_api.hip); a null `shared`
   creates one private stream per set, else set k runs on shared[k] */
int vsg_ctx_init(VSGContext* c, int device, int width, int height, int bpp, int blurSize, const VSGStageDesc* coarse,
                 const VSGStageDesc* fine, int maxBatch, const hipStream_t* shared);
/* all streams of the context idle */
int vsg_ctx_sync(VSGContext* c);
/* plan with the lowest modelled cost; returns workgroups per job */
int vsg_plan_model(VSGGrid* g, int S, int step, int maxShift, int bpp);
/* picks each stage's plan (vsg_tune.hip) */
int vsg_ctx_tune(VSGContext* c);
/* the learned form (vsg_learned.hip): quarter-resolution planes in the
   front stage, the proposal coarse and fine stages, their statistics,
   teardown */
int vsg_learned_front(VSGContext* c, int set, int n, int refSlot, hipStream_t st);
bool vsg_learned_net_busy(VSGContext* c);
int vsg_learned_net(VSGContext* c, int set, int n, hipStream_t st);
int vsg_learned_coarse(VSGContext* c, int set, int n, hipStream_t st);
bool vsg_learned_fine_enabled(const VSGContext* c);
int vsg_learned_fine(VSGContext* c, int set, int njobs, hipStream_t st);
void vsg_learned_account(VSGContext* c, int set, int stage);
void vsg_learned_free(VSGContext* c);
void vsg_ctx_free(VSGContext* c);
/* frees a set's buffers that cross PCIe (descriptors, selection and coarse
   results, fine jobs and results; device and pinned host) */
void vsg_ctx_release_transfer(VSGContext* c, int set);
/* enqueue helpers (no host synchronisation) */
int vsg_ctx_upload_front(VSGContext* c, int set, int n);       /* H2D frame descriptors */
/* (H2D descriptors unless upload is false), blur, fields, select */
int vsg_ctx_enqueue_front(VSGContext* c, int set, int n, bool upload = true);
int vsg_ctx_enqueue_coarse_readback(VSGContext* c, int set, int n);
VSGSearchArgs vsg_ctx_search_args(VSGContext* c, int stage, const VSGJob* jobs, int njobs,
                                  unsigned long long* best, unsigned* done, VSGResult* res);
int vsg_fail(const char* what, hipError_t e);

#endif /* VSG_CONTEXT_H */
