/*
 * vs_motiondetect.c -- the detect-pass driver behind the vid.stab C API
 * (reference: src/motiondetect.c:60-305, 314-358, 752-807).
 *
 * Per batch of frames (one frame for vsMotionDetection):
 *   GPU  (vsg_submit)      upload, box blur (packed RGB: a copy), per-field min/max, field
 *                          selection of both grids, coarse grid search and
 *                          refinement for every field of every frame
 *   host (after the wait)  coarse LocalMotions, vsSimpleMotionsToTransform
 *                          (glibc libm), fine offsets with the border discard
 *   GPU  (vsg_submit_fine) fine search of every frame of the batch
 *   host                   match cleanup, coarse ++ fine output
 * Batches rotate through the GPU context's VSG_NSETS slot sets: batch k+1 is
 * already queued when the host works on batch k's coarse results, and batch
 * k+2 is queued as soon as batch k is finished, so the device is kept busy
 * while the host does its double-precision steps.  Frames of a batch are
 * independent given their previous plane, so the output is exactly that of
 * one vsMotionDetection call per frame.
 *
 * VSMotionDetect is caller-owned with a frozen layout; the GPU context of a
 * detector lives in a side table keyed by the struct's address.
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "localmotion2transform.h"
#include "motiondetect.h"
#include "transformtype_operations.h"
#include "vidstab_amd.h"
#include "vidstabdefines.h"
#include "vs_detect_internal.h"
#include "vs_pool.h"
#include "vsg.h"

#define MD_MAX_BATCH 128

/* ---- per-detector state ------------------------------------------------- */

typedef struct Batch {
  int set, n, slot0;
  int fnum0;                   /* frame number of the batch's first frame */
  uint8_t* draw[MD_MAX_BATCH]; /* show > 0: plane 0 of the caller's frames (null: no overlays) */
  int drawLs[MD_MAX_BATCH];
  int drawDevice;              /* the HIP device holding the frames drawn into, -1: host frames */
  int prevSlot[MD_MAX_BATCH];
  uint8_t detect[MD_MAX_BATCH];
  LocalMotions* out;
  LocalMotions coarse[MD_MAX_BATCH];
  int jobStart[MD_MAX_BATCH + 1];
  int jobCount[MD_MAX_BATCH];
  VSTransform offset[MD_MAX_BATCH];
  int jobCap;                  /* fine jobs per frame (fine field count) */
  VSGJob* jobs;                /* fine jobs of the batch (frame order) */
  int* jobSel;
  int njobs;
  int active;
} Batch;

/* batches in flight: k (cleanup), k+1 (fine stage), k+2 (front and coarse stage) */
#define MD_INFLIGHT 3

/* host phases of a detect call, timed when VIDSTAB_AMD_CALL_TIMES is set */
enum { PH_SUBMIT, PH_STAGE, PH_WAIT_COARSE, PH_COARSE_HOST, PH_SUBMIT_FINE, PH_WAIT_FINE, PH_FINE_HOST, PH_CALL, PH_COUNT };

typedef struct MDState {
  const VSMotionDetect* key;
  struct MDState* next;
  VSGContext* gpu;
  int device;      /* HIP device of gpu */
  VSPool* pool;    /* per-frame host steps of a batch */
  VSPool* stagePool;   /* copies of a few host frames into pinned staging (at most 4 threads) */
  int maxBatch;
  int nextSet;     /* slot set of the next batch */
  int lastSlot;    /* global slot holding the previous frame's blurred plane */
  /* the latest VSG_NSETS batches, newest at histHead: their planes stay on
     the device until their set is submitted again */
  int histSlot0[VSG_NSETS], histN[VSG_NSETS];
  int histCount, histHead;
  /* per-call host state, allocated once: the batches in flight and their
     fine job lists (maxBatch x fine fields each) */
  Batch* bs;
  VSGJob* jobs;
  int* jobSel;
  /* VIDSTAB_AMD_CALL_TIMES: summed host time per phase, logged at cleanup */
  int timed;
  long long calls;
  double phaseUs[PH_COUNT];
} MDState;

static MDState* g_states = NULL;
static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;

static MDState* state_find(const VSMotionDetect* md) {
  pthread_mutex_lock(&g_lock);
  MDState* s = g_states;
  while (s && s->key != md) s = s->next;
  pthread_mutex_unlock(&g_lock);
  return s;
}

static void state_free(MDState* s) {
  if (!s) return;
  if (s->timed && s->calls > 0) {
    static const char* names[PH_COUNT] = {"submit (staging, enqueue)", "  of which host-frame staging",
                                          "wait for coarse results", "coarse host step", "submit fine",
                                          "wait for fine results", "fine host step", "whole call"};
    for (int k = 0; k < PH_COUNT; k++)
      vs_log_info("vidstab_amd", "call time: %-26s %9.1f us per call (%lld calls)\n", names[k],
                  s->phaseUs[k] / (double)s->calls, s->calls);
  }
  if (s->gpu) vsg_destroy(s->gpu);
  vs_pool_destroy(s->pool);
  vs_pool_destroy(s->stagePool);
  free(s->bs);
  free(s->jobs);
  free(s->jobSel);
  free(s);
}

static double now_us(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return 1e6 * (double)ts.tv_sec + 1e-3 * (double)ts.tv_nsec;
}

static void state_remove(const VSMotionDetect* md) {
  pthread_mutex_lock(&g_lock);
  MDState** pp = &g_states;
  MDState* found = NULL;
  while (*pp) {
    if ((*pp)->key == md) {
      found = *pp;
      *pp = found->next;
      break;
    }
    pp = &(*pp)->next;
  }
  pthread_mutex_unlock(&g_lock);
  state_free(found);
}

static void state_insert(MDState* s) {
  pthread_mutex_lock(&g_lock);
  s->next = g_states;
  g_states = s;
  pthread_mutex_unlock(&g_lock);
}

static int env_int(const char* name, int def) {
  const char* v = getenv(name);
  if (!v || !*v) return def;
  char* end = NULL;
  long x = strtol(v, &end, 10);
  return (end && *end == 0) ? (int)x : def;
}

/* ---- configuration ------------------------------------------------------- */

VS_API VSMotionDetectConfig vsMotionDetectGetDefaultConfig(const char* modName) {
  VSMotionDetectConfig conf;
  memset(&conf, 0, sizeof(conf));
  conf.stepSize = 6;
  conf.accuracy = 15;
  conf.shakiness = 5;
  conf.virtualTripod = 0;
  conf.contrastThreshold = 0.25;
  conf.show = 0;
  conf.modName = modName;
  conf.numThreads = 0;
  return conf;
}

VS_API void vsMotionDetectGetConfig(VSMotionDetectConfig* conf, const VSMotionDetect* md) {
  if (md && conf) *conf = md->conf;
}

VS_API const VSFrameInfo* vsMotionDetectGetFrameInfo(const VSMotionDetect* md) { return &md->fi; }

/* src/motiondetect.c:314-358: rows x cols grid of fields kept clear of the
   border by size/2 + maxShift + stepSize */
static int init_fields(VSMotionDetect* md, VSMotionDetectFields* fs, int size, int maxShift,
                       int stepSize, int keepBorder, int spacing, double contrastThreshold) {
  fs->fieldSize = size;
  fs->maxShift = maxShift;
  fs->stepSize = stepSize;
  fs->useOffset = 0;
  fs->contrastThreshold = contrastThreshold;
  const int rows = VS_MAX(3, (md->fi.height - fs->maxShift * 2) / (size + spacing) - 1);
  const int cols = VS_MAX(3, (md->fi.width - fs->maxShift * 2) / (size + spacing) - 1);
  fs->fieldNum = rows * cols;
  fs->fieldRows = rows;
  fs->fields = (Field*)vs_malloc(sizeof(Field) * (size_t)fs->fieldNum);
  if (!fs->fields) {
    vs_log_error(md->conf.modName, "malloc failed!\n");
    return 0;
  }
  const int border = keepBorder ? size / 2 + fs->maxShift + fs->stepSize : fs->stepSize;
  const int step_x = (md->fi.width - 2 * border) / VS_MAX(cols - 1, 1);
  const int step_y = (md->fi.height - 2 * border) / VS_MAX(rows - 1, 1);
  for (int j = 0; j < rows; j++)
    for (int i = 0; i < cols; i++) {
      Field* f = &fs->fields[j * cols + i];
      f->x = (int16_t)(border + i * step_x);
      f->y = (int16_t)(border + j * step_y);
      f->size = (int16_t)size;
    }
  fs->maxFields = md->conf.accuracy * fs->fieldNum / 15;
  vs_log_info(md->conf.modName, "Fieldsize: %i, Maximal translation: %i pixel\n", fs->fieldSize,
              fs->maxShift);
  vs_log_info(md->conf.modName, "Number of used measurement fields: %i out of %i\n", fs->maxFields,
              fs->fieldNum);
  return 1;
}

/* How far a coarse search reads from its field's position: neg px to the
   left / up, pos px to the right / down.  Planar frames: the spiral's n =
   maxShift / step rings, then refinement levels of step/2, step/4, ...
   around the best shift (src/motiondetect.c:466-518, 545-567); packed
   frames: the grid from -maxShift in steps of step, then one level of
   step - 1 (src/motiondetect.c:613-643). */
static void search_reach(int maxShift, int step, int packed, int* neg, int* pos) {
  if (packed) {
    const int r = step > 1 ? step - 1 : 0;
    *neg = maxShift + r;
    *pos = -maxShift + (2 * maxShift / step) * step + r;
    return;
  }
  *neg = *pos = (maxShift / step) * step;
  for (int st = step; st > 1; st /= 2) {
    const int ns = st / 2, r = st - ns;
    *neg += r;
    *pos += -r + (2 * r / ns) * ns;
  }
}

/* 1 when every field of `fs`, shifted up to neg px left / up and pos px
   right / down, lies inside the frame; otherwise logs the first field that
   does not */
static int fields_inside(const VSMotionDetect* md, const VSMotionDetectFields* fs, int neg, int pos,
                         const char* stage) {
  for (int i = 0; i < fs->fieldNum; i++) {
    const Field* f = &fs->fields[i];
    const int lo = f->size / 2 + neg, hi = f->size - f->size / 2 - 1 + pos;
    if (f->x - lo < 0 || f->x + hi > md->fi.width - 1 || f->y - lo < 0 || f->y + hi > md->fi.height - 1) {
      vs_log_error(md->conf.modName,
                   "frame %ix%i is too small: the %s field at (%i, %i), %i px searched %i px left and up and "
                   "%i px right and down, leaves it\n",
                   md->fi.width, md->fi.height, stage, f->x, f->y, f->size, neg, pos);
      return 0;
    }
  }
  return 1;
}

int vsd_blur_size(const VSMotionDetect* md) {
  /* packed frames are searched unblurred (src/motiondetect.c:197-201) */
  if (md->fi.pFormat > PF_PACKED) return 1;
  int size = md->conf.stepSize;
  if (size < 2) return 1;
  size = (size / 2) * 2 + 1;
  return VS_CLAMP(size, 3, VS_MIN(md->fi.height / 2, md->fi.width / 2));
}

VSGStageDesc vsd_stage_desc(const VSMotionDetectFields* fs) {
  VSGStageDesc d;
  d.fieldSize = fs->fieldSize;
  d.maxShift = fs->maxShift;
  d.stepSize = fs->stepSize;
  d.nfields = fs->fieldNum;
  d.fieldRows = fs->fieldRows;
  d.maxFields = fs->maxFields;
  d.contrastThreshold = fs->contrastThreshold;
  d.fields = fs->fields;
  return d;
}

int vsd_configure(VSMotionDetect* md, const VSMotionDetectConfig* conf, const VSFrameInfo* fi) {
  if (!md || !conf || !fi) return VS_ERROR;
  md->conf = *conf;
  md->fi = *fi;
  md->fieldscoarse.fields = NULL;
  md->fieldsfine.fields = NULL;
  if (fi->pFormat <= PF_NONE || fi->pFormat == PF_PACKED || fi->pFormat >= PF_NUMBER) {
    vs_log_warn(md->conf.modName, "unsupported Pixel Format (%i)\n", md->fi.pFormat);
    return VS_ERROR;
  }
  if (md->conf.stepSize < 1) {
    vs_log_error(md->conf.modName, "stepsize must be at least 1\n");
    return VS_ERROR;
  }
  vsFrameNull(&md->prev);
  vsFrameNull(&md->curr);
  vsFrameNull(&md->currorig);
  vsFrameNull(&md->currtmp);
  md->hasSeenOneFrame = 0;
  md->frameNum = 0;
  if (md->serializationMode != ASCII_SERIALIZATION_MODE &&
      md->serializationMode != BINARY_SERIALIZATION_MODE)
    md->serializationMode = BINARY_SERIALIZATION_MODE;

  md->conf.shakiness = VS_MIN(10, VS_MAX(1, md->conf.shakiness));
  md->conf.accuracy = VS_MIN(15, VS_MAX(1, md->conf.accuracy));
  if (md->conf.accuracy < md->conf.shakiness / 2) {
    vs_log_info(md->conf.modName, "Accuracy should not be lower than shakiness/2 -- fixed");
    md->conf.accuracy = md->conf.shakiness / 2;
  }
  if (md->conf.accuracy > 9 && md->conf.stepSize > 6) {
    vs_log_info(md->conf.modName, "For high accuracy use lower stepsize  -- set to 6 now");
    md->conf.stepSize = 6;
  }
  const int minDim = VS_MIN(md->fi.width, md->fi.height);
  const int maxShift = VS_MAX(16, minDim / 7);
  int fieldSize = VS_MAX(16, minDim / 10);
  int fieldSizeFine = VS_MAX(6, minDim / 60);
  /* always rounded up to the next multiple of 16, as the reference does */
  fieldSize = (fieldSize / 16 + 1) * 16;
  fieldSizeFine = (fieldSizeFine / 16 + 1) * 16;
  if (!init_fields(md, &md->fieldscoarse, fieldSize, maxShift, md->conf.stepSize, 1, 0,
                   md->conf.contrastThreshold))
    return VS_ERROR;
  if (!init_fields(md, &md->fieldsfine, fieldSizeFine, fieldSizeFine, 2, 1, fieldSizeFine,
                   md->conf.contrastThreshold / 2))
    return VS_ERROR;
  /* A fine field's contrast is taken at its own position, and its searches
     are dropped per frame when they would leave the frame (vsd_frame_coarse).
     On frames so small that the grid's spacing turns negative
     (src/motiondetect.c:343-346) coarse search windows leave the frame and
     the reference reads outside the frame's rows; such frames are refused. */
  int neg, pos;
  search_reach(maxShift, md->conf.stepSize, md->fi.pFormat > PF_PACKED, &neg, &pos);
  if (!fields_inside(md, &md->fieldscoarse, neg, pos, "coarse") || !fields_inside(md, &md->fieldsfine, 0, 0, "fine"))
    return VS_ERROR;
  return VS_OK;
}

static void free_fields(VSMotionDetect* md) {
  if (md->fieldscoarse.fields) {
    vs_free(md->fieldscoarse.fields);
    md->fieldscoarse.fields = 0;
  }
  if (md->fieldsfine.fields) {
    vs_free(md->fieldsfine.fields);
    md->fieldsfine.fields = 0;
  }
}

VS_API int vsMotionDetectInit(VSMotionDetect* md, const VSMotionDetectConfig* conf, const VSFrameInfo* fi) {
  if (!md || !conf || !fi) return VS_ERROR;
  state_remove(md);   /* a re-initialised detector drops its old device state */
  if (vsd_configure(md, conf, fi) != VS_OK) {
    free_fields(md);
    return VS_ERROR;
  }
  MDState* s = (MDState*)calloc(1, sizeof(MDState));
  if (!s) {
    free_fields(md);
    return VS_ERROR;
  }
  s->key = md;
  s->maxBatch = VS_MIN(MD_MAX_BATCH, VS_MAX(1, env_int("VIDSTAB_AMD_MAX_BATCH", MD_MAX_BATCH)));
  s->lastSlot = -1;
  const int device = env_int("VIDSTAB_AMD_DEVICE", 0);
  VSGStageDesc dc = vsd_stage_desc(&md->fieldscoarse), df = vsd_stage_desc(&md->fieldsfine);
  const int bpp = md->fi.pFormat > PF_PACKED ? md->fi.bytesPerPixel : 1;
  s->device = device;
  if (vsg_create_fmt(&s->gpu, device, md->fi.width, md->fi.height, bpp, vsd_blur_size(md), &dc, &df,
                     s->maxBatch) != 0) {
    vs_log_error(md->conf.modName, "GPU initialisation failed: %s\n", vsg_last_error());
    state_free(s);
    free_fields(md);
    return VS_ERROR;
  }
  /* host threads: numThreads if set, else 0.8 x the online cores as the
     reference does (src/motiondetect.c:98-100), at most 16 */
  int nth = md->conf.numThreads;
  if (nth <= 0) {
    const long cores = sysconf(_SC_NPROCESSORS_ONLN);
    nth = (int)(cores * 0.8);
  }
  nth = VS_MIN(16, VS_MAX(1, env_int("VIDSTAB_AMD_HOST_THREADS", nth)));
  s->pool = vs_pool_create(nth);
  s->stagePool = vs_pool_create(VS_MIN(4, nth));
  const size_t nfj = (size_t)s->maxBatch * (size_t)md->fieldsfine.fieldNum;
  s->bs = (Batch*)calloc(MD_INFLIGHT, sizeof(Batch));
  s->jobs = (VSGJob*)malloc(sizeof(VSGJob) * nfj * MD_INFLIGHT);
  s->jobSel = (int*)malloc(sizeof(int) * nfj * MD_INFLIGHT);
  if (!s->pool || !s->stagePool || !s->bs || !s->jobs || !s->jobSel) {
    vs_log_error(md->conf.modName, "out of host memory\n");
    state_free(s);
    free_fields(md);
    return VS_ERROR;
  }
  for (int k = 0; k < MD_INFLIGHT; k++) {
    s->bs[k].jobs = s->jobs + nfj * k;
    s->bs[k].jobSel = s->jobSel + nfj * k;
    s->bs[k].jobCap = md->fieldsfine.fieldNum;
  }
  s->timed = getenv("VIDSTAB_AMD_CALL_TIMES") != NULL;
  for (int k = 0; k < 2; k++) {
    int pl[7];
    float tm[6];
    if (vsg_get_plan(s->gpu, k, pl, tm) == 0)
      vs_log_info(md->conf.modName,
                  "%s search plan: J=%i K=%i D=%i CW=%i workgroup %i x %i per field, field %s "
                  "(batch timed at K=4/6/8: scalar %.3f/%.3f/%.3f ms, LDS %.3f/%.3f/%.3f ms)\n",
                  k == 0 ? "coarse" : "fine", pl[0], pl[1], pl[2], pl[3], pl[4], pl[5], pl[6] ? "in LDS" : "scalar",
                  tm[0], tm[1], tm[2], tm[3], tm[4], tm[5]);
  }
  vs_log_info(md->conf.modName, "GPU: HIP device %i, batch %i, %i host threads\n", device, s->maxBatch, nth);
  state_insert(s);
  md->initialized = 2;
  return VS_OK;
}

VS_API void vsMotionDetectionCleanup(VSMotionDetect* md) {
  state_remove(md);
  free_fields(md);
  vsFrameFree(&md->prev);
  vsFrameFree(&md->curr);
  vsFrameFree(&md->currtmp);
  md->initialized = 0;
}

/* ---- per-frame host steps ------------------------------------------------------ */

/* src/motiondetect.c:573-585 (planar) and :648-658 (packed).  The planar
   rule reads stepSize after the refinement halved it to 1, hence the
   asymmetric bound; the packed one uses the stage's step on both axes. */
static int make_motion(const VSMotionDetectFields* fs, int packed, const Field* f, const VSGResult* r, int ox,
                       int oy, double contrast, LocalMotion* lm) {
  if (packed) {
    const int lim = fs->maxShift + fs->stepSize - 1;
    if (abs(r->tx) >= lim || abs(r->ty) >= lim) return 0;
  } else {
    int st = fs->stepSize;
    while (st > 1) st /= 2;
    if (abs(r->tx) >= fs->maxShift + st - 1 || abs(r->ty) >= fs->maxShift + st) return 0;
  }
  memset(lm, 0, sizeof(*lm));
  lm->f = *f;
  lm->v.x = (int16_t)(r->tx + ox);
  lm->v.y = (int16_t)(r->ty + oy);
  lm->match = ((double)r->minerror) / (f->size * f->size);
  lm->contrast = contrast;
  return 1;
}

/* src/motiondetect.c:184-189 */
static short lm_match_better(void* thresh, void* lm) {
  return ((LocalMotion*)lm)->match <= *((double*)thresh) ? 1 : 0;
}

int vsd_frame_coarse(const VSMotionDetect* md, const VSGBatchView* v, int i, int fnum, int slot, int prevSlot,
                     LocalMotions* coarse, VSTransform* offset, VSGJob* jobs, int* jobSel) {
  const char* mod = md->conf.modName;
  const VSMotionDetectFields* fc = &md->fieldscoarse;
  const VSMotionDetectFields* ff = &md->fieldsfine;
  vs_vector_init(coarse, fc->maxFields);
  const int16_t* idx = v->idx + (size_t)i * v->stride;
  const double* con = v->contrast + (size_t)i * v->stride;
  const VSGResult* res = v->coarse + (size_t)i * v->ncoarse;
  for (int q = 0; q < v->count[2 * i]; q++) {
    LocalMotion lm;
    if (make_motion(fc, md->fi.pFormat > PF_PACKED, &fc->fields[idx[q]], &res[q], 0, 0, con[q], &lm))
      vs_vector_append_dup(coarse, &lm, sizeof(LocalMotion));
  }
  if (vs_vector_size(coarse) < 1) {
    vs_log_warn(mod, "too low contrast. (no translations are detected in frame %i)\n", fnum);
    return 0;
  }
  /* src/motiondetect.c:238-240, 418-434 */
  *offset = vsSimpleMotionsToTransform(md->fi, mod, coarse);
  PreparedTransform pt = prepare_transform(offset, &md->fi);
  const int16_t* fidx = idx + v->ncoarse;
  int nj = 0;
  for (int q = 0; q < v->count[2 * i + 1]; q++) {
    const Field* f = &ff->fields[fidx[q]];
    Vec pos = {f->x, f->y};
    Vec off = sub_vec(transform_vec(&pt, &pos), pos);
    const int border = f->size / 2 + ff->maxShift + ff->stepSize;
    const int x = pos.x + off.x, y = pos.y + off.y;
    if (x - border < 0 || x + border >= md->fi.width || y - border < 0 || y + border >= md->fi.height)
      continue;
    VSGJob* j = &jobs[nj];
    j->slot = slot;
    j->prev = prevSlot;
    j->field = fidx[q];
    j->ox = off.x;
    j->oy = off.y;
    jobSel[nj] = q;
    nj++;
  }
  return nj;
}

/* the overlays of a frame in device memory: its plane 0 comes down to the
   host, is drawn by the host-frame code and goes back (vsg_copy_plane,
   synchronous; the frame's own kernels finished before its fine results came
   back, and later batches read only their own frames).  The LocalMotions are
   rewritten as for a host frame even if a copy fails. */
static void draw_device_frame(const VSMotionDetect* md, int device, uint8_t* plane, int linesize,
                              LocalMotions* coarse, LocalMotions* fine) {
  const int rowBytes = md->fi.width * (md->fi.pFormat > PF_PACKED ? md->fi.bytesPerPixel : 1);
  const int rows = md->fi.height;
  uint8_t* h = (uint8_t*)calloc((size_t)rowBytes * (size_t)rows, 1);
  if (!h) {
    vs_log_error(md->conf.modName, "show: out of host memory for a device frame\n");
    return;
  }
  const int down = vsg_copy_plane(device, h, rowBytes, plane, linesize, rowBytes, rows) == 0;
  vsd_draw_overlays(md, h, rowBytes, coarse, fine);
  if (!down || vsg_copy_plane(device, plane, linesize, h, rowBytes, rowBytes, rows) != 0)
    vs_log_error(md->conf.modName, "show: device frame copy failed: %s\n", vsg_last_error());
  free(h);
}

void vsd_frame_fine(const VSMotionDetect* md, const VSGBatchView* v, int i, LocalMotions* coarse,
                    const VSGJob* jobs, const int* jobSel, const VSGResult* res, int njobs, uint8_t* draw,
                    int drawLinesize, int drawDevice, LocalMotions* out) {
  const VSMotionDetectFields* ff = &md->fieldsfine;
  LocalMotions fine;
  vs_vector_init(&fine, 0);
  if (vs_vector_size(coarse) >= 1) {
    /* "cleanmotions": keep fine motions no worse than the trimmed mean of the
       coarse matches (src/motiondetect.c:249-258) */
    VSArray mq = localmotionsGetMatch(coarse);
    double meanMatch = cleanmean(mq.dat, mq.len, NULL, NULL);
    vs_array_free(mq);
    const double* fcon = v->contrast + (size_t)i * v->stride + v->ncoarse;
    for (int k = 0; k < njobs; k++) {
      LocalMotion lm;
      if (!make_motion(ff, md->fi.pFormat > PF_PACKED, &ff->fields[jobs[k].field], &res[k], jobs[k].ox, jobs[k].oy,
                       fcon[jobSel[k]], &lm))
        continue;
      if (lm_match_better(&meanMatch, &lm)) vs_vector_append_dup(&fine, &lm, sizeof(LocalMotion));
    }
  }
  if (draw && md->conf.show > 0) {
    if (drawDevice >= 0)
      draw_device_frame(md, drawDevice, draw, drawLinesize, coarse, &fine);
    else
      vsd_draw_overlays(md, draw, drawLinesize, coarse, &fine);
  }
  *out = vs_vector_concat(coarse, &fine);
  vs_vector_fini(coarse);
  vs_vector_fini(&fine);
}

/* ---- batches ------------------------------------------------------------------ */


/* decides, for the next n frames, which are measured and against which plane
   (src/motiondetect.c:211-218, 294-303), and submits them */
/* host frames into the pinned staging buffers, row bands of rows
   [y0, y0 + rows) of the frames spread over the pool */
typedef struct {
  const uint8_t* const* luma;
  const int* linesize;
  uint8_t** stage;
  int pitch, rowBytes, rows, bands, y0;
} StageWork;

static void stage_fn(void* ctx, int t) {
  const StageWork* w = (const StageWork*)ctx;
  const int i = t / w->bands, band = t - i * w->bands;
  const int r0 = w->y0 + (int)((long long)w->rows * band / w->bands);
  const int r1 = w->y0 + (int)((long long)w->rows * (band + 1) / w->bands);
  const uint8_t* src = w->luma[i];
  uint8_t* dst = w->stage[i];
  if (w->linesize[i] == w->pitch) {
    memcpy(dst + (size_t)r0 * w->pitch, src + (size_t)r0 * w->pitch,
           (size_t)(r1 - r0 - 1) * w->pitch + (size_t)w->rowBytes);
  } else {
    for (int y = r0; y < r1; y++)
      memcpy(dst + (size_t)y * w->pitch, src + (size_t)y * w->linesize[i], (size_t)w->rowBytes);
  }
}

/* chunks a frame's rows are staged and sent up in when a few frames are
   staged: each chunk's upload (a copy kernel on the compute queue) runs
   while the host stages the next chunk.  One vsMotionDetection call per
   1080p host frame (C driver): 2 chunks 3,298-3,303 frames/s, 4 chunks
   3,177-3,340, 8 chunks 3,124-3,133; one upload after the whole frame
   3,180-3,250 */
#define STAGE_CHUNKS 2

/* copies the n host frames into the set's pinned staging buffers; returns
   the vsg_submit mode for them: VSG_FRAMES_UPLOADED when this already sent
   them up (a few frames, in row chunks), else VSG_FRAMES_STAGED; -1 on
   failure */
static int stage_host_frames(VSMotionDetect* md, MDState* s, int set, int n, const uint8_t* const* luma,
                             const int* linesize) {
  uint8_t* stage[MD_MAX_BATCH];
  int pitch = 0;
  if (vsg_stage_frames(s->gpu, set, n, stage, &pitch) != 0) return -1;
  const int bpp = md->fi.pFormat > PF_PACKED ? md->fi.bytesPerPixel : 1;
  StageWork w = {luma, linesize, stage, pitch, md->fi.width * bpp, md->fi.height, 1, 0};
  if (n >= 4) {
    if (w.bands > w.rows) w.bands = w.rows;
    vs_pool_run(s->pool, n * w.bands, stage_fn, &w);
    return VSG_FRAMES_STAGED;
  }
  /* a few frames: their rows split over a pool of its own, of at most four
     threads.  The copy is bound by host memory bandwidth, not by threads:
     one 1080p plane took 51 us on such a pool, 76-85 us on the 12 threads of
     the batch pool (threads that only wait slow down the ones that copy).
     Frame by frame, chunk by chunk: the chunk's rows of the frame, then its
     upload */
  const size_t plane = (size_t)pitch * (size_t)md->fi.height;
  const int rows = md->fi.height;
  for (int i = 0; i < n; i++) {
    StageWork f = w;
    f.luma = luma + i;
    f.linesize = linesize + i;
    f.stage = stage + i;
    for (int c = 0; c < STAGE_CHUNKS; c++) {
      f.y0 = (int)((long long)rows * c / STAGE_CHUNKS);
      f.rows = (int)((long long)rows * (c + 1) / STAGE_CHUNKS) - f.y0;
      if (f.rows <= 0) continue;
      f.bands = VS_MIN(vs_pool_threads(s->stagePool), f.rows);
      vs_pool_run(s->stagePool, f.bands, stage_fn, &f);
      if (vsg_stage_upload(s->gpu, set, plane * (size_t)i + (size_t)f.y0 * (size_t)pitch,
                           (size_t)f.rows * (size_t)pitch) != 0)
        return -1;
    }
  }
  return VSG_FRAMES_UPLOADED;
}

static int batch_submit(VSMotionDetect* md, MDState* s, Batch* b, const uint8_t* const* luma,
                        const int* linesize, int onDevice, uint8_t* const* draw, int n, LocalMotions* out) {
  const int tripod = md->conf.virtualTripod;
  b->set = s->nextSet;
  s->nextSet = (s->nextSet + 1) % VSG_NSETS;
  b->n = n;
  b->slot0 = b->set * s->maxBatch;
  b->fnum0 = md->frameNum;
  b->out = out;
  b->njobs = 0;
  int refSlot = -1;
  int seen = md->hasSeenOneFrame;
  for (int i = 0; i < n; i++) {
    const int fnum = md->frameNum + i;
    const int leadIn = tripod >= 1 && fnum < tripod;
    b->detect[i] = (uint8_t)(seen && !leadIn);
    if (!b->detect[i]) seen = 1;
    /* tripod: every measured frame compares against frame tripod-1, kept as
       the stored reference plane; otherwise against the frame before */
    b->prevSlot[i] = tripod >= 1 ? -1 : (i == 0 ? s->lastSlot : b->slot0 + i - 1);
    b->draw[i] = draw ? draw[i] : NULL;
    b->drawLs[i] = linesize[i];
    b->drawDevice = onDevice ? s->device : -1;
    if (tripod >= 1 && fnum == tripod - 1) refSlot = b->slot0 + i;
  }
  int mode = VSG_FRAMES_DEVICE;
  const double t0 = s->timed ? now_us() : 0;
  if (!onDevice) mode = stage_host_frames(md, s, b->set, n, luma, linesize);
  if (s->timed) s->phaseUs[PH_STAGE] += now_us() - t0;
  if (mode < 0 || vsg_submit(s->gpu, b->set, n, luma, linesize, mode, b->prevSlot, b->detect, refSlot) != 0) {
    vs_log_error(md->conf.modName, "GPU detect failed: %s\n", vsg_last_error());
    return VS_ERROR;
  }
  md->hasSeenOneFrame = (short)seen;
  md->frameNum += n;
  s->lastSlot = b->slot0 + n - 1;
  s->histHead = (s->histHead + 1) % VSG_NSETS;
  s->histSlot0[s->histHead] = b->slot0;
  s->histN[s->histHead] = n;
  if (s->histCount < VSG_NSETS) s->histCount++;
  b->active = 1;
  return VS_OK;
}

typedef struct {
  VSMotionDetect* md;
  Batch* b;
  const VSGBatchView* v;
  const VSGResult* res;
} BatchWork;

/* frame i: coarse LocalMotions, transform and fine jobs, into the frame's own
   region of the job list */
static void frame_coarse_fn(void* ctx, int i) {
  BatchWork* w = (BatchWork*)ctx;
  Batch* b = w->b;
  vs_vector_init(&b->coarse[i], 0);
  b->jobCount[i] = 0;
  if (!b->detect[i]) return;
  const size_t at = (size_t)i * b->jobCap;
  b->jobCount[i] = vsd_frame_coarse(w->md, w->v, i, b->fnum0 + i, b->slot0 + i, b->prevSlot[i], &b->coarse[i],
                                    &b->offset[i], b->jobs + at, b->jobSel + at);
}

static void frame_fine_fn(void* ctx, int i) {
  BatchWork* w = (BatchWork*)ctx;
  Batch* b = w->b;
  if (!b->detect[i]) {
    vs_vector_fini(&b->coarse[i]);
    vs_vector_init(&b->out[i], 1);   /* the reference's empty "dummy" vector */
    return;
  }
  const int j0 = b->jobStart[i];
  vsd_frame_fine(w->md, w->v, i, &b->coarse[i], b->jobs + j0, b->jobSel + j0, w->res + j0,
                 b->jobStart[i + 1] - j0, b->draw[i], b->drawLs[i], b->drawDevice, &b->out[i]);
}

/* coarse results -> coarse LocalMotions, transform, fine jobs; submits them */
static int batch_coarse(VSMotionDetect* md, MDState* s, Batch* b) {
  VSGBatchView v;
  double t = s->timed ? now_us() : 0;
  if (vsg_wait_coarse(s->gpu, b->set, &v) != 0) {
    vs_log_error(md->conf.modName, "GPU detect failed: %s\n", vsg_last_error());
    return VS_ERROR;
  }
  if (s->timed) {
    const double t2 = now_us();
    s->phaseUs[PH_WAIT_COARSE] += t2 - t;
    t = t2;
  }
  BatchWork w = {md, b, &v, NULL};
  vs_pool_run(s->pool, b->n, frame_coarse_fn, &w);
  /* pack the per-frame job regions into one list (destinations never pass
     their sources, so front-to-back moves are safe) */
  int nj = 0, last = -1;
  for (int i = 0; i < b->n; i++) {
    const size_t at = (size_t)i * b->jobCap;
    b->jobStart[i] = nj;
    if (b->jobCount[i] > 0 && (size_t)nj != at) {
      memmove(b->jobs + nj, b->jobs + at, sizeof(VSGJob) * (size_t)b->jobCount[i]);
      memmove(b->jobSel + nj, b->jobSel + at, sizeof(int) * (size_t)b->jobCount[i]);
    }
    nj += b->jobCount[i];
    if (vs_vector_size(&b->coarse[i]) >= 1) last = i;
  }
  b->jobStart[b->n] = nj;
  b->njobs = nj;
  if (last >= 0) {
    /* as after the reference's frame: the last transform stays in fieldsfine */
    md->fieldsfine.offset = b->offset[last];
    md->fieldsfine.useOffset = 1;
  }
  if (s->timed) {
    const double t2 = now_us();
    s->phaseUs[PH_COARSE_HOST] += t2 - t;
    t = t2;
  }
  if (vsg_submit_fine(s->gpu, b->set, nj, b->jobs) != 0) {
    vs_log_error(md->conf.modName, "GPU fine search failed: %s\n", vsg_last_error());
    return VS_ERROR;
  }
  if (s->timed) s->phaseUs[PH_SUBMIT_FINE] += now_us() - t;
  return VS_OK;
}

/* fine results -> cleanup and output (src/motiondetect.c:249-291) */
static int batch_fine(VSMotionDetect* md, MDState* s, Batch* b) {
  const VSGResult* res = NULL;
  VSGBatchView v;
  double t = s->timed ? now_us() : 0;
  if (vsg_wait_fine(s->gpu, b->set, &res) != 0 || vsg_wait_coarse(s->gpu, b->set, &v) != 0) {
    vs_log_error(md->conf.modName, "GPU fine search failed: %s\n", vsg_last_error());
    return VS_ERROR;
  }
  if (s->timed) {
    const double t2 = now_us();
    s->phaseUs[PH_WAIT_FINE] += t2 - t;
    t = t2;
  }
  BatchWork w = {md, b, &v, res};
  vs_pool_run(s->pool, b->n, frame_fine_fn, &w);
  b->active = 0;
  if (s->timed) s->phaseUs[PH_FINE_HOST] += now_us() - t;
  return VS_OK;
}

/* batch_submit, timed */
static int batch_submit_t(VSMotionDetect* md, MDState* s, Batch* b, const uint8_t* const* luma, const int* linesize,
                          int onDevice, uint8_t* const* draw, int n, LocalMotions* out) {
  const double t = s->timed ? now_us() : 0;
  const int r = batch_submit(md, s, b, luma, linesize, onDevice, draw, n, out);
  if (s->timed) s->phaseUs[PH_SUBMIT] += now_us() - t;
  return r;
}

/* Runs n frames through as few batches of at most maxBatch as there can be,
   all of one size but the last (300 frames: 3 x 100, not 2 x 128 + 44 --
   1080p 31,250 frames/s against 29,790-29,880 with 4 x 64 + 44; 4 x 75:
   29,440-29,740), three in flight: batch k+2's
   front stages and coarse search are queued, then batch k+1's coarse results
   become its fine jobs, and only then does the host wait for batch k's fine
   search and clean it up -- so the device always holds the next batch's
   fine search and the batch after next. */
static int detect_frames(VSMotionDetect* md, LocalMotions* motions, const uint8_t* const* luma,
                         const int* linesize, int onDevice, uint8_t* const* draw, int n) {
  if (!md || md->initialized != 2 || !motions || n < 0) return VS_ERROR;
  MDState* s = state_find(md);
  if (!s) {
    vs_log_error(md->conf.modName, "detector has no GPU state (not initialised?)\n");
    return VS_ERROR;
  }
  if (n == 0) return VS_OK;
  const double tcall = s->timed ? now_us() : 0;
  const int nb = (n + s->maxBatch - 1) / s->maxBatch;
  const int B = (n + nb - 1) / nb;   /* <= maxBatch, and (nb - 1) * B < n */
  Batch* bs = s->bs;
  for (int k = 0; k < MD_INFLIGHT; k++) bs[k].active = 0;
  int ok = 1, done = 0;
  for (int k = 0; k < nb && k < 2 && ok; k++) {
    const int f0 = k * B, cnt = VS_MIN(B, n - f0);
    ok = batch_submit_t(md, s, &bs[k], luma + f0, linesize + f0, onDevice, draw ? draw + f0 : NULL, cnt,
                        motions + f0) == VS_OK;
  }
  /* batch k's coarse results are turned into its fine jobs before the host
     waits for batch k-1's fine results, so the device gets batch k's fine
     stage as soon as its coarse stage is back (and the last batch's does not
     wait behind the previous batch's cleanup) */
  if (ok && nb > 0 && batch_coarse(md, s, &bs[0]) != VS_OK) ok = 0;
  for (int k = 0; k < nb && ok; k++) {
    Batch* b = &bs[k % MD_INFLIGHT];
    if (k + 2 < nb) {
      /* the slot set of batch k-1, whose fine stage is done */
      const int f0 = (k + 2) * B, cnt = VS_MIN(B, n - f0);
      if (batch_submit_t(md, s, &bs[(k + 2) % MD_INFLIGHT], luma + f0, linesize + f0, onDevice,
                         draw ? draw + f0 : NULL, cnt, motions + f0) != VS_OK) {
        ok = 0;
        break;
      }
    }
    if (k + 1 < nb && batch_coarse(md, s, &bs[(k + 1) % MD_INFLIGHT]) != VS_OK) {
      ok = 0;
      break;
    }
    if (batch_fine(md, s, b) != VS_OK) {
      ok = 0;
      break;
    }
    done += b->n;
  }
  if (!ok) {
    for (int k = 0; k < MD_INFLIGHT; k++)
      if (bs[k].active)
        for (int i = 0; i < bs[k].n; i++) vs_vector_del(&bs[k].coarse[i]);
    for (int i = 0; i < done; i++) vs_vector_del(&motions[i]);
  }
  if (s->timed) {
    s->phaseUs[PH_CALL] += now_us() - tcall;
    s->calls++;
  }
  return ok ? VS_OK : VS_ERROR;
}

VS_API int vsMotionDetection(VSMotionDetect* md, LocalMotions* motions, VSFrame* frame) {
  if (!md || !frame || md->initialized != 2) return VS_ERROR;
  md->currorig = *frame;
  const uint8_t* luma = frame->data[0];
  uint8_t* draw = frame->data[0];
  int ls = frame->linesize[0];
  return detect_frames(md, motions, &luma, &ls, 0, md->conf.show > 0 ? &draw : NULL, 1);
}

VS_API int vsMotionDetectionBatch(VSMotionDetect* md, LocalMotions* motions, VSFrame* frames, int n) {
  if (!md || !frames || n < 0) return VS_ERROR;
  if (n == 0) return VS_OK;
  const uint8_t** luma = (const uint8_t**)malloc(sizeof(uint8_t*) * (size_t)n);
  uint8_t** draw = (uint8_t**)malloc(sizeof(uint8_t*) * (size_t)n);
  int* ls = (int*)malloc(sizeof(int) * (size_t)n);
  if (!luma || !draw || !ls) {
    free(luma);
    free(draw);
    free(ls);
    return VS_ERROR;
  }
  for (int i = 0; i < n; i++) {
    luma[i] = frames[i].data[0];
    draw[i] = frames[i].data[0];
    ls[i] = frames[i].linesize[0];
  }
  md->currorig = frames[n - 1];
  int r = detect_frames(md, motions, luma, ls, 0, md->conf.show > 0 ? draw : NULL, n);
  free(luma);
  free(draw);
  free(ls);
  return r;
}

VS_API int vsMotionDetectionDevice(VSMotionDetect* md, LocalMotions* motions, const uint8_t* const* lumaDev,
                                   const int* linesize, int n) {
  if (n == 0) return VS_OK;
  /* show > 0: the overlays are drawn into the device frames, as into a host
     frame by vsMotionDetection */
  return detect_frames(md, motions, lumaDev, linesize, 1, md->conf.show > 0 ? (uint8_t* const*)lumaDev : NULL, n);
}

VS_API int vsMotionDetectSetProposal(VSMotionDetect* md, const char* weightsFile) {
  MDState* s = state_find(md);
  if (!s || !weightsFile) return VS_ERROR;
  FILE* f = fopen(weightsFile, "rb");
  if (!f) {
    vs_log_error(md->conf.modName, "cannot open proposal weights %s\n", weightsFile);
    return VS_ERROR;
  }
  char magic[4];
  int32_t hdr[2], dils[8];
  float thr = 0;
  float* w = NULL;
  int ok = fread(magic, 1, 4, f) == 4 && memcmp(magic, "VSGP", 4) == 0 && fread(hdr, 4, 2, f) == 2 &&
           hdr[0] == 1 && hdr[1] >= 1 && hdr[1] <= 8;
  const size_t perLayer = 9 * 32 * 32 + 32;
  if (ok) ok = fread(dils, 4, (size_t)hdr[1], f) == (size_t)hdr[1] && fread(&thr, 4, 1, f) == 1;
  if (ok) {
    w = (float*)malloc(sizeof(float) * perLayer * (size_t)hdr[1]);
    ok = w && fread(w, sizeof(float), perLayer * (size_t)hdr[1], f) == perLayer * (size_t)hdr[1];
  }
  fclose(f);
  if (!ok) {
    vs_log_error(md->conf.modName, "bad proposal weights file %s\n", weightsFile);
    free(w);
    return VS_ERROR;
  }
  const int rc = vsg_set_proposal(s->gpu, hdr[1], dils, w, thr);
  free(w);
  if (rc != 0) {
    vs_log_error(md->conf.modName, "proposal network: %s\n", vsg_last_error());
    return VS_ERROR;
  }
  vs_log_info(md->conf.modName, "learned proposals: %i layers, residual threshold %.3f\n", hdr[1], thr);
  return VS_OK;
}

VS_API int vsMotionDetectProposalStageStats(const VSMotionDetect* md, int stage, long long* fields,
                                            long long* fallbacks) {
  MDState* s = state_find(md);
  if (!s) return VS_ERROR;
  return vsg_proposal_stage_stats(s->gpu, stage, fields, fallbacks) == 0 ? VS_OK : VS_ERROR;
}

VS_API int vsMotionDetectProposalStats(const VSMotionDetect* md, long long* fields, long long* fallbacks) {
  MDState* s = state_find(md);
  if (!s) return VS_ERROR;
  if (vsg_proposal_stats(s->gpu, fields, fallbacks) != 0) return VS_ERROR;
  if (getenv("VIDSTAB_AMD_PROPOSAL_LOG")) {
    for (int stage = 0; stage < 2; ++stage) {
      long long r = 0, e = 0, t = 0;
      vsg_proposal_fallbacks(s->gpu, stage, &r, &e, &t);
      vs_log_info(md->conf.modName, "%s proposal fallbacks: %lld out of range, %lld on the window edge, %lld "
                  "above the residual threshold\n", stage ? "fine" : "coarse", r, e, t);
    }
  }
  return VS_OK;
}

VS_API int vsMotionDetectGetBlurred(const VSMotionDetect* md, uint8_t* dst, int dstStride) {
  return vsMotionDetectGetBlurredAt(md, 0, dst, dstStride);
}

VS_API int vsMotionDetectGetBlurredAt(const VSMotionDetect* md, int back, uint8_t* dst, int dstStride) {
  MDState* s = state_find(md);
  if (!s || s->lastSlot < 0 || back < 0) return VS_ERROR;
  for (int k = 0, h = s->histHead; k < s->histCount; k++, h = (h + VSG_NSETS - 1) % VSG_NSETS) {
    if (back < s->histN[h])
      return vsg_read_blurred(s->gpu, s->histSlot0[h] + s->histN[h] - 1 - back, dst, dstStride) == 0 ? VS_OK
                                                                                                      : VS_ERROR;
    back -= s->histN[h];
  }
  return VS_ERROR;   /* that frame's plane has been overwritten */
}
