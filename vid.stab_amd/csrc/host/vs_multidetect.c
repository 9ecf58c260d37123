/*
 * vs_multidetect.c -- host driver of the batched multi-clip form
 * (vsMotionDetectMulti, declared in vidstab_amd.h).
 *
 * Each clip is an independent detect stream with its own VSMotionDetect
 * state (configured exactly as vsMotionDetectInit does, src/motiondetect.c:
 * 83-165) and its own previous-frame chain, or, with virtualTripod >= 1, its
 * own frozen reference frame.  A batch takes up to B frames of
 * every clip; batches rotate through the VSG_NSETS slot sets in the same
 * three-in-flight order as the single-clip driver (vs_motiondetect.c).  The
 * per-frame host steps (coarse LocalMotions, vsSimpleMotionsToTransform,
 * fine job lists, match cleanup) are the single-clip ones and run over all
 * (clip, frame) pairs of a batch on a thread pool.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "motiondetect.h"
#include "vidstab_amd.h"
#include "vidstabdefines.h"
#include "vs_detect_internal.h"
#include "vs_pool.h"
#include "vsg.h"

#define MULTI_MAX_B 64
#define MULTI_INFLIGHT 3

typedef struct {
  VSMotionDetect md;
  int lastSlot;          /* slot of the previous frame, -1 before the first */
  int next;              /* next frame to submit */
  int jobCap;            /* fine fields */
} MClip;

typedef struct {
  int n, fnum0, slot0;
  int prevSlot[MULTI_MAX_B];
  uint8_t detect[MULTI_MAX_B];
  const uint8_t* luma[MULTI_MAX_B];
  LocalMotions coarse[MULTI_MAX_B];
  VSTransform offset[MULTI_MAX_B];
  int jobStart[MULTI_MAX_B + 1];
  int jobCount[MULTI_MAX_B];
  VSGJob* jobs;
  int* jobSel;
  VSGBatchView view;
  const VSGResult* res;
} MClipBatch;

typedef struct {
  int set;
  int active;
  MClipBatch* cb;        /* [nclips] */
  int* itemClip;         /* (clip, frame) items of the batch */
  int* itemFrame;
  int nitems;
} MBatch;

typedef struct {
  int nclips, B;
  int device;            /* the HIP device of the clips' frames */
  MClip* clips;
  VSGMulti* gpu;
  VSPool* pool;
  LocalMotions* const* motions;
  const uint8_t* const* const* luma;
  const int* nframes;
  const int* linesize;
  /* per clip scratch for the multi ABI */
  int* n;
  const uint8_t* const** rows;
  const int** prevs;
  const uint8_t** dets;
  int* refs;
  const VSGJob** jobs;
} Multi;

typedef struct {
  Multi* M;
  MBatch* b;
} Work;

static int env_int(const char* name, int def) {
  const char* v = getenv(name);
  if (!v || !*v) return def;
  char* end = NULL;
  long x = strtol(v, &end, 10);
  return (end && *end == 0) ? (int)x : def;
}

static void coarse_fn(void* ctx, int k) {
  Work* w = (Work*)ctx;
  const int c = w->b->itemClip[k], i = w->b->itemFrame[k];
  MClip* cl = &w->M->clips[c];
  MClipBatch* cb = &w->b->cb[c];
  vs_vector_init(&cb->coarse[i], 0);
  cb->jobCount[i] = 0;
  if (!cb->detect[i]) return;
  const size_t at = (size_t)i * cl->jobCap;
  cb->jobCount[i] = vsd_frame_coarse(&cl->md, &cb->view, i, cb->fnum0 + i, cb->slot0 + i, cb->prevSlot[i],
                                     &cb->coarse[i], &cb->offset[i], cb->jobs + at, cb->jobSel + at);
}

static void fine_fn(void* ctx, int k) {
  Work* w = (Work*)ctx;
  const int c = w->b->itemClip[k], i = w->b->itemFrame[k];
  MClip* cl = &w->M->clips[c];
  MClipBatch* cb = &w->b->cb[c];
  LocalMotions* out = &w->M->motions[c][cb->fnum0 + i];
  if (!cb->detect[i]) {
    vs_vector_fini(&cb->coarse[i]);
    vs_vector_init(out, 1);   /* the reference's empty "dummy" vector */
    return;
  }
  const int j0 = cb->jobStart[i];
  /* show > 0: the overlays go into the clip's device frame, as into a host
     frame by vsMotionDetection */
  uint8_t* draw = cl->md.conf.show > 0 ? (uint8_t*)cb->luma[i] : NULL;
  vsd_frame_fine(&cl->md, &cb->view, i, &cb->coarse[i], cb->jobs + j0, cb->jobSel + j0, cb->res + j0,
                 cb->jobStart[i + 1] - j0, draw, w->M->linesize[c], w->M->device, out);
}

/* next frames of every clip into set `set` */
static int batch_submit(Multi* M, MBatch* b, int set) {
  b->set = set;
  b->nitems = 0;
  for (int c = 0; c < M->nclips; c++) {
    MClip* cl = &M->clips[c];
    MClipBatch* cb = &b->cb[c];
    const int n = VS_MIN(M->B, M->nframes[c] - cl->next);
    cb->n = n;
    cb->fnum0 = cl->next;
    cb->slot0 = set * M->B;
    const int tripod = cl->md.conf.virtualTripod;
    M->refs[c] = -1;
    for (int i = 0; i < n; i++) {
      const int fnum = cl->next + i;
      /* every frame after a clip's first is measured against the frame
         before; with virtualTripod >= 1 the frames before the reference
         frame (index virtualTripod-1) are a lead-in that emits no motion,
         the reference frame's blurred plane is frozen as the clip's stored
         reference, and every later frame is measured against it
         (src/motiondetect.c:211-218, 294-303) */
      cb->detect[i] = (uint8_t)(tripod >= 1 ? fnum >= tripod : fnum > 0);
      cb->prevSlot[i] =
          tripod >= 1 ? -1 : (i == 0 ? (cl->lastSlot >= 0 ? cl->lastSlot : cb->slot0) : cb->slot0 + i - 1);
      if (tripod >= 1 && fnum == tripod - 1) M->refs[c] = cb->slot0 + i;
      cb->luma[i] = M->luma[c][fnum];
      b->itemClip[b->nitems] = c;
      b->itemFrame[b->nitems] = i;
      b->nitems++;
    }
    if (n > 0) {
      cl->next += n;
      cl->lastSlot = cb->slot0 + n - 1;
    }
    M->n[c] = n;
    M->rows[c] = cb->luma;
    M->prevs[c] = cb->prevSlot;
    M->dets[c] = cb->detect;
  }
  if (vsg_multi_submit(M->gpu, set, M->n, M->rows, M->linesize, M->prevs, M->dets, M->refs) != 0) {
    vs_log_error("vidstab_amd", "GPU multi-clip detect failed: %s\n", vsg_last_error());
    return VS_ERROR;
  }
  b->active = 1;
  return VS_OK;
}

static int batch_coarse(Multi* M, MBatch* b) {
  if (vsg_multi_wait_coarse(M->gpu, b->set) != 0) goto fail;
  for (int c = 0; c < M->nclips; c++)
    if (vsg_wait_coarse(vsg_multi_clip(M->gpu, c), b->set, &b->cb[c].view) != 0) goto fail;
  {
    Work w = {M, b};
    vs_pool_run(M->pool, b->nitems, coarse_fn, &w);
  }
  int* nj = M->n;
  for (int c = 0; c < M->nclips; c++) {
    MClipBatch* cb = &b->cb[c];
    const int cap = M->clips[c].jobCap;
    int n = 0;
    for (int i = 0; i < cb->n; i++) {
      const size_t at = (size_t)i * cap;
      cb->jobStart[i] = n;
      if (cb->jobCount[i] > 0 && (size_t)n != at) {
        memmove(cb->jobs + n, cb->jobs + at, sizeof(VSGJob) * (size_t)cb->jobCount[i]);
        memmove(cb->jobSel + n, cb->jobSel + at, sizeof(int) * (size_t)cb->jobCount[i]);
      }
      n += cb->jobCount[i];
    }
    cb->jobStart[cb->n] = n;
    nj[c] = n;
    M->jobs[c] = cb->jobs;
  }
  if (vsg_multi_submit_fine(M->gpu, b->set, nj, M->jobs) != 0) goto fail;
  return VS_OK;
fail:
  vs_log_error("vidstab_amd", "GPU multi-clip detect failed: %s\n", vsg_last_error());
  return VS_ERROR;
}

static int batch_fine(Multi* M, MBatch* b) {
  if (vsg_multi_wait_fine(M->gpu, b->set) != 0) goto fail;
  for (int c = 0; c < M->nclips; c++) {
    if (vsg_wait_coarse(vsg_multi_clip(M->gpu, c), b->set, &b->cb[c].view) != 0 ||
        vsg_wait_fine(vsg_multi_clip(M->gpu, c), b->set, &b->cb[c].res) != 0)
      goto fail;
  }
  {
    Work w = {M, b};
    vs_pool_run(M->pool, b->nitems, fine_fn, &w);
  }
  b->active = 0;
  return VS_OK;
fail:
  vs_log_error("vidstab_amd", "GPU multi-clip fine search failed: %s\n", vsg_last_error());
  return VS_ERROR;
}

/* ---- the previous call's device state ----------------------------------------
   Setting up the contexts of many clips (buffers, graphs) and the host pool
   costs more than detecting a few short clips, so the last call's state is
   kept for a next call over clips of the same shapes (device, batch, frame
   sizes, blur boxes, field grids): then only the frames change.  A call that
   finds the kept state busy (another thread inside) sets up its own. */
typedef struct {
  int device, B, nclips;
  int *w, *h, *blur;
  VSGStageDesc *dc, *df;
} MultiShape;

static pthread_mutex_t g_keptMu = PTHREAD_MUTEX_INITIALIZER;
static VSGMulti* g_keptGpu = NULL;
static VSPool* g_keptPool = NULL;
static MultiShape g_keptShape;

static int same_desc(const VSGStageDesc* a, const VSGStageDesc* b) {
  if (a->fieldSize != b->fieldSize || a->maxShift != b->maxShift || a->stepSize != b->stepSize ||
      a->nfields != b->nfields || a->fieldRows != b->fieldRows || a->maxFields != b->maxFields ||
      a->contrastThreshold != b->contrastThreshold)
    return 0;
  for (int f = 0; f < a->nfields; f++)
    if (a->fields[f].x != b->fields[f].x || a->fields[f].y != b->fields[f].y || a->fields[f].size != b->fields[f].size)
      return 0;
  return 1;
}

static int same_shape(const MultiShape* a, const MultiShape* b) {
  if (a->device != b->device || a->B != b->B || a->nclips != b->nclips) return 0;
  for (int c = 0; c < a->nclips; c++)
    if (a->w[c] != b->w[c] || a->h[c] != b->h[c] || a->blur[c] != b->blur[c] || !same_desc(&a->dc[c], &b->dc[c]) ||
        !same_desc(&a->df[c], &b->df[c]))
      return 0;
  return 1;
}

static void shape_free(MultiShape* s) {
  if (s->dc)
    for (int c = 0; c < s->nclips; c++) {
      free((void*)s->dc[c].fields);
      free((void*)s->df[c].fields);
    }
  free(s->w);
  free(s->h);
  free(s->blur);
  free(s->dc);
  free(s->df);
  memset(s, 0, sizeof(*s));
}

/* a deep copy of the shape (the field arrays belong to the call's detectors) */
static int shape_copy(MultiShape* dst, const MultiShape* src) {
  memset(dst, 0, sizeof(*dst));
  dst->device = src->device;
  dst->B = src->B;
  dst->nclips = src->nclips;
  const size_t n = (size_t)src->nclips;
  dst->w = (int*)malloc(sizeof(int) * n);
  dst->h = (int*)malloc(sizeof(int) * n);
  dst->blur = (int*)malloc(sizeof(int) * n);
  dst->dc = (VSGStageDesc*)calloc(n, sizeof(VSGStageDesc));
  dst->df = (VSGStageDesc*)calloc(n, sizeof(VSGStageDesc));
  if (!dst->w || !dst->h || !dst->blur || !dst->dc || !dst->df) {
    shape_free(dst);
    return 0;
  }
  memcpy(dst->w, src->w, sizeof(int) * n);
  memcpy(dst->h, src->h, sizeof(int) * n);
  memcpy(dst->blur, src->blur, sizeof(int) * n);
  for (size_t c = 0; c < n; c++) {
    const VSGStageDesc* sd[2] = {&src->dc[c], &src->df[c]};
    VSGStageDesc* dd[2] = {&dst->dc[c], &dst->df[c]};
    for (int k = 0; k < 2; k++) {
      *dd[k] = *sd[k];
      Field* f = (Field*)malloc(sizeof(Field) * (size_t)(sd[k]->nfields ? sd[k]->nfields : 1));
      dd[k]->fields = f;
      if (!f) {
        shape_free(dst);
        return 0;
      }
      memcpy(f, sd[k]->fields, sizeof(Field) * (size_t)sd[k]->nfields);
    }
  }
  return 1;
}

/* takes the kept state if it fits `shape` and nobody else holds it */
static void kept_take(const MultiShape* shape, VSGMulti** gpu, VSPool** pool) {
  *gpu = NULL;
  *pool = NULL;
  if (pthread_mutex_trylock(&g_keptMu) != 0) return;
  if (g_keptGpu && same_shape(&g_keptShape, shape)) {
    *gpu = g_keptGpu;
    g_keptGpu = NULL;
  }
  *pool = g_keptPool;
  g_keptPool = NULL;
  pthread_mutex_unlock(&g_keptMu);
}

/* keeps a call's state for the next call (replacing what was kept); false
   when it could not be kept, and then the caller releases it */
static int kept_put(const MultiShape* shape, VSGMulti* gpu, VSPool* pool) {
  if (pthread_mutex_trylock(&g_keptMu) != 0) return 0;
  MultiShape copy;
  int ok = !gpu || shape_copy(&copy, shape);
  if (ok) {
    if (gpu) {
      if (g_keptGpu) vsg_multi_destroy(g_keptGpu);
      shape_free(&g_keptShape);
      g_keptGpu = gpu;
      g_keptShape = copy;
    }
    if (pool) {
      if (g_keptPool) vs_pool_destroy(g_keptPool);
      g_keptPool = pool;
    }
  }
  pthread_mutex_unlock(&g_keptMu);
  return ok;
}

static void multi_free(Multi* M, MBatch* bs) {
  if (bs) {
    for (int k = 0; k < MULTI_INFLIGHT; k++) {
      if (bs[k].cb) {
        for (int c = 0; c < M->nclips; c++) {
          if (bs[k].active)
            for (int i = 0; i < bs[k].cb[c].n; i++) vs_vector_del(&bs[k].cb[c].coarse[i]);
          free(bs[k].cb[c].jobs);
          free(bs[k].cb[c].jobSel);
        }
        free(bs[k].cb);
      }
      free(bs[k].itemClip);
      free(bs[k].itemFrame);
    }
    free(bs);
  }
  if (M->gpu) vsg_multi_destroy(M->gpu);
  if (M->pool) vs_pool_destroy(M->pool);
  if (M->clips) {
    for (int c = 0; c < M->nclips; c++) {
      VSMotionDetect* md = &M->clips[c].md;
      if (md->fieldscoarse.fields) vs_free(md->fieldscoarse.fields);
      if (md->fieldsfine.fields) vs_free(md->fieldsfine.fields);
    }
    free(M->clips);
  }
  free(M->n);
  free((void*)M->rows);
  free((void*)M->prevs);
  free((void*)M->dets);
  free(M->refs);
  free((void*)M->jobs);
}

VS_API int vsMotionDetectMulti(const VSMotionDetectConfig* confs, const VSFrameInfo* fis, int nclips,
                               const int* nframes, const uint8_t* const* const* lumaDev, const int* linesize,
                               LocalMotions* const* motions) {
  if (!confs || !fis || nclips <= 0 || !nframes || !lumaDev || !linesize || !motions) return VS_ERROR;
  Multi M;
  memset(&M, 0, sizeof(M));
  M.nclips = nclips;
  /* frames of every clip per device batch (at most MULTI_MAX_B = 64): 64
     when the longest clip has 96 frames or more, else 16.  Long clips gain
     from fewer, larger batches (fewer front launches and host steps per
     frame); short clips from more batches in flight.  Measured on config 4
     (64 clips, 30-120 frames; frames/s of the whole call, against one clip
     after another in the same run): 16: 6,599-6,615 (0.92-0.94x), 32:
     7,349-7,369 (1.00x), 48: 7,602-7,607 (1.04x), 64: 7,633-7,668
     (1.05x); on 64 clips of 30 frames: 16: 88,600-91,600, 32:
     79,700-83,400.  Device memory grows with B: three slot sets of B planes
     per clip. */
  int longest = 0;
  for (int c = 0; c < nclips; c++) longest = VS_MAX(longest, nframes[c]);
  M.B = VS_MIN(MULTI_MAX_B, VS_MAX(1, env_int("VIDSTAB_AMD_MULTI_BATCH", longest >= 96 ? 64 : 16)));
  M.motions = motions;
  M.luma = lumaDev;
  M.nframes = nframes;
  M.linesize = linesize;
  M.clips = (MClip*)calloc((size_t)nclips, sizeof(MClip));
  M.n = (int*)calloc((size_t)nclips, sizeof(int));
  M.rows = (const uint8_t* const**)calloc((size_t)nclips, sizeof(void*));
  M.prevs = (const int**)calloc((size_t)nclips, sizeof(void*));
  M.dets = (const uint8_t**)calloc((size_t)nclips, sizeof(void*));
  M.refs = (int*)calloc((size_t)nclips, sizeof(int));
  M.jobs = (const VSGJob**)calloc((size_t)nclips, sizeof(void*));
  MBatch* bs = (MBatch*)calloc(MULTI_INFLIGHT, sizeof(MBatch));
  int* w = (int*)calloc((size_t)nclips, sizeof(int));
  int* h = (int*)calloc((size_t)nclips, sizeof(int));
  int* blur = (int*)calloc((size_t)nclips, sizeof(int));
  VSGStageDesc* dc = (VSGStageDesc*)calloc((size_t)nclips, sizeof(VSGStageDesc));
  VSGStageDesc* df = (VSGStageDesc*)calloc((size_t)nclips, sizeof(VSGStageDesc));
  int ok = M.clips && M.n && M.rows && M.prevs && M.dets && M.refs && M.jobs && bs && w && h && blur && dc && df;
  int maxFrames = 0;
  for (int c = 0; c < nclips && ok; c++) {
    MClip* cl = &M.clips[c];
    if (nframes[c] < 0 || vsd_configure(&cl->md, &confs[c], &fis[c]) != VS_OK) {
      ok = 0;
      break;
    }
    if (cl->md.fi.pFormat > PF_PACKED) {
      vs_log_error(confs[c].modName, "the multi-clip form takes planar frames\n");
      ok = 0;
      break;
    }
    cl->lastSlot = -1;
    cl->jobCap = cl->md.fieldsfine.fieldNum;
    w[c] = cl->md.fi.width;
    h[c] = cl->md.fi.height;
    blur[c] = vsd_blur_size(&cl->md);
    dc[c] = vsd_stage_desc(&cl->md.fieldscoarse);
    df[c] = vsd_stage_desc(&cl->md.fieldsfine);
    maxFrames = VS_MAX(maxFrames, nframes[c]);
  }
  MultiShape shape = {env_int("VIDSTAB_AMD_DEVICE", 0), M.B, nclips, w, h, blur, dc, df};
  M.device = shape.device;
  if (ok) kept_take(&shape, &M.gpu, &M.pool);
  if (ok && !M.gpu && vsg_multi_create(&M.gpu, shape.device, nclips, w, h, blur, dc, df, M.B) != 0) {
    vs_log_error("vidstab_amd", "GPU multi-clip initialisation failed: %s\n", vsg_last_error());
    ok = 0;
  }
  if (ok && !M.pool) {
    const long cores = sysconf(_SC_NPROCESSORS_ONLN);
    M.pool = vs_pool_create(VS_MIN(16, VS_MAX(1, env_int("VIDSTAB_AMD_HOST_THREADS", (int)(cores * 0.8)))));
  }
  if (ok) {
    for (int k = 0; k < MULTI_INFLIGHT && ok; k++) {
      bs[k].cb = (MClipBatch*)calloc((size_t)nclips, sizeof(MClipBatch));
      bs[k].itemClip = (int*)malloc(sizeof(int) * (size_t)nclips * M.B);
      bs[k].itemFrame = (int*)malloc(sizeof(int) * (size_t)nclips * M.B);
      ok = bs[k].cb && bs[k].itemClip && bs[k].itemFrame;
      for (int c = 0; c < nclips && ok; c++) {
        const size_t cap = (size_t)M.B * M.clips[c].jobCap;
        bs[k].cb[c].jobs = (VSGJob*)malloc(sizeof(VSGJob) * (cap ? cap : 1));
        bs[k].cb[c].jobSel = (int*)malloc(sizeof(int) * (cap ? cap : 1));
        ok = bs[k].cb[c].jobs && bs[k].cb[c].jobSel;
      }
    }
  }
  const int nb = (maxFrames + M.B - 1) / M.B;
  int done = 0;
  for (int k = 0; k < nb && k < 2 && ok; k++) ok = batch_submit(&M, &bs[k], k % VSG_NSETS) == VS_OK;
  for (int k = 0; k < nb && ok; k++) {
    MBatch* b = &bs[k % MULTI_INFLIGHT];
    if (batch_coarse(&M, b) != VS_OK) {
      ok = 0;
      break;
    }
    if (k + 2 < nb && batch_submit(&M, &bs[(k + 2) % MULTI_INFLIGHT], (k + 2) % VSG_NSETS) != VS_OK) {
      ok = 0;
      break;
    }
    if (batch_fine(&M, b) != VS_OK) {
      ok = 0;
      break;
    }
    done = k + 1;
  }
  if (!ok) {
    /* release what was produced: the first `done` batches of every clip */
    for (int c = 0; c < nclips; c++) {
      const int upto = VS_MIN(nframes[c], done * M.B);
      for (int t = 0; t < upto; t++) vs_vector_del(&motions[c][t]);
    }
  }
  /* a call that went through keeps its device state and pool for the next */
  if (ok && kept_put(&shape, M.gpu, M.pool)) {
    M.gpu = NULL;
    M.pool = NULL;
  }
  free(w);
  free(h);
  free(blur);
  free(dc);
  free(df);
  multi_free(&M, bs);
  return ok ? VS_OK : VS_ERROR;
}
