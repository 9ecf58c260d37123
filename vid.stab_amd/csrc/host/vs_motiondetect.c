/*
 * vs_motiondetect.c -- the detect-pass driver behind the vid.stab C API
 * (reference: src/motiondetect.c:60-305, 314-358, 665-807).
 *
 * Per batch of frames (one frame for vsMotionDetection):
 *   1. GPU: upload, box blur, per-field min/max        (vsg_frames)
 *   2. host: contrast, threshold and stable selection  (selectfields)
 *   3. GPU: coarse grid search + refinement            (vsg_search, stage 0)
 *   4. host: coarse LocalMotions, simple transform, fine offsets
 *   5. GPU: fine search at the offsets                 (vsg_search, stage 1)
 *   6. host: match cleanup ("keep fine with match <= cleanmean(coarse)"),
 *            coarse ++ fine output, previous-frame bookkeeping
 * Frames of one batch are independent given their previous frame, so each
 * GPU stage runs every field of every frame of the batch in one launch.
 *
 * VSMotionDetect is caller-owned with a frozen layout, so the GPU context of
 * a detector lives in a side table keyed by the struct's address.
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "localmotion2transform.h"
#include "motiondetect.h"
#include "transformtype_operations.h"
#include "vidstab_amd.h"
#include "vidstabdefines.h"
#include "vsg.h"

/* ---- side table --------------------------------------------------------- */

typedef struct {
  double contrast;
  int index;
} SelEntry;

typedef struct MDState {
  const VSMotionDetect* key;
  struct MDState* next;
  VSGContext* gpu;
  int maxBatch;
  int lastSlot;
  /* per-batch host scratch */
  uint8_t* mmC;            /* [maxBatch][ncoarse][2] */
  uint8_t* mmF;            /* [maxBatch][nfine][2] */
  SelEntry* selC;          /* [maxBatch][ncoarse] */
  SelEntry* selF;          /* [maxBatch][nfine] */
  int* nselC;
  int* nselF;
  VSGJob* jobs;            /* [maxBatch * max(nc, nf)] */
  VSGResult* res;
  int* jobFrame;           /* frame of each fine job */
  int* jobSel;             /* selection index of each fine job */
  SelEntry* sortTmp;
  double* contrast;
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
  if (s->gpu) vsg_destroy(s->gpu);
  free(s->mmC); free(s->mmF); free(s->selC); free(s->selF); free(s->nselC); free(s->nselF);
  free(s->jobs); free(s->res); free(s->jobFrame); free(s->jobSel); free(s->sortTmp);
  free(s->contrast);
  free(s);
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

/* src/boxblur.c:109-122 */
static int blur_box_size(const VSMotionDetect* md) {
  int size = md->conf.stepSize;
  if (size < 2) return 1;
  size = (size / 2) * 2 + 1;
  return VS_CLAMP(size, 3, VS_MIN(md->fi.height / 2, md->fi.width / 2));
}

static void* xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

VS_API int vsMotionDetectInit(VSMotionDetect* md, const VSMotionDetectConfig* conf, const VSFrameInfo* fi) {
  if (!md || !conf || !fi) return VS_ERROR;
  md->conf = *conf;
  md->fi = *fi;
  if (fi->pFormat <= PF_NONE || fi->pFormat == PF_PACKED || fi->pFormat >= PF_NUMBER) {
    vs_log_warn(md->conf.modName, "unsupported Pixel Format (%i)\n", md->fi.pFormat);
    return VS_ERROR;
  }
  if (fi->pFormat > PF_PACKED) {
    vs_log_error(md->conf.modName, "packed pixel formats are not supported by the GPU detector (%i)\n",
                 md->fi.pFormat);
    return VS_ERROR;
  }
  if (md->conf.stepSize < 1) {
    vs_log_error(md->conf.modName, "stepsize must be at least 1\n");
    return VS_ERROR;
  }
  state_remove(md);   /* a re-initialised detector drops its old device state */

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
  md->fieldscoarse.fields = NULL;
  md->fieldsfine.fields = NULL;
  if (!init_fields(md, &md->fieldscoarse, fieldSize, maxShift, md->conf.stepSize, 1, 0,
                   md->conf.contrastThreshold))
    return VS_ERROR;
  if (!init_fields(md, &md->fieldsfine, fieldSizeFine, fieldSizeFine, 2, 1, fieldSizeFine,
                   md->conf.contrastThreshold / 2))
    return VS_ERROR;

  MDState* s = (MDState*)calloc(1, sizeof(MDState));
  if (!s) return VS_ERROR;
  s->key = md;
  s->maxBatch = VS_MAX(1, env_int("VIDSTAB_AMD_MAX_BATCH", 16));
  const int device = env_int("VIDSTAB_AMD_DEVICE", 0);
  const int nc = md->fieldscoarse.fieldNum, nf = md->fieldsfine.fieldNum;
  const int B = s->maxBatch, nmax = VS_MAX(nc, nf);
  s->mmC = (uint8_t*)xcalloc((size_t)B * nc * 2, 1);
  s->mmF = (uint8_t*)xcalloc((size_t)B * nf * 2, 1);
  s->selC = (SelEntry*)xcalloc((size_t)B * nc, sizeof(SelEntry));
  s->selF = (SelEntry*)xcalloc((size_t)B * nf, sizeof(SelEntry));
  s->nselC = (int*)xcalloc((size_t)B, sizeof(int));
  s->nselF = (int*)xcalloc((size_t)B, sizeof(int));
  s->jobs = (VSGJob*)xcalloc((size_t)B * nmax, sizeof(VSGJob));
  s->res = (VSGResult*)xcalloc((size_t)B * nmax, sizeof(VSGResult));
  s->jobFrame = (int*)xcalloc((size_t)B * nmax, sizeof(int));
  s->jobSel = (int*)xcalloc((size_t)B * nmax, sizeof(int));
  s->sortTmp = (SelEntry*)xcalloc((size_t)nmax, sizeof(SelEntry));
  s->contrast = (double*)xcalloc((size_t)nmax, sizeof(double));
  if (!s->mmC || !s->mmF || !s->selC || !s->selF || !s->nselC || !s->nselF || !s->jobs || !s->res ||
      !s->jobFrame || !s->jobSel || !s->sortTmp || !s->contrast) {
    state_free(s);
    vs_log_error(md->conf.modName, "malloc failed");
    return VS_ERROR;
  }
  VSGStageDesc dc = {md->fieldscoarse.fieldSize, md->fieldscoarse.maxShift, md->fieldscoarse.stepSize,
                     nc, md->fieldscoarse.fields};
  VSGStageDesc df = {md->fieldsfine.fieldSize, md->fieldsfine.maxShift, md->fieldsfine.stepSize, nf,
                     md->fieldsfine.fields};
  if (vsg_create(&s->gpu, device, md->fi.width, md->fi.height, blur_box_size(md), &dc, &df, B) != 0) {
    vs_log_error(md->conf.modName, "GPU initialisation failed: %s\n", vsg_last_error());
    state_free(s);
    return VS_ERROR;
  }
  vs_log_info(md->conf.modName, "GPU: HIP device %i, batch %i\n", device, B);
  state_insert(s);
  md->initialized = 2;
  return VS_OK;
}

VS_API void vsMotionDetectionCleanup(VSMotionDetect* md) {
  state_remove(md);
  if (md->fieldscoarse.fields) {
    vs_free(md->fieldscoarse.fields);
    md->fieldscoarse.fields = 0;
  }
  if (md->fieldsfine.fields) {
    vs_free(md->fieldsfine.fields);
    md->fieldsfine.fields = 0;
  }
  vsFrameFree(&md->prev);
  vsFrameFree(&md->curr);
  vsFrameFree(&md->currtmp);
  md->initialized = 0;
}

/* ---- field selection: src/motiondetect.c:665-743 ---------------------------
   The reference sorts with glibc qsort, a stable merge sort, so equal
   contrasts keep ascending index order; the merge sort below is stable too. */

static void sort_desc_stable(SelEntry* a, int n, SelEntry* tmp) {
  for (int width = 1; width < n; width *= 2) {
    for (int lo = 0; lo < n; lo += 2 * width) {
      const int mid = VS_MIN(lo + width, n), hi = VS_MIN(lo + 2 * width, n);
      int p = lo, q = mid, o = lo;
      while (p < mid && q < hi) tmp[o++] = (a[q].contrast > a[p].contrast) ? a[q++] : a[p++];
      while (p < mid) tmp[o++] = a[p++];
      while (q < hi) tmp[o++] = a[q++];
    }
    memcpy(a, tmp, sizeof(SelEntry) * (size_t)n);
  }
}

/* mm: (min,max) per field -> ordered selection in out; returns its length */
static int select_fields(const VSMotionDetectFields* fs, const uint8_t* mm, SelEntry* out,
                         MDState* s) {
  const int n = fs->fieldNum;
  double* c = s->contrast;
  SelEntry* seg = (SelEntry*)malloc(sizeof(SelEntry) * (size_t)n);
  if (!seg) return 0;
  for (int i = 0; i < n; i++) {
    const int mn = mm[2 * i], mx = mm[2 * i + 1];
    double v = (double)(mx - mn) / ((double)(mx + mn) + 0.1);
    c[i] = v < fs->contrastThreshold ? 0.0 : v;
    seg[i].contrast = c[i];
    seg[i].index = i;
  }
  int cnt = 0;
  const int numsegms = fs->fieldRows + 1;
  const int segmlen = n / (fs->fieldRows + 1) + 1;
  for (int g = 0; g < numsegms; g++) {
    const int start = segmlen * g, end = VS_MIN(segmlen * (g + 1), n);
    if (end > start) sort_desc_stable(seg + start, end - start, s->sortTmp);
    for (int j = 0; j < fs->maxFields / numsegms; j++) {
      if (start + j >= end) continue;
      if (seg[start + j].contrast > 0) {
        out[cnt].index = seg[start + j].index;
        out[cnt].contrast = c[seg[start + j].index];
        cnt++;
        seg[start + j].contrast = 0;
      }
    }
  }
  const int remaining = fs->maxFields - cnt;
  if (remaining > 0) {
    sort_desc_stable(seg, n, s->sortTmp);
    for (int j = 0; j < remaining && j < n; j++)
      if (seg[j].contrast > 0) out[cnt++] = seg[j];
  }
  free(seg);
  return cnt;
}

/* ---- per-field result -> LocalMotion: src/motiondetect.c:573-585 ---------- */

static int make_motion(const VSMotionDetectFields* fs, const Field* f, const VSGResult* r, int ox,
                       int oy, double contrast, LocalMotion* lm) {
  int st = fs->stepSize;
  while (st > 1) st /= 2;
  if (abs(r->tx) >= fs->maxShift + st - 1 || abs(r->ty) >= fs->maxShift + st) return 0;
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

/* ---- the batch core --------------------------------------------------------- */

static int detect_chunk(VSMotionDetect* md, MDState* s, LocalMotions* motions,
                        const uint8_t* const* luma, const int* linesize, int onDevice, int nb) {
  const char* mod = md->conf.modName;
  VSMotionDetectFields* fc = &md->fieldscoarse;
  VSMotionDetectFields* ff = &md->fieldsfine;
  const int nc = fc->fieldNum, nf = ff->fieldNum;
  const int tripod = md->conf.virtualTripod;

  if (vsg_frames(s->gpu, nb, luma, linesize, onDevice, s->mmC, s->mmF) != 0) {
    vs_log_error(mod, "GPU blur/contrast failed: %s\n", vsg_last_error());
    return VS_ERROR;
  }

  /* which frames are measured, and against which plane */
  int detect[64], prevSlot[64];
  int lastKept = -1;   /* last slot of this batch that became the stored previous plane */
  int seen = md->hasSeenOneFrame;
  for (int i = 0; i < nb; i++) {
    const int fnum = md->frameNum + i;
    const int leadIn = tripod >= 1 && fnum < tripod;
    detect[i] = seen && !leadIn;
    prevSlot[i] = lastKept;
    if (!detect[i]) seen = 1;
    if (tripod < 1 || fnum < tripod) lastKept = i;
  }

  /* coarse selection and jobs */
  int njobs = 0;
  for (int i = 0; i < nb; i++) {
    s->nselC[i] = 0;
    if (!detect[i]) continue;
    SelEntry* sel = s->selC + (size_t)i * nc;
    s->nselC[i] = select_fields(fc, s->mmC + (size_t)i * nc * 2, sel, s);
    for (int q = 0; q < s->nselC[i]; q++) {
      VSGJob* j = &s->jobs[njobs++];
      j->slot = i;
      j->prev = prevSlot[i];
      j->field = sel[q].index;
      j->ox = 0;
      j->oy = 0;
    }
  }
  if (vsg_search(s->gpu, 0, njobs, s->jobs, s->res) != 0) {
    vs_log_error(mod, "GPU coarse search failed: %s\n", vsg_last_error());
    return VS_ERROR;
  }

  /* coarse LocalMotions, transform, fine jobs */
  LocalMotions coarse[64];
  int jc = 0, njf = 0;
  for (int i = 0; i < nb; i++) {
    vs_vector_init(&coarse[i], 0);
    if (!detect[i]) continue;
    vs_vector_init(&coarse[i], fc->maxFields);
    const SelEntry* sel = s->selC + (size_t)i * nc;
    for (int q = 0; q < s->nselC[i]; q++, jc++) {
      LocalMotion lm;
      const Field* f = &fc->fields[sel[q].index];
      if (make_motion(fc, f, &s->res[jc], 0, 0, sel[q].contrast, &lm))
        vs_vector_append_dup(&coarse[i], &lm, sizeof(LocalMotion));
    }
    if (vs_vector_size(&coarse[i]) < 1) {
      vs_log_warn(mod, "too low contrast. (no translations are detected in frame %i)\n", md->frameNum + i);
      continue;
    }
    VSTransform t = vsSimpleMotionsToTransform(md->fi, mod, &coarse[i]);
    ff->offset = t;
    ff->useOffset = 1;
    PreparedTransform pt = prepare_transform(&ff->offset, &md->fi);
    SelEntry* sel2 = s->selF + (size_t)i * nf;
    s->nselF[i] = select_fields(ff, s->mmF + (size_t)i * nf * 2, sel2, s);
    for (int q = 0; q < s->nselF[i]; q++) {
      const Field* f = &ff->fields[sel2[q].index];
      Vec pos = {f->x, f->y};
      Vec off = sub_vec(transform_vec(&pt, &pos), pos);
      /* src/motiondetect.c:418-434: discard fields whose window leaves the frame */
      const int border = f->size / 2 + ff->maxShift + ff->stepSize;
      const int x = pos.x + off.x, y = pos.y + off.y;
      if (x - border < 0 || x + border >= md->fi.width || y - border < 0 || y + border >= md->fi.height)
        continue;
      VSGJob* j = &s->jobs[njf];
      j->slot = i;
      j->prev = prevSlot[i];
      j->field = sel2[q].index;
      j->ox = off.x;
      j->oy = off.y;
      s->jobFrame[njf] = i;
      s->jobSel[njf] = q;
      njf++;
    }
  }
  if (vsg_search(s->gpu, 1, njf, s->jobs, s->res) != 0) {
    vs_log_error(mod, "GPU fine search failed: %s\n", vsg_last_error());
    for (int i = 0; i < nb; i++) vs_vector_del(&coarse[i]);
    return VS_ERROR;
  }

  /* fine LocalMotions, match cleanup, output */
  int jf = 0;
  for (int i = 0; i < nb; i++) {
    if (!detect[i]) {
      vs_vector_fini(&coarse[i]);
      vs_vector_init(&motions[i], 1);   /* the reference's empty "dummy" vector */
      continue;
    }
    LocalMotions fine;
    vs_vector_init(&fine, 0);
    if (vs_vector_size(&coarse[i]) >= 1) {
      VSArray mq = localmotionsGetMatch(&coarse[i]);
      double meanMatch = cleanmean(mq.dat, mq.len, NULL, NULL);
      vs_array_free(mq);
      const SelEntry* sel2 = s->selF + (size_t)i * nf;
      for (; jf < njf && s->jobFrame[jf] == i; jf++) {
        const VSGJob* j = &s->jobs[jf];
        LocalMotion lm;
        const int q = s->jobSel[jf];
        if (!make_motion(ff, &ff->fields[j->field], &s->res[jf], j->ox, j->oy, sel2[q].contrast, &lm))
          continue;
        if (lm_match_better(&meanMatch, &lm)) vs_vector_append_dup(&fine, &lm, sizeof(LocalMotion));
      }
    }
    motions[i] = vs_vector_concat(&coarse[i], &fine);
    vs_vector_fini(&coarse[i]);
    vs_vector_fini(&fine);
  }

  if (lastKept >= 0 && vsg_keep_prev(s->gpu, lastKept) != 0) {
    vs_log_error(mod, "GPU: %s\n", vsg_last_error());
    return VS_ERROR;
  }
  s->lastSlot = nb - 1;
  md->hasSeenOneFrame = (short)seen;
  md->frameNum += nb;
  return VS_OK;
}

static int detect_frames(VSMotionDetect* md, LocalMotions* motions, const uint8_t* const* luma,
                         const int* linesize, int onDevice, int n) {
  if (!md || md->initialized != 2 || !motions || n < 0) return VS_ERROR;
  MDState* s = state_find(md);
  if (!s) {
    vs_log_error(md->conf.modName, "detector has no GPU state (not initialised?)\n");
    return VS_ERROR;
  }
  const int B = VS_MIN(s->maxBatch, 64);
  for (int done = 0; done < n;) {
    const int nb = VS_MIN(B, n - done);
    if (detect_chunk(md, s, motions + done, luma + done, linesize + done, onDevice, nb) != VS_OK) {
      for (int i = 0; i < done; i++) vs_vector_del(&motions[i]);
      return VS_ERROR;
    }
    done += nb;
  }
  return VS_OK;
}

VS_API int vsMotionDetection(VSMotionDetect* md, LocalMotions* motions, VSFrame* frame) {
  if (!md || !frame || md->initialized != 2) return VS_ERROR;
  md->currorig = *frame;
  const uint8_t* luma = frame->data[0];
  int ls = frame->linesize[0];
  return detect_frames(md, motions, &luma, &ls, 0, 1);
}

VS_API int vsMotionDetectionBatch(VSMotionDetect* md, LocalMotions* motions, VSFrame* frames, int n) {
  if (!md || !frames || n < 0) return VS_ERROR;
  if (n == 0) return VS_OK;
  const uint8_t** luma = (const uint8_t**)malloc(sizeof(uint8_t*) * (size_t)n);
  int* ls = (int*)malloc(sizeof(int) * (size_t)n);
  if (!luma || !ls) {
    free(luma);
    free(ls);
    return VS_ERROR;
  }
  for (int i = 0; i < n; i++) {
    luma[i] = frames[i].data[0];
    ls[i] = frames[i].linesize[0];
  }
  md->currorig = frames[n - 1];
  int r = detect_frames(md, motions, luma, ls, 0, n);
  free(luma);
  free(ls);
  return r;
}

VS_API int vsMotionDetectionDevice(VSMotionDetect* md, LocalMotions* motions, const uint8_t* const* lumaDev,
                                   const int* linesize, int n) {
  if (n == 0) return VS_OK;
  return detect_frames(md, motions, lumaDev, linesize, 1, n);
}

VS_API int vsMotionDetectGetBlurred(const VSMotionDetect* md, uint8_t* dst, int dstStride) {
  MDState* s = state_find(md);
  if (!s) return VS_ERROR;
  return vsg_read_blurred(s->gpu, s->lastSlot, dst, dstStride) == 0 ? VS_OK : VS_ERROR;
}
