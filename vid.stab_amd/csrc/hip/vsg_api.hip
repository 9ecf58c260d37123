/*
 * vsg_api.hip -- the thin C ABI over the detect kernels (declared in vsg.h):
 * device buffers, the per-set HIP streams of one detector, uploads and read-backs.
 *
 * HBM layout of one context (pitch = width * bytes per pixel rounded up to
 * 64 bytes; VSG_NSETS sets of B = maxBatch slots, global slot = set*B + i):
 *   raw[3B]       uploaded planes, pitch x H (host-frame batches only): the
 *                 luma, or plane 0 of a packed RGB24 / BGR24 / RGBA frame
 *   blur[3B], ref blurred luma (packed: an unblurred copy), pitch x
 *                 (H + padRows); the pad rows keep the grid kernel's
 *                 idle-candidate reads inside the buffer
 *   patches[3B]   an aligned, dense copy of every coarse and fine field
 *   minmax[3B]    8-bit (min, max) per field (packed: per colour channel)
 *   per set       selection lists, coarse/fine job lists, best[] keys, results
 *
 * The sets share one compute stream by default, which orders every batch's
 * stages by submission.  With a stream per set (VIDSTAB_AMD_STREAMS=3, or
 * streams handed in by the multi-clip form) cross-set hazards are ordered
 * with events: a batch's coarse search waits for the previous batch's
 * planesReady (which in turn waited for the one before, so the tripod plane
 * is covered), and a set is overwritten only after the batch that followed
 * its previous use has finished its latest search (it read the set's last
 * plane as its previous frame).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <mutex>
#include <new>
#include <vector>

#include "vsg_context.h"

static thread_local char g_err[512];

int vsg_fail(const char* what, hipError_t e) {
  snprintf(g_err, sizeof(g_err), "%s: %s", what, e == hipSuccess ? "invalid argument" : hipGetErrorString(e));
  return -1;
}

#define VSG_CHECK(call)                                  \
  do {                                                   \
    hipError_t e_ = (call);                              \
    if (e_ != hipSuccess) return vsg_fail(#call, e_);    \
  } while (0)

extern "C" const char* vsg_last_error(void) { return g_err; }
extern "C" int vsg_spiral_rank(int a, int b) { return vsg_rank(a, b); }
extern "C" void vsg_spiral_unrank(int rank, int* a, int* b) { vsg_unrank(rank, a, b); }

/* ---- optional per-kernel timing ------------------------------------------- */

static std::atomic<int> g_profOn{0};
static std::mutex g_profMu;
static VSGProfile g_prof;

extern "C" void vsg_profile_enable(int on) { g_profOn.store(on ? 1 : 0); }
extern "C" void vsg_profile_reset(void) {
  std::lock_guard<std::mutex> lk(g_profMu);
  memset(&g_prof, 0, sizeof(g_prof));
}
extern "C" void vsg_profile_get(VSGProfile* out) {
  std::lock_guard<std::mutex> lk(g_profMu);
  if (out) *out = g_prof;
}

static void prof_add(int k, hipEvent_t a, hipEvent_t b, double work) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return;
  std::lock_guard<std::mutex> lk(g_profMu);
  g_prof.launches[k] += 1;
  g_prof.ms[k] += ms;
  g_prof.work[k] += work;
}

/* ---- lifetime ---------------------------------------------------------------- */

template <class T>
static void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}
template <class T>
static void hfree(T*& p) {
  if (p) (void)hipHostFree(p);
  p = nullptr;
}

void vsg_ctx_release_transfer(VSGContext* c, int set) {
  VSGSet& s = c->set[set];
  dfree(s.d_selIdx); dfree(s.d_selC); dfree(s.d_count); dfree(s.d_jobsF); dfree(s.d_resC); dfree(s.d_resF);
  dfree(s.d_desc);
  hfree(s.h_selIdx); hfree(s.h_selC); hfree(s.h_count); hfree(s.h_resC); hfree(s.h_jobsF); hfree(s.h_resF);
  hfree(s.h_desc);
  s.d_src = nullptr; s.d_prev = nullptr; s.d_detect = nullptr;
  s.h_src = nullptr; s.h_prev = nullptr; s.h_detect = nullptr;
  s.m_selIdx = nullptr; s.m_selC = nullptr; s.m_count = nullptr; s.m_resC = nullptr;
  s.m_jobsF = nullptr; s.m_resF = nullptr; s.m_desc = nullptr;
  s.zc = false;
}

void vsg_ctx_free(VSGContext* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  for (VSGSet& s : c->set)
    if (s.stream) (void)hipStreamSynchronize(s.stream);
  vsg_learned_free(c);
  /* the streams this context created, each once (the sets may share one);
     destroyed after every event recorded on them */
  hipStream_t own[VSG_NSETS + 1] = {};
  int nown = 0;
  if (c->ownStreams)
    for (const VSGSet& s : c->set) {
      bool seen = s.stream == nullptr;
      for (int q = 0; q < nown && !seen; ++q) seen = own[q] == s.stream;
      if (!seen) own[nown++] = s.stream;
    }
  if (c->upStream) own[nown++] = c->upStream;
  c->upStream = nullptr;
  for (VSGStage& s : c->st) {
    dfree(s.d_fx);
    dfree(s.d_fy);
  }
  dfree(c->d_allx); dfree(c->d_ally); dfree(c->d_allsz); dfree(c->d_patchOff); dfree(c->d_ctab);
  dfree(c->d_raw); dfree(c->d_blur); dfree(c->d_ref); dfree(c->d_tmp); dfree(c->d_patches);
  dfree(c->d_minmax);
  for (int k = 0; k < VSG_NSETS; ++k) {
    VSGSet& s = c->set[k];
    if (!s.borrowed) vsg_ctx_release_transfer(c, k);
    dfree(s.d_jobsC); dfree(s.d_best); dfree(s.d_done); hfree(s.h_stage);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.doneF) (void)hipEventDestroy(s.doneF);
    if (s.planesReady) (void)hipEventDestroy(s.planesReady);
    if (s.searchDone) (void)hipEventDestroy(s.searchDone);
    if (s.uploaded) (void)hipEventDestroy(s.uploaded);
    for (hipEvent_t& e : s.ev)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t& e : s.evf)
      if (e) (void)hipEventDestroy(e);
    s = VSGSet();
  }
  for (int q = 0; q < nown; ++q) (void)hipStreamDestroy(own[q]);
}

int vsg_ctx_sync(VSGContext* c) {
  VSG_CHECK(hipSetDevice(c->device));
  for (VSGSet& s : c->set) VSG_CHECK(hipStreamSynchronize(s.stream));
  return 0;
}

static void free_ctx(VSGContext* c) {
  if (!c) return;
  vsg_ctx_free(c);
  delete c;
}

extern "C" void vsg_destroy(VSGContext* ctx) { free_ctx(ctx); }

/* bytes of a set's descriptor block for B frames */
/* batches of at least this many host frames upload beside the compute stream */
constexpr int kUploadAsideMin = 8;

static size_t desc_bytes(int B) { return (sizeof(VSGBlurSrc) + sizeof(int) + 1) * (size_t)B; }

template <class T>
static hipError_t dalloc(T** p, size_t n) {
  return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (n ? n : 1));
}
template <class T>
static hipError_t halloc(T** p, size_t n) {
  return hipHostMalloc(reinterpret_cast<void**>(p), sizeof(T) * (n ? n : 1), hipHostMallocDefault);
}
/* the device address of pinned host memory */
template <class T>
static hipError_t hmap(T** dev, T* host) {
  return hipHostGetDevicePointer(reinterpret_cast<void**>(dev), host, 0);
}

/* batches of up to zero_copy_max_batch() frames skip the small copies
   across PCIe: their descriptors and fine jobs are read and their selection
   and search results written by the kernels in pinned host memory, which
   saves the copy launches and the gaps between them at every batch
   boundary.  Host frames of small batches (fewer than kUploadAsideMin: a
   single frame per vsMotionDetection call) also go up by a kernel on the
   compute queue; those of larger batches by DMA on a stream beside it.
   VIDSTAB_AMD_ZERO_COPY=0 keeps every copy; VIDSTAB_AMD_ZERO_COPY_MAX_BATCH
   sets the largest batch that skips them (default: all). */
static bool zero_copy_enabled() {
  static const bool on = [] {
    const char* e = getenv("VIDSTAB_AMD_ZERO_COPY");
    return !(e && atoi(e) == 0);
  }();
  return on;
}

static int zero_copy_max_batch() {
  static const int n = [] {
    const char* e = getenv("VIDSTAB_AMD_ZERO_COPY_MAX_BATCH");
    return e && *e ? atoi(e) : 1 << 30;
  }();
  return n;
}

int vsg_ctx_init(VSGContext* c, int device, int width, int height, int bpp, int blurSize, const VSGStageDesc* coarse,
                 const VSGStageDesc* fine, int B, const hipStream_t* shared) {
  if (bpp != 1 && bpp != 3 && bpp != 4) return vsg_fail("vsg_create: bytes per pixel", hipSuccess);
  if (bpp != 1 && blurSize >= 2) return vsg_fail("vsg_create: packed frames are not blurred", hipSuccess);
  c->device = device;
  VSG_CHECK(hipSetDevice(device));
  c->ownStreams = shared == nullptr;
  /* the slot sets share one compute stream: the device runs one batch stage
     at a time, in submission order, each with the whole chip.  Measured on
     the 1080p and 4K clips, the throughput equals that of one stream per set
     (whose stages overlap and stretch each other), and every search launch
     runs unshared.  VIDSTAB_AMD_STREAMS=3 gives each set its own stream. */
  const char* ns = getenv("VIDSTAB_AMD_STREAMS");
  const bool oneStream = !(ns && atoi(ns) == VSG_NSETS);
  for (int k = 0; k < VSG_NSETS; ++k) {
    if (shared)
      c->set[k].stream = shared[k];
    else if (k > 0 && oneStream)
      c->set[k].stream = c->set[0].stream;
    else
      VSG_CHECK(hipStreamCreateWithFlags(&c->set[k].stream, hipStreamNonBlocking));
  }
  c->oneStream = true;
  for (int k = 1; k < VSG_NSETS; ++k) c->oneStream = c->oneStream && c->set[k].stream == c->set[0].stream;
  c->W = width;
  c->H = height;
  c->bpp = bpp;
  c->Wb = width * bpp;
  c->nmm = bpp == 1 ? 2 : 6;
  c->pitch = (c->Wb + 63) & ~63;
  c->blurSize = blurSize;
  c->B = B;
  const VSGStageDesc* sd[2] = {coarse, fine};
  int maxStep = 1;
  size_t patchBytes = 0;
  std::vector<int16_t> ax, ay, asz;
  std::vector<uint32_t> poff;
  for (int k = 0; k < 2; ++k) {
    VSGStage& S = c->st[k];
    S.S = sd[k]->fieldSize;
    S.maxShift = sd[k]->maxShift;
    S.step = sd[k]->stepSize;
    S.nf = sd[k]->nfields;
    S.rows = sd[k]->fieldRows;
    S.maxFields = sd[k]->maxFields;
    S.ct = sd[k]->contrastThreshold;
    if (S.S % 16 != 0 || S.S <= 0 || S.S > 256 || S.step < 1 || S.nf <= 0 || S.nf > 4096 || S.maxShift < 0 ||
        S.rows < 1 || S.rows + 1 > 256)
      return vsg_fail("vsg_create: field grid", hipSuccess);
    maxStep = S.step > maxStep ? S.step : maxStep;
    S.tiles = vsg_plan_model(&S.g, S.S, S.step, S.maxShift, bpp);
    S.candsPerJob = (double)S.g.G * S.g.G;
    int rc = 0;
    if (bpp == 1) {
      for (int st = S.step; st > 1; st /= 2) {
        const int ns = st / 2, r = st - ns, ni = (2 * r) / ns + 1;
        rc += ni * ni - ((r % ns) == 0 ? 1 : 0);
      }
    } else {
      /* the unshifted position, then (2r+1)^2 - 1 shifts, r = step - 1 */
      rc = 1 + (2 * S.step - 1) * (2 * S.step - 1) - 1;
    }
    S.refinePerJob = rc;
    S.patchOff = patchBytes;
    std::vector<int16_t> fx(S.nf), fy(S.nf);
    for (int i = 0; i < S.nf; ++i) {
      const Field& f = sd[k]->fields[i];
      if (f.size != S.S) return vsg_fail("vsg_create: field size mismatch", hipSuccess);
      fx[i] = f.x;
      fy[i] = f.y;
      ax.push_back(f.x);
      ay.push_back(f.y);
      asz.push_back(f.size);
      poff.push_back((uint32_t)(patchBytes + (size_t)i * S.S * S.S * bpp));
    }
    patchBytes += (size_t)S.nf * S.S * S.S * bpp;
    VSG_CHECK(dalloc(&S.d_fx, S.nf));
    VSG_CHECK(dalloc(&S.d_fy, S.nf));
    VSG_CHECK(hipMemcpy(S.d_fx, fx.data(), sizeof(int16_t) * S.nf, hipMemcpyHostToDevice));
    VSG_CHECK(hipMemcpy(S.d_fy, fy.data(), sizeof(int16_t) * S.nf, hipMemcpyHostToDevice));
  }
  c->nfTotal = (int)ax.size();
  VSG_CHECK(dalloc(&c->d_allx, c->nfTotal));
  VSG_CHECK(dalloc(&c->d_ally, c->nfTotal));
  VSG_CHECK(dalloc(&c->d_allsz, c->nfTotal));
  VSG_CHECK(dalloc(&c->d_patchOff, c->nfTotal));
  VSG_CHECK(hipMemcpy(c->d_allx, ax.data(), sizeof(int16_t) * c->nfTotal, hipMemcpyHostToDevice));
  VSG_CHECK(hipMemcpy(c->d_ally, ay.data(), sizeof(int16_t) * c->nfTotal, hipMemcpyHostToDevice));
  VSG_CHECK(hipMemcpy(c->d_allsz, asz.data(), sizeof(int16_t) * c->nfTotal, hipMemcpyHostToDevice));
  VSG_CHECK(hipMemcpy(c->d_patchOff, poff.data(), sizeof(uint32_t) * c->nfTotal, hipMemcpyHostToDevice));

  /* contrast of every (min, max) pair, evaluated on the host exactly as the
     reference writes it: (max - min) / (max + min + 0.1) */
  std::vector<double> ctab(65536);
  for (int mn = 0; mn < 256; ++mn)
    for (int mx = 0; mx < 256; ++mx) ctab[mn * 256 + mx] = (double)(mx - mn) / ((double)(mx + mn) + 0.1);
  VSG_CHECK(dalloc(&c->d_ctab, 65536));
  VSG_CHECK(hipMemcpy(c->d_ctab, ctab.data(), sizeof(double) * 65536, hipMemcpyHostToDevice));

  /* pad rows: a lane's K <= 8 candidate rows may reach (K-1)*step rows below
     a search window, plus column overrun into the next row */
  c->rowsAlloc = height + 8 * maxStep + 16;
  c->slotStride = (size_t)c->pitch * c->rowsAlloc;
  const int slots = VSG_NSETS * B;
  VSG_CHECK(dalloc(&c->d_raw, (size_t)slots * c->pitch * height));
  VSG_CHECK(dalloc(&c->d_blur, (size_t)slots * c->slotStride));
  VSG_CHECK(dalloc(&c->d_ref, c->slotStride));
  VSG_CHECK(hipMemset(c->d_blur, 0, (size_t)slots * c->slotStride));
  VSG_CHECK(hipMemset(c->d_ref, 0, c->slotStride));
  /* the two-pass blur of wide boxes keeps one intermediate plane per slot:
     the sets' streams blur concurrently, each into its own slots' planes */
  if (vsg_blur_needs_tmp(c->Wb, blurSize)) VSG_CHECK(dalloc(&c->d_tmp, (size_t)slots * c->pitch * height));
  c->patchSlotStride = (patchBytes + 255) & ~(size_t)255;
  VSG_CHECK(dalloc(&c->d_patches, (size_t)slots * c->patchSlotStride));
  VSG_CHECK(dalloc(&c->d_minmax, (size_t)slots * c->nfTotal * c->nmm));
  const int nc = c->st[0].nf, nf = c->st[1].nf, nmax = nc > nf ? nc : nf;
  for (VSGSet& s : c->set) {
    VSG_CHECK(dalloc(&s.d_selIdx, (size_t)B * c->nfTotal));
    VSG_CHECK(dalloc(&s.d_selC, (size_t)B * c->nfTotal));
    VSG_CHECK(dalloc(&s.d_count, (size_t)B * 2));
    VSG_CHECK(dalloc(&s.d_jobsC, (size_t)B * nc));
    VSG_CHECK(dalloc(&s.d_jobsF, (size_t)B * nf));
    VSG_CHECK(dalloc(&s.d_best, (size_t)B * nmax));
    VSG_CHECK(hipMemset(s.d_best, 0xFF, sizeof(unsigned long long) * (size_t)B * nmax));
    VSG_CHECK(dalloc(&s.d_done, (size_t)B * nmax));
    VSG_CHECK(hipMemset(s.d_done, 0, sizeof(unsigned) * (size_t)B * nmax));

    VSG_CHECK(dalloc(&s.d_resC, (size_t)B * nc));
    VSG_CHECK(dalloc(&s.d_resF, (size_t)B * nf));
    /* the per-frame descriptors (source plane, previous slot, detect flag)
       are one block, uploaded with one copy per batch */
    VSG_CHECK(dalloc(&s.d_desc, desc_bytes(B)));
    s.d_src = reinterpret_cast<VSGBlurSrc*>(s.d_desc);
    s.d_prev = reinterpret_cast<int*>(s.d_desc + sizeof(VSGBlurSrc) * B);
    s.d_detect = s.d_desc + (sizeof(VSGBlurSrc) + sizeof(int)) * B;
    VSG_CHECK(halloc(&s.h_selIdx, (size_t)B * c->nfTotal));
    VSG_CHECK(halloc(&s.h_selC, (size_t)B * c->nfTotal));
    VSG_CHECK(halloc(&s.h_count, (size_t)B * 2));
    VSG_CHECK(halloc(&s.h_resC, (size_t)B * nc));
    VSG_CHECK(halloc(&s.h_jobsF, (size_t)B * nf));
    VSG_CHECK(halloc(&s.h_resF, (size_t)B * nf));
    VSG_CHECK(halloc(&s.h_desc, desc_bytes(B)));
    s.h_src = reinterpret_cast<VSGBlurSrc*>(s.h_desc);
    s.h_prev = reinterpret_cast<int*>(s.h_desc + sizeof(VSGBlurSrc) * B);
    s.h_detect = s.h_desc + (sizeof(VSGBlurSrc) + sizeof(int)) * B;
    VSG_CHECK(hmap(&s.m_selIdx, s.h_selIdx));
    VSG_CHECK(hmap(&s.m_selC, s.h_selC));
    VSG_CHECK(hmap(&s.m_count, s.h_count));
    VSG_CHECK(hmap(&s.m_resC, s.h_resC));
    VSG_CHECK(hmap(&s.m_jobsF, s.h_jobsF));
    VSG_CHECK(hmap(&s.m_resF, s.h_resF));
    VSG_CHECK(hmap(&s.m_desc, s.h_desc));
    VSG_CHECK(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
    VSG_CHECK(hipEventCreateWithFlags(&s.doneF, hipEventDisableTiming));
    VSG_CHECK(hipEventCreateWithFlags(&s.planesReady, hipEventDisableTiming));
    VSG_CHECK(hipEventCreateWithFlags(&s.searchDone, hipEventDisableTiming));
    VSG_CHECK(hipEventCreateWithFlags(&s.uploaded, hipEventDisableTiming));
    for (hipEvent_t& e : s.ev) VSG_CHECK(hipEventCreate(&e));
    for (hipEvent_t& e : s.evf) VSG_CHECK(hipEventCreate(&e));
  }
  VSG_CHECK(hipMemset(c->d_patches, 0, (size_t)slots * c->patchSlotStride));
  VSG_CHECK(hipDeviceSynchronize());
  return vsg_ctx_tune(c);
}

extern "C" int vsg_create_fmt(VSGContext** out, int device, int width, int height, int bytesPerPixel,
                              int blurSize, const VSGStageDesc* coarse, const VSGStageDesc* fine, int maxBatch) {
  if (!out || !coarse || !fine || width <= 0 || height <= 0 || maxBatch <= 0)
    return vsg_fail("vsg_create", hipSuccess);
  *out = nullptr;
  VSGContext* c = new (std::nothrow) VSGContext();
  if (!c) return vsg_fail("vsg_create: out of host memory", hipSuccess);
  if (vsg_ctx_init(c, device, width, height, bytesPerPixel, blurSize, coarse, fine, maxBatch, nullptr) != 0) {
    char keep[sizeof(g_err)];
    memcpy(keep, g_err, sizeof(keep));
    free_ctx(c);
    memcpy(g_err, keep, sizeof(keep));
    return -1;
  }
  *out = c;
  return 0;
}

extern "C" int vsg_create(VSGContext** out, int device, int width, int height, int blurSize,
                          const VSGStageDesc* coarse, const VSGStageDesc* fine, int maxBatch) {
  return vsg_create_fmt(out, device, width, height, 1, blurSize, coarse, fine, maxBatch);
}

/* ---- batches ------------------------------------------------------------------ */

VSGSearchArgs vsg_ctx_search_args(VSGContext* c, int stage, const VSGJob* jobs, int njobs,
                                  unsigned long long* best, unsigned* done, VSGResult* res) {
  const VSGStage& S = c->st[stage];
  VSGSearchArgs a;
  a.jobs = jobs;
  a.njobs = njobs;
  a.fx = S.d_fx;
  a.fy = S.d_fy;
  a.blurred = c->d_blur;
  a.slotStride = c->slotStride;
  a.ref = c->d_ref;
  a.pitch = c->pitch;
  a.patches = c->d_patches;
  a.patchSlotStride = c->patchSlotStride;
  a.stageOff = S.patchOff;
  a.g = S.g;
  a.tiles = S.tiles;
  /* one frame's worth of jobs per XCD chunk: its planes stay in one L2 */
  a.chunk = S.nf * S.tiles;
  a.best = best;
  a.done = done;
  a.results = res;
  return a;
}

/* the pinned per-frame descriptors of set `set` (h_src / h_prev / h_detect)
   must be filled for n frames; enqueues their upload, the blur, the field pass
   and the selection */
int vsg_ctx_upload_front(VSGContext* c, int set, int n) {
  VSGSet& s = c->set[set];
  (void)n;   /* the whole block: one small copy instead of three */
  VSG_CHECK(hipMemcpyAsync(s.d_desc, s.h_desc, desc_bytes(c->B), hipMemcpyHostToDevice, s.stream));
  return 0;
}

int vsg_ctx_enqueue_front(VSGContext* c, int set, int n, bool upload) {
  VSGSet& s = c->set[set];
  hipStream_t stream = s.stream;
  const int slot0 = set * c->B;
  /* zero-copy batch: the kernels read the descriptors in pinned memory */
  const VSGBlurSrc* src = s.zc ? reinterpret_cast<const VSGBlurSrc*>(s.m_desc) : s.d_src;
  const int* prevSlot = s.zc ? reinterpret_cast<const int*>(s.m_desc + sizeof(VSGBlurSrc) * c->B) : s.d_prev;
  const uint8_t* detect = s.zc ? s.m_desc + (sizeof(VSGBlurSrc) + sizeof(int)) * c->B : s.d_detect;
  if (upload && !s.zc && vsg_ctx_upload_front(c, set, n) != 0) return -1;
  if (s.prof) VSG_CHECK(hipEventRecord(s.ev[VSG_EV_BLUR0], stream));
  /* packed frames are not blurred (src/motiondetect.c:197-201): their blurSize
     is below 2, which copies Wb = W*bpp bytes a row */
  uint8_t* tmp = c->d_tmp ? c->d_tmp + (size_t)slot0 * c->pitch * c->H : nullptr;
  VSG_CHECK(vsg_launch_blur(src, n, c->d_blur + (size_t)slot0 * c->slotStride, c->slotStride, c->pitch,
                            c->Wb, c->H, c->blurSize, tmp, stream));
  if (s.prof) VSG_CHECK(hipEventRecord(s.ev[VSG_EV_BLUR1], stream));
  VSGFieldArgs fa;
  fa.blurred = c->d_blur + (size_t)slot0 * c->slotStride;
  fa.slotStride = c->slotStride;
  fa.pitch = c->pitch;
  fa.fx = c->d_allx;
  fa.fy = c->d_ally;
  fa.fsz = c->d_allsz;
  fa.patchOff = c->d_patchOff;
  fa.nfields = c->nfTotal;
  fa.patches = c->d_patches + (size_t)slot0 * c->patchSlotStride;
  fa.patchSlotStride = c->patchSlotStride;
  fa.minmax = c->d_minmax + (size_t)slot0 * c->nfTotal * c->nmm;
  fa.src = src;
  fa.bpp = c->bpp;
  fa.nmm = c->nmm;
  VSG_CHECK(vsg_launch_fields(fa, n, stream));
  if (s.prof) VSG_CHECK(hipEventRecord(s.ev[VSG_EV_FIELDS1], stream));
  VSGSelectArgs sa;
  sa.minmax = c->d_minmax;
  sa.nfTotal = c->nfTotal;
  sa.nmm = c->nmm;
  sa.ctab = c->d_ctab;
  for (int k = 0; k < 2; ++k) {
    sa.st[k].nf = c->st[k].nf;
    sa.st[k].rows = c->st[k].rows;
    sa.st[k].maxFields = c->st[k].maxFields;
    sa.st[k].mmOff = k == 0 ? 0 : c->st[0].nf;
    sa.st[k].ct = c->st[k].ct;
  }
  sa.slot0 = slot0;
  sa.prevSlot = prevSlot;
  sa.detect = detect;
  sa.selIdx = s.zc ? s.m_selIdx : s.d_selIdx;
  sa.selC = s.zc ? s.m_selC : s.d_selC;
  sa.count = s.zc ? s.m_count : s.d_count;
  sa.jobs = s.d_jobsC;
  VSG_CHECK(vsg_launch_select(sa, n, stream));
  if (s.prof) VSG_CHECK(hipEventRecord(s.ev[VSG_EV_SELECT1], stream));
  return 0;
}

int vsg_ctx_enqueue_coarse_readback(VSGContext* c, int set, int n) {
  VSGSet& s = c->set[set];
  hipStream_t stream = s.stream;
  const int nc = c->st[0].nf;
  VSG_CHECK(hipMemcpyAsync(s.h_count, s.d_count, sizeof(int) * 2 * n, hipMemcpyDeviceToHost, stream));
  VSG_CHECK(hipMemcpyAsync(s.h_selIdx, s.d_selIdx, sizeof(int16_t) * (size_t)n * c->nfTotal,
                           hipMemcpyDeviceToHost, stream));
  VSG_CHECK(hipMemcpyAsync(s.h_selC, s.d_selC, sizeof(double) * (size_t)n * c->nfTotal,
                           hipMemcpyDeviceToHost, stream));
  VSG_CHECK(hipMemcpyAsync(s.h_resC, s.d_resC, sizeof(VSGResult) * (size_t)n * nc, hipMemcpyDeviceToHost,
                           stream));
  return 0;
}

extern "C" int vsg_stage_frames(VSGContext* c, int set, int n, uint8_t** stage, int* pitch) {
  if (!c || set < 0 || set >= VSG_NSETS || n <= 0 || n > c->B || !stage || !pitch)
    return vsg_fail("vsg_stage_frames", hipSuccess);
  VSGSet& s = c->set[set];
  VSG_CHECK(hipSetDevice(c->device));
  /* the set's previous batch (its uploads and read-backs) must be over */
  VSG_CHECK(hipEventSynchronize(s.done));
  VSG_CHECK(hipEventSynchronize(s.doneF));
  const size_t plane = (size_t)c->pitch * c->H;
  if (!s.h_stage) {
    VSG_CHECK(halloc(&s.h_stage, plane * c->B));
    VSG_CHECK(hmap(&s.m_stage, s.h_stage));
  }
  for (int i = 0; i < n; ++i) stage[i] = s.h_stage + plane * i;
  *pitch = c->pitch;
  return 0;
}

extern "C" int vsg_stage_upload(VSGContext* c, int set, size_t offset, size_t bytes) {
  if (!c || set < 0 || set >= VSG_NSETS) return vsg_fail("vsg_stage_upload", hipSuccess);
  VSGSet& s = c->set[set];
  const size_t plane = (size_t)c->pitch * c->H;
  if (!s.m_stage || offset % (size_t)c->pitch || bytes % (size_t)c->pitch || offset + bytes > plane * c->B)
    return vsg_fail("vsg_stage_upload: range", hipSuccess);
  VSG_CHECK(hipSetDevice(c->device));
  VSG_CHECK(vsg_launch_upload(s.m_stage + offset, c->d_raw + (size_t)set * c->B * plane + offset, bytes, s.stream));
  return 0;
}

extern "C" int vsg_submit(VSGContext* c, int set, int n, const uint8_t* const* luma, const int* linesize,
                          int onDevice, const int* prevSlot, const uint8_t* detect, int refSlot) {
  if (!c || set < 0 || set >= VSG_NSETS || n <= 0 || n > c->B || !luma || !linesize || !prevSlot || !detect)
    return vsg_fail("vsg_submit", hipSuccess);
  const int slot0 = set * c->B;
  if (refSlot >= 0 && (refSlot < slot0 || refSlot >= slot0 + n))
    return vsg_fail("vsg_submit: refSlot outside the batch", hipSuccess);
  VSGSet& s = c->set[set];
  hipStream_t stream = s.stream;
  VSG_CHECK(hipSetDevice(c->device));
  /* this set's pinned buffers are reused below: wait until the previous batch
     of the set (its coarse read-back and its fine stage) has completed */
  VSG_CHECK(hipEventSynchronize(s.done));
  VSG_CHECK(hipEventSynchronize(s.doneF));
  for (int i = 0; i < n; ++i) {
    if (linesize[i] < c->Wb) return vsg_fail("vsg_submit: linesize < width * bytes per pixel", hipSuccess);
    if (prevSlot[i] >= VSG_NSETS * c->B) return vsg_fail("vsg_submit: prevSlot", hipSuccess);
  }
  /* the set's planes are overwritten below; the batch that followed the set's
     previous batch reads its last plane as the previous frame, so its latest
     search must be over first */
  if (s.follower >= 0 && !c->oneStream) VSG_CHECK(hipStreamWaitEvent(stream, c->set[s.follower].searchDone, 0));
  /* host frames of a large batch go up on a stream of their own, beside the
     compute stream's earlier batches (the set's raw planes are free: its
     previous batch completed above); a few frames go in order */
  s.zc = zero_copy_enabled() && n <= zero_copy_max_batch() && !c->learned && !s.borrowed;
  /* the frames of a zero-copy batch of fewer than kUploadAsideMin frames go
     up by the copy kernel */
  const bool kernelUpload = s.zc && n < kUploadAsideMin;
  hipStream_t up = stream;
  if (onDevice != VSG_FRAMES_DEVICE && onDevice != VSG_FRAMES_UPLOADED && n >= kUploadAsideMin) {
    if (!c->upStream) VSG_CHECK(hipStreamCreateWithFlags(&c->upStream, hipStreamNonBlocking));
    up = c->upStream;
  }
  for (int i = 0; i < n; ++i) {
    if (onDevice == VSG_FRAMES_DEVICE) {
      s.h_src[i].p = luma[i];
      s.h_src[i].pitch = linesize[i];
    } else if (onDevice == VSG_FRAMES_STAGED || onDevice == VSG_FRAMES_UPLOADED) {
      /* the caller copied the frame into the set's staging area (and, for
         VSG_FRAMES_UPLOADED, queued its upload with vsg_stage_upload) */
      const size_t plane = (size_t)c->pitch * c->H;
      if (!s.h_stage) return vsg_fail("vsg_submit: frames not staged", hipSuccess);
      uint8_t* dst = c->d_raw + (size_t)(slot0 + i) * plane;
      if (!kernelUpload && onDevice == VSG_FRAMES_STAGED)
        VSG_CHECK(hipMemcpyAsync(dst, s.h_stage + plane * i, plane, hipMemcpyHostToDevice, up));
      s.h_src[i].p = dst;
      s.h_src[i].pitch = c->pitch;
    } else {
      /* pageable host frame: rows into the set's pinned staging area (the
         set's previous batch is complete), then one DMA of the plane */
      const size_t plane = (size_t)c->pitch * c->H;
      if (!s.h_stage) {
        VSG_CHECK(halloc(&s.h_stage, plane * c->B));
        VSG_CHECK(hmap(&s.m_stage, s.h_stage));
      }
      uint8_t* stg = s.h_stage + plane * i;
      if (linesize[i] == c->pitch) {
        memcpy(stg, luma[i], plane - (c->pitch - c->Wb));
      } else {
        for (int y = 0; y < c->H; ++y) memcpy(stg + (size_t)y * c->pitch, luma[i] + (size_t)y * linesize[i], c->Wb);
      }
      uint8_t* dst = c->d_raw + (size_t)(slot0 + i) * plane;
      if (!kernelUpload) VSG_CHECK(hipMemcpyAsync(dst, stg, plane, hipMemcpyHostToDevice, up));
      s.h_src[i].p = dst;
      s.h_src[i].pitch = c->pitch;
    }
    s.h_src[i].pad = 0;
    s.h_prev[i] = prevSlot[i];
    s.h_detect[i] = detect[i] ? 1 : 0;
  }
  if (up != stream) {
    VSG_CHECK(hipEventRecord(s.uploaded, up));
    VSG_CHECK(hipStreamWaitEvent(stream, s.uploaded, 0));
  }
  if (kernelUpload && (onDevice == VSG_FRAMES_STAGED || onDevice == VSG_FRAMES_HOST)) {
    /* the staged planes are consecutive, as are the batch's raw slots */
    const size_t plane = (size_t)c->pitch * c->H;
    VSG_CHECK(vsg_launch_upload(s.m_stage, c->d_raw + (size_t)slot0 * plane, plane * n, stream));
  }
  s.prof = g_profOn.load() != 0;
  /* learned form: the quarter-resolution planes first; then, with no
     earlier network still running, the batch's network goes ahead of its
     blur, field copies and selection (which need not wait for it); else the
     front stage runs beside the earlier network and this one follows it.
     Either way the network waits only for the previous network (whose
     batch's planes hold this batch's first previous frame), not for other
     sets' blurred planes. */
  if (c->learned) {
    if (vsg_ctx_upload_front(c, set, n) != 0 || vsg_learned_front(c, set, n, refSlot, stream) != 0) return -1;
    const bool netFirst = !vsg_learned_net_busy(c);
    if (netFirst && vsg_learned_net(c, set, n, stream) != 0) return -1;
    if (vsg_ctx_enqueue_front(c, set, n, false) != 0) return -1;
    if (!netFirst && vsg_learned_net(c, set, n, stream) != 0) return -1;
  } else if (vsg_ctx_enqueue_front(c, set, n, true) != 0) {
    return -1;
  }
  if (refSlot >= 0) {
    /* the frozen tripod plane: after the blur, before any search reads it */
    VSG_CHECK(hipMemcpyAsync(c->d_ref, c->d_blur + (size_t)refSlot * c->slotStride, c->slotStride,
                             hipMemcpyDeviceToDevice, stream));
  }
  /* planes of earlier batches (previous frames, the tripod plane) come from
     other sets' streams: planesReady chains them, so one wait covers all.
     On one shared stream the stream order already covers them, and the
     cross-set events are neither recorded nor waited for. */
  if (c->lastSet >= 0 && c->lastSet != set) {
    if (!c->oneStream) VSG_CHECK(hipStreamWaitEvent(stream, c->set[c->lastSet].planesReady, 0));
    c->set[c->lastSet].follower = set;
  }
  if (!c->oneStream) VSG_CHECK(hipEventRecord(s.planesReady, stream));
  c->lastSet = set;
  s.follower = -1;
  if (s.prof) VSG_CHECK(hipEventRecord(s.ev[VSG_EV_GRID0], stream));
  const int nc = c->st[0].nf;
  if (c->learned) {
    if (vsg_learned_coarse(c, set, n, stream) != 0) return -1;
  } else {
    VSGSearchArgs ga = vsg_ctx_search_args(c, 0, s.d_jobsC, n * nc, s.d_best, s.d_done, s.zc ? s.m_resC : s.d_resC);
    VSG_CHECK(vsg_launch_search_any(ga, stream));
  }
  if (s.prof) VSG_CHECK(hipEventRecord(s.ev[VSG_EV_GRID1], stream));
  if (!c->oneStream) VSG_CHECK(hipEventRecord(s.searchDone, stream));
  if (!s.zc && vsg_ctx_enqueue_coarse_readback(c, set, n) != 0) return -1;
  VSG_CHECK(hipEventRecord(s.done, stream));
  s.n = n;
  return 0;
}

extern "C" int vsg_wait_coarse(VSGContext* c, int set, VSGBatchView* view) {
  if (!c || set < 0 || set >= VSG_NSETS || !view) return vsg_fail("vsg_wait_coarse", hipSuccess);
  VSGSet& s = c->set[set];
  VSG_CHECK(hipSetDevice(c->device));
  VSG_CHECK(hipEventSynchronize(s.done));
  vsg_learned_account(c, set, 0);
  if (s.prof) {
    s.prof = false;
    double fieldBytes = 0;
    for (const VSGStage& S : c->st) fieldBytes += (double)S.nf * S.S * S.S * c->bpp;
    int ncoarse = 0;
    for (int i = 0; i < s.n; ++i) ncoarse += s.h_count[2 * i];
    prof_add(VSG_K_BLUR, s.ev[VSG_EV_BLUR0], s.ev[VSG_EV_BLUR1], 2.0 * c->Wb * c->H * s.n);
    prof_add(VSG_K_FIELDS, s.ev[VSG_EV_BLUR1], s.ev[VSG_EV_FIELDS1], 2.0 * fieldBytes * s.n);
    prof_add(VSG_K_SELECT, s.ev[VSG_EV_FIELDS1], s.ev[VSG_EV_SELECT1], (double)c->nfTotal * s.n);
    const double px = (double)c->st[0].S * c->st[0].S * c->bpp;
    prof_add(VSG_K_GRID_COARSE, s.ev[VSG_EV_GRID0], s.ev[VSG_EV_GRID1],
             ncoarse * (c->st[0].candsPerJob + c->st[0].refinePerJob) * px);
  }
  view->count = s.h_count;
  view->idx = s.h_selIdx;
  view->contrast = s.h_selC;
  view->coarse = s.h_resC;
  view->stride = c->nfTotal;
  view->ncoarse = c->st[0].nf;
  return 0;
}

extern "C" int vsg_submit_fine(VSGContext* c, int set, int njobs, const VSGJob* jobs) {
  if (!c || set < 0 || set >= VSG_NSETS || njobs < 0) return vsg_fail("vsg_submit_fine", hipSuccess);
  VSGSet& s = c->set[set];
  const VSGStage& S = c->st[1];
  if (njobs > c->B * S.nf) return vsg_fail("vsg_submit_fine: too many jobs", hipSuccess);
  for (int i = 0; i < njobs; ++i) {
    const VSGJob& j = jobs[i];
    if (j.slot < 0 || j.slot >= VSG_NSETS * c->B || j.prev >= VSG_NSETS * c->B || j.field >= S.nf)
      return vsg_fail("vsg_submit_fine: job out of range", hipSuccess);
  }
  VSG_CHECK(hipSetDevice(c->device));
  s.njobsF = njobs;
  s.profF = false;
  if (njobs == 0) {
    VSG_CHECK(hipEventRecord(s.doneF, s.stream));
    return 0;
  }
  if (jobs != s.h_jobsF) memcpy(s.h_jobsF, jobs, sizeof(VSGJob) * njobs);
  if (!s.zc) VSG_CHECK(hipMemcpyAsync(s.d_jobsF, s.h_jobsF, sizeof(VSGJob) * njobs, hipMemcpyHostToDevice, s.stream));
  s.profF = g_profOn.load() != 0;
  if (s.profF) VSG_CHECK(hipEventRecord(s.evf[VSG_EVF_0], s.stream));
  if (vsg_learned_fine_enabled(c)) {
    if (vsg_learned_fine(c, set, njobs, s.stream) != 0) return -1;
  } else {
    VSGSearchArgs fa = vsg_ctx_search_args(c, 1, s.zc ? s.m_jobsF : s.d_jobsF, njobs, s.d_best, s.d_done,
                                           s.zc ? s.m_resF : s.d_resF);
    VSG_CHECK(vsg_launch_search_any(fa, s.stream));
  }
  if (s.profF) VSG_CHECK(hipEventRecord(s.evf[VSG_EVF_GRID1], s.stream));
  if (!c->oneStream) VSG_CHECK(hipEventRecord(s.searchDone, s.stream));
  if (!s.zc)
    VSG_CHECK(hipMemcpyAsync(s.h_resF, s.d_resF, sizeof(VSGResult) * njobs, hipMemcpyDeviceToHost, s.stream));
  VSG_CHECK(hipEventRecord(s.doneF, s.stream));
  return 0;
}

extern "C" int vsg_wait_fine(VSGContext* c, int set, const VSGResult** results) {
  if (!c || set < 0 || set >= VSG_NSETS || !results) return vsg_fail("vsg_wait_fine", hipSuccess);
  VSGSet& s = c->set[set];
  VSG_CHECK(hipSetDevice(c->device));
  VSG_CHECK(hipEventSynchronize(s.doneF));
  if (s.njobsF > 0) vsg_learned_account(c, set, 1);
  if (s.profF) {
    s.profF = false;
    const VSGStage& S = c->st[1];
    const double px = (double)S.S * S.S * c->bpp;
    prof_add(VSG_K_GRID_FINE, s.evf[VSG_EVF_0], s.evf[VSG_EVF_GRID1],
             s.njobsF * (S.candsPerJob + S.refinePerJob) * px);
  }
  *results = s.h_resF;
  return 0;
}

extern "C" int vsg_copy_plane(int device, uint8_t* dst, int dstPitch, const uint8_t* src, int srcPitch, int rowBytes,
                              int rows) {
  if (!dst || !src || rowBytes <= 0 || rows <= 0 || dstPitch < rowBytes || srcPitch < rowBytes)
    return vsg_fail("vsg_copy_plane", hipSuccess);
  VSG_CHECK(hipSetDevice(device));
  VSG_CHECK(hipMemcpy2D(dst, dstPitch, src, srcPitch, rowBytes, rows, hipMemcpyDefault));
  return 0;
}

extern "C" int vsg_read_blurred(VSGContext* c, int slot, uint8_t* dst, int dstStride) {
  if (!c || slot >= VSG_NSETS * c->B || !dst || dstStride < c->Wb) return vsg_fail("vsg_read_blurred", hipSuccess);
  VSG_CHECK(hipSetDevice(c->device));
  const uint8_t* src = slot < 0 ? c->d_ref : c->d_blur + (size_t)slot * c->slotStride;
  if (vsg_ctx_sync(c) != 0) return -1;
  VSG_CHECK(hipMemcpy2D(dst, dstStride, src, c->pitch, c->Wb, c->H, hipMemcpyDeviceToHost));
  return 0;
}
