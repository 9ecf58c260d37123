/*
 * vsg_api.hip -- the thin C ABI over the detect kernels (declared in vsg.h):
 * device buffers, the HIP stream of one detector, uploads and read-backs.
 *
 * HBM layout of one context (pitch = width rounded up to 64 bytes):
 *   raw[maxBatch]      uploaded luma planes, pitch x H
 *   blur[maxBatch]     blurred planes, pitch x (H + padRows); the pad rows keep
 *                      the grid kernel's idle-candidate reads inside the buffer
 *   ref                the stored previous / tripod reference plane (same shape)
 *   patches[maxBatch]  an aligned copy of every coarse and fine field
 *   minmax[maxBatch]   8-bit (min, max) per field
 *   jobs / best / results   per search job
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <mutex>
#include <new>
#include <vector>

#include "vsg_internal.h"

struct VSGContext {
  int device = 0;
  hipStream_t stream = nullptr;
  int W = 0, H = 0, pitch = 0, rowsAlloc = 0;
  size_t slotStride = 0;
  int blurSize = 0;
  int maxBatch = 0;
  struct Stage {
    int S = 0, maxShift = 0, step = 0, nf = 0;
    int16_t* d_fx = nullptr;
    int16_t* d_fy = nullptr;
    VSGGrid g{};
    int tiles = 0;
    size_t patchOff = 0;   /* start of this stage's patches inside a slot */
  } st[2];
  int nfTotal = 0;
  int16_t* d_allx = nullptr;
  int16_t* d_ally = nullptr;
  int16_t* d_allsz = nullptr;
  uint32_t* d_patchOff = nullptr;
  uint8_t* d_raw = nullptr;
  uint8_t* d_blur = nullptr;
  uint8_t* d_ref = nullptr;
  uint8_t* d_tmp = nullptr;
  uint8_t* d_patches = nullptr;
  size_t patchSlotStride = 0;
  uint8_t* d_minmax = nullptr;
  uint8_t* h_minmax = nullptr;
  int maxJobs = 0;
  VSGJob* d_jobs = nullptr;
  VSGJob* h_jobs = nullptr;
  unsigned long long* d_best = nullptr;
  VSGResult* d_res = nullptr;
  VSGResult* h_res = nullptr;
  VSGBlurSrc* d_src = nullptr;
  VSGBlurSrc* h_src = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
};

static thread_local char g_err[512];

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

/* after the stream has been synchronised: account ev[0]->ev[1] to kind k0 and
   ev[1]->ev[2] to kind k1 */
static void prof_account(VSGContext* c, int k0, double w0, int k1, double w1) {
  float t0 = 0.f, t1 = 0.f;
  if (hipEventElapsedTime(&t0, c->ev[0], c->ev[1]) != hipSuccess) return;
  if (hipEventElapsedTime(&t1, c->ev[1], c->ev[2]) != hipSuccess) return;
  std::lock_guard<std::mutex> lk(g_profMu);
  g_prof.launches[k0] += 1;
  g_prof.ms[k0] += t0;
  g_prof.work[k0] += w0;
  g_prof.launches[k1] += 1;
  g_prof.ms[k1] += t1;
  g_prof.work[k1] += w1;
}

static int fail(const char* what, hipError_t e) {
  snprintf(g_err, sizeof(g_err), "%s: %s", what, e == hipSuccess ? "invalid argument" : hipGetErrorString(e));
  return -1;
}

#define VSG_CHECK(call)                                  \
  do {                                                   \
    hipError_t e_ = (call);                              \
    if (e_ != hipSuccess) return fail(#call, e_);        \
  } while (0)

extern "C" const char* vsg_last_error(void) { return g_err; }

extern "C" int vsg_spiral_rank(int a, int b) { return vsg_rank(a, b); }
extern "C" void vsg_spiral_unrank(int rank, int* a, int* b) { vsg_unrank(rank, a, b); }

static void free_ctx(VSGContext* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int s = 0; s < 2; ++s) {
    (void)hipFree(c->st[s].d_fx);
    (void)hipFree(c->st[s].d_fy);
  }
  (void)hipFree(c->d_allx); (void)hipFree(c->d_ally); (void)hipFree(c->d_allsz); (void)hipFree(c->d_patchOff);
  (void)hipFree(c->d_raw); (void)hipFree(c->d_blur); (void)hipFree(c->d_ref); (void)hipFree(c->d_tmp);
  (void)hipFree(c->d_patches); (void)hipFree(c->d_minmax); (void)hipFree(c->d_jobs); (void)hipFree(c->d_best);
  (void)hipFree(c->d_res); (void)hipFree(c->d_src);
  (void)hipHostFree(c->h_minmax); (void)hipHostFree(c->h_jobs); (void)hipHostFree(c->h_res); (void)hipHostFree(c->h_src);
  for (int i = 0; i < 3; ++i)
    if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" void vsg_destroy(VSGContext* ctx) { free_ctx(ctx); }

static int create(VSGContext* c, int device, int width, int height, int blurSize,
                  const VSGStageDesc* coarse, const VSGStageDesc* fine, int maxBatch) {
  c->device = device;
  VSG_CHECK(hipSetDevice(device));
  VSG_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (int i = 0; i < 3; ++i) VSG_CHECK(hipEventCreate(&c->ev[i]));
  c->W = width;
  c->H = height;
  c->pitch = (width + 63) & ~63;
  c->blurSize = blurSize;
  c->maxBatch = maxBatch;
  const VSGStageDesc* sd[2] = {coarse, fine};
  int maxStep = 1;
  size_t patchBytes = 0;
  std::vector<int16_t> ax, ay, asz;
  std::vector<uint32_t> poff;
  for (int s = 0; s < 2; ++s) {
    VSGContext::Stage& S = c->st[s];
    S.S = sd[s]->fieldSize;
    S.maxShift = sd[s]->maxShift;
    S.step = sd[s]->stepSize;
    S.nf = sd[s]->nfields;
    if (S.S % 16 != 0 || S.S <= 0 || S.step < 1 || S.nf <= 0 || S.maxShift < 0)
      return fail("vsg_create: field grid", hipSuccess);
    maxStep = S.step > maxStep ? S.step : maxStep;
    S.tiles = vsg_plan_grid(&S.g, S.S, S.step, S.maxShift);
    S.patchOff = patchBytes;
    std::vector<int16_t> fx(S.nf), fy(S.nf);
    for (int i = 0; i < S.nf; ++i) {
      const Field& f = sd[s]->fields[i];
      if (f.size != S.S) return fail("vsg_create: field size mismatch", hipSuccess);
      fx[i] = f.x;
      fy[i] = f.y;
      ax.push_back(f.x);
      ay.push_back(f.y);
      asz.push_back(f.size);
      poff.push_back((uint32_t)(patchBytes + (size_t)i * S.S * S.S));
    }
    patchBytes += (size_t)S.nf * S.S * S.S;
    VSG_CHECK(hipMalloc(&S.d_fx, sizeof(int16_t) * S.nf));
    VSG_CHECK(hipMalloc(&S.d_fy, sizeof(int16_t) * S.nf));
    VSG_CHECK(hipMemcpy(S.d_fx, fx.data(), sizeof(int16_t) * S.nf, hipMemcpyHostToDevice));
    VSG_CHECK(hipMemcpy(S.d_fy, fy.data(), sizeof(int16_t) * S.nf, hipMemcpyHostToDevice));
  }
  c->nfTotal = (int)ax.size();
  VSG_CHECK(hipMalloc(&c->d_allx, sizeof(int16_t) * c->nfTotal));
  VSG_CHECK(hipMalloc(&c->d_ally, sizeof(int16_t) * c->nfTotal));
  VSG_CHECK(hipMalloc(&c->d_allsz, sizeof(int16_t) * c->nfTotal));
  VSG_CHECK(hipMalloc(&c->d_patchOff, sizeof(uint32_t) * c->nfTotal));
  VSG_CHECK(hipMemcpy(c->d_allx, ax.data(), sizeof(int16_t) * c->nfTotal, hipMemcpyHostToDevice));
  VSG_CHECK(hipMemcpy(c->d_ally, ay.data(), sizeof(int16_t) * c->nfTotal, hipMemcpyHostToDevice));
  VSG_CHECK(hipMemcpy(c->d_allsz, asz.data(), sizeof(int16_t) * c->nfTotal, hipMemcpyHostToDevice));
  VSG_CHECK(hipMemcpy(c->d_patchOff, poff.data(), sizeof(uint32_t) * c->nfTotal, hipMemcpyHostToDevice));

  /* pad rows: the grid kernel's K candidate rows per lane may reach
     (K-1)*step rows below a search window, plus one row of column overrun */
  c->rowsAlloc = height + 4 * maxStep + 16;
  c->slotStride = (size_t)c->pitch * c->rowsAlloc;
  VSG_CHECK(hipMalloc(&c->d_raw, (size_t)maxBatch * c->pitch * height));
  VSG_CHECK(hipMalloc(&c->d_blur, (size_t)maxBatch * c->slotStride));
  VSG_CHECK(hipMalloc(&c->d_ref, c->slotStride));
  VSG_CHECK(hipMemset(c->d_blur, 0, (size_t)maxBatch * c->slotStride));
  VSG_CHECK(hipMemset(c->d_ref, 0, c->slotStride));
  if (blurSize >= 2) {
    const int s2 = blurSize / 2;
    const size_t lds = (size_t)(32 + 2 * s2) * (((256 + 2 * s2) + 3) & ~3) + (size_t)(32 + 2 * s2) * 256;
    if (lds > 64 * 1024) VSG_CHECK(hipMalloc(&c->d_tmp, (size_t)maxBatch * c->pitch * height));
  }
  c->patchSlotStride = (patchBytes + 255) & ~(size_t)255;
  VSG_CHECK(hipMalloc(&c->d_patches, (size_t)maxBatch * c->patchSlotStride));
  VSG_CHECK(hipMalloc(&c->d_minmax, (size_t)maxBatch * c->nfTotal * 2));
  VSG_CHECK(hipHostMalloc(&c->h_minmax, (size_t)maxBatch * c->nfTotal * 2, hipHostMallocDefault));
  c->maxJobs = maxBatch * (c->st[0].nf > c->st[1].nf ? c->st[0].nf : c->st[1].nf);
  VSG_CHECK(hipMalloc(&c->d_jobs, sizeof(VSGJob) * c->maxJobs));
  VSG_CHECK(hipHostMalloc(&c->h_jobs, sizeof(VSGJob) * c->maxJobs, hipHostMallocDefault));
  VSG_CHECK(hipMalloc(&c->d_best, sizeof(unsigned long long) * c->maxJobs));
  VSG_CHECK(hipMemset(c->d_best, 0xFF, sizeof(unsigned long long) * c->maxJobs));
  VSG_CHECK(hipMalloc(&c->d_res, sizeof(VSGResult) * c->maxJobs));
  VSG_CHECK(hipHostMalloc(&c->h_res, sizeof(VSGResult) * c->maxJobs, hipHostMallocDefault));
  VSG_CHECK(hipMalloc(&c->d_src, sizeof(VSGBlurSrc) * maxBatch));
  VSG_CHECK(hipHostMalloc(&c->h_src, sizeof(VSGBlurSrc) * maxBatch, hipHostMallocDefault));
  VSG_CHECK(hipDeviceSynchronize());
  return 0;
}

extern "C" int vsg_create(VSGContext** out, int device, int width, int height, int blurSize,
                          const VSGStageDesc* coarse, const VSGStageDesc* fine, int maxBatch) {
  if (!out || !coarse || !fine || width <= 0 || height <= 0 || maxBatch <= 0)
    return fail("vsg_create", hipSuccess);
  *out = nullptr;
  VSGContext* c = new (std::nothrow) VSGContext();
  if (!c) return fail("vsg_create: out of host memory", hipSuccess);
  if (create(c, device, width, height, blurSize, coarse, fine, maxBatch) != 0) {
    char keep[sizeof(g_err)];
    memcpy(keep, g_err, sizeof(keep));
    free_ctx(c);
    memcpy(g_err, keep, sizeof(keep));
    return -1;
  }
  *out = c;
  return 0;
}

extern "C" int vsg_frames(VSGContext* c, int n, const uint8_t* const* luma, const int* linesize,
                          int onDevice, uint8_t* mmCoarse, uint8_t* mmFine) {
  if (!c || n <= 0 || n > c->maxBatch || !luma || !linesize) return fail("vsg_frames", hipSuccess);
  VSG_CHECK(hipSetDevice(c->device));
  for (int i = 0; i < n; ++i) {
    if (linesize[i] < c->W) return fail("vsg_frames: linesize < width", hipSuccess);
    if (onDevice) {
      c->h_src[i].p = luma[i];
      c->h_src[i].pitch = linesize[i];
    } else {
      uint8_t* dst = c->d_raw + (size_t)i * c->pitch * c->H;
      VSG_CHECK(hipMemcpy2DAsync(dst, c->pitch, luma[i], linesize[i], c->W, c->H,
                                 hipMemcpyHostToDevice, c->stream));
      c->h_src[i].p = dst;
      c->h_src[i].pitch = c->pitch;
    }
    c->h_src[i].pad = 0;
  }
  VSG_CHECK(hipMemcpyAsync(c->d_src, c->h_src, sizeof(VSGBlurSrc) * n, hipMemcpyHostToDevice, c->stream));
  const int prof = g_profOn.load();
  if (prof) VSG_CHECK(hipEventRecord(c->ev[0], c->stream));
  VSG_CHECK(vsg_launch_blur(c->d_src, n, c->d_blur, c->slotStride, c->pitch, c->W, c->H, c->blurSize,
                            c->d_tmp, c->stream));
  if (prof) VSG_CHECK(hipEventRecord(c->ev[1], c->stream));
  VSGFieldArgs fa;
  fa.blurred = c->d_blur;
  fa.slotStride = c->slotStride;
  fa.pitch = c->pitch;
  fa.fx = c->d_allx;
  fa.fy = c->d_ally;
  fa.fsz = c->d_allsz;
  fa.patchOff = c->d_patchOff;
  fa.nfields = c->nfTotal;
  fa.patches = c->d_patches;
  fa.patchSlotStride = c->patchSlotStride;
  fa.minmax = c->d_minmax;
  VSG_CHECK(vsg_launch_fields(fa, n, c->stream));
  if (prof) VSG_CHECK(hipEventRecord(c->ev[2], c->stream));
  const size_t mmBytes = (size_t)n * c->nfTotal * 2;
  VSG_CHECK(hipMemcpyAsync(c->h_minmax, c->d_minmax, mmBytes, hipMemcpyDeviceToHost, c->stream));
  VSG_CHECK(hipStreamSynchronize(c->stream));
  if (prof) {
    double fieldBytes = 0;
    for (int s = 0; s < 2; ++s) fieldBytes += (double)c->st[s].nf * c->st[s].S * c->st[s].S;
    prof_account(c, VSG_K_BLUR, 2.0 * c->W * c->H * n, VSG_K_FIELDS, 2.0 * fieldBytes * n);
  }
  const int nc = c->st[0].nf, nf = c->st[1].nf;
  for (int i = 0; i < n; ++i) {
    const uint8_t* src = c->h_minmax + (size_t)i * c->nfTotal * 2;
    if (mmCoarse) memcpy(mmCoarse + (size_t)i * nc * 2, src, (size_t)nc * 2);
    if (mmFine) memcpy(mmFine + (size_t)i * nf * 2, src + (size_t)nc * 2, (size_t)nf * 2);
  }
  return 0;
}

extern "C" int vsg_search(VSGContext* c, int stage, int njobs, const VSGJob* jobs, VSGResult* out) {
  if (!c || stage < 0 || stage > 1 || njobs < 0 || njobs > c->maxJobs) return fail("vsg_search", hipSuccess);
  if (njobs == 0) return 0;
  const VSGContext::Stage& S = c->st[stage];
  for (int i = 0; i < njobs; ++i) {
    const VSGJob& j = jobs[i];
    if (j.slot < 0 || j.slot >= c->maxBatch || j.prev >= c->maxBatch || j.field < 0 || j.field >= S.nf)
      return fail("vsg_search: job out of range", hipSuccess);
  }
  VSG_CHECK(hipSetDevice(c->device));
  memcpy(c->h_jobs, jobs, sizeof(VSGJob) * njobs);
  VSG_CHECK(hipMemcpyAsync(c->d_jobs, c->h_jobs, sizeof(VSGJob) * njobs, hipMemcpyHostToDevice, c->stream));
  VSGSearchArgs a;
  a.jobs = c->d_jobs;
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
  a.best = c->d_best;
  a.results = c->d_res;
  const int prof = g_profOn.load();
  if (prof) VSG_CHECK(hipEventRecord(c->ev[0], c->stream));
  VSG_CHECK(vsg_launch_grid(a, c->stream));
  if (prof) VSG_CHECK(hipEventRecord(c->ev[1], c->stream));
  VSG_CHECK(vsg_launch_refine(a, c->stream));
  if (prof) VSG_CHECK(hipEventRecord(c->ev[2], c->stream));
  VSG_CHECK(hipMemcpyAsync(c->h_res, c->d_res, sizeof(VSGResult) * njobs, hipMemcpyDeviceToHost, c->stream));
  VSG_CHECK(hipStreamSynchronize(c->stream));
  if (prof) {
    /* algorithmic work: every grid candidate over the whole field, and the
       refinement candidates the reference visits (8 + 24 for step 6 ...) */
    const double px = (double)S.S * S.S;
    const double G = 2.0 * S.g.n + 1.0;
    int refineCands = 0;
    for (int st = S.step; st > 1; st /= 2) {
      const int ns = st / 2, r = st - ns, ni = (2 * r) / ns + 1;
      refineCands += ni * ni - ((r % ns) == 0 ? 1 : 0);
    }
    prof_account(c, stage == 0 ? VSG_K_GRID_COARSE : VSG_K_GRID_FINE, njobs * G * G * px, VSG_K_REFINE,
                 njobs * (double)refineCands * px);
  }
  memcpy(out, c->h_res, sizeof(VSGResult) * njobs);
  return 0;
}

extern "C" int vsg_keep_prev(VSGContext* c, int slot) {
  if (!c || slot < 0 || slot >= c->maxBatch) return fail("vsg_keep_prev", hipSuccess);
  /* device-to-device copy on the detector's stream, ordered after its searches */
  uint8_t* p = c->d_blur + (size_t)slot * c->slotStride;
  VSG_CHECK(hipSetDevice(c->device));
  VSG_CHECK(hipMemcpyAsync(c->d_ref, p, c->slotStride, hipMemcpyDeviceToDevice, c->stream));
  return 0;
}

extern "C" int vsg_read_blurred(VSGContext* c, int slot, uint8_t* dst, int dstStride) {
  if (!c || slot >= c->maxBatch || !dst || dstStride < c->W) return fail("vsg_read_blurred", hipSuccess);
  VSG_CHECK(hipSetDevice(c->device));
  const uint8_t* src = slot < 0 ? c->d_ref : c->d_blur + (size_t)slot * c->slotStride;
  VSG_CHECK(hipMemcpy2DAsync(dst, dstStride, src, c->pitch, c->W, c->H, hipMemcpyDeviceToHost, c->stream));
  VSG_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}
