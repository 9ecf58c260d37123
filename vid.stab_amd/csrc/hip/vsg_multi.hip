/*
 * vsg_multi.hip -- the batched multi-clip form of the detect pass.
 *
 * Many independent clips (any mix of resolutions) advance together, up to B
 * frames of every clip per batch.  Each clip keeps its own VSGContext (its
 * planes, fields and selection lists); all contexts share one stream.  The
 * searches of a stage run as ONE persistent launch of k_sad_queue
 * (vsg_search.hip) over a single flattened queue of the units of every clip,
 * whatever their plan shapes -- plus, when some clips' jobs are very small
 * (360p fields), a second persistent launch over the split units of
 * those clips.
 *
 * Every batch has a fixed shape -- B frame slots per clip, idle slots marked
 * by a null source (skipped by the blur and field kernels, detect = 0 for the
 * selection, field = -1 for the searches) and fine job lists padded to their
 * capacity with field = -1 -- so each slot set's front + coarse stage and its
 * fine stage are captured once as HIP graphs and replayed for every batch
 * that uses the set.  Only pinned host descriptors change between replays.
 * Virtual-tripod clips keep their reference frame in the context's stored
 * reference plane: each clip's descriptor block carries the slot to freeze
 * this batch (-1: none), and one k_copy_refs launch between the fronts and
 * the coarse queue copies the planes of the clips that name one.
 * What crosses PCIe for all clips of a batch sits in four per-set arenas, so
 * each crossing is one copy node: the frame descriptors, the selection and
 * coarse results, the fine jobs and the fine results.  The contexts' buffers
 * for them point into the arenas.  No memset nodes are recorded: every
 * buffer a graph reads is written by the host or by an earlier node, and the
 * queue counters are left zero by the kernel itself.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "vsg_context.h"

namespace {

#define MCHECK(call)                                      \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) return vsg_fail(#call, e_);     \
  } while (0)

/* host-visible regions of all clips' per-set buffers, one copy each */
struct Arena {
  uint8_t* d = nullptr;
  uint8_t* h = nullptr;   /* pinned */
  size_t bytes = 0;
};

enum { AR_DESC, AR_COARSE, AR_JOBS, AR_RESF, AR_COUNT };

/* one clip's tripod copy: blurred slot *refSlot -> ref, when >= 0 */
struct RefCopy {
  const uint8_t* blurred;
  uint8_t* ref;
  const int* refSlot;   /* in the set's descriptor arena */
  unsigned long long slotStride;
};

constexpr int kRefBlocks = 32;   /* workgroups per clip */

/* grid (kRefBlocks, clips) x 256: 16-byte copies of whole slots (slot
   strides are multiples of 64 bytes); clips with no slot to freeze exit */
__global__ __launch_bounds__(256) void k_copy_refs(const RefCopy* __restrict__ tab) {
  const RefCopy t = tab[blockIdx.y];
  const int slot = *t.refSlot;
  if (slot < 0) return;
  const uint4* src = reinterpret_cast<const uint4*>(t.blurred + (size_t)slot * t.slotStride);
  uint4* dst = reinterpret_cast<uint4*>(t.ref);
  const size_t n = t.slotStride / 16;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)kRefBlocks * 256) dst[i] = src[i];
}

/* descriptor block of a clip: src[B], prev[B], detect[B], then the slot to
   freeze as the tripod reference at the next 4-byte boundary */
size_t desc_ref_offset(int B) { return ((sizeof(VSGBlurSrc) + sizeof(int) + 1) * (size_t)B + 3) & ~(size_t)3; }

}  // namespace

struct VSGMulti {
  int device = 0;
  int B = 0;
  hipStream_t stream = nullptr;
  std::vector<VSGContext*> ctx;
  std::vector<int> order[2];                /* clips by descending field size, per stage */
  /* per kind of unit (0: workgroup-wide tiles, 1: split 64-lane tiles) */
  VSGMultiSeg* d_segs[2][2][VSG_NSETS] = {};   /* [kind][stage][set] */
  int nseg[2][2] = {};                          /* [kind][stage] */
  VSGQueueTotals totals[2][2] = {};             /* [kind][stage]: units per XCD queue */
  std::vector<int> direct[2];   /* per stage: clips searched by their own tuned kernel, not a queue */
  unsigned* d_queue = nullptr;                  /* [kind][stage][set][xcd][2] counters */
  Arena ar[VSG_NSETS][AR_COUNT];
  std::vector<int> jobsFilled[VSG_NSETS];   /* per clip: fine entries that may hold jobs */
  std::vector<int*> h_refSlot[VSG_NSETS];   /* per clip: pinned slot to freeze as the tripod plane */
  RefCopy* d_refTab[VSG_NSETS] = {};
  hipGraphExec_t gA[VSG_NSETS] = {};
  hipGraphExec_t gF[VSG_NSETS] = {};
  hipEvent_t doneA[VSG_NSETS] = {};
  hipEvent_t doneF[VSG_NSETS] = {};
};

static void multi_free(VSGMulti* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (int s = 0; s < VSG_NSETS; ++s) {
    if (m->gA[s]) (void)hipGraphExecDestroy(m->gA[s]);
    if (m->gF[s]) (void)hipGraphExecDestroy(m->gF[s]);
    if (m->doneA[s]) (void)hipEventDestroy(m->doneA[s]);
    if (m->doneF[s]) (void)hipEventDestroy(m->doneF[s]);
    for (int kind = 0; kind < 2; ++kind)
      for (int k = 0; k < 2; ++k)
        if (m->d_segs[kind][k][s]) (void)hipFree(m->d_segs[kind][k][s]);
    if (m->d_refTab[s]) (void)hipFree(m->d_refTab[s]);
    for (Arena& a : m->ar[s]) {
      if (a.d) (void)hipFree(a.d);
      if (a.h) (void)hipHostFree(a.h);
    }
  }
  if (m->d_queue) (void)hipFree(m->d_queue);
  for (VSGContext* c : m->ctx) {
    if (!c) continue;
    vsg_ctx_free(c);   /* skips the buffers that point into the arenas */
    delete c;
  }
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

static size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

/* per clip and set: the bytes each arena holds for it */
static void clip_arena_bytes(const VSGMulti* m, const VSGContext* x, size_t (&b)[AR_COUNT]) {
  const size_t B = (size_t)m->B, nft = (size_t)x->nfTotal, nc = (size_t)x->st[0].nf, nf = (size_t)x->st[1].nf;
  b[AR_DESC] = up256(desc_ref_offset(m->B) + sizeof(int));
  b[AR_COARSE] = up256(sizeof(double) * B * nft + sizeof(VSGResult) * B * nc + sizeof(int) * 2 * B +
                       sizeof(int16_t) * B * nft);
  b[AR_JOBS] = up256(sizeof(VSGJob) * B * nf);
  b[AR_RESF] = up256(sizeof(VSGResult) * B * nf);
}

/* frees a context's own transfer buffers of every set and points them into
   the arenas (the context's tuning, at creation, used its own) */
static int bind_arenas(VSGMulti* m) {
  const int B = m->B;
  size_t total[AR_COUNT] = {0, 0, 0, 0};
  std::vector<std::vector<size_t>> off(m->ctx.size(), std::vector<size_t>(AR_COUNT));
  for (size_t c = 0; c < m->ctx.size(); ++c) {
    size_t b[AR_COUNT];
    clip_arena_bytes(m, m->ctx[c], b);
    for (int k = 0; k < AR_COUNT; ++k) {
      off[c][k] = total[k];
      total[k] += b[k];
    }
  }
  for (int set = 0; set < VSG_NSETS; ++set)
    for (int k = 0; k < AR_COUNT; ++k) {
      Arena& a = m->ar[set][k];
      a.bytes = total[k];
      MCHECK(hipMalloc(&a.d, a.bytes));
      MCHECK(hipHostMalloc(reinterpret_cast<void**>(&a.h), a.bytes, hipHostMallocDefault));
      memset(a.h, 0, a.bytes);
      MCHECK(hipMemset(a.d, 0, a.bytes));
    }
  std::vector<RefCopy> tab[VSG_NSETS];
  for (size_t c = 0; c < m->ctx.size(); ++c) {
    VSGContext* x = m->ctx[c];
    const size_t nft = (size_t)x->nfTotal, nc = (size_t)x->st[0].nf;
    for (int set = 0; set < VSG_NSETS; ++set) {
      VSGSet& s = x->set[set];
      vsg_ctx_release_transfer(x, set);
      s.borrowed = true;
      /* descriptors: src[B], prev[B], detect[B] */
      uint8_t* dd = m->ar[set][AR_DESC].d + off[c][AR_DESC];
      uint8_t* hd = m->ar[set][AR_DESC].h + off[c][AR_DESC];
      s.d_desc = dd;
      s.h_desc = hd;
      s.d_src = reinterpret_cast<VSGBlurSrc*>(dd);
      s.h_src = reinterpret_cast<VSGBlurSrc*>(hd);
      s.d_prev = reinterpret_cast<int*>(dd + sizeof(VSGBlurSrc) * B);
      s.h_prev = reinterpret_cast<int*>(hd + sizeof(VSGBlurSrc) * B);
      s.d_detect = dd + (sizeof(VSGBlurSrc) + sizeof(int)) * B;
      s.h_detect = hd + (sizeof(VSGBlurSrc) + sizeof(int)) * B;
      int* hr = reinterpret_cast<int*>(hd + desc_ref_offset(B));
      *hr = -1;
      m->h_refSlot[set].push_back(hr);
      tab[set].push_back(RefCopy{x->d_blur, x->d_ref, reinterpret_cast<const int*>(dd + desc_ref_offset(B)),
                                 (unsigned long long)x->slotStride});
      /* coarse out: selC, resC, count, selIdx (alignment order) */
      uint8_t* dc = m->ar[set][AR_COARSE].d + off[c][AR_COARSE];
      uint8_t* hc = m->ar[set][AR_COARSE].h + off[c][AR_COARSE];
      size_t o = 0;
      s.d_selC = reinterpret_cast<double*>(dc + o);
      s.h_selC = reinterpret_cast<double*>(hc + o);
      o += sizeof(double) * B * nft;
      s.d_resC = reinterpret_cast<VSGResult*>(dc + o);
      s.h_resC = reinterpret_cast<VSGResult*>(hc + o);
      o += sizeof(VSGResult) * B * nc;
      s.d_count = reinterpret_cast<int*>(dc + o);
      s.h_count = reinterpret_cast<int*>(hc + o);
      o += sizeof(int) * 2 * B;
      s.d_selIdx = reinterpret_cast<int16_t*>(dc + o);
      s.h_selIdx = reinterpret_cast<int16_t*>(hc + o);
      s.d_jobsF = reinterpret_cast<VSGJob*>(m->ar[set][AR_JOBS].d + off[c][AR_JOBS]);
      s.h_jobsF = reinterpret_cast<VSGJob*>(m->ar[set][AR_JOBS].h + off[c][AR_JOBS]);
      s.d_resF = reinterpret_cast<VSGResult*>(m->ar[set][AR_RESF].d + off[c][AR_RESF]);
      s.h_resF = reinterpret_cast<VSGResult*>(m->ar[set][AR_RESF].h + off[c][AR_RESF]);
    }
  }
  for (int set = 0; set < VSG_NSETS; ++set) {
    MCHECK(hipMalloc(&m->d_refTab[set], sizeof(RefCopy) * tab[set].size()));
    MCHECK(hipMemcpy(m->d_refTab[set], tab[set].data(), sizeof(RefCopy) * tab[set].size(), hipMemcpyHostToDevice));
  }
  return 0;
}

/* the queue launches' shape: workgroup lanes and candidate rows per lane of
   each stage, and the queue count (1, or one per XCD); VIDSTAB_AMD_QUEUE_WG,
   VIDSTAB_AMD_QUEUE_K (coarse,fine) and VIDSTAB_AMD_QUEUES override */
static int env_pair(const char* name, int stage, int def) {
  const char* v = getenv(name);
  if (!v || !*v) return def;
  int a = def, b = def;
  if (sscanf(v, "%d,%d", &a, &b) < 2) b = a;
  return stage == 0 ? a : b;
}
/* defaults measured on config 4's 64 clips (frames/s of the whole call):
   one queue 6,013, one per XCD 4,961 (the eighths finish unevenly with no
   stealing between them); 256 / 128 lanes 6,013, 256 / 64 5,651, 256 / 256
   5,053; K 6 / 6 6,013, 4 / 6 5,651, 4 / 8 5,255 */
static int queue_wg(int stage) { return env_pair("VIDSTAB_AMD_QUEUE_WG", stage, stage == 0 ? 256 : 128); }
static int queue_k(int stage) { return env_pair("VIDSTAB_AMD_QUEUE_K", stage, 6); }
/* a clip whose stage launch holds at least this many lanes (B frames' jobs
   x workgroups per job x workgroup size, under its own tuned plan) keeps
   the chip busy on its own: it is searched by its own k_sad_search launch
   (the autotuned plan: the field staged in LDS where that is faster, the
   units of a frame on one XCD) inside the multi-clip call, and only the
   smaller clips share the queues.  Measured on config 4's 64 clips (frames/s
   of the whole call, two runs each): every clip queued 5,973-5,978; direct
   from 262,144 lanes (4 waves per SIMD) 6,325-6,452, from 131,072
   6,469-6,510, from 65,536 (one wave per SIMD) 6,505-6,642, from 32,768
   6,425-6,513 with 64 small clips down from 81,000-94,000 to 62,000.
   VIDSTAB_AMD_QUEUE_DIRECT_LANES overrides (0: every clip queued). */
static long long queue_direct_lanes(void) {
  const char* v = getenv("VIDSTAB_AMD_QUEUE_DIRECT_LANES");
  if (!v || !*v) return 1LL << 16;
  const long long n = atoll(v);
  return n > 0 ? n : (1LL << 62);
}
/* VIDSTAB_AMD_QUEUE_SPLIT=0 keeps workgroup-wide units for every job */
static bool queue_split_on(void) {
  const char* v = getenv("VIDSTAB_AMD_QUEUE_SPLIT");
  return !(v && atoi(v) == 0);
}
/* jobs of at most this many lanes (at the wide queue's K) go to the split
   queue (VIDSTAB_AMD_QUEUE_SPLIT_LANES overrides).  Measured on config 4's
   64 clips (frames/s of the whole call) and on 64 clips of 640x360:
   32 lanes (only the 360p clips' jobs, 18 lanes) 6,038 / 96,714; 64 (also
   720p coarse and the 1080p-1440p fine jobs) 5,799 / 96,194; 128 (also
   1080p coarse) 5,298 / 92,557: in the wide queue a small job's idle waves
   leave their SIMDs to other workgroups' waves, while the split queue runs
   after it and adds its whole length */
static int queue_split_lanes(int wg) {
  (void)wg;
  const char* v = getenv("VIDSTAB_AMD_QUEUE_SPLIT_LANES");
  return v && atoi(v) > 0 ? atoi(v) : 32;
}
static int queue_count(void) {
  const char* v = getenv("VIDSTAB_AMD_QUEUES");
  return v && atoi(v) == VSG_XCDS ? VSG_XCDS : 1;
}

/* the queues' segment tables of one stage for one set: every clip with all
   B frame slots (fixed shape), larger fields first; XCD queue x takes the
   x-th eighth of each clip's job list.  A clip whose jobs fill at most
   queue_split_lanes lanes (360p clips) goes to the split queue instead, in
   64-lane tiles whose rows the workgroup's waves share (K = 4 there) */
static int build_segments(VSGMulti* m, int stage, int set) {
  std::vector<VSGMultiSeg> segs[2];
  const int queues = queue_count();
  const int wg = queue_wg(stage);
  for (int kind = 0; kind < 2; ++kind) {
    VSGQueueTotals& tot = m->totals[kind][stage];
    tot = VSGQueueTotals{};
    tot.queues = queues;
    tot.chunk = 1;
  }
  /* kind 0: wide queue, 1: split queue, 2: its own launch */
  std::vector<int> kindOf(m->ctx.size());
  m->direct[stage].clear();
  for (int c : m->order[stage]) {
    const VSGStage& st = m->ctx[c]->st[stage];
    const long long lanes = (long long)m->B * st.nf * st.tiles * st.g.wg;
    if (lanes >= queue_direct_lanes()) {
      kindOf[c] = 2;
      m->direct[stage].push_back(c);
      continue;
    }
    VSGSearchArgs a = vsg_queue_args(vsg_ctx_search_args(m->ctx[c], stage, nullptr, 0, nullptr, nullptr, nullptr),
                                     wg, queue_k(stage));
    kindOf[c] = queue_split_on() && a.g.U * a.g.V <= queue_split_lanes(wg) ? 1 : 0;
  }
  for (int kind = 0; kind < 2; ++kind) {
    int n = 0;
    for (int c : m->order[stage]) n += kindOf[c] == kind;
    m->nseg[kind][stage] = n;
    segs[kind].resize((size_t)queues * n);
  }
  for (int x = 0; x < queues; ++x) {
    int units[2] = {0, 0}, k[2] = {0, 0};
    for (int c : m->order[stage]) {
      const int kind = kindOf[c];
      if (kind == 2) continue;
      VSGContext* x_ = m->ctx[c];
      VSGSet& s = x_->set[set];
      const int njobs = m->B * x_->st[stage].nf;
      VSGMultiSeg sg;
      sg.a = vsg_queue_args(vsg_ctx_search_args(x_, stage, stage == 0 ? s.d_jobsC : s.d_jobsF, njobs, s.d_best,
                                                s.d_done, stage == 0 ? s.d_resC : s.d_resF),
                            wg, kind ? 4 : queue_k(stage));
      if (kind) sg.a.tiles = (sg.a.g.U * sg.a.g.V + 63) / 64;
      sg.variant = vsg_queue_variant(sg.a.g);
      if (sg.variant < 0) return vsg_fail("vsg_multi: search plan without a queue variant", hipSuccess);
      const int j0 = (int)((long long)njobs * x / queues), j1 = (int)((long long)njobs * (x + 1) / queues);
      sg.jobBase = j0;
      sg.unitBase = units[kind];
      sg.pad = 0;
      units[kind] += (j1 - j0) * sg.a.tiles;
      segs[kind][(size_t)x * m->nseg[kind][stage] + k[kind]++] = sg;
    }
    for (int kind = 0; kind < 2; ++kind) m->totals[kind][stage].units[x] = units[kind];
  }
  for (int kind = 0; kind < 2; ++kind) {
    if (segs[kind].empty()) continue;
    VSGMultiSeg*& d = m->d_segs[kind][stage][set];
    if (!d) MCHECK(hipMalloc(&d, sizeof(VSGMultiSeg) * segs[kind].size()));
    MCHECK(hipMemcpy(d, segs[kind].data(), sizeof(VSGMultiSeg) * segs[kind].size(), hipMemcpyHostToDevice));
  }
  return 0;
}

/* the stage's launches of the clips that fill the chip alone, then its
   wide-unit queue, then its split-unit queue (the two queues run as two
   branches of the graph, side by side, measured the same: config 4 5,894
   against 5,843 frames/s, 64 small clips 95,251 against 94,705) */
static int launch_queue(VSGMulti* m, int stage, int set) {
  for (int c : m->direct[stage]) {
    VSGContext* x = m->ctx[c];
    VSGSet& s = x->set[set];
    const VSGSearchArgs a = vsg_ctx_search_args(x, stage, stage == 0 ? s.d_jobsC : s.d_jobsF, m->B * x->st[stage].nf,
                                                s.d_best, s.d_done, stage == 0 ? s.d_resC : s.d_resF);
    MCHECK(vsg_launch_search(a, m->stream));
  }
  for (int kind = 0; kind < 2; ++kind) {
    if (m->nseg[kind][stage] == 0) continue;
    unsigned* q = m->d_queue + 2 * VSG_XCDS * ((kind * 2 + stage) * VSG_NSETS + set);
    MCHECK(vsg_launch_queue(m->d_segs[kind][stage][set], m->nseg[kind][stage], m->totals[kind][stage],
                            queue_wg(stage), kind ? 4 : queue_k(stage), kind == 1, q, m->stream));
  }
  return 0;
}

/* descriptors upload + front of every clip + tripod planes + coarse queue +
   one read-back */
static int capture_front(VSGMulti* m, int set) {
  if (build_segments(m, 0, set) != 0) return -1;
  hipGraph_t g = nullptr;
  MCHECK(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  const Arena& ad = m->ar[set][AR_DESC];
  if (hipMemcpyAsync(ad.d, ad.h, ad.bytes, hipMemcpyHostToDevice, m->stream) != hipSuccess)
    rc = vsg_fail("descriptor upload capture", hipErrorUnknown);
  for (VSGContext* x : m->ctx) {
    if (rc != 0) break;
    x->set[set].prof = false;
    rc = vsg_ctx_enqueue_front(x, set, m->B, false);
  }
  if (rc == 0) {
    /* after every clip's blur, before any search reads a tripod plane */
    hipLaunchKernelGGL(k_copy_refs, dim3(kRefBlocks, (unsigned)m->ctx.size()), dim3(256), 0, m->stream,
                       m->d_refTab[set]);
    if (hipGetLastError() != hipSuccess) rc = vsg_fail("tripod copy capture", hipErrorUnknown);
  }
  if (rc == 0) rc = launch_queue(m, 0, set);
  const Arena& ac = m->ar[set][AR_COARSE];
  if (rc == 0 && hipMemcpyAsync(ac.h, ac.d, ac.bytes, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
    rc = vsg_fail("coarse read-back capture", hipErrorUnknown);
  hipError_t e = hipStreamEndCapture(m->stream, &g);
  if (rc != 0) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  MCHECK(e);
  e = hipGraphInstantiate(&m->gA[set], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  MCHECK(e);
  return 0;
}

/* fine jobs upload + fine queue + results read-back, fixed shape */
static int capture_fine(VSGMulti* m, int set) {
  if (build_segments(m, 1, set) != 0) return -1;
  hipGraph_t g = nullptr;
  MCHECK(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  const Arena& aj = m->ar[set][AR_JOBS];
  const Arena& ar = m->ar[set][AR_RESF];
  if (hipMemcpyAsync(aj.d, aj.h, aj.bytes, hipMemcpyHostToDevice, m->stream) != hipSuccess)
    rc = vsg_fail("fine job upload capture", hipErrorUnknown);
  if (rc == 0) rc = launch_queue(m, 1, set);
  if (rc == 0 && hipMemcpyAsync(ar.h, ar.d, ar.bytes, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
    rc = vsg_fail("fine read-back capture", hipErrorUnknown);
  hipError_t e = hipStreamEndCapture(m->stream, &g);
  if (rc != 0) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  MCHECK(e);
  e = hipGraphInstantiate(&m->gF[set], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  MCHECK(e);
  return 0;
}

extern "C" int vsg_multi_create(VSGMulti** out, int device, int nclips, const int* width, const int* height,
                                const int* blurSize, const VSGStageDesc* coarse, const VSGStageDesc* fine,
                                int maxBatch) {
  if (!out || nclips <= 0 || !width || !height || !blurSize || !coarse || !fine || maxBatch <= 0)
    return vsg_fail("vsg_multi_create", hipSuccess);
  *out = nullptr;
  VSGMulti* m = new (std::nothrow) VSGMulti();
  if (!m) return vsg_fail("vsg_multi_create: out of host memory", hipSuccess);
  m->device = device;
  m->B = maxBatch;
  int rc = 0;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
    multi_free(m);
    return vsg_fail("vsg_multi_create: stream", hipSuccess);
  }
  const hipStream_t shared[VSG_NSETS] = {m->stream, m->stream, m->stream};
  for (int c = 0; c < nclips && rc == 0; ++c) {
    VSGContext* x = new (std::nothrow) VSGContext();
    if (!x) {
      rc = vsg_fail("vsg_multi_create: out of host memory", hipSuccess);
      break;
    }
    m->ctx.push_back(x);
    rc = vsg_ctx_init(x, device, width[c], height[c], 1, blurSize[c], &coarse[c], &fine[c], maxBatch, shared);
  }
  /* queue order per stage: larger fields (longer units) first */
  for (int k = 0; k < 2 && rc == 0; ++k) {
    for (int c = 0; c < nclips; ++c) m->order[k].push_back(c);
    std::stable_sort(m->order[k].begin(), m->order[k].end(),
                     [&](int a, int b) { return m->ctx[a]->st[k].S > m->ctx[b]->st[k].S; });
  }
  if (rc == 0) rc = bind_arenas(m);
  const size_t qbytes = sizeof(unsigned) * 2 * VSG_XCDS * 2 * 2 * VSG_NSETS;
  if (rc == 0 && (hipMalloc(&m->d_queue, qbytes) != hipSuccess || hipMemset(m->d_queue, 0, qbytes) != hipSuccess ||
                  hipDeviceSynchronize() != hipSuccess))
    rc = vsg_fail("vsg_multi_create: queue counters", hipSuccess);
  for (int s = 0; s < VSG_NSETS && rc == 0; ++s) {
    if (hipEventCreateWithFlags(&m->doneA[s], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&m->doneF[s], hipEventDisableTiming) != hipSuccess)
      rc = vsg_fail("vsg_multi_create: events", hipSuccess);
    m->jobsFilled[s].assign(nclips, -1);
  }
  if (rc != 0) {
    char keep[512];
    strncpy(keep, vsg_last_error(), sizeof(keep) - 1);
    keep[sizeof(keep) - 1] = 0;
    multi_free(m);
    vsg_fail(keep, hipSuccess);
    return -1;
  }
  *out = m;
  return 0;
}

extern "C" void vsg_multi_destroy(VSGMulti* m) { multi_free(m); }

extern "C" VSGContext* vsg_multi_clip(VSGMulti* m, int clip) {
  if (!m || clip < 0 || clip >= (int)m->ctx.size()) return nullptr;
  return m->ctx[clip];
}

extern "C" int vsg_multi_groups(VSGMulti* m, int stage) {
  if (!m || stage < 0 || stage > 1) return -1;
  return 1;   /* one queue launch per stage for all clips */
}

extern "C" int vsg_multi_submit(VSGMulti* m, int set, const int* n, const uint8_t* const* const* lumaDev,
                                const int* linesize, const int* const* prevSlot, const uint8_t* const* detect,
                                const int* refSlot) {
  if (!m || set < 0 || set >= VSG_NSETS || !n || !lumaDev || !linesize || !prevSlot || !detect)
    return vsg_fail("vsg_multi_submit", hipSuccess);
  MCHECK(hipSetDevice(m->device));
  /* the set's pinned buffers are rewritten below: its previous batch must be
     complete (front, coarse read-back and fine stage) */
  MCHECK(hipEventSynchronize(m->doneA[set]));
  MCHECK(hipEventSynchronize(m->doneF[set]));
  const int nc = (int)m->ctx.size();
  for (int c = 0; c < nc; ++c) {
    VSGContext* x = m->ctx[c];
    VSGSet& s = x->set[set];
    if (n[c] < 0 || n[c] > m->B || (n[c] > 0 && linesize[c] < x->W))
      return vsg_fail("vsg_multi_submit: frame count or linesize", hipSuccess);
    const int ref = refSlot ? refSlot[c] : -1;
    if (ref >= 0 && (ref < set * m->B || ref >= set * m->B + n[c]))
      return vsg_fail("vsg_multi_submit: refSlot outside the clip's batch", hipSuccess);
    *m->h_refSlot[set][c] = ref < 0 ? -1 : ref;
    for (int i = 0; i < m->B; ++i) {
      if (i < n[c]) {
        if (prevSlot[c][i] < -1 || prevSlot[c][i] >= VSG_NSETS * m->B)
          return vsg_fail("vsg_multi_submit: prevSlot", hipSuccess);
        s.h_src[i].p = lumaDev[c][i];
        s.h_src[i].pitch = linesize[c];
        s.h_prev[i] = prevSlot[c][i];
        s.h_detect[i] = detect[c][i] ? 1 : 0;
      } else {
        s.h_src[i].p = nullptr;   /* idle slot */
        s.h_src[i].pitch = 0;
        s.h_prev[i] = set * m->B;
        s.h_detect[i] = 0;
      }
      s.h_src[i].pad = 0;
    }
    s.n = n[c];
  }
  if (!m->gA[set] && capture_front(m, set) != 0) return -1;
  MCHECK(hipGraphLaunch(m->gA[set], m->stream));
  MCHECK(hipEventRecord(m->doneA[set], m->stream));
  return 0;
}

extern "C" int vsg_multi_wait_coarse(VSGMulti* m, int set) {
  if (!m || set < 0 || set >= VSG_NSETS) return vsg_fail("vsg_multi_wait_coarse", hipSuccess);
  MCHECK(hipSetDevice(m->device));
  MCHECK(hipEventSynchronize(m->doneA[set]));
  return 0;
}

extern "C" int vsg_multi_submit_fine(VSGMulti* m, int set, const int* njobs, const VSGJob* const* jobs) {
  if (!m || set < 0 || set >= VSG_NSETS || !njobs || !jobs) return vsg_fail("vsg_multi_submit_fine", hipSuccess);
  MCHECK(hipSetDevice(m->device));
  const int nc = (int)m->ctx.size();
  for (int c = 0; c < nc; ++c) {
    VSGContext* x = m->ctx[c];
    VSGSet& s = x->set[set];
    const int cap = m->B * x->st[1].nf;
    if (njobs[c] < 0 || njobs[c] > cap) return vsg_fail("vsg_multi_submit_fine: too many jobs", hipSuccess);
    for (int i = 0; i < njobs[c]; ++i) {
      const VSGJob& j = jobs[c][i];
      if (j.slot < 0 || j.slot >= VSG_NSETS * m->B || j.prev < -1 || j.prev >= VSG_NSETS * m->B ||
          j.field >= x->st[1].nf)
        return vsg_fail("vsg_multi_submit_fine: job out of range", hipSuccess);
    }
    if (njobs[c] > 0 && jobs[c] != s.h_jobsF) memcpy(s.h_jobsF, jobs[c], sizeof(VSGJob) * njobs[c]);
    /* pad to capacity with skipped jobs: only entries that may still hold
       jobs from an earlier batch need rewriting */
    int& filled = m->jobsFilled[set][c];
    const int from = njobs[c], to = filled < 0 ? cap : filled;
    for (int i = from; i < to; ++i) s.h_jobsF[i].field = -1;
    filled = njobs[c];
    s.njobsF = njobs[c];
  }
  if (!m->gF[set] && capture_fine(m, set) != 0) return -1;
  MCHECK(hipGraphLaunch(m->gF[set], m->stream));
  MCHECK(hipEventRecord(m->doneF[set], m->stream));
  return 0;
}

extern "C" int vsg_multi_wait_fine(VSGMulti* m, int set) {
  if (!m || set < 0 || set >= VSG_NSETS) return vsg_fail("vsg_multi_wait_fine", hipSuccess);
  MCHECK(hipSetDevice(m->device));
  MCHECK(hipEventSynchronize(m->doneF[set]));
  return 0;
}
