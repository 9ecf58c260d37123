/*
 * vsg_multi.hip -- the batched multi-clip form of the detect pass.
 *
 * Many independent clips (any mix of resolutions) advance together, up to B
 * frames of every clip per batch.  Each clip keeps its own VSGContext (its
 * planes, fields and selection lists); all contexts share one stream.  What
 * changes is the search: the coarse (and later the fine) jobs of all clips
 * whose search plans agree (same field size, step and grid, so the same
 * workgroup shape) form one flattened unit list and run as ONE launch of
 * k_sad_multi.  The plans come from the autotuner of each context, so clips
 * of one geometry always land in the same launch.
 *
 * Every batch has a fixed shape -- B frame slots per clip, idle slots marked
 * by a null source (skipped by the blur and field kernels, detect = 0 for the
 * selection, field = -1 for the searches) and fine job lists padded to their
 * capacity with field = -1 -- so each slot set's front + coarse stage and its
 * fine stage are captured once as HIP graphs and replayed for every batch
 * that uses the set.  Only pinned host descriptors change between replays.
 * No memset nodes are recorded: every buffer a graph reads is written by the
 * host or by an earlier node.
 */
#include <stdlib.h>
#include <string.h>

#include <map>
#include <new>
#include <tuple>
#include <vector>

#include "vsg_context.h"

namespace {

#define MCHECK(call)                                      \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) return vsg_fail(#call, e_);     \
  } while (0)

struct Group {
  VSGGrid g{};
  int tiles = 0;
  std::vector<int> clips;
  VSGMultiSeg* d_segs[VSG_NSETS] = {};
  int units[VSG_NSETS] = {};
};

}  // namespace

struct VSGMulti {
  int device = 0;
  int B = 0;
  hipStream_t stream = nullptr;
  std::vector<VSGContext*> ctx;
  std::vector<Group> grp[2];            /* per stage */
  std::vector<int> jobsFilled[VSG_NSETS];   /* per clip: fine entries that may hold jobs */
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
  }
  for (auto& gs : m->grp)
    for (Group& g : gs)
      for (int s = 0; s < VSG_NSETS; ++s)
        if (g.d_segs[s]) (void)hipFree(g.d_segs[s]);
  for (VSGContext* c : m->ctx) {
    if (!c) continue;
    vsg_ctx_free(c);
    delete c;
  }
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

/* segment tables of one stage for one set: every member clip with all B
   frame slots (fixed shape) */
static int build_segments(VSGMulti* m, int stage, int set) {
  for (Group& gr : m->grp[stage]) {
    std::vector<VSGMultiSeg> segs;
    int units = 0;
    for (int c : gr.clips) {
      VSGContext* x = m->ctx[c];
      VSGSet& s = x->set[set];
      const int njobs = m->B * x->st[stage].nf;
      VSGMultiSeg sg;
      sg.a = vsg_ctx_search_args(x, stage, stage == 0 ? s.d_jobsC : s.d_jobsF, njobs, s.d_best, s.d_done,
                                 stage == 0 ? s.d_resC : s.d_resF);
      sg.unitBase = units;
      sg.pad = 0;
      units += njobs * sg.a.tiles;
      segs.push_back(sg);
    }
    if (!gr.d_segs[set]) MCHECK(hipMalloc(&gr.d_segs[set], sizeof(VSGMultiSeg) * segs.size()));
    MCHECK(hipMemcpy(gr.d_segs[set], segs.data(), sizeof(VSGMultiSeg) * segs.size(), hipMemcpyHostToDevice));
    gr.units[set] = units;
  }
  return 0;
}

static int launch_groups(VSGMulti* m, int stage, int set) {
  for (Group& gr : m->grp[stage])
    MCHECK(vsg_launch_multi(gr.d_segs[set], (int)gr.clips.size(), gr.units[set], gr.g, m->stream));
  return 0;
}

/* front + coarse search + coarse read-back of every clip, fixed shape */
static int capture_front(VSGMulti* m, int set) {
  if (build_segments(m, 0, set) != 0) return -1;
  hipGraph_t g = nullptr;
  MCHECK(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (VSGContext* x : m->ctx) {
    x->set[set].prof = false;
    if ((rc = vsg_ctx_enqueue_front(x, set, m->B)) != 0) break;
  }
  if (rc == 0) rc = launch_groups(m, 0, set);
  if (rc == 0)
    for (VSGContext* x : m->ctx)
      if ((rc = vsg_ctx_enqueue_coarse_readback(x, set, m->B)) != 0) break;
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

/* fine jobs upload + fine search + results read-back, fixed shape */
static int capture_fine(VSGMulti* m, int set) {
  if (build_segments(m, 1, set) != 0) return -1;
  hipGraph_t g = nullptr;
  MCHECK(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (VSGContext* x : m->ctx) {
    VSGSet& s = x->set[set];
    const size_t cap = (size_t)m->B * x->st[1].nf;
    if (hipMemcpyAsync(s.d_jobsF, s.h_jobsF, sizeof(VSGJob) * cap, hipMemcpyHostToDevice, m->stream) !=
        hipSuccess) {
      rc = vsg_fail("fine job upload capture", hipErrorUnknown);
      break;
    }
  }
  if (rc == 0) rc = launch_groups(m, 1, set);
  if (rc == 0)
    for (VSGContext* x : m->ctx) {
      VSGSet& s = x->set[set];
      const size_t cap = (size_t)m->B * x->st[1].nf;
      if (hipMemcpyAsync(s.h_resF, s.d_resF, sizeof(VSGResult) * cap, hipMemcpyDeviceToHost, m->stream) !=
          hipSuccess) {
        rc = vsg_fail("fine read-back capture", hipErrorUnknown);
        break;
      }
    }
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

static bool same_plan(const VSGStage& a, const VSGStage& b) {
  return a.S == b.S && a.step == b.step && a.g.n == b.g.n && a.g.J == b.g.J && a.g.K == b.g.K &&
         a.g.D == b.g.D && a.g.CW == b.g.CW && a.g.wg == b.g.wg && a.tiles == b.tiles;
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
  /* group clips by plan, per stage */
  for (int k = 0; k < 2 && rc == 0; ++k)
    for (int c = 0; c < nclips; ++c) {
      const VSGStage& S = m->ctx[c]->st[k];
      Group* hit = nullptr;
      for (Group& gr : m->grp[k])
        if (same_plan(m->ctx[gr.clips[0]]->st[k], S)) hit = &gr;
      if (!hit) {
        m->grp[k].push_back(Group());
        hit = &m->grp[k].back();
        hit->g = S.g;
        hit->tiles = S.tiles;
      }
      hit->clips.push_back(c);
    }
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
  return (int)m->grp[stage].size();
}

extern "C" int vsg_multi_submit(VSGMulti* m, int set, const int* n, const uint8_t* const* const* lumaDev,
                                const int* linesize, const int* const* prevSlot, const uint8_t* const* detect) {
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
    for (int i = 0; i < m->B; ++i) {
      if (i < n[c]) {
        if (prevSlot[c][i] < 0 || prevSlot[c][i] >= VSG_NSETS * m->B)
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
      if (j.slot < 0 || j.slot >= VSG_NSETS * m->B || j.prev < 0 || j.prev >= VSG_NSETS * m->B ||
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
