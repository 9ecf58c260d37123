/*
 * vsg_tune.hip -- the search-plan autotuner.
 *
 * A plan fixes how a job's (2n+1)^2 candidate grid is spread over lanes:
 * K candidate rows per lane (4, 6 or 8; J columns follow from the step: 4
 * where the candidates of a lane sit whole words apart), or, for grids whose
 * neighbouring columns are one word apart, a tile of 3 columns by 7 rows
 * (K = 7), which wastes fewer lane slots on the fine stage's 33 x 33 and
 * 49 x 49 grids than 4 x 6; with it the workgroup size and workgroups per
 * job; the column chunk CW
 * follows from the field size.  For fields up to 128 px a plan may also
 * stage the field in LDS (read by broadcast into a VGPR operand) instead of
 * reading it with scalar loads (SGPR operand), which trades scalar-cache
 * misses for LDS reads.  Every plan computes the same
 * argmin, so the choice only moves time.  For each stage of a new context the
 * tuner times every plan on the device over one full batch of that stage's
 * jobs (all fields of maxBatch frames, zero offsets, the context's own
 * buffers) and keeps the fastest; results are cached per device and stage
 * geometry for the life of the process.
 *
 *   VIDSTAB_AMD_SEARCH_K=4|6|7|8 force a plan (tests, sweeps; 7 only where
 *                                the three-column tile applies, else 4)
 *   VIDSTAB_AMD_AUTOTUNE=0       no timing: the cost model's plan
 */
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "vsg_context.h"

namespace {

constexpr int kPlans[3] = {4, 6, 8};

typedef std::tuple<int, int, int, int, int, int, int> TuneKey;   /* device, S, step, maxShift, B, nf, bpp */
std::mutex g_tuneMu;
std::map<TuneKey, int> g_tuned;   /* K + 16 * lds */
constexpr int kMaxLdsPatch = 128 * 128 * 3;   /* bytes of a field staged in LDS */

int env_int(const char* name, int def) {
  const char* v = getenv(name);
  if (!v || !*v) return def;
  return atoi(v);
}

#define TUNE_CHECK(call)                                  \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) return vsg_fail(#call, e_);     \
  } while (0)

/* times plan (K, lds) on stage k (ms, best of 2 after a warm-up) */
int time_plan(VSGContext* c, int k, int K, int lds, const VSGJob* dJobs, int njobs, float* ms) {
  VSGStage& S = c->st[k];
  VSGSet& s = c->set[0];
  VSGStage trial = S;
  trial.tiles = vsg_plan_grid(&trial.g, S.S, S.step, S.maxShift, K, c->bpp);
  trial.g.lds = lds;
  VSGSearchArgs a = vsg_ctx_search_args(c, k, dJobs, njobs, s.d_best, s.d_done, k == 0 ? s.d_resC : s.d_resF);
  a.g = trial.g;
  a.tiles = trial.tiles;
  a.chunk = S.nf * trial.tiles;
  hipEvent_t e0, e1;
  TUNE_CHECK(hipEventCreate(&e0));
  TUNE_CHECK(hipEventCreate(&e1));
  float best = 0;
  hipError_t err = hipSuccess;
  for (int rep = 0; rep < 3 && err == hipSuccess; ++rep) {
    err = hipEventRecord(e0, s.stream);
    if (err == hipSuccess) err = vsg_launch_search(a, s.stream);
    if (err == hipSuccess) err = hipEventRecord(e1, s.stream);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float t = 0;
    if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
    if (rep > 0 && (best == 0 || t < best)) best = t;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (err != hipSuccess) return vsg_fail("search plan timing", err);
  *ms = best;
  return 0;
}

}  // namespace

int vsg_plan_model(VSGGrid* g, int S, int step, int maxShift, int bpp) {
  int bestTiles = 0;
  double bestCost = 0;
  for (int K : kPlans) {   /* the four-column plans */
    VSGGrid t;
    const int tiles = vsg_plan_grid(&t, S, step, maxShift, K, bpp);
    const double cost = vsg_plan_cost(t, tiles);
    if (bestTiles == 0 || cost < bestCost) {
      bestCost = cost;
      bestTiles = tiles;
      *g = t;
    }
  }
  return bestTiles;
}

int vsg_ctx_tune(VSGContext* c) {
  const int forced = env_int("VIDSTAB_AMD_SEARCH_K", 0);
  const int forcedLds = env_int("VIDSTAB_AMD_SEARCH_LDS", -1);
  const bool timing = env_int("VIDSTAB_AMD_AUTOTUNE", 1) != 0;
  for (int k = 0; k < 2; ++k) {
    VSGStage& S = c->st[k];
    int K = 0, lds = 0;
    const bool ldsOk = S.S <= 128 && (size_t)S.S * S.S * c->bpp <= (size_t)kMaxLdsPatch;
    if (forced == 4 || forced == 6 || forced == 7 || forced == 8) {
      K = forced;
      lds = forcedLds == 1 && ldsOk;
    } else if (timing) {
      const TuneKey key(c->device, S.S, S.step, S.maxShift, c->B, S.nf, c->bpp);
      {
        std::lock_guard<std::mutex> lk(g_tuneMu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
          K = it->second & 15;
          lds = it->second >> 4;
        }
      }
      if (K == 0) {
        /* one full batch of the stage's jobs on set 0's buffers */
        const int njobs = c->B * S.nf;
        std::vector<VSGJob> jobs((size_t)njobs);
        for (int i = 0; i < c->B; ++i)
          for (int f = 0; f < S.nf; ++f) {
            VSGJob& j = jobs[(size_t)i * S.nf + f];
            j.slot = i;
            j.prev = (i + 1) % c->B;
            j.field = f;
            j.ox = 0;
            j.oy = 0;
          }
        VSGJob* dJobs = k == 0 ? c->set[0].d_jobsC : c->set[0].d_jobsF;
        TUNE_CHECK(hipMemcpy(dJobs, jobs.data(), sizeof(VSGJob) * (size_t)njobs, hipMemcpyHostToDevice));
        float bestMs = 0;
        for (int l = 0; l < 2; ++l) {
          if (l == 1 && (!ldsOk || forcedLds == 0)) break;
          if (l == 0 && forcedLds == 1 && ldsOk) continue;
          /* K = 4, 6, 8, then the three-column tile where it applies (its
             time is not among tuneMs) */
          VSGGrid probe;
          vsg_plan_grid(&probe, S.S, S.step, S.maxShift, 7, c->bpp);
          const int nplans = probe.J == 3 ? 4 : 3;
          for (int q = 0; q < nplans; ++q) {
            const int Kq = q < 3 ? kPlans[q] : 7;
            float ms = 0;
            if (time_plan(c, k, Kq, l, dJobs, njobs, &ms) != 0) return -1;
            if (q < 3) S.tuneMs[3 * l + q] = ms;
            if (K == 0 || ms < bestMs) {
              bestMs = ms;
              K = Kq;
              lds = l;
            }
          }
        }
        std::lock_guard<std::mutex> lk(g_tuneMu);
        g_tuned[key] = K + 16 * lds;
      }
    }
    if (K != 0) {
      S.tiles = vsg_plan_grid(&S.g, S.S, S.step, S.maxShift, K, c->bpp);
      S.g.lds = lds;
    }
  }
  return 0;
}

extern "C" int vsg_get_plan(VSGContext* c, int stage, int* out, float* tuneMs) {
  if (!c || stage < 0 || stage > 1 || !out) return vsg_fail("vsg_get_plan", hipSuccess);
  const VSGStage& S = c->st[stage];
  if (tuneMs)
    for (int q = 0; q < 6; ++q) tuneMs[q] = S.tuneMs[q];
  out[0] = S.g.J;
  out[1] = S.g.K;
  out[2] = S.g.D;
  out[3] = S.g.CW;
  out[4] = S.g.wg;
  out[5] = S.tiles;
  out[6] = S.g.lds;
  return 0;
}
