/*
 * vsg_learned.hip -- the learned-proposal coarse stage of a context.
 *
 * With a proposal network attached (vsg_set_proposal), a batch's coarse
 * stage becomes:
 *   1. every frame's raw luma is 4x4-downsampled once, in the front stage,
 *      into a per-slot quarter-resolution BF16 plane;
 *   2. the network's first layer reads frame i's previous plane (channel 0)
 *      and its own (channel 1) straight from those planes;
 *   3. the 6-layer network (vsg_cnn.hip) runs over the batch, its last layer
 *      writing only the two displacement channels, and each coarse field's
 *      proposal is their mean over the field's window (field size / 4 at
 *      the field centre / 4);
 *   4. k_sad_window searches the 5x5 window of 1-px shifts around each
 *      proposal; a minimum inside the window whose SAD is within the typical
 *      residual of the exact search is the field's result;
 *   5. every other field is listed for the exact full search (the same
 *      grid search and refinement as the plain form, k_sad_fallback spreading
 *      each listed job over more waves) -- the fallback that keeps the form
 *      honest.
 * The fine stage works the same way: the network output is also pooled over
 * every fine field's window when the network runs, and the fine search of
 * a field becomes a 5x5 window search around that displacement less the
 * field's search offset (which the host derives from the coarse result),
 * with the exact fine search as the fallback.  VIDSTAB_AMD_PROPOSAL_FINE=0
 * keeps the exact fine search for every field.  Everything on the host is
 * unchanged.
 */
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "vsg_cnn.h"
#include "vsg_context.h"

namespace {

#define LCHECK(call)                                      \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) return vsg_fail(#call, e_);     \
  } while (0)

constexpr int kC = 32;
constexpr int kMaxLayers = 8;

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

/* frame z of the batch -> quarter-resolution plane of its slot */
__global__ void k_down_slots(const VSGBlurSrc* __restrict__ src, int Wq, int Hq, uint16_t* __restrict__ down,
                             size_t planeStride) {
  const VSGBlurSrc s = src[blockIdx.z];
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (!s.p || x >= Wq) return;
  const uint8_t* p = s.p + (size_t)(4 * y) * s.pitch + 4 * x;
  uint32_t sum = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const uint8_t* q = p + (size_t)r * s.pitch;
    sum += (uint32_t)q[0] + q[1] + q[2] + q[3];
  }
  down[blockIdx.z * planeStride + (size_t)y * Wq + x] = f2bf(((float)sum * (1.0f / 16.0f) - 128.0f) * (1.0f / 64.0f));
}

__global__ void k_zero_u32(unsigned* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

struct VSGLearned;
static uint32_t max_sad(const VSGLearned* L, int S);

struct VSGLearned {
  int nl = 0;
  int dil[kMaxLayers] = {};
  float thr = 0;
  bool fuse01 = true;   /* layers 0 and 1 in one pass (vsgc_conv16_pair2) */
  uint16_t* d_pk[kMaxLayers] = {};   /* K = 16 packings */
  float* d_bias[kMaxLayers] = {};
  int Wq = 0, Hq = 0;
  size_t planeStride = 0;
  uint16_t* d_down = nullptr;      /* [slots][Hq][Wq] */
  uint16_t* d_downRef = nullptr;   /* [Hq][Wq] */
  uint16_t* d_act[2] = {};         /* [B][Hq][Wq][32] */
  float* d_out = nullptr;          /* [B][Hq][Wq][2] */
  int* d_win = nullptr;            /* [B][nc][5] */
  float* d_pred[VSG_NSETS] = {};   /* [B][nc][2]: kept per set until its coarse windows */
  bool fine = true;                /* proposals for the fine stage too */
  int walk = 4;                    /* windows a search may walk before the full search (1: none) */
  int* d_winF = nullptr;           /* [B][nff][5] */
  float* d_predF[VSG_NSETS] = {};  /* [B][nff][2]: kept per set until its fine stage */
  int* d_flist[VSG_NSETS] = {};    /* coarse jobs left to the exact search */
  int* d_flistF[VSG_NSETS] = {};   /* fine jobs left to the exact search */
  unsigned* d_stats[VSG_NSETS] = {};   /* 5 counters per stage: coarse, then fine */
  unsigned* h_stats[VSG_NSETS] = {};
  long long fields[2] = {0, 0}, fallbacks[2] = {0, 0}, why[2][3] = {{0, 0, 0}, {0, 0, 0}};
  /* the network buffers (d_act, d_out) are shared by the slot sets, whose
     streams run concurrently: each batch's network waits for the previous
     batch's (netDone, recorded after its pooling); the proposals it pools
     are per set, so the next network need not wait for the windows */
  hipEvent_t netDone = nullptr;
  bool netRecorded = false;
};

/* the acceptance bound of a window minimum: the residual threshold (a match
   value, SAD per pixel) times the field area */
static uint32_t max_sad(const VSGLearned* L, int S) {
  const double m = (double)L->thr * S * S;
  return (uint32_t)(m < 4294967295.0 ? m : 4294967295.0);
}

void vsg_learned_free(VSGContext* c) {
  VSGLearned* L = c->learned;
  if (!L) return;
  for (int k = 0; k < kMaxLayers; ++k) {
    if (L->d_pk[k]) (void)hipFree(L->d_pk[k]);
    if (L->d_bias[k]) (void)hipFree(L->d_bias[k]);
  }
  void* bufs[] = {L->d_down, L->d_downRef, L->d_act[0], L->d_act[1], L->d_out, L->d_win, L->d_winF};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (L->netDone) (void)hipEventDestroy(L->netDone);
  for (int s = 0; s < VSG_NSETS; ++s) {
    if (L->d_pred[s]) (void)hipFree(L->d_pred[s]);
    if (L->d_flist[s]) (void)hipFree(L->d_flist[s]);
    if (L->d_flistF[s]) (void)hipFree(L->d_flistF[s]);
    if (L->d_predF[s]) (void)hipFree(L->d_predF[s]);
    if (L->d_stats[s]) (void)hipFree(L->d_stats[s]);
    if (L->h_stats[s]) (void)hipHostFree(L->h_stats[s]);
  }
  delete L;
  c->learned = nullptr;
}

extern "C" int vsg_set_proposal(VSGContext* c, int nlayers, const int* dil, const float* weights, float threshold) {
  if (!c || nlayers < 1 || nlayers > kMaxLayers || !dil || !weights || !(threshold > 0.0f))
    return vsg_fail("vsg_set_proposal", hipSuccess);
  if (c->bpp != 1) return vsg_fail("vsg_set_proposal: the learned form takes planar frames", hipSuccess);
  if (c->W < 64 || c->H < 64) return vsg_fail("vsg_set_proposal: frame too small", hipSuccess);
  LCHECK(hipSetDevice(c->device));
  if (vsg_ctx_sync(c) != 0) return -1;
  vsg_learned_free(c);
  VSGLearned* L = new VSGLearned();
  c->learned = L;
  L->nl = nlayers;
  L->thr = threshold;
  L->Wq = c->W / 4;
  L->Hq = c->H / 4;
  L->planeStride = (size_t)L->Wq * L->Hq;
  const int B = c->B, nc = c->st[0].nf, nff = c->st[1].nf, slots = VSG_NSETS * B;
  const char* fineEnv = getenv("VIDSTAB_AMD_PROPOSAL_FINE");
  L->fine = !(fineEnv && atoi(fineEnv) == 0);
  const char* walkEnv = getenv("VIDSTAB_AMD_PROPOSAL_WALK");
  L->walk = walkEnv && *walkEnv ? (atoi(walkEnv) < 1 ? 1 : (atoi(walkEnv) > 8 ? 8 : atoi(walkEnv))) : 4;
  const char* fuseEnv = getenv("VIDSTAB_AMD_CONV_FUSE");
  L->fuse01 = !(fuseEnv && atoi(fuseEnv) == 0);
  const size_t mapElems = (size_t)B * L->Hq * L->Wq * kC;
  float* d_w = nullptr;
  const size_t perLayer = 9 * kC * kC + kC;
  int rc = 0;
  do {
    if (hipMalloc(&d_w, sizeof(float) * 9 * kC * kC) != hipSuccess) { rc = vsg_fail("proposal weights", hipErrorOutOfMemory); break; }
    for (int k = 0; k < nlayers && rc == 0; ++k) {
      L->dil[k] = dil[k];
      /* the tiled convolution (dilations 1..8) runs the fused first layer and
         the two-channel last layer */
      if (dil[k] < 1 || dil[k] > 8) { rc = vsg_fail("vsg_set_proposal: dilation outside 1..8", hipSuccess); break; }
      if (hipMalloc(&L->d_pk[k], sizeof(uint16_t) * 9 * 2 * 64 * 8) != hipSuccess ||
          hipMalloc(&L->d_bias[k], sizeof(float) * kC) != hipSuccess ||
          hipMemcpy(d_w, weights + k * perLayer, sizeof(float) * 9 * kC * kC, hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(L->d_bias[k], weights + k * perLayer + 9 * kC * kC, sizeof(float) * kC, hipMemcpyHostToDevice) !=
              hipSuccess ||
          vsgc_pack_weights16(d_w, L->d_pk[k], nullptr) != 0) {
        rc = vsg_fail("proposal weights", hipErrorOutOfMemory);
        break;
      }
    }
    if (rc) break;
    if (hipMalloc(&L->d_down, sizeof(uint16_t) * L->planeStride * slots) != hipSuccess ||
        hipMalloc(&L->d_downRef, sizeof(uint16_t) * L->planeStride) != hipSuccess ||
        hipMalloc(&L->d_act[0], sizeof(uint16_t) * mapElems) != hipSuccess ||
        hipMalloc(&L->d_act[1], sizeof(uint16_t) * mapElems) != hipSuccess ||
        hipMalloc(&L->d_out, sizeof(float) * 2 * (size_t)B * L->Hq * L->Wq) != hipSuccess ||
        hipMalloc(&L->d_win, sizeof(int) * 5 * (size_t)B * nc) != hipSuccess ||
        hipMalloc(&L->d_winF, sizeof(int) * 5 * (size_t)B * nff) != hipSuccess) {
      rc = vsg_fail("proposal buffers", hipErrorOutOfMemory);
      break;
    }
    if (hipMemset(L->d_down, 0, sizeof(uint16_t) * L->planeStride * slots) != hipSuccess ||
        hipMemset(L->d_downRef, 0, sizeof(uint16_t) * L->planeStride) != hipSuccess) {
      rc = vsg_fail("proposal buffers", hipErrorUnknown);
      break;
    }
    /* windows: field size / 4 around the field centre / 4, per frame slot,
       for the coarse fields (d_win) and the fine fields (d_winF) */
    for (int stage = 0; stage < 2 && rc == 0; ++stage) {
      const VSGStage& S = c->st[stage];
      std::vector<int16_t> fx(S.nf), fy(S.nf);
      if (hipMemcpy(fx.data(), S.d_fx, sizeof(int16_t) * S.nf, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(fy.data(), S.d_fy, sizeof(int16_t) * S.nf, hipMemcpyDeviceToHost) != hipSuccess) {
        rc = vsg_fail("proposal windows", hipErrorUnknown);
        break;
      }
      std::vector<int> win((size_t)5 * B * S.nf);
      const int h = S.S / 8;
      for (int z = 0; z < B; ++z)
        for (int f = 0; f < S.nf; ++f) {
          int* w = &win[5 * ((size_t)z * S.nf + f)];
          const int cx = fx[f] / 4, cy = fy[f] / 4;
          w[0] = z;
          w[1] = cx - h < 0 ? 0 : cx - h;
          w[2] = cy - h < 0 ? 0 : cy - h;
          w[3] = cx + h > L->Wq ? L->Wq : cx + h;
          w[4] = cy + h > L->Hq ? L->Hq : cy + h;
        }
      if (hipMemcpy(stage ? L->d_winF : L->d_win, win.data(), sizeof(int) * win.size(), hipMemcpyHostToDevice) !=
          hipSuccess)
        rc = vsg_fail("proposal windows", hipErrorUnknown);
    }
    if (rc) break;
    if (hipEventCreateWithFlags(&L->netDone, hipEventDisableTiming) != hipSuccess) {
      rc = vsg_fail("proposal event", hipErrorUnknown);
      break;
    }
    for (int s = 0; s < VSG_NSETS; ++s) {
      if (hipMalloc(&L->d_pred[s], sizeof(float) * 2 * (size_t)B * nc) != hipSuccess ||
          hipMalloc(&L->d_flist[s], sizeof(int) * (size_t)B * nc) != hipSuccess ||
          hipMalloc(&L->d_flistF[s], sizeof(int) * (size_t)B * nff) != hipSuccess ||
          hipMalloc(&L->d_predF[s], sizeof(float) * 2 * (size_t)B * nff) != hipSuccess ||
          hipMalloc(&L->d_stats[s], sizeof(unsigned) * 10) != hipSuccess ||
          hipHostMalloc(&L->h_stats[s], sizeof(unsigned) * 10, hipHostMallocDefault) != hipSuccess) {
        rc = vsg_fail("proposal buffers", hipErrorOutOfMemory);
        break;
      }
      for (int q = 0; q < 10; ++q) L->h_stats[s][q] = 0;
    }
  } while (0);
  if (d_w) (void)hipFree(d_w);
  if (rc == 0 && hipDeviceSynchronize() != hipSuccess) rc = vsg_fail("vsg_set_proposal", hipErrorUnknown);
  if (rc != 0) vsg_learned_free(c);
  return rc;
}

extern "C" int vsg_proposal_stats(VSGContext* c, long long* fields, long long* fallbacks) {
  return vsg_proposal_stage_stats(c, 0, fields, fallbacks);
}

extern "C" int vsg_proposal_stage_stats(VSGContext* c, int stage, long long* fields, long long* fallbacks) {
  if (!c || !c->learned || stage < 0 || stage > 1)
    return vsg_fail("vsg_proposal_stage_stats: no proposal network", hipSuccess);
  if (fields) *fields = c->learned->fields[stage];
  if (fallbacks) *fallbacks = c->learned->fallbacks[stage];
  return 0;
}

extern "C" int vsg_proposal_fallbacks(VSGContext* c, int stage, long long* outOfRange, long long* edge,
                                      long long* residual) {
  if (!c || !c->learned || stage < 0 || stage > 1)
    return vsg_fail("vsg_proposal_fallbacks: no proposal network", hipSuccess);
  if (outOfRange) *outOfRange = c->learned->why[stage][0];
  if (edge) *edge = c->learned->why[stage][1];
  if (residual) *residual = c->learned->why[stage][2];
  return 0;
}

/* front-stage part: quarter-resolution planes of the batch's frames */
int vsg_learned_front(VSGContext* c, int set, int n, int refSlot, hipStream_t st) {
  VSGLearned* L = c->learned;
  const int slot0 = set * c->B;
  hipLaunchKernelGGL(k_down_slots, dim3((L->Wq + 255) / 256, L->Hq, n), dim3(256), 0, st, c->set[set].d_src, L->Wq,
                     L->Hq, L->d_down + (size_t)slot0 * L->planeStride, L->planeStride);
  LCHECK(hipGetLastError());
  if (refSlot >= 0)
    LCHECK(hipMemcpyAsync(L->d_downRef, L->d_down + (size_t)refSlot * L->planeStride,
                          sizeof(uint16_t) * L->planeStride, hipMemcpyDeviceToDevice, st));
  return 0;
}

/* whether a network launched earlier may still run (then the next batch's
   blur, field copies and selection go ahead of its own network, so they
   overlap the running one; otherwise the network goes first) */
bool vsg_learned_net_busy(VSGContext* c) {
  VSGLearned* L = c->learned;
  return L && L->netRecorded && hipEventQuery(L->netDone) == hipErrorNotReady;
}

/* the network over the batch's planes and the pooling of its output into
   the set's coarse (and fine) proposals */
int vsg_learned_net(VSGContext* c, int set, int n, hipStream_t st) {
  VSGLearned* L = c->learned;
  VSGSet& s = c->set[set];
  const int slot0 = set * c->B, nc = c->st[0].nf;
  if (L->netRecorded) LCHECK(hipStreamWaitEvent(st, L->netDone, 0));
  const uint16_t* x = nullptr;
  /* the first two layers run fused (layer 0's map stays on chip) unless
     VIDSTAB_AMD_CONV_FUSE=0 */
  const bool fuse = L->nl >= 3 && L->fuse01;   /* layer 1 has a BF16 output */
  for (int k = 0; k < L->nl; ++k) {
    const bool last = k == L->nl - 1;
    uint16_t* y = last ? nullptr : L->d_act[k & 1];
    float* yf = last ? L->d_out : nullptr;
    const int mode = last ? 3 : 0;
    int rc;
    if (k == 0 && fuse) {
      continue;
    } else if (k == 1 && fuse) {
      rc = vsgc_conv16_pair2(L->d_down, L->planeStride, L->d_downRef, s.d_prev, slot0, n, L->Hq, L->Wq, L->d_pk[0],
                             L->d_bias[0], L->dil[0], L->d_pk[1], L->d_bias[1], L->dil[1], mode, y, yf, st);
    } else if (k == 0) {
      rc = vsgc_conv16_pair(L->d_down, L->planeStride, L->d_downRef, s.d_prev, slot0, n, L->Hq, L->Wq, L->d_pk[k],
                            L->d_bias[k], L->dil[k], mode, y, yf, st);
    } else {
      rc = vsgc_conv16(x, n, L->Hq, L->Wq, L->d_pk[k], L->d_bias[k], L->dil[k], mode, 32, y, yf, st);
    }
    if (rc != 0) return vsg_fail("proposal network", hipErrorLaunchFailure);
    x = y;
  }
  if (vsgc_field_pool2(L->d_out, L->Hq, L->Wq, L->d_win, n * nc, L->d_pred[set], st) != 0 ||
      (L->fine && vsgc_field_pool2(L->d_out, L->Hq, L->Wq, L->d_winF, n * c->st[1].nf, L->d_predF[set], st) != 0))
    return vsg_fail("proposal pooling", hipErrorLaunchFailure);
  LCHECK(hipEventRecord(L->netDone, st));
  L->netRecorded = true;
  return 0;
}

/* the coarse stage: window search around the network's proposals, exact
   fallback */
int vsg_learned_coarse(VSGContext* c, int set, int n, hipStream_t st) {
  VSGLearned* L = c->learned;
  VSGSet& s = c->set[set];
  const int slot0 = set * c->B, nc = c->st[0].nf;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, L->d_stats[set], 10);
  VSGSearchArgs ga = vsg_ctx_search_args(c, 0, s.d_jobsC, n * nc, s.d_best, s.d_done, s.d_resC);
  LCHECK(vsg_launch_window(ga, L->d_pred[set], nc, slot0, max_sad(L, c->st[0].S), L->d_flist[set], L->d_stats[set],
                           L->walk, st));
  LCHECK(vsg_launch_fallback(ga, L->d_flist[set], L->d_stats[set] + 1, st));
  LCHECK(hipMemcpyAsync(L->h_stats[set], L->d_stats[set], sizeof(unsigned) * 5, hipMemcpyDeviceToHost, st));
  return 0;
}

bool vsg_learned_fine_enabled(const VSGContext* c) { return c->learned && c->learned->fine; }

/* the fine stage over the host's job list: window search around each fine
   field's pooled displacement less its offset, exact search for the rest */
int vsg_learned_fine(VSGContext* c, int set, int njobs, hipStream_t st) {
  VSGLearned* L = c->learned;
  VSGSet& s = c->set[set];
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, L->d_stats[set] + 5, 5);
  VSGSearchArgs fa = vsg_ctx_search_args(c, 1, s.d_jobsF, njobs, s.d_best, s.d_done, s.d_resF);
  LCHECK(vsg_launch_window(fa, L->d_predF[set], c->st[1].nf, set * c->B, max_sad(L, c->st[1].S), L->d_flistF[set],
                           L->d_stats[set] + 5, L->walk, st));
  LCHECK(vsg_launch_fallback(fa, L->d_flistF[set], L->d_stats[set] + 6, st));
  LCHECK(hipMemcpyAsync(L->h_stats[set] + 5, L->d_stats[set] + 5, sizeof(unsigned) * 5, hipMemcpyDeviceToHost, st));
  return 0;
}

/* after the set's coarse (stage 0) or fine (stage 1) read-back completed */
void vsg_learned_account(VSGContext* c, int set, int stage) {
  VSGLearned* L = c->learned;
  if (!L) return;
  unsigned* h = L->h_stats[set] + 5 * stage;
  L->fields[stage] += h[0];
  L->fallbacks[stage] += h[1];
  for (int q = 0; q < 3; ++q) L->why[stage][q] += h[2 + q];
  for (int q = 0; q < 5; ++q) h[q] = 0;
}
