/*
 * vsg_search.hip -- the field search kernel of the detect pass on gfx950.
 *
 *   k_sad_search  exhaustive grid search of the reference's spiral
 *                 (src/motiondetect.c:466-518, SAD :971-996); the lane body is
 *                 vsg_sad_lane (vsg_device.h): candidates across lanes,
 *                 v_sad_u8 with the field as a wave-uniform SGPR operand,
 *                 argmin on (sad, spiral rank); then the dyadic refinement
 *                 around the grid winner, replayed in the reference's visit
 *                 order (src/motiondetect.c:545-567), by the job's last
 *                 workgroup
 *
 * k_sad_queue runs the same body, one persistent launch per stage, over one
 * flattened queue of the units of every clip of a batch (the batched
 * multi-clip form).
 *
 * The launch is a 1-D grid whose block -> work mapping is XCD-aware: the
 * hardware hands workgroups to the 8 XCDs round-robin (block b runs on XCD
 * b % 8), so units of work are grouped in chunks (normally the fields of one
 * frame) and every chunk is placed on a single XCD.  A frame's blurred planes
 * then stay in that XCD's L2 while all of its fields are searched.
 */
#include "vsg_device.h"
#include "vsg_internal.h"

namespace {

/* ---- refinement --------------------------------------------------------------
   Run by one whole workgroup (1-4 waves) per job.  Thread t handles word
   w = t % (S/4) of field rows t / (S/4), + rows-per-pass, ...: its loads of
   the field patch and of the prev rows are consecutive dwords across the
   wave (coalesced; prev words at any byte offset, which gfx950 global loads
   accept).  A level's ni x ni candidates (spacing ns around the current best)
   accumulate per thread, are summed across each wave and then across the
   waves in LDS, and every thread scans them in the reference's order
   (i outer, j inner, centre skipped) with the strict "<" against the running
   minimum -- exactly the sequential loop. */
typedef uint32_t __attribute__((aligned(1))) u32_any;
constexpr int kMaxNi = 5;   /* ni = 2r/ns + 1 <= 5 for every step */
constexpr int kNc = kMaxNi * kMaxNi;

struct RefineLds {
  uint32_t part[16][kNc];   /* per wave of a workgroup of up to 1,024 threads */
  uint32_t sums[kNc];
  unsigned long long wmin[4];
  unsigned long long key;
  int last;
};

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off);
  return v;
}

/* L.sums[ti*5 + tj] = SAD of candidate (x0 + ns*ti, y0 + ns*tj) for ti < nx,
   tj < ny (<= 5), summed by the whole workgroup; ends with a barrier.  A
   field row is S*bpp bytes; the workgroup has at least that many words'
   threads (the plan guarantees nthr >= S*bpp/4). */
__device__ void level_sums(const VSGSearchArgs& a, const VSGJob& jb, int x0, int y0, int ns, int nx, int ny,
                           RefineLds& L) {
  const int nthr = blockDim.x, nwaves = nthr >> 6;
  const int S = a.g.S, bpp = a.g.bpp, W4 = (S * bpp) >> 2;
  const int rowsPerPass = nthr / W4;
  const int w = threadIdx.x % W4, r0 = threadIdx.x / W4;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int fxc = a.fx[jb.field], fyc = a.fy[jb.field];
  const uint8_t* prev = jb.prev < 0 ? a.ref : a.blurred + (size_t)jb.prev * a.slotStride;
  const uint32_t* patch = reinterpret_cast<const uint32_t*>(a.patches + (size_t)jb.slot * a.patchSlotStride +
                                                            a.stageOff + (size_t)jb.field * S * S * bpp);
  const int nsb = ns * bpp;
  uint32_t acc[kNc];
#pragma unroll
  for (int c = 0; c < kNc; ++c) acc[c] = 0;
  if (r0 < rowsPerPass) {
    const uint8_t* colBase = prev + (fxc - S / 2 + x0 + jb.ox) * bpp + 4 * w;
    const int rowBase = fyc - S / 2 + y0 + jb.oy;
#pragma unroll 2
    for (int y = r0; y < S; y += rowsPerPass) {
      const uint32_t c = patch[y * W4 + w];
#pragma unroll
      for (int tj = 0; tj < kMaxNi; ++tj) {
        if (tj < ny) {
          const uint8_t* p = colBase + (size_t)(rowBase + y + ns * tj) * a.pitch;
#pragma unroll
          for (int ti = 0; ti < kMaxNi; ++ti)
            if (ti < nx)
              acc[ti * kMaxNi + tj] =
                  __builtin_amdgcn_sad_u8(*reinterpret_cast<const u32_any*>(p + nsb * ti), c, acc[ti * kMaxNi + tj]);
        }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < kNc; ++c) {
    const uint32_t v = wave_sum(acc[c]);
    if (lane == 0) L.part[wave][c] = v;
  }
  __syncthreads();
  if (threadIdx.x < kNc) {
    uint32_t t = 0;
    for (int q = 0; q < nwaves; ++q) t += L.part[q][threadIdx.x];
    L.sums[threadIdx.x] = t;
  }
  __syncthreads();
}

/* Packed frames (src/motiondetect.c:613-645): the unshifted position first
   (rank 0, ahead of every grid point), then one level of 1-px shifts within
   r = step-1 of the grid winner, x outer / y inner, centre skipped, strict
   "<".  The (2r+1)^2 square is summed in blocks of at most 5 x 5; the
   sequential scan's winner is the minimum of (sad, visit position), with the
   grid winner at position 0. */
__device__ void refine_packed(const VSGSearchArgs& a, const VSGJob& jb, unsigned long long key, int job,
                              RefineLds& L) {
  const int step = a.g.step;
  level_sums(a, jb, 0, 0, 1, 1, 1, L);
  const unsigned long long k0 = (unsigned long long)L.sums[0] << 32;
  key = k0 < key ? k0 : key;
  __syncthreads();
  const int q = (int)(uint32_t)key;
  int tx = 0, ty = 0;
  if (q > 0) {
    tx = a.g.org + ((q - 1) / a.g.G) * step;
    ty = a.g.org + ((q - 1) % a.g.G) * step;
  }
  unsigned long long best = key & 0xFFFFFFFF00000000ull;
  const int r = step - 1, ni = 2 * r + 1;
  if (r > 0) {
    for (int bi = 0; bi < ni; bi += kMaxNi)
      for (int bj = 0; bj < ni; bj += kMaxNi) {
        const int nx = min(kMaxNi, ni - bi), ny = min(kMaxNi, ni - bj);
        level_sums(a, jb, tx - r + bi, ty - r + bj, 1, nx, ny, L);
        for (int ti = 0; ti < nx; ++ti)
          for (int tj = 0; tj < ny; ++tj) {
            const int gi = bi + ti, gj = bj + tj;
            if (gi == r && gj == r) continue;
            const unsigned long long kk =
                ((unsigned long long)L.sums[ti * kMaxNi + tj] << 32) | (uint32_t)(1 + gi * ni + gj);
            best = kk < best ? kk : best;
          }
        __syncthreads();   /* sums[] is rewritten by the next block */
      }
  }
  const int p = (int)(uint32_t)best;
  if (p > 0) {
    tx += (p - 1) / ni - r;
    ty += (p - 1) % ni - r;
  }
  if (threadIdx.x == 0) {
    VSGResult res;
    res.tx = (int16_t)tx;
    res.ty = (int16_t)ty;
    res.minerror = (uint32_t)(best >> 32);
    a.results[job] = res;
  }
}

__device__ void refine_job(const VSGSearchArgs& a, const VSGJob& jb, unsigned long long key, int job,
                           RefineLds& L) {
  if (a.g.bpp != 1) {
    refine_packed(a, jb, key, job, L);
    return;
  }
  int ga, gb;
  vsg_unrank((int)(uint32_t)key, &ga, &gb);
  int tx = ga * a.g.step, ty = gb * a.g.step;
  uint32_t minerr = (uint32_t)(key >> 32);
  for (int st = a.g.step; st > 1; st /= 2) {
    const int ns = st / 2, r = st - ns;
    const int ni = (2 * r) / ns + 1;
    level_sums(a, jb, tx - r, ty - r, ns, ni, ni, L);
    int btx = tx, bty = ty;
    for (int ti = 0; ti < ni; ++ti)
      for (int tj = 0; tj < ni; ++tj) {
        const int i = tx - r + ns * ti, j = ty - r + ns * tj;
        const uint32_t e = L.sums[ti * kMaxNi + tj];
        if (!(i == tx && j == ty) && e < minerr) {
          minerr = e;
          btx = i;
          bty = j;
        }
      }
    tx = btx;
    ty = bty;
    __syncthreads();   /* part[] / sums[] are rewritten by the next level */
  }
  if (threadIdx.x == 0) {
    VSGResult res;
    res.tx = (int16_t)tx;
    res.ty = (int16_t)ty;
    res.minerror = minerr;
    a.results[job] = res;
  }
}

/* job jb's search geometry (the LDS patch pointer is left null) */
__device__ __forceinline__ VSGTile make_tile(const VSGSearchArgs& a, const VSGJob& jb) {
  VSGTile t;
  t.S = a.g.S;
  t.s = a.g.step;
  t.alo = a.g.alo;
  t.G = a.g.G;
  t.org = a.g.org;
  t.bpp = a.g.bpp;
  t.m = a.g.m;
  t.U = a.g.U;
  t.V = a.g.V;
  t.fxc = a.fx[jb.field];
  t.fyc = a.fy[jb.field];
  t.ox = jb.ox;
  t.oy = jb.oy;
  t.pitch = a.pitch;
  t.prev = jb.prev < 0 ? a.ref : a.blurred + (size_t)jb.prev * a.slotStride;
  t.patch = (vsg_cword_ptr)(a.patches + (size_t)jb.slot * a.patchSlotStride + a.stageOff +
                            (size_t)jb.field * t.S * t.S * t.bpp);
  t.lpatch = nullptr;
  return t;
}

/* ---- grid search + refinement ---------------------------------------------------
   Workgroup (job, tile) searches its share of the job's candidate grid.  With
   one tile per job the workgroup's argmin is reduced in LDS; otherwise every
   wave folds its argmin into best[job] with a u64 atomic min, and the
   workgroup that finishes the job last (counted in done[job]) reads the
   winner and runs the refinement, then resets best / done for the next
   launch.  The refinement of early jobs overlaps the grid of later ones. */
template <int J, int K, int D, int CW, bool LDSP>
__device__ __forceinline__ void search_wg(const VSGSearchArgs& a, int job, int tile, RefineLds& L) {
  extern __shared__ uint32_t lpatch[];   /* LDSP: the field, S*S*bpp/4 words */
  const VSGJob jb = a.jobs[job];
  if (jb.field < 0) return;   /* unselected slot of a fixed-size job list */
  VSGTile t = make_tile(a, jb);
  t.lpatch = lpatch;
  const size_t patchBytes = (size_t)t.S * t.S * t.bpp;
  if (LDSP) {
    /* the whole workgroup reads the same field words: stage them once, and
       every lane then reads them by LDS broadcast instead of scalar loads
       that miss a scalar cache shared by many resident fields */
    const uint4* src = reinterpret_cast<const uint4*>(a.patches + (size_t)jb.slot * a.patchSlotStride + a.stageOff +
                                                      (size_t)jb.field * patchBytes);
    uint4* dst = reinterpret_cast<uint4*>(lpatch);
    for (int i = threadIdx.x; i < (int)(patchBytes / 16); i += blockDim.x) dst[i] = src[i];
    __syncthreads();
  }
  unsigned long long best = vsg_sad_lane<J, K, D, CW, LDSP>(t, tile * blockDim.x + threadIdx.x);
  best = vsg_wave_min_u64(best);
  const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  if (a.tiles == 1) {
    if ((threadIdx.x & 63) == 0) L.wmin[wave] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long k = L.wmin[0];
      for (int q = 1; q < nwaves; ++q) k = L.wmin[q] < k ? L.wmin[q] : k;
      L.key = k;
    }
    __syncthreads();
  } else {
    if ((threadIdx.x & 63) == 0) {
      if (best != ~0ull) atomicMin(&a.best[job], best);
      __threadfence();   /* the min is visible before this workgroup counts as done */
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const unsigned before = atomicAdd(&a.done[job], 1u);
      L.last = before == (unsigned)a.tiles - 1;
      if (L.last) {
        __threadfence();
        L.key = __hip_atomic_load(&a.best[job], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        a.best[job] = ~0ull;   /* ready for the next launch */
        a.done[job] = 0;
      }
    }
    __syncthreads();
    if (!L.last) return;
  }
  refine_job(a, jb, L.key, job, L);
}

template <int J, int K, int D, int CW, bool LDSP>
__global__ __launch_bounds__(256) void k_sad_search(VSGSearchArgs a) {
  __shared__ RefineLds L;
  const int unit = vsg_xcd_unit(blockIdx.x, a.chunk, a.njobs * a.tiles);
  if (unit < 0) return;
  const int job = unit / a.tiles;
  search_wg<J, K, D, CW, LDSP>(a, job, unit - job * a.tiles, L);
}

/* ---- the multi-clip queue kernel -----------------------------------------------
   One persistent launch per stage for ALL clips of a batch: the units of
   every clip and every plan shape form one flattened queue.  A unit is one
   workgroup-wide tile of a job's candidates (256 lanes for the coarse stage,
   whose waves then share the field's scalar loads, 128 for the fine one),
   or, for a job that fills a small part of a workgroup (a split segment), one
   64-lane tile whose prev rows the workgroup's waves share out.  A
   workgroup pulls the next units (a few consecutive ones) with an atomic
   counter, runs the grid body of the unit's plan (selected by the
   segment's variant: J, D and the column chunk CW; K candidate rows per lane
   is the launch's), folds its argmin into best[job], and the job's last
   unit refines it -- the same code as k_sad_search.  No workgroup ever waits
   for another: the only barriers are inside a workgroup, reached by all its
   threads for every unit, and a job's refinement is run by whichever unit
   finishes it last.  There is one queue per XCD (workgroup b serves queue
   b % 8): queue x holds the x-th eighth of every clip's job list,
   consecutive jobs of consecutive frames, so a frame's planes are searched
   from one L2 (or a single queue for the whole grid, balanced across the
   XCDs).  When its queue is empty a workgroup counts itself out; the last
   one out of a queue resets that queue's counters for the next launch
   (every pull of every member happened before its own count, so no pull can
   follow the reset).  Segments are ordered by descending field size, so the
   largest units are pulled first and the tail is made of small ones. */
constexpr int kQueueWaves = 256 * 16;   /* resident waves: 256 CUs x 16 */

/* a split unit: one 64-lane tile of a job whose candidates fill a small part
   of the workgroup, its prev rows shared out over all W waves of the
   workgroup (each sums 1/W of them; the partial sums meet in LDS), so no
   wave idles and the tile's latency is 1/W of one lane's walk.  With one
   tile per job the wave minimum is the job's grid winner; otherwise it is
   folded into best[job] and the job's last unit refines, as in search_wg. */
constexpr int kQueueMaxWaves = 4;   /* 256-lane workgroups */

template <int J, int K, int D, int CW>
__device__ __forceinline__ void split_unit(const VSGSearchArgs& a, int job, int tile, RefineLds& L,
                                           uint32_t (*part)[K * 4][64]) {
  const VSGJob jb = a.jobs[job];
  if (jb.field < 0) return;   /* unselected slot: uniform over the workgroup */
  const VSGTile t = make_tile(a, jb);
  const int W = blockDim.x >> 6, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int prows = t.S + (K - 1) * t.s;
  VSGLane<J, K, D, CW, false> ln;
  vsg_sad_acc<J, K, D, CW, false>(t, tile * 64 + lane, wave * prows / W, (wave + 1) * prows / W, ln);
  if (wave > 0) {
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int j = 0; j < J; ++j) part[wave - 1][k * J + j][lane] = ln.acc[k][j];
  }
  __syncthreads();
  if (wave == 0) {
    for (int w = 0; w < W - 1; ++w)
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int j = 0; j < J; ++j) ln.acc[k][j] += part[w][k * J + j][lane];
    const unsigned long long best = vsg_wave_min_u64(vsg_lane_best<J, K, D, CW, false>(t, ln));
    if (lane == 0) {
      if (a.tiles == 1) {
        L.key = best;
        L.last = 1;
      } else {
        if (best != ~0ull) atomicMin(&a.best[job], best);
        __threadfence();   /* the min is visible before this unit counts as done */
        const unsigned before = atomicAdd(&a.done[job], 1u);
        L.last = before == (unsigned)a.tiles - 1;
        if (L.last) {
          __threadfence();
          L.key = __hip_atomic_load(&a.best[job], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          a.best[job] = ~0ull;   /* ready for the next launch */
          a.done[job] = 0;
        }
      }
    }
  }
  __syncthreads();
  if (L.last) refine_job(a, jb, L.key, job, L);
}

template <int J, int K, int D, int CW, bool SPLIT>
__device__ __forceinline__ void queue_unit(const VSGSearchArgs& a, int job, int tile, RefineLds& L,
                                           uint32_t (*part)[K * 4][64]) {
  if constexpr (SPLIT)
    split_unit<J, K, D, CW>(a, job, tile, L, part);
  else
    search_wg<J, K, D, CW, false>(a, job, tile, L);
}

/* SPLIT: the queue of split units (K = 4), whose partial sums need LDS the
   wide units' kernel does without (it keeps its occupancy) */
template <int K, bool SPLIT>
__global__ __launch_bounds__(256) void k_sad_queue(const VSGMultiSeg* __restrict__ segs, int nseg,
                                                   VSGQueueTotals totals, unsigned* __restrict__ queues) {
  __shared__ RefineLds L;
  __shared__ int pulled;
  __shared__ uint32_t partS[SPLIT ? kQueueMaxWaves - 1 : 1][SPLIT ? K * 4 : 1][SPLIT ? 64 : 1];
  uint32_t(*part)[K * 4][64] = SPLIT ? reinterpret_cast<uint32_t(*)[K * 4][64]>(&partS[0][0][0]) : nullptr;
  /* workgroup b serves queue b % nq; with 8 queues, the queue of its XCD */
  const int nq = totals.queues;
  const int x = (int)(blockIdx.x % (unsigned)nq);
  const VSGMultiSeg* qs = segs + (size_t)x * nseg;
  const int totalUnits = totals.units[x], chunk = totals.chunk;
  unsigned* queue = queues + 2 * x;
  int seg = 0;   /* segment of the previous unit: consecutive units mostly share it */
  for (;;) {
    /* a pull takes `chunk` consecutive units: fewer atomics on one address */
    if (threadIdx.x == 0) pulled = (int)atomicAdd(&queue[0], (unsigned)chunk);
    __syncthreads();
    const int first = pulled;
    __syncthreads();   /* `pulled` is rewritten by the next pull */
    if (first >= totalUnits) break;
    const int last = min(first + chunk, totalUnits);
    for (int unit = first; unit < last; ++unit) {
      const int end = seg + 1 < nseg ? qs[seg + 1].unitBase : totalUnits;
      if (unit < qs[seg].unitBase || unit >= end) {
        int lo = 0, hi = nseg - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (qs[mid].unitBase <= unit)
            lo = mid;
          else
            hi = mid - 1;
        }
        seg = lo;
      }
      const VSGMultiSeg* sg = qs + seg;
      const int local = unit - sg->unitBase;
      const int job = sg->jobBase + local / sg->a.tiles, tile = local % sg->a.tiles;
      switch (sg->variant) {
#define VSG_QUEUE_CASE(jd, cwi, J, D, CW) \
  case (jd) * 6 + (cwi): queue_unit<J, K, D, CW, SPLIT>(sg->a, job, tile, L, part); break;
        VSG_QUEUE_CASE(0, 0, 4, 1, 28) VSG_QUEUE_CASE(0, 1, 4, 1, 20) VSG_QUEUE_CASE(0, 2, 4, 1, 16)
        VSG_QUEUE_CASE(0, 3, 4, 1, 12) VSG_QUEUE_CASE(0, 4, 4, 1, 8) VSG_QUEUE_CASE(0, 5, 4, 1, 4)
        VSG_QUEUE_CASE(1, 0, 4, 3, 28) VSG_QUEUE_CASE(1, 1, 4, 3, 20) VSG_QUEUE_CASE(1, 2, 4, 3, 16)
        VSG_QUEUE_CASE(1, 3, 4, 3, 12) VSG_QUEUE_CASE(1, 4, 4, 3, 8) VSG_QUEUE_CASE(1, 5, 4, 3, 4)
        VSG_QUEUE_CASE(2, 0, 1, 1, 28) VSG_QUEUE_CASE(2, 1, 1, 1, 20) VSG_QUEUE_CASE(2, 2, 1, 1, 16)
        VSG_QUEUE_CASE(2, 3, 1, 1, 12) VSG_QUEUE_CASE(2, 4, 1, 1, 8) VSG_QUEUE_CASE(2, 5, 1, 1, 4)
#undef VSG_QUEUE_CASE
        default: break;
      }
      __syncthreads();   /* L is rewritten by the next unit */
    }
  }
  if (threadIdx.x == 0) {
    /* the workgroups serving queue x: b = x, x + nq, ... < gridDim.x */
    const unsigned members = (gridDim.x - (unsigned)x + (unsigned)nq - 1) / (unsigned)nq;
    __threadfence();
    if (atomicAdd(&queue[1], 1u) == members - 1) {
      queue[0] = 0;
      queue[1] = 0;
    }
  }
}

/* ---- exact search of a fallback list ----------------------------------------------
   A fallback list holds few jobs, so its launch lasts as long as one job's
   search.  The job's candidates are cut into 64-lane tiles with the plan's K
   forced to 4 (the most lanes per job), and the WPT waves of a workgroup
   share one tile, each summing 1/WPT of the prev rows; the partial sums meet
   in LDS, so a tile's latency is 1/WPT of a grid-kernel lane's.  A fixed
   grid walks the units (list entry, tile) of the count that the window
   search left; the job's last workgroup refines it with all its threads,
   as in the grid kernel.  WPT = 4 for the learned form's fallback lists,
   8 for plain batches too small to fill the GPU (a single frame). */
constexpr int kFallbackBlocks = 2048;

template <int J, int D, int CW, int WPT>
__global__ __launch_bounds__(64 * WPT) void k_sad_fallback(VSGSearchArgs a, const int* __restrict__ list,
                                                           const unsigned* __restrict__ count) {
  constexpr int K = 4;
  __shared__ RefineLds L;
  __shared__ uint32_t part[WPT - 1][K * J][64];
  /* no list: every job of a.jobs (a whole small batch), skipping unselected
     slots (field < 0; uniform over the workgroup, which holds one job) */
  const unsigned n = list ? min(*count, (unsigned)a.njobs) : (unsigned)a.njobs;
  const unsigned units = n * (unsigned)a.tiles;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (unsigned u = blockIdx.x; u < units; u += gridDim.x) {
    const int job = list ? list[u / (unsigned)a.tiles] : (int)(u / (unsigned)a.tiles);
    const int tile = (int)(u % (unsigned)a.tiles);
    const VSGJob jb = a.jobs[job];
    if (jb.field < 0) continue;
    const VSGTile t = make_tile(a, jb);
    const int prows = t.S + (K - 1) * t.s;
    VSGLane<J, K, D, CW, false> ln;
    vsg_sad_acc<J, K, D, CW, false>(t, tile * 64 + lane, wave * prows / WPT, (wave + 1) * prows / WPT, ln);
    if (wave > 0) {
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int j = 0; j < J; ++j) part[wave - 1][k * J + j][lane] = ln.acc[k][j];
    }
    __syncthreads();
    if (wave == 0) {
      for (int w = 0; w < WPT - 1; ++w)
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
          for (int j = 0; j < J; ++j) ln.acc[k][j] += part[w][k * J + j][lane];
      const unsigned long long best = vsg_wave_min_u64(vsg_lane_best<J, K, D, CW, false>(t, ln));
      if (lane == 0) {
        if (best != ~0ull) atomicMin(&a.best[job], best);
        __threadfence();   /* the min is visible before this workgroup counts as done */
        const unsigned before = atomicAdd(&a.done[job], 1u);
        L.last = before == (unsigned)a.tiles - 1;
        if (L.last) {
          __threadfence();
          L.key = __hip_atomic_load(&a.best[job], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          a.best[job] = ~0ull;   /* ready for the next launch */
          a.done[job] = 0;
        }
      }
    }
    __syncthreads();
    if (L.last) refine_job(a, jb, L.key, job, L);
    __syncthreads();   /* part[] and L are rewritten by the next unit */
  }
}

/* ---- proposal window (learned form) -----------------------------------------------
   Job j's field gets proposal p = round(4 * pred) - (ox, oy): the network's
   displacement (quarter-resolution units) of the field, less the job's
   search offset, as a shift of the stage's search.  pred is indexed by the
   job's frame in the batch (slot - slot0) and its field, npred fields per
   frame.  The 5x5 window of 1-px shifts around p is searched exhaustively,
   keeping the reference refinement's rule (the centre first, then x outer /
   y inner, strict "<"); while its minimum lies on the window's edge the
   window moves to that minimum (a descent of at most `steps` windows; 1
   falls back at the first edge minimum).
   The result is accepted when the minimum is inside the window and its SAD
   is at most maxSad; then it is written to results[j], otherwise j is
   appended to the fallback list for the exact full search.

   A job runs on WJ waves (1 for fields of at most 64 px, else 4; a 256-
   thread workgroup holds 4 / WJ jobs).  Thread t of a job takes word
   w = t % (S/4) of a run of consecutive field rows (run g = t / (S/4)), so
   the 5 previous-frame rows a field row meets at the window's 5 vertical
   shifts are shared with the next row: each field row costs one patch word,
   two (unaligned) dwords of one new previous row, 4 alignbytes for the 5
   horizontal shifts and the 25 v_sad_u8.  The 25 sums are reduced across
   the wave by a transposing butterfly (a lane keeps half of its values and
   sends the other half at each of 5 steps: 31 exchanges), after which lane l
   holds candidate l / 2's wave total, and the scan's winner is the minimum
   of (sad, scan rank) over the lanes. */

struct WindowLds {
  uint32_t part[2][4][32];   /* [iteration parity][wave][candidate] (4-wave jobs) */
  unsigned cnt[5];           /* the workgroup's outcome counts (cnt[1] unused) */
};

__device__ __forceinline__ uint32_t ld_any(const uint8_t* p) { return *reinterpret_cast<const u32_any*>(p); }

/* rank of candidate c = ti*5 + tj in the reference's scan: the centre first,
   then x outer / y inner */
__device__ __forceinline__ uint32_t window_rank(int c) { return c == 12 ? 0u : (c < 12 ? (uint32_t)c + 1u : (uint32_t)c); }

/* the job's 25 window SADs at shifts (x0 + ti, y0 + tj), left in lane l as
   candidate (l >> 1)'s total over the wave (lanes of candidates >= 25: 0) */
template <int WJ>
__device__ uint32_t window_wave_sums(const VSGSearchArgs& a, const VSGJob& jb, int x0, int y0, int t) {
  const int S = a.g.S, W4 = S >> 2;
  const int groups = (64 * WJ) / W4;
  const int run = (S + groups - 1) / groups;
  const int w = t % W4, g = t / W4;
  const int r0 = g * run;
  const int len = g < groups ? min(run, S - r0) : 0;
  uint32_t acc[32];
#pragma unroll
  for (int c = 0; c < 32; ++c) acc[c] = 0;
  if (len > 0) {
    const uint8_t* prev = jb.prev < 0 ? a.ref : a.blurred + (size_t)jb.prev * a.slotStride;
    const uint32_t* patch = reinterpret_cast<const uint32_t*>(a.patches + (size_t)jb.slot * a.patchSlotStride +
                                                              a.stageOff + (size_t)jb.field * S * S) +
                            (size_t)r0 * W4 + w;
    const uint8_t* row = prev + (size_t)(a.fy[jb.field] - S / 2 + y0 + jb.oy + r0) * a.pitch +
                         (a.fx[jb.field] - S / 2 + x0 + jb.ox) + 4 * w;
    const int pitch = a.pitch;
    /* sw[q % 5][i]: previous row q (relative to the run) shifted by i bytes */
    uint32_t sw[5][5];
/* v_alignbyte_b32 takes the shift modulo 4: shift 4 is the high dword itself */
#define WIN_SHIFTS(slot, lo, hi)                                               \
  _Pragma("unroll") for (int i = 0; i < 4; ++i) sw[slot][i] = __builtin_amdgcn_alignbyte(hi, lo, i); \
  sw[slot][4] = hi;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t lo = ld_any(row + (size_t)q * pitch), hi = ld_any(row + (size_t)q * pitch + 4);
      WIN_SHIFTS(q, lo, hi)
    }
    uint32_t nlo = ld_any(row + (size_t)4 * pitch), nhi = ld_any(row + (size_t)4 * pitch + 4);
    uint32_t nc = patch[0];
    for (int yb = 0; yb < len; yb += 5) {
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const int y = yb + k;
        if (y < len) {
          WIN_SHIFTS((k + 4) % 5, nlo, nhi)
          const uint32_t c = nc;
          if (y + 1 < len) {
            const uint8_t* pr = row + (size_t)(y + 5) * pitch;
            nlo = ld_any(pr);
            nhi = ld_any(pr + 4);
            nc = patch[(size_t)(y + 1) * W4];
          }
#pragma unroll
          for (int tj = 0; tj < 5; ++tj)
#pragma unroll
            for (int ti = 0; ti < 5; ++ti)
              acc[ti * 5 + tj] = __builtin_amdgcn_sad_u8(sw[(k + tj) % 5][ti], c, acc[ti * 5 + tj]);
        }
      }
    }
#undef WIN_SHIFTS
  }
  /* transposing butterfly: after the step over lane bit m (32 .. 2), a lane
     holds the half of the candidates whose index bit matches its lane bit */
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int h = 16; h >= 1; h >>= 1) {
    const bool up = (lane & (2 * h)) != 0;
#pragma unroll
    for (int k = 0; k < h; ++k) {
      const uint32_t keep = up ? acc[k + h] : acc[k];
      const uint32_t send = up ? acc[k] : acc[k + h];
      acc[k] = keep + (uint32_t)__shfl_xor((int)send, 2 * h);
    }
  }
  return acc[0] + (uint32_t)__shfl_xor((int)acc[0], 1);
}

__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, off));
  return v;
}

template <int WJ>
__global__ __launch_bounds__(256) void k_sad_window(VSGSearchArgs a, const float* __restrict__ pred, int npred,
                                                    int slot0, uint32_t maxSad, int* __restrict__ list,
                                                    unsigned* __restrict__ stats, int steps) {
  __shared__ WindowLds L;
  constexpr int JB = 4 / WJ;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int jw = wave / WJ;                      /* the job's index in the workgroup */
  const int t = threadIdx.x - jw * 64 * WJ;      /* thread index within the job */
  const int job = blockIdx.x * JB + jw;
  if (threadIdx.x < 5) L.cnt[threadIdx.x] = 0;
  __syncthreads();
  /* job and everything derived from it are uniform over the job's waves */
  const bool live = job < a.njobs;
  VSGJob jb;
  jb.field = -1;
  if (live) jb = a.jobs[job];
  int outcome = -1;   /* -1: skipped, 0: accepted, 2: out of range, 3: edge, 4: residual */
  if (live && jb.field >= 0) {
    const float* pp = pred + 2 * ((size_t)(jb.slot - slot0) * npred + jb.field);
    const float fx = rintf(4.0f * pp[0]) - (float)jb.ox, fy = rintf(4.0f * pp[1]) - (float)jb.oy;
    const float lim = (float)(a.g.n * a.g.step - 2);   /* the full search's reach, less the window */
    bool inRange = fx >= -lim && fx <= lim && fy >= -lim && fy <= lim, interior = false;
    int bx = 0, by = 0;
    uint32_t minerr = 0;
    if (inRange) {
      /* a minimum on the window's edge moves the window there, up to
         `steps` windows in all */
      int px = (int)fx, py = (int)fy;
      for (int it = 0; it < steps; ++it) {
        uint32_t tot = window_wave_sums<WJ>(a, jb, px - 2, py - 2, t);
        const int cand = lane >> 1;
        if (WJ > 1) {
          uint32_t(*part)[32] = L.part[it & 1];
          if ((lane & 1) == 0) part[wave][cand] = tot;
          __syncthreads();
          tot = 0;
#pragma unroll
          for (int q = 0; q < WJ; ++q) tot += part[q][cand];
        }
        const uint32_t key = wave_min(cand < 25 ? (tot << 5) | window_rank(cand) : 0xffffffffu);
        const int c = (int)(key & 31u) == 0 ? 12 : ((int)(key & 31u) <= 12 ? (int)(key & 31u) - 1 : (int)(key & 31u));
        minerr = key >> 5;
        bx = px - 2 + c / 5;
        by = py - 2 + c % 5;
        interior = bx > px - 2 && bx < px + 2 && by > py - 2 && by < py + 2;
        if (interior) break;
        px = bx;
        py = by;
        inRange = px >= -lim && px <= lim && py >= -lim && py <= lim;
        if (!inRange) break;
      }
    }
    const bool accept = inRange && interior && minerr <= maxSad;
    outcome = accept ? 0 : (!inRange ? 2 : (!interior ? 3 : 4));
    if (t == 0 && accept) {
      VSGResult res;
      res.tx = (int16_t)bx;
      res.ty = (int16_t)by;
      res.minerror = minerr;
      a.results[job] = res;
    }
  }
  if (live && t == 0 && outcome >= 0) {
    atomicAdd(&L.cnt[0], 1u);
    if (outcome > 0) {
      list[atomicAdd(&stats[1], 1u)] = job;
      atomicAdd(&L.cnt[outcome], 1u);
    }
  }
  __syncthreads();
  if (threadIdx.x < 5 && threadIdx.x != 1 && L.cnt[threadIdx.x]) atomicAdd(&stats[threadIdx.x], L.cnt[threadIdx.x]);
}

template <int J, int K, int D, int CW>
hipError_t launch_grid(const VSGSearchArgs& a, hipStream_t st) {
  const int nb = vsg_xcd_blocks(a.njobs * a.tiles, a.chunk);
  if (a.g.lds)
    hipLaunchKernelGGL((k_sad_search<J, K, D, CW, true>), dim3(nb), dim3(a.g.wg), vsg_patch_bytes(a.g), st, a);
  else
    hipLaunchKernelGGL((k_sad_search<J, K, D, CW, false>), dim3(nb), dim3(a.g.wg), 0, st, a);
  return hipGetLastError();
}

template <int J, int K, int D>
hipError_t launch_grid_cw(const VSGSearchArgs& a, hipStream_t st) {
  switch (a.g.CW) {
    case 28: return launch_grid<J, K, D, 28>(a, st);
    case 20: return launch_grid<J, K, D, 20>(a, st);
    case 16: return launch_grid<J, K, D, 16>(a, st);
    case 12: return launch_grid<J, K, D, 12>(a, st);
    case 8: return launch_grid<J, K, D, 8>(a, st);
    default: return launch_grid<J, K, D, 4>(a, st);
  }
}

template <int J, int D>
hipError_t launch_grid_k(const VSGSearchArgs& a, hipStream_t st) {
  switch (a.g.K) {
    case 8: return launch_grid_cw<J, 8, D>(a, st);
    case 6: return launch_grid_cw<J, 6, D>(a, st);
    default: return launch_grid_cw<J, 4, D>(a, st);
  }
}

template <int J, int D, int WPT>
hipError_t launch_fallback(const VSGSearchArgs& b, const int* list, const unsigned* count, hipStream_t st) {
  const int nb = min(kFallbackBlocks, b.njobs * b.tiles);
  const dim3 wg(64 * WPT);
  switch (b.g.CW) {
    case 28: hipLaunchKernelGGL((k_sad_fallback<J, D, 28, WPT>), dim3(nb), wg, 0, st, b, list, count); break;
    case 20: hipLaunchKernelGGL((k_sad_fallback<J, D, 20, WPT>), dim3(nb), wg, 0, st, b, list, count); break;
    case 16: hipLaunchKernelGGL((k_sad_fallback<J, D, 16, WPT>), dim3(nb), wg, 0, st, b, list, count); break;
    case 12: hipLaunchKernelGGL((k_sad_fallback<J, D, 12, WPT>), dim3(nb), wg, 0, st, b, list, count); break;
    case 8: hipLaunchKernelGGL((k_sad_fallback<J, D, 8, WPT>), dim3(nb), wg, 0, st, b, list, count); break;
    default: hipLaunchKernelGGL((k_sad_fallback<J, D, 4, WPT>), dim3(nb), wg, 0, st, b, list, count); break;
  }
  return hipGetLastError();
}

/* a grid on the three-column tile (J = 3, K = 7: a plan of the per-job
   search only) back on four columns, as the split and queue kernels take it */
void four_columns(VSGGrid& g) {
  if (g.J != 3) return;
  g.J = 4;
  g.U = g.m * (((g.G + g.m - 1) / g.m + 3) / 4);
}

/* the split kernel with WPT waves per tile */
template <int WPT>
hipError_t launch_split(const VSGSearchArgs& a, const int* list, const unsigned* count, hipStream_t st) {
  if (a.njobs <= 0) return hipSuccess;
  if ((!list) != (!count) || a.g.S > 256 || a.g.S * a.g.bpp > 4 * 256) return hipErrorInvalidValue;
  VSGSearchArgs b = a;
  four_columns(b.g);
  b.g.K = 4;
  b.g.V = (b.g.G + 3) / 4;
  b.tiles = (b.g.U * b.g.V + 63) / 64;
  if (b.g.J == 4 && b.g.D == 1) return launch_fallback<4, 1, WPT>(b, list, count, st);
  if (b.g.J == 4 && b.g.D == 3) return launch_fallback<4, 3, WPT>(b, list, count, st);
  return launch_fallback<1, 1, WPT>(b, list, count, st);
}

constexpr int kSplitCus = 256;   /* MI355X compute units */

int gcd_int(int x, int y) {
  while (y) {
    int t = x % y;
    x = y;
    y = t;
  }
  return x;
}

}  // namespace

int vsg_plan_grid(VSGGrid* g, int S, int step, int maxShift, int K, int bpp) {
  g->S = S;
  g->step = step;
  g->bpp = bpp;
  g->n = maxShift / step;
  if (bpp == 1) {
    g->G = 2 * g->n + 1;
    g->alo = -g->n;
    g->org = 0;
  } else {
    g->G = 2 * maxShift / step + 1;
    g->alo = 0;
    g->org = -maxShift;
  }
  const int G = g->G;
  const int stepBytes = step * bpp;
  const int lcm4 = 4 * stepBytes / gcd_int(4, stepBytes);
  const int D = lcm4 / 4;
  g->K = (K == 6 || K == 7 || K == 8) ? K : 4;
  g->lds = 0;
  /* the widest column chunk (<= 28 words) that divides a field row: one prev
     row segment load then feeds J*K*CW v_sad_u8 */
  const int rowWords = S * bpp / 4;
  static const int kCW[] = {28, 20, 16, 12, 8, 4};
  g->CW = 4;
  for (int cw : kCW)
    if (rowWords % cw == 0) {
      g->CW = cw;
      break;
    }
  if (D == 1 || D == 3) {
    /* K = 7 is the three-column tile, for grids one word apart (D = 1) */
    g->J = g->K == 7 && D == 1 ? 3 : 4;
    g->D = D;
    g->m = lcm4 / stepBytes;
  } else {
    g->J = 1;   /* any step: one column per lane, no horizontal sharing */
    g->D = 1;
    g->m = 1;
  }
  if (g->K == 7 && g->J != 3) g->K = 4;
  const int perClass = (G + g->m - 1) / g->m;
  g->U = g->m * ((perClass + g->J - 1) / g->J);
  g->V = (G + g->K - 1) / g->K;
  /* a job that needs fewer than 256 lanes gets a workgroup of just enough
     whole waves, so idle lanes stay below one wave */
  const int lanes = g->U * g->V;
  g->wg = lanes >= 256 ? 256 : 64 * ((lanes + 63) / 64);
  /* the refinement gives each word of a field row its own thread */
  const int rowThreads = 64 * ((rowWords + 63) / 64);
  if (g->wg < rowThreads) g->wg = rowThreads;
  return (lanes + g->wg - 1) / g->wg;
}

double vsg_plan_cost(const VSGGrid& g, int tiles) {
  /* instructions one job issues: per lane and prev row, NP+1 loads, NP
     alignbytes and a few address / scalar-load slots; per lane the J*K*S*CW
     v_sad_u8 of each column chunk; times the lanes the job occupies */
  const int NP = g.CW + g.D * (g.J - 1);
  const int chunks = (g.S * g.bpp / 4) / g.CW;
  const double rows = g.S + (g.K - 1) * g.step;
  const double perLane = chunks * (rows * (2.0 * NP + 5.0) + (double)g.K * g.J * g.S * g.CW);
  return perLane * (double)tiles * g.wg;
}

hipError_t vsg_launch_search_any(const VSGSearchArgs& a, hipStream_t st) {
  /* a batch too small to give every CU two grid workgroups is latency-bound:
     the split kernel spreads each job's rows over four or eight waves.
     Eight halve each wave's rows, but (175 VGPRs: 2 waves a SIMD) only one
     such workgroup fits a CU: they are used while every unit finds a CU at
     once (a single frame's coarse jobs), four otherwise (its fine jobs).
     Sixteen (1,024 threads, at most 128 VGPRs) spill 712 bytes a lane and
     made a per-frame call slower (0.315 against 0.298 ms).
     VIDSTAB_AMD_SPLIT_WAVES=4 or 8 forces one. */
  static const int forced = [] {
    const char* e = getenv("VIDSTAB_AMD_SPLIT_WAVES");
    const int v = e ? atoi(e) : 0;
    return v == 4 || v == 8 ? v : 0;
  }();
  if (a.njobs > 0 && (long long)a.njobs * a.tiles < 2LL * kSplitCus) {
    VSGGrid g4 = a.g;
    four_columns(g4);
    const long long units = (long long)a.njobs * ((g4.U * ((g4.G + 3) / 4) + 63) / 64);
    const int wpt = forced ? forced : (units <= 2LL * kSplitCus ? 8 : 4);
    return wpt == 4 ? launch_split<4>(a, nullptr, nullptr, st) : launch_split<8>(a, nullptr, nullptr, st);
  }
  return vsg_launch_search(a, st);
}

hipError_t vsg_launch_search(const VSGSearchArgs& a, hipStream_t st) {
  if (a.njobs <= 0) return hipSuccess;
  if (a.chunk <= 0 || a.g.S > 256 || a.g.wg < 64 || a.g.wg > 256 || a.g.S * a.g.bpp > 4 * a.g.wg)
    return hipErrorInvalidValue;
  const VSGGrid& g = a.g;
  if (g.J == 3 && g.D == 1 && g.K == 7) return launch_grid_cw<3, 7, 1>(a, st);
  if (g.J == 4 && g.D == 1) return launch_grid_k<4, 1>(a, st);
  if (g.J == 4 && g.D == 3) return launch_grid_k<4, 3>(a, st);
  return launch_grid_k<1, 1>(a, st);
}

int vsg_queue_variant(const VSGGrid& g) {
  static const int kCWs[6] = {28, 20, 16, 12, 8, 4};
  const int jd = g.J == 4 && g.D == 1 ? 0 : (g.J == 4 && g.D == 3 ? 1 : (g.J == 1 && g.D == 1 ? 2 : -1));
  if (jd < 0 || g.bpp != 1) return -1;
  for (int i = 0; i < 6; ++i)
    if (g.CW == kCWs[i]) return jd * 6 + i;
  return -1;
}

VSGSearchArgs vsg_queue_args(const VSGSearchArgs& a, int wg, int K) {
  VSGSearchArgs b = a;
  four_columns(b.g);
  b.g.K = K;
  b.g.V = (b.g.G + K - 1) / K;
  b.g.wg = wg;
  b.g.lds = 0;
  b.tiles = (b.g.U * b.g.V + wg - 1) / wg;
  return b;
}

hipError_t vsg_launch_queue(const VSGMultiSeg* dSegs, int nseg, const VSGQueueTotals& totals, int wg, int K,
                            bool split, unsigned* queues, hipStream_t st) {
  int sum = 0, most = 0;
  if (totals.queues != 1 && totals.queues != VSG_XCDS) return hipErrorInvalidValue;
  for (int x = 0; x < totals.queues; ++x) {
    sum += totals.units[x];
    most = totals.units[x] > most ? totals.units[x] : most;
  }
  if (sum <= 0 || nseg <= 0) return hipSuccess;
  if (!dSegs || !queues || (wg != 64 && wg != 128 && wg != 64 * kQueueMaxWaves) || (K != 4 && K != 6 && K != 8) ||
      (split && K != 4))
    return hipErrorInvalidValue;
  /* every queue gets workgroups: at least 8, at most the resident count */
  const int resident = kQueueWaves / (wg / 64);
  int nb = sum < resident ? sum : resident;
  nb = nb < totals.queues ? totals.queues : nb;
  VSGQueueTotals t = totals;
  /* about 16 pulls per workgroup, at most 16 units a pull */
  const int perQueueWg = nb / totals.queues;
  int chunk = most / (perQueueWg * 16);
  t.chunk = chunk < 1 ? 1 : (chunk > 16 ? 16 : chunk);
  if (split)
    hipLaunchKernelGGL((k_sad_queue<4, true>), dim3(nb), dim3(wg), 0, st, dSegs, nseg, t, queues);
  else if (K == 4)
    hipLaunchKernelGGL((k_sad_queue<4, false>), dim3(nb), dim3(wg), 0, st, dSegs, nseg, t, queues);
  else if (K == 6)
    hipLaunchKernelGGL((k_sad_queue<6, false>), dim3(nb), dim3(wg), 0, st, dSegs, nseg, t, queues);
  else
    hipLaunchKernelGGL((k_sad_queue<8, false>), dim3(nb), dim3(wg), 0, st, dSegs, nseg, t, queues);
  return hipGetLastError();
}

hipError_t vsg_launch_window(const VSGSearchArgs& a, const float* pred, int npred, int slot0, uint32_t maxSad,
                             int* list, unsigned* stats, int steps, hipStream_t st) {
  if (a.njobs <= 0) return hipSuccess;
  if (steps < 1) return hipErrorInvalidValue;
  if (a.g.S > 256 || a.g.S < 16 || (a.g.S & 3) || npred <= 0 || slot0 < 0 || a.g.bpp != 1)
    return hipErrorInvalidValue;
  if (a.g.S <= 64)
    hipLaunchKernelGGL(k_sad_window<1>, dim3((a.njobs + 3) / 4), dim3(256), 0, st, a, pred, npred, slot0, maxSad,
                       list, stats, steps);
  else
    hipLaunchKernelGGL(k_sad_window<4>, dim3(a.njobs), dim3(256), 0, st, a, pred, npred, slot0, maxSad, list,
                       stats, steps);
  return hipGetLastError();
}

hipError_t vsg_launch_fallback(const VSGSearchArgs& a, const int* list, const unsigned* count, hipStream_t st) {
  return launch_split<4>(a, list, count, st);
}
