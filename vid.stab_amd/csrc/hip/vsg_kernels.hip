/*
 * vsg_kernels.hip -- gfx950 kernels of the detect pass.
 *
 *   k_blur        separable integer box blur, row pass and column pass through
 *                 LDS in one kernel (reference: src/boxblur.c:190-332)
 *   k_fields      per-field 8-bit min/max (src/motiondetect.c:389-407) and an
 *                 aligned copy of the field for the SAD kernels
 *   k_sad_grid    exhaustive grid search of the reference's spiral
 *                 (src/motiondetect.c:466-518, SAD :971-996): candidate shifts
 *                 spread over lanes, packed-u8 SAD (v_sad_u8) with the current
 *                 field as a wave-uniform operand, wave argmin on
 *                 (sad, spiral rank), one 64-bit atomic min per wave
 *   k_sad_refine  the dyadic refinement around the grid winner, replayed in
 *                 the reference's visit order (src/motiondetect.c:545-567)
 *
 * No MFMA: SAD is not a dense contraction.  All arithmetic is integer.
 */
#include "vsg_internal.h"

namespace {

constexpr int kBlurTW = 256;   /* output tile width (px) */
constexpr int kBlurTH = 32;    /* output tile height (rows) */

/* ---- box blur ----------------------------------------------------------------
   Output pixel = floor(sum over the clamped box / size), first along rows
   (8-bit result) and then along columns, which is what the reference's two
   running-sum passes compute. */
__global__ __launch_bounds__(256) void k_blur(const VSGBlurSrc* __restrict__ srcs,
                                              uint8_t* __restrict__ dst, size_t slotStride,
                                              int dstPitch, int W, int H, int size, uint32_t mul) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int s2 = size >> 1;
  const int x0 = blockIdx.x * kBlurTW, y0 = blockIdx.y * kBlurTH;
  const VSGBlurSrc src = srcs[blockIdx.z];
  uint8_t* out = dst + (size_t)blockIdx.z * slotStride;
  const int rowsIn = kBlurTH + 2 * s2;
  const int colsIn = kBlurTW + 2 * s2;
  const int inPitch = (colsIn + 3) & ~3;
  uint8_t* sIn = lds;
  uint8_t* sHor = lds + rowsIn * inPitch;

  /* stage the input tile with edge replication */
  for (int i = threadIdx.x; i < rowsIn * colsIn; i += 256) {
    int r = i / colsIn, c = i - r * colsIn;
    int gy = min(max(y0 - s2 + r, 0), H - 1);
    int gx = min(max(x0 - s2 + c, 0), W - 1);
    sIn[r * inPitch + c] = src.p[(size_t)gy * src.pitch + gx];
  }
  __syncthreads();

  /* row pass: 4 adjacent outputs per item with a sliding sum */
  for (int i = threadIdx.x; i < rowsIn * (kBlurTW / 4); i += 256) {
    int r = i / (kBlurTW / 4), g = i - r * (kBlurTW / 4);
    const uint8_t* row = sIn + r * inPitch + 4 * g;
    uint32_t acc = 0;
    for (int k = 0; k < size; ++k) acc += row[k];
    uint32_t o = __umulhi(acc, mul);
    acc += (uint32_t)row[size] - row[0];
    o |= __umulhi(acc, mul) << 8;
    acc += (uint32_t)row[size + 1] - row[1];
    o |= __umulhi(acc, mul) << 16;
    acc += (uint32_t)row[size + 2] - row[2];
    o |= __umulhi(acc, mul) << 24;
    *reinterpret_cast<uint32_t*>(sHor + r * kBlurTW + 4 * g) = o;
  }
  __syncthreads();

  /* column pass: 4 columns x (kBlurTH/4) rows per thread */
  const int g = threadIdx.x & 63, band = threadIdx.x >> 6;
  const int ox = x0 + 4 * g;
  if (ox >= W) return;
  const int r0 = band * (kBlurTH / 4);
  uint32_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  for (int k = 0; k < size; ++k) {
    uint32_t w = *reinterpret_cast<const uint32_t*>(sHor + (r0 + k) * kBlurTW + 4 * g);
    a0 += w & 0xff; a1 += (w >> 8) & 0xff; a2 += (w >> 16) & 0xff; a3 += w >> 24;
  }
  for (int rr = 0; rr < kBlurTH / 4; ++rr) {
    if (rr > 0) {
      uint32_t wn = *reinterpret_cast<const uint32_t*>(sHor + (r0 + rr + size - 1) * kBlurTW + 4 * g);
      uint32_t wo = *reinterpret_cast<const uint32_t*>(sHor + (r0 + rr - 1) * kBlurTW + 4 * g);
      a0 += (wn & 0xff) - (wo & 0xff);
      a1 += ((wn >> 8) & 0xff) - ((wo >> 8) & 0xff);
      a2 += ((wn >> 16) & 0xff) - ((wo >> 16) & 0xff);
      a3 += (wn >> 24) - (wo >> 24);
    }
    int oy = y0 + r0 + rr;
    if (oy < H) {
      uint32_t o = __umulhi(a0, mul) | (__umulhi(a1, mul) << 8) | (__umulhi(a2, mul) << 16) |
                   (__umulhi(a3, mul) << 24);
      /* the destination pitch is a multiple of 64, so the dword never leaves the row */
      *reinterpret_cast<uint32_t*>(out + (size_t)oy * dstPitch + ox) = o;
    }
  }
}

/* fallback for boxes too wide for the LDS tile: one pass per kernel */
__global__ __launch_bounds__(256) void k_blur_rows(const VSGBlurSrc* __restrict__ srcs,
                                                   uint8_t* __restrict__ tmp, int tmpPitch,
                                                   int W, int H, int size, uint32_t mul) {
  const VSGBlurSrc src = srcs[blockIdx.z];
  int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= W) return;
  int s2 = size >> 1;
  uint32_t acc = 0;
  for (int k = -s2; k <= s2; ++k) acc += src.p[(size_t)y * src.pitch + min(max(x + k, 0), W - 1)];
  tmp[(size_t)blockIdx.z * tmpPitch * H + (size_t)y * tmpPitch + x] = (uint8_t)__umulhi(acc, mul);
}

__global__ __launch_bounds__(256) void k_blur_cols(const uint8_t* __restrict__ tmp, int tmpPitch,
                                                   uint8_t* __restrict__ dst, size_t slotStride,
                                                   int dstPitch, int W, int H, int size, uint32_t mul) {
  int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= W) return;
  int s2 = size >> 1;
  const uint8_t* t = tmp + (size_t)blockIdx.z * tmpPitch * H;
  uint32_t acc = 0;
  for (int k = -s2; k <= s2; ++k) acc += t[(size_t)min(max(y + k, 0), H - 1) * tmpPitch + x];
  dst[(size_t)blockIdx.z * slotStride + (size_t)y * dstPitch + x] = (uint8_t)__umulhi(acc, mul);
}

/* stepSize < 2: the reference copies the plane unchanged */
__global__ __launch_bounds__(256) void k_copy_plane(const VSGBlurSrc* __restrict__ srcs,
                                                    uint8_t* __restrict__ dst, size_t slotStride,
                                                    int dstPitch, int W, int H) {
  const VSGBlurSrc src = srcs[blockIdx.z];
  int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= W) return;
  dst[(size_t)blockIdx.z * slotStride + (size_t)y * dstPitch + x] = src.p[(size_t)y * src.pitch + x];
}

/* ---- field min/max + patch ------------------------------------------------ */

__device__ __forceinline__ uint32_t load_u32_unaligned(const uint8_t* p) {
  uintptr_t a = reinterpret_cast<uintptr_t>(p);
  const uint32_t* w = reinterpret_cast<const uint32_t*>(a & ~(uintptr_t)3);
  return __builtin_amdgcn_alignbyte(w[1], w[0], (uint32_t)(a & 3));
}

/* one wave per (field, slot) */
__global__ __launch_bounds__(256) void k_fields(VSGFieldArgs a) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int f = blockIdx.x * 4 + wave;
  if (f >= a.nfields) return;
  const int slot = blockIdx.y;
  const int S = a.fsz[f];
  const int words = S >> 2;
  const int left = a.fx[f] - S / 2, top = a.fy[f] - S / 2;
  const uint8_t* img = a.blurred + (size_t)slot * a.slotStride;
  uint32_t* patch = reinterpret_cast<uint32_t*>(a.patches + (size_t)slot * a.patchSlotStride + a.patchOff[f]);
  uint32_t lo = 255, hi = 0;
  for (int i = lane; i < S * words; i += 64) {
    int y = i / words, w = i - y * words;
    uint32_t v = load_u32_unaligned(img + (size_t)(top + y) * a.pitch + left + 4 * w);
    patch[i] = v;
    uint32_t b0 = v & 0xff, b1 = (v >> 8) & 0xff, b2 = (v >> 16) & 0xff, b3 = v >> 24;
    lo = min(lo, min(min(b0, b1), min(b2, b3)));
    hi = max(hi, max(max(b0, b1), max(b2, b3)));
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo = min(lo, (uint32_t)__shfl_xor((int)lo, off));
    hi = max(hi, (uint32_t)__shfl_xor((int)hi, off));
  }
  if (lane == 0) {
    uint8_t* mm = a.minmax + ((size_t)slot * a.nfields + f) * 2;
    mm[0] = (uint8_t)lo;
    mm[1] = (uint8_t)hi;
  }
}

/* ---- SAD grid search ---------------------------------------------------------
   Lane (u, v) of a job owns the candidates a = a0 + m*j (j < J) and
   b = b0 + k (k < K) of the (2n+1)^2 grid, a0 = -n + (u % m) + m*J*(u / m),
   b0 = -n + K*v.  m*step is a multiple of 4, so the J candidates of a lane
   sit D = m*step/4 words apart in the same prev row: one load of a prev row
   segment feeds J*K candidates, each prev word is shifted into the field's
   alignment once (v_alignbyte), and the current field's words come in as
   wave-uniform scalar loads (SGPR operand of v_sad_u8).  Iterating over prev
   rows p, candidate k pairs prev row p with field row p - k*step. */

typedef const __attribute__((address_space(4))) uint32_t* cword_ptr;   /* scalar loads */
typedef const __attribute__((address_space(1))) uint32_t* gword_ptr;   /* global loads */

template <int J, int K, int D, int CW>
__global__ __launch_bounds__(256) void k_sad_grid(VSGSearchArgs a) {
  const int job = blockIdx.y;
  const VSGJob jb = a.jobs[job];
  if (jb.field < 0) return;   /* unselected slot of a fixed-size job list */
  const int S = a.g.S, s = a.g.step, n = a.g.n, m = a.g.m, U = a.g.U, V = a.g.V;
  const int lid = blockIdx.x * blockDim.x + threadIdx.x;
  const int u = lid % U;
  int v = lid / U;
  const bool laneOk = v < V;
  if (!laneOk) v = V - 1;
  const int a0 = -n + (u % m) + m * J * (u / m);
  const int b0 = -n + K * v;
  const int fxc = a.fx[jb.field], fyc = a.fy[jb.field];
  const uint8_t* prev = jb.prev < 0 ? a.ref : a.blurred + (size_t)jb.prev * a.slotStride;
  cword_ptr patch = (cword_ptr)(a.patches + (size_t)jb.slot * a.patchSlotStride + a.stageOff +
                                (size_t)jb.field * S * S);
  const int prow0 = fyc - S / 2 + b0 * s + jb.oy;
  const int pcol0 = fxc - S / 2 + a0 * s + jb.ox;
  const int wordsPerRow = S >> 2;
  const int pitchWords = a.pitch >> 2;

  uint32_t acc[K][J];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int j = 0; j < J; ++j) acc[k][j] = 0;

  constexpr int NP = CW + D * (J - 1);   /* shifted prev words per row segment */
  const int nchunks = wordsPerRow / CW;
  const int prows = S + (K - 1) * s;
  for (int c = 0; c < nchunks; ++c) {
    const uintptr_t start = reinterpret_cast<uintptr_t>(prev + (size_t)prow0 * a.pitch + pcol0 + 4 * CW * c);
    gword_ptr wp = (gword_ptr)(start & ~(uintptr_t)3);
    const uint32_t sh = (uint32_t)(start & 3);
    cword_ptr cbase = patch + CW * c;
    for (int p = 0; p < prows; ++p) {
      uint32_t raw[NP + 1];
#pragma unroll
      for (int q = 0; q <= NP; ++q) raw[q] = wp[q];
      uint32_t P[NP];
#pragma unroll
      for (int q = 0; q < NP; ++q) P[q] = __builtin_amdgcn_alignbyte(raw[q + 1], raw[q], sh);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int y = p - k * s;
        if (y >= 0 && y < S) {
          cword_ptr crow = cbase + y * wordsPerRow;
#pragma unroll
          for (int w = 0; w < CW; ++w) {
            const uint32_t cw = crow[w];
#pragma unroll
            for (int j = 0; j < J; ++j) acc[k][j] = __builtin_amdgcn_sad_u8(P[w + D * j], cw, acc[k][j]);
          }
        }
      }
      wp += pitchWords;
    }
  }

  unsigned long long best = ~0ull;
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int ga = a0 + m * j, gb = b0 + k;
      if (laneOk && ga <= n && gb <= n) {
        unsigned long long key = ((unsigned long long)acc[k][j] << 32) | (uint32_t)vsg_rank(ga, gb);
        best = key < best ? key : best;
      }
    }
  for (int off = 32; off > 0; off >>= 1) {
    uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)best, off);
    uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(best >> 32), off);
    unsigned long long o = ((unsigned long long)hi << 32) | lo;
    best = o < best ? o : best;
  }
  if ((threadIdx.x & 63) == 0 && best != ~0ull) atomicMin(&a.best[job], best);
}

/* ---- refinement --------------------------------------------------------------
   One workgroup per job.  A level's candidates form an ni x ni grid (spacing
   ns around the current best); an item is one (candidate row, field row)
   pair, whose prev row serves all ni candidates of that row: 16-byte loads of
   the prev row at each candidate's byte offset (gfx950 global loads accept any
   byte alignment) against 16-byte loads of the aligned field patch, four
   v_sad_u8 per load.  Row sums meet in LDS; then one lane scans the
   candidates in the reference's order (i outer, j inner, centre skipped) with
   the strict "<" against the running minimum, which is exactly the
   sequential loop. */
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef u32x4 __attribute__((aligned(1))) u32x4_any;

__global__ __launch_bounds__(256) void k_sad_refine(VSGSearchArgs a) {
  constexpr int kMaxNi = 5;   /* ni = 2r/ns + 1 <= 5 for every step */
  __shared__ uint32_t sads[kMaxNi * kMaxNi];
  __shared__ int shTx, shTy;
  __shared__ uint32_t shMin;
  const int job = blockIdx.x;
  const VSGJob jb = a.jobs[job];
  if (jb.field < 0) return;
  const int S = a.g.S, w16 = S >> 4;
  const int fxc = a.fx[jb.field], fyc = a.fy[jb.field];
  const uint8_t* prev = jb.prev < 0 ? a.ref : a.blurred + (size_t)jb.prev * a.slotStride;
  const u32x4* patch = reinterpret_cast<const u32x4*>(a.patches + (size_t)jb.slot * a.patchSlotStride +
                                                      a.stageOff + (size_t)jb.field * S * S);
  if (threadIdx.x == 0) {
    unsigned long long key = a.best[job];
    int ga, gb;
    vsg_unrank((int)(uint32_t)key, &ga, &gb);
    shTx = ga * a.g.step;
    shTy = gb * a.g.step;
    shMin = (uint32_t)(key >> 32);
  }
  __syncthreads();
  int tx = shTx, ty = shTy;
  uint32_t minerr = shMin;
  int st = a.g.step;
  while (st > 1) {
    const int ns = st / 2, r = st - ns;
    const int ni = (2 * r) / ns + 1;
    if (threadIdx.x < kMaxNi * kMaxNi) sads[threadIdx.x] = 0;
    __syncthreads();
    const int col0 = fxc - S / 2 + tx - r + jb.ox;
    for (int item = threadIdx.x; item < ni * S; item += 256) {
      const int tj = item / S, y = item - tj * S;
      const int j = ty - r + ns * tj;
      const uint8_t* prow = prev + (size_t)(fyc - S / 2 + y + j + jb.oy) * a.pitch + col0;
      const u32x4* crow = patch + y * w16;
      uint32_t acc[kMaxNi] = {0, 0, 0, 0, 0};
      for (int w = 0; w < w16; ++w) {
        const u32x4 c = crow[w];
#pragma unroll
        for (int t = 0; t < kMaxNi; ++t) {
          if (t < ni) {
            const u32x4 p = *reinterpret_cast<const u32x4_any*>(prow + ns * t + 16 * w);
            uint32_t s = acc[t];
            s = __builtin_amdgcn_sad_u8(p.x, c.x, s);
            s = __builtin_amdgcn_sad_u8(p.y, c.y, s);
            s = __builtin_amdgcn_sad_u8(p.z, c.z, s);
            acc[t] = __builtin_amdgcn_sad_u8(p.w, c.w, s);
          }
        }
      }
#pragma unroll
      for (int t = 0; t < kMaxNi; ++t)
        if (t < ni) atomicAdd(&sads[t * ni + tj], acc[t]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int btx = tx, bty = ty;
      for (int ti = 0; ti < ni; ++ti)
        for (int tj = 0; tj < ni; ++tj) {
          const int i = tx - r + ns * ti, j = ty - r + ns * tj;
          if (i == tx && j == ty) continue;
          const uint32_t e = sads[ti * ni + tj];
          if (e < minerr) { minerr = e; btx = i; bty = j; }
        }
      shTx = btx;
      shTy = bty;
      shMin = minerr;
    }
    __syncthreads();
    tx = shTx;
    ty = shTy;
    minerr = shMin;
    st /= 2;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    VSGResult res;
    res.tx = (int16_t)tx;
    res.ty = (int16_t)ty;
    res.minerror = minerr;
    a.results[job] = res;
    a.best[job] = ~0ull;   /* ready for the next launch */
  }
}

template <int J, int K, int D, int CW>
hipError_t launch_grid(const VSGSearchArgs& a, hipStream_t st) {
  hipLaunchKernelGGL((k_sad_grid<J, K, D, CW>), dim3(a.tiles, a.njobs), dim3(a.g.wg), 0, st, a);
  return hipGetLastError();
}

/* ---- field selection -----------------------------------------------------------
   selectfields (src/motiondetect.c:676-743) for one (frame, stage) per workgroup.
   The reference sorts with glibc's stable qsort, so equal contrasts keep index
   order; that order is the total order (contrast desc, index asc), and each
   element's place in it is computed directly as the number of elements that
   precede it.  Segment picks: the first maxFields/(rows+1) positive entries of
   each index segment, in segment order.  Leftovers: the stable re-sort of the
   whole array with the picks zeroed keeps equal contrasts in index order
   again, so the same rule applies.  Contrast doubles come from a host-built
   table of (max-min)/(max+min+0.1) over all 65536 (min, max) pairs, so no
   device double division is involved. */
__global__ __launch_bounds__(256) void k_select(VSGSelectArgs a) {
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ int segCnt[256];
  __shared__ int segBase[256];
  __shared__ int total, nleft;
  const int f = blockIdx.x, stage = blockIdx.y;
  const VSGSelectStage st = a.st[stage];
  const int n = st.nf;
  const size_t rowOff = (size_t)f * a.nfTotal + (stage ? a.st[0].nf : 0);
  int16_t* outIdx = a.selIdx + rowOff;
  double* outC = a.selC + rowOff;
  const int nc = a.st[0].nf;
  if (!a.detect[f]) {
    if (threadIdx.x == 0) a.count[f * 2 + stage] = 0;
    if (stage == 0)
      for (int q = threadIdx.x; q < nc; q += 256) a.jobs[(size_t)f * nc + q].field = -1;
    return;
  }
  double* c = reinterpret_cast<double*>(lds);
  int* pick = reinterpret_cast<int*>(c + n);
  const uint8_t* mm = a.minmax + ((size_t)(a.slot0 + f) * a.nfTotal + st.mmOff) * 2;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double v = a.ctab[mm[2 * i] * 256 + mm[2 * i + 1]];
    c[i] = v < st.ct ? 0.0 : v;
    pick[i] = 0;
  }
  const int numsegms = st.rows + 1;
  const int segmlen = n / numsegms + 1;
  const int perSeg = st.maxFields / numsegms;
  for (int s = threadIdx.x; s < numsegms; s += 256) segCnt[s] = 0;
  if (threadIdx.x == 0) nleft = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 256) {
    const double ci = c[i];
    if (!(ci > 0.0)) continue;
    const int s = i / segmlen, start = s * segmlen, end = min(start + segmlen, n);
    int pos = 0;
    for (int j = start; j < end; ++j) pos += (c[j] > ci || (c[j] == ci && j < i)) ? 1 : 0;
    if (pos < perSeg) {
      pick[i] = pos + 1;
      atomicAdd(&segCnt[s], 1);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int s = 0; s < numsegms; ++s) {
      segBase[s] = acc;
      acc += segCnt[s];
    }
    total = acc;
  }
  __syncthreads();
  const int picked = total;
  const int remaining = st.maxFields - picked;
  for (int i = threadIdx.x; i < n; i += 256) {
    if (pick[i]) {
      const int o = segBase[i / segmlen] + pick[i] - 1;
      outIdx[o] = (int16_t)i;
      outC[o] = c[i];
    } else if (remaining > 0 && c[i] > 0.0) {
      const double ci = c[i];
      int rank = 0;
      for (int j = 0; j < n && rank < remaining; ++j) {
        const double cj = pick[j] ? 0.0 : c[j];
        rank += (cj > ci || (cj == ci && j < i)) ? 1 : 0;
      }
      if (rank < remaining) {
        outIdx[picked + rank] = (int16_t)i;
        outC[picked + rank] = ci;
        atomicAdd(&nleft, 1);
      }
    }
  }
  __syncthreads();
  const int cnt = picked + nleft;
  if (threadIdx.x == 0) a.count[f * 2 + stage] = cnt;
  if (stage == 0) {
    for (int q = threadIdx.x; q < nc; q += 256) {
      VSGJob j;
      j.slot = a.slot0 + f;
      j.prev = a.prevSlot[f];
      j.field = q < cnt ? outIdx[q] : -1;
      j.ox = 0;
      j.oy = 0;
      a.jobs[(size_t)f * nc + q] = j;
    }
  }
}

}  // namespace

hipError_t vsg_launch_select(const VSGSelectArgs& a, int nframes, hipStream_t st) {
  if (nframes <= 0) return hipSuccess;
  const int nmax = a.st[0].nf > a.st[1].nf ? a.st[0].nf : a.st[1].nf;
  const size_t lds = (size_t)nmax * (sizeof(double) + sizeof(int));
  if (lds > 64 * 1024 || a.st[0].rows + 1 > 256 || a.st[1].rows + 1 > 256) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_select, dim3(nframes, 2), dim3(256), lds, st, a);
  return hipGetLastError();
}

/* ---- launchers ------------------------------------------------------------- */

hipError_t vsg_launch_blur(const VSGBlurSrc* d_src, int n, uint8_t* dst, size_t slotStride,
                           int dstPitch, int W, int H, int size, uint8_t* tmp, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (size < 2) {
    hipLaunchKernelGGL(k_copy_plane, dim3((W + 255) / 256, H, n), dim3(256), 0, st, d_src, dst,
                       slotStride, dstPitch, W, H);
    return hipGetLastError();
  }
  const uint32_t mul = vsg_div_magic(size);
  const int s2 = size / 2;
  const size_t lds = (size_t)(kBlurTH + 2 * s2) * (((kBlurTW + 2 * s2) + 3) & ~3) +
                     (size_t)(kBlurTH + 2 * s2) * kBlurTW;
  if (lds <= 64 * 1024) {
    hipLaunchKernelGGL(k_blur, dim3((W + kBlurTW - 1) / kBlurTW, (H + kBlurTH - 1) / kBlurTH, n),
                       dim3(256), lds, st, d_src, dst, slotStride, dstPitch, W, H, size, mul);
    return hipGetLastError();
  }
  if (!tmp) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_blur_rows, dim3((W + 255) / 256, H, n), dim3(256), 0, st, d_src, tmp, dstPitch,
                     W, H, size, mul);
  hipLaunchKernelGGL(k_blur_cols, dim3((W + 255) / 256, H, n), dim3(256), 0, st, tmp, dstPitch, dst,
                     slotStride, dstPitch, W, H, size, mul);
  return hipGetLastError();
}

hipError_t vsg_launch_fields(const VSGFieldArgs& a, int nslots, hipStream_t st) {
  if (nslots <= 0 || a.nfields <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fields, dim3((a.nfields + 3) / 4, nslots), dim3(256), 0, st, a);
  return hipGetLastError();
}

static int gcd_int(int x, int y) {
  while (y) { int t = x % y; x = y; y = t; }
  return x;
}

int vsg_plan_grid(VSGGrid* g, int S, int step, int maxShift) {
  g->S = S;
  g->step = step;
  g->n = maxShift / step;
  const int G = 2 * g->n + 1;
  const int lcm4 = 4 * step / gcd_int(4, step);
  const int D = lcm4 / 4;
  g->K = 4;
  /* the widest column chunk (<= 28 words) that divides a field row: one prev
     row segment load then feeds J*K*CW v_sad_u8 */
  static const int kCW[] = {28, 20, 16, 12, 8, 4};
  g->CW = 4;
  for (int cw : kCW)
    if ((S / 4) % cw == 0) {
      g->CW = cw;
      break;
    }
  if (D == 1 || D == 3) {
    g->J = 4;
    g->D = D;
    g->m = lcm4 / step;
  } else {
    g->J = 1;   /* any step: one column per lane, no horizontal sharing */
    g->D = 1;
    g->m = 1;
  }
  const int perClass = (G + g->m - 1) / g->m;
  g->U = g->m * ((perClass + g->J - 1) / g->J);
  g->V = (G + g->K - 1) / g->K;
  /* a job that needs fewer than 256 lanes gets a workgroup of just enough
     whole waves, so idle lanes stay below one wave */
  const int lanes = g->U * g->V;
  g->wg = lanes >= 256 ? 256 : 64 * ((lanes + 63) / 64);
  return (lanes + g->wg - 1) / g->wg;
}

template <int J, int D>
static hipError_t launch_grid_cw(const VSGSearchArgs& a, hipStream_t st) {
  switch (a.g.CW) {
    case 28: return launch_grid<J, 4, D, 28>(a, st);
    case 20: return launch_grid<J, 4, D, 20>(a, st);
    case 16: return launch_grid<J, 4, D, 16>(a, st);
    case 12: return launch_grid<J, 4, D, 12>(a, st);
    case 8: return launch_grid<J, 4, D, 8>(a, st);
    default: return launch_grid<J, 4, D, 4>(a, st);
  }
}

hipError_t vsg_launch_grid(const VSGSearchArgs& a, hipStream_t st) {
  if (a.njobs <= 0) return hipSuccess;
  const VSGGrid& g = a.g;
  if (g.J == 4 && g.D == 1) return launch_grid_cw<4, 1>(a, st);
  if (g.J == 4 && g.D == 3) return launch_grid_cw<4, 3>(a, st);
  return launch_grid_cw<1, 1>(a, st);
}

hipError_t vsg_launch_refine(const VSGSearchArgs& a, hipStream_t st) {
  if (a.njobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sad_refine, dim3(a.njobs), dim3(256), 0, st, a);
  return hipGetLastError();
}
