/*
 * vsg_kernels.hip -- gfx950 kernels of the detect pass that precede the
 * field searches (the searches are in vsg_search.hip).
 *
 *   k_blur_walk   separable integer box blur, row pass and column pass in
 *                 one kernel, a wave walking down a strip of rows with both
 *                 passes' running sums in registers (reference:
 *                 src/boxblur.c:190-332; lane body in vsg_blur.h)
 *   k_fields      per-field 8-bit min/max (src/motiondetect.c:389-407) and an
 *                 aligned copy of the field for the SAD kernels
 *   k_select      selectfields (src/motiondetect.c:676-743) for every frame
 *                 and both field grids, writing the coarse job lists
 *
 * No MFMA: none of this is a dense contraction.  All arithmetic is integer.
 */
#include "vsg_blur.h"
#include "vsg_device.h"
#include "vsg_internal.h"

namespace {

/* ---- box blur ----------------------------------------------------------------
   k_blur_walk: one wave per (plane, band, 1024-px strip), the lane body in
   vsg_blur.h.  Loads and stores are plain global accesses: gfx950 takes
   dword and 16-byte loads at any byte address, and the 16-byte stores land
   on 16-byte-aligned columns of 64-byte-pitched rows (columns past the
   width stay inside the row's pitch). */
typedef uint32_t __attribute__((aligned(1))) u32_unaligned;

typedef const __attribute__((address_space(1))) uint8_t* g8_ptr;
typedef const __attribute__((address_space(1))) u32_unaligned* g32_ptr;
typedef uint32_t u32x4_vec __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) u32x4_vec* g128_out;

/* global (not flat) loads and stores: the planes are device memory */
struct BlurDev {
  static __device__ __forceinline__ uint32_t ld8(const uint8_t* p) { return *(g8_ptr)p; }
  static __device__ __forceinline__ uint32_t ld32(const uint8_t* p) { return *(g32_ptr)p; }
  static __device__ __forceinline__ void ld128(const uint8_t* p, uint32_t (&m)[4]) {
    const g32_ptr q = (g32_ptr)p;
    m[0] = q[0];
    m[1] = q[1];
    m[2] = q[2];
    m[3] = q[3];
  }
  static __device__ __forceinline__ void st128(uint8_t* p, const uint32_t (&o)[4]) {
    u32x4_vec v;
    v.x = o[0];
    v.y = o[1];
    v.z = o[2];
    v.w = o[3];
    *(g128_out)p = v;
  }
  static __device__ __forceinline__ uint32_t perm(uint32_t a, uint32_t b, uint32_t sel) {
    return __builtin_amdgcn_perm(a, b, sel);
  }
  static __device__ __forceinline__ uint32_t mul24(uint32_t a, uint32_t b) { return __umul24(a, b); }
};

template <int S2>
__global__ __launch_bounds__(256) void k_blur_walk(const VSGBlurSrc* __restrict__ srcs, uint8_t* __restrict__ dst,
                                                   size_t slotStride, int dstPitch, int W, int H, uint32_t M,
                                                   VSGBlurPlan pl) {
  const int u = vsg_xcd_unit(blockIdx.x, pl.chunk, pl.units);
  if (u < 0) return;
  const int plane = u / pl.blocksPerPlane;
  const int unit = (u - plane * pl.blocksPerPlane) * 4 + (int)(threadIdx.x >> 6);
  if (unit >= pl.strips * pl.bands) return;
  const VSGBlurSrc src = srcs[plane];
  if (!src.p) return;   /* idle frame slot of a fixed-shape (graph) launch */
  const int band = unit / pl.strips, strip = unit - band * pl.strips;
  const int x = strip * kBlurStrip + 16 * (int)(threadIdx.x & 63);
  if (x >= W) return;
  const int y0 = band * pl.R, y1 = min(H, y0 + pl.R);
  uint8_t* out = dst + (size_t)plane * slotStride;
  if ((W & 15) == 0)
    VSGBlurLane<S2, BlurDev>::template walk<true>(src.p, src.pitch, out, dstPitch, W, H, x, y0, y1, M);
  else
    VSGBlurLane<S2, BlurDev>::template walk<false>(src.p, src.pitch, out, dstPitch, W, H, x, y0, y1, M);
}

/* fallback for boxes too wide for the LDS tile: one pass per kernel */
__global__ __launch_bounds__(256) void k_blur_rows(const VSGBlurSrc* __restrict__ srcs,
                                                   uint8_t* __restrict__ tmp, int tmpPitch,
                                                   int W, int H, int size, uint32_t mul) {
  const VSGBlurSrc src = srcs[blockIdx.z];
  int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= W || !src.p) return;
  int s2 = size >> 1;
  uint32_t acc = 0;
  for (int k = -s2; k <= s2; ++k) acc += src.p[(size_t)y * src.pitch + min(max(x + k, 0), W - 1)];
  tmp[(size_t)blockIdx.z * tmpPitch * H + (size_t)y * tmpPitch + x] = (uint8_t)__umulhi(acc, mul);
}

__global__ __launch_bounds__(256) void k_blur_cols(const VSGBlurSrc* __restrict__ srcs,
                                                   const uint8_t* __restrict__ tmp, int tmpPitch,
                                                   uint8_t* __restrict__ dst, size_t slotStride,
                                                   int dstPitch, int W, int H, int size, uint32_t mul) {
  int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= W || !srcs[blockIdx.z].p) return;
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
  if (x >= W || !src.p) return;
  dst[(size_t)blockIdx.z * slotStride + (size_t)y * dstPitch + x] = src.p[(size_t)y * src.pitch + x];
}

/* ---- field min/max + patch ------------------------------------------------ */

__device__ __forceinline__ uint32_t load_u32_unaligned(const uint8_t* p) {
  uintptr_t a = reinterpret_cast<uintptr_t>(p);
  const uint32_t* w = reinterpret_cast<const uint32_t*>(a & ~(uintptr_t)3);
  return __builtin_amdgcn_alignbyte(w[1], w[0], (uint32_t)(a & 3));
}

/* packed pixels: min/max of each colour channel.  Byte b of field-row word w
   is byte 4w + b of a row that starts on a pixel, channel (4w + b) % bpp */
__device__ void fields_packed(const VSGFieldArgs& a, int f, int slot, int lane) {
  const int S = a.fsz[f], bpp = a.bpp;
  const int words = (S * bpp) >> 2;
  const int left = (a.fx[f] - S / 2) * bpp, top = a.fy[f] - S / 2;
  const uint8_t* img = a.blurred + (size_t)slot * a.slotStride;
  uint32_t* patch = reinterpret_cast<uint32_t*>(a.patches + (size_t)slot * a.patchSlotStride + a.patchOff[f]);
  uint32_t lo0 = 255, lo1 = 255, lo2 = 255, hi0 = 0, hi1 = 0, hi2 = 0;
  for (int i = lane; i < S * words; i += 64) {
    const int y = i / words, w = i - y * words;
    const uint32_t v = load_u32_unaligned(img + (size_t)(top + y) * a.pitch + left + 4 * w);
    patch[i] = v;
    const int c0 = (4 * w) % bpp;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const uint32_t x = (v >> (8 * b)) & 0xff;
      int ch = c0 + b;
      ch = ch >= bpp ? ch - bpp : ch;
      ch = ch >= bpp ? ch - bpp : ch;
      if (ch == 0) {
        lo0 = min(lo0, x);
        hi0 = max(hi0, x);
      } else if (ch == 1) {
        lo1 = min(lo1, x);
        hi1 = max(hi1, x);
      } else if (ch == 2) {
        lo2 = min(lo2, x);
        hi2 = max(hi2, x);
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo0 = min(lo0, (uint32_t)__shfl_xor((int)lo0, off));
    hi0 = max(hi0, (uint32_t)__shfl_xor((int)hi0, off));
    lo1 = min(lo1, (uint32_t)__shfl_xor((int)lo1, off));
    hi1 = max(hi1, (uint32_t)__shfl_xor((int)hi1, off));
    lo2 = min(lo2, (uint32_t)__shfl_xor((int)lo2, off));
    hi2 = max(hi2, (uint32_t)__shfl_xor((int)hi2, off));
  }
  if (lane == 0) {
    uint8_t* mm = a.minmax + ((size_t)slot * a.nfields + f) * 6;
    mm[0] = (uint8_t)lo0;
    mm[1] = (uint8_t)hi0;
    mm[2] = (uint8_t)lo1;
    mm[3] = (uint8_t)hi1;
    mm[4] = (uint8_t)lo2;
    mm[5] = (uint8_t)hi2;
  }
}

/* host frames of a small batch: pinned staging -> raw planes, by the
   compute queue itself.  A copy-engine transfer would hand the stream from
   the DMA engine to the compute queue before the blur (tens of microseconds
   per batch: most of a single-frame call's upload).  Each thread keeps four
   16-byte loads in flight; n16 16-byte words, both ends 16-byte aligned. */
__global__ __launch_bounds__(256) void k_upload(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
  const size_t step = (size_t)gridDim.x * 1024;
  for (size_t i = (size_t)blockIdx.x * 1024 + threadIdx.x; i < n16; i += step) {
    uint4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (i + 256 * k < n16) v[k] = src[i + 256 * k];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (i + 256 * k < n16) dst[i + 256 * k] = v[k];
  }
}

/* one wave per (field, slot) */
__global__ __launch_bounds__(256) void k_fields(VSGFieldArgs a) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int f = blockIdx.x * 4 + wave;
  if (f >= a.nfields) return;
  const int slot = blockIdx.y;
  if (a.src && !a.src[slot].p) return;   /* idle frame slot */
  if (a.bpp != 1) {
    fields_packed(a, f, slot, lane);
    return;
  }
  /* the field in 16-byte chunks (field sizes are multiples of 16; patches
     start at multiples of S*S bytes), four chunks of a lane in flight at once */
  typedef uint4 __attribute__((aligned(1))) u128_any;
  const int S = a.fsz[f];
  const int cpr = S >> 4, total = S * cpr;
  const int left = a.fx[f] - S / 2, top = a.fy[f] - S / 2;
  const uint8_t* img = a.blurred + (size_t)slot * a.slotStride;
  uint4* patch = reinterpret_cast<uint4*>(a.patches + (size_t)slot * a.patchSlotStride + a.patchOff[f]);
  uint32_t lo = 255, hi = 0;
  for (int i0 = 0; i0 < total; i0 += 4 * 64) {
    uint4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + u * 64 + lane;
      const int ic = i < total ? i : 0;
      const int y = ic / cpr, c = ic - y * cpr;
      v[u] = *reinterpret_cast<const u128_any*>(img + (size_t)(top + y) * a.pitch + left + 16 * c);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + u * 64 + lane;
      if (i < total) {
        patch[i] = v[u];
        const uint32_t w4[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const uint32_t w = w4[k];
          const uint32_t b0 = w & 0xff, b1 = (w >> 8) & 0xff, b2 = (w >> 16) & 0xff, b3 = w >> 24;
          lo = min(lo, min(min(b0, b1), min(b2, b3)));
          hi = max(hi, max(max(b0, b1), max(b2, b3)));
        }
      }
    }
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
   device double division is involved.  A packed field's contrast is the mean
   of its three channels' (src/motiondetect.c:370-377), which the reference's
   -ffast-math build evaluates as ((c0 + c1) + c2) * (1/3); the same two
   IEEE additions and one multiplication by the rounded 1/3 run here. */
constexpr int kSelectThreads = 512;

__global__ __launch_bounds__(kSelectThreads) void k_select(VSGSelectArgs a) {
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ int segCnt[256];
  __shared__ int segBase[256];
  __shared__ int total, nleft, nunp;
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
      for (int q = threadIdx.x; q < nc; q += kSelectThreads) a.jobs[(size_t)f * nc + q].field = -1;
    return;
  }
  double* c = reinterpret_cast<double*>(lds);
  int* pick = reinterpret_cast<int*>(c + n);
  const int nmm = a.nmm;
  const uint8_t* mm = a.minmax + ((size_t)(a.slot0 + f) * a.nfTotal + st.mmOff) * nmm;
  for (int i = threadIdx.x; i < n; i += kSelectThreads) {
    const uint8_t* q = mm + (size_t)nmm * i;
    double v = a.ctab[q[0] * 256 + q[1]];
    if (nmm == 6) {
      const double sum = (v + a.ctab[q[2] * 256 + q[3]]) + a.ctab[q[4] * 256 + q[5]];
      v = __dmul_rn(sum, 1.0 / 3.0);
    }
    c[i] = v < st.ct ? 0.0 : v;
    pick[i] = 0;
  }
  const int numsegms = st.rows + 1;
  const int segmlen = n / numsegms + 1;
  const int perSeg = st.maxFields / numsegms;
  for (int s = threadIdx.x; s < numsegms; s += kSelectThreads) segCnt[s] = 0;
  if (threadIdx.x == 0) {
    nleft = 0;
    nunp = 0;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += kSelectThreads) {
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
  /* picks go to their slots; unpicked positive fields are gathered into a
     short list (the zeroed picks can never precede a positive contrast in the
     leftover order, so only this list needs ranking) */
  int* unp = pick + n;
  for (int i = threadIdx.x; i < n; i += kSelectThreads) {
    if (pick[i]) {
      const int o = segBase[i / segmlen] + pick[i] - 1;
      outIdx[o] = (int16_t)i;
      outC[o] = c[i];
    } else if (remaining > 0 && c[i] > 0.0) {
      unp[atomicAdd(&nunp, 1)] = i;
    }
  }
  __syncthreads();
  const int L = nunp;
  for (int k = threadIdx.x; k < L; k += kSelectThreads) {
    const int i = unp[k];
    const double ci = c[i];
    int rank = 0;
    for (int m = 0; m < L; ++m) {
      const int j = unp[m];
      rank += (c[j] > ci || (c[j] == ci && j < i)) ? 1 : 0;
    }
    if (rank < remaining) {
      outIdx[picked + rank] = (int16_t)i;
      outC[picked + rank] = ci;
      atomicAdd(&nleft, 1);
    }
  }
  __syncthreads();
  const int cnt = picked + nleft;
  if (threadIdx.x == 0) a.count[f * 2 + stage] = cnt;
  if (stage == 0) {
    for (int q = threadIdx.x; q < nc; q += kSelectThreads) {
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
  const size_t lds = (size_t)nmax * (sizeof(double) + 2 * sizeof(int));
  if (lds > 64 * 1024 || a.st[0].rows + 1 > 256 || a.st[1].rows + 1 > 256) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_select, dim3(nframes, 2), dim3(kSelectThreads), lds, st, a);
  return hipGetLastError();
}

/* ---- launchers ------------------------------------------------------------- */

bool vsg_blur_needs_tmp(int W, int size) { return size >= 2 && (size / 2 > kBlurMaxS2 || W < 16); }

template <int S2>
static void launch_walk(const VSGBlurSrc* d_src, int n, uint8_t* dst, size_t slotStride, int dstPitch, int W, int H,
                        int size, hipStream_t st) {
  const VSGBlurPlan pl = vsg_blur_plan(W, H, n);
  hipLaunchKernelGGL(k_blur_walk<S2>, dim3(vsg_xcd_blocks(pl.units, pl.chunk)), dim3(256), 0, st, d_src, dst,
                     slotStride, dstPitch, W, H, vsg_blur_mul(size), pl);
}

hipError_t vsg_launch_blur(const VSGBlurSrc* d_src, int n, uint8_t* dst, size_t slotStride,
                           int dstPitch, int W, int H, int size, uint8_t* tmp, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (size < 2) {
    hipLaunchKernelGGL(k_copy_plane, dim3((W + 255) / 256, H, n), dim3(256), 0, st, d_src, dst,
                       slotStride, dstPitch, W, H);
    return hipGetLastError();
  }
  if (!vsg_blur_needs_tmp(W, size)) {
    switch (size / 2) {
      case 1: launch_walk<1>(d_src, n, dst, slotStride, dstPitch, W, H, size, st); break;
      case 2: launch_walk<2>(d_src, n, dst, slotStride, dstPitch, W, H, size, st); break;
      case 3: launch_walk<3>(d_src, n, dst, slotStride, dstPitch, W, H, size, st); break;
      case 4: launch_walk<4>(d_src, n, dst, slotStride, dstPitch, W, H, size, st); break;
      case 5: launch_walk<5>(d_src, n, dst, slotStride, dstPitch, W, H, size, st); break;
      case 6: launch_walk<6>(d_src, n, dst, slotStride, dstPitch, W, H, size, st); break;
      case 7: launch_walk<7>(d_src, n, dst, slotStride, dstPitch, W, H, size, st); break;
      default: launch_walk<8>(d_src, n, dst, slotStride, dstPitch, W, H, size, st); break;
    }
    return hipGetLastError();
  }
  if (!tmp) return hipErrorInvalidValue;
  const uint32_t mul = vsg_div_magic(size);
  hipLaunchKernelGGL(k_blur_rows, dim3((W + 255) / 256, H, n), dim3(256), 0, st, d_src, tmp, dstPitch,
                     W, H, size, mul);
  hipLaunchKernelGGL(k_blur_cols, dim3((W + 255) / 256, H, n), dim3(256), 0, st, d_src, tmp, dstPitch, dst,
                     slotStride, dstPitch, W, H, size, mul);
  return hipGetLastError();
}

hipError_t vsg_launch_upload(const uint8_t* src, uint8_t* dst, size_t bytes, hipStream_t st) {
  if (bytes == 0) return hipSuccess;
  if ((bytes & 15) || ((uintptr_t)src & 15) || ((uintptr_t)dst & 15)) return hipErrorInvalidValue;
  const size_t n16 = bytes / 16, blocks = (n16 + 1023) / 1024;
  hipLaunchKernelGGL(k_upload, dim3(blocks < 2048 ? (unsigned)blocks : 2048u), dim3(256), 0, st,
                     reinterpret_cast<const uint4*>(src), reinterpret_cast<uint4*>(dst), n16);
  return hipGetLastError();
}

hipError_t vsg_launch_fields(const VSGFieldArgs& a, int nslots, hipStream_t st) {
  if (nslots <= 0 || a.nfields <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fields, dim3((a.nfields + 3) / 4, nslots), dim3(256), 0, st, a);
  return hipGetLastError();
}

