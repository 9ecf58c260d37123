/*
 * vsg_cnn.hip -- the proposal network of the learned detect form: a stack
 * of 3x3 convolutions, 32 channels, BF16 activations and weights with FP32
 * accumulation on the CDNA matrix cores (v_mfma_f32_32x32x8_bf16), and the
 * kernels that train it on the device (data and weight gradients, field
 * pooling and its loss, Adam).  Declared in vsg_cnn.h.
 *
 * Layouts.  Activations are NHWC with C = 32 in BF16 (raw 16-bit patterns):
 * index ((n*H + y)*W + x)*32 + c.  Narrower layers (the 2-channel input, the
 * 2-channel output) are carried as 32 channels with the rest zero, so one
 * kernel shape serves every layer.  Master weights are FP32 [tap][ci][co]
 * with tap = (dy+1)*3 + (dx+1); the convolution reads them pre-packed as the
 * MFMA B operand: [tap][kstep][lane][4], lane l holding rows
 * k = 4*(l/32) + 0..3 of column n = l%32 for input channels kstep*8 + k.
 *
 * The convolution is an implicit GEMM, M = 32 pixels of one image row,
 * N = 32 output channels, K = 9 taps x 32 channels: each wave keeps the 36
 * B fragments of a layer in registers and walks row tiles; the A fragment of
 * a lane is 4 consecutive channels of one (shifted) pixel, one 8-byte load,
 * zero outside the image (padding = dilation).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>

#include "vsg_cnn.h"

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

/* round to nearest even (finite values) */
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

constexpr int kC = 32;

/* ---- weight packing -------------------------------------------------------------
   forward: B[k][n] = W[tap][ks*8+k][n]; data gradient (transposed, flipped):
   B[k][n] = W[8-tap][n][ks*8+k] */
__global__ void k_pack(const float* __restrict__ w, uint16_t* __restrict__ pk, int transposed) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;   /* over 9*4*64*4 */
  if (i >= 9 * 4 * 64 * 4) return;
  const int e = i & 3, lane = (i >> 2) & 63, ks = (i >> 8) & 3, tap = i >> 10;
  const int k = ks * 8 + 4 * (lane >> 5) + e, n = lane & 31;
  const float v = transposed ? w[((8 - tap) * kC + n) * kC + k] : w[(tap * kC + k) * kC + n];
  pk[i] = f2bf(v);
}

/* ---- convolution ------------------------------------------------------------------
   out = act(conv(in, W) + bias); mode: 0 bf16 + ReLU, 1 bf16, 2 fp32.
   KS input-channel steps of 8: 4 for a 32-channel input, 1 when only
   channels 0..7 can be non-zero (the network's 2-channel input), which
   skips MFMAs whose A operand is all zero.  All A fragments of a tile are
   loaded before its MFMA chain, so the loads of a tile overlap each other
   instead of each waiting in front of its MFMA. */
template <int KS>
__global__ __launch_bounds__(256) void k_conv(const uint16_t* __restrict__ in, int N, int H, int W,
                                              const uint16_t* __restrict__ pk, const float* __restrict__ bias,
                                              int dil, int mode, uint16_t* __restrict__ outb,
                                              float* __restrict__ outf) {
  const int lane = threadIdx.x & 63;
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nw = (gridDim.x * blockDim.x) >> 6;
  v4s b[9 * KS];
#pragma unroll
  for (int tap = 0; tap < 9; ++tap)
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
      b[tap * KS + ks] = *reinterpret_cast<const v4s*>(pk + ((size_t)(tap * 4 + ks) * 64 + lane) * 4);
  const int co = lane & 31, kq = lane >> 5;
  const float bco = bias ? bias[co] : 0.0f;
  const int tilesX = (W + 31) >> 5;
  const int ntiles = N * H * tilesX;
  for (int t = gw; t < ntiles; t += nw) {
    const int xt = t % tilesX, r = t / tilesX, y = r % H, n = r / H;
    const int x = xt * 32 + (lane & 31);
    v4s a[9 * KS];
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int yy = y + (tap / 3 - 1) * dil, xx = x + (tap % 3 - 1) * dil;
      const bool ok = yy >= 0 && yy < H && xx >= 0 && xx < W;
      const uint16_t* p = in + (((size_t)n * H + (ok ? yy : 0)) * W + (ok ? xx : 0)) * kC + 4 * kq;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        v4s v = {0, 0, 0, 0};
        if (ok) v = *reinterpret_cast<const v4s*>(p + ks * 8);
        a[tap * KS + ks] = v;
      }
    }
    v16f acc = {};
#pragma unroll
    for (int q = 0; q < 9 * KS; ++q) acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a[q], b[q], acc, 0, 0, 0);
    const size_t rowBase = ((size_t)n * H + y) * W;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int m = 8 * (v >> 2) + 4 * kq + (v & 3);
      const int px = xt * 32 + m;
      if (px < W) {
        float val = acc[v] + bco;
        const size_t o = (rowBase + px) * kC + co;
        if (mode == 2) {
          outf[o] = val;
        } else {
          if (mode == 0) val = val > 0.0f ? val : 0.0f;
          outb[o] = f2bf(val);
        }
      }
    }
  }
}

/* ---- convolution on the gfx950 K = 16 matrix instruction -----------------------------
   v_mfma_f32_32x32x16_bf16: lane l holds 8 consecutive K values (rows
   8*(l/32) + 0..7) of A row / B column l%32, so an A fragment is one 16-byte
   load of 8 channels of one pixel; a tile takes 9 x KS MFMAs (KS = 2 for 32
   input channels, 1 for at most 16).  B is packed [tap][ks][lane][8]. */
__global__ void k_pack16(const float* __restrict__ w, uint16_t* __restrict__ pk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;   /* over 9*2*64*8 */
  if (i >= 9 * 2 * 64 * 8) return;
  const int e = i & 7, lane = (i >> 3) & 63, ks = (i >> 9) & 1, tap = i >> 10;
  const int k = ks * 16 + 8 * (lane >> 5) + e, n = lane & 31;
  pk[i] = f2bf(w[(tap * kC + k) * kC + n]);
}

template <int KS>
__global__ __launch_bounds__(256) void k_conv16(const uint16_t* __restrict__ in, int N, int H, int W,
                                                const uint16_t* __restrict__ pk, const float* __restrict__ bias,
                                                int dil, int mode, uint16_t* __restrict__ outb,
                                                float* __restrict__ outf) {
  const int lane = threadIdx.x & 63;
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nw = (gridDim.x * blockDim.x) >> 6;
  v8bf b[9 * KS];
#pragma unroll
  for (int tap = 0; tap < 9; ++tap)
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
      b[tap * KS + ks] = *reinterpret_cast<const v8bf*>(pk + ((size_t)(tap * 2 + ks) * 64 + lane) * 8);
  const int co = lane & 31, kq = lane >> 5;
  const float bco = bias ? bias[co] : 0.0f;
  const int tilesX = (W + 31) >> 5;
  const int ntiles = N * H * tilesX;
  for (int t = gw; t < ntiles; t += nw) {
    const int xt = t % tilesX, r = t / tilesX, y = r % H, n = r / H;
    const int x = xt * 32 + (lane & 31);
    v8bf a[9 * KS];
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int yy = y + (tap / 3 - 1) * dil, xx = x + (tap % 3 - 1) * dil;
      const bool ok = yy >= 0 && yy < H && xx >= 0 && xx < W;
      const uint16_t* p = in + (((size_t)n * H + (ok ? yy : 0)) * W + (ok ? xx : 0)) * kC + 8 * kq;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        v8bf v = {};
        if (ok) v = *reinterpret_cast<const v8bf*>(p + ks * 16);
        a[tap * KS + ks] = v;
      }
    }
    v16f acc = {};
#pragma unroll
    for (int q = 0; q < 9 * KS; ++q) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[q], b[q], acc, 0, 0, 0);
    const size_t rowBase = ((size_t)n * H + y) * W;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int m = 8 * (v >> 2) + 4 * kq + (v & 3);
      const int px = xt * 32 + m;
      if (px < W) {
        float val = acc[v] + bco;
        const size_t o = (rowBase + px) * kC + co;
        if (mode == 2) {
          outf[o] = val;
        } else {
          if (mode == 0) val = val > 0.0f ? val : 0.0f;
          outb[o] = f2bf(val);
        }
      }
    }
  }
}

/* ---- tiled convolution (inference) -----------------------------------------------------
   A workgroup of 4 waves computes R output rows y0 + k*dil (k < R) of a
   128-px column band of one image; wave w takes the 32 px x0 + 32w + (0..31).
   The R + 2 input rows those rows read (y0 + (j-1)*dil, j < R + 2) and the
   band plus a dil-px halo on either side are staged in LDS once, so the
   workgroup reads each input pixel from HBM about (R+2)/R times instead of
   once per tap.  Every staging load of a thread is issued before the first
   LDS write, so a workgroup waits for one memory latency, not one per load.
   Rows spaced by the dilation make the vertical taps of different output rows
   read the same input row: an A fragment read from LDS (one ds_read_b128:
   8 channels of one pixel per lane) feeds the MFMAs of up to three output
   rows, (R+2)*3*KS reads for R*9*KS MFMAs.
   LDS pixel p (64 bytes, 4 chunks of 8 channels) keeps chunk c at position
   c ^ ((p >> 2) & 3): the 16 consecutive pixels that 16 lanes of a
   ds_read_b128 read together then cover all 64 banks (256 bytes) once.
   BF16 outputs go back through LDS (the staged rows are dead by then) and
   leave as 16-byte stores of 8 channels; FP32 outputs are stored directly.
   With PAIR the input is not an NHWC map: channel 0 of sample z is the
   quarter-resolution plane prevIdx[z] (or refPlane when it is negative),
   channel 1 the plane curIdx0 + z, every other channel zero (the network's
   first layer, KS = 1).
   Block b runs logical tile (b mod 8) * ceil(tiles / 8) + b / 8: blocks on
   one XCD take consecutive row tiles of one band, whose staged rows overlap,
   so the overlap is served by that XCD's L2. */
constexpr int kTileW = 128;
constexpr int kTileR = 4;
constexpr int kMaxTileDil = 8;
constexpr int kRowPad = kTileW + 2 * kMaxTileDil;   /* staged row stride in pixels, a multiple of 16 */

__device__ __forceinline__ int lds_swz(int p) { return (p >> 2) & 3; }
__device__ __forceinline__ int lds_chunk(int p, int c) { return p * 64 + 16 * (c ^ lds_swz(p)); }

/* read by staging loads that fall outside the image (the zero padding): the
   loads need no select afterwards, so none of them waits before the last */
__device__ __align__(16) uint32_t kZeroPad[4] = {0, 0, 0, 0};

struct ConvTileArgs {
  const uint16_t* in;          /* NHWC-32 input (PAIR: unused) */
  const uint16_t* planes;      /* PAIR: quarter-resolution planes, planeStride elements apart */
  size_t planeStride;
  const uint16_t* refPlane;
  const int* prevIdx;
  int curIdx0;
  int N, H, W;
  const uint16_t* pk;
  const float* bias;
  int dil, mode;
  uint16_t* outb;
  float* outf;
  int rowTiles, xTiles, tiles, perXcd;
};

/* the staged tile (55 KB) allows two workgroups per CU, so two waves per
   SIMD is the occupancy to plan registers for */
template <int KS, bool PAIR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_conv16t(const ConvTileArgs a) {
  constexpr int R = kTileR;
  extern __shared__ __align__(16) uint8_t tile[];
  const int lt = (int)(blockIdx.x & 7) * a.perXcd + (int)(blockIdx.x >> 3);
  if (lt >= a.tiles) return;
  const int dil = a.dil, H = a.H, W = a.W;
  const int Wt = kTileW + 2 * dil;
  int b = lt;
  const int rt = b % a.rowTiles;
  b /= a.rowTiles;
  const int xt = b % a.xTiles;
  b /= a.xTiles;
  const int ph = b % dil;
  const int n = b / dil;
  const int y0 = rt * R * dil + ph;
  const int x0 = xt * kTileW;

  /* the layer's B fragments (weights) and bias, loaded with the staging loads */
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, half = lane >> 5;
  v8bf bq[9 * KS];
#pragma unroll
  for (int tap = 0; tap < 9; ++tap)
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
      bq[tap * KS + ks] = *reinterpret_cast<const v8bf*>(a.pk + ((size_t)(tap * 2 + ks) * 64 + lane) * 8);
  const int co = lane & 31;
  const float bco = a.bias ? a.bias[co] : 0.0f;

  /* stage R + 2 input rows of Wt pixels; zeros outside the image are the
     convolution's padding */
  if constexpr (PAIR) {
    constexpr int PPR = (kTileW + 2 * kMaxTileDil + 255) / 256;
    const int pv = a.prevIdx[n];
    const uint16_t* p0 = pv < 0 ? a.refPlane : a.planes + (size_t)pv * a.planeStride;
    const uint16_t* p1 = a.planes + (size_t)(a.curIdx0 + n) * a.planeStride;
    uint32_t v[(R + 2) * PPR];
#pragma unroll
    for (int j = 0; j < R + 2; ++j) {
      const int y = y0 + (j - 1) * dil;
      const bool rowOk = y >= 0 && y < H;
#pragma unroll
      for (int u = 0; u < PPR; ++u) {
        const int px = threadIdx.x + 256 * u, x = x0 - dil + px;
        /* every load is unconditional (outside the image it reads kZeroPad),
           so all of them are in flight together */
        const bool ok = rowOk && px < Wt && x >= 0 && x < W;
        const size_t o = (size_t)y * W + x;
        const uint16_t* q0 = ok ? p0 + o : reinterpret_cast<const uint16_t*>(kZeroPad);
        const uint16_t* q1 = ok ? p1 + o : reinterpret_cast<const uint16_t*>(kZeroPad);
        v[j * PPR + u] = (uint32_t)*q0 | ((uint32_t)*q1 << 16);
      }
    }
#pragma unroll
    for (int j = 0; j < R + 2; ++j)
#pragma unroll
      for (int u = 0; u < PPR; ++u) {
        const int px = threadIdx.x + 256 * u;
        if (px < Wt) {
          uint8_t* row = tile + (size_t)j * kRowPad * 64;
          *reinterpret_cast<uint4*>(row + lds_chunk(px, 0)) = make_uint4(v[j * PPR + u], 0, 0, 0);
          *reinterpret_cast<uint4*>(row + lds_chunk(px, 1)) = make_uint4(0, 0, 0, 0);
        }
      }
  } else {
    /* LDS DMA (global_load_lds_dwordx4): one instruction moves 16 pixels of a
       row, lane l loading the chunk that the swizzle puts at its LDS slot
       l * 16 (pixel l / 4, position l % 4); nothing passes through VGPRs,
       so every load of a wave is in flight at once */
    constexpr int NG = kRowPad / 16;
    for (int t = wave; t < (R + 2) * NG; t += 4) {
      const int j = t / NG, g = t - j * NG;
      const int y = y0 + (j - 1) * dil;
      const int p = g * 16 + (lane >> 2);
      const int c = (lane & 3) ^ lds_swz(p);
      const int x = x0 - dil + p;
      const bool ok = y >= 0 && y < H && p < Wt && x >= 0 && x < W;
      const uint16_t* q =
          ok ? a.in + (((size_t)n * H + y) * W + x) * kC + 8 * c : reinterpret_cast<const uint16_t*>(kZeroPad);
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)q,
                                       (__attribute__((address_space(3))) void*)(tile + (j * kRowPad + g * 16) * 64),
                                       16, 0, 0);
    }
  }
  __syncthreads();

  v16f acc[R];
#pragma unroll
  for (int k = 0; k < R; ++k) acc[k] = v16f{};
  const int pxBase = wave * 32 + (lane & 31);   /* staged pixel of tap dx = -dil */
#pragma unroll
  for (int j = 0; j < R + 2; ++j) {
    const uint8_t* row = tile + (size_t)j * kRowPad * 64;
    v8bf af[3 * KS];
#pragma unroll
    for (int tx = 0; tx < 3; ++tx)
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
        af[tx * KS + ks] = *reinterpret_cast<const v8bf*>(row + lds_chunk(pxBase + tx * dil, 2 * ks + half));
#pragma unroll
    for (int tx = 0; tx < 3; ++tx)
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int ty = 0; ty < 3; ++ty) {
          const int k = j - ty;   /* output row k reads input row k + ty */
          if (k >= 0 && k < R)
            acc[k] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[tx * KS + ks], bq[(ty * 3 + tx) * KS + ks], acc[k], 0,
                                                              0, 0);
        }
  }

  if (a.mode <= 1) {
    /* BF16 through LDS: row k of the tile at k * kTileW * 64 bytes */
    __syncthreads();
#pragma unroll
    for (int k = 0; k < R; ++k)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int pl = wave * 32 + 8 * (v >> 2) + 4 * half + (v & 3);
        float val = acc[k][v] + bco;
        if (a.mode == 0) val = val > 0.0f ? val : 0.0f;
        *reinterpret_cast<uint16_t*>(tile + k * kTileW * 64 + lds_chunk(pl, co >> 3) + 2 * (co & 7)) = f2bf(val);
      }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int y = y0 + k * dil;
      if (y >= H) break;
      uint16_t* orow = a.outb + ((size_t)n * H + y) * W * kC;
#pragma unroll
      for (int u = 0; u < kTileW * 4 / 256; ++u) {
        const int i = threadIdx.x + 256 * u;
        const int pl = i >> 2, c = i & 3, x = x0 + pl;
        if (x < W)
          *reinterpret_cast<uint4*>(orow + (size_t)x * kC + 8 * c) =
              *reinterpret_cast<const uint4*>(tile + k * kTileW * 64 + lds_chunk(pl, c));
      }
    }
  } else {
    /* FP32: every channel (mode 2, NHWC-32) or channels 0, 1 (mode 3, NHWC-2) */
    const int oc = a.mode == 2 ? kC : 2;
    if (co < oc) {
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int y = y0 + k * dil;
        if (y >= H) break;
        const size_t rowBase = ((size_t)n * H + y) * W;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int px = x0 + wave * 32 + 8 * (v >> 2) + 4 * half + (v & 3);
          if (px < W) a.outf[(rowBase + px) * oc + co] = acc[k][v] + bco;
        }
      }
    }
  }
}

/* ---- band walker (inference, 32 input channels) ---------------------------------------
   The tiled kernel above reads a tile's rows, computes and stores in turn, so
   a workgroup has memory traffic in flight only between its phases.  Here a
   persistent workgroup of 2 waves walks down one 64-px column band of one
   (image, dilation phase): step t computes the R output rows y = ph + m*dil,
   m = Rt .. Rt+R-1, from the R+2 input rows m = Rt-1 .. Rt+R, which sit in
   an LDS ring of R+2 rows (row m in slot (m+1) mod (R+2)).  Each step needs
   R new rows; every thread loads its share of the rows of step t+2 into
   registers at the start of step t and writes those of step t+1 to the ring
   after step t's MFMAs (to the slots of the R rows step t was the last to
   read), so each load has two steps of compute to arrive in.  A band's steps
   are split into segments (each with a 2-row warm-up) so that enough
   workgroups run; block b takes unit (b mod 8) * ceil(units / 8) + b / 8, so
   the segments of one band stay on one XCD.  The taps accumulate in the
   tiled kernel's order: the results are identical. */
constexpr int kWalkW = 64;
constexpr int kWalkThreads = 128;
constexpr int kWalkPad = kWalkW + 2 * kMaxTileDil;   /* staged row stride in pixels */
constexpr int kWalkRowChunks = kWalkPad * 4;           /* 16-byte chunks of a staged row */

struct ConvWalkArgs {
  const uint16_t* in;
  int N, H, W;
  const uint16_t* pk;
  const float* bias;
  int dil, mode;
  uint16_t* outb;
  float* outf;
  int steps;     /* R-row steps of a phase's row sequence (the longest phase) */
  int xTiles, segs, segLen, units, perXcd;
  /* PAIR0: the ring rows are the ReLU output of a first layer over the pair
     of quarter-resolution planes (as k_conv16t<1, true> reads them) */
  const uint16_t* planes;
  size_t planeStride;
  const uint16_t* refPlane;
  const int* prevIdx;
  int curIdx0;
  const uint16_t* pk0;
  const float* bias0;
  int dil0;
};

/* PAIR0: a wave stages the first layer's input for ring row m: the 3 plane
   rows y - dil0, y, y + dil0 of its taps, wp = 64 + 2 dil + 2 dil0 pixels
   from x0 - dil - dil0 on, as (prev | cur << 16) pairs at row stride kPairW,
   zero outside the image (the first layer's padding).  pair_fetch issues
   the loads (kPairLd per lane), pair_put writes them to LDS. */
constexpr int kPairW = 112;   /* >= 64 + 2*8 + 2*8 staged pixels, + 16 read past the end by idle pixels */
constexpr int kPairLd = (3 * (kWalkW + 4 * kMaxTileDil) + 63) / 64;

struct PairFetch {
  uint16_t lo[kPairLd], hi[kPairLd];
};

__device__ __forceinline__ void pair_fetch(const ConvWalkArgs& a, const uint16_t* pl0, const uint16_t* pl1, int ph,
                                           int x0, int m, PairFetch& f) {
  const int lane = threadIdx.x & 63;
  const int y = ph + m * a.dil, xs = x0 - a.dil - a.dil0, wp = kWalkW + 2 * a.dil + 2 * a.dil0;
#pragma unroll
  for (int j = 0; j < kPairLd; ++j) {
    const int idx = lane + 64 * j;
    const int k = idx / wp, p = idx - k * wp;
    const int yy = y + (k - 1) * a.dil0, xx = xs + p;
    const bool ok = k < 3 && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
    const size_t o = (size_t)yy * a.W + xx;
    f.lo[j] = *(ok ? pl0 + o : reinterpret_cast<const uint16_t*>(kZeroPad));
    f.hi[j] = *(ok ? pl1 + o : reinterpret_cast<const uint16_t*>(kZeroPad));
  }
}

__device__ __forceinline__ void pair_put(const ConvWalkArgs& a, const PairFetch& f, uint32_t* pairS) {
  const int lane = threadIdx.x & 63, wp = kWalkW + 2 * a.dil + 2 * a.dil0;
#pragma unroll
  for (int j = 0; j < kPairLd; ++j) {
    const int idx = lane + 64 * j;
    const int k = idx / wp, p = idx - k * wp;
    if (k < 3) pairS[k * kPairW + p] = (uint32_t)f.lo[j] | ((uint32_t)f.hi[j] << 16);
  }
}

/* the first layer's B fragments [tap][lane], then each wave's staged rows */
constexpr int kPair0Lds = 9 * 64 * 16 + 2 * 3 * kPairW * 4;

/* PAIR0: the first layer at tile i of ring row m (taps accumulated in order
   0..8 from zero, then bias and ReLU, as k_conv16t<1, true> does), stored as
   BF16 into the row's ring slot; pixels outside the image are the second
   layer's zero padding */
template <int RING>
__device__ __forceinline__ void pair_store(const ConvWalkArgs& a, const uint32_t* pairS, const uint8_t* b0S,
                                           float bco0, uint8_t* ring, int ph, int x0, int m, int i) {
  const int lane = threadIdx.x & 63, half = lane >> 5, co = lane & 31;
  const int pl0 = 32 * i + (lane & 31);   /* the lane's ring pixel (A row) */
  v16f acc = {};
#pragma unroll
  for (int tap = 0; tap < 9; ++tap) {
    /* channels 0, 1 of the tap's pixel in lanes 0..31; the upper K half is zero */
    const uint32_t pr = pairS[(tap / 3) * kPairW + pl0 + (tap % 3) * a.dil0];
    const uint4 u = make_uint4(half ? 0u : pr, 0u, 0u, 0u);
    const v8bf b = *reinterpret_cast<const v8bf*>(b0S + (tap * 64 + lane) * 16);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(*reinterpret_cast<const v8bf*>(&u), b, acc, 0, 0, 0);
  }
  const int y = ph + m * a.dil;
  const bool rowOk = y >= 0 && y < a.H;
  uint8_t* row = ring + (size_t)((m + 1) % RING) * kWalkPad * 64;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int pl = 32 * i + 8 * (e >> 2) + 4 * half + (e & 3), x = x0 - a.dil + pl;
    if (pl < kWalkW + 2 * a.dil) {
      float val = acc[e] + bco0;
      val = val > 0.0f ? val : 0.0f;
      const uint16_t bv = rowOk && x >= 0 && x < a.W ? f2bf(val) : (uint16_t)0;
      *reinterpret_cast<uint16_t*>(row + lds_chunk(pl, co >> 3) + 2 * (co & 7)) = bv;
    }
  }
}

/* chunk u of a thread's share of rows m0 .. m0+NR-1 of phase ph: the address
   to load (kZeroPad outside the image) and the ring position to store it at */
template <int NR>
__device__ __forceinline__ const uint16_t* walk_src(const ConvWalkArgs& a, int n, int ph, int x0, int m0, int u) {
  const int i = (int)threadIdx.x + kWalkThreads * u;
  const int r = i / kWalkRowChunks, rem = i - r * kWalkRowChunks;
  const int p = rem >> 2, c = rem & 3;
  const int y = ph + (m0 + r) * a.dil, x = x0 - a.dil + p;
  const bool ok = i < NR * kWalkRowChunks && p < kWalkW + 2 * a.dil && y >= 0 && y < a.H && x >= 0 && x < a.W;
  return ok ? a.in + (((size_t)n * a.H + y) * a.W + x) * kC + 8 * c
            : reinterpret_cast<const uint16_t*>(kZeroPad);
}

template <int NR, int RING>
__device__ __forceinline__ void walk_put(uint8_t* ring, int m0, int u, const uint4& v) {
  const int i = (int)threadIdx.x + kWalkThreads * u;
  if (i < NR * kWalkRowChunks) {
    const int r = i / kWalkRowChunks, rem = i - r * kWalkRowChunks;
    const int slot = (m0 + r + 1) % RING;
    *reinterpret_cast<uint4*>(ring + (size_t)slot * kWalkPad * 64 + lds_chunk(rem >> 2, rem & 3)) = v;
  }
}

template <int R, bool PAIR0>
__global__ __launch_bounds__(kWalkThreads) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_conv16w(
    const ConvWalkArgs a) {
  static_assert(!PAIR0 || R == 2, "PAIR0: wave w computes ring row m + w of each step");
  constexpr int KS = 2, RING = R + 2;
  extern __shared__ __align__(16) uint8_t lds[];
  uint8_t* ring = lds;                              /* RING rows x kWalkPad px x 64 B */
  uint8_t* outS = lds + RING * kWalkPad * 64;       /* R rows x kWalkW px x 64 B (BF16 output) */
  uint8_t* b0S = outS + R * kWalkW * 64;            /* PAIR0: kPair0Lds bytes */
  uint32_t* pairS = reinterpret_cast<uint32_t*>(b0S + 9 * 64 * 16) + (threadIdx.x >> 6) * 3 * kPairW;
  const int unit = (int)(blockIdx.x & 7) * a.perXcd + (int)(blockIdx.x >> 3);
  if (unit >= a.units) return;
  int b = unit;
  const int seg = b % a.segs;
  b /= a.segs;
  const int xt = b % a.xTiles;
  b /= a.xTiles;
  const int ph = b % a.dil;
  const int n = b / a.dil;
  const int t0 = seg * a.segLen, t1 = min(a.steps, t0 + a.segLen);
  if (t0 >= t1) return;
  const int dil = a.dil, H = a.H, W = a.W, x0 = xt * kWalkW;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, half = lane >> 5, co = lane & 31;

  v8bf bq[9 * KS];
#pragma unroll
  for (int tap = 0; tap < 9; ++tap)
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
      bq[tap * KS + ks] = *reinterpret_cast<const v8bf*>(a.pk + ((size_t)(tap * 2 + ks) * 64 + lane) * 8);
  const float bco = a.bias ? a.bias[co] : 0.0f;
  float bco0 = 0.0f;
  const uint16_t* pl0 = nullptr;
  const uint16_t* pl1 = nullptr;
  if constexpr (PAIR0) {
    /* the first layer's K step 0 fragments (its input has 2 channels) */
    for (int i = threadIdx.x; i < 9 * 64; i += kWalkThreads)
      *reinterpret_cast<uint4*>(b0S + i * 16) =
          *reinterpret_cast<const uint4*>(a.pk0 + ((size_t)((i >> 6) * 2) * 64 + (i & 63)) * 8);
    __syncthreads();
    bco0 = a.bias0 ? a.bias0[co] : 0.0f;
    const int pv = a.prevIdx[n];
    pl0 = pv < 0 ? a.refPlane : a.planes + (size_t)pv * a.planeStride;
    pl1 = a.planes + (size_t)(a.curIdx0 + n) * a.planeStride;
    /* warm-up: ring rows R*t0 - 1 .. R*t0 + R; wave w computes rows 2w, 2w + 1 */
    const int m0 = R * t0 - 1;
#pragma nounroll
    for (int r = 0; r < 2; ++r) {
      PairFetch f;
      pair_fetch(a, pl0, pl1, ph, x0, m0 + 2 * wave + r, f);
      pair_put(a, f, pairS);
      __syncthreads();
      pair_store<RING>(a, pairS, b0S, bco0, ring, ph, x0, m0 + 2 * wave + r, 0);
      pair_store<RING>(a, pairS, b0S, bco0, ring, ph, x0, m0 + 2 * wave + r, 1);
      pair_store<RING>(a, pairS, b0S, bco0, ring, ph, x0, m0 + 2 * wave + r, 2);
      __syncthreads();   /* pairS is rewritten for the next row */
    }
  } else {
    /* warm-up: the segment's first RING rows, in two halves of R + 2 rows'
       worth of chunks at most 10 a thread */
    static_assert((RING * kWalkRowChunks + kWalkThreads - 1) / kWalkThreads <= 10, "warm-up chunks");
    const int m0 = R * t0 - 1;
#define WALK_LD(u) const uint4 f##u = *reinterpret_cast<const uint4*>(walk_src<RING>(a, n, ph, x0, m0, u))
#define WALK_ST(u) walk_put<RING, RING>(ring, m0, u, f##u)
    WALK_LD(0); WALK_LD(1); WALK_LD(2); WALK_LD(3); WALK_LD(4);
    WALK_LD(5); WALK_LD(6); WALK_LD(7); WALK_LD(8); WALK_LD(9);
    WALK_ST(0); WALK_ST(1); WALK_ST(2); WALK_ST(3); WALK_ST(4);
    WALK_ST(5); WALK_ST(6); WALK_ST(7); WALK_ST(8); WALK_ST(9);
#undef WALK_LD
#undef WALK_ST
  }
  __syncthreads();

  const int pxBase = wave * 32 + (lane & 31);   /* staged pixel of tap dx = -dil */
  /* the new rows of step t + 1 sit in registers (q*) while step t runs; the
     loads for step t + 2 (p*) are issued at the start of step t, so each
     row load has two steps of MFMAs to arrive in.  Loads past the segment
     read rows that exist or the zero pad and are dropped.  PAIR0 instead
     loads the plane pixels of step t + 1's new rows at the start of step t
     and runs the first layer on them after step t's MFMAs. */
  static_assert((R * kWalkRowChunks + kWalkThreads - 1) / kWalkThreads <= 5, "prefetch chunks");
#define WALK_LD(v, u, m) const uint4 v##u = *reinterpret_cast<const uint4*>(walk_src<R>(a, n, ph, x0, m, u))
  uint4 q0, q1, q2, q3, q4;
  if constexpr (!PAIR0) {
    const int m1 = R * t0 + R + 1;
    WALK_LD(l, 0, m1); WALK_LD(l, 1, m1); WALK_LD(l, 2, m1); WALK_LD(l, 3, m1); WALK_LD(l, 4, m1);
    q0 = l0; q1 = l1; q2 = l2; q3 = l3; q4 = l4;
  }
  for (int t = t0; t < t1; ++t) {
    uint4 p0, p1, p2, p3, p4;
    PairFetch pf;
    if constexpr (PAIR0) {
      pair_fetch(a, pl0, pl1, ph, x0, R * t + R + 1 + wave, pf);   /* this wave's new row of step t + 1 */
    } else {
      const int m2 = R * t + 2 * R + 1;   /* first new row of step t + 2 */
      WALK_LD(l, 0, m2); WALK_LD(l, 1, m2); WALK_LD(l, 2, m2); WALK_LD(l, 3, m2); WALK_LD(l, 4, m2);
      p0 = l0; p1 = l1; p2 = l2; p3 = l3; p4 = l4;
    }

    v16f acc[R];
#pragma unroll
    for (int k = 0; k < R; ++k) acc[k] = v16f{};
#pragma unroll
    for (int j = 0; j < R + 2; ++j) {
      const uint8_t* row = ring + (size_t)((R * t + j) % RING) * kWalkPad * 64;   /* input row m = Rt - 1 + j */
      v8bf af[3 * KS];
#pragma unroll
      for (int tx = 0; tx < 3; ++tx)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
          af[tx * KS + ks] = *reinterpret_cast<const v8bf*>(row + lds_chunk(pxBase + tx * dil, 2 * ks + half));
#pragma unroll
      for (int tx = 0; tx < 3; ++tx)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
          for (int ty = 0; ty < 3; ++ty) {
            const int k = j - ty;   /* output row k reads input row k + ty */
            if (k >= 0 && k < R)
              acc[k] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[tx * KS + ks], bq[(ty * 3 + tx) * KS + ks], acc[k],
                                                                0, 0, 0);
          }
    }

    if (a.mode <= 1) {
#pragma unroll
      for (int k = 0; k < R; ++k)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int pl = wave * 32 + 8 * (v >> 2) + 4 * half + (v & 3);
          float val = acc[k][v] + bco;
          if (a.mode == 0) val = val > 0.0f ? val : 0.0f;
          *reinterpret_cast<uint16_t*>(outS + k * kWalkW * 64 + lds_chunk(pl, co >> 3) + 2 * (co & 7)) = f2bf(val);
        }
    }
    __syncthreads();   /* every wave is done with the ring rows this step read last */
    if (a.mode <= 1) {
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int y = ph + (R * t + k) * dil;
        if (y < H) {
          uint16_t* orow = a.outb + ((size_t)n * H + y) * W * kC;
#pragma unroll
          for (int u = 0; u < kWalkW * 4 / kWalkThreads; ++u) {
            const int i = threadIdx.x + kWalkThreads * u;
            const int pl = i >> 2, c = i & 3, x = x0 + pl;
            if (x < W)
              *reinterpret_cast<uint4*>(orow + (size_t)x * kC + 8 * c) =
                  *reinterpret_cast<const uint4*>(outS + k * kWalkW * 64 + lds_chunk(pl, c));
          }
        }
      }
    } else if constexpr (!PAIR0) {
      const int oc = a.mode == 2 ? kC : 2;
      if (co < oc) {
#pragma unroll
        for (int k = 0; k < R; ++k) {
          const int y = ph + (R * t + k) * dil;
          if (y < H) {
            const size_t rowBase = ((size_t)n * H + y) * W;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
              const int px = x0 + wave * 32 + 8 * (v >> 2) + 4 * half + (v & 3);
              if (px < W) a.outf[(rowBase + px) * oc + co] = acc[k][v] + bco;
            }
          }
        }
      }
    }
    if (t + 1 < t1) {
      const int m1 = R * t + R + 1;
      if constexpr (PAIR0) {
        /* wave w's new row of step t + 1 */
        pair_put(a, pf, pairS);
        __syncthreads();
        pair_store<RING>(a, pairS, b0S, bco0, ring, ph, x0, m1 + wave, 0);
        pair_store<RING>(a, pairS, b0S, bco0, ring, ph, x0, m1 + wave, 1);
        pair_store<RING>(a, pairS, b0S, bco0, ring, ph, x0, m1 + wave, 2);
      } else {
        walk_put<R, RING>(ring, m1, 0, q0);
        walk_put<R, RING>(ring, m1, 1, q1);
        walk_put<R, RING>(ring, m1, 2, q2);
        walk_put<R, RING>(ring, m1, 3, q3);
        walk_put<R, RING>(ring, m1, 4, q4);
      }
    }
    if constexpr (!PAIR0) {
      q0 = p0; q1 = p1; q2 = p2; q3 = p3; q4 = p4;
    }
    __syncthreads();   /* the new rows staged; the output rows read out */
  }
#undef WALK_LD
}

/* ---- weight gradient ----------------------------------------------------------------
   dW[tap][ci][co] += sum_p in[p + tap*dil][ci] * dz[p][co] for taps
   tap0 .. tap0+2: an MFMA GEMM with M = ci, N = co, K = pixels (8 per step). */
__global__ __launch_bounds__(256) void k_wgrad(const uint16_t* __restrict__ in, const uint16_t* __restrict__ dz,
                                               int N, int H, int W, int dil, int tap0, float* __restrict__ dW) {
  __shared__ float red[3 * kC * kC];   /* the workgroup's partial sums */
  for (int i = threadIdx.x; i < 3 * kC * kC; i += blockDim.x) red[i] = 0.0f;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nw = (gridDim.x * blockDim.x) >> 6;
  const int c = lane & 31, kq = lane >> 5;
  const long long npix = (long long)N * H * W;
  const long long ngroups = (npix + 7) / 8;
  v16f acc[3] = {};
  for (long long g = gw; g < ngroups; g += nw) {
    v4s bfr;
    int py[4], px[4], pn[4];
    bool pv[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const long long p = g * 8 + 4 * kq + e;
      pv[e] = p < npix;
      const long long q = pv[e] ? p : 0;
      px[e] = (int)(q % W);
      const long long r = q / W;
      py[e] = (int)(r % H);
      pn[e] = (int)(r / H);
      bfr[e] = pv[e] ? (short)dz[q * kC + c] : (short)0;
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int tap = tap0 + j;
      const int dy = (tap / 3 - 1) * dil, dx = (tap % 3 - 1) * dil;
      v4s a;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int yy = py[e] + dy, xx = px[e] + dx;
        const bool ok = pv[e] && yy >= 0 && yy < H && xx >= 0 && xx < W;
        a[e] = ok ? (short)in[(((size_t)pn[e] * H + yy) * W + xx) * kC + c] : (short)0;
      }
      acc[j] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, bfr, acc[j], 0, 0, 0);
    }
  }
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int ci = 8 * (v >> 2) + 4 * kq + (v & 3);
      atomicAdd(&red[(j * kC + ci) * kC + c], acc[j][v]);
    }
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * kC * kC; i += blockDim.x) atomicAdd(&dW[tap0 * kC * kC + i], red[i]);
}

/* dB[co] += sum_p dz[p][co]: each half-wave walks pixels, lane = channel */
__global__ __launch_bounds__(256) void k_bgrad(const uint16_t* __restrict__ dz, long long npix, float* __restrict__ dB) {
  const int c = threadIdx.x & 31;
  const long long stride = (long long)gridDim.x * (blockDim.x >> 5);
  float s = 0.0f;
  for (long long p = blockIdx.x * (blockDim.x >> 5) + (threadIdx.x >> 5); p < npix; p += stride)
    s += bf2f(dz[p * kC + c]);
  atomicAdd(&dB[c], s);
}

/* grad *= (act > 0), bf16 in place */
__global__ void k_relu_mask(uint16_t* __restrict__ g, const uint16_t* __restrict__ act, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(bf2f(act[i]) > 0.0f)) g[i] = 0;
}

__global__ void k_f2bf(const float* __restrict__ f, uint16_t* __restrict__ b, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = f2bf(f[i]);
}

/* ---- 4x4 box downsample of a luma plane into one channel of an NHWC-32 map:
   (mean - 128) / 64 */
__global__ void k_down4(const uint8_t* __restrict__ src, int pitch, int Wq, int Hq, uint16_t* __restrict__ out,
                        int chan) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= Wq || y >= Hq) return;
  const uint8_t* p = src + (size_t)(4 * y) * pitch + 4 * x;
  uint32_t s = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const uint32_t w = *reinterpret_cast<const uint32_t*>(p + (size_t)r * pitch);
    s += (w & 0xff) + ((w >> 8) & 0xff) + ((w >> 16) & 0xff) + (w >> 24);
  }
  out[((size_t)y * Wq + x) * kC + chan] = f2bf(((float)s * (1.0f / 16.0f) - 128.0f) * (1.0f / 64.0f));
}

/* ---- field pooling: pred[f] = mean over the field's window of channels 0, 1
   of an fp32 NHWC map of cs channels; one wave per field */
__global__ void k_pool(const float* __restrict__ map, int H, int W, int cs, const int* __restrict__ win, int nf,
                       float* __restrict__ pred) {
  const int f = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (f >= nf) return;
  const int n = win[5 * f];
  const int x0 = max(win[5 * f + 1], 0), y0 = max(win[5 * f + 2], 0);
  const int x1 = min(win[5 * f + 3], W), y1 = min(win[5 * f + 4], H);   /* clipped to the map */
  const int w = x1 > x0 ? x1 - x0 : 0, cnt = y1 > y0 ? w * (y1 - y0) : 0;
  float sx = 0.0f, sy = 0.0f;
  for (int i = lane; i < cnt; i += 64) {
    const int y = y0 + i / w, x = x0 + i % w;
    const float* p = map + (((size_t)n * H + y) * W + x) * cs;
    sx += p[0];
    sy += p[1];
  }
  for (int off = 32; off > 0; off >>= 1) {
    sx += __shfl_xor(sx, off);
    sy += __shfl_xor(sy, off);
  }
  if (lane == 0) {
    pred[2 * f] = cnt ? sx / cnt : 0.0f;
    pred[2 * f + 1] = cnt ? sy / cnt : 0.0f;
  }
}

/* d map (fp32, zeroed by the caller) += dpred[f] / area over each window */
__global__ void k_pool_bwd(const float* __restrict__ dpred, int H, int W, const int* __restrict__ win, int nf,
                           float* __restrict__ dmap) {
  const int f = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (f >= nf) return;
  const int n = win[5 * f];
  const int x0 = max(win[5 * f + 1], 0), y0 = max(win[5 * f + 2], 0);
  const int x1 = min(win[5 * f + 3], W), y1 = min(win[5 * f + 4], H);   /* clipped to the map */
  const int w = x1 > x0 ? x1 - x0 : 0, cnt = y1 > y0 ? w * (y1 - y0) : 0;
  if (!cnt) return;
  const float gx = dpred[2 * f] / cnt, gy = dpred[2 * f + 1] / cnt;
  for (int i = lane; i < cnt; i += 64) {
    const int y = y0 + i / w, x = x0 + i % w;
    float* p = dmap + (((size_t)n * H + y) * W + x) * kC;
    atomicAdd(p, gx);
    atomicAdd(p + 1, gy);
  }
}

/* mean squared error over valid fields: loss += sum w*(pred-label)^2 / nvalid,
   dpred = 2 w (pred - label) / nvalid */
__global__ void k_loss(const float* __restrict__ pred, const float* __restrict__ label,
                       const float* __restrict__ weight, int nf, float inv, float* __restrict__ dpred,
                       float* __restrict__ loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < 2 * nf) {
    const float d = pred[i] - label[i], w = weight[i >> 1];
    dpred[i] = 2.0f * w * d * inv;
    l = w * d * d * inv;
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
  if ((threadIdx.x & 63) == 0 && l != 0.0f) atomicAdd(loss, l);
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

int blocks_for(long long n, int t) { return (int)((n + t - 1) / t); }

int conv_grid(void) {
  static int g = 0;
  if (g == 0) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
    g = cus * 4;   /* 16 waves per CU walking row tiles */
  }
  return g;
}

}  // namespace

#define CNN_RET return (int)hipGetLastError()

extern "C" int vsgc_pack_weights(const float* w, uint16_t* packed, int transposed, void* stream) {
  hipLaunchKernelGGL(k_pack, dim3(blocks_for(9 * 4 * 64 * 4, 256)), dim3(256), 0, (hipStream_t)stream, w, packed,
                     transposed);
  CNN_RET;
}

extern "C" int vsgc_conv(const uint16_t* in, int n, int h, int w, const uint16_t* packed, const float* bias,
                         int dil, int mode, uint16_t* outb, float* outf, void* stream) {
  return vsgc_conv_k(in, n, h, w, packed, bias, dil, mode, 32, outb, outf, stream);
}

extern "C" int vsgc_conv_k(const uint16_t* in, int n, int h, int w, const uint16_t* packed, const float* bias,
                           int dil, int mode, int inChannels, uint16_t* outb, float* outf, void* stream) {
  if (n <= 0 || h <= 0 || w <= 0 || dil < 1 || mode < 0 || mode > 2) return (int)hipErrorInvalidValue;
  if ((mode == 2 && !outf) || (mode != 2 && !outb)) return (int)hipErrorInvalidValue;
  if (inChannels <= 8)
    hipLaunchKernelGGL(k_conv<1>, dim3(conv_grid()), dim3(256), 0, (hipStream_t)stream, in, n, h, w, packed, bias,
                       dil, mode, outb, outf);
  else
    hipLaunchKernelGGL(k_conv<4>, dim3(conv_grid()), dim3(256), 0, (hipStream_t)stream, in, n, h, w, packed, bias,
                       dil, mode, outb, outf);
  CNN_RET;
}

extern "C" int vsgc_conv_wgrad(const uint16_t* in, const uint16_t* dz, int n, int h, int w, int dil, float* dW,
                               float* dB, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  for (int tap0 = 0; tap0 < 9; tap0 += 3)
    hipLaunchKernelGGL(k_wgrad, dim3(128), dim3(256), 0, st, in, dz, n, h, w, dil, tap0, dW);
  if (dB) hipLaunchKernelGGL(k_bgrad, dim3(256), dim3(256), 0, st, dz, (long long)n * h * w, dB);
  CNN_RET;
}

extern "C" int vsgc_relu_mask(uint16_t* grad, const uint16_t* act, long long count, void* stream) {
  hipLaunchKernelGGL(k_relu_mask, dim3(blocks_for(count, 256)), dim3(256), 0, (hipStream_t)stream, grad, act, count);
  CNN_RET;
}

extern "C" int vsgc_f32_to_bf16(const float* f, uint16_t* b, long long count, void* stream) {
  hipLaunchKernelGGL(k_f2bf, dim3(blocks_for(count, 256)), dim3(256), 0, (hipStream_t)stream, f, b, count);
  CNN_RET;
}

extern "C" int vsgc_downsample(const uint8_t* plane, int pitch, int width, int height, uint16_t* out, int chan,
                               void* stream) {
  const int wq = width / 4, hq = height / 4;
  if (wq <= 0 || hq <= 0 || chan < 0 || chan >= kC) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_down4, dim3(blocks_for(wq, 256), hq), dim3(256), 0, (hipStream_t)stream, plane, pitch, wq, hq,
                     out, chan);
  CNN_RET;
}

extern "C" int vsgc_field_pool(const float* map, int h, int w, const int* windows, int nf, float* pred,
                               void* stream) {
  if (nf <= 0) return 0;
  hipLaunchKernelGGL(k_pool, dim3(blocks_for((long long)nf * 64, 256)), dim3(256), 0, (hipStream_t)stream, map, h, w,
                     kC, windows, nf, pred);
  CNN_RET;
}

extern "C" int vsgc_field_pool2(const float* map, int h, int w, const int* windows, int nf, float* pred,
                                void* stream) {
  if (nf <= 0) return 0;
  hipLaunchKernelGGL(k_pool, dim3(blocks_for((long long)nf * 64, 256)), dim3(256), 0, (hipStream_t)stream, map, h, w,
                     2, windows, nf, pred);
  CNN_RET;
}

extern "C" int vsgc_field_pool_bwd(const float* dpred, int h, int w, const int* windows, int nf, float* dmap,
                                   void* stream) {
  if (nf <= 0) return 0;
  hipLaunchKernelGGL(k_pool_bwd, dim3(blocks_for((long long)nf * 64, 256)), dim3(256), 0, (hipStream_t)stream, dpred,
                     h, w, windows, nf, dmap);
  CNN_RET;
}

extern "C" int vsgc_field_loss(const float* pred, const float* label, const float* weight, int nf, float inv,
                               float* dpred, float* loss, void* stream) {
  if (nf <= 0) return 0;
  hipLaunchKernelGGL(k_loss, dim3(blocks_for(2LL * nf, 256)), dim3(256), 0, (hipStream_t)stream, pred, label, weight,
                     nf, inv, dpred, loss);
  CNN_RET;
}

extern "C" int vsgc_adam(float* w, const float* g, float* m, float* v, int count, float lr, float b1, float b2,
                         float eps, int step, void* stream) {
  const float c1 = 1.0f / (1.0f - powf(b1, (float)step)), c2 = 1.0f / (1.0f - powf(b2, (float)step));
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(count, 256)), dim3(256), 0, (hipStream_t)stream, w, g, m, v, count, lr,
                     b1, b2, eps, c1, c2);
  CNN_RET;
}

extern "C" int vsgc_pack_weights16(const float* w, uint16_t* packed, void* stream) {
  hipLaunchKernelGGL(k_pack16, dim3(blocks_for(9 * 2 * 64 * 8, 256)), dim3(256), 0, (hipStream_t)stream, w, packed);
  CNN_RET;
}

namespace {

/* geometry of a tiled launch; false when the grid would not fit */
bool tiled_grid(ConvTileArgs& a, unsigned& blocks, size_t& lds) {
  a.rowTiles = (a.H + kTileR * a.dil - 1) / (kTileR * a.dil);
  a.xTiles = (a.W + kTileW - 1) / kTileW;
  const long long tiles = (long long)a.N * a.rowTiles * a.dil * a.xTiles;
  if (tiles > 0x7FFFFFF0LL) return false;
  a.tiles = (int)tiles;
  a.perXcd = (a.tiles + 7) / 8;
  blocks = 8u * (unsigned)a.perXcd;
  lds = (size_t)(kTileR + 2) * kRowPad * 64;
  return true;
}

/* the walker's grid: segments of a band's steps so that about 2048 workgroups
   run (8 per CU at the walker's LDS use of 47 KB allows 3 resident) */
constexpr int kWalkR = 2;   /* output rows per walker step */

bool walk_grid(ConvWalkArgs& w, unsigned& blocks, size_t& lds) {
  w.steps = (w.H + kWalkR * w.dil - 1) / (kWalkR * w.dil);
  w.xTiles = (w.W + kWalkW - 1) / kWalkW;
  const long long bands = (long long)w.N * w.dil * w.xTiles;
  long long segs = (2048 + bands - 1) / bands;
  if (segs < 1) segs = 1;
  if (segs > w.steps) segs = w.steps;
  w.segLen = (int)((w.steps + segs - 1) / segs);
  w.segs = (w.steps + w.segLen - 1) / w.segLen;
  const long long units = bands * w.segs;
  if (units > 0x7FFFFFF0LL) return false;
  w.units = (int)units;
  w.perXcd = (w.units + 7) / 8;
  blocks = 8u * (unsigned)w.perXcd;
  lds = (size_t)(kWalkR + 2) * kWalkPad * 64 + (size_t)kWalkR * kWalkW * 64;
  return true;
}

/* VIDSTAB_AMD_CONV_WALK=0 runs the 32-channel layers on the tiled kernel */
bool conv_walker_off() {
  static int off = -1;
  if (off < 0) {
    const char* e = getenv("VIDSTAB_AMD_CONV_WALK");
    off = e && atoi(e) == 0 ? 1 : 0;
  }
  return off == 1;
}

}  // namespace

extern "C" int vsgc_conv16(const uint16_t* in, int n, int h, int w, const uint16_t* packed16, const float* bias,
                           int dil, int mode, int inChannels, uint16_t* outb, float* outf, void* stream) {
  if (n <= 0 || h <= 0 || w <= 0 || dil < 1 || mode < 0 || mode > 3) return (int)hipErrorInvalidValue;
  if ((mode >= 2 && !outf) || (mode < 2 && !outb)) return (int)hipErrorInvalidValue;
  if (dil <= kMaxTileDil) {
    ConvTileArgs a = {};
    a.in = in;
    a.N = n;
    a.H = h;
    a.W = w;
    a.pk = packed16;
    a.bias = bias;
    a.dil = dil;
    a.mode = mode;
    a.outb = outb;
    a.outf = outf;
    unsigned blocks;
    size_t lds;
    if (inChannels > 16 && !conv_walker_off()) {
      ConvWalkArgs wk = {};
      wk.in = in;
      wk.N = n;
      wk.H = h;
      wk.W = w;
      wk.pk = packed16;
      wk.bias = bias;
      wk.dil = dil;
      wk.mode = mode;
      wk.outb = outb;
      wk.outf = outf;
      if (!walk_grid(wk, blocks, lds)) return (int)hipErrorInvalidValue;
      hipLaunchKernelGGL((k_conv16w<kWalkR, false>), dim3(blocks), dim3(kWalkThreads), lds, (hipStream_t)stream, wk);
      CNN_RET;
    }
    if (!tiled_grid(a, blocks, lds)) return (int)hipErrorInvalidValue;
    if (inChannels <= 16)
      hipLaunchKernelGGL((k_conv16t<1, false>), dim3(blocks), dim3(256), lds, (hipStream_t)stream, a);
    else
      hipLaunchKernelGGL((k_conv16t<2, false>), dim3(blocks), dim3(256), lds, (hipStream_t)stream, a);
    CNN_RET;
  }
  if (mode == 3) return (int)hipErrorInvalidValue;
  return vsgc_conv16_direct(in, n, h, w, packed16, bias, dil, mode, inChannels, outb, outf, stream);
}

extern "C" int vsgc_conv16_pair(const uint16_t* planes, size_t planeStride, const uint16_t* refPlane,
                                const int* prevIdx, int curIdx0, int n, int h, int w, const uint16_t* packed16,
                                const float* bias, int dil, int mode, uint16_t* outb, float* outf, void* stream) {
  if (!planes || !prevIdx || n <= 0 || h <= 0 || w <= 0 || dil < 1 || dil > kMaxTileDil || mode < 0 || mode > 3 ||
      curIdx0 < 0 || planeStride < (size_t)h * w)
    return (int)hipErrorInvalidValue;
  if ((mode >= 2 && !outf) || (mode < 2 && !outb)) return (int)hipErrorInvalidValue;
  ConvTileArgs a = {};
  a.planes = planes;
  a.planeStride = planeStride;
  a.refPlane = refPlane ? refPlane : planes;
  a.prevIdx = prevIdx;
  a.curIdx0 = curIdx0;
  a.N = n;
  a.H = h;
  a.W = w;
  a.pk = packed16;
  a.bias = bias;
  a.dil = dil;
  a.mode = mode;
  a.outb = outb;
  a.outf = outf;
  unsigned blocks;
  size_t lds;
  if (!tiled_grid(a, blocks, lds)) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL((k_conv16t<1, true>), dim3(blocks), dim3(256), lds, (hipStream_t)stream, a);
  CNN_RET;
}

extern "C" int vsgc_conv16_pair2(const uint16_t* planes, size_t planeStride, const uint16_t* refPlane,
                                 const int* prevIdx, int curIdx0, int n, int h, int w, const uint16_t* packed0,
                                 const float* bias0, int dil0, const uint16_t* packed1, const float* bias1, int dil1,
                                 int mode, uint16_t* outb, float* outf, void* stream) {
  if (!planes || !prevIdx || !packed0 || !packed1 || n <= 0 || h <= 0 || w <= 0 || dil0 < 1 || dil1 < 1 ||
      dil0 > kMaxTileDil || dil1 > kMaxTileDil || mode < 0 || mode > 1 || !outb || curIdx0 < 0 ||
      planeStride < (size_t)h * w)
    return (int)hipErrorInvalidValue;
  ConvWalkArgs wk = {};
  wk.N = n;
  wk.H = h;
  wk.W = w;
  wk.pk = packed1;
  wk.bias = bias1;
  wk.dil = dil1;
  wk.mode = mode;
  wk.outb = outb;
  wk.outf = outf;
  wk.planes = planes;
  wk.planeStride = planeStride;
  wk.refPlane = refPlane ? refPlane : planes;
  wk.prevIdx = prevIdx;
  wk.curIdx0 = curIdx0;
  wk.pk0 = packed0;
  wk.bias0 = bias0;
  wk.dil0 = dil0;
  unsigned blocks;
  size_t lds;
  if (!walk_grid(wk, blocks, lds)) return (int)hipErrorInvalidValue;
  lds += kPair0Lds;
  hipLaunchKernelGGL((k_conv16w<kWalkR, true>), dim3(blocks), dim3(kWalkThreads), lds, (hipStream_t)stream, wk);
  CNN_RET;
}

extern "C" int vsgc_conv16_direct(const uint16_t* in, int n, int h, int w, const uint16_t* packed16,
                                  const float* bias, int dil, int mode, int inChannels, uint16_t* outb, float* outf,
                                  void* stream) {
  if (n <= 0 || h <= 0 || w <= 0 || dil < 1 || mode < 0 || mode > 2) return (int)hipErrorInvalidValue;
  if ((mode == 2 && !outf) || (mode != 2 && !outb)) return (int)hipErrorInvalidValue;
  if (inChannels <= 16)
    hipLaunchKernelGGL(k_conv16<1>, dim3(conv_grid()), dim3(256), 0, (hipStream_t)stream, in, n, h, w, packed16,
                       bias, dil, mode, outb, outf);
  else
    hipLaunchKernelGGL(k_conv16<2>, dim3(conv_grid()), dim3(256), 0, (hipStream_t)stream, in, n, h, w, packed16,
                       bias, dil, mode, outb, outf);
  CNN_RET;
}
