/*
 * vsg_device.h -- device-side building blocks shared by the blur, field and
 * selection kernels (vsg_kernels.hip) and the SAD searches (vsg_search.hip:
 * the per-job searches and the persistent queue kernel of the multi-clip
 * form).
 */
#ifndef VSG_DEVICE_H
#define VSG_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vsg_common.h"

typedef const __attribute__((address_space(4))) uint32_t* vsg_cword_ptr;   /* scalar loads */
typedef const __attribute__((address_space(1))) uint32_t* vsg_gword_ptr;   /* global loads */

/* one field's search: geometry of its candidate grid and where its planes are */
struct VSGTile {
  const uint8_t* prev;     /* previous blurred plane */
  int pitch;               /* bytes per row of prev (multiple of 64) */
  vsg_cword_ptr patch;     /* the current field, S rows of S*bpp bytes, 4-byte aligned */
  const uint32_t* lpatch;  /* the same field staged in LDS (plans with an LDS patch) */
  int S, s, m, U, V;       /* field size, grid step, lane layout */
  int alo, G, org, bpp;    /* grid indices [alo, alo+G), shift org + index*s, bytes per pixel */
  int fxc, fyc;            /* field centre */
  int ox, oy;              /* search offset */
};

/* rank of grid point (a, b) in the reference's visit order */
__device__ __forceinline__ int vsg_grid_rank(const VSGTile& t, int a, int b) {
  return t.bpp == 1 ? vsg_rank(a, b) : 1 + a * t.G + b;
}

/* ---- SAD grid search, one lane ------------------------------------------------
   Lane lid = (u, v) of a job owns the candidates a = a0 + m*j (j < J) and
   b = b0 + k (k < K) of the G x G grid, a0 = alo + (u % m) + m*J*(u / m),
   b0 = alo + K*v.  m*step*bpp is a multiple of 4, so the J candidates of a
   lane sit D = m*step*bpp/4 words apart in the same prev row: one load of a
   prev row segment feeds J*K candidates, each prev word is shifted into the
   field's alignment once (v_alignbyte), and the current field's words come in
   as wave-uniform scalar loads (SGPR operand of v_sad_u8).  Iterating over
   prev rows p, candidate k pairs prev row p with field row p - k*step.  A
   packed pixel is bpp consecutive bytes; the SAD runs over all of them
   (src/motiondetect.c:984-990).
   vsg_sad_lane returns the lane's best (sad << 32 | rank), or ~0 for idle
   lanes; vsg_sad_acc sums a range of prev rows only (the fallback kernel
   splits a lane's rows over several waves), vsg_lane_best ranks the sums. */
template <int J, int K, int D, int CW, bool LDSP>
struct VSGLane {
  int a0, b0;          /* the lane's first candidate */
  bool ok;             /* the lane owns candidates */
  uint32_t acc[K][J];  /* SAD sums */
};

/* the lane's partial sums over prev rows [p0, p1) of every column chunk */
template <int J, int K, int D, int CW, bool LDSP>
__device__ __forceinline__ void vsg_sad_acc(const VSGTile& t, int lid, int p0, int p1, VSGLane<J, K, D, CW, LDSP>& L) {
  const int S = t.S, s = t.s, m = t.m;
  const int u = lid % t.U;
  int v = lid / t.U;
  L.ok = v < t.V;
  if (!L.ok) v = t.V - 1;
  const int a0 = t.alo + (u % m) + m * J * (u / m);
  const int b0 = t.alo + K * v;
  L.a0 = a0;
  L.b0 = b0;
  const int prow0 = t.fyc - S / 2 + t.org + b0 * s + t.oy;
  const int pcol0 = (t.fxc - S / 2 + t.org + a0 * s + t.ox) * t.bpp;
  const int wordsPerRow = (S * t.bpp) >> 2;
  const int pitchWords = t.pitch >> 2;

#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int j = 0; j < J; ++j) L.acc[k][j] = 0;

  constexpr int NP = CW + D * (J - 1);   /* shifted prev words per row segment */
  const int nchunks = wordsPerRow / CW;
  for (int c = 0; c < nchunks; ++c) {
    const uintptr_t start =
        reinterpret_cast<uintptr_t>(t.prev + (size_t)(prow0 + p0) * t.pitch + pcol0 + 4 * CW * c);
    vsg_gword_ptr wp = (vsg_gword_ptr)(start & ~(uintptr_t)3);
    const uint32_t sh = (uint32_t)(start & 3);
    vsg_cword_ptr cbase = t.patch + CW * c;
    const uint32_t* lbase = t.lpatch + CW * c;
    for (int p = p0; p < p1; ++p) {
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
          if constexpr (LDSP) {
            /* the chunk's field words (same address in every lane: broadcast
               32-bit reads), all issued before the SADs that use them */
            const uint32_t* lrow = lbase + y * wordsPerRow;
            uint32_t cw[CW];
#pragma unroll
            for (int w = 0; w < CW; ++w) cw[w] = lrow[w];
#pragma unroll
            for (int w = 0; w < CW; ++w)
#pragma unroll
              for (int j = 0; j < J; ++j) L.acc[k][j] = __builtin_amdgcn_sad_u8(P[w + D * j], cw[w], L.acc[k][j]);
          } else {
            /* the same from wave-uniform scalar loads (SGPR operands) */
            vsg_cword_ptr crow = cbase + y * wordsPerRow;
            uint32_t cw[CW];
#pragma unroll
            for (int w = 0; w < CW; ++w) cw[w] = crow[w];
#pragma unroll
            for (int w = 0; w < CW; ++w)
#pragma unroll
              for (int j = 0; j < J; ++j) L.acc[k][j] = __builtin_amdgcn_sad_u8(P[w + D * j], cw[w], L.acc[k][j]);
          }
        }
      }
      wp += pitchWords;
    }
  }
}

/* the lane's best (sad << 32 | rank) over its complete sums, ~0 for idle lanes */
template <int J, int K, int D, int CW, bool LDSP>
__device__ __forceinline__ unsigned long long vsg_lane_best(const VSGTile& t, const VSGLane<J, K, D, CW, LDSP>& L) {
  const int ahi = t.alo + t.G - 1;
  unsigned long long best = ~0ull;
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int ga = L.a0 + t.m * j, gb = L.b0 + k;
      if (L.ok && ga <= ahi && gb <= ahi) {
        const unsigned long long key = ((unsigned long long)L.acc[k][j] << 32) | (uint32_t)vsg_grid_rank(t, ga, gb);
        best = key < best ? key : best;
      }
    }
  return best;
}

template <int J, int K, int D, int CW, bool LDSP>
__device__ __forceinline__ unsigned long long vsg_sad_lane(const VSGTile& t, int lid) {
  VSGLane<J, K, D, CW, LDSP> L;
  vsg_sad_acc<J, K, D, CW, LDSP>(t, lid, 0, t.S + (K - 1) * t.s, L);
  return vsg_lane_best<J, K, D, CW, LDSP>(t, L);
}

/* minimum over the 64 lanes of a wave */
__device__ __forceinline__ unsigned long long vsg_wave_min_u64(unsigned long long best) {
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)best, off);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(best >> 32), off);
    const unsigned long long o = ((unsigned long long)hi << 32) | lo;
    best = o < best ? o : best;
  }
  return best;
}

#endif /* VSG_DEVICE_H */
