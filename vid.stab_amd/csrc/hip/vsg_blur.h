/*
 * vsg_blur.h -- the lane body of the box blur kernel (k_blur_walk in
 * vsg_kernels.hip): the reference's two running-sum passes, boxblur_hori_C
 * then boxblur_vert_C (src/boxblur.c:190-332), with edge replication and an
 * 8-bit floor after each pass.
 *
 * One wave walks one (plane, 1024-px strip, band of rows) unit down the
 * rows; lane l owns the 16 columns x = strip*1024 + 16*l .. x+15 and keeps
 * everything of them in registers:
 *   row pass     the 16 + 2*S2 source bytes around its columns (the 16-byte
 *                chunk plus S2 <= 4*HW bytes each side, loaded as dwords;
 *                columns outside the plane replicate the edge pixel), a
 *                sliding 32-bit sum, and the floor division;
 *   column pass  a ring of the last SIZE = 2*S2+1 row-pass results (16
 *                bytes each), and the 16 column sums as packed 16-bit pairs
 *                (v_pk_add_u16 / v_pk_sub_u16 update two columns at once;
 *                a sum is at most 255*SIZE < 65536).
 * The walk covers rows yFirst-S2 .. yEnd-1+S2 (clamped to the plane: the
 * reference's replication of the first and last row) and stores rows
 * yFirst .. yEnd-1 once SIZE rows are summed.  No lane reads another
 * lane's data, so the body runs the same for any subset of lanes.
 *
 * Division: floor(acc / size) = (acc * M) >> 24 with M = ceil(2^24 / size),
 * exact for every acc <= 255*size when size <= 256 (acc*M < 2^32, and with
 * e = M*size - 2^24 < size the error acc*e/2^24 stays below 1); the product
 * is one v_mul_u32_u24 and the quotient is its top byte, which the packing
 * picks with v_perm_b32.
 *
 * Memory access goes through the policy class P (the kernel's plain loads,
 * or a bounds-checking host emulation in tests): P::ld32 / P::ld128 read a
 * dword / 16 bytes at any byte address, P::ld8 one byte, P::st128 writes 16
 * bytes, P::perm is v_perm_b32 (bytes 0-3 of the 64-bit {a, b} from b,
 * 4-7 from a; selector 0x0c gives 0), P::mul24 is v_mul_u32_u24.
 */
#ifndef VSG_BLUR_H
#define VSG_BLUR_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define VSG_BLUR_FN __host__ __device__ __forceinline__
#else
#define VSG_BLUR_FN inline
#endif

typedef unsigned short vsg_u16x2 __attribute__((ext_vector_type(2)));

constexpr int kBlurStrip = 1024;   /* columns per wave: 64 lanes x 16 */
constexpr int kBlurMaxS2 = 8;      /* boxes up to 17 px; wider ones take the two-kernel path */

/* exact floor(acc / size) multiplier, see above */
VSG_BLUR_FN uint32_t vsg_blur_mul(int size) { return (uint32_t)(((1u << 24) + (uint32_t)size - 1u) / (uint32_t)size); }

/* launch geometry of k_blur_walk over n planes of W x H: units (plane,
   band of R rows, strip) four to a 256-thread workgroup, the workgroups of
   a plane consecutive; about 2,048 waves in all (two per SIMD), bands of at
   least 8 rows; chunk = workgroups kept on one XCD (a plane's, or a share of
   one when there are fewer than 8 planes) */
typedef struct {
  int strips, R, bands;
  int blocksPerPlane;   /* workgroups of one plane */
  int units;            /* workgroups of all planes */
  int chunk;
} VSGBlurPlan;

VSG_BLUR_FN VSGBlurPlan vsg_blur_plan(int W, int H, int n) {
  VSGBlurPlan p;
  p.strips = (W + kBlurStrip - 1) / kBlurStrip;
  int want = (2048 + n * p.strips - 1) / (n * p.strips);   /* bands per plane */
  if (want < 1) want = 1;
  p.R = (H + want - 1) / want;
  if (p.R < 8) p.R = 8;
  p.bands = (H + p.R - 1) / p.R;
  p.blocksPerPlane = (p.strips * p.bands + 3) / 4;
  p.units = n * p.blocksPerPlane;
  const int perPlaneXcds = n >= 8 ? 1 : (8 + n - 1) / n;
  p.chunk = (p.blocksPerPlane + perPlaneXcds - 1) / perPlaneXcds;
  return p;
}

template <int S2, class P>
struct VSGBlurLane {
  static constexpr int SIZE = 2 * S2 + 1;
  static constexpr int HW = (S2 + 3) / 4;   /* halo dwords each side */
  static constexpr int ND = 4 + 2 * HW;     /* dwords of a source row segment */

  /* dword of row pixels pos .. pos+3, edge pixels replicated */
  static VSG_BLUR_FN uint32_t edge32(const uint8_t* row, int pos, int W) {
    uint32_t v = 0;
    for (int j = 0; j < 4; ++j) {
      int p = pos + j;
      p = p < 0 ? 0 : (p > W - 1 ? W - 1 : p);
      v |= (uint32_t)P::ld8(row + p) << (8 * j);
    }
    return v;
  }

  /* Where a lane's row segment comes from, fixed for the whole walk.  For
     widths that are multiples of 16 (every workload size) the 16-byte chunk
     at x lies inside the row, and a halo dword either does too or lies
     entirely beyond one edge: it is then read at the nearest in-row dword
     and its bytes are replaced by the edge pixel with one v_perm_b32
     (selector 0x00000000: pixel 0 at the left, 0x03030303: pixel W-1 at the
     right; 0x03020100 keeps the dword).  Other widths take the generic
     loader, which builds every dword that leaves the row from clamped
     bytes. */
  struct Src {
    int off[2 * HW];         /* byte offsets of the halo dwords in the row */
    uint32_t sel[2 * HW];
  };

  static VSG_BLUR_FN Src halo_plan(int x, int W) {
    Src h;
#pragma unroll
    for (int i = 0; i < HW; ++i) {
      const int pl = x - 4 * HW + 4 * i, pr = x + 16 + 4 * i;
      h.off[i] = pl < 0 ? 0 : pl;
      h.sel[i] = pl < 0 ? 0x00000000u : 0x03020100u;
      h.off[HW + i] = pr + 4 > W ? W - 4 : pr;
      h.sel[HW + i] = pr + 4 > W ? 0x03030303u : 0x03020100u;
    }
    return h;
  }

  /* d[i] = pixels x - 4*HW + 4*i .. +3 of one source row (width % 16 == 0) */
  static VSG_BLUR_FN void load_fast(const uint8_t* row, int x, const Src& h, uint32_t (&d)[ND]) {
    uint32_t m[4];
    P::ld128(row + x, m);
#pragma unroll
    for (int i = 0; i < 4; ++i) d[HW + i] = m[i];
#pragma unroll
    for (int i = 0; i < HW; ++i) {
      d[i] = P::ld32(row + h.off[i]);
      d[HW + 4 + i] = P::ld32(row + h.off[HW + i]);
    }
  }

  /* the same for any width >= 16 */
  static VSG_BLUR_FN void load_any(const uint8_t* row, int x, int W, uint32_t (&d)[ND]) {
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      const int pos = x - 4 * HW + 4 * i;
      d[i] = pos >= 0 && pos + 4 <= W ? P::ld32(row + pos) : edge32(row, pos, W);
    }
  }

  /* the halo bytes of a fast load that lie beyond an edge */
  static VSG_BLUR_FN void fix_halo(const Src& h, uint32_t (&d)[ND]) {
#pragma unroll
    for (int i = 0; i < HW; ++i) {
      d[i] = P::perm(d[i], d[i], h.sel[i]);
      d[HW + 4 + i] = P::perm(d[HW + 4 + i], d[HW + 4 + i], h.sel[HW + i]);
    }
  }

  /* top bytes of four products -> one dword, in column order */
  static VSG_BLUR_FN uint32_t pack4(uint32_t p0, uint32_t p1, uint32_t p2, uint32_t p3) {
    const uint32_t lo = P::perm(p1, p0, 0x0c0c0703u);   /* [p0.b3, p1.b3, 0, 0] */
    const uint32_t hi = P::perm(p3, p2, 0x0c0c0703u);   /* [p2.b3, p3.b3, 0, 0] */
    return P::perm(hi, lo, 0x05040100u);                /* [lo.b0, lo.b1, hi.b0, hi.b1] */
  }

  /* the row pass of one source row: 16 outputs as 4 dwords */
  static VSG_BLUR_FN void row_pass(const uint32_t (&d)[ND], uint32_t M, uint32_t (&h)[4]) {
    constexpr int NV = 16 + 2 * S2;
    uint32_t v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const int b = 4 * HW - S2 + k;   /* byte of d[] holding pixel x - S2 + k */
      v[k] = (d[b >> 2] >> (8 * (b & 3))) & 0xffu;
    }
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < SIZE; ++k) acc += v[k];
    uint32_t p[16];
    p[0] = P::mul24(acc, M);
#pragma unroll
    for (int j = 1; j < 16; ++j) {
      acc += v[j + 2 * S2] - v[j - 1];
      p[j] = P::mul24(acc, M);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) h[k] = pack4(p[4 * k], p[4 * k + 1], p[4 * k + 2], p[4 * k + 3]);
  }

  /* the column sums' update with one new and one leaving row-pass row:
     A[2k] holds columns (4k, 4k+2), A[2k+1] columns (4k+1, 4k+3) */
  static VSG_BLUR_FN void col_update(vsg_u16x2 (&A)[8], const uint32_t (&hn)[4], const uint32_t (&ho)[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t n02 = P::perm(0u, hn[k], 0x0c020c00u), n13 = P::perm(0u, hn[k], 0x0c030c01u);
      const uint32_t o02 = P::perm(0u, ho[k], 0x0c020c00u), o13 = P::perm(0u, ho[k], 0x0c030c01u);
      A[2 * k] = (A[2 * k] + as_pair(n02)) - as_pair(o02);
      A[2 * k + 1] = (A[2 * k + 1] + as_pair(n13)) - as_pair(o13);
    }
  }

  static VSG_BLUR_FN vsg_u16x2 as_pair(uint32_t w) { return __builtin_bit_cast(vsg_u16x2, w); }

  /* the column pass's output row: 16 bytes as 4 dwords */
  static VSG_BLUR_FN void col_out(const vsg_u16x2 (&A)[8], uint32_t M, uint32_t (&o)[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t p0 = P::mul24(A[2 * k].x, M), p2 = P::mul24(A[2 * k].y, M);
      const uint32_t p1 = P::mul24(A[2 * k + 1].x, M), p3 = P::mul24(A[2 * k + 1].y, M);
      o[k] = pack4(p0, p1, p2, p3);
    }
  }

  static VSG_BLUR_FN const uint8_t* row_ptr(const uint8_t* src, int srcPitch, int r, int H) {
    return src + (size_t)(r < 0 ? 0 : (r > H - 1 ? H - 1 : r)) * srcPitch;
  }

  template <bool FAST>
  static VSG_BLUR_FN void load(const uint8_t* row, int x, int W, const Src& h, uint32_t (&d)[ND]) {
    if (FAST)
      load_fast(row, x, h, d);
    else
      load_any(row, x, W, d);
  }

  /* one lane's walk: output rows [yFirst, yEnd) of columns x .. x+15
     (x < W; FAST needs W % 16 == 0).  Source rows are loaded two steps
     ahead of their arithmetic.  The walk is padded to whole rounds of the
     ring (SIZE steps) so that every step issues its load unconditionally
     (rows past the plane read its last row; their results are not stored):
     the compiler then waits only for the row it is about to use. */
  template <bool FAST>
  static VSG_BLUR_FN void walk(const uint8_t* src, int srcPitch, uint8_t* dst, int dstPitch, int W, int H, int x,
                               int yFirst, int yEnd, uint32_t M) {
    vsg_u16x2 A[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) A[k] = as_pair(0u);
    uint32_t ring[SIZE][4];
#pragma unroll
    for (int s = 0; s < SIZE; ++s)
#pragma unroll
      for (int k = 0; k < 4; ++k) ring[s][k] = 0;
    const Src h = halo_plan(x, W);
    const int r0 = yFirst - S2;
    const int steps = (yEnd - yFirst) + 2 * S2;   /* source rows walked, >= SIZE */
    uint32_t cur[ND], nx1[ND];
    load<FAST>(row_ptr(src, srcPitch, r0, H), x, W, h, cur);
    load<FAST>(row_ptr(src, srcPitch, r0 + 1, H), x, W, h, nx1);
    for (int k0 = 0; k0 < steps; k0 += SIZE) {
#pragma unroll
      for (int s = 0; s < SIZE; ++s) {
        const int k = k0 + s;
        uint32_t nx2[ND];
        load<FAST>(row_ptr(src, srcPitch, r0 + k + 2, H), x, W, h, nx2);
        if (FAST) fix_halo(h, cur);
        uint32_t hr[4];
        row_pass(cur, M, hr);
        col_update(A, hr, ring[s]);
#pragma unroll
        for (int q = 0; q < 4; ++q) ring[s][q] = hr[q];
        if (k >= 2 * S2 && k < steps) {
          uint32_t o[4];
          col_out(A, M, o);
          P::st128(dst + (size_t)(r0 + k - S2) * dstPitch + x, o);
        }
#pragma unroll
        for (int q = 0; q < ND; ++q) {
          cur[q] = nx1[q];
          nx1[q] = nx2[q];
        }
      }
    }
  }
};

#endif /* VSG_BLUR_H */
