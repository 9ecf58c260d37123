/*
 * vsg_common.h -- small helpers shared by the gfx950 kernels and the host
 * side of the thin C ABI (compiled by hipcc for both).
 */
#ifndef VSG_COMMON_H
#define VSG_COMMON_H

#include <stdint.h>

#if defined(__HIPCC__)
#define VSG_HD __host__ __device__ __forceinline__
#else
#define VSG_HD static inline
#endif

/* Position of grid point (a, b) in the reference's outgoing spiral
   (src/motiondetect.c:466-518): ring k = max(|a|,|b|) occupies ranks
   (2k-1)^2 .. (2k+1)^2-1 and is walked right column downwards, bottom row
   leftwards, left column upwards, top row rightwards, starting at (k, 1-k). */
VSG_HD int vsg_rank(int a, int b) {
  int aa = a < 0 ? -a : a, bb = b < 0 ? -b : b;
  int k = aa > bb ? aa : bb;
  if (k == 0) return 0;
  int base = (2 * k - 1) * (2 * k - 1);
  if (a == k && b > -k) return base + (b + k - 1);
  if (b == k) return base + 2 * k - 1 + (k - a);
  if (a == -k) return base + 4 * k - 1 + (k - b);
  return base + 6 * k - 1 + (a + k);
}

/* inverse of vsg_rank */
VSG_HD void vsg_unrank(int r, int* a, int* b) {
  if (r <= 0) { *a = 0; *b = 0; return; }
  int k = 1;
  while ((2 * k + 1) * (2 * k + 1) <= r) ++k;     /* k <= ~500 in practice */
  int t = r - (2 * k - 1) * (2 * k - 1);
  int seg = t / (2 * k), pos = t - seg * (2 * k);
  switch (seg) {
    case 0: *a = k;           *b = pos + 1 - k; break;
    case 1: *a = k - 1 - pos; *b = k;           break;
    case 2: *a = -k;          *b = k - 1 - pos; break;
    default:*a = pos + 1 - k; *b = -k;          break;
  }
}

/* exact floor(acc / size) for acc <= 255*size and 2 <= size < 4096:
   q = mulhi(acc, ceil(2^32/size)).  With e = mul*size - 2^32 < size the error
   term acc*e/2^32/size stays below 1/size, which cannot carry past the next
   integer because the remainder is at most size-1. */
VSG_HD uint32_t vsg_div_magic(int size) {
  return (uint32_t)((((uint64_t)1 << 32) + (uint64_t)size - 1) / (uint64_t)size);
}

/* XCD-aware block mapping.  The hardware hands workgroups to the 8 XCDs
   round-robin (block b runs on XCD b % 8); units of work are grouped in
   chunks and every chunk is placed on a single XCD (XCD x takes chunks x,
   x+8, ...), so the data a chunk shares stays in that XCD's L2.
   vsg_xcd_unit: the unit of block b, or -1 past nunits; vsg_xcd_blocks: the
   grid size that covers nunits. */
#define VSG_XCDS 8
VSG_HD int vsg_xcd_unit(int b, int chunk, int nunits) {
  const int x = b % VSG_XCDS, i = b / VSG_XCDS;
  const int k = x + VSG_XCDS * (i / chunk);
  const int u = k * chunk + (i % chunk);
  return u < nunits ? u : -1;
}
VSG_HD int vsg_xcd_blocks(int nunits, int chunk) {
  const int nchunks = (nunits + chunk - 1) / chunk;
  return VSG_XCDS * ((nchunks + VSG_XCDS - 1) / VSG_XCDS) * chunk;
}

/* grid geometry of one search stage.  Candidate (a, b), a and b in
   [alo, alo + G), is the shift (org + a*step, org + b*step) px.
   Planar frames: the spiral's square, alo = -n, org = 0, ranked by vsg_rank.
   Packed frames (bpp 3 or 4): the raster i = -maxShift, -maxShift + step, ...
   <= maxShift (src/motiondetect.c:618-619), alo = 0, org = -maxShift, ranked
   1 + a*G + b (x outer, y inner); rank 0 is the unshifted position, which the
   reference evaluates first (:613). */
typedef struct {
  int S;        /* field edge, multiple of 16 */
  int step;     /* grid step in px */
  int n;        /* planar: grid half width, maxShift / step */
  int alo, G;   /* first grid index and grid points per axis */
  int org;      /* shift of grid index 0 in px */
  int bpp;      /* bytes per pixel: 1 planar luma, 3 or 4 packed */
  int m;        /* grid columns between a lane's horizontal candidates */
  int U, V;     /* lanes per job along x and y */
  int J, K;     /* candidates per lane along x and y */
  int D;        /* words between a lane's horizontal candidates (m*step*bpp/4) */
  int CW;       /* curr words per column chunk */
  int wg;       /* threads per workgroup of the grid kernel (multiple of 64) */
  int lds;      /* 1: the field patch is staged in LDS (VGPR operand), 0: scalar loads (SGPR operand) */
} VSGGrid;

#endif /* VSG_COMMON_H */
