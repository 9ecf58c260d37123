/*
 * vsg_lens.hip -- the second pass's lens model on the GPU: the per-frame
 * similarity (or rotational) fit through a radial lens that the clip-wide
 * distortion estimate minimises over, its outlier masks, and the per-frame
 * lens transforms (reference: src/lensdistortion.c).
 *
 * One 256-thread workgroup per frame, one kernel per mode and camera model
 * (similarity, or rotational when a focal length is given).  The frame's
 * points sit in LDS, in arrays as long as the clip's largest frame (at most
 * 2,048 LocalMotions): the field position p and the observed destination
 * q = p + v (integers), the active flag, the undistorted, centred source a,
 * and in the masking modes a residual per point and the sort buffer; thread
 * t handles points j = t, t + 256, ....  A 1080p clip's frames (about 300
 * points) take 8 KB in the objective's mode, so many frames share a CU and a
 * clip of a few hundred frames runs in one round.
 * Sums over a frame's points (means, the closed-form start, the 4x4
 * Gauss-Newton normal equations, the residual) are reduced across the wave
 * by shuffles and across the 4 waves in LDS, in a fixed order, and every
 * thread then carries the frame's scalar state (parameters, solve) alike,
 * in registers.
 * Outlier masks take the median and the median absolute deviation of the
 * active points' residuals: the residuals are sorted in LDS by a bitonic
 * network (+inf pads to a power of two).
 * The per-frame outputs go straight into pinned host memory (mapped into
 * the device's address space): an evaluation of the estimator's objective
 * is one launch and one stream synchronisation.
 *
 * The reference sums sequentially in point order, so frame residuals agree
 * with it to rounding, not bit for bit; the estimate's Brent search (host,
 * vs_lensdistortion.c) then converges to the reference's k within its
 * tolerance.
 *
 * Per-frame modes (one launch over all frames of the clip):
 *   OBJ    the fit at strength k with the current mask: r[f] = RMS residual
 *          (vsLensFitSimilarityFov, :202-365; the estimator's objective
 *          lensObjective, :455-478, sums r^2 n over frames on the host)
 *   MASK   the masking pass of the estimator (:562-581): fit at k with the
 *          mask and the similarity model, residuals of every point
 *          (vsLensMatchResiduals, :695-723), then lensMaskOutliers (:412-453)
 *   FINAL  vsLensMotionsToTransform (:725-786): match-based mask, fit,
 *          residual-based mask, refit
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string.h>

#include <new>
#include <vector>

#include "transformtype.h"
#include "vsg_context.h"

namespace {

constexpr int kLensThreads = 256;
constexpr int kLensWaves = kLensThreads / 64;
constexpr int kLensCap = 2048;                            /* points per frame */

enum { LENS_OBJ = 0, LENS_MASK = 1, LENS_FINAL = 2 };
enum { LENS_OK = 0, LENS_FAIL = 1, LENS_SKIP = 2 };

struct LensArgs {
  const short4* lm;        /* per motion: field x, y, motion x, y */
  const double* match;     /* per motion */
  const int* first;        /* per frame: first motion */
  const int* count;        /* per frame: motions, at most cap */
  int nframes;
  int cap;                 /* the clip's largest frame: the LDS arrays' length */
  int np;                  /* cap rounded up to a power of two: the sort buffer's length */
  uint8_t* flags;          /* per motion: active (estimator mask); null: all active */
  double k, f, cx, cy, rho;
  int gn, reject;
  double stddevs;
  double* r;               /* OBJ / FINAL: per frame residual (host memory) */
  int* status;             /* per frame (host memory) */
  int* cut;                /* MASK: per frame, points newly masked (host memory) */
  VSTransform* out;        /* FINAL (host memory) */
};

struct Lens {
  double k, cx, cy, rho;
};


/* vsLensUndistortPoint (:47-63); false outside the model's domain */
__device__ bool undistort(const Lens& ld, double xi, double yi, double* xo, double* yo) {
  const double dx = xi - ld.cx, dy = yi - ld.cy;
  const double r2 = (dx * dx + dy * dy) / (ld.rho * ld.rho);
  const double denom = 1.0 + ld.k * r2;
  if (denom <= 0) return false;
  *xo = ld.cx + dx / denom;
  *yo = ld.cy + dy / denom;
  return true;
}

/* vsLensDistortPoint (:65-82) */
__device__ bool distort(const Lens& ld, double xi, double yi, double* xo, double* yo) {
  const double dx = xi - ld.cx, dy = yi - ld.cy;
  const double r2 = (dx * dx + dy * dy) / (ld.rho * ld.rho);
  const double disc = 1.0 - 4.0 * ld.k * r2;
  if (disc < 0) return false;
  const double scale = 2.0 / (1.0 + sqrt(disc));
  *xo = ld.cx + dx * scale;
  *yo = ld.cy + dy * scale;
  return true;
}

/* lensDistortJacobian (:88-109) at the centred point (wx, wy) */
__device__ bool distort_jacobian(const Lens& ld, double wx, double wy, double J[4]) {
  const double inv = 1.0 / ld.rho;
  const double ux = wx * inv, uy = wy * inv;
  const double t = ux * ux + uy * uy;
  const double disc = 1.0 - 4.0 * ld.k * t;
  if (disc < 0) return false;
  const double root = sqrt(disc);
  const double g = 2.0 / (1.0 + root);
  const double gp = (root > 1e-12) ? (4.0 * ld.k) / ((1.0 + root) * (1.0 + root) * root) : 0.0;
  J[0] = g + 2.0 * gp * ux * ux;
  J[1] = 2.0 * gp * ux * uy;
  J[2] = 2.0 * gp * uy * ux;
  J[3] = g + 2.0 * gp * uy * uy;
  return true;
}

/* rotation_matrix_backward (src/transformtype.c:133-152): Ry(-yaw) Rx(pitch) Rz(roll) */
__device__ void rotation_backward(double yaw, double pitch, double roll, double r[9]) {
  const double cy = cos(-yaw), sy = sin(-yaw), cp = cos(pitch), sp = sin(pitch), cr = cos(roll), sr = sin(roll);
  const double a[9] = {cy, sy * sp, sy * cp, 0.0, cp, -sp, -sy, cy * sp, cy * cp};   /* Ry Rx */
  const double b[9] = {cr, -sr, 0.0, sr, cr, 0.0, 0.0, 0.0, 1.0};                     /* Rz */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
}

/* LensFwd (:138-196): the camera's forward map on a centred point */
struct Fwd {
  double f, c, s, tx, ty, z;
  double rf[9];
};

__device__ void fwd_init(Fwd& w, double f, double tx, double ty, double c, double s) {
  w.f = f;
  w.c = c;
  w.s = s;
  w.tx = tx;
  w.ty = ty;
  w.z = 1.0;
  for (int i = 0; i < 9; ++i) w.rf[i] = (i % 4 == 0) ? 1.0 : 0.0;
  if (f > 0.0) {
    double b[9];
    w.z = sqrt(c * c + s * s);
    rotation_backward(tx / f, ty / f, atan2(s, c), b);
    w.rf[0] = b[0]; w.rf[1] = b[3]; w.rf[2] = b[6];
    w.rf[3] = b[1]; w.rf[4] = b[4]; w.rf[5] = b[7];
    w.rf[6] = b[2]; w.rf[7] = b[5]; w.rf[8] = b[8];
  }
}

struct LensLds {
  double red[kLensWaves][16];
  Fwd fw[5];   /* rotational model: the current parameters' forward map and its four perturbations */
  int cnt;
};

/* the camera's forward map on a centred point; ROT: the rotational model
   (f > 0), else the similarity */
template <bool ROT>
__device__ void fwd_apply(const Fwd& w, double ax, double ay, double* wx, double* wy) {
  if constexpr (ROT) {
    const double X = w.rf[0] * ax + w.rf[1] * ay + w.rf[2] * w.f;
    const double Y = w.rf[3] * ax + w.rf[4] * ay + w.rf[5] * w.f;
    const double Z = w.rf[6] * ax + w.rf[7] * ay + w.rf[8] * w.f;
    *wx = w.z * w.f * X / Z;
    *wy = w.z * w.f * Y / Z;
  } else {
    *wx = w.c * ax + w.s * ay + w.tx;
    *wy = -w.s * ax + w.c * ay + w.ty;
  }
}

/* the similarity's forward map, held in registers by every thread */
__device__ Fwd fwd_similarity(double tx, double ty, double c, double s) {
  Fwd w;
  w.f = 0.0;
  w.c = c;
  w.s = s;
  w.tx = tx;
  w.ty = ty;
  w.z = 1.0;
  return w;
}

/* the sum over the workgroup of K values per thread, the same in every
   thread: wave shuffles, then the wave totals in wave order */
template <int K>
__device__ void block_sum(double (&v)[K], LensLds& L) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) L.red[wave][k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kLensWaves; ++w) t += L.red[w][k];
    v[k] = t;
  }
  __syncthreads();
}

__device__ int block_count(int c, LensLds& L) {
  double v[1] = {(double)c};
  block_sum<1>(v, L);
  return (int)v[0];
}

/* lensSolve4 (:111-152): Gaussian elimination with partial pivoting (the
   first largest pivot candidate wins).  Every index is a compile-time
   constant once the loops unroll and rows are exchanged under uniform
   branches, so the system stays in registers. */
__device__ bool solve4(double A[16], double b[4], double x[4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int p = i;
    double best = fabs(A[i * 4 + i]);
#pragma unroll
    for (int j = i + 1; j < 4; ++j) {
      const double v = fabs(A[j * 4 + i]);
      if (v > best) {
        best = v;
        p = j;
      }
    }
    if (best < 1e-12) return false;
#pragma unroll
    for (int j = i + 1; j < 4; ++j) {
      if (p == j) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double t = A[i * 4 + q];
          A[i * 4 + q] = A[j * 4 + q];
          A[j * 4 + q] = t;
        }
        const double t = b[i];
        b[i] = b[j];
        b[j] = t;
      }
    }
#pragma unroll
    for (int j = i + 1; j < 4; ++j) {
      const double f = A[j * 4 + i] / A[i * 4 + i];
      if (f == 0) continue;
#pragma unroll
      for (int q = i; q < 4; ++q) A[j * 4 + q] -= f * A[i * 4 + q];
      b[j] -= f * b[i];
    }
  }
#pragma unroll
  for (int i = 3; i >= 0; --i) {
    double sum = b[i];
#pragma unroll
    for (int q = i + 1; q < 4; ++q) sum -= A[i * 4 + q] * x[q];
    x[i] = sum / A[i * 4 + i];
  }
  return true;
}

/* one frame's points in the workgroup's LDS: the field position p and the
   observed destination q = p + v (integer pixels), the undistorted centred
   source a, the active flag; the masking modes add a residual per point and
   the sort buffer */
struct FrameLds {
  short4* pq;      /* p.x, p.y, q.x, q.y */
  double* ax;
  double* ay;
  double* res;
  double* sbuf;
  uint8_t* act;
};

/* bytes of dynamic LDS a mode's workgroup takes for frames of at most cap
   points (np: cap rounded up to a power of two) */
__host__ __device__ constexpr size_t lens_lds_bytes(int mode, int cap, int np) {
  return (size_t)cap * (sizeof(short4) + 2 * sizeof(double) + 1) +
         (mode == LENS_OBJ ? 0 : (size_t)cap * sizeof(double) + (size_t)np * sizeof(double));
}

struct FitOut {
  bool ok;
  double tx, ty, c, s;
  double res2;
  int nAct;
};

__device__ int count_active(int n, const FrameLds& F, LensLds& L) {
  int a = 0;
  for (int j = threadIdx.x; j < n; j += kLensThreads) a += F.act[j];
  return block_count(a, L);
}

/* vsLensFitSimilarityFov (:202-365) over the frame's active points; ROT:
   the rotational model at focal length f */
template <bool ROT>
__device__ FitOut lens_fit(int n, const FrameLds& F, const Lens& ld, int gn, double f, LensLds& L) {
  FitOut o;
  o.ok = false;
  o.res2 = 0.0;
  o.nAct = count_active(n, F, L);
  if (n < 2 || o.nAct < 3) return o;
  /* a: the undistorted, centred sources, fixed throughout; the undistorted
     destinations (for the closed-form start) are recomputed where needed */
  bool bad = false;
  double m[4] = {0.0, 0.0, 0.0, 0.0};   /* sums of ax, ay, ubx, uby */
  for (int j = threadIdx.x; j < n; j += kLensThreads) {
    if (!F.act[j]) continue;
    const short4 pq = F.pq[j];
    double ux, uy, vx, vy;
    if (!undistort(ld, pq.x, pq.y, &ux, &uy) || !undistort(ld, pq.z, pq.w, &vx, &vy)) {
      bad = true;
      continue;
    }
    F.ax[j] = ux - ld.cx;
    F.ay[j] = uy - ld.cy;
    m[0] += F.ax[j];
    m[1] += F.ay[j];
    m[2] += vx - ld.cx;
    m[3] += vy - ld.cy;
  }
  if (__syncthreads_or(bad)) return o;
  block_sum<4>(m, L);
  const double mAx = m[0] / o.nAct, mAy = m[1] / o.nAct, mBx = m[2] / o.nAct, mBy = m[3] / o.nAct;
  double sc[3] = {0.0, 0.0, 0.0};   /* sumRR, num_c, num_s */
  for (int j = threadIdx.x; j < n; j += kLensThreads) {
    if (!F.act[j]) continue;
    const short4 pq = F.pq[j];
    double vx, vy;
    undistort(ld, pq.z, pq.w, &vx, &vy);
    const double dax = F.ax[j] - mAx, day = F.ay[j] - mAy, dbx = (vx - ld.cx) - mBx, dby = (vy - ld.cy) - mBy;
    sc[0] += dax * dax + day * day;
    sc[1] += dax * dbx + day * dby;
    sc[2] += day * dbx - dax * dby;
  }
  block_sum<3>(sc, L);
  if (sc[0] < 1e-12) return o;
  double c = sc[1] / sc[0], s = sc[2] / sc[0];
  double tx = mBx - (c * mAx + s * mAy);
  double ty = mBy - (-s * mAx + c * mAy);
  const double kHT = 1e-3, kHCS = 1e-6;   /* LENS_H_T, LENS_H_CS (:196-199) */
  for (int it = 0; it < gn; ++it) {
    Fwd w0;
    if constexpr (ROT) {
      /* built once per iteration, by one thread, into LDS; the previous
         iteration's readers are past block_sum's barriers */
      if (threadIdx.x == 0) {
        fwd_init(L.fw[0], f, tx, ty, c, s);
        fwd_init(L.fw[1], f, tx + kHT, ty, c, s);
        fwd_init(L.fw[2], f, tx, ty + kHT, c, s);
        fwd_init(L.fw[3], f, tx, ty, c + kHCS, s);
        fwd_init(L.fw[4], f, tx, ty, c, s + kHCS);
      }
      __syncthreads();
    } else {
      w0 = fwd_similarity(tx, ty, c, s);
    }
    double acc[14];   /* g[4], then H's upper triangle row by row */
#pragma unroll
    for (int i = 0; i < 14; ++i) acc[i] = 0.0;
    for (int j = threadIdx.x; j < n; j += kLensThreads) {
      if (!F.act[j]) continue;
      const short4 pq = F.pq[j];
      const double ax = F.ax[j], ay = F.ay[j];
      double wx, wy, dpx, dpy, J[4];
      if constexpr (ROT)
        fwd_apply<true>(L.fw[0], ax, ay, &wx, &wy);
      else
        fwd_apply<false>(w0, ax, ay, &wx, &wy);
      if (!distort(ld, wx + ld.cx, wy + ld.cy, &dpx, &dpy) || !distort_jacobian(ld, wx, wy, J)) {
        bad = true;
        continue;
      }
      const double ex = (pq.z - ld.cx) - (dpx - ld.cx), ey = (pq.w - ld.cy) - (dpy - ld.cy);
      double Jp[4][2];
      if constexpr (ROT) {
        const double H[4] = {kHT, kHT, kHCS, kHCS};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          double pwx, pwy;
          fwd_apply<true>(L.fw[1 + i], ax, ay, &pwx, &pwy);
          const double dwx = (pwx - wx) / H[i], dwy = (pwy - wy) / H[i];
          Jp[i][0] = J[0] * dwx + J[1] * dwy;
          Jp[i][1] = J[2] * dwx + J[3] * dwy;
        }
      } else {
        Jp[0][0] = J[0];
        Jp[0][1] = J[2];
        Jp[1][0] = J[1];
        Jp[1][1] = J[3];
        Jp[2][0] = J[0] * ax + J[1] * ay;
        Jp[2][1] = J[2] * ax + J[3] * ay;
        Jp[3][0] = J[0] * ay - J[1] * ax;
        Jp[3][1] = J[2] * ay - J[3] * ax;
      }
      int h = 4;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        acc[i] += Jp[i][0] * ex + Jp[i][1] * ey;
#pragma unroll
        for (int q = i; q < 4; ++q) acc[h++] += Jp[i][0] * Jp[q][0] + Jp[i][1] * Jp[q][1];
      }
    }
    if (__syncthreads_or(bad)) return o;
    block_sum<14>(acc, L);
    double Hm[16], g[4], delta[4] = {0.0, 0.0, 0.0, 0.0};
    int h = 4;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      g[i] = acc[i];
#pragma unroll
      for (int q = i; q < 4; ++q) Hm[i * 4 + q] = acc[h++];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int q = 0; q < i; ++q) Hm[i * 4 + q] = Hm[q * 4 + i];
    if (!solve4(Hm, g, delta)) break;   /* keep the current estimate */
    tx += delta[0];
    ty += delta[1];
    c += delta[2];
    s += delta[3];
  }
  Fwd w0;
  if constexpr (ROT) {
    if (threadIdx.x == 0) fwd_init(L.fw[0], f, tx, ty, c, s);
    __syncthreads();
  } else {
    w0 = fwd_similarity(tx, ty, c, s);
  }
  double r2[1] = {0.0};
  for (int j = threadIdx.x; j < n; j += kLensThreads) {
    if (!F.act[j]) continue;
    const short4 pq = F.pq[j];
    double wx, wy, dpx, dpy;
    if constexpr (ROT)
      fwd_apply<true>(L.fw[0], F.ax[j], F.ay[j], &wx, &wy);
    else
      fwd_apply<false>(w0, F.ax[j], F.ay[j], &wx, &wy);
    if (!distort(ld, wx + ld.cx, wy + ld.cy, &dpx, &dpy)) {
      bad = true;
      continue;
    }
    const double ex = (pq.z - ld.cx) - (dpx - ld.cx), ey = (pq.w - ld.cy) - (dpy - ld.cy);
    r2[0] += ex * ex + ey * ey;
  }
  if (__syncthreads_or(bad)) return o;
  block_sum<1>(r2, L);
  o.ok = true;
  o.tx = tx;
  o.ty = ty;
  o.c = c;
  o.s = s;
  o.res2 = r2[0];
  return o;
}

/* vsLensMatchResidualsFov (:700-723): |q - D(fwd(U(p)))| of every point
   (masked ones too) into F.res under transform t; false if a point leaves
   the model's domain */
template <bool ROT>
__device__ bool match_residuals(int n, const FrameLds& F, const Lens& ld, double f, const VSTransform& t,
                                LensLds& L) {
  const double z = 1.0 + t.zoom / 100.0;
  Fwd w0;
  if constexpr (ROT) {
    __syncthreads();   /* every thread is done with the forward maps in LDS */
    if (threadIdx.x == 0) fwd_init(L.fw[0], f, t.x, t.y, z * cos(t.alpha), z * sin(t.alpha));
    __syncthreads();
  } else {
    w0 = fwd_similarity(t.x, t.y, z * cos(t.alpha), z * sin(t.alpha));
  }
  bool bad = false;
  for (int j = threadIdx.x; j < n; j += kLensThreads) {
    const short4 pq = F.pq[j];
    double ux, uy, wx, wy, dpx, dpy;
    F.res[j] = 0.0;
    if (!undistort(ld, pq.x, pq.y, &ux, &uy)) {
      bad = true;
      continue;
    }
    if constexpr (ROT)
      fwd_apply<true>(L.fw[0], ux - ld.cx, uy - ld.cy, &wx, &wy);
    else
      fwd_apply<false>(w0, ux - ld.cx, uy - ld.cy, &wx, &wy);
    if (!distort(ld, wx + ld.cx, wy + ld.cy, &dpx, &dpy)) {
      bad = true;
      continue;
    }
    const double ex = pq.z - dpx, ey = pq.w - dpy;
    F.res[j] = sqrt(ex * ex + ey * ey);
  }
  return !__syncthreads_or(bad);
}

/* bitonic sort of sbuf[0, np), np a power of two */
__device__ void sort_lds(double* sbuf, int np) {
  for (int size = 2; size <= np; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < np; i += kLensThreads) {
        const int j = i ^ stride;
        if (j > i) {
          const bool up = (i & size) == 0;
          const double a = sbuf[i], b = sbuf[j];
          if ((a > b) == up) {
            sbuf[i] = b;
            sbuf[j] = a;
          }
        }
      }
      __syncthreads();
    }
}

/* lensMaskOutliers (:412-453) on F.res: masks active points whose residual
   exceeds median + stddevs * 1.4826 * MAD; returns how many were newly
   masked */
__device__ int mask_outliers(int n, const FrameLds& F, double stddevs, LensLds& L) {
  const int m = count_active(n, F, L);
  if (m < 8) return 0;
  if (threadIdx.x == 0) L.cnt = 0;
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += kLensThreads)
    if (F.act[j]) F.sbuf[atomicAdd(&L.cnt, 1)] = F.res[j];
  int np = 1;
  while (np < m) np <<= 1;
  __syncthreads();
  for (int i = m + (int)threadIdx.x; i < np; i += kLensThreads) F.sbuf[i] = __builtin_inf();
  __syncthreads();
  sort_lds(F.sbuf, np);
  const double median = F.sbuf[m / 2];
  __syncthreads();
  for (int i = threadIdx.x; i < m; i += kLensThreads) F.sbuf[i] = fabs(F.sbuf[i] - median);
  __syncthreads();
  sort_lds(F.sbuf, np);
  const double mad = F.sbuf[m / 2];
  __syncthreads();
  if (mad <= 1e-9) return 0;
  const double thresh = median + stddevs * 1.4826 * mad;
  int cut = 0;
  for (int j = threadIdx.x; j < n; j += kLensThreads)
    if (F.act[j] && F.res[j] > thresh) {
      F.act[j] = 0;
      cut++;
    }
  return block_count(cut, L);
}

__device__ VSTransform to_transform(const FitOut& o, double k) {
  const double z = sqrt(o.c * o.c + o.s * o.s);
  VSTransform t;
  t.x = o.tx;
  t.y = o.ty;
  t.alpha = atan2(o.s, o.c);
  t.zoom = (z - 1.0) * 100.0;
  t.barrel = k;
  t.rshutter = 0;
  t.extra = 0;
  return t;
}

/* MODE: LENS_OBJ, LENS_MASK or LENS_FINAL; ROT: the rotational model at
   focal length a.f (the masking pass always fits the similarity) */
template <int MODE, bool ROT>
__global__ __launch_bounds__(kLensThreads) void k_lens_frame(LensArgs a) {
  __shared__ LensLds L;
  extern __shared__ __align__(16) uint8_t lds[];
  const int fr = blockIdx.x;
  if (fr >= a.nframes) return;
  FrameLds F;
  F.pq = reinterpret_cast<short4*>(lds);
  F.ax = reinterpret_cast<double*>(lds + sizeof(short4) * (size_t)a.cap);
  F.ay = F.ax + a.cap;
  if constexpr (MODE == LENS_OBJ) {
    F.res = F.sbuf = nullptr;
    F.act = reinterpret_cast<uint8_t*>(F.ay + a.cap);
  } else {
    F.res = F.ay + a.cap;
    F.sbuf = F.res + a.cap;
    F.act = reinterpret_cast<uint8_t*>(F.sbuf + a.np);
  }
  const int n = a.count[fr];   /* at most a.cap (the host takes the largest) */
  const int first = a.first[fr];
  for (int j = threadIdx.x; j < n; j += kLensThreads) {
    const short4 m = a.lm[first + j];
    F.pq[j] = make_short4(m.x, m.y, (short)(m.x + m.z), (short)(m.y + m.w));
    F.act[j] = (!a.flags || a.flags[first + j]) ? 1 : 0;
  }
  __syncthreads();
  Lens ld;
  ld.k = a.k;
  ld.cx = a.cx;
  ld.cy = a.cy;
  ld.rho = a.rho;
  if constexpr (MODE == LENS_OBJ) {
    int st = LENS_SKIP;
    double r = 0.0;
    if (n >= 3) {
      const FitOut o = lens_fit<ROT>(n, F, ld, a.gn, a.f, L);
      st = o.ok ? LENS_OK : LENS_FAIL;
      if (o.ok) r = sqrt(o.res2 / o.nAct);
    }
    if (threadIdx.x == 0) {
      a.status[fr] = st;
      a.r[fr] = r;
    }
  }
  if constexpr (MODE == LENS_MASK) {
    /* the estimator's masking pass fits the similarity model (f = 0) */
    int cut = 0;
    const FitOut o = lens_fit<false>(n, F, ld, a.gn, 0.0, L);
    if (o.ok && match_residuals<false>(n, F, ld, 0.0, to_transform(o, ld.k), L)) {
      cut = mask_outliers(n, F, a.stddevs, L);
      __syncthreads();
      for (int j = threadIdx.x; j < n; j += kLensThreads) a.flags[first + j] = F.act[j];
    }
    if (threadIdx.x == 0) a.cut[fr] = cut;
  }
  if constexpr (MODE == LENS_FINAL) {
    /* vsLensMotionsToTransform */
    VSTransform t = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0};
    double residual = 0.0;
    int st = LENS_OK;
    if (n < 3) {
      t.extra = 1;
      st = LENS_SKIP;
    } else {
      if (a.reject) {
        /* stage one: the matcher's own confidence */
        for (int j = threadIdx.x; j < n; j += kLensThreads) F.res[j] = a.match[first + j];
        __syncthreads();
        mask_outliers(n, F, a.stddevs, L);
        __syncthreads();
      }
      const FitOut o = lens_fit<ROT>(n, F, ld, a.gn, a.f, L);
      if (!o.ok) {
        t.extra = 1;
        st = LENS_FAIL;
      } else {
        t = to_transform(o, ld.k);
        residual = sqrt(o.res2 / o.nAct);
        /* stage two: the residual at the known k */
        if (a.reject && match_residuals<ROT>(n, F, ld, a.f, t, L) && mask_outliers(n, F, a.stddevs, L) > 0) {
          __syncthreads();
          const FitOut o2 = lens_fit<ROT>(n, F, ld, a.gn, a.f, L);
          if (o2.ok) {
            t = to_transform(o2, ld.k);
            residual = sqrt(o2.res2 / o2.nAct);
          }
        }
        t.barrel = ld.k;
      }
    }
    if (threadIdx.x == 0) {
      a.out[fr] = t;
      a.r[fr] = residual;
      a.status[fr] = st;
    }
  }
}

}  // namespace

/* ---- host side: one clip's motions on the device ------------------------------ */

#define LCHECK(call)                                     \
  do {                                                   \
    hipError_t e_ = (call);                              \
    if (e_ != hipSuccess) return vsg_fail(#call, e_);    \
  } while (0)

struct VSGLens {
  int device = 0;
  int nframes = 0;
  long long total = 0;
  int cap = 1, np = 1;   /* the largest frame's points, and that rounded up to a power of two */
  double cx = 0, cy = 0, rho = 0;
  hipStream_t st = nullptr;
  short4* d_lm = nullptr;
  double* d_match = nullptr;
  int* d_first = nullptr;
  int* d_count = nullptr;
  uint8_t* d_flags = nullptr;
  /* the per-frame outputs: pinned host memory the kernels write through its
     device mapping (m_*) */
  double* h_r = nullptr;
  int* h_status = nullptr;
  int* h_cut = nullptr;
  VSTransform* h_out = nullptr;
  double* m_r = nullptr;
  int* m_status = nullptr;
  int* m_cut = nullptr;
  VSTransform* m_out = nullptr;
};

extern "C" void vsg_lens_destroy(VSGLens* L) {
  if (!L) return;
  (void)hipSetDevice(L->device);
  if (L->st) (void)hipStreamSynchronize(L->st);
  void* d[] = {L->d_lm, L->d_match, L->d_first, L->d_count, L->d_flags};
  for (void* p : d)
    if (p) (void)hipFree(p);
  void* h[] = {L->h_r, L->h_status, L->h_cut, L->h_out};
  for (void* p : h)
    if (p) (void)hipHostFree(p);
  if (L->st) (void)hipStreamDestroy(L->st);
  delete L;
}

template <typename T>
static hipError_t pinned_mapped(T** host, T** dev, size_t n) {
  hipError_t e = hipHostMalloc(reinterpret_cast<void**>(host), sizeof(T) * n, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostGetDevicePointer(reinterpret_cast<void**>(dev), *host, 0);
  return e;
}

extern "C" int vsg_lens_create(VSGLens** out, int device, int width, int height, const int16_t* lm,
                               const double* match, const int* count, int nframes) {
  if (!out || width <= 0 || height <= 0 || nframes <= 0 || !count) return vsg_fail("vsg_lens_create", hipSuccess);
  *out = nullptr;
  std::vector<int> first(nframes);
  long long total = 0;
  int cap = 1;
  for (int f = 0; f < nframes; ++f) {
    if (count[f] < 0 || count[f] > kLensCap)
      return vsg_fail("vsg_lens_create: a frame holds more than 2048 LocalMotions", hipSuccess);
    first[f] = (int)total;
    total += count[f];
    if (count[f] > cap) cap = count[f];
  }
  if (total > 0x7FFFFFFFLL || (total > 0 && (!lm || !match))) return vsg_fail("vsg_lens_create: motions", hipSuccess);
  int np = 1;
  while (np < cap) np <<= 1;
  int lmax = 0;
  LCHECK(hipSetDevice(device));
  LCHECK(hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
  if (lens_lds_bytes(LENS_FINAL, cap, np) + sizeof(LensLds) > (size_t)lmax)
    return vsg_fail("vsg_lens_create: a frame's points do not fit the workgroup's LDS", hipSuccess);
  VSGLens* L = new (std::nothrow) VSGLens();
  if (!L) return vsg_fail("vsg_lens_create: out of host memory", hipSuccess);
  L->device = device;
  L->nframes = nframes;
  L->total = total;
  L->cap = cap;
  L->np = np;
  L->cx = width / 2.0;
  L->cy = height / 2.0;
  L->rho = 0.5 * sqrt((double)width * width + (double)height * height);
  const size_t nm = total > 0 ? (size_t)total : 1;
  hipError_t e = hipStreamCreateWithFlags(&L->st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&L->d_lm, sizeof(short4) * nm);
  if (e == hipSuccess) e = hipMalloc(&L->d_match, sizeof(double) * nm);
  if (e == hipSuccess) e = hipMalloc(&L->d_flags, nm);
  if (e == hipSuccess) e = hipMalloc(&L->d_first, sizeof(int) * nframes);
  if (e == hipSuccess) e = hipMalloc(&L->d_count, sizeof(int) * nframes);
  if (e == hipSuccess) e = pinned_mapped(&L->h_r, &L->m_r, nframes);
  if (e == hipSuccess) e = pinned_mapped(&L->h_status, &L->m_status, nframes);
  if (e == hipSuccess) e = pinned_mapped(&L->h_cut, &L->m_cut, nframes);
  if (e == hipSuccess) e = pinned_mapped(&L->h_out, &L->m_out, nframes);
  if (e == hipSuccess && total > 0)
    e = hipMemcpyAsync(L->d_lm, lm, sizeof(short4) * total, hipMemcpyHostToDevice, L->st);
  if (e == hipSuccess && total > 0)
    e = hipMemcpyAsync(L->d_match, match, sizeof(double) * total, hipMemcpyHostToDevice, L->st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(L->d_first, first.data(), sizeof(int) * nframes, hipMemcpyHostToDevice, L->st);
  if (e == hipSuccess) e = hipMemcpyAsync(L->d_count, count, sizeof(int) * nframes, hipMemcpyHostToDevice, L->st);
  if (e == hipSuccess) e = hipStreamSynchronize(L->st);
  if (e != hipSuccess) {
    vsg_lens_destroy(L);
    return vsg_fail("vsg_lens_create", e);
  }
  *out = L;
  return 0;
}

static LensArgs lens_args(VSGLens* L, double k, double f, int gn) {
  LensArgs a;
  a.lm = L->d_lm;
  a.match = L->d_match;
  a.first = L->d_first;
  a.count = L->d_count;
  a.nframes = L->nframes;
  a.cap = L->cap;
  a.np = L->np;
  a.flags = nullptr;
  a.k = k;
  a.f = f;
  a.cx = L->cx;
  a.cy = L->cy;
  a.rho = L->rho;
  a.gn = gn;
  a.reject = 0;
  a.stddevs = 0.0;
  a.r = L->m_r;
  a.status = L->m_status;
  a.cut = L->m_cut;
  a.out = L->m_out;
  return a;
}

/* one launch over the clip's frames, then the stream's synchronisation: the
   outputs are in the host buffers when it returns 0 */
template <int MODE>
static int lens_run(VSGLens* L, const LensArgs& a, const char* what) {
  const size_t lds = lens_lds_bytes(MODE, a.cap, a.np);
  if constexpr (MODE == LENS_MASK) {
    hipLaunchKernelGGL((k_lens_frame<MODE, false>), dim3(a.nframes), dim3(kLensThreads), lds, L->st, a);
  } else {
    if (a.f > 0.0)
      hipLaunchKernelGGL((k_lens_frame<MODE, true>), dim3(a.nframes), dim3(kLensThreads), lds, L->st, a);
    else
      hipLaunchKernelGGL((k_lens_frame<MODE, false>), dim3(a.nframes), dim3(kLensThreads), lds, L->st, a);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(L->st);
  return e == hipSuccess ? 0 : vsg_fail(what, e);
}

extern "C" int vsg_lens_reset_mask(VSGLens* L) {
  if (!L) return vsg_fail("vsg_lens_reset_mask", hipSuccess);
  LCHECK(hipSetDevice(L->device));
  LCHECK(hipMemsetAsync(L->d_flags, 1, L->total > 0 ? (size_t)L->total : 1, L->st));
  LCHECK(hipStreamSynchronize(L->st));
  return 0;
}

extern "C" int vsg_lens_fit(VSGLens* L, double k, double f, int gnSteps, int useMask, double* r, int* status) {
  if (!L || !r || !status || gnSteps < 0) return vsg_fail("vsg_lens_fit", hipSuccess);
  LCHECK(hipSetDevice(L->device));
  LensArgs a = lens_args(L, k, f, gnSteps);
  if (useMask) a.flags = L->d_flags;
  if (lens_run<LENS_OBJ>(L, a, "vsg_lens_fit")) return -1;
  memcpy(r, L->h_r, sizeof(double) * L->nframes);
  memcpy(status, L->h_status, sizeof(int) * L->nframes);
  return 0;
}

extern "C" int vsg_lens_mask(VSGLens* L, double k, int gnSteps, double stddevs, int* cut) {
  if (!L || !cut || gnSteps < 0) return vsg_fail("vsg_lens_mask", hipSuccess);
  LCHECK(hipSetDevice(L->device));
  LensArgs a = lens_args(L, k, 0.0, gnSteps);
  a.flags = L->d_flags;
  a.stddevs = stddevs;
  if (lens_run<LENS_MASK>(L, a, "vsg_lens_mask")) return -1;
  memcpy(cut, L->h_cut, sizeof(int) * L->nframes);
  return 0;
}

extern "C" int vsg_lens_rejected(VSGLens* L, long long* rejected) {
  if (!L || !rejected) return vsg_fail("vsg_lens_rejected", hipSuccess);
  LCHECK(hipSetDevice(L->device));
  std::vector<uint8_t> fl(L->total > 0 ? (size_t)L->total : 1);
  LCHECK(hipMemcpyAsync(fl.data(), L->d_flags, fl.size(), hipMemcpyDeviceToHost, L->st));
  LCHECK(hipStreamSynchronize(L->st));
  long long n = 0;
  for (long long i = 0; i < L->total; ++i) n += fl[(size_t)i] ? 0 : 1;
  *rejected = n;
  return 0;
}

extern "C" int vsg_lens_transforms(VSGLens* L, double k, double f, int gnSteps, int reject, double stddevs,
                                   VSTransform* out, double* residual, int* status) {
  if (!L || !out || gnSteps < 0) return vsg_fail("vsg_lens_transforms", hipSuccess);
  LCHECK(hipSetDevice(L->device));
  LensArgs a = lens_args(L, k, f, gnSteps);
  a.reject = reject ? 1 : 0;
  a.stddevs = stddevs;
  if (lens_run<LENS_FINAL>(L, a, "vsg_lens_transforms")) return -1;
  memcpy(out, L->h_out, sizeof(VSTransform) * L->nframes);
  if (residual) memcpy(residual, L->h_r, sizeof(double) * L->nframes);
  if (status) memcpy(status, L->h_status, sizeof(int) * L->nframes);
  return 0;
}
