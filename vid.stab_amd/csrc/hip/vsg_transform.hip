/*
 * vsg_transform.hip -- the second pass's per-frame global transform on the
 * GPU: vsMotionsToTransform (src/localmotion2transform.c:213-283) for every
 * frame of a clip in one launch, with the similarity model (fov 0) or the
 * rotational one (fov > 0, src/transformtype.c:123-209), no lens.
 *
 * One wave per frame.  The frame's LocalMotions sit in LDS with the
 * "missmatch" array of the reference (the last error of every enabled field,
 * -1 for disabled ones) and a compacted, ascending list of the enabled
 * fields.  A quality evaluation (calcTransformQuality, :139-171) computes
 * the fields' errors across the lanes, then every lane adds them in field
 * order, reading each 64-field chunk back from LDS -- the reference's loop
 * calls transform_vec_double in another translation unit, so its sum is
 * strictly sequential.  The
 * gradient descent (vsGradientDescent, :288-328) is scalar code that every
 * lane runs identically.  Every double expression is evaluated in the order
 * the reference's -ffast-math x86 build uses (the host restatement in the
 * oracle documents each); the file is compiled without FMA contraction.
 * sin and cos are the device's: they may differ from glibc's in the last
 * bit, which the finite-difference gradient amplifies, so results agree
 * with the reference to a tolerance, not bit for bit.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <vector>

#include "transformtype.h"
#include "vsg_context.h"

namespace {

struct M2TArgs {
  const short4* lm;       /* per motion: field x, y, motion x, y */
  const double* match;    /* per motion */
  const int* first;       /* per frame: index of its first motion (null: summed from count) */
  const int* count;       /* per frame: motions */
  int nframes, cx, cy, smoothZoom, cap;
  double focal;           /* 0: similarity model; > 0: rotation about the optical centre */
  VSTransform* out;
  double* residual;
  int* rounds;
};

struct Frame {
  const short4* lm;
  double* wsum;     /* [64]: one chunk of values on their way to an ordered sum */
  double* miss;     /* [n] */
  double* ecomp;    /* [n]: the match values, for the first disable pass */
  int* eidx;        /* [n]: the enabled fields, ascending */
  int n, ne;
  double cx, cy;
  double f;         /* focal length in px (VSTransformConfig.fov), 0 for the similarity model */
};

/* ascending list of the fields with miss >= 0 (a wave-wide ordered compaction) */
__device__ void compact(Frame& F) {
  const int lane = threadIdx.x;
  int base = 0;
  for (int i0 = 0; i0 < F.n; i0 += 64) {
    const int i = i0 + lane;
    const bool on = i < F.n && F.miss[i] >= 0.0;
    const uint64_t mask = __ballot(on);
    if (on) F.eidx[base + __popcll(mask & ((1ull << lane) - 1ull))] = i;
    base += __popcll(mask);
  }
  F.ne = base;
  __syncthreads();
}

/* acc plus the wave's 64 values, one after another in lane order (lane 0
   first), the same in every lane: the values pass through LDS and every lane
   reads all of them back (broadcast reads) before the dependent adds */
__device__ __forceinline__ double ordered_sum(double acc, double v, double* buf) {
  buf[threadIdx.x] = v;
  __syncthreads();
  const double2* b = reinterpret_cast<const double2*>(buf);
#pragma unroll
  for (int l = 0; l < 32; ++l) {
    const double2 p = b[l];
    acc += p.x;
    acc += p.y;
  }
  __syncthreads();   /* every lane has read buf before it is written again */
  return acc;
}

/* the same for two interleaved sums: even lanes into even, odd lanes into odd */
__device__ __forceinline__ void ordered_pair_sums(double& even, double& odd, double v, double* buf) {
  buf[threadIdx.x] = v;
  __syncthreads();
  const double2* b = reinterpret_cast<const double2*>(buf);
#pragma unroll
  for (int l = 0; l < 32; ++l) {
    const double2 p = b[l];
    even += p.x;
    odd += p.y;
  }
  __syncthreads();
}

/* the error sum of calcTransformQuality under the rotational model
   (VSTransformConfig.fov > 0): prepare_transform_fov and the f > 0 branch of
   transform_vec_double (src/transformtype.c:133-209), every expression in
   the order the reference's -ffast-math x86 build evaluates it.  That build
   computes rotation_matrix_backward's two 3x3 products with the constant
   zeros and ones of Ry, Rx and Rz folded into pairwise sums (the products
   by 0.0 stay: they are not foldable in IEEE arithmetic), takes cos(yaw)
   for cos(-yaw), and sin(-yaw) as such.  Lane l computes enabled field
   j0 + l of each 64-field chunk and the wave adds the errors in field
   order, as in quality(). */
__device__ double quality_fov_sum(Frame& F, const double p[4]) {
  const double f = F.f;
  const double z = p[3] * 0.01 + 1.0;
  const double yaw = p[0] / f, pitch = p[1] / f;
  double sr, cr, sp, cp;
  sincos(p[2], &sr, &cr);
  sincos(pitch, &sp, &cp);
  const double cy = cos(yaw), sy = sin(-yaw), nsy = -sy;
  const double A = cp * 0.0 + sy * sp, B = cp + 0.0 * sp, spcy = sp * cy;
  /* b = Ry(-yaw) Rx(pitch) Rz(roll), row major */
  const double b0 = sr * A + cr * cy, b1 = cr * A - sr * cy, b2 = sy * cp - sp * 0.0;
  const double b3 = sr * B + cr * 0.0, b4 = cr * B - sr * 0.0, b5 = 0.0 * cp - sp;
  const double b6 = nsy * cr + sr * spcy, b7 = cr * spcy - nsy * sr, b8 = cy * cp;
  /* the forward rotation is its transpose */
  const double r0 = b0, r1 = b3, r2 = b6, r3 = b1, r4 = b4, r5 = b7, r6 = b2, r7 = b5, r8 = b8;
  const double fz = f * z;
  double err = 0.0;
  for (int j0 = 0; j0 < F.ne; j0 += 64) {
    const int j = j0 + (int)threadIdx.x;
    double e = 0.0;
    if (j < F.ne) {
      const int i = F.eidx[j];
      const short4 m = F.lm[i];
      const double vx = (double)m.x, vy = (double)m.y;
      const double rx = vx - F.cx, ry = vy - F.cy;
      const double X = (r0 * rx + r1 * ry) + r2 * f;
      const double Y = (r3 * rx + r4 * ry) + r5 * f;
      const double Z = (r6 * rx + r7 * ry) + r8 * f;
      const double x = (X * fz) / Z + F.cx, y = (fz * Y) / Z + F.cy;
      const double dx = (x - vx) - (double)m.z, dy = (y - vy) - (double)m.w;
      e = dy * dy + dx * dx;
      F.miss[i] = e;
    }
    err = ordered_sum(err, e, F.wsum);
  }
  return err;
}

/* calcTransformQuality for params p: updates miss[] of the enabled fields.
   Lane l computes the error of enabled field j0 + l of each 64-field chunk;
   every lane then adds the chunk's 64 errors in field order (ordered_sum;
   lanes past the last field hold +0.0, which leaves a sum of non-negative
   terms unchanged), so the total is the reference's sequential sum. */
__device__ double quality(Frame& F, const double p[4]) {
  const double reg = fabs(p[2]) * 0.2 + fabs(p[3]) * 0.002;
  if (F.f > 0.0) {
    const double err = quality_fov_sum(F, p);
    return reg + (F.n > 0 ? err / (double)(F.ne + 1) : 0.0);
  }
  double sn, cs;
  sincos(p[2], &sn, &cs);
  const double z = p[3] * 0.01 + 1.0;
  const double zcos = cs * z, zsin = z * sn;
  double err = 0.0;
  for (int j0 = 0; j0 < F.ne; j0 += 64) {
    const int j = j0 + (int)threadIdx.x;
    double e = 0.0;
    if (j < F.ne) {
      const int i = F.eidx[j];
      const short4 m = F.lm[i];
      const double vx = (double)m.x, vy = (double)m.y;
      const double rx = vx - F.cx, ry = vy - F.cy;
      const double X = (rx * zcos + zsin * ry) + (p[0] + F.cx);
      const double Y = ((F.cx - vx) * zsin + ry * zcos) + (p[1] + F.cy);
      const double dx = (X - vx) - (double)m.z, dy = (Y - vy) - (double)m.w;
      e = dy * dy + dx * dx;
      F.miss[i] = e;
    }
    err = ordered_sum(err, e, F.wsum);
  }
  return reg + (F.n > 0 ? err / (double)(F.ne + 1) : 0.0);
}

/* disableFields (:180-211) of q against mask = miss; q may be miss itself.
   mean: two interleaved partial sums over the pairs (even and odd indices),
   added, then the odd tail; stddev: one sequential sum -- both taken in
   index order as in quality(), with +0.0 past the end */
__device__ void disable_fields(Frame& F, const double* q, double stddevs) {
  __syncthreads();   /* every lane's miss[] writes visible to every lane */
  const int len = F.n, lane = threadIdx.x, paired = 2 * (len / 2);
  double even = 0.0, odd = 0.0;
  for (int i0 = 0; i0 < paired; i0 += 64) {
    const int i = i0 + lane;
    const double v = i < paired ? q[i] : 0.0;
    ordered_pair_sums(even, odd, v, F.wsum);
  }
  double sum = odd + even;
  if (len & 1) sum += q[len - 1];
  const double mu = sum / (double)len;
  double ss = 0.0;
  for (int i0 = 0; i0 < len; i0 += 64) {
    const int i = i0 + lane;
    const double d = i < len ? q[i] - mu : 0.0;
    const double d2 = d * d;
    ss = ordered_sum(ss, d2, F.wsum);
  }
  const double thresh = sqrt(ss / (double)len) * stddevs + mu;
  __syncthreads();
  for (int i = lane; i < F.n; i += 64)
    if (q[i] > thresh) F.miss[i] = -1.0;
  __syncthreads();
  compact(F);
}

/* vsGradientDescent with N = 16, threshold 0.01 and the fixed step sizes;
   the parameter index is selected with unrolled compares, so the four-entry
   arrays stay in registers */
__device__ double descend(Frame& F, double x[4]) {
  double step[4] = {0.2, 0.2, 0.00005, 0.1};
  double v = quality(F, x);
  for (int i = 0; i < 16 * 4 && v > 0.01; ++i) {
    const int k = i % 4;
    const double h = ((i / 4) % 2) ? 1e-6 : -1e-6;
    double x2[4], sk = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      x2[j] = j == k ? h + x[j] : x[j];
      sk = j == k ? step[j] : sk;
    }
    double v2 = quality(F, x2);
    const double g = (v - v2) / h;
#pragma unroll
    for (int j = 0; j < 4; ++j) x2[j] = x[j] + (j == k ? g : 0.0) * sk;
    v2 = quality(F, x2);
    const bool better = v2 < v;
    if (better) {
#pragma unroll
      for (int j = 0; j < 4; ++j) x[j] = x2[j];
      v = v2;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (j == k) step[j] = better ? sk * 1.2 : sk * 0.5;
  }
  return v;
}

__global__ __launch_bounds__(64) void k_motions_to_transform(M2TArgs a) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int f = blockIdx.x;
  if (f >= a.nframes) return;
  Frame F;
  F.n = a.count[f];
  F.wsum = reinterpret_cast<double*>(lds);
  F.miss = F.wsum + 64;
  F.ecomp = F.miss + a.cap;
  short4* lmS = reinterpret_cast<short4*>(F.ecomp + a.cap);
  F.lm = lmS;
  F.eidx = reinterpret_cast<int*>(lmS + a.cap);
  F.cx = (double)a.cx;
  F.cy = (double)a.cy;
  F.f = a.focal;
  VSTransform t = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0};
  double res = 0.0;
  int rounds = 0;
  if (F.n <= 0) {
    t.extra = 1;
  } else {
    /* the frame's first motion: given, or the sum of the earlier frames' counts */
    int first;
    if (a.first) {
      first = a.first[f];
    } else {
      first = 0;
      for (int g = threadIdx.x; g < f; g += 64) first += a.count[g];
      for (int off = 32; off > 0; off >>= 1) first += __shfl_xor(first, off);
    }
    /* the frame's motions and match values into LDS; miss = 0 */
    const short4* lmG = a.lm + first;
    const double* match = a.match + first;
    int sx = 0, sy = 0;
    for (int i = threadIdx.x; i < F.n; i += 64) {
      const short4 m = lmG[i];
      lmS[i] = m;
      sx += m.z;
      sy += m.w;
      F.miss[i] = 0.0;
      F.ecomp[i] = match[i];
    }
    /* meanMotions (:189-203): integer sums, exact in double */
    for (int off = 32; off > 0; off >>= 1) {
      sx += __shfl_xor(sx, off);
      sy += __shfl_xor(sy, off);
    }
    __syncthreads();
    disable_fields(F, F.ecomp, 1.5);
    double params[4] = {(double)sx / (double)F.n, (double)sy / (double)F.n, 0.0, 0.0};
    int k;
    for (k = 0; k < 3; ++k) {
      double x[4] = {params[0], params[1], params[2], params[3]};
      res = descend(F, x);
      for (int j = 0; j < 4; ++j) params[j] = x[j];
      if ((k == 0 && res > 0.1) || (k == 1 && res > 20))
        disable_fields(F, F.miss, 1.0);
      else
        break;
    }
    rounds = k + 1;
    t.x = params[0];
    t.y = params[1];
    t.alpha = params[2];
    t.zoom = a.smoothZoom ? params[3] : 0.0;
    t.extra = res > 100 ? 1 : 0;
  }
  if (threadIdx.x == 0) {
    a.out[f] = t;
    if (a.residual) a.residual[f] = res;
    if (a.rounds) a.rounds[f] = rounds;
  }
}

/* LDS bytes for frames of up to cap motions */
size_t m2t_lds(int cap) { return 64 * sizeof(double) + (size_t)cap * (2 * sizeof(double) + sizeof(short4) + sizeof(int)); }

}  // namespace

/* ---- the allocation-free variant ---------------------------------------------
   Staging for the inputs and outputs lives in one static device array per
   GPU: the motions and match values of up to kStaticMotions motions, the
   counts and the transform, residual and round count of up to kStaticFrames
   frames (4.8 MB).  Its device address is looked up once per GPU; the host
   name of the array is not a device address and is never passed to HIP.
   Nothing in it needs clearing: the kernel writes every output slot of the
   frames it is given and reads only what the call uploaded.  Calls on one GPU
   share the array and would race if they overlapped, so a per-device lock
   serializes them; the array keeps the last call's contents. */
namespace {

constexpr int kStaticMotions = 1 << 18;
constexpr int kStaticFrames = 8192;
constexpr size_t kOffMatch = sizeof(short4) * kStaticMotions;
constexpr size_t kOffCount = kOffMatch + sizeof(double) * kStaticMotions;
constexpr size_t kOffOut = kOffCount + sizeof(int) * kStaticFrames;
constexpr size_t kOffRes = kOffOut + sizeof(VSTransform) * kStaticFrames;
constexpr size_t kOffRounds = kOffRes + sizeof(double) * kStaticFrames;
constexpr size_t kScratchBytes = kOffRounds + sizeof(int) * kStaticFrames;
static_assert(kOffOut % 8 == 0 && kOffRes % 8 == 0, "aligned scratch sections");

constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

std::mutex g_scratchMu[kMaxDevices];
uint8_t* g_scratchPtr[kMaxDevices];   /* device address of scratch on each GPU, once looked up */

}  // namespace

extern "C" int vsg_motions_to_transform_static(int device, int width, int height, const int16_t* lm,
                                               const double* match, const int* count, int nframes, int smoothZoom,
                                               VSTransform* out, double* residual, int* rounds) {
  if (width <= 0 || height <= 0 || nframes < 0 || (nframes > 0 && (!count || !out)) || device < 0 ||
      device >= kMaxDevices)
    return vsg_fail("vsg_motions_to_transform_static: invalid arguments", hipSuccess);
  if (nframes == 0) return 0;
  if (nframes > kStaticFrames) return vsg_fail("vsg_motions_to_transform_static: more than 8192 frames", hipSuccess);
  long long total = 0;
  int cap = 1;
  for (int f = 0; f < nframes; ++f) {
    if (count[f] < 0) return vsg_fail("vsg_motions_to_transform_static: negative motion count", hipSuccess);
    total += count[f];
    if (count[f] > cap) cap = count[f];
  }
  if (total > kStaticMotions)
    return vsg_fail("vsg_motions_to_transform_static: more than 262144 motions", hipSuccess);
  if (total > 0 && (!lm || !match)) return vsg_fail("vsg_motions_to_transform_static: no motions", hipSuccess);
  std::lock_guard<std::mutex> lock(g_scratchMu[device]);
  int lmax = 0;
  if (hipSetDevice(device) != hipSuccess ||
      hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
    return vsg_fail("vsg_motions_to_transform_static: device", hipSuccess);
  if (m2t_lds(cap) > (size_t)lmax)
    return vsg_fail("vsg_motions_to_transform_static: too many motions in a frame", hipSuccess);
  if (!g_scratchPtr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return vsg_fail("vsg_motions_to_transform_static: hipGetSymbolAddress", e);
    g_scratchPtr[device] = static_cast<uint8_t*>(p);
  }
  uint8_t* base = g_scratchPtr[device];
  /* the null stream: no stream is created, and its pageable-memory copies
     return when the data has left or reached host memory */
  hipStream_t st = nullptr;
  if ((total > 0 && (hipMemcpyAsync(base, lm, sizeof(short4) * total, hipMemcpyHostToDevice, st) != hipSuccess ||
                     hipMemcpyAsync(base + kOffMatch, match, sizeof(double) * total, hipMemcpyHostToDevice, st) !=
                         hipSuccess)) ||
      hipMemcpyAsync(base + kOffCount, count, sizeof(int) * nframes, hipMemcpyHostToDevice, st) != hipSuccess)
    return vsg_fail("vsg_motions_to_transform_static: upload", hipSuccess);
  M2TArgs a;
  a.lm = reinterpret_cast<const short4*>(base);
  a.match = reinterpret_cast<const double*>(base + kOffMatch);
  a.first = nullptr;
  a.count = reinterpret_cast<const int*>(base + kOffCount);
  a.nframes = nframes;
  a.cx = width / 2;
  a.cy = height / 2;
  a.smoothZoom = smoothZoom ? 1 : 0;
  a.cap = cap;
  a.focal = 0.0;
  a.out = reinterpret_cast<VSTransform*>(base + kOffOut);
  a.residual = reinterpret_cast<double*>(base + kOffRes);
  a.rounds = reinterpret_cast<int*>(base + kOffRounds);
  hipLaunchKernelGGL(k_motions_to_transform, dim3(nframes), dim3(64), m2t_lds(cap), st, a);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(out, a.out, sizeof(VSTransform) * nframes, hipMemcpyDeviceToHost, st) != hipSuccess ||
      (residual && hipMemcpyAsync(residual, a.residual, sizeof(double) * nframes, hipMemcpyDeviceToHost, st) !=
                       hipSuccess) ||
      (rounds && hipMemcpyAsync(rounds, a.rounds, sizeof(int) * nframes, hipMemcpyDeviceToHost, st) != hipSuccess) ||
      hipStreamSynchronize(st) != hipSuccess)
    return vsg_fail("vsg_motions_to_transform_static: kernel", hipSuccess);
  return 0;
}

extern "C" int vsg_motions_to_transform(int device, int width, int height, const int16_t* lm, const double* match,
                                        const int* count, int nframes, int smoothZoom, VSTransform* out,
                                        double* residual, int* rounds) {
  return vsg_motions_to_transform_fov(device, width, height, lm, match, count, nframes, smoothZoom, 0.0, out,
                                      residual, rounds);
}

extern "C" int vsg_motions_to_transform_fov(int device, int width, int height, const int16_t* lm,
                                            const double* match, const int* count, int nframes, int smoothZoom,
                                            double focal, VSTransform* out, double* residual, int* rounds) {
  if (width <= 0 || height <= 0 || nframes < 0 || (nframes > 0 && (!count || !out)) || !(focal >= 0.0))
    return vsg_fail("vsg_motions_to_transform: invalid arguments", hipSuccess);
  if (nframes == 0) return 0;
  long long total = 0;
  int cap = 1;
  for (int f = 0; f < nframes; ++f) {
    if (count[f] < 0) return vsg_fail("vsg_motions_to_transform: negative motion count", hipSuccess);
    total += count[f];
    if (count[f] > cap) cap = count[f];
  }
  if (total > 0 && (!lm || !match)) return vsg_fail("vsg_motions_to_transform: no motions", hipSuccess);
  int lmax = 0;
  if (hipSetDevice(device) != hipSuccess ||
      hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
    return vsg_fail("vsg_motions_to_transform: device", hipSuccess);
  if (m2t_lds(cap) > (size_t)lmax) return vsg_fail("vsg_motions_to_transform: too many motions in a frame", hipSuccess);
  std::vector<int> first(nframes);
  long long acc = 0;
  for (int f = 0; f < nframes; ++f) {
    first[f] = (int)acc;
    acc += count[f];
  }
  if (acc > 0x7FFFFFFFLL) return vsg_fail("vsg_motions_to_transform: too many motions", hipSuccess);
  const size_t nm = total > 0 ? (size_t)total : 1;
  void *d_lm = nullptr, *d_match = nullptr, *d_first = nullptr, *d_count = nullptr, *d_out = nullptr,
       *d_res = nullptr, *d_rounds = nullptr;
  hipStream_t st = nullptr;
  int rc = 0;
  do {
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&d_lm, sizeof(short4) * nm) != hipSuccess || hipMalloc(&d_match, sizeof(double) * nm) != hipSuccess ||
        hipMalloc(&d_first, sizeof(int) * nframes) != hipSuccess ||
        hipMalloc(&d_count, sizeof(int) * nframes) != hipSuccess ||
        hipMalloc(&d_out, sizeof(VSTransform) * nframes) != hipSuccess ||
        hipMalloc(&d_res, sizeof(double) * nframes) != hipSuccess ||
        hipMalloc(&d_rounds, sizeof(int) * nframes) != hipSuccess) {
      rc = vsg_fail("vsg_motions_to_transform: out of device memory", hipSuccess);
      break;
    }
    if ((total > 0 && (hipMemcpyAsync(d_lm, lm, sizeof(short4) * total, hipMemcpyHostToDevice, st) != hipSuccess ||
                       hipMemcpyAsync(d_match, match, sizeof(double) * total, hipMemcpyHostToDevice, st) !=
                           hipSuccess)) ||
        hipMemcpyAsync(d_first, first.data(), sizeof(int) * nframes, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d_count, count, sizeof(int) * nframes, hipMemcpyHostToDevice, st) != hipSuccess) {
      rc = vsg_fail("vsg_motions_to_transform: upload", hipSuccess);
      break;
    }
    M2TArgs a;
    a.lm = (const short4*)d_lm;
    a.match = (const double*)d_match;
    a.first = (const int*)d_first;
    a.count = (const int*)d_count;
    a.nframes = nframes;
    a.cx = width / 2;
    a.cy = height / 2;
    a.smoothZoom = smoothZoom ? 1 : 0;
    a.cap = cap;
    a.focal = focal;
    a.out = (VSTransform*)d_out;
    a.residual = (double*)d_res;
    a.rounds = (int*)d_rounds;
    hipLaunchKernelGGL(k_motions_to_transform, dim3(nframes), dim3(64), m2t_lds(cap), st, a);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(out, d_out, sizeof(VSTransform) * nframes, hipMemcpyDeviceToHost, st) != hipSuccess ||
        (residual &&
         hipMemcpyAsync(residual, d_res, sizeof(double) * nframes, hipMemcpyDeviceToHost, st) != hipSuccess) ||
        (rounds && hipMemcpyAsync(rounds, d_rounds, sizeof(int) * nframes, hipMemcpyDeviceToHost, st) != hipSuccess) ||
        hipStreamSynchronize(st) != hipSuccess) {
      rc = vsg_fail("vsg_motions_to_transform: kernel", hipSuccess);
      break;
    }
  } while (0);
  void* bufs[] = {d_lm, d_match, d_first, d_count, d_out, d_res, d_rounds};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (st) (void)hipStreamDestroy(st);
  return rc;
}
