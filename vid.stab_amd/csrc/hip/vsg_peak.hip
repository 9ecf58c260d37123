/*
 * vsg_peak.hip -- measurement helpers.
 *
 *   vsg_sad_peak   the packed-u8 SAD issue rate (v_sad_u8): every lane runs
 *                  16 independent accumulator chains whose second operand is
 *                  a wave-uniform word from a scalar load (as in the grid
 *                  search), with no vector memory traffic.  The measured
 *                  ceiling the search kernel's abs-diff rate is compared with.
 *   vsg_blur_time  the batch box blur alone on an idle device: inside the
 *                  detect pipeline it shares the CUs with other batches'
 *                  searches, which stretches its event interval.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "vsg.h"
#include "vsg_internal.h"

namespace {

constexpr int kChains = 16;

typedef const __attribute__((address_space(4))) uint32_t* cword_ptr;

__global__ __launch_bounds__(256) void k_sad_peak(const uint32_t* __restrict__ words, int iters,
                                                  uint32_t* __restrict__ out) {
  cword_ptr w = (cword_ptr)words;
  uint32_t x[kChains], acc[kChains];
#pragma unroll
  for (int i = 0; i < kChains; ++i) {
    x[i] = (threadIdx.x * 2654435761u) ^ (i * 0x9E3779B9u) ^ blockIdx.x;
    acc[i] = 0;
  }
#pragma unroll 8
  for (int it = 0; it < iters; ++it) {
    const uint32_t c = w[it & 255];
#pragma unroll
    for (int i = 0; i < kChains; ++i) acc[i] = __builtin_amdgcn_sad_u8(x[i], c, acc[i]);
  }
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kChains; ++i) s += acc[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

extern "C" int vsg_sad_peak(int device, double* absdiffPerSec, double* ms) {
  if (!absdiffPerSec) return -1;
  if (hipSetDevice(device) != hipSuccess) return -1;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return -1;
  const int blocks = prop.multiProcessorCount * 32, threads = 256, iters = 8192;
  uint32_t* words = nullptr;
  uint32_t* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = -1;
  uint32_t h[256];
  for (int i = 0; i < 256; ++i) h[i] = 0x01020304u * (uint32_t)(i + 1);
  if (hipMalloc(&words, sizeof(h)) != hipSuccess) goto done;
  if (hipMalloc(&out, sizeof(uint32_t) * (size_t)blocks * threads) != hipSuccess) goto done;
  if (hipMemcpy(words, h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess) goto done;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto done;
  hipLaunchKernelGGL(k_sad_peak, dim3(blocks), dim3(threads), 0, 0, words, iters / 8, out);   /* warm-up */
  if (hipEventRecord(e0, 0) != hipSuccess) goto done;
  hipLaunchKernelGGL(k_sad_peak, dim3(blocks), dim3(threads), 0, 0, words, iters, out);
  if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) goto done;
  {
    float t = 0;
    if (hipEventElapsedTime(&t, e0, e1) != hipSuccess || t <= 0) goto done;
    const double sads = (double)blocks * threads * iters * kChains;
    *absdiffPerSec = 4.0 * sads / (t * 1e-3);
    if (ms) *ms = t;
    rc = 0;
  }
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (words) (void)hipFree(words);
  if (out) (void)hipFree(out);
  return rc;
}

extern "C" int vsg_blur_time(int device, const uint8_t* const* lumaDev, const int* linesize, int n, int width,
                             int height, int blurSize, int reps, double* msPerLaunch) {
  if (!lumaDev || !linesize || !msPerLaunch || n <= 0 || width <= 0 || height <= 0 || reps <= 0) return -1;
  for (int i = 0; i < n; ++i)
    if (!lumaDev[i] || linesize[i] < width) return -1;
  if (hipSetDevice(device) != hipSuccess) return -1;
  /* destination slots as the detector lays them out: 64-byte-aligned pitch */
  const int pitch = (width + 63) & ~63;
  const size_t slotStride = (size_t)pitch * height;
  std::vector<VSGBlurSrc> h((size_t)n);
  for (int i = 0; i < n; ++i) h[(size_t)i] = VSGBlurSrc{lumaDev[i], linesize[i], 0};
  VSGBlurSrc* d_src = nullptr;
  uint8_t* dst = nullptr;
  uint8_t* tmp = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = -1;
  if (hipMalloc(&d_src, sizeof(VSGBlurSrc) * (size_t)n) != hipSuccess) goto done;
  if (hipMalloc(&dst, slotStride * (size_t)n) != hipSuccess) goto done;
  if (vsg_blur_needs_tmp(width, blurSize) && hipMalloc(&tmp, slotStride * (size_t)n) != hipSuccess) goto done;
  if (hipMemcpy(d_src, h.data(), sizeof(VSGBlurSrc) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess) goto done;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto done;
  if (vsg_launch_blur(d_src, n, dst, slotStride, pitch, width, height, blurSize, tmp, 0) != hipSuccess) goto done;
  if (hipEventRecord(e0, 0) != hipSuccess) goto done;
  for (int r = 0; r < reps; ++r)
    if (vsg_launch_blur(d_src, n, dst, slotStride, pitch, width, height, blurSize, tmp, 0) != hipSuccess) goto done;
  if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) goto done;
  {
    float t = 0;
    if (hipEventElapsedTime(&t, e0, e1) != hipSuccess || t <= 0) goto done;
    *msPerLaunch = t / reps;
    rc = 0;
  }
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (d_src) (void)hipFree(d_src);
  if (dst) (void)hipFree(dst);
  if (tmp) (void)hipFree(tmp);
  return rc;
}
