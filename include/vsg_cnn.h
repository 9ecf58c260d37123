/*
 * vsg_cnn.h -- C ABI of the proposal network's gfx950 kernels (learned
 * detect form): BF16 3x3 convolutions on the matrix cores with FP32
 * accumulation, their data / weight gradients, field pooling, loss and Adam.
 * The learned form has no counterpart in the reference; these entry points
 * serve its training driver (vidstab_amd/learned.py) and the learned detect
 * path.  All pointers are device pointers; `stream` is a hipStream_t (NULL:
 * the default stream).  Functions return 0 or a HIP error code.
 *
 * Activations: NHWC, C = 32, BF16 bit patterns, index ((n*H+y)*W+x)*32+c.
 * Weights: FP32 [9 taps][32 in][32 out], tap = (dy+1)*3 + (dx+1); bias FP32 [32].
 */
#ifndef VSG_CNN_H
#define VSG_CNN_H

#include <stddef.h>
#include <stdint.h>
#include "vidstab_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* MFMA B-operand packing of one layer (9*4*64*4 BF16 values): forward, or
   (transposed != 0) flipped and transposed for the data gradient */
VS_API int vsgc_pack_weights(const float* w, uint16_t* packed, int transposed, void* stream);

/* out = act(conv3x3_dilated(in) + bias): mode 0 -> BF16 + ReLU into outb,
   1 -> BF16 into outb, 2 -> FP32 into outf; padding = dilation; bias may be
   NULL (the data gradient: conv of dz with the transposed packing) */
VS_API int vsgc_conv(const uint16_t* in, int n, int h, int w, const uint16_t* packed, const float* bias,
                     int dil, int mode, uint16_t* outb, float* outf, void* stream);
/* as vsgc_conv, for an input whose channels >= inChannels are zero; with
   inChannels <= 8 only the first 8 channels enter the MFMA chain */
VS_API int vsgc_conv_k(const uint16_t* in, int n, int h, int w, const uint16_t* packed, const float* bias,
                       int dil, int mode, int inChannels, uint16_t* outb, float* outf, void* stream);

/* The same convolution on the gfx950 K = 16 matrix instruction
   (v_mfma_f32_32x32x16_bf16), used by the learned detect path: weights
   packed by vsgc_pack_weights16 (9*2*64*8 BF16 values, forward only);
   inChannels <= 16 uses one K step per tap.  Dilations up to 8 run the
   LDS-tiled kernel (4 output rows spaced by the dilation per workgroup) or,
   for 32 input channels, the band walker (a persistent workgroup walking
   down a 64-px band with its input rows in an LDS ring;
   VIDSTAB_AMD_CONV_WALK=0 selects the tiled kernel), larger ones the direct
   kernel; all accumulate the taps in the same order, so their results are
   identical.  mode 3 (tiled dilations only) writes FP32
   channels 0 and 1 alone, as an NHWC-2 map [n][h][w][2]. */
VS_API int vsgc_pack_weights16(const float* w, uint16_t* packed16, void* stream);
VS_API int vsgc_conv16(const uint16_t* in, int n, int h, int w, const uint16_t* packed16, const float* bias,
                       int dil, int mode, int inChannels, uint16_t* outb, float* outf, void* stream);
/* vsgc_conv16 of the network's first layer with its input built on the fly
   from quarter-resolution BF16 planes (planeStride elements apart, each
   h x w): sample z has channel 0 = plane prevIdx[z] (refPlane when
   prevIdx[z] < 0; prevIdx is a device array of n ints), channel 1 = plane
   curIdx0 + z and every other channel zero -- the result equals vsgc_conv16
   with inChannels = 2 on that NHWC-32 map.  dil <= 8. */
VS_API int vsgc_conv16_pair(const uint16_t* planes, size_t planeStride, const uint16_t* refPlane,
                            const int* prevIdx, int curIdx0, int n, int h, int w, const uint16_t* packed16,
                            const float* bias, int dil, int mode, uint16_t* outb, float* outf, void* stream);
/* the first two layers in one pass: layer 0 (packed0, bias0, dilation
   dil0, BF16 + ReLU) on the plane pair as vsgc_conv16_pair reads it, and
   layer 1 (packed1, bias1, dil1 <= 8, 32 input channels, BF16 output:
   mode 0 or 1, into outb) on its result, which never leaves the chip: the
   band walker computes each staged row of layer 1's input with layer 0's
   MFMAs.  dil0 <= 8.  The result equals vsgc_conv16_pair (mode 0) followed
   by vsgc_conv16 (inChannels 32), bit for bit. */
VS_API int vsgc_conv16_pair2(const uint16_t* planes, size_t planeStride, const uint16_t* refPlane,
                             const int* prevIdx, int curIdx0, int n, int h, int w, const uint16_t* packed0,
                             const float* bias0, int dil0, const uint16_t* packed1, const float* bias1, int dil1,
                             int mode, uint16_t* outb, float* outf, void* stream);
/* vsgc_conv16 on the direct kernel (one 32-px row tile per wave, every A
   fragment loaded from global memory) at any dilation */
VS_API int vsgc_conv16_direct(const uint16_t* in, int n, int h, int w, const uint16_t* packed16,
                              const float* bias, int dil, int mode, int inChannels, uint16_t* outb, float* outf,
                              void* stream);

/* dW += d(loss)/dW and dB += d(loss)/dB of a layer from its input and the
   gradient dz at its pre-activation (accumulates; dB may be NULL) */
VS_API int vsgc_conv_wgrad(const uint16_t* in, const uint16_t* dz, int n, int h, int w, int dil, float* dW,
                           float* dB, void* stream);

/* grad[i] = 0 where act[i] <= 0 (ReLU backward, in place) */
VS_API int vsgc_relu_mask(uint16_t* grad, const uint16_t* act, long long count, void* stream);
VS_API int vsgc_f32_to_bf16(const float* f, uint16_t* b, long long count, void* stream);

/* channel `chan` of an NHWC-32 map (width/4 x height/4) = (4x4 mean - 128)/64 */
VS_API int vsgc_downsample(const uint8_t* plane, int pitch, int width, int height, uint16_t* out, int chan,
                           void* stream);

/* windows: nf x {sample, x0, y0, x1, y1} (half-open) on the map;
   pred[f] = mean of channels 0, 1 of the FP32 map over window f */
VS_API int vsgc_field_pool(const float* map, int h, int w, const int* windows, int nf, float* pred, void* stream);
/* the same over an FP32 NHWC-2 map (vsgc_conv16 mode 3) */
VS_API int vsgc_field_pool2(const float* map, int h, int w, const int* windows, int nf, float* pred, void* stream);
/* dmap (FP32, zeroed by the caller) += dpred[f] / area over window f */
VS_API int vsgc_field_pool_bwd(const float* dpred, int h, int w, const int* windows, int nf, float* dmap,
                               void* stream);
/* loss += sum_f weight[f] * |pred - label|^2 * inv; dpred = its gradient */
VS_API int vsgc_field_loss(const float* pred, const float* label, const float* weight, int nf, float inv,
                           float* dpred, float* loss, void* stream);
/* one Adam step (bias-corrected, step >= 1) */
VS_API int vsgc_adam(float* w, const float* g, float* m, float* v, int count, float lr, float b1, float b2,
                     float eps, int step, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* VSG_CNN_H */
