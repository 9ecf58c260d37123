/*
 * vidstabdetect_like.c -- the call sequence of ffmpeg's vf_vidstabdetect
 * filter, in C, against this library's detect ABI: the drop-in check one
 * level below the Python bindings.
 *
 *   vsFrameInfoInit -> vsMotionDetectGetDefaultConfig -> vsMotionDetectInit
 *   -> vsMotionDetectGetConfig -> vsPrepareFile
 *   -> per frame: fill a VSFrame (luma with its own linesize, constant
 *      chroma) -> vsMotionDetection -> vsWriteToFile -> vs_vector_del
 *   -> vsMotionDetectionCleanup
 *
 * (the same sequence as the reference's tests/test_store_restore.c:11-57)
 *
 * usage: vidstabdetect_like FRAMES.raw WIDTH HEIGHT NFRAMES STRIDE SHAKINESS
 *        ACCURACY STEPSIZE MINCONTRAST_BITS TRIPOD MODE OUT.trf [WARMUP]
 *   FRAMES.raw      NFRAMES luma planes of HEIGHT rows of STRIDE bytes
 *   MINCONTRAST_BITS the IEEE-754 bits of the contrast threshold, in hex
 *   MODE            1 ASCII, 2 binary .trf (md->serializationMode)
 *   WARMUP          if given: the frames after the first WARMUP are timed
 *                   (each one's vsMotionDetection, vsWriteToFile and
 *                   vs_vector_del, as the filter does per frame) and one line
 *                   "per-frame <frames> <ms per frame> <frames/s>" is printed
 * Exits 0 on success; every failing call is reported on stderr.
 */
#include <inttypes.h>
#include <stdio.h>
#include <time.h>
#include <stdlib.h>
#include <string.h>

#include "libvidstab.h"   /* what ffmpeg includes (as <vid.stab/libvidstab.h>) */

static int fail(const char* what) {
  fprintf(stderr, "vidstabdetect_like: %s failed\n", what);
  return 1;
}

int main(int argc, char** argv) {
  if (argc != 13 && argc != 14) {
    fprintf(stderr, "usage: %s FRAMES.raw W H N STRIDE SHAKINESS ACCURACY STEPSIZE MINCONTRAST_BITS TRIPOD MODE OUT\n",
            argv[0]);
    return 2;
  }
  const int w = atoi(argv[2]), h = atoi(argv[3]), n = atoi(argv[4]), stride = atoi(argv[5]);
  const uint64_t bits = strtoull(argv[9], NULL, 16);
  double mincontrast;
  memcpy(&mincontrast, &bits, sizeof mincontrast);
  if (w <= 0 || h <= 0 || n <= 0 || stride < w) return fail("argument check");

  const size_t plane = (size_t)stride * h;
  uint8_t* frames = (uint8_t*)malloc(plane * (size_t)n);
  uint8_t* chroma = (uint8_t*)malloc((size_t)(w / 2) * (h / 2));
  if (!frames || !chroma) return fail("malloc");
  memset(chroma, 128, (size_t)(w / 2) * (h / 2));
  FILE* in = fopen(argv[1], "rb");
  if (!in || fread(frames, plane, (size_t)n, in) != (size_t)n) return fail("reading the frames");
  fclose(in);

  VSFrameInfo fi;
  if (!vsFrameInfoInit(&fi, w, h, PF_YUV420P)) return fail("vsFrameInfoInit");   /* 1 on success */
  VSMotionDetect md;
  memset(&md, 0, sizeof md);   /* ffmpeg's filter context is zero-allocated */
  VSMotionDetectConfig conf = vsMotionDetectGetDefaultConfig("vidstabdetect");
  conf.shakiness = atoi(argv[6]);
  conf.accuracy = atoi(argv[7]);
  conf.stepSize = atoi(argv[8]);
  conf.contrastThreshold = mincontrast;
  conf.virtualTripod = atoi(argv[10]);
  conf.show = 0;
  if (vsMotionDetectInit(&md, &conf, &fi) != VS_OK) return fail("vsMotionDetectInit");
  vsMotionDetectGetConfig(&conf, &md);
  md.serializationMode = atoi(argv[11]);

  FILE* f = fopen(argv[12], "wb");
  if (!f) return fail("fopen");
  if (vsPrepareFile(&md, f) != VS_OK) return fail("vsPrepareFile");
  const int warmup = argc == 14 ? atoi(argv[13]) : -1;
  struct timespec t0, t1;
  for (int t = 0; t < n; ++t) {
    if (t == warmup) clock_gettime(CLOCK_MONOTONIC, &t0);
    VSFrame frame;
    memset(&frame, 0, sizeof frame);
    frame.data[0] = frames + plane * (size_t)t;
    frame.linesize[0] = stride;
    frame.data[1] = frame.data[2] = chroma;
    frame.linesize[1] = frame.linesize[2] = w / 2;
    LocalMotions localmotions;
    if (vsMotionDetection(&md, &localmotions, &frame) != VS_OK) return fail("vsMotionDetection");
    if (vsWriteToFile(&md, f, &localmotions) != VS_OK) return fail("vsWriteToFile");
    vs_vector_del(&localmotions);
  }
  if (warmup >= 0 && warmup < n) {
    clock_gettime(CLOCK_MONOTONIC, &t1);
    const double s = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    printf("per-frame %d %.4f %.2f\n", n - warmup, 1000.0 * s / (n - warmup), (n - warmup) / s);
  }
  vsMotionDetectionCleanup(&md);
  fclose(f);
  free(frames);
  free(chroma);
  return 0;
}
