/*
 * two_libraries.c -- one process holding both vid.stab libraries, as one
 * ffmpeg binary with both vidstab filters would: detection through whichever
 * library the dynamic linker binds first, the transform side
 * (vsReadLocalMotionsFile, vsLocalmotions2Transforms, vsPreprocessTransforms)
 * from the reference library.
 *
 * Built twice from this file (oracle/Makefile, against the reference's
 * headers): linked with -lvidstab_amd ahead of the reference's -lvidstab, and
 * with the reference's -lvidstab alone.  Like ffmpeg's vidstabutils it first
 * replaces the allocation and log hooks and remaps VS_OK / VS_ERROR to 1 / 0
 * (libavfilter/vidstabutils.c, recalled: ffmpeg is not in this tree), then
 * runs the detect filter's call sequence on a synthetic clip, writes the .trf
 * stream, reads it back with the transform filter's calls and prints what
 * both builds must agree on: the .trf bytes, every transform, and that every
 * hook allocation was freed.  It also prints which library vsMotionDetection
 * resolved to.
 *
 * usage: two_libraries <out.trf> <frames>
 */
#define _GNU_SOURCE 1
#include <dlfcn.h>
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "libvidstab.h"
#include "vsynth.h"

static long g_live = 0, g_allocs = 0;

static void* hook_malloc(size_t n) {
  void* p = malloc(n);
  if (p) g_live++, g_allocs++;
  return p;
}
static void* hook_zalloc(size_t n) {
  void* p = calloc(1, n ? n : 1);
  if (p) g_live++, g_allocs++;
  return p;
}
static void* hook_realloc(void* p, size_t n) {
  void* q = realloc(p, n);
  if (q && !p) g_live++, g_allocs++;
  return q;
}
static void hook_free(void* p) {
  if (p) g_live--;
  free(p);
}
static int hook_log(int type, const char* tag, const char* format, ...) {
  (void)type;
  (void)tag;
  (void)format;
  return 0;
}

static uint64_t fnv(const void* p, size_t n) {
  uint64_t h = 0xCBF29CE484222325ull;
  for (size_t i = 0; i < n; i++) h = (h ^ ((const uint8_t*)p)[i]) * 0x100000001B3ull;
  return h;
}

static uint64_t bits(double d) {
  uint64_t u;
  memcpy(&u, &d, 8);
  return u;
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const int nframes = atoi(argv[2]);
  vs_malloc = hook_malloc;
  vs_zalloc = hook_zalloc;
  vs_realloc = hook_realloc;
  vs_free = hook_free;
  vs_log = hook_log;
  VS_ERROR = 0;
  VS_OK = 1;

  Dl_info where;
  if (dladdr((void*)vsMotionDetection, &where) && where.dli_fname) {
    const char* base = strrchr(where.dli_fname, '/');
    printf("vsMotionDetection from %s\n", base ? base + 1 : where.dli_fname);
  }
  if (dladdr((void*)vsLocalmotions2Transforms, &where) && where.dli_fname) {
    const char* base = strrchr(where.dli_fname, '/');
    printf("vsLocalmotions2Transforms from %s\n", base ? base + 1 : where.dli_fname);
  }

  VSynSpec sp;
  VSynClip clip;
  vsyn_spec(1, 50, &sp);
  VSynMotion* path = (VSynMotion*)calloc((size_t)sp.nframes, sizeof(VSynMotion));
  vsyn_open(&sp, &clip, path);
  const int W = sp.width, H = sp.height;

  /* the detect filter: vsFrameInfoInit -> vsMotionDetectInit -> vsMotionDetectGetConfig
     -> vsPrepareFile -> {vsMotionDetection; vsWriteToFile; vs_vector_del} -> cleanup */
  VSFrameInfo fi;
  if (!vsFrameInfoInit(&fi, W, H, PF_YUV420P)) return 3;
  VSMotionDetect md;
  memset(&md, 0, sizeof(md));
  VSMotionDetectConfig conf = vsMotionDetectGetDefaultConfig("two_libraries");
  conf.shakiness = sp.shakiness;
  conf.accuracy = sp.accuracy;
  conf.stepSize = sp.stepsize;
  conf.contrastThreshold = sp.mincontrast;
  if (vsMotionDetectInit(&md, &conf, &fi) != VS_OK) {
    printf("vsMotionDetectInit failed\n");
    return 4;
  }
  vsMotionDetectGetConfig(&conf, &md);
  FILE* f = fopen(argv[1], "wb");
  if (!f || vsPrepareFile(&md, f) != VS_OK) return 5;
  uint8_t* Y = (uint8_t*)malloc((size_t)W * H);
  uint8_t* C = (uint8_t*)malloc((size_t)(W / 2) * (H / 2));
  memset(C, 128, (size_t)(W / 2) * (H / 2));
  for (int t = 0; t < nframes; t++) {
    vsyn_render(&clip, path, t, Y, W);
    VSFrame fr;
    memset(&fr, 0, sizeof(fr));
    fr.data[0] = Y;
    fr.linesize[0] = W;
    fr.data[1] = fr.data[2] = C;
    fr.linesize[1] = fr.linesize[2] = W / 2;
    LocalMotions lms;
    if (vsMotionDetection(&md, &lms, &fr) != VS_OK) {
      printf("vsMotionDetection failed at frame %d\n", t);
      return 6;
    }
    if (vsWriteToFile(&md, f, &lms) != VS_OK) return 7;
    vs_vector_del(&lms);
  }
  vsMotionDetectionCleanup(&md);
  fclose(f);

  /* the transform filter's reading of the file */
  f = fopen(argv[1], "rb");
  if (!f) return 8;
  fseek(f, 0, SEEK_END);
  const long flen = ftell(f);
  fseek(f, 0, SEEK_SET);
  uint8_t* bytes = (uint8_t*)malloc((size_t)flen);
  if (fread(bytes, 1, (size_t)flen, f) != (size_t)flen) return 9;
  printf("trf %ld bytes fnv %016" PRIx64 "\n", flen, fnv(bytes, (size_t)flen));
  free(bytes);
  fseek(f, 0, SEEK_SET);
  VSManyLocalMotions mlms;
  if (vsReadLocalMotionsFile(f, &mlms) != VS_OK) {
    printf("vsReadLocalMotionsFile failed\n");
    return 10;
  }
  fclose(f);
  VSTransformData td;
  memset(&td, 0, sizeof(td));
  VSTransformConfig tconf = vsTransformGetDefaultConfig("two_libraries");
  if (vsTransformDataInit(&td, &tconf, &fi, &fi) != VS_OK) return 11;
  VSTransformations trans;
  vsTransformationsInit(&trans);
  if (vsLocalmotions2Transforms(&td, &mlms, &trans) != VS_OK) {
    printf("vsLocalmotions2Transforms failed\n");
    return 12;
  }
  printf("frames read %d, transforms %d\n", vs_vector_size(&mlms), trans.len);
  for (int i = 0; i < trans.len; i++)
    printf("raw %d %016" PRIx64 " %016" PRIx64 " %016" PRIx64 " %016" PRIx64 " %d\n", i, bits(trans.ts[i].x),
           bits(trans.ts[i].y), bits(trans.ts[i].alpha), bits(trans.ts[i].zoom), trans.ts[i].extra);
  if (vsPreprocessTransforms(&td, &trans) != VS_OK) return 13;
  for (int i = 0; i < trans.len; i++)
    printf("smoothed %d %016" PRIx64 " %016" PRIx64 " %016" PRIx64 " %016" PRIx64 "\n", i, bits(trans.ts[i].x),
           bits(trans.ts[i].y), bits(trans.ts[i].alpha), bits(trans.ts[i].zoom));
  vsTransformationsCleanup(&trans);
  vsTransformDataCleanup(&td);
  for (int i = 0; i < vs_vector_size(&mlms); i++) {
    LocalMotions* lm = (LocalMotions*)vs_vector_get(&mlms, i);
    if (lm) vs_vector_del(lm);
  }
  vs_vector_del(&mlms);
  printf("hook allocations %s\n", g_allocs > 0 && g_live == 0 ? "all freed" : "NOT BALANCED");
  vsyn_clip_free(&clip);
  free(path);
  free(Y);
  free(C);
  return 0;
}
