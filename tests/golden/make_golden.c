/*
 * make_golden.c -- writes the detect-pass golden fixtures from the REFERENCE
 * implementation (oracle/_ref/libvsref.so, built from /root/reference/src
 * with the reference's own release flags).
 *
 * Run only where /root/reference exists (see make_golden.sh).  For each case
 * it regenerates the seeded synthetic frames, runs the reference's
 * vsMotionDetectInit / vsMotionDetection / vsPrepareFile / vsWriteToFile, and
 * records per frame: a hash of the input luma (pins the generator), a hash of
 * the reference's blurred plane (md->curr), every LocalMotion with the exact
 * bits of its doubles, and hashes of the binary and ASCII .trf streams.
 * In "transforms" mode it runs the planar cases' detection again and records
 * the reference's second-pass vsMotionsToTransform of every frame's
 * LocalMotions (default transform config: fov 0, no lens; smoothZoom 1 so
 * the fitted zoom is kept).
 */
#define _GNU_SOURCE 1
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "localmotion2transform.h"
#include "motiondetect.h"
#include "serialize.h"
#include "transform.h"
#include "vsynth.h"

static uint64_t fnv(const void* p, size_t n, uint64_t h) {
  const uint8_t* b = (const uint8_t*)p;
  for (size_t i = 0; i < n; i++) { h ^= b[i]; h *= 0x100000001B3ull; }
  return h;
}

static uint64_t dbits(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }

typedef struct {
  const char* name;
  int config, clip;
  int width, height;     /* 0: from the spec */
  int nframes;           /* frames to run */
  int stepsize;          /* 0: from the spec */
  int accuracy;          /* 0: from the spec */
  int shakiness;         /* 0: from the spec */
  int tripod;
  int stride_pad;
  double mincontrast;    /* <0: from the spec */
  int pformat;           /* 0: PF_YUV420P; else a packed format (vsyn_pack of the luma) */
  int show;              /* conf.show: overlays drawn into the caller's frame */
} Case;

static void run_case(FILE* out, const Case* cs, int first) {
  VSynSpec sp;
  vsyn_spec(cs->config, cs->clip, &sp);
  if (cs->width) { sp.width = cs->width; sp.height = cs->height; }
  if (cs->stepsize) sp.stepsize = cs->stepsize;
  if (cs->accuracy) sp.accuracy = cs->accuracy;
  if (cs->shakiness) sp.shakiness = cs->shakiness;
  if (cs->mincontrast >= 0) sp.mincontrast = cs->mincontrast;
  int nfr = cs->nframes < sp.nframes ? cs->nframes : sp.nframes;
  int total = sp.nframes > nfr ? sp.nframes : nfr;
  VSynClip clip;
  VSynMotion* path = (VSynMotion*)calloc((size_t)total, sizeof(VSynMotion));
  VSynSpec sp2 = sp; sp2.nframes = total;
  vsyn_open(&sp2, &clip, path);

  VSFrameInfo fi;
  vsFrameInfoInit(&fi, sp.width, sp.height, cs->pformat ? (VSPixelFormat)cs->pformat : PF_YUV420P);
  const int bpp = cs->pformat ? fi.bytesPerPixel : 1;
  VSMotionDetect md;
  memset(&md, 0, sizeof(md));
  VSMotionDetectConfig conf = vsMotionDetectGetDefaultConfig("golden");
  conf.shakiness = sp.shakiness;
  conf.accuracy = sp.accuracy;
  conf.stepSize = sp.stepsize;
  conf.contrastThreshold = sp.mincontrast;
  conf.virtualTripod = cs->tripod;
  conf.show = cs->show;
  conf.numThreads = 1;
  if (vsMotionDetectInit(&md, &conf, &fi) != VS_OK) { fprintf(stderr, "init failed\n"); exit(1); }

  char* bbuf = NULL; size_t blen = 0;
  char* abuf = NULL; size_t alen = 0;
  FILE* fb = open_memstream(&bbuf, &blen);
  FILE* fa = open_memstream(&abuf, &alen);
  vsPrepareFile(&md, fb);
  md.serializationMode = ASCII_SERIALIZATION_MODE;
  vsPrepareFile(&md, fa);
  md.serializationMode = BINARY_SERIALIZATION_MODE;

  int ls = sp.width + cs->stride_pad;
  const int pls = sp.width * bpp + cs->stride_pad;   /* packed frames: plane 0 stride */
  uint8_t* P = cs->pformat ? (uint8_t*)malloc((size_t)pls * sp.height) : NULL;
  if (P) memset(P, 0x5A, (size_t)pls * sp.height);
  uint8_t* Y = (uint8_t*)malloc((size_t)ls * sp.height);
  uint8_t* U = (uint8_t*)malloc((size_t)(sp.width / 2) * (sp.height / 2));
  uint8_t* V = (uint8_t*)malloc((size_t)(sp.width / 2) * (sp.height / 2));
  memset(Y, 0x5A, (size_t)ls * sp.height);
  memset(U, 128, (size_t)(sp.width / 2) * (sp.height / 2));
  memset(V, 128, (size_t)(sp.width / 2) * (sp.height / 2));

  fprintf(out, "%s{\"name\":\"%s\",\"pformat\":%d,\"bpp\":%d,\"config\":%d,\"clip\":%d,\"width\":%d,\"height\":%d,"
               "\"stride\":%d,\"shakiness\":%d,\"accuracy\":%d,\"stepsize\":%d,"
               "\"mincontrast\":\"%016" PRIx64 "\",\"tripod\":%d,\"show\":%d,\"md_conf\":[%d,%d,%d],"
               "\"coarse\":[%d,%d,%d,%d],\"fine\":[%d,%d,%d,%d],\"frames\":[",
          first ? "" : ",\n", cs->name, cs->pformat, bpp, cs->config, cs->clip, sp.width, sp.height,
          cs->pformat ? pls : ls,
          sp.shakiness, sp.accuracy, sp.stepsize, dbits(sp.mincontrast), cs->tripod, cs->show,
          md.conf.shakiness, md.conf.accuracy, md.conf.stepSize,
          md.fieldscoarse.fieldSize, md.fieldscoarse.maxShift, md.fieldscoarse.fieldNum,
          md.fieldscoarse.maxFields, md.fieldsfine.fieldSize, md.fieldsfine.maxShift,
          md.fieldsfine.fieldNum, md.fieldsfine.maxFields);
  for (int t = 0; t < nfr; t++) {
    vsyn_render(&clip, path, t, Y, ls);
    VSFrame fr;
    memset(&fr, 0, sizeof(fr));
    if (P) {
      vsyn_pack(Y, sp.width, sp.height, ls, bpp, P, pls);
      fr.data[0] = P; fr.linesize[0] = pls;
    } else {
      fr.data[0] = Y; fr.linesize[0] = ls;
      fr.data[1] = U; fr.linesize[1] = sp.width / 2;
      fr.data[2] = V; fr.linesize[2] = sp.width / 2;
    }
    LocalMotions lms;
    uint64_t hin = P ? vsyn_hash_plane(P, sp.width * bpp, sp.height, pls) : vsyn_hash_plane(Y, sp.width, sp.height, ls);
    vsMotionDetection(&md, &lms, &fr);
    /* with show > 0 the reference draws into the frame it was handed */
    uint64_t hdr = P ? vsyn_hash_plane(P, sp.width * bpp, sp.height, pls) : vsyn_hash_plane(Y, sp.width, sp.height, ls);
    uint64_t hbl = vsyn_hash_plane(md.curr.data[0], sp.width * bpp, sp.height, md.curr.linesize[0]);
    vsWriteToFile(&md, fb, &lms);
    md.serializationMode = ASCII_SERIALIZATION_MODE;
    vsWriteToFile(&md, fa, &lms);
    md.serializationMode = BINARY_SERIALIZATION_MODE;
    int n = vs_vector_size(&lms);
    fprintf(out, "%s\n {\"in\":\"%016" PRIx64 "\",\"drawn\":\"%016" PRIx64 "\",\"blur\":\"%016" PRIx64
                 "\",\"lm\":[",
            t ? "," : "", hin, hdr, hbl);
    for (int i = 0; i < n; i++) {
      LocalMotion* m = (LocalMotion*)vs_vector_get(&lms, i);
      fprintf(out, "%s[%d,%d,%d,%d,%d,\"%016" PRIx64 "\",\"%016" PRIx64 "\"]", i ? "," : "",
              m->v.x, m->v.y, m->f.x, m->f.y, m->f.size, dbits(m->contrast), dbits(m->match));
    }
    fprintf(out, "]}");
    vs_vector_del(&lms);
  }
  fflush(fb);
  fflush(fa);
  fprintf(out, "],\"trf_bin\":\"%016" PRIx64 "\",\"trf_bin_len\":%zu,"
               "\"trf_ascii\":\"%016" PRIx64 "\",\"trf_ascii_len\":%zu}",
          fnv(bbuf, blen, 0xCBF29CE484222325ull), blen,
          fnv(abuf, alen, 0xCBF29CE484222325ull), alen);
  fclose(fb);
  fclose(fa);
  free(bbuf);
  free(abuf);
  vsMotionDetectionCleanup(&md);
  vsyn_clip_free(&clip);
  free(path);
  free(Y); free(U); free(V); free(P);
  fprintf(stderr, "case %s done (%d frames)\n", cs->name, nfr);
}

/* FNV-1a over one frame's LocalMotions in output order: per motion the
   little-endian int16 v.x, v.y, f.x, f.y, f.size, then the IEEE bits of
   contrast and match (26 bytes; the struct's padding is left out) */
static uint64_t hash_motions(const LocalMotions* lms) {
  uint64_t h = 0xCBF29CE484222325ull;
  for (int i = 0; i < vs_vector_size(lms); i++) {
    const LocalMotion* m = (const LocalMotion*)vs_vector_get(lms, i);
    int16_t v[5] = {m->v.x, m->v.y, m->f.x, m->f.y, m->f.size};
    uint64_t c = dbits(m->contrast), q = dbits(m->match);
    h = fnv(v, sizeof(v), h);
    h = fnv(&c, 8, h);
    h = fnv(&q, 8, h);
  }
  return h;
}

/* Whole clips, hashes only: per frame the input luma hash, the number of
   LocalMotions, hash_motions and the blurred-plane hash.  For the full-length
   workload clips that are too long to keep every motion in a fixture. */
static void run_case_hashes(FILE* out, const Case* cs, int first, int threads) {
  VSynSpec sp;
  vsyn_spec(cs->config, cs->clip, &sp);
  if (cs->stepsize) sp.stepsize = cs->stepsize;
  int nfr = cs->nframes > 0 && cs->nframes < sp.nframes ? cs->nframes : sp.nframes;
  VSynClip clip;
  VSynMotion* path = (VSynMotion*)calloc((size_t)sp.nframes, sizeof(VSynMotion));
  vsyn_open(&sp, &clip, path);
  VSFrameInfo fi;
  vsFrameInfoInit(&fi, sp.width, sp.height, PF_YUV420P);
  VSMotionDetect md;
  memset(&md, 0, sizeof(md));
  VSMotionDetectConfig conf = vsMotionDetectGetDefaultConfig("golden");
  conf.shakiness = sp.shakiness;
  conf.accuracy = sp.accuracy;
  conf.stepSize = sp.stepsize;
  conf.contrastThreshold = sp.mincontrast;
  conf.virtualTripod = cs->tripod;
  conf.numThreads = threads;
  if (vsMotionDetectInit(&md, &conf, &fi) != VS_OK) { fprintf(stderr, "init failed\n"); exit(1); }
  const int ls = sp.width;
  uint8_t* Y = (uint8_t*)malloc((size_t)ls * sp.height);
  uint8_t* U = (uint8_t*)malloc((size_t)(sp.width / 2) * (sp.height / 2));
  uint8_t* V = (uint8_t*)malloc((size_t)(sp.width / 2) * (sp.height / 2));
  memset(U, 128, (size_t)(sp.width / 2) * (sp.height / 2));
  memset(V, 128, (size_t)(sp.width / 2) * (sp.height / 2));
  fprintf(out, "%s{\"name\":\"%s\",\"config\":%d,\"clip\":%d,\"width\":%d,\"height\":%d,\"stride\":%d,"
               "\"shakiness\":%d,\"accuracy\":%d,\"stepsize\":%d,\"mincontrast\":\"%016" PRIx64 "\","
               "\"tripod\":%d,\"frames\":[",
          first ? "" : ",\n", cs->name, cs->config, cs->clip, sp.width, sp.height, ls, sp.shakiness, sp.accuracy,
          sp.stepsize, dbits(sp.mincontrast), cs->tripod);
  long total = 0;
  for (int t = 0; t < nfr; t++) {
    vsyn_render(&clip, path, t, Y, ls);
    VSFrame fr;
    memset(&fr, 0, sizeof(fr));
    fr.data[0] = Y; fr.linesize[0] = ls;
    fr.data[1] = U; fr.linesize[1] = sp.width / 2;
    fr.data[2] = V; fr.linesize[2] = sp.width / 2;
    LocalMotions lms;
    const uint64_t hin = vsyn_hash_plane(Y, sp.width, sp.height, ls);
    vsMotionDetection(&md, &lms, &fr);
    const uint64_t hbl = vsyn_hash_plane(md.curr.data[0], sp.width, sp.height, md.curr.linesize[0]);
    const int n = vs_vector_size(&lms);
    total += n;
    fprintf(out, "%s[\"%016" PRIx64 "\",%d,\"%016" PRIx64 "\",\"%016" PRIx64 "\"]", t ? "," : "", hin, n,
            hash_motions(&lms), hbl);
    vs_vector_del(&lms);
  }
  fprintf(out, "]}");
  fflush(out);
  vsMotionDetectionCleanup(&md);
  vsyn_clip_free(&clip);
  free(path);
  free(Y); free(U); free(V);
  fprintf(stderr, "hashes of %s done (%d frames, %ld motions)\n", cs->name, nfr, total);
}

/* the reference's vsMotionsToTransform of each frame of a planar case */
static void run_case_transforms(FILE* out, const Case* cs, int first) {
  VSynSpec sp;
  vsyn_spec(cs->config, cs->clip, &sp);
  if (cs->width) { sp.width = cs->width; sp.height = cs->height; }
  if (cs->stepsize) sp.stepsize = cs->stepsize;
  if (cs->accuracy) sp.accuracy = cs->accuracy;
  if (cs->shakiness) sp.shakiness = cs->shakiness;
  if (cs->mincontrast >= 0) sp.mincontrast = cs->mincontrast;
  int nfr = cs->nframes < sp.nframes ? cs->nframes : sp.nframes;
  int total = sp.nframes > nfr ? sp.nframes : nfr;
  VSynClip clip;
  VSynMotion* path = (VSynMotion*)calloc((size_t)total, sizeof(VSynMotion));
  VSynSpec sp2 = sp; sp2.nframes = total;
  vsyn_open(&sp2, &clip, path);
  VSFrameInfo fi;
  vsFrameInfoInit(&fi, sp.width, sp.height, PF_YUV420P);
  VSMotionDetect md;
  memset(&md, 0, sizeof(md));
  VSMotionDetectConfig conf = vsMotionDetectGetDefaultConfig("golden");
  conf.shakiness = sp.shakiness;
  conf.accuracy = sp.accuracy;
  conf.stepSize = sp.stepsize;
  conf.contrastThreshold = sp.mincontrast;
  conf.virtualTripod = cs->tripod;
  conf.numThreads = 1;
  if (vsMotionDetectInit(&md, &conf, &fi) != VS_OK) { fprintf(stderr, "init failed\n"); exit(1); }
  VSTransformData td;
  memset(&td, 0, sizeof(td));
  td.fiSrc = fi;
  td.conf = vsTransformGetDefaultConfig("golden");
  td.conf.smoothZoom = 1;
  int ls = sp.width + cs->stride_pad;
  uint8_t* Y = (uint8_t*)malloc((size_t)ls * sp.height);
  uint8_t* U = (uint8_t*)malloc((size_t)(sp.width / 2) * (sp.height / 2));
  uint8_t* V = (uint8_t*)malloc((size_t)(sp.width / 2) * (sp.height / 2));
  memset(Y, 0x5A, (size_t)ls * sp.height);
  memset(U, 128, (size_t)(sp.width / 2) * (sp.height / 2));
  memset(V, 128, (size_t)(sp.width / 2) * (sp.height / 2));
  fprintf(out, "%s{\"name\":\"%s\",\"width\":%d,\"height\":%d,\"frames\":[", first ? "" : ",\n", cs->name,
          sp.width, sp.height);
  for (int t = 0; t < nfr; t++) {
    vsyn_render(&clip, path, t, Y, ls);
    VSFrame fr;
    memset(&fr, 0, sizeof(fr));
    fr.data[0] = Y; fr.linesize[0] = ls;
    fr.data[1] = U; fr.linesize[1] = sp.width / 2;
    fr.data[2] = V; fr.linesize[2] = sp.width / 2;
    LocalMotions lms;
    vsMotionDetection(&md, &lms, &fr);
    VSTransform tr = vsMotionsToTransform(&td, &lms, NULL);
    fprintf(out, "%s[\"%016" PRIx64 "\",\"%016" PRIx64 "\",\"%016" PRIx64 "\",\"%016" PRIx64 "\",%d]",
            t ? "," : "", dbits(tr.x), dbits(tr.y), dbits(tr.alpha), dbits(tr.zoom), tr.extra);
    vs_vector_del(&lms);
  }
  fprintf(out, "]}");
  vsMotionDetectionCleanup(&md);
  vsyn_clip_free(&clip);
  free(path);
  free(Y); free(U); free(V);
  fprintf(stderr, "transforms of case %s done (%d frames)\n", cs->name, nfr);
}

int main(int argc, char** argv) {
  const char* path = argc > 1 ? argv[1] : "detect_golden.json";
  vs_log_level = 0;
  static const Case planar[] = {
    /* name          cfg clip  w     h    nfr step acc shak tripod pad  contrast */
    { "c1_full",      1, 0,    0,    0,   60,  0,   0,  0,   0,    0,  -1 },
    { "c1_stride",    1, 1,    0,    0,   8,   0,   0,  0,   0,    40, -1 },
    { "c1_tripod",    1, 2,    0,    0,   10,  0,   0,  0,   3,    0,  -1 },
    { "c1_step5",     1, 3,    0,    0,   6,   5,   0,  0,   0,    0,  -1 },
    { "c1_step3",     1, 4,    0,    0,   6,   3,   0,  0,   0,    0,  -1 },
    { "c1_step8acc6", 1, 5,    0,    0,   6,   8,   6,  0,   0,    0,  -1 },
    { "c1_step1",     1, 6,    0,    0,   4,   1,   0,  0,   0,    0,  -1 },
    { "c1_acc4",      1, 7,    0,    0,   6,   0,   4,  3,   0,    0,  -1 },
    { "odd_502x298",  1, 8,    502,  298, 6,   0,   0,  0,   0,    3,  -1 },
    { "c2_1080p",     2, 0,    0,    0,   5,   0,   0,  0,   0,    0,  -1 },
    { "c3_4k",        3, 0,    0,    0,   3,   0,   0,  0,   0,    0,  -1 },
    { "c4_clip0",     4, 0,    0,    0,   4,   0,   0,  0,   0,    0,  -1 },
    { "c4_clip1",     4, 1,    0,    0,   4,   0,   0,  0,   0,    0,  -1 },
    { "c4_clip2",     4, 2,    0,    0,   4,   0,   0,  0,   0,    0,  -1 },
    { "c5_4kblur",    5, 0,    0,    0,   2,   0,   0,  0,   0,    0,  -1 },
  };
  /* packed RGB24 / BGR24 / RGBA frames: vsyn_pack of the rendered luma */
  static const Case packed[] = {
    /* name            cfg clip  w     h    nfr step acc shak tripod pad  contrast  format */
    { "p_rgb24_c1",      1, 20,   0,    0,   8,   0,   0,  0,   0,    0,  -1, PF_RGB24 },
    { "p_bgr24_pad",     1, 21,   0,    0,   5,   0,   0,  0,   0,    40, -1, PF_BGR24 },
    { "p_rgba_c1",       1, 22,   0,    0,   6,   0,   0,  0,   0,    0,  -1, PF_RGBA },
    { "p_rgb24_tripod",  1, 23,   0,    0,   6,   0,   0,  0,   2,    0,  -1, PF_RGB24 },
    { "p_rgb24_step4",   1, 24,   0,    0,   5,   4,   0,  0,   0,    0,  -1, PF_RGB24 },
    { "p_rgba_step8acc6",1, 25,   0,    0,   5,   8,   6,  0,   0,    0,  -1, PF_RGBA },
    { "p_rgb24_step1",   1, 26,   0,    0,   3,   1,   0,  0,   0,    0,  -1, PF_RGB24 },
    { "p_rgba_odd",      1, 27,   502,  298, 5,   5,   0,  0,   0,    12, -1, PF_RGBA },
    { "p_rgb24_1080p",   2, 20,   0,    0,   3,   0,   0,  0,   0,    0,  -1, PF_RGB24 },
  };
  /* show = 1 / 2: the overlays the reference draws into the caller's frame */
  static const Case shows[] = {
    /* name            cfg clip  w     h    nfr step acc shak tripod pad  contrast  format    show */
    { "s_c1_show1",      1, 60,   0,    0,   6,   0,   0,  0,   0,    0,  -1, 0,        1 },
    { "s_c1_show2",      1, 61,   0,    0,   6,   0,   0,  0,   0,    24, -1, 0,        2 },
    { "s_c1_step3_show2",1, 62,   0,    0,   4,   3,   0,  0,   0,    0,  -1, 0,        2 },
    { "s_odd_show2",     1, 63,   502,  298, 5,   0,   0,  0,   0,    0,  -1, 0,        2 },
    { "s_rgba_show2",    1, 64,   0,    0,   5,   0,   0,  0,   0,    0,  -1, PF_RGBA,  2 },
    { "s_bgr24_show1",   1, 65,   0,    0,   5,   0,   0,  0,   0,    12, -1, PF_BGR24, 1 },
    { "s_1080p_show2",   2, 60,   0,    0,   3,   0,   0,  0,   0,    0,  -1, 0,        2 },
  };
  if (argc > 2 && strcmp(argv[2], "hashes") == 0) {
    /* make_golden out.json hashes <threads> [config clip]...: whole clips of
       the workload specs; default: the full 1080p (300 frames) and 4K step-4
       (240 frames) clips and all 64 clips of the multi-clip workload */
    const int threads = argc > 3 ? atoi(argv[3]) : 8;
    FILE* out = fopen(path, "w");
    if (!out) return 1;
    fprintf(out, "{\"source\":\"reference vsMotionDetection (oracle/_ref/libvsref.so), whole clips\",\"cases\":[\n");
    int first = 1;
    char name[64];
    if (argc > 5) {
      for (int a = 4; a + 1 < argc; a += 2) {
        Case c = {0};
        c.config = atoi(argv[a]);
        c.clip = atoi(argv[a + 1]);
        snprintf(name, sizeof(name), "w%d_clip%d", c.config, c.clip);
        c.name = name;
        run_case_hashes(out, &c, first, threads);
        first = 0;
      }
    } else {
      Case c2 = {0}, c3 = {0};
      c2.name = "w2_full"; c2.config = 2;
      c3.name = "w3_full"; c3.config = 3;
      run_case_hashes(out, &c2, 1, threads);
      run_case_hashes(out, &c3, 0, threads);
      for (int k = 0; k < 64; k++) {
        Case c = {0};
        c.config = 4;
        c.clip = k;
        snprintf(name, sizeof(name), "w4_clip%d", k);
        c.name = name;
        run_case_hashes(out, &c, 0, threads);
      }
    }
    fprintf(out, "\n]}\n");
    fclose(out);
    return 0;
  }
  const int useTransforms = argc > 2 && strcmp(argv[2], "transforms") == 0;
  if (useTransforms) {
    FILE* out = fopen(path, "w");
    if (!out) return 1;
    fprintf(out, "{\"source\":\"reference vsMotionsToTransform (oracle/_ref/libvsref.so)\",\"cases\":[\n");
    for (size_t i = 0; i < sizeof(planar) / sizeof(planar[0]); i++) run_case_transforms(out, &planar[i], i == 0);
    fprintf(out, "\n]}\n");
    fclose(out);
    return 0;
  }
  const int usePacked = argc > 2 && strcmp(argv[2], "packed") == 0;
  const int useShow = argc > 2 && strcmp(argv[2], "show") == 0;
  const Case* cases = useShow ? shows : usePacked ? packed : planar;
  const size_t ncases = useShow     ? sizeof(shows) / sizeof(shows[0])
                        : usePacked ? sizeof(packed) / sizeof(packed[0])
                                    : sizeof(planar) / sizeof(planar[0]);
  FILE* out = fopen(path, "w");
  if (!out) return 1;
  fprintf(out, "{\"source\":\"reference vsMotionDetection (oracle/_ref/libvsref.so)\",\"cases\":[\n");
  int only = argc > 3 ? atoi(argv[3]) : -1;
  int first = 1;
  for (size_t i = 0; i < ncases; i++) {
    if (only >= 0 && (int)i != only) continue;
    run_case(out, &cases[i], first);
    first = 0;
  }
  fprintf(out, "\n]}\n");
  fclose(out);
  return 0;
}
