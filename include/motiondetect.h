/*
 * motiondetect.h -- the detect-pass C API, served by MI355X kernels.
 *
 * Drop-in for the reference's src/motiondetect.h.  The struct layouts below
 * are frozen ABI: ffmpeg's vidstabdetect filter embeds VSMotionDetect by value
 * (496 bytes on x86-64) and reads md->fi directly, so the GPU state of a
 * detector lives in a side table keyed by the VSMotionDetect address, never
 * inside the struct.
 */
#ifndef MOTIONDETECT_H
#define MOTIONDETECT_H

#include <stddef.h>
#include <stdlib.h>

#include "transformtype.h"
#include "vidstabdefines.h"
#include "vsvector.h"
#include "frameinfo.h"
#include "vidstab_api.h"

/* src/motiondetect.h:42-43 */
#define ASCII_SERIALIZATION_MODE 1
#define BINARY_SERIALIZATION_MODE 2

/* 48 bytes; src/motiondetect.h:45-58 */
typedef struct _vsmotiondetectconfig {
  int         shakiness;         /* 1..10; only raises the accuracy floor */
  int         accuracy;          /* 1..15; share of fields used */
  int         stepSize;          /* grid step of the coarse search */
  int         algo;              /* unused */
  int         virtualTripod;     /* >0: compare against frame virtualTripod-1 */
  int         show;              /* 1,2: draw fields into the caller's frame */
  double      contrastThreshold; /* fields below this contrast are skipped */
  const char* modName;           /* log tag */
  int         numThreads;        /* kept for ABI; the GPU ignores it */
} VSMotionDetectConfig;

/* 104 bytes; src/motiondetect.h:61-73 */
typedef struct _vsmotiondetectfields {
  int maxShift;
  int stepSize;
  int fieldNum;
  int maxFields;
  double contrastThreshold;
  int fieldSize;
  int fieldRows;
  Field* fields;
  short useOffset;
  VSTransform offset;
} VSMotionDetectFields;

/* 496 bytes; src/motiondetect.h:76-93.  curr/currtmp/prev stay null in this
   library: the blurred planes live in device memory. */
typedef struct _vsmotiondetect {
  VSFrameInfo fi;
  VSMotionDetectConfig conf;
  VSMotionDetectFields fieldscoarse;
  VSMotionDetectFields fieldsfine;
  VSFrame curr;
  VSFrame currorig;
  VSFrame currtmp;
  VSFrame prev;
  short hasSeenOneFrame;
  int initialized;
  int serializationMode;
  int frameNum;
} VSMotionDetect;

/* defaults: step 6, accuracy 15, shakiness 5, contrast 0.25;
   src/motiondetect.h:121, src/motiondetect.c:60-71 */
VS_API VSMotionDetectConfig vsMotionDetectGetDefaultConfig(const char* modName);

/* Validates the format, clamps the config, lays out the coarse and fine field
   grids and creates the GPU context.  VS_OK or VS_ERROR; VS_ERROR also for
   frames too small for the grids (a coarse field moved as far as its search
   reaches, or a fine field, would leave the frame).
   src/motiondetect.h:123-128, src/motiondetect.c:83-165 */
VS_API int vsMotionDetectInit(VSMotionDetect* md, const VSMotionDetectConfig* conf,
                              const VSFrameInfo* fi);

/* One detect step on a host frame (caller stride).  *motions receives
   coarse ++ cleaned fine LocalMotions, each element vs_malloc'd; the caller
   frees them with vs_vector_del.  src/motiondetect.h:130-136,
   src/motiondetect.c:191-305 */
VS_API int vsMotionDetection(VSMotionDetect* md, LocalMotions* motions, VSFrame* frame);

/* Releases host and device state.  src/motiondetect.h:138-141,
   src/motiondetect.c:167-182 */
VS_API void vsMotionDetectionCleanup(VSMotionDetect* md);

/* src/motiondetect.h:143-147, src/motiondetect.c:73-81 */
VS_API void vsMotionDetectGetConfig(VSMotionDetectConfig* conf, const VSMotionDetect* md);
VS_API const VSFrameInfo* vsMotionDetectGetFrameInfo(const VSMotionDetect* md);

#endif /* MOTIONDETECT_H */
