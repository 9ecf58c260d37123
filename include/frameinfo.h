/*
 * frameinfo.h -- frame description types of the vid.stab ABI.
 *
 * Layouts and enum values are fixed by the reference (src/frameinfo.h:34-67):
 * ffmpeg fills VSFrame from an AVFrame and reads VSFrameInfo out of the
 * caller-owned VSMotionDetect, so neither may change.
 */
#ifndef FRAMEINFO_H
#define FRAMEINFO_H

#include <stddef.h>
#include <stdlib.h>
#include <inttypes.h>
#include "vidstab_api.h"

/* pixel formats -- values as in src/frameinfo.h:34-48 */
typedef enum {
  PF_NONE = -1,
  PF_GRAY8,      /* Y only, 8 bpp */
  PF_YUV420P,    /* planar 4:2:0 */
  PF_YUV422P,    /* planar 4:2:2 */
  PF_YUV444P,    /* planar 4:4:4 */
  PF_YUV410P,    /* planar 4:1:0 */
  PF_YUV411P,    /* planar 4:1:1 */
  PF_YUV440P,    /* planar 4:4:0 */
  PF_YUVA420P,   /* planar 4:2:0 plus alpha plane */
  PF_PACKED,     /* marker: packed formats follow */
  PF_RGB24,
  PF_BGR24,
  PF_RGBA,
  PF_NUMBER      /* number of formats */
} VSPixelFormat;

/* 28 bytes; src/frameinfo.h:53-60 */
typedef struct vsframeinfo {
  int width, height;
  int planes;        /* 1 luma, 2/3 chroma, 4 alpha */
  int log2ChromaW;   /* horizontal chroma subsampling shift */
  int log2ChromaH;   /* vertical chroma subsampling shift */
  VSPixelFormat pFormat;
  int bytesPerPixel; /* packed formats only */
} VSFrameInfo;

/* 48 bytes; src/frameinfo.h:63-67.  linesize is in bytes and may exceed the
   plane width (caller stride). */
typedef struct vsframe {
  uint8_t* data[4];
  int linesize[4];
} VSFrame;

/* chroma plane size with round-up */
#define CHROMA_SIZE(width,log2sub)  (-(-(width) >> (log2sub)))

/* Returns 1 on success and 0 on an unknown format or dimensions that do not
   fit the chroma subsampling (a boolean, not VS_OK/VS_ERROR).
   Reference: src/frameinfo.h:72-84, src/frameinfo.c:32-115. */
VS_API int vsFrameInfoInit(VSFrameInfo* fi, int width, int height, VSPixelFormat pFormat);
/* src/frameinfo.c:117-123 */
VS_API int vsGetPlaneWidthSubS(const VSFrameInfo* fi, int plane);
VS_API int vsGetPlaneHeightSubS(const VSFrameInfo* fi, int plane);
/* src/frameinfo.c:125-137 */
VS_API void vsFrameNull(VSFrame* frame);
VS_API int vsFrameIsNull(const VSFrame* frame);
VS_API int vsFramesEqual(const VSFrame* frame1, const VSFrame* frame2);
/* zero-filled planes with linesize == plane width; src/frameinfo.c:139-161 */
VS_API void vsFrameAllocate(VSFrame* frame, const VSFrameInfo* fi);
/* src/frameinfo.c:163-188 */
VS_API void vsFrameCopyPlane(VSFrame* dest, const VSFrame* src, const VSFrameInfo* fi, int plane);
VS_API void vsFrameCopy(VSFrame* dest, const VSFrame* src, const VSFrameInfo* fi);
/* points the planes into one linear buffer, no copy; src/frameinfo.c:190-202 */
VS_API void vsFrameFillFromBuffer(VSFrame* frame, uint8_t* img, const VSFrameInfo* fi);
/* src/frameinfo.c:204-211 */
VS_API void vsFrameFree(VSFrame* frame);

#endif /* FRAMEINFO_H */
