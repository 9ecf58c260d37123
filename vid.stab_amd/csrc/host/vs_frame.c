/*
 * vs_frame.c -- frame descriptions of the detect ABI (interface:
 * include/frameinfo.h, which keeps the reference's src/frameinfo.h:34-133).
 *
 * A pixel format is a row of a table: how many planes it has, the chroma
 * subsampling shifts of planes 1 and 2, and the bytes per pixel of a packed
 * format (whose pixels all live in plane 0).  Every plane helper derives its
 * row bytes and row count from that row.  These helpers only ever touch host
 * memory: the detect pass keeps its working planes on the GPU.
 */
#include <string.h>

#include "frameinfo.h"
#include "vidstabdefines.h"

typedef struct {
  VSPixelFormat format;
  int planes;
  int log2w, log2h;   /* chroma subsampling of planes 1 and 2 */
  int bpp;            /* bytes per pixel (packed formats), else 1 */
} FormatRow;

static const FormatRow kFormats[] = {
    {PF_GRAY8, 1, 0, 0, 1},    {PF_YUV420P, 3, 1, 1, 1}, {PF_YUV422P, 3, 1, 0, 1},
    {PF_YUV444P, 3, 0, 0, 1},  {PF_YUV410P, 3, 2, 2, 1}, {PF_YUV411P, 3, 2, 0, 1},
    {PF_YUV440P, 3, 0, 1, 1},  {PF_YUVA420P, 4, 1, 1, 1}, {PF_RGB24, 1, 0, 0, 3},
    {PF_BGR24, 1, 0, 0, 3},    {PF_RGBA, 1, 0, 0, 4},
};

static const FormatRow* format_row(VSPixelFormat f) {
  for (size_t i = 0; i < sizeof(kFormats) / sizeof(kFormats[0]); ++i)
    if (kFormats[i].format == f) return &kFormats[i];
  return NULL;
}

/* Returns 1, or 0 for an unknown format or a size the subsampling cannot
   split; a failed call leaves pFormat at 0, as the reference's does. */
VS_API int vsFrameInfoInit(VSFrameInfo* fi, int width, int height, VSPixelFormat pFormat) {
  const FormatRow* r = format_row(pFormat);
  fi->width = width;
  fi->height = height;
  fi->pFormat = pFormat;
  fi->planes = r ? r->planes : 3;
  fi->log2ChromaW = r ? r->log2w : 0;
  fi->log2ChromaH = r ? r->log2h : 0;
  fi->bytesPerPixel = r ? r->bpp : 1;
  if (!r) {
    fi->pFormat = 0;
    return 0;
  }
  const int xdiv = 1 << r->log2w, ydiv = 1 << r->log2h;
  if (width <= 0 || height <= 0 || width % xdiv || height % ydiv) {
    vs_log_error("vid.stab", "frame size %ix%i does not fit pixel format %i (needs positive multiples of %i x %i)\n",
                 width, height, (int)pFormat, xdiv, ydiv);
    fi->pFormat = 0;
    return 0;
  }
  return 1;
}

VS_API int vsGetPlaneWidthSubS(const VSFrameInfo* fi, int plane) {
  return plane == 1 || plane == 2 ? fi->log2ChromaW : 0;
}

VS_API int vsGetPlaneHeightSubS(const VSFrameInfo* fi, int plane) {
  return plane == 1 || plane == 2 ? fi->log2ChromaH : 0;
}

/* bytes of image data in one row of `plane`, and its number of rows */
static void plane_shape(const VSFrameInfo* fi, int plane, int* rowBytes, int* rows) {
  *rowBytes = (fi->width >> vsGetPlaneWidthSubS(fi, plane)) * fi->bytesPerPixel;
  *rows = fi->height >> vsGetPlaneHeightSubS(fi, plane);
}

static int plane_count(const VSFrameInfo* fi) { return fi->planes < 4 ? fi->planes : 4; }

VS_API int vsFrameIsNull(const VSFrame* frame) { return !frame || !frame->data[0]; }

VS_API int vsFramesEqual(const VSFrame* a, const VSFrame* b) {
  if (!a || !b) return 0;
  return a == b || a->data[0] == b->data[0];
}

VS_API void vsFrameNull(VSFrame* frame) { memset(frame, 0, sizeof(*frame)); }

/* zero-filled planes, rows packed (linesize = row bytes); packed formats get
   their single plane 0 */
VS_API void vsFrameAllocate(VSFrame* frame, const VSFrameInfo* fi) {
  vsFrameNull(frame);
  for (int p = 0; p < plane_count(fi); ++p) {
    int rowBytes, rows;
    plane_shape(fi, p, &rowBytes, &rows);
    frame->linesize[p] = rowBytes;
    frame->data[p] = (uint8_t*)vs_zalloc((size_t)rowBytes * (size_t)rows);
    if (!frame->data[p]) vs_log_error("vid.stab", "no memory for a %i x %i frame plane\n", rowBytes, rows);
  }
}

VS_API void vsFrameCopyPlane(VSFrame* dest, const VSFrame* src, const VSFrameInfo* fi, int plane) {
  int rowBytes, rows;
  plane_shape(fi, plane, &rowBytes, &rows);
  const int ds = dest->linesize[plane], ss = src->linesize[plane];
  if (ds == ss) {
    /* same layout: one block, padding included */
    memcpy(dest->data[plane], src->data[plane], (size_t)ss * (size_t)rows);
    return;
  }
  for (int y = 0; y < rows; ++y)
    memcpy(dest->data[plane] + (size_t)y * ds, src->data[plane] + (size_t)y * ss, (size_t)rowBytes);
}

VS_API void vsFrameCopy(VSFrame* dest, const VSFrame* src, const VSFrameInfo* fi) {
  for (int p = 0; p < plane_count(fi); ++p) vsFrameCopyPlane(dest, src, fi, p);
}

/* planes back to back in one caller buffer, no copy */
VS_API void vsFrameFillFromBuffer(VSFrame* frame, uint8_t* img, const VSFrameInfo* fi) {
  vsFrameNull(frame);
  size_t at = 0;
  for (int p = 0; p < plane_count(fi); ++p) {
    int rowBytes, rows;
    plane_shape(fi, p, &rowBytes, &rows);
    frame->data[p] = img + at;
    frame->linesize[p] = rowBytes;
    at += (size_t)rowBytes * (size_t)rows;
  }
}

VS_API void vsFrameFree(VSFrame* frame) {
  for (int p = 0; p < 4; ++p)
    if (frame->data[p]) vs_free(frame->data[p]);
  vsFrameNull(frame);
}
