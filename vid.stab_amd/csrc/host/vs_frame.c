/*
 * vs_frame.c -- frame description helpers of the vid.stab ABI
 * (reference: src/frameinfo.c:32-211).  Host memory only; the detect pass
 * keeps its working planes in device memory.
 */
#include <assert.h>
#include <string.h>

#include "frameinfo.h"
#include "vidstabdefines.h"

/* src/frameinfo.c:32-115: fills fi; 1 on success, 0 on a bad format/size */
VS_API int vsFrameInfoInit(VSFrameInfo* fi, int width, int height, VSPixelFormat pFormat) {
  fi->pFormat = pFormat;
  fi->width = width;
  fi->height = height;
  fi->planes = 3;
  fi->log2ChromaW = 0;
  fi->log2ChromaH = 0;
  fi->bytesPerPixel = 1;
  switch (pFormat) {
    case PF_GRAY8:    fi->planes = 1; break;
    case PF_YUV420P:  fi->log2ChromaW = 1; fi->log2ChromaH = 1; break;
    case PF_YUV422P:  fi->log2ChromaW = 1; break;
    case PF_YUV444P:  break;
    case PF_YUV410P:  fi->log2ChromaW = 2; fi->log2ChromaH = 2; break;
    case PF_YUV411P:  fi->log2ChromaW = 2; break;
    case PF_YUV440P:  fi->log2ChromaH = 1; break;
    case PF_YUVA420P: fi->log2ChromaW = 1; fi->log2ChromaH = 1; fi->planes = 4; break;
    case PF_RGB24:
    case PF_BGR24:    fi->bytesPerPixel = 3; fi->planes = 1; break;
    case PF_RGBA:     fi->bytesPerPixel = 4; fi->planes = 1; break;
    default:
      fi->pFormat = 0;
      return 0;
  }
  if (width <= 0 || height <= 0) {
    vs_log_error("vid.stab", "invalid frame dimensions %ix%i: width and height must be positive\n",
                 width, height);
    fi->pFormat = 0;
    return 0;
  }
  if (width % (1 << fi->log2ChromaW) != 0 || height % (1 << fi->log2ChromaH) != 0) {
    vs_log_error("vid.stab", "invalid frame dimensions %ix%i for pixel format %i: "
                 "width must be a multiple of %i and height a multiple of %i\n",
                 width, height, (int)pFormat, 1 << fi->log2ChromaW, 1 << fi->log2ChromaH);
    fi->pFormat = 0;
    return 0;
  }
  return 1;
}

VS_API int vsGetPlaneWidthSubS(const VSFrameInfo* fi, int plane) {
  return (plane == 1 || plane == 2) ? fi->log2ChromaW : 0;
}

VS_API int vsGetPlaneHeightSubS(const VSFrameInfo* fi, int plane) {
  return (plane == 1 || plane == 2) ? fi->log2ChromaH : 0;
}

VS_API int vsFrameIsNull(const VSFrame* frame) { return frame == 0 || frame->data[0] == 0; }

VS_API int vsFramesEqual(const VSFrame* a, const VSFrame* b) {
  return a && b && (a == b || a->data[0] == b->data[0]);
}

VS_API void vsFrameNull(VSFrame* frame) {
  memset(frame->data, 0, sizeof(frame->data));
  memset(frame->linesize, 0, sizeof(frame->linesize));
}

VS_API void vsFrameAllocate(VSFrame* frame, const VSFrameInfo* fi) {
  vsFrameNull(frame);
  if (fi->pFormat < PF_PACKED) {
    for (int i = 0; i < fi->planes && i < 4; i++) {
      int w = fi->width >> vsGetPlaneWidthSubS(fi, i);
      int h = fi->height >> vsGetPlaneHeightSubS(fi, i);
      frame->data[i] = (uint8_t*)vs_zalloc((size_t)w * h);
      frame->linesize[i] = w;
      if (!frame->data[i]) vs_log_error("vid.stab", "out of memory: cannot allocated buffer");
    }
  } else {
    int w = fi->width, h = fi->height;
    frame->data[0] = (uint8_t*)vs_zalloc((size_t)w * h * fi->bytesPerPixel);
    frame->linesize[0] = w * fi->bytesPerPixel;
    if (!frame->data[0]) vs_log_error("vid.stab", "out of memory: cannot allocated buffer");
  }
}

VS_API void vsFrameCopyPlane(VSFrame* dest, const VSFrame* src, const VSFrameInfo* fi, int plane) {
  assert(src->data[plane]);
  int h = fi->height >> vsGetPlaneHeightSubS(fi, plane);
  if (src->linesize[plane] == dest->linesize[plane]) {
    memcpy(dest->data[plane], src->data[plane], (size_t)src->linesize[plane] * h);
  } else {
    int w = (fi->width >> vsGetPlaneWidthSubS(fi, plane)) * fi->bytesPerPixel;
    uint8_t* d = dest->data[plane];
    const uint8_t* s = src->data[plane];
    for (; h > 0; h--) {
      memcpy(d, s, (size_t)w);
      d += dest->linesize[plane];
      s += src->linesize[plane];
    }
  }
}

VS_API void vsFrameCopy(VSFrame* dest, const VSFrame* src, const VSFrameInfo* fi) {
  for (int plane = 0; plane < fi->planes && plane < 4; plane++) vsFrameCopyPlane(dest, src, fi, plane);
}

VS_API void vsFrameFillFromBuffer(VSFrame* frame, uint8_t* img, const VSFrameInfo* fi) {
  vsFrameNull(frame);
  long offset = 0;
  for (int i = 0; i < fi->planes && i < 4; i++) {
    int w = fi->width >> vsGetPlaneWidthSubS(fi, i);
    int h = fi->height >> vsGetPlaneHeightSubS(fi, i);
    frame->data[i] = img + offset;
    frame->linesize[i] = w * fi->bytesPerPixel;
    offset += (long)h * w * fi->bytesPerPixel;
  }
}

VS_API void vsFrameFree(VSFrame* frame) {
  for (int plane = 0; plane < 4; plane++) {
    if (frame->data[plane]) vs_free(frame->data[plane]);
    frame->data[plane] = 0;
    frame->linesize[plane] = 0;
  }
}
