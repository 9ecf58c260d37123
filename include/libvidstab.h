/*
 * libvidstab.h -- the aggregate header of the vid.stab C API as this library
 * provides it (reference: src/libvidstab.h:28-38).  ffmpeg's vidstab filters
 * include it as <vid.stab/libvidstab.h>.
 *
 * This library implements the detect side: frame descriptions, the detector,
 * the LocalMotions container, the .trf writer and the per-frame motion ->
 * transform fit.  The reference's transform.h (the warping pass, used by
 * ffmpeg's vidstabtransform) is not part of it; a build that also uses the
 * warping pass compiles against the reference's headers, whose declarations
 * of the functions below are identical, and links this library ahead of the
 * reference library.
 */
#ifndef LIBVIDSTAB_H
#define LIBVIDSTAB_H

/* the interface version whose struct layouts this library keeps
   (src/libvidstab_version.h.in with version.cmake) */
#ifndef LIBVIDSTAB_VERSION_MAJOR
#define LIBVIDSTAB_VERSION_MAJOR 1
#define LIBVIDSTAB_VERSION_MINOR 4
#define LIBVIDSTAB_VERSION_MICRO 0
#define LIBVIDSTAB_VERSION_MAKE(a, b, c) (((a) << 16) | ((b) << 8) | (c))
#define LIBVIDSTAB_VERSION_INT \
  LIBVIDSTAB_VERSION_MAKE(LIBVIDSTAB_VERSION_MAJOR, LIBVIDSTAB_VERSION_MINOR, LIBVIDSTAB_VERSION_MICRO)
#define LIBVIDSTAB_VERSION "v1.4.0"
#endif

#include "frameinfo.h"
#include "motiondetect.h"
#include "vsvector.h"
#include "serialize.h"
#include "localmotion2transform.h"

#endif /* LIBVIDSTAB_H */
