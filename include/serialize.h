/*
 * serialize.h -- .trf writer for the detect pass (the consumer of
 * LocalMotions that ffmpeg's vidstabdetect calls after every frame).
 *
 * The byte format is the reference's (docs/trf-format.md:81-144,
 * src/serialize.c:144-190, 252-280, 355-409): binary files start with
 * "TRF" + version byte + int32 accuracy, shakiness, stepsize + double
 * mincontrast; each frame is int32 frameNum, int32 n, then n packed 26-byte
 * records (int16 v.x v.y f.x f.y f.size, double contrast, double match),
 * little endian.  The text format is the reference's "VID.STAB 1" layout.
 */
#ifndef __SERIALIZE_H
#define __SERIALIZE_H

#include <stdio.h>
#include "transformtype.h"
#include "motiondetect.h"
#include "vidstab_api.h"

/* src/serialize.h:30 */
#define LIBVIDSTAB_FILE_FORMAT_VERSION 1

/* the LocalMotions of every frame of a clip (elements are LocalMotions);
   src/serialize.h:59-62 */
typedef VSVector VSManyLocalMotions;
#define VSMLMGet(manylocalmotions,index) \
    ((LocalMotions*)vs_vector_get(manylocalmotions,index))

/* one frame's LocalMotions in the file's mode; src/serialize.c:252-280 */
VS_API int vsStoreLocalmotions(FILE* f, const LocalMotions* lms, const int serializationMode);
/* reads one LocalMotions list back; src/serialize.c:282-350 */
VS_API LocalMotions vsRestoreLocalmotions(FILE* f, const int serializationMode);

/* header; VS_OK / VS_ERROR.  src/serialize.h:74, src/serialize.c:355-382 */
VS_API int vsPrepareFile(const VSMotionDetect* td, FILE* f);
/* one frame record; src/serialize.h:77, src/serialize.c:384-409 */
VS_API int vsWriteToFile(const VSMotionDetect* td, FILE* f, const LocalMotions* lms);

#endif /* __SERIALIZE_H */
