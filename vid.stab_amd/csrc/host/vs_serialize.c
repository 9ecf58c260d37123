/*
 * vs_serialize.c -- the .trf writer the detect caller uses after every frame,
 * plus the LocalMotions list reader for round trips (reference:
 * src/serialize.c:144-190, 252-350, 355-409; format: docs/trf-format.md).
 *
 * Binary records are written field by field (26 packed bytes), little endian,
 * which is the host order on x86-64.
 */
#include <stdio.h>
#include <string.h>

#include "serialize.h"
#include "transformtype_operations.h"
#include "vidstabdefines.h"

static const char* kMod = "vid.stab - serialization";

#define VS_MAX_LOCALMOTIONS_PER_FRAME (1 << 20)

static int put16(int16_t v, FILE* f) { return (int)fwrite(&v, sizeof(v), 1, f); }
static int put32(int32_t v, FILE* f) { return (int)fwrite(&v, sizeof(v), 1, f); }
static int putd(double v, FILE* f) { return (int)fwrite(&v, sizeof(v), 1, f); }

static int store_lm_binary(FILE* f, const LocalMotion* lm) {
  return put16(lm->v.x, f) > 0 && put16(lm->v.y, f) > 0 && put16(lm->f.x, f) > 0 &&
         put16(lm->f.y, f) > 0 && put16(lm->f.size, f) > 0 && putd(lm->contrast, f) > 0 &&
         putd(lm->match, f) > 0;
}

static int store_lm_text(FILE* f, const LocalMotion* lm) {
  return fprintf(f, "(LM %hi %hi %hi %hi %hi %lf %lf)", lm->v.x, lm->v.y, lm->f.x, lm->f.y,
                 lm->f.size, lm->contrast, lm->match);
}

VS_API int vsStoreLocalmotions(FILE* f, const LocalMotions* lms, const int mode) {
  const int len = vs_vector_size(lms);
  if (mode == BINARY_SERIALIZATION_MODE) {
    if (put32(len, f) <= 0) return 0;
    for (int i = 0; i < len; i++)
      if (!store_lm_binary(f, LMGet(lms, i))) return 0;
    return 1;
  }
  fprintf(f, "List %i [", len);
  for (int i = 0; i < len; i++) {
    if (i > 0) fprintf(f, ",");
    if (store_lm_text(f, LMGet(lms, i)) <= 0) return 0;
  }
  fprintf(f, "]");
  return 1;
}

static int read_lm_binary(FILE* f, LocalMotion* lm) {
  memset(lm, 0, sizeof(*lm));
  return fread(&lm->v.x, 2, 1, f) == 1 && fread(&lm->v.y, 2, 1, f) == 1 &&
         fread(&lm->f.x, 2, 1, f) == 1 && fread(&lm->f.y, 2, 1, f) == 1 &&
         fread(&lm->f.size, 2, 1, f) == 1 && fread(&lm->contrast, 8, 1, f) == 1 &&
         fread(&lm->match, 8, 1, f) == 1;
}

static LocalMotion read_lm_text(FILE* f) {
  LocalMotion lm = null_localmotion();
  int c;
  if (fscanf(f, "(LM %hi %hi %hi %hi %hi %lf %lf", &lm.v.x, &lm.v.y, &lm.f.x, &lm.f.y, &lm.f.size,
             &lm.contrast, &lm.match) != 7) {
    vs_log_error(kMod, "Cannot parse localmotion!\n");
    return null_localmotion();
  }
  while ((c = fgetc(f)) && c != ')' && c != EOF) {
  }
  return lm;
}

VS_API LocalMotions vsRestoreLocalmotions(FILE* f, const int mode) {
  LocalMotions lms;
  int len = 0;
  vs_vector_init(&lms, 0);
  if (mode == BINARY_SERIALIZATION_MODE) {
    if (fread(&len, 4, 1, f) != 1) {
      vs_log_error(kMod, "Cannot parse localmotions list: unexpected end of file\n");
      return lms;
    }
    if (len < 0 || len > VS_MAX_LOCALMOTIONS_PER_FRAME) {
      vs_log_error(kMod, "Implausible number of localmotions (%i)\n", len);
      return lms;
    }
    if (len > 0) vs_vector_init(&lms, len);
    for (int i = 0; i < len; i++) {
      LocalMotion lm;
      if (!read_lm_binary(f, &lm)) {
        vs_log_error(kMod, "Cannot parse localmotion: unexpected end of file\n");
        break;
      }
      vs_vector_append_dup(&lms, &lm, sizeof(LocalMotion));
    }
    return lms;
  }
  int c;
  if (fscanf(f, "List %i [", &len) != 1) {
    vs_log_error(kMod, "Cannot parse localmotions list expect 'List len ['!\n");
    return lms;
  }
  if (len > 0) {
    vs_vector_init(&lms, len);
    for (int i = 0; i < len; i++) {
      if (i > 0)
        while ((c = fgetc(f)) && c != ',' && c != EOF) {
        }
      LocalMotion lm = read_lm_text(f);
      vs_vector_append_dup(&lms, &lm, sizeof(LocalMotion));
    }
  }
  while ((c = fgetc(f)) && c != ']' && c != EOF) {
  }
  return lms;
}

VS_API int vsPrepareFile(const VSMotionDetect* md, FILE* f) {
  if (!f) return VS_ERROR;
  if (md->serializationMode == BINARY_SERIALIZATION_MODE) {
    fprintf(f, "TRF%hhu", (unsigned char)LIBVIDSTAB_FILE_FORMAT_VERSION);
    put32(md->conf.accuracy, f);
    put32(md->conf.shakiness, f);
    put32(md->conf.stepSize, f);
    putd(md->conf.contrastThreshold, f);
    return VS_OK;
  }
  fprintf(f, "VID.STAB %i\n", LIBVIDSTAB_FILE_FORMAT_VERSION);
  fprintf(f, "#      accuracy = %d\n", md->conf.accuracy);
  fprintf(f, "#     shakiness = %d\n", md->conf.shakiness);
  fprintf(f, "#      stepsize = %d\n", md->conf.stepSize);
  fprintf(f, "#   mincontrast = %f\n", md->conf.contrastThreshold);
  return VS_OK;
}

VS_API int vsWriteToFile(const VSMotionDetect* md, FILE* f, const LocalMotions* lms) {
  if (!f || !lms) return VS_ERROR;
  if (md->serializationMode == BINARY_SERIALIZATION_MODE) {
    if (put32(md->frameNum, f) <= 0) return VS_ERROR;
    if (vsStoreLocalmotions(f, lms, BINARY_SERIALIZATION_MODE) <= 0) return VS_ERROR;
    return VS_OK;
  }
  if (fprintf(f, "Frame %i (", md->frameNum) > 0 &&
      vsStoreLocalmotions(f, lms, ASCII_SERIALIZATION_MODE) > 0 && fprintf(f, ")\n"))
    return VS_OK;
  return VS_ERROR;
}
