/*
 * vidstab_api.h -- symbol visibility for the MI355X detect library.
 *
 * Same macro name and meaning as the reference's src/vidstab_api.h:42-61:
 * every exported entry point is tagged VS_API, so a caller compiled against
 * the reference headers links against this library unchanged.  Only the
 * ELF/GCC branch matters here (the library is built for Linux/ROCm).
 */
#ifndef VIDSTAB_API_H
#define VIDSTAB_API_H

#if defined(VIDSTAB_STATIC_DEFINE)
#  define VS_API
#elif defined(__GNUC__) && __GNUC__ >= 4
#  define VS_API __attribute__((visibility("default")))
#else
#  define VS_API
#endif

#endif /* VIDSTAB_API_H */
