/*
 * util.h - the two helpers libxm exports from its private util.h and that
 * its test driver links against (libxm src/util.h:32-35).
 */
#ifndef UTIL_H_INCLUDED
#define UTIL_H_INCLUDED

#include "dim.h"

#ifndef __dead
#if defined(__GNUC__)
#define __dead __attribute__((__noreturn__))
#else
#define __dead
#endif
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* "libxm: <func>: <msg>" on stderr, then exit(1) (libxm src/util.h:32). */
#define fatal(x) xm_fatal("%s: %s", __func__, (x))

/* printf-style message prefixed by "libxm: ", then exit(1)
 * (libxm src/util.h:34). */
void xm_fatal(const char *, ...) __dead;

/* Pair every (i, j) with s1[i] == s2[j], in s1 order
 * (libxm src/util.h:35). */
void xm_make_masks(const char *, const char *, xm_dim_t *, xm_dim_t *);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* UTIL_H_INCLUDED */
