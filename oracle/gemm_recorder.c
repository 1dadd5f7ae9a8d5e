/*
 * gemm_recorder.c - TEST INFRASTRUCTURE ONLY.
 *
 * Exports the Fortran BLAS symbols sgemm_/cgemm_/dgemm_/zgemm_ with libxm's
 * calling convention (libxm src/contract.c:32-47) and forwards every call to
 * the BLAS named by $XM_CPU_BLAS or MKL's libmkl_rt.  While recording is on,
 * it appends (type, m, n, k, alpha) of each call to a table that the test
 * suite reads back.  Loaded with RTLD_GLOBAL before the reference library,
 * it sits in front of the reference's ?gemm_ calls, which pins the block
 * pairs (and their merged factors) the reference multiplies.
 */
#define _GNU_SOURCE
#include <complex.h>
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define API __attribute__((visibility("default")))

struct rec { int type; long m, n, k; double re, im; };

static struct rec *recs;
static size_t nrecs, caprecs;
static int recording;
static void *real[4];

static void
load_real(void)
{
	static const char *names[4] = { "sgemm_", "cgemm_", "dgemm_", "zgemm_" };
	const char *path = getenv("XM_CPU_BLAS");
	void *h = NULL;
	int i;

	if (real[0])
		return;
	if (path && *path)
		h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
	if (h == NULL)
		h = dlopen("/opt/conda/lib/libmkl_rt.so", RTLD_NOW | RTLD_LOCAL);
	if (h == NULL) {
		fprintf(stderr, "gemm_recorder: no BLAS to forward to\n");
		exit(1);
	}
	for (i = 0; i < 4; i++)
		if ((real[i] = dlsym(h, names[i])) == NULL) {
			fprintf(stderr, "gemm_recorder: %s missing\n", names[i]);
			exit(1);
		}
}

static void
add(int type, long m, long n, long k, double re, double im)
{
	if (!recording)
		return;
#pragma omp critical(gemm_recorder)
	{
		if (nrecs == caprecs) {
			caprecs = caprecs ? 2 * caprecs : 1024;
			recs = realloc(recs, caprecs * sizeof *recs);
		}
		recs[nrecs].type = type;
		recs[nrecs].m = m;
		recs[nrecs].n = n;
		recs[nrecs].k = k;
		recs[nrecs].re = re;
		recs[nrecs].im = im;
		nrecs++;
	}
}

typedef void (*gemm_t)(char *, char *, long *, long *, long *, void *, void *,
    long *, void *, long *, void *, void *, long *);

#define FORWARD(NAME, IDX, T, RE, IM)					\
API void NAME(char *ta, char *tb, long *m, long *n, long *k, T *al,	\
    T *a, long *lda, T *b, long *ldb, T *bt, T *c, long *ldc)		\
{									\
	load_real();							\
	add(IDX, *m, *n, *k, RE, IM);					\
	((gemm_t)real[IDX])(ta, tb, m, n, k, al, a, lda, b, ldb, bt, c, ldc); \
}

FORWARD(sgemm_, 0, float, (double)*al, 0.0)
FORWARD(cgemm_, 1, float complex, (double)crealf(*al), (double)cimagf(*al))
FORWARD(dgemm_, 2, double, *al, 0.0)
FORWARD(zgemm_, 3, double complex, creal(*al), cimag(*al))

API void gemm_recorder_start(void) { nrecs = 0; recording = 1; }
API void gemm_recorder_stop(void) { recording = 0; }
API size_t gemm_recorder_count(void) { return nrecs; }

/* Copy record i into out[0..5] = type, m, n, k, re(alpha), im(alpha). */
API void
gemm_recorder_get(size_t i, double *out)
{
	out[0] = recs[i].type;
	out[1] = (double)recs[i].m;
	out[2] = (double)recs[i].n;
	out[3] = (double)recs[i].k;
	out[4] = recs[i].re;
	out[5] = recs[i].im;
}
