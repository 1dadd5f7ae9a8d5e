/*
 * xm_oracle.c - CPU oracle for the libxm contraction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the algorithm
 * of libxm's CPU implementation (block spaces, block tensors with zero /
 * canonical / derivative blocks, xm_contract with its pair enumeration and
 * merge, unfold/fold, xm_set, xm_add, xm_copy, xm_mul, xm_div, xm_dot) so
 * that the MI355X library can be checked against it.  Only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark may load it; the
 * product never does.
 *
 * It exports the same C ABI as the product (include/xm.h), so one binding
 * drives both.  Storage is host RAM (data_ptr = host address); the disk
 * pagefile of libxm is not restated.  GEMMs go to an optimized BLAS found at
 * run time ($XM_CPU_BLAS, MKL's libmkl_rt, numpy's OpenBLAS) called exactly
 * as libxm calls ?gemm_ ('T','N', beta = 1), sequential inside each OpenMP
 * thread; without one, a plain triple loop is used and reported as such.
 *
 * Every function cites the libxm file:line whose behaviour it follows.
 */
#define _GNU_SOURCE
#include <complex.h>
#include <dlfcn.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "xm.h"
#include "util.h"

#define ORACLE_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* util (libxm src/util.c:24-53)                                       */
/* ------------------------------------------------------------------ */

/* libxm src/util.c:24-35 */
ORACLE_API void
xm_fatal(const char *fmt, ...)
{
	va_list ap;

	fputs("libxm: ", stderr);
	va_start(ap, fmt);
	vfprintf(stderr, fmt, ap);
	va_end(ap);
	fputc('\n', stderr);
	exit(1);
}

/* libxm src/util.c:37-53: every (i, j) with s1[i] == s2[j], s1 order. */
ORACLE_API void
xm_make_masks(const char *s1, const char *s2, xm_dim_t *m1, xm_dim_t *m2)
{
	size_t i, j, n1 = strlen(s1), n2 = strlen(s2);

	m1->n = m2->n = 0;
	for (i = 0; i < n1; i++)
		for (j = 0; j < n2; j++)
			if (s1[i] == s2[j]) {
				m1->i[m1->n++] = i;
				m2->i[m2->n++] = j;
			}
}

/* ------------------------------------------------------------------ */
/* dim (libxm src/dim.c)                                               */
/* ------------------------------------------------------------------ */

/* libxm src/dim.c:20-36 */
ORACLE_API xm_dim_t
xm_dim_same(size_t ndim, size_t dim)
{
	xm_dim_t r;

	if (ndim > XM_MAX_DIM)
		fatal("too many dimensions");
	for (r.n = 0; r.n < ndim; r.n++)
		r.i[r.n] = dim;
	return r;
}

ORACLE_API xm_dim_t xm_dim_zero(size_t ndim) { return xm_dim_same(ndim, 0); }

static xm_dim_t
dim_make(size_t n, const size_t *v)
{
	xm_dim_t r;
	size_t k;

	r.n = n;
	for (k = 0; k < n; k++)
		r.i[k] = v[k];
	return r;
}

/* libxm src/dim.c:38-166 */
ORACLE_API xm_dim_t xm_dim_1(size_t a)
{ size_t v[] = {a}; return dim_make(1, v); }
ORACLE_API xm_dim_t xm_dim_2(size_t a, size_t b)
{ size_t v[] = {a, b}; return dim_make(2, v); }
ORACLE_API xm_dim_t xm_dim_3(size_t a, size_t b, size_t c)
{ size_t v[] = {a, b, c}; return dim_make(3, v); }
ORACLE_API xm_dim_t xm_dim_4(size_t a, size_t b, size_t c, size_t d)
{ size_t v[] = {a, b, c, d}; return dim_make(4, v); }
ORACLE_API xm_dim_t xm_dim_5(size_t a, size_t b, size_t c, size_t d, size_t e)
{ size_t v[] = {a, b, c, d, e}; return dim_make(5, v); }
ORACLE_API xm_dim_t xm_dim_6(size_t a, size_t b, size_t c, size_t d, size_t e,
    size_t f)
{ size_t v[] = {a, b, c, d, e, f}; return dim_make(6, v); }
ORACLE_API xm_dim_t xm_dim_7(size_t a, size_t b, size_t c, size_t d, size_t e,
    size_t f, size_t g)
{ size_t v[] = {a, b, c, d, e, f, g}; return dim_make(7, v); }
ORACLE_API xm_dim_t xm_dim_8(size_t a, size_t b, size_t c, size_t d, size_t e,
    size_t f, size_t g, size_t h)
{ size_t v[] = {a, b, c, d, e, f, g, h}; return dim_make(8, v); }

/* libxm src/dim.c:168-184 */
ORACLE_API int
xm_dim_eq(const xm_dim_t *a, const xm_dim_t *b)
{
	size_t k;

	if (a->n != b->n)
		return 0;
	for (k = 0; k < a->n; k++)
		if (a->i[k] != b->i[k])
			return 0;
	return 1;
}

ORACLE_API int xm_dim_ne(const xm_dim_t *a, const xm_dim_t *b)
{ return !xm_dim_eq(a, b); }

/* libxm src/dim.c:186-197 */
ORACLE_API int
xm_dim_less(const xm_dim_t *idx, const xm_dim_t *dim)
{
	size_t k;

	for (k = 0; k < idx->n; k++)
		if (idx->i[k] >= dim->i[k])
			return 0;
	return 1;
}

/* libxm src/dim.c:199-207 */
ORACLE_API size_t
xm_dim_dot(const xm_dim_t *dim)
{
	size_t k, p = 1;

	for (k = 0; k < dim->n; k++)
		p *= dim->i[k];
	return p;
}

/* libxm src/dim.c:209-230: first index fastest. */
ORACLE_API size_t
xm_dim_offset(const xm_dim_t *idx, const xm_dim_t *dim)
{
	size_t k, off = 0, stride = 1;

	for (k = 0; k < idx->n; k++) {
		off += idx->i[k] * stride;
		stride *= dim->i[k];
	}
	return off;
}

/* libxm src/dim.c:232-275 */
ORACLE_API xm_dim_t
xm_dim_from_offset(size_t offset, const xm_dim_t *dim)
{
	xm_dim_t r;
	size_t k;

	r.n = dim->n;
	for (k = 0; k < dim->n; k++) {
		if (k + 1 == dim->n) {
			r.i[k] = offset;
		} else {
			r.i[k] = offset % dim->i[k];
			offset /= dim->i[k];
		}
	}
	return r;
}

/* libxm src/dim.c:277-291: past the end, idx becomes dim. */
ORACLE_API void
xm_dim_inc(xm_dim_t *idx, const xm_dim_t *dim)
{
	size_t k, carry = 1;

	for (k = 0; carry && k < idx->n; k++) {
		idx->i[k] += carry;
		carry = idx->i[k] / dim->i[k];
		idx->i[k] %= dim->i[k];
	}
	if (carry)
		*idx = *dim;
}

/* libxm src/dim.c:293-300 */
ORACLE_API void
xm_dim_zero_mask(xm_dim_t *dim, const xm_dim_t *mask)
{
	size_t k;

	for (k = 0; k < mask->n; k++)
		dim->i[mask->i[k]] = 0;
}

/* libxm src/dim.c:302-312 */
ORACLE_API void
xm_dim_set_mask(xm_dim_t *a, const xm_dim_t *ma, const xm_dim_t *b,
    const xm_dim_t *mb)
{
	size_t k;

	for (k = 0; k < ma->n; k++)
		a->i[ma->i[k]] = b->i[mb->i[k]];
}

/* libxm src/dim.c:314-322 */
ORACLE_API size_t
xm_dim_dot_mask(const xm_dim_t *dim, const xm_dim_t *mask)
{
	size_t k, p = 1;

	for (k = 0; k < mask->n; k++)
		p *= dim->i[mask->i[k]];
	return p;
}

/* libxm src/dim.c:324-336: no wrap to dim at the end. */
ORACLE_API void
xm_dim_inc_mask(xm_dim_t *idx, const xm_dim_t *dim, const xm_dim_t *mask)
{
	size_t k, carry = 1;

	for (k = 0; carry && k < mask->n; k++) {
		size_t d = mask->i[k];
		idx->i[d] += carry;
		carry = idx->i[d] / dim->i[d];
		idx->i[d] %= dim->i[d];
	}
}

/* libxm src/dim.c:338-348 */
ORACLE_API xm_dim_t
xm_dim_identity_permutation(size_t ndim)
{
	xm_dim_t r;

	for (r.n = 0; r.n < ndim; r.n++)
		r.i[r.n] = r.n;
	return r;
}

/* libxm src/dim.c:350-369: r.i[p.i[k]] = idx.i[k]. */
ORACLE_API xm_dim_t
xm_dim_permute(const xm_dim_t *idx, const xm_dim_t *p)
{
	xm_dim_t r;
	size_t k;

	r.n = idx->n;
	for (k = 0; k < idx->n; k++)
		r.i[p->i[k]] = idx->i[k];
	return r;
}

/* ------------------------------------------------------------------ */
/* block space (libxm src/blockspace.c)                                */
/* ------------------------------------------------------------------ */

struct xm_block_space {
	xm_dim_t dims, nblocks;
	size_t *splits[XM_MAX_DIM];	/* nblocks+1 ascending points */
};

/* libxm src/blockspace.c:30-51 */
ORACLE_API xm_block_space_t *
xm_block_space_create(xm_dim_t dims)
{
	xm_block_space_t *bs;
	size_t k;

	if ((bs = calloc(1, sizeof *bs)) == NULL)
		return NULL;
	bs->dims = dims;
	bs->nblocks = xm_dim_same(dims.n, 1);
	for (k = 0; k < dims.n; k++) {
		if ((bs->splits[k] = malloc(2 * sizeof(size_t))) == NULL) {
			xm_block_space_free(bs);
			return NULL;
		}
		bs->splits[k][0] = 0;
		bs->splits[k][1] = dims.i[k];
	}
	return bs;
}

/* libxm src/blockspace.c:53-73 */
ORACLE_API xm_block_space_t *
xm_block_space_clone(const xm_block_space_t *src)
{
	xm_block_space_t *bs;
	size_t k, n;

	if ((bs = calloc(1, sizeof *bs)) == NULL)
		return NULL;
	bs->dims = src->dims;
	bs->nblocks = src->nblocks;
	for (k = 0; k < bs->dims.n; k++) {
		n = (bs->nblocks.i[k] + 1) * sizeof(size_t);
		if ((bs->splits[k] = malloc(n)) == NULL) {
			xm_block_space_free(bs);
			return NULL;
		}
		memcpy(bs->splits[k], src->splits[k], n);
	}
	return bs;
}

/* libxm src/blockspace.c:75-97: dimension j of src becomes p[j]. */
ORACLE_API xm_block_space_t *
xm_block_space_permute_clone(const xm_block_space_t *src, xm_dim_t p)
{
	xm_block_space_t *bs;
	size_t j, n;

	if ((bs = calloc(1, sizeof *bs)) == NULL)
		return NULL;
	bs->dims = xm_dim_permute(&src->dims, &p);
	bs->nblocks = xm_dim_permute(&src->nblocks, &p);
	for (j = 0; j < bs->dims.n; j++) {
		n = (src->nblocks.i[j] + 1) * sizeof(size_t);
		if ((bs->splits[p.i[j]] = malloc(n)) == NULL) {
			xm_block_space_free(bs);
			return NULL;
		}
		memcpy(bs->splits[p.i[j]], src->splits[j], n);
	}
	return bs;
}

ORACLE_API size_t xm_block_space_get_ndims(const xm_block_space_t *bs)
{ return bs->dims.n; }
ORACLE_API xm_dim_t xm_block_space_get_abs_dims(const xm_block_space_t *bs)
{ return bs->dims; }
ORACLE_API xm_dim_t xm_block_space_get_nblocks(const xm_block_space_t *bs)
{ return bs->nblocks; }

/* libxm src/blockspace.c:117-139: insert x, duplicates ignored. */
ORACLE_API void
xm_block_space_split(xm_block_space_t *bs, size_t dim, size_t x)
{
	size_t pos, nb = bs->nblocks.i[dim], *s;

	for (pos = 0; pos <= nb; pos++) {
		if (bs->splits[dim][pos] == x)
			return;
		if (bs->splits[dim][pos] > x)
			break;
	}
	if ((s = realloc(bs->splits[dim], (nb + 2) * sizeof *s)) == NULL)
		return;
	memmove(s + pos + 1, s + pos, (nb + 1 - pos) * sizeof *s);
	s[pos] = x;
	bs->splits[dim] = s;
	bs->nblocks.i[dim]++;
}

/* libxm src/blockspace.c:141-163: blocks of about 32, even sizes. */
ORACLE_API void
xm_block_space_autosplit(xm_block_space_t *bs)
{
	const size_t target = 32;
	size_t j, t, pos;

	for (j = 0; j < bs->dims.n; j++) {
		size_t left = bs->dims.i[j];
		size_t nb = (left + target - 1) / target;

		for (t = 0, pos = 0; t + 1 < nb; t++) {
			size_t sz = left / (nb - t);
			if (sz > 1 && (sz & 1))
				sz = sz < target ? sz + 1 : sz - 1;
			left -= sz;
			pos += sz;
			xm_block_space_split(bs, j, pos);
		}
	}
}

ORACLE_API size_t
xm_block_space_get_split(xm_block_space_t *bs, size_t dim, size_t i)
{
	return bs->splits[dim][i];
}

/* libxm src/blockspace.c:174-189 */
ORACLE_API xm_dim_t
xm_block_space_get_block_dims(const xm_block_space_t *bs, xm_dim_t blkidx)
{
	xm_dim_t d;
	size_t k;

	d.n = bs->dims.n;
	for (k = 0; k < d.n; k++)
		d.i[k] = bs->splits[k][blkidx.i[k] + 1] -
		    bs->splits[k][blkidx.i[k]];
	return d;
}

/* libxm src/blockspace.c:191-204 */
ORACLE_API size_t
xm_block_space_get_block_size(const xm_block_space_t *bs, xm_dim_t blkidx)
{
	xm_dim_t d = xm_block_space_get_block_dims(bs, blkidx);
	return xm_dim_dot(&d);
}

/* libxm src/blockspace.c:206-221: product of per-dimension maxima. */
ORACLE_API size_t
xm_block_space_get_largest_block_size(const xm_block_space_t *bs)
{
	size_t k, j, prod = 1;

	for (k = 0; k < bs->dims.n; k++) {
		size_t mx = 0;
		for (j = 0; j < bs->nblocks.i[k]; j++) {
			size_t w = bs->splits[k][j + 1] - bs->splits[k][j];
			if (w > mx)
				mx = w;
		}
		prod *= mx;
	}
	return prod;
}

/* libxm src/blockspace.c:223-240 */
ORACLE_API void
xm_block_space_decompose_index(const xm_block_space_t *bs, xm_dim_t idx,
    xm_dim_t *blkidx, xm_dim_t *elidx)
{
	size_t k, j;

	*blkidx = xm_dim_zero(bs->dims.n);
	*elidx = xm_dim_zero(bs->dims.n);
	for (k = 0; k < bs->dims.n; k++)
		for (j = 0; j < bs->nblocks.i[k]; j++)
			if (bs->splits[k][j + 1] > idx.i[k]) {
				blkidx->i[k] = j;
				elidx->i[k] = idx.i[k] - bs->splits[k][j];
				break;
			}
}

/* libxm src/blockspace.c:255-270 */
ORACLE_API int
xm_block_space_eq1(const xm_block_space_t *a, size_t da,
    const xm_block_space_t *b, size_t db)
{
	size_t j;

	if (a->nblocks.i[da] != b->nblocks.i[db])
		return 0;
	for (j = 0; j <= a->nblocks.i[da]; j++)
		if (a->splits[da][j] != b->splits[db][j])
			return 0;
	return 1;
}

/* libxm src/blockspace.c:242-253 */
ORACLE_API int
xm_block_space_eq(const xm_block_space_t *a, const xm_block_space_t *b)
{
	size_t k;

	if (a->dims.n != b->dims.n)
		return 0;
	for (k = 0; k < a->dims.n; k++)
		if (!xm_block_space_eq1(a, k, b, k))
			return 0;
	return 1;
}

/* libxm src/blockspace.c:272-282 */
ORACLE_API void
xm_block_space_free(xm_block_space_t *bs)
{
	size_t k;

	if (bs == NULL)
		return;
	for (k = 0; k < bs->dims.n; k++)
		free(bs->splits[k]);
	free(bs);
}

/* ------------------------------------------------------------------ */
/* scalar helpers (libxm src/scalar.c)                                 */
/* ------------------------------------------------------------------ */

/* libxm src/scalar.c:20-43 */
ORACLE_API int
xm_scalar_check_type(xm_scalar_type_t t)
{
	return t == XM_SCALAR_FLOAT || t == XM_SCALAR_FLOAT_COMPLEX ||
	    t == XM_SCALAR_DOUBLE || t == XM_SCALAR_DOUBLE_COMPLEX;
}

ORACLE_API size_t
xm_scalar_sizeof(xm_scalar_type_t t)
{
	switch (t) {
	case XM_SCALAR_FLOAT: return sizeof(float);
	case XM_SCALAR_FLOAT_COMPLEX: return sizeof(float complex);
	case XM_SCALAR_DOUBLE: return sizeof(double);
	case XM_SCALAR_DOUBLE_COMPLEX: return sizeof(double complex);
	}
	fatal("unexpected scalar type");
}

/* libxm src/scalar.c:45-95: rounded to the element type. */
ORACLE_API xm_scalar_t
xm_scalar_add(xm_scalar_t a, xm_scalar_t b, xm_scalar_type_t t)
{
	switch (t) {
	case XM_SCALAR_FLOAT: return (float)a + (float)b;
	case XM_SCALAR_FLOAT_COMPLEX:
		return (float complex)a + (float complex)b;
	case XM_SCALAR_DOUBLE: return (double)a + (double)b;
	case XM_SCALAR_DOUBLE_COMPLEX: return a + b;
	}
	fatal("unexpected scalar type");
}

ORACLE_API xm_scalar_t
xm_scalar_mul(xm_scalar_t a, xm_scalar_t b, xm_scalar_type_t t)
{
	switch (t) {
	case XM_SCALAR_FLOAT: return (float)a * (float)b;
	case XM_SCALAR_FLOAT_COMPLEX:
		return (float complex)a * (float complex)b;
	case XM_SCALAR_DOUBLE: return (double)a * (double)b;
	case XM_SCALAR_DOUBLE_COMPLEX: return a * b;
	}
	fatal("unexpected scalar type");
}

/* libxm src/scalar.c:97-121 */
ORACLE_API xm_scalar_t
xm_scalar_get_element(void *x, size_t i, xm_scalar_type_t t)
{
	switch (t) {
	case XM_SCALAR_FLOAT: return ((float *)x)[i];
	case XM_SCALAR_FLOAT_COMPLEX: return ((float complex *)x)[i];
	case XM_SCALAR_DOUBLE: return ((double *)x)[i];
	case XM_SCALAR_DOUBLE_COMPLEX: return ((double complex *)x)[i];
	}
	fatal("unexpected scalar type");
}

/* The element loops below are written once per element type through a
 * macro; each restates one loop of libxm src/scalar.c. */
#define FOR_TYPES(t, ...)						\
	switch (t) {							\
	case XM_SCALAR_FLOAT: { typedef float T; __VA_ARGS__; return; }	\
	case XM_SCALAR_FLOAT_COMPLEX: { typedef float complex T; __VA_ARGS__; return; } \
	case XM_SCALAR_DOUBLE: { typedef double T; __VA_ARGS__; return; }	\
	case XM_SCALAR_DOUBLE_COMPLEX: { typedef double complex T; __VA_ARGS__; return; } \
	}								\
	fatal("unexpected scalar type")

/* libxm src/scalar.c:123-156 */
ORACLE_API void
xm_scalar_set(void *x, xm_scalar_t a, size_t len, xm_scalar_type_t t)
{
	size_t i;
	FOR_TYPES(t, { T *xx = x; T v = (T)a;
	    for (i = 0; i < len; i++) xx[i] = v; });
}

/* libxm src/scalar.c:158-191 */
ORACLE_API void
xm_scalar_scale(void *x, xm_scalar_t a, size_t len, xm_scalar_type_t t)
{
	size_t i;
	FOR_TYPES(t, { T *xx = x; T v = (T)a;
	    for (i = 0; i < len; i++) xx[i] *= v; });
}

/* libxm src/scalar.c:193-239: x = a*x + b*y, no fused multiply-add. */
ORACLE_API void
xm_scalar_axpy(void *x, xm_scalar_t a, const void *y, xm_scalar_t b,
    size_t len, xm_scalar_type_t t)
{
	size_t i;
	FOR_TYPES(t, { T *xx = x; const T *yy = y; T aa = (T)a, bb = (T)b;
	    for (i = 0; i < len; i++) xx[i] = aa * xx[i] + bb * yy[i]; });
}

/* libxm src/scalar.c:241-283: x = (a*x)*y */
ORACLE_API void
xm_scalar_vec_mul(void *x, xm_scalar_t a, const void *y, size_t len,
    xm_scalar_type_t t)
{
	size_t i;
	FOR_TYPES(t, { T *xx = x; const T *yy = y; T aa = (T)a;
	    for (i = 0; i < len; i++) xx[i] = aa * xx[i] * yy[i]; });
}

/* libxm src/scalar.c:285-327: x = x/(a*y) */
ORACLE_API void
xm_scalar_vec_div(void *x, xm_scalar_t a, const void *y, size_t len,
    xm_scalar_type_t t)
{
	size_t i;
	FOR_TYPES(t, { T *xx = x; const T *yy = y; T aa = (T)a;
	    for (i = 0; i < len; i++) xx[i] = xx[i] / (aa * yy[i]); });
}

/* libxm src/scalar.c:329-371: plain sum in the element type. */
ORACLE_API xm_scalar_t
xm_scalar_dot(const void *x, const void *y, size_t len, xm_scalar_type_t t)
{
	size_t i;

	switch (t) {
	case XM_SCALAR_FLOAT: { const float *a = x, *b = y; float s = 0;
		for (i = 0; i < len; i++) { s += a[i] * b[i]; } return s; }
	case XM_SCALAR_FLOAT_COMPLEX: { const float complex *a = x, *b = y;
		float complex s = 0;
		for (i = 0; i < len; i++) { s += a[i] * b[i]; } return s; }
	case XM_SCALAR_DOUBLE: { const double *a = x, *b = y; double s = 0;
		for (i = 0; i < len; i++) { s += a[i] * b[i]; } return s; }
	case XM_SCALAR_DOUBLE_COMPLEX: { const double complex *a = x, *b = y;
		double complex s = 0;
		for (i = 0; i < len; i++) { s += a[i] * b[i]; } return s; }
	}
	fatal("unexpected scalar type");
}

/* libxm src/scalar.c:373-414 */
ORACLE_API void
xm_scalar_convert(void *x, const void *y, size_t len, xm_scalar_type_t xt,
    xm_scalar_type_t yt)
{
	size_t i;

	if (xt == yt) {
		memcpy(x, y, len * xm_scalar_sizeof(xt));
		return;
	}
#define CONV(XT, YT, TX, TY) \
	if (xt == XT && yt == YT) { TX *a = x; const TY *b = y; \
		for (i = 0; i < len; i++) { a[i] = (TX)b[i]; } return; }
	CONV(XM_SCALAR_DOUBLE, XM_SCALAR_FLOAT, double, float)
	CONV(XM_SCALAR_FLOAT, XM_SCALAR_DOUBLE, float, double)
	CONV(XM_SCALAR_DOUBLE_COMPLEX, XM_SCALAR_FLOAT_COMPLEX,
	    double complex, float complex)
	CONV(XM_SCALAR_FLOAT_COMPLEX, XM_SCALAR_DOUBLE_COMPLEX,
	    float complex, double complex)
#undef CONV
	fatal("unsupported scalar conversion");
}

/* ------------------------------------------------------------------ */
/* allocator: host RAM only (libxm src/alloc.c:147-347, path == NULL)  */
/* ------------------------------------------------------------------ */

struct xm_allocator {
	char *path;
};

/* libxm src/alloc.c:147-207 */
ORACLE_API xm_allocator_t *
xm_allocator_create(const char *path)
{
	xm_allocator_t *a;

	if ((a = calloc(1, sizeof *a)) == NULL) {
		perror("calloc");
		return NULL;
	}
	if (path && (a->path = strdup(path)) == NULL) {
		perror("strdup");
		free(a);
		return NULL;
	}
	return a;
}

ORACLE_API const char *
xm_allocator_get_path(xm_allocator_t *a)
{
	return a->path;
}

/* libxm src/alloc.c:215-247 (RAM branch :233-238) */
ORACLE_API uint64_t
xm_allocator_allocate(xm_allocator_t *a, size_t bytes)
{
	void *p;

	(void)a;
	if ((p = malloc(bytes ? bytes : 1)) == NULL) {
		perror("malloc");
		return XM_NULL_PTR;
	}
	return (uint64_t)(uintptr_t)p;
}

/* libxm src/alloc.c:252-275 (RAM branch :261-264) */
ORACLE_API void
xm_allocator_read(xm_allocator_t *a, uint64_t ptr, void *mem, size_t bytes)
{
	(void)a;
	if (ptr == XM_NULL_PTR)
		fatal("data pointer is NULL");
	memcpy(mem, (const void *)(uintptr_t)ptr, bytes);
}

/* libxm src/alloc.c:277-300 (RAM branch :286-289) */
ORACLE_API void
xm_allocator_write(xm_allocator_t *a, uint64_t ptr, const void *mem,
    size_t bytes)
{
	(void)a;
	if (ptr == XM_NULL_PTR)
		fatal("data pointer is NULL");
	memcpy((void *)(uintptr_t)ptr, mem, bytes);
}

/* libxm src/alloc.c:302-326 */
ORACLE_API void
xm_allocator_deallocate(xm_allocator_t *a, uint64_t ptr)
{
	(void)a;
	if (ptr != XM_NULL_PTR)
		free((void *)(uintptr_t)ptr);
}

/* libxm src/alloc.c:328-347 */
ORACLE_API void
xm_allocator_destroy(xm_allocator_t *a)
{
	if (a == NULL)
		return;
	free(a->path);
	free(a);
}

/* ------------------------------------------------------------------ */
/* block tensor (libxm src/tensor.c)                                   */
/* ------------------------------------------------------------------ */

struct block {
	xm_block_type_t type;
	xm_dim_t perm;
	xm_scalar_t scalar;
	uint64_t data;		/* canonical: data_ptr; derivative: offset of
				   the source block in the block grid */
};

struct xm_tensor {
	xm_scalar_type_t type;
	xm_block_space_t *bs;
	xm_allocator_t *alloc;
	struct block *blocks;
};

static struct block *
blk(const xm_tensor_t *t, xm_dim_t idx)
{
	return &t->blocks[xm_dim_offset(&idx, &t->bs->nblocks)];
}

/* libxm src/tensor.c:243-253 */
ORACLE_API void
xm_tensor_set_zero_block(xm_tensor_t *t, xm_dim_t idx)
{
	struct block *b = blk(t, idx);

	b->type = XM_BLOCK_TYPE_ZERO;
	b->perm = xm_dim_identity_permutation(idx.n);
	b->scalar = 0;
	b->data = XM_NULL_PTR;
}

/* libxm src/tensor.c:53-81 */
ORACLE_API xm_tensor_t *
xm_tensor_create(const xm_block_space_t *bs, xm_scalar_type_t type,
    xm_allocator_t *alloc)
{
	xm_tensor_t *t;
	xm_dim_t idx, nb;

	if (!xm_scalar_check_type(type))
		fatal("unexpected scalar type");
	if ((t = calloc(1, sizeof *t)) == NULL)
		fatal("out of memory");
	if ((t->bs = xm_block_space_clone(bs)) == NULL)
		fatal("out of memory");
	t->type = type;
	t->alloc = alloc;
	nb = bs->nblocks;
	if ((t->blocks = calloc(xm_dim_dot(&nb), sizeof *t->blocks)) == NULL)
		fatal("out of memory");
	for (idx = xm_dim_zero(nb.n); xm_dim_ne(&idx, &nb);
	    xm_dim_inc(&idx, &nb))
		xm_tensor_set_zero_block(t, idx);
	return t;
}

/* libxm src/tensor.c:83-98 */
ORACLE_API xm_tensor_t *
xm_tensor_create_canonical(const xm_block_space_t *bs, xm_scalar_type_t type,
    xm_allocator_t *alloc)
{
	xm_tensor_t *t = xm_tensor_create(bs, type, alloc);
	xm_dim_t idx, nb = bs->nblocks;

	for (idx = xm_dim_zero(nb.n); xm_dim_ne(&idx, &nb);
	    xm_dim_inc(&idx, &nb))
		xm_tensor_set_canonical_block(t, idx);
	return t;
}

/* libxm src/tensor.c:100-122: derivative entries are copied as they are
 * (their source offsets stay valid); canonical blocks get new storage. */
ORACLE_API xm_tensor_t *
xm_tensor_create_structure(const xm_tensor_t *src, xm_scalar_type_t type,
    xm_allocator_t *alloc)
{
	xm_tensor_t *t;
	xm_dim_t idx, nb = src->bs->nblocks;

	if (alloc == NULL)
		alloc = src->alloc;
	t = xm_tensor_create(src->bs, type, alloc);
	for (idx = xm_dim_zero(nb.n); xm_dim_ne(&idx, &nb);
	    xm_dim_inc(&idx, &nb)) {
		size_t o = xm_dim_offset(&idx, &nb);
		t->blocks[o] = src->blocks[o];
		if (t->blocks[o].type == XM_BLOCK_TYPE_CANONICAL) {
			t->blocks[o].type = XM_BLOCK_TYPE_ZERO;
			xm_tensor_set_canonical_block(t, idx);
		}
	}
	return t;
}

ORACLE_API const xm_block_space_t *
xm_tensor_get_block_space(const xm_tensor_t *t) { return t->bs; }
ORACLE_API xm_scalar_type_t
xm_tensor_get_scalar_type(const xm_tensor_t *t) { return t->type; }
ORACLE_API xm_allocator_t *
xm_tensor_get_allocator(const xm_tensor_t *t) { return t->alloc; }
ORACLE_API xm_dim_t
xm_tensor_get_abs_dims(const xm_tensor_t *t) { return t->bs->dims; }
ORACLE_API xm_dim_t
xm_tensor_get_nblocks(const xm_tensor_t *t) { return t->bs->nblocks; }

ORACLE_API xm_block_type_t
xm_tensor_get_block_type(const xm_tensor_t *t, xm_dim_t idx)
{
	return blk(t, idx)->type;
}

ORACLE_API xm_dim_t
xm_tensor_get_block_dims(const xm_tensor_t *t, xm_dim_t idx)
{
	return xm_block_space_get_block_dims(t->bs, idx);
}

ORACLE_API size_t
xm_tensor_get_block_size(const xm_tensor_t *t, xm_dim_t idx)
{
	return xm_block_space_get_block_size(t->bs, idx);
}

ORACLE_API size_t
xm_tensor_get_block_bytes(const xm_tensor_t *t, xm_dim_t idx)
{
	return xm_tensor_get_block_size(t, idx) * xm_scalar_sizeof(t->type);
}

ORACLE_API size_t
xm_tensor_get_largest_block_size(const xm_tensor_t *t)
{
	return xm_block_space_get_largest_block_size(t->bs);
}

ORACLE_API size_t
xm_tensor_get_largest_block_bytes(const xm_tensor_t *t)
{
	return xm_tensor_get_largest_block_size(t) * xm_scalar_sizeof(t->type);
}

/* libxm src/tensor.c:218-229 */
ORACLE_API uint64_t
xm_tensor_get_block_data_ptr(const xm_tensor_t *t, xm_dim_t idx)
{
	struct block *b = blk(t, idx);

	if (b->type == XM_BLOCK_TYPE_CANONICAL)
		return b->data;
	if (b->type == XM_BLOCK_TYPE_DERIVATIVE)
		return t->blocks[b->data].data;
	return XM_NULL_PTR;
}

ORACLE_API xm_dim_t
xm_tensor_get_block_permutation(const xm_tensor_t *t, xm_dim_t idx)
{
	return blk(t, idx)->perm;
}

ORACLE_API xm_scalar_t
xm_tensor_get_block_scalar(const xm_tensor_t *t, xm_dim_t idx)
{
	return blk(t, idx)->scalar;
}

/* libxm src/tensor.c:270-285 */
ORACLE_API void
xm_tensor_set_canonical_block_raw(xm_tensor_t *t, xm_dim_t idx, uint64_t ptr)
{
	struct block *b;

	if (ptr == XM_NULL_PTR)
		fatal("unexpected null data pointer");
	if (xm_tensor_get_block_type(t, idx) != XM_BLOCK_TYPE_ZERO)
		fatal("block must be zero");
	b = blk(t, idx);
	b->type = XM_BLOCK_TYPE_CANONICAL;
	b->perm = xm_dim_identity_permutation(idx.n);
	b->scalar = 1;
	b->data = ptr;
}

/* libxm src/tensor.c:255-268 */
ORACLE_API void
xm_tensor_set_canonical_block(xm_tensor_t *t, xm_dim_t idx)
{
	uint64_t ptr;

	if (xm_tensor_get_block_type(t, idx) != XM_BLOCK_TYPE_ZERO)
		fatal("block must be zero");
	ptr = xm_allocator_allocate(t->alloc, xm_tensor_get_block_bytes(t, idx));
	if (ptr == XM_NULL_PTR)
		fatal("unable to allocate block data");
	xm_tensor_set_canonical_block_raw(t, idx, ptr);
}

/* libxm src/tensor.c:287-311 */
ORACLE_API void
xm_tensor_set_derivative_block(xm_tensor_t *t, xm_dim_t idx, xm_dim_t src,
    xm_dim_t perm, xm_scalar_t s)
{
	struct block *b;
	xm_dim_t d1, d2;

	if (xm_tensor_get_block_type(t, idx) != XM_BLOCK_TYPE_ZERO)
		fatal("block must be zero");
	if (xm_tensor_get_block_type(t, src) != XM_BLOCK_TYPE_CANONICAL)
		fatal("derivative blocks must have canonical source blocks");
	d1 = xm_block_space_get_block_dims(t->bs, idx);
	d1 = xm_dim_permute(&d1, &perm);
	d2 = xm_block_space_get_block_dims(t->bs, src);
	if (xm_dim_ne(&d1, &d2))
		fatal("invalid block permutation");
	b = blk(t, idx);
	b->type = XM_BLOCK_TYPE_DERIVATIVE;
	b->perm = perm;
	b->scalar = s;
	b->data = xm_dim_offset(&src, &t->bs->nblocks);
}

/* libxm src/tensor.c:313-336 */
ORACLE_API void
xm_tensor_get_canonical_block_list(const xm_tensor_t *t, xm_dim_t **list,
    size_t *nlist)
{
	xm_dim_t idx, nb = t->bs->nblocks, *l;
	size_t n = 0, cap = xm_dim_dot(&nb);

	if ((l = malloc((cap ? cap : 1) * sizeof *l)) == NULL)
		fatal("out of memory");
	for (idx = xm_dim_zero(nb.n); xm_dim_ne(&idx, &nb);
	    xm_dim_inc(&idx, &nb))
		if (xm_tensor_get_block_type(t, idx) == XM_BLOCK_TYPE_CANONICAL)
			l[n++] = idx;
	*list = l;
	*nlist = n;
}

/* libxm src/tensor.c:338-351 */
ORACLE_API void
xm_tensor_read_block(const xm_tensor_t *t, xm_dim_t idx, void *buf)
{
	if (xm_tensor_get_block_type(t, idx) == XM_BLOCK_TYPE_ZERO)
		fatal("cannot read data from zero-blocks");
	xm_allocator_read(t->alloc, xm_tensor_get_block_data_ptr(t, idx), buf,
	    xm_tensor_get_block_bytes(t, idx));
}

/* libxm src/tensor.c:353-366 */
ORACLE_API void
xm_tensor_write_block(xm_tensor_t *t, xm_dim_t idx, const void *buf)
{
	if (xm_tensor_get_block_type(t, idx) != XM_BLOCK_TYPE_CANONICAL)
		fatal("can only write to canonical blocks");
	xm_allocator_write(t->alloc, xm_tensor_get_block_data_ptr(t, idx), buf,
	    xm_tensor_get_block_bytes(t, idx));
}

/* libxm src/tensor.c:154-178 */
ORACLE_API xm_scalar_t
xm_tensor_get_element(const xm_tensor_t *t, xm_dim_t idx)
{
	struct block *b;
	xm_dim_t bi, ei, bd;
	xm_scalar_t v;
	void *buf;

	xm_block_space_decompose_index(t->bs, idx, &bi, &ei);
	b = blk(t, bi);
	if (b->type == XM_BLOCK_TYPE_ZERO)
		return 0;
	ei = xm_dim_permute(&ei, &b->perm);
	bd = xm_tensor_get_block_dims(t, bi);
	if ((buf = malloc(xm_tensor_get_block_bytes(t, bi))) == NULL)
		fatal("out of memory");
	bd = xm_dim_permute(&bd, &b->perm);
	xm_tensor_read_block(t, bi, buf);
	v = xm_scalar_get_element(buf, xm_dim_offset(&ei, &bd), t->type);
	free(buf);
	return xm_scalar_mul(b->scalar, v, t->type);
}

/*
 * Unfold / fold (libxm src/tensor.c:368-639).  Matrix element (i, j) of the
 * unfolded form lives at to[j*stride + i]; i runs over the block indices in
 * mask_i (first fastest), j over those in mask_j.  The source element is
 * found through the block permutation; the block scalar is NOT applied.
 * Runs along the leading mask_i index are copied with the source stride of
 * that index, exactly as libxm's per-run kernels do.
 */
static void
move_runs(const xm_tensor_t *t, xm_dim_t idx, xm_dim_t mi, xm_dim_t mj,
    const void *from, void *to, size_t stride, int fold)
{
	xm_dim_t dims, pdims, perm, e, pe;
	size_t esz = xm_scalar_sizeof(t->type);
	size_t ni, nj, ii, jj, k, inc = 1, run = 1, lead;

	if (from == NULL || to == NULL || from == to)
		fatal("invalid argument");
	if (mi.n + mj.n != idx.n)
		fatal("invalid mask dimensions");
	dims = xm_tensor_get_block_dims(t, idx);
	ni = xm_dim_dot_mask(&dims, &mi);
	nj = xm_dim_dot_mask(&dims, &mj);
	perm = fold ? xm_dim_identity_permutation(idx.n) :
	    xm_tensor_get_block_permutation(t, idx);
	pdims = xm_dim_permute(&dims, &perm);
	if (mi.n > 0) {
		lead = mi.i[0];
		for (k = 0; k < perm.i[lead]; k++)
			inc *= pdims.i[k];
		run = dims.i[lead];
		for (k = 0; k + 1 < mi.n; k++)
			mi.i[k] = mi.i[k + 1];
		mi.n--;
	}
	e = xm_dim_zero(dims.n);
	for (jj = 0; jj < nj; jj++) {
		xm_dim_zero_mask(&e, &mi);
		for (ii = 0; ii < ni; ii += run) {
			size_t off, mat = jj * stride + ii;
			pe = xm_dim_permute(&e, &perm);
			off = xm_dim_offset(&pe, &pdims);
			for (k = 0; k < run; k++) {
				char *dst;
				const char *src;
				if (fold) {
					dst = (char *)to + (off + k * inc) * esz;
					src = (const char *)from + (mat + k) * esz;
				} else {
					dst = (char *)to + (mat + k) * esz;
					src = (const char *)from + (off + k * inc) * esz;
				}
				memcpy(dst, src, esz);
			}
			xm_dim_inc_mask(&e, &dims, &mi);
		}
		xm_dim_inc_mask(&e, &dims, &mj);
	}
}

/* libxm src/tensor.c:572-639 */
ORACLE_API void
xm_tensor_unfold_block(const xm_tensor_t *t, xm_dim_t idx, xm_dim_t mi,
    xm_dim_t mj, const void *from, void *to, size_t stride)
{
	move_runs(t, idx, mi, mj, from, to, stride, 0);
}

/* libxm src/tensor.c:436-505 */
ORACLE_API void
xm_tensor_fold_block(const xm_tensor_t *t, xm_dim_t idx, xm_dim_t mi,
    xm_dim_t mj, const void *from, void *to, size_t stride)
{
	if (from == NULL || to == NULL || from == to)
		fatal("invalid argument");
	if (mi.n + mj.n != idx.n)
		fatal("invalid mask dimensions");
	if (xm_tensor_get_block_type(t, idx) != XM_BLOCK_TYPE_CANONICAL)
		fatal("can only fold canonical blocks");
	move_runs(t, idx, mi, mj, from, to, stride, 1);
}

/* libxm src/tensor.c:641-658 */
ORACLE_API void
xm_tensor_free_block_data(xm_tensor_t *t)
{
	xm_dim_t idx, nb = t->bs->nblocks;

	for (idx = xm_dim_zero(nb.n); xm_dim_ne(&idx, &nb);
	    xm_dim_inc(&idx, &nb)) {
		if (xm_tensor_get_block_type(t, idx) == XM_BLOCK_TYPE_CANONICAL)
			xm_allocator_deallocate(t->alloc,
			    xm_tensor_get_block_data_ptr(t, idx));
		xm_tensor_set_zero_block(t, idx);
	}
}

/* libxm src/tensor.c:660-669 */
ORACLE_API void
xm_tensor_free(xm_tensor_t *t)
{
	if (t) {
		xm_block_space_free(t->bs);
		free(t->blocks);
		free(t);
	}
}

/* ------------------------------------------------------------------ */
/* BLAS found at run time                                              */
/* ------------------------------------------------------------------ */

typedef void (*gemm_lp64_t)(const char *, const char *, const int *,
    const int *, const int *, const void *, const void *, const int *,
    const void *, const int *, const void *, void *, const int *);
typedef void (*gemm_ilp64_t)(const char *, const char *, const long *,
    const long *, const long *, const void *, const void *, const long *,
    const void *, const long *, const void *, void *, const long *);

static struct {
	int ready, ilp64;
	void *fn[4];		/* indexed by xm_scalar_type_t */
	char name[256];
} blas;

static void
blas_try(const char *path, const char *const *names, int ilp64)
{
	void *h;
	int k;

	if (blas.ready || path == NULL || *path == '\0')
		return;
	if ((h = dlopen(path, RTLD_NOW | RTLD_LOCAL)) == NULL)
		return;
	for (k = 0; k < 4; k++)
		if ((blas.fn[k] = dlsym(h, names[k])) == NULL) {
			dlclose(h);
			return;
		}
	if (!ilp64) {
		/* MKL: sequential inside each OpenMP thread, as the
		 * release build of libxm links it (libxm Makefile:7). */
		int (*setlayer)(int) = (int (*)(int))dlsym(h,
		    "MKL_Set_Threading_Layer");
		if (setlayer)
			setlayer(1 /* MKL_THREADING_SEQUENTIAL */);
	} else {
		void (*setn)(int) = (void (*)(int))dlsym(h,
		    "scipy_openblas_set_num_threads64_");
		if (setn)
			setn(1);
	}
	blas.ilp64 = ilp64;
	blas.ready = 1;
	snprintf(blas.name, sizeof blas.name, "%s", path);
}

static void
blas_init(void)
{
	static const char *const lp64[] = { "sgemm_", "cgemm_", "dgemm_",
	    "zgemm_" };
	static const char *const ilp64[] = { "scipy_sgemm_64_",
	    "scipy_cgemm_64_", "scipy_dgemm_64_", "scipy_zgemm_64_" };
	static const char *const np_openblas[] = {
	    "/usr/local/lib/python3.10/dist-packages/numpy.libs/"
	    "libscipy_openblas64_-56d6093b.so", NULL };
	const char *env = getenv("XM_CPU_BLAS");

	if (blas.ready)
		return;
	if (env && strcmp(env, "naive") == 0) {
		blas.ready = -1;
		snprintf(blas.name, sizeof blas.name, "naive (unoptimized)");
		return;
	}
	blas_try(env, lp64, 0);
	blas_try(env, ilp64, 1);
	blas_try("/opt/conda/lib/libmkl_rt.so", lp64, 0);
	blas_try("libmkl_rt.so", lp64, 0);
	blas_try(np_openblas[0], ilp64, 1);
	if (!blas.ready) {
		blas.ready = -1;
		snprintf(blas.name, sizeof blas.name, "naive (unoptimized)");
	}
}

/* Name of the BLAS the oracle uses (test and benchmark reporting). */
ORACLE_API const char *
xm_oracle_blas_name(void)
{
#ifdef _OPENMP
#pragma omp critical(xm_oracle_blas)
#endif
	blas_init();
	return blas.name;
}

/* C(m x n, ldc) += al * A^T * B, A is k x m (lda), B is k x n (ldb): the
 * only form libxm uses (libxm src/contract.c:207-216). */
#define NAIVE_GEMM(T)							\
	do {								\
		const T *A = a, *B = b; T *C = c, s, al = *(const T *)alp; \
		for (j = 0; j < n; j++)					\
			for (i = 0; i < m; i++) {			\
				s = 0;					\
				for (p = 0; p < k; p++)			\
					s += A[p + i * lda] * B[p + j * ldb]; \
				C[i + j * ldc] += al * s;		\
			}						\
	} while (0)

static void
gemm_tn(xm_scalar_type_t type, size_t m, size_t n, size_t k, xm_scalar_t al,
    const void *a, size_t lda, const void *b, size_t ldb, void *c, size_t ldc)
{
	float fal = (float)al, fone = 1;
	float complex cal = (float complex)al, cone = 1;
	double dal = (double)al, done = 1;
	double complex zal = al, zone = 1;
	const void *alp, *one;
	size_t i, j, p;

	switch (type) {
	case XM_SCALAR_FLOAT: alp = &fal; one = &fone; break;
	case XM_SCALAR_FLOAT_COMPLEX: alp = &cal; one = &cone; break;
	case XM_SCALAR_DOUBLE: alp = &dal; one = &done; break;
	default: alp = &zal; one = &zone; break;
	}
	if (blas.ready == 0)
		xm_oracle_blas_name();
	if (blas.ready > 0 && !blas.ilp64) {
		int im = (int)m, in = (int)n, ik = (int)k, ia = (int)lda,
		    ib = (int)ldb, ic = (int)ldc;
		((gemm_lp64_t)blas.fn[type])("T", "N", &im, &in, &ik, alp, a,
		    &ia, b, &ib, one, c, &ic);
		return;
	}
	if (blas.ready > 0) {
		long lm = (long)m, ln = (long)n, lk = (long)k, la = (long)lda,
		    lb = (long)ldb, lc = (long)ldc;
		((gemm_ilp64_t)blas.fn[type])("T", "N", &lm, &ln, &lk, alp, a,
		    &la, b, &lb, one, c, &lc);
		return;
	}
	switch (type) {
	case XM_SCALAR_FLOAT: NAIVE_GEMM(float); break;
	case XM_SCALAR_FLOAT_COMPLEX: NAIVE_GEMM(float complex); break;
	case XM_SCALAR_DOUBLE: NAIVE_GEMM(double); break;
	default: NAIVE_GEMM(double complex); break;
	}
}

/* ------------------------------------------------------------------ */
/* contraction (libxm src/contract.c)                                  */
/* ------------------------------------------------------------------ */

struct pair {
	xm_dim_t a, b;
	xm_scalar_t alpha;
};

struct masks {
	xm_dim_t cidxa, aidxa, cidxb, aidxb, cidxc, aidxc;
};

/* libxm src/contract.c:86-115: same source data and same permutation at
 * the free positions of a and of b. */
static int
same_contraction(const struct pair *p, size_t i, size_t j,
    const struct masks *mk, const xm_tensor_t *a, const xm_tensor_t *b)
{
	xm_dim_t pia, pja, pib, pjb;
	size_t k;

	if (xm_tensor_get_block_data_ptr(a, p[i].a) !=
	    xm_tensor_get_block_data_ptr(a, p[j].a) ||
	    xm_tensor_get_block_data_ptr(b, p[i].b) !=
	    xm_tensor_get_block_data_ptr(b, p[j].b))
		return 0;
	pia = xm_tensor_get_block_permutation(a, p[i].a);
	pja = xm_tensor_get_block_permutation(a, p[j].a);
	pib = xm_tensor_get_block_permutation(b, p[i].b);
	pjb = xm_tensor_get_block_permutation(b, p[j].b);
	for (k = 0; k < mk->aidxa.n; k++)
		if (pia.i[mk->aidxa.i[k]] != pja.i[mk->aidxa.i[k]])
			return 0;
	for (k = 0; k < mk->aidxb.n; k++)
		if (pib.i[mk->aidxb.i[k]] != pjb.i[mk->aidxb.i[k]])
			return 0;
	return 1;
}

/* libxm src/contract.c:161-192: enumerate the contracted block
 * combinations of one C block, factor s_a*s_b, then merge.  Returns the
 * number of combinations; *npre counts those with both blocks nonzero. */
static size_t
enumerate_pairs(const xm_tensor_t *a, const xm_tensor_t *b, xm_dim_t blkc,
    const struct masks *mk, struct pair *p, size_t *npre)
{
	xm_scalar_type_t type = a->type;
	xm_dim_t nba = a->bs->nblocks, nbb = b->bs->nblocks, ia, ib;
	size_t i, j, n = xm_dim_dot_mask(&nba, &mk->cidxa);

	ia = xm_dim_zero(nba.n);
	ib = xm_dim_zero(nbb.n);
	xm_dim_set_mask(&ia, &mk->aidxa, &blkc, &mk->cidxc);
	xm_dim_set_mask(&ib, &mk->aidxb, &blkc, &mk->aidxc);
	*npre = 0;
	for (i = 0; i < n; i++) {
		p[i].a = ia;
		p[i].b = ib;
		p[i].alpha = 0;
		if (xm_tensor_get_block_type(a, ia) != XM_BLOCK_TYPE_ZERO &&
		    xm_tensor_get_block_type(b, ib) != XM_BLOCK_TYPE_ZERO) {
			p[i].alpha = xm_scalar_mul(
			    xm_tensor_get_block_scalar(a, ia),
			    xm_tensor_get_block_scalar(b, ib), type);
			(*npre)++;
		}
		xm_dim_inc_mask(&ia, &nba, &mk->cidxa);
		xm_dim_inc_mask(&ib, &nbb, &mk->cidxb);
	}
	for (i = 0; i < n; i++) {
		if (p[i].alpha == 0)
			continue;
		for (j = i + 1; j < n; j++) {
			if (p[j].alpha == 0)
				continue;
			if (same_contraction(p, i, j, mk, a, b)) {
				p[i].alpha = xm_scalar_add(p[i].alpha,
				    p[j].alpha, type);
				p[j].alpha = 0;
			}
		}
	}
	return n;
}

/* libxm src/contract.c:117-225 */
static void
compute_block(xm_scalar_t alpha, const xm_tensor_t *a, const xm_tensor_t *b,
    xm_scalar_t beta, xm_tensor_t *c, const struct masks *mk, xm_dim_t blkc,
    struct pair *p, char *buf)
{
	size_t mba = xm_tensor_get_largest_block_bytes(a);
	size_t mbb = xm_tensor_get_largest_block_bytes(b);
	size_t mbc = xm_tensor_get_largest_block_bytes(c);
	char *a1 = buf, *a2 = a1 + mba, *b1 = a2 + mba, *b2 = b1 + mbb;
	char *c1 = b2 + mbb, *c2 = c1 + mbc;
	xm_scalar_type_t type = c->type;
	xm_dim_t dc = xm_tensor_get_block_dims(c, blkc), da;
	size_t m = xm_dim_dot_mask(&dc, &mk->cidxc);
	size_t n = xm_dim_dot_mask(&dc, &mk->aidxc);
	int b_first = mk->aidxc.n > 0 && mk->aidxc.i[0] == 0;
	size_t i, k, npairs, npre;

	xm_tensor_read_block(c, blkc, c2);
	if (b_first)
		xm_tensor_unfold_block(c, blkc, mk->aidxc, mk->cidxc, c2, c1, n);
	else
		xm_tensor_unfold_block(c, blkc, mk->cidxc, mk->aidxc, c2, c1, m);
	if (beta == 0)
		xm_scalar_set(c1, 0, m * n, type);
	else
		xm_scalar_scale(c1, beta, m * n, type);
	if (alpha != 0) {
		npairs = enumerate_pairs(a, b, blkc, mk, p, &npre);
		for (i = 0; i < npairs; i++) {
			xm_scalar_t al;
			if (p[i].alpha == 0)
				continue;
			da = xm_tensor_get_block_dims(a, p[i].a);
			k = xm_dim_dot_mask(&da, &mk->cidxa);
			xm_tensor_read_block(a, p[i].a, a1);
			xm_tensor_unfold_block(a, p[i].a, mk->cidxa, mk->aidxa,
			    a1, a2, k);
			xm_tensor_read_block(b, p[i].b, b1);
			xm_tensor_unfold_block(b, p[i].b, mk->cidxb, mk->aidxb,
			    b1, b2, k);
			al = xm_scalar_mul(alpha, p[i].alpha, type);
			if (b_first)
				gemm_tn(type, n, m, k, al, b2, k, a2, k, c1, n);
			else
				gemm_tn(type, m, n, k, al, a2, k, b2, k, c1, m);
		}
	}
	if (b_first)
		xm_tensor_fold_block(c, blkc, mk->aidxc, mk->cidxc, c1, c2, n);
	else
		xm_tensor_fold_block(c, blkc, mk->cidxc, mk->aidxc, c1, c2, m);
	xm_tensor_write_block(c, blkc, c2);
}

/* libxm src/contract.c:237-280: the checks and masks of xm_contract.
 * `fn` names the API function for the fatal messages. */
static void
contract_masks(const char *fn, const xm_tensor_t *a, const xm_tensor_t *b,
    const xm_tensor_t *c, const char *idxa, const char *idxb,
    const char *idxc, struct masks *mk)
{
	size_t k;

#define CFATAL(msg) xm_fatal("%s: %s", fn, msg)
	if (a->alloc != c->alloc || b->alloc != c->alloc)
		CFATAL("tensors must use same allocator");
	if (a->type != c->type || b->type != c->type)
		CFATAL("tensors must have same scalar type");
	if (strlen(idxa) != a->bs->dims.n || strlen(idxb) != b->bs->dims.n ||
	    strlen(idxc) != c->bs->dims.n)
		CFATAL("bad contraction indices");
	xm_make_masks(idxa, idxb, &mk->cidxa, &mk->cidxb);
	xm_make_masks(idxc, idxa, &mk->cidxc, &mk->aidxa);
	xm_make_masks(idxc, idxb, &mk->aidxc, &mk->aidxb);
	if (mk->aidxa.n + mk->cidxa.n != a->bs->dims.n ||
	    mk->aidxb.n + mk->cidxb.n != b->bs->dims.n ||
	    mk->aidxc.n + mk->cidxc.n != c->bs->dims.n)
		CFATAL("bad contraction indices");
	if (!(mk->aidxc.n > 0 && mk->aidxc.i[0] == 0) &&
	    !(mk->cidxc.n > 0 && mk->cidxc.i[0] == 0))
		CFATAL("bad contraction indices");
	for (k = 0; k < mk->cidxa.n; k++)
		if (!xm_block_space_eq1(a->bs, mk->cidxa.i[k], b->bs,
		    mk->cidxb.i[k]))
			CFATAL("inconsistent a and b tensor block-spaces");
	for (k = 0; k < mk->cidxc.n; k++)
		if (!xm_block_space_eq1(c->bs, mk->cidxc.i[k], a->bs,
		    mk->aidxa.i[k]))
			CFATAL("inconsistent a and c tensor block-spaces");
	for (k = 0; k < mk->aidxc.n; k++)
		if (!xm_block_space_eq1(c->bs, mk->aidxc.i[k], b->bs,
		    mk->aidxb.i[k]))
			CFATAL("inconsistent b and c tensor block-spaces");
#undef CFATAL
}

static double last_flops, last_flops_pre;
static size_t last_npairs, last_npairs_pre, last_nblocks;

/* libxm src/contract.c:227-314 */
ORACLE_API void
xm_contract(xm_scalar_t alpha, const xm_tensor_t *a, const xm_tensor_t *b,
    xm_scalar_t beta, xm_tensor_t *c, const char *idxa, const char *idxb,
    const char *idxc)
{
	struct masks mk;
	xm_dim_t nba, *list;
	size_t nlist, nk, bufbytes;
	long i;

	contract_masks("xm_contract", a, b, c, idxa, idxb, idxc, &mk);
	xm_oracle_blas_name();
	nba = a->bs->nblocks;
	nk = xm_dim_dot_mask(&nba, &mk.cidxa);
	bufbytes = 2 * (xm_tensor_get_largest_block_bytes(a) +
	    xm_tensor_get_largest_block_bytes(b) +
	    xm_tensor_get_largest_block_bytes(c));
	xm_tensor_get_canonical_block_list(c, &list, &nlist);
#ifdef _OPENMP
#pragma omp parallel
#endif
	{
		struct pair *p = malloc((nk ? nk : 1) * sizeof *p);
		char *buf = malloc(bufbytes ? bufbytes : 1);

		if (p == NULL || buf == NULL)
			fatal("out of memory");
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
		for (i = 0; i < (long)nlist; i++)
			compute_block(alpha, a, b, beta, c, &mk, list[i], p, buf);
		free(buf);
		free(p);
	}
	free(list);
}

/* The merged pair list in libxm's order, with the GEMM work it implies
 * (same contract as the product's xm_contract_get_pairs). */
ORACLE_API size_t
xm_contract_get_pairs(const xm_tensor_t *a, const xm_tensor_t *b,
    const xm_tensor_t *c, const char *idxa, const char *idxb,
    const char *idxc, xm_contract_pair_t *out, size_t maxout)
{
	struct masks mk;
	xm_dim_t nba = a->bs->nblocks, *list, dc, da;
	size_t nlist, nk, i, j, cnt = 0, npre, mult;
	struct pair *p;

	contract_masks("xm_contract_get_pairs", a, b, c, idxa, idxb, idxc, &mk);
	nk = xm_dim_dot_mask(&nba, &mk.cidxa);
	if ((p = malloc((nk ? nk : 1) * sizeof *p)) == NULL)
		fatal("out of memory");
	mult = (c->type == XM_SCALAR_FLOAT_COMPLEX ||
	    c->type == XM_SCALAR_DOUBLE_COMPLEX) ? 4 : 1;
	xm_tensor_get_canonical_block_list(c, &list, &nlist);
	last_flops = last_flops_pre = 0;
	last_npairs = last_npairs_pre = 0;
	last_nblocks = nlist;
	for (i = 0; i < nlist; i++) {
		size_t n = enumerate_pairs(a, b, list[i], &mk, p, &npre), m, nn;
		dc = xm_tensor_get_block_dims(c, list[i]);
		m = xm_dim_dot_mask(&dc, &mk.cidxc);
		nn = xm_dim_dot_mask(&dc, &mk.aidxc);
		for (j = 0; j < n; j++) {
			size_t k;
			if (xm_tensor_get_block_type(a, p[j].a) ==
			    XM_BLOCK_TYPE_ZERO ||
			    xm_tensor_get_block_type(b, p[j].b) ==
			    XM_BLOCK_TYPE_ZERO)
				continue;
			da = xm_tensor_get_block_dims(a, p[j].a);
			k = xm_dim_dot_mask(&da, &mk.cidxa);
			last_npairs_pre++;
			last_flops_pre += 2.0 * mult * m * nn * k;
			if (p[j].alpha == 0)
				continue;
			last_npairs++;
			last_flops += 2.0 * mult * m * nn * k;
			if (cnt < maxout) {
				out[cnt].blkidxc = list[i];
				out[cnt].blkidxa = p[j].a;
				out[cnt].blkidxb = p[j].b;
				out[cnt].alpha = p[j].alpha;
				out[cnt].m = m;
				out[cnt].n = nn;
				out[cnt].k = k;
			}
			cnt++;
		}
	}
	free(list);
	free(p);
	return cnt;
}

/* Work counted by the last xm_contract_get_pairs call. */
ORACLE_API void
xm_contract_get_stats(xm_contract_stats_t *s)
{
	memset(s, 0, sizeof *s);
	s->flops = last_flops;
	s->flops_pre = last_flops_pre;
	s->npairs = last_npairs;
	s->npairs_pre = last_npairs_pre;
	s->nblocks_c = last_nblocks;
}

/* ------------------------------------------------------------------ */
/* tensor-wide operations (libxm src/xm.c)                             */
/* ------------------------------------------------------------------ */

/* libxm src/xm.c:28-64 */
ORACLE_API void
xm_set(xm_tensor_t *a, xm_scalar_t x)
{
	xm_dim_t *list;
	size_t nlist;
	long i;
	void *buf;

	if ((buf = malloc(xm_tensor_get_largest_block_bytes(a) + 1)) == NULL)
		fatal("out of memory");
	xm_scalar_set(buf, x, xm_tensor_get_largest_block_size(a), a->type);
	xm_tensor_get_canonical_block_list(a, &list, &nlist);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic)
#endif
	for (i = 0; i < (long)nlist; i++)
		xm_tensor_write_block(a, list[i], buf);
	free(buf);
	free(list);
}

/* The checks shared by copy/add/mul/div/dot (libxm src/xm.c:169-189):
 * the index strings must pair every dimension of a with one of b. */
static void
pair_masks(const char *fn, const xm_tensor_t *a, const xm_tensor_t *b,
    const char *idxa, const char *idxb, int same_type, xm_dim_t *ca,
    xm_dim_t *cb)
{
	size_t k;

	if (a->alloc != b->alloc)
		xm_fatal("%s: %s", fn, "tensors must use same allocator");
	if (same_type && a->type != b->type)
		xm_fatal("%s: %s", fn, "tensors must have same scalar type");
	if (strlen(idxa) != a->bs->dims.n)
		xm_fatal("%s: %s", fn, "idxa does not match tensor dimensions");
	if (strlen(idxb) != b->bs->dims.n)
		xm_fatal("%s: %s", fn, "idxb does not match tensor dimensions");
	xm_make_masks(idxa, idxb, ca, cb);
	if (ca->n != a->bs->dims.n || cb->n != b->bs->dims.n)
		xm_fatal("%s: %s", fn, "index spaces do not match");
	for (k = 0; k < ca->n; k++)
		if (!xm_block_space_eq1(a->bs, ca->i[k], b->bs, cb->i[k]))
			xm_fatal("%s: %s", fn, "inconsistent block-spaces");
}

/* libxm src/xm.c:66-157 */
ORACLE_API void
xm_copy(xm_tensor_t *a, xm_scalar_t s, const xm_tensor_t *b,
    const char *idxa, const char *idxb)
{
	xm_dim_t ca, cb, zero = xm_dim_zero(0), *list;
	size_t nlist, mba, mbb;
	long i;

	pair_masks("xm_copy", a, b, idxa, idxb, 0, &ca, &cb);
	mba = xm_tensor_get_largest_block_bytes(a);
	mbb = xm_tensor_get_largest_block_bytes(b);
	xm_tensor_get_canonical_block_list(a, &list, &nlist);
#ifdef _OPENMP
#pragma omp parallel
#endif
	{
		char *a1 = malloc(mba + 1), *a2 = malloc(mba + 1);
		char *b1 = malloc(mbb + 1), *b2 = malloc(mbb + 1);
		xm_dim_t ib = xm_dim_zero(cb.n);

		if (!a1 || !a2 || !b1 || !b2)
			fatal("out of memory");
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
		for (i = 0; i < (long)nlist; i++) {
			xm_dim_t ia = list[i];
			size_t n;
			xm_dim_set_mask(&ib, &cb, &ia, &ca);
			n = xm_tensor_get_block_size(b, ib);
			if (s == 0 || xm_tensor_get_block_type(b, ib) ==
			    XM_BLOCK_TYPE_ZERO) {
				memset(a2, 0, mba);
			} else {
				xm_scalar_t f = xm_scalar_mul(s,
				    xm_tensor_get_block_scalar(b, ib), b->type);
				xm_tensor_read_block(b, ib, b2);
				xm_tensor_unfold_block(b, ib, cb, zero, b2, b1, n);
				xm_scalar_scale(b1, f, n, b->type);
				xm_scalar_convert(a1, b1, n, a->type, b->type);
				xm_tensor_fold_block(a, ia, ca, zero, a1, a2, n);
			}
			xm_tensor_write_block(a, ia, a2);
		}
		free(a1); free(a2); free(b1); free(b2);
	}
	free(list);
}

/* libxm src/xm.c:159-249 */
ORACLE_API void
xm_add(xm_scalar_t alpha, xm_tensor_t *a, xm_scalar_t beta,
    const xm_tensor_t *b, const char *idxa, const char *idxb)
{
	xm_dim_t ca, cb, zero = xm_dim_zero(0), *list;
	size_t nlist, mb;
	long i;

	pair_masks("xm_add", a, b, idxa, idxb, 1, &ca, &cb);
	mb = xm_tensor_get_largest_block_bytes(a);
	xm_tensor_get_canonical_block_list(a, &list, &nlist);
#ifdef _OPENMP
#pragma omp parallel
#endif
	{
		char *t1 = malloc(mb + 1), *t2 = malloc(mb + 1);
		xm_dim_t ib = xm_dim_zero(cb.n);

		if (!t1 || !t2)
			fatal("out of memory");
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
		for (i = 0; i < (long)nlist; i++) {
			xm_dim_t ia = list[i];
			size_t n;
			xm_dim_set_mask(&ib, &cb, &ia, &ca);
			n = xm_tensor_get_block_size(b, ib);
			if (beta == 0 || xm_tensor_get_block_type(b, ib) ==
			    XM_BLOCK_TYPE_ZERO) {
				memset(t2, 0, mb);
			} else {
				xm_scalar_t f = xm_scalar_mul(beta,
				    xm_tensor_get_block_scalar(b, ib), a->type);
				xm_tensor_read_block(b, ib, t2);
				xm_tensor_unfold_block(b, ib, cb, zero, t2, t1, n);
				xm_scalar_scale(t1, f, n, a->type);
				xm_tensor_fold_block(a, ia, ca, zero, t1, t2, n);
			}
			if (alpha == 0) {
				xm_tensor_write_block(a, ia, t2);
			} else {
				xm_tensor_read_block(a, ia, t1);
				xm_scalar_axpy(t1, alpha, t2, 1, n, a->type);
				xm_tensor_write_block(a, ia, t1);
			}
		}
		free(t1); free(t2);
	}
	free(list);
}

/* libxm src/xm.c:251-332 (mul) and 334-413 (div). */
static void
mul_div(const char *fn, xm_tensor_t *a, const xm_tensor_t *b,
    const char *idxa, const char *idxb, int div)
{
	xm_dim_t ca, cb, zero = xm_dim_zero(0), *list;
	size_t nlist, mb;
	long i;

	pair_masks(fn, a, b, idxa, idxb, 1, &ca, &cb);
	mb = xm_tensor_get_largest_block_bytes(a);
	xm_tensor_get_canonical_block_list(a, &list, &nlist);
#ifdef _OPENMP
#pragma omp parallel
#endif
	{
		char *t1 = malloc(mb + 1), *t2 = malloc(mb + 1);
		xm_dim_t ib = xm_dim_zero(cb.n);

		if (!t1 || !t2)
			fatal("out of memory");
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
		for (i = 0; i < (long)nlist; i++) {
			xm_dim_t ia = list[i];
			size_t n;
			xm_dim_set_mask(&ib, &cb, &ia, &ca);
			n = xm_tensor_get_block_size(b, ib);
			if (xm_tensor_get_block_type(b, ib) ==
			    XM_BLOCK_TYPE_ZERO) {
				if (div)
					xm_fatal("%s: %s", fn,
					    "division by zero");
				memset(t1, 0, mb);
			} else {
				xm_scalar_t s = xm_tensor_get_block_scalar(b, ib);
				xm_tensor_read_block(b, ib, t1);
				xm_tensor_unfold_block(b, ib, cb, zero, t1, t2, n);
				xm_tensor_read_block(a, ia, t1);
				if (div)
					xm_scalar_vec_div(t1, s, t2, n, a->type);
				else
					xm_scalar_vec_mul(t1, s, t2, n, a->type);
			}
			xm_tensor_write_block(a, ia, t1);
		}
		free(t1); free(t2);
	}
	free(list);
}

ORACLE_API void
xm_mul(xm_tensor_t *a, const xm_tensor_t *b, const char *idxa,
    const char *idxb)
{
	mul_div("xm_mul", a, b, idxa, idxb, 0);
}

ORACLE_API void
xm_div(xm_tensor_t *a, const xm_tensor_t *b, const char *idxa,
    const char *idxb)
{
	mul_div("xm_div", a, b, idxa, idxb, 1);
}

/* libxm src/xm.c:415-508: per-block dot, scaled by s_a*s_b, summed in the
 * element type.  Blocks are visited in grid order on one thread so the
 * sum is reproducible. */
ORACLE_API xm_scalar_t
xm_dot(const xm_tensor_t *a, const xm_tensor_t *b, const char *idxa,
    const char *idxb)
{
	xm_dim_t ca, cb, zero = xm_dim_zero(0), nb = a->bs->nblocks, ia, ib;
	size_t mb, i, total;
	xm_scalar_t dot = 0;
	char *t1, *t2, *t3;

	pair_masks("xm_dot", a, b, idxa, idxb, 1, &ca, &cb);
	mb = xm_tensor_get_largest_block_bytes(a);
	t1 = malloc(mb + 1);
	t2 = malloc(mb + 1);
	t3 = malloc(mb + 1);
	if (!t1 || !t2 || !t3)
		fatal("out of memory");
	total = xm_dim_dot(&nb);
	ib = xm_dim_zero(cb.n);
	for (i = 0; i < total; i++) {
		xm_scalar_t sa, sb;
		size_t n;
		ia = xm_dim_from_offset(i, &nb);
		xm_dim_set_mask(&ib, &cb, &ia, &ca);
		if (xm_tensor_get_block_type(a, ia) == XM_BLOCK_TYPE_ZERO ||
		    xm_tensor_get_block_type(b, ib) == XM_BLOCK_TYPE_ZERO)
			continue;
		n = xm_tensor_get_block_size(b, ib);
		xm_tensor_read_block(b, ib, t1);
		xm_tensor_unfold_block(b, ib, cb, zero, t1, t2, n);
		xm_tensor_read_block(a, ia, t1);
		xm_tensor_unfold_block(a, ia, ca, zero, t1, t3, n);
		sa = xm_tensor_get_block_scalar(a, ia);
		sb = xm_tensor_get_block_scalar(b, ib);
		sa = xm_scalar_mul(sa, sb, a->type);
		sb = xm_scalar_dot(t2, t3, n, a->type);
		sa = xm_scalar_mul(sa, sb, a->type);
		dot = xm_scalar_add(dot, sa, a->type);
	}
	free(t1); free(t2); free(t3);
	return dot;
}

/* libxm src/xm.c:510-517 */
ORACLE_API void
xm_print_banner(void)
{
	printf("libxm CPU oracle (restatement of libxm 2.0 for testing)\n");
}
