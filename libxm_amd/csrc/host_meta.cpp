/*
 * host_meta.cpp - host-side metadata of the libxm API: error convention,
 * index masks, multi-index algebra, block spaces, scalar helpers and the
 * block map of tensors.  Block DATA lives in the arena (arena.cpp); this
 * file only touches it through the allocator's read/write helpers, a
 * hipMemcpy or a pread.
 */
#include "internal.h"

#include <atomic>
#include <map>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>

/* ------------------------------------------------------------------ */
/* util                                                                */
/* ------------------------------------------------------------------ */

/* libxm src/util.c:24-35: message on stderr, exit status 1. */
XM_API void
xm_fatal(const char *fmt, ...)
{
	va_list ap;

	fputs("libxm: ", stderr);
	va_start(ap, fmt);
	vfprintf(stderr, fmt, ap);
	va_end(ap);
	fputc('\n', stderr);
	fflush(stderr);
	exit(1);
}

/* libxm src/util.c:37-53: all (i, j) with s1[i] == s2[j], in s1 order. */
XM_API void
xm_make_masks(const char *s1, const char *s2, xm_dim_t *m1, xm_dim_t *m2)
{
	size_t n1 = strlen(s1), n2 = strlen(s2);

	m1->n = m2->n = 0;
	for (size_t i = 0; i < n1; i++)
		for (size_t j = 0; j < n2; j++)
			if (s1[i] == s2[j] && m1->n < XM_MAX_DIM) {
				m1->i[m1->n++] = i;
				m2->i[m2->n++] = j;
			}
}

/* ------------------------------------------------------------------ */
/* dim (libxm src/dim.c)                                               */
/* ------------------------------------------------------------------ */

XM_API xm_dim_t
xm_dim_same(size_t ndim, size_t v)
{
	xm_dim_t d;

	if (ndim > XM_MAX_DIM)
		XM_FATAL("xm_dim_same", "too many dimensions");
	d.n = ndim;
	for (size_t k = 0; k < XM_MAX_DIM; k++)
		d.i[k] = k < ndim ? v : 0;
	return d;
}

XM_API xm_dim_t xm_dim_zero(size_t ndim) { return xm_dim_same(ndim, 0); }

static xm_dim_t
mkdim(std::initializer_list<size_t> v)
{
	xm_dim_t d = xm_dim_same(0, 0);
	for (size_t x : v)
		d.i[d.n++] = x;
	return d;
}

XM_API xm_dim_t xm_dim_1(size_t a) { return mkdim({a}); }
XM_API xm_dim_t xm_dim_2(size_t a, size_t b) { return mkdim({a, b}); }
XM_API xm_dim_t xm_dim_3(size_t a, size_t b, size_t c)
{ return mkdim({a, b, c}); }
XM_API xm_dim_t xm_dim_4(size_t a, size_t b, size_t c, size_t d)
{ return mkdim({a, b, c, d}); }
XM_API xm_dim_t xm_dim_5(size_t a, size_t b, size_t c, size_t d, size_t e)
{ return mkdim({a, b, c, d, e}); }
XM_API xm_dim_t xm_dim_6(size_t a, size_t b, size_t c, size_t d, size_t e,
    size_t f)
{ return mkdim({a, b, c, d, e, f}); }
XM_API xm_dim_t xm_dim_7(size_t a, size_t b, size_t c, size_t d, size_t e,
    size_t f, size_t g)
{ return mkdim({a, b, c, d, e, f, g}); }
XM_API xm_dim_t xm_dim_8(size_t a, size_t b, size_t c, size_t d, size_t e,
    size_t f, size_t g, size_t h)
{ return mkdim({a, b, c, d, e, f, g, h}); }

XM_API int
xm_dim_eq(const xm_dim_t *a, const xm_dim_t *b)
{
	if (a->n != b->n)
		return 0;
	for (size_t k = 0; k < a->n; k++)
		if (a->i[k] != b->i[k])
			return 0;
	return 1;
}

XM_API int xm_dim_ne(const xm_dim_t *a, const xm_dim_t *b)
{ return !xm_dim_eq(a, b); }

XM_API int
xm_dim_less(const xm_dim_t *idx, const xm_dim_t *dim)
{
	for (size_t k = 0; k < idx->n; k++)
		if (idx->i[k] >= dim->i[k])
			return 0;
	return 1;
}

XM_API size_t
xm_dim_dot(const xm_dim_t *d)
{
	size_t p = 1;
	for (size_t k = 0; k < d->n; k++)
		p *= d->i[k];
	return p;
}

/* First index fastest (libxm src/dim.c:209-230). */
XM_API size_t
xm_dim_offset(const xm_dim_t *idx, const xm_dim_t *dim)
{
	size_t off = 0, s = 1;
	for (size_t k = 0; k < idx->n; k++) {
		off += idx->i[k] * s;
		s *= dim->i[k];
	}
	return off;
}

/* libxm src/dim.c:232-275 */
XM_API xm_dim_t
xm_dim_from_offset(size_t off, const xm_dim_t *dim)
{
	xm_dim_t r = xm_dim_same(dim->n, 0);
	for (size_t k = 0; k < dim->n; k++) {
		if (k + 1 == dim->n) {
			r.i[k] = off;
		} else {
			r.i[k] = off % dim->i[k];
			off /= dim->i[k];
		}
	}
	return r;
}

/* libxm src/dim.c:277-291 */
XM_API void
xm_dim_inc(xm_dim_t *idx, const xm_dim_t *dim)
{
	size_t carry = 1;
	for (size_t k = 0; carry && k < idx->n; k++) {
		idx->i[k] += carry;
		carry = idx->i[k] / dim->i[k];
		idx->i[k] %= dim->i[k];
	}
	if (carry)
		*idx = *dim;
}

XM_API void
xm_dim_zero_mask(xm_dim_t *d, const xm_dim_t *mask)
{
	for (size_t k = 0; k < mask->n; k++)
		d->i[mask->i[k]] = 0;
}

XM_API void
xm_dim_set_mask(xm_dim_t *a, const xm_dim_t *ma, const xm_dim_t *b,
    const xm_dim_t *mb)
{
	for (size_t k = 0; k < ma->n; k++)
		a->i[ma->i[k]] = b->i[mb->i[k]];
}

XM_API size_t
xm_dim_dot_mask(const xm_dim_t *d, const xm_dim_t *mask)
{
	size_t p = 1;
	for (size_t k = 0; k < mask->n; k++)
		p *= d->i[mask->i[k]];
	return p;
}

/* libxm src/dim.c:324-336 (no wrap to dim at the end). */
XM_API void
xm_dim_inc_mask(xm_dim_t *idx, const xm_dim_t *dim, const xm_dim_t *mask)
{
	size_t carry = 1;
	for (size_t k = 0; carry && k < mask->n; k++) {
		size_t d = mask->i[k];
		idx->i[d] += carry;
		carry = idx->i[d] / dim->i[d];
		idx->i[d] %= dim->i[d];
	}
}

XM_API xm_dim_t
xm_dim_identity_permutation(size_t ndim)
{
	xm_dim_t r = xm_dim_same(ndim, 0);
	for (size_t k = 0; k < ndim; k++)
		r.i[k] = k;
	return r;
}

/* r.i[p.i[k]] = idx.i[k] (libxm src/dim.c:350-369). */
XM_API xm_dim_t
xm_dim_permute(const xm_dim_t *idx, const xm_dim_t *p)
{
	xm_dim_t r = xm_dim_same(idx->n, 0);
	for (size_t k = 0; k < idx->n; k++)
		r.i[p->i[k]] = idx->i[k];
	return r;
}

/* ------------------------------------------------------------------ */
/* block space (libxm src/blockspace.c)                                */
/* ------------------------------------------------------------------ */

XM_API xm_block_space_t *
xm_block_space_create(xm_dim_t dims)
{
	if (dims.n == 0 || dims.n > XM_MAX_DIM)
		XM_FATAL("xm_block_space_create", "bad number of dimensions");
	xm_block_space_t *bs = new (std::nothrow) xm_block_space_t;
	if (bs == nullptr)
		return nullptr;
	bs->dims = dims;
	bs->nblocks = xm_dim_same(dims.n, 1);
	for (size_t k = 0; k < dims.n; k++)
		bs->splits[k] = {0, dims.i[k]};
	return bs;
}

XM_API xm_block_space_t *
xm_block_space_clone(const xm_block_space_t *src)
{
	return new (std::nothrow) xm_block_space_t(*src);
}

/* Dimension j of src becomes dimension p[j] (libxm src/blockspace.c:75-97). */
XM_API xm_block_space_t *
xm_block_space_permute_clone(const xm_block_space_t *src, xm_dim_t p)
{
	xm_block_space_t *bs = new (std::nothrow) xm_block_space_t;
	if (bs == nullptr)
		return nullptr;
	bs->dims = xm_dim_permute(&src->dims, &p);
	bs->nblocks = xm_dim_permute(&src->nblocks, &p);
	for (size_t j = 0; j < src->dims.n; j++)
		bs->splits[p.i[j]] = src->splits[j];
	return bs;
}

XM_API size_t xm_block_space_get_ndims(const xm_block_space_t *bs)
{ return bs->dims.n; }
XM_API xm_dim_t xm_block_space_get_abs_dims(const xm_block_space_t *bs)
{ return bs->dims; }
XM_API xm_dim_t xm_block_space_get_nblocks(const xm_block_space_t *bs)
{ return bs->nblocks; }

/* Insert a split point; existing points are ignored
 * (libxm src/blockspace.c:117-139). */
XM_API void
xm_block_space_split(xm_block_space_t *bs, size_t dim, size_t x)
{
	if (dim >= bs->dims.n || x > bs->dims.i[dim])
		XM_FATAL("xm_block_space_split", "split point out of range");
	std::vector<size_t> &s = bs->splits[dim];
	auto it = std::lower_bound(s.begin(), s.end(), x);
	if (it != s.end() && *it == x)
		return;
	s.insert(it, x);
	bs->nblocks.i[dim]++;
}

/* Blocks of about 32 elements, odd sizes nudged to even
 * (libxm src/blockspace.c:141-163). */
XM_API void
xm_block_space_autosplit(xm_block_space_t *bs)
{
	const size_t target = 32;

	for (size_t j = 0; j < bs->dims.n; j++) {
		size_t left = bs->dims.i[j];
		size_t nb = (left + target - 1) / target;
		size_t pos = 0;

		for (size_t t = 0; t + 1 < nb; t++) {
			size_t sz = left / (nb - t);
			if (sz > 1 && (sz & 1))
				sz = sz < target ? sz + 1 : sz - 1;
			left -= sz;
			pos += sz;
			xm_block_space_split(bs, j, pos);
		}
	}
}

XM_API size_t
xm_block_space_get_split(xm_block_space_t *bs, size_t dim, size_t i)
{
	if (dim >= bs->dims.n || i > bs->nblocks.i[dim])
		XM_FATAL("xm_block_space_get_split", "index out of range");
	return bs->splits[dim][i];
}

XM_API xm_dim_t
xm_block_space_get_block_dims(const xm_block_space_t *bs, xm_dim_t blk)
{
	xm_dim_t d = xm_dim_same(bs->dims.n, 0);
	for (size_t k = 0; k < d.n; k++)
		d.i[k] = bs->splits[k][blk.i[k] + 1] - bs->splits[k][blk.i[k]];
	return d;
}

XM_API size_t
xm_block_space_get_block_size(const xm_block_space_t *bs, xm_dim_t blk)
{
	xm_dim_t d = xm_block_space_get_block_dims(bs, blk);
	return xm_dim_dot(&d);
}

/* Product of the per-dimension largest block extents
 * (libxm src/blockspace.c:206-221). */
XM_API size_t
xm_block_space_get_largest_block_size(const xm_block_space_t *bs)
{
	size_t prod = 1;
	for (size_t k = 0; k < bs->dims.n; k++) {
		size_t mx = 0;
		for (size_t j = 0; j < bs->nblocks.i[k]; j++)
			mx = std::max(mx, bs->splits[k][j + 1] - bs->splits[k][j]);
		prod *= mx;
	}
	return prod;
}

XM_API void
xm_block_space_decompose_index(const xm_block_space_t *bs, xm_dim_t idx,
    xm_dim_t *blk, xm_dim_t *el)
{
	*blk = xm_dim_zero(bs->dims.n);
	*el = xm_dim_zero(bs->dims.n);
	for (size_t k = 0; k < bs->dims.n; k++) {
		const std::vector<size_t> &s = bs->splits[k];
		for (size_t j = 0; j < bs->nblocks.i[k]; j++)
			if (s[j + 1] > idx.i[k]) {
				blk->i[k] = j;
				el->i[k] = idx.i[k] - s[j];
				break;
			}
	}
}

XM_API int
xm_block_space_eq1(const xm_block_space_t *a, size_t da,
    const xm_block_space_t *b, size_t db)
{
	return a->nblocks.i[da] == b->nblocks.i[db] &&
	    a->splits[da] == b->splits[db];
}

XM_API int
xm_block_space_eq(const xm_block_space_t *a, const xm_block_space_t *b)
{
	if (a->dims.n != b->dims.n)
		return 0;
	for (size_t k = 0; k < a->dims.n; k++)
		if (!xm_block_space_eq1(a, k, b, k))
			return 0;
	return 1;
}

XM_API void
xm_block_space_free(xm_block_space_t *bs)
{
	delete bs;
}

/* ------------------------------------------------------------------ */
/* scalars (libxm src/scalar.c); host vectors                          */
/* ------------------------------------------------------------------ */

XM_API int
xm_scalar_check_type(xm_scalar_type_t t)
{
	return t == XM_SCALAR_FLOAT || t == XM_SCALAR_FLOAT_COMPLEX ||
	    t == XM_SCALAR_DOUBLE || t == XM_SCALAR_DOUBLE_COMPLEX;
}

XM_API size_t
xm_scalar_sizeof(xm_scalar_type_t t)
{
	switch (t) {
	case XM_SCALAR_FLOAT: return 4;
	case XM_SCALAR_FLOAT_COMPLEX: return 8;
	case XM_SCALAR_DOUBLE: return 8;
	case XM_SCALAR_DOUBLE_COMPLEX: return 16;
	}
	XM_FATAL("xm_scalar_sizeof", "unexpected scalar type");
}

size_t elem_size(xm_scalar_type_t t) { return xm_scalar_sizeof(t); }

bool
is_complex(xm_scalar_type_t t)
{
	return t == XM_SCALAR_FLOAT_COMPLEX || t == XM_SCALAR_DOUBLE_COMPLEX;
}

/* Complex products as (ac - bd) + i(ad + bc) with every operation rounded
 * separately: what C99 complex multiplication compiles to for finite
 * operands on x86-64 without FMA. */
static inline std::complex<float>
cmulf(std::complex<float> x, std::complex<float> y)
{
	volatile float ac = x.real() * y.real(), bd = x.imag() * y.imag();
	volatile float ad = x.real() * y.imag(), bc = x.imag() * y.real();
	return {ac - bd, ad + bc};
}

static inline zdouble
cmul(zdouble x, zdouble y)
{
	volatile double ac = x.real() * y.real(), bd = x.imag() * y.imag();
	volatile double ad = x.real() * y.imag(), bc = x.imag() * y.real();
	return {ac - bd, ad + bc};
}

xm_scalar_t
scalar_mul(xm_scalar_t a, xm_scalar_t b, xm_scalar_type_t t)
{
	switch (t) {
	case XM_SCALAR_FLOAT:
		return (double)((float)a.real() * (float)b.real());
	case XM_SCALAR_FLOAT_COMPLEX: {
		std::complex<float> r = cmulf(std::complex<float>(a),
		    std::complex<float>(b));
		return zdouble(r.real(), r.imag());
	}
	case XM_SCALAR_DOUBLE:
		return a.real() * b.real();
	case XM_SCALAR_DOUBLE_COMPLEX:
		return cmul(a, b);
	}
	XM_FATAL("xm_scalar_mul", "unexpected scalar type");
}

xm_scalar_t
scalar_add(xm_scalar_t a, xm_scalar_t b, xm_scalar_type_t t)
{
	switch (t) {
	case XM_SCALAR_FLOAT:
		return (double)((float)a.real() + (float)b.real());
	case XM_SCALAR_FLOAT_COMPLEX:
		return zdouble((float)a.real() + (float)b.real(),
		    (float)a.imag() + (float)b.imag());
	case XM_SCALAR_DOUBLE:
		return a.real() + b.real();
	case XM_SCALAR_DOUBLE_COMPLEX:
		return a + b;
	}
	XM_FATAL("xm_scalar_add", "unexpected scalar type");
}

XM_API xm_scalar_t
xm_scalar_add(xm_scalar_t a, xm_scalar_t b, xm_scalar_type_t t)
{
	return scalar_add(a, b, t);
}

XM_API xm_scalar_t
xm_scalar_mul(xm_scalar_t a, xm_scalar_t b, xm_scalar_type_t t)
{
	return scalar_mul(a, b, t);
}

template <typename T> static inline T to_elem(xm_scalar_t x);
template <> inline float to_elem<float>(xm_scalar_t x) { return (float)x.real(); }
template <> inline double to_elem<double>(xm_scalar_t x) { return x.real(); }
template <> inline std::complex<float> to_elem<std::complex<float>>(xm_scalar_t x)
{ return std::complex<float>((float)x.real(), (float)x.imag()); }
template <> inline zdouble to_elem<zdouble>(xm_scalar_t x) { return x; }

static inline float emul(float a, float b) { return a * b; }
static inline double emul(double a, double b) { return a * b; }
static inline std::complex<float> emul(std::complex<float> a,
    std::complex<float> b) { return cmulf(a, b); }
static inline zdouble emul(zdouble a, zdouble b) { return cmul(a, b); }

#define ON_TYPE(t, fn, ...)						\
	switch (t) {							\
	case XM_SCALAR_FLOAT: fn<float>(__VA_ARGS__); return;		\
	case XM_SCALAR_FLOAT_COMPLEX: fn<std::complex<float>>(__VA_ARGS__); return; \
	case XM_SCALAR_DOUBLE: fn<double>(__VA_ARGS__); return;		\
	case XM_SCALAR_DOUBLE_COMPLEX: fn<zdouble>(__VA_ARGS__); return; \
	}

XM_API xm_scalar_t
xm_scalar_get_element(void *x, size_t i, xm_scalar_type_t t)
{
	switch (t) {
	case XM_SCALAR_FLOAT: return (double)((float *)x)[i];
	case XM_SCALAR_FLOAT_COMPLEX: {
		std::complex<float> v = ((std::complex<float> *)x)[i];
		return zdouble(v.real(), v.imag());
	}
	case XM_SCALAR_DOUBLE: return ((double *)x)[i];
	case XM_SCALAR_DOUBLE_COMPLEX: return ((zdouble *)x)[i];
	}
	XM_FATAL("xm_scalar_get_element", "unexpected scalar type");
}

template <typename T> static void
h_set(void *x, xm_scalar_t a, size_t n)
{
	T v = to_elem<T>(a), *p = (T *)x;
	for (size_t i = 0; i < n; i++)
		p[i] = v;
}

template <typename T> static void
h_scale(void *x, xm_scalar_t a, size_t n)
{
	T v = to_elem<T>(a), *p = (T *)x;
	for (size_t i = 0; i < n; i++)
		p[i] = emul(p[i], v);
}

template <typename T> static void
h_axpy(void *x, xm_scalar_t a, const void *y, xm_scalar_t b, size_t n)
{
	T aa = to_elem<T>(a), bb = to_elem<T>(b), *p = (T *)x;
	const T *q = (const T *)y;
	for (size_t i = 0; i < n; i++) {
		T u = emul(aa, p[i]), w = emul(bb, q[i]);
		p[i] = u + w;
	}
}

template <typename T> static void
h_vmul(void *x, xm_scalar_t a, const void *y, size_t n)
{
	T aa = to_elem<T>(a), *p = (T *)x;
	const T *q = (const T *)y;
	for (size_t i = 0; i < n; i++)
		p[i] = emul(emul(aa, p[i]), q[i]);
}

template <typename T> static void
h_vdiv(void *x, xm_scalar_t a, const void *y, size_t n)
{
	T aa = to_elem<T>(a), *p = (T *)x;
	const T *q = (const T *)y;
	for (size_t i = 0; i < n; i++)
		p[i] = p[i] / emul(aa, q[i]);
}

XM_API void
xm_scalar_set(void *x, xm_scalar_t a, size_t n, xm_scalar_type_t t)
{
	ON_TYPE(t, h_set, x, a, n);
	XM_FATAL("xm_scalar_set", "unexpected scalar type");
}

XM_API void
xm_scalar_scale(void *x, xm_scalar_t a, size_t n, xm_scalar_type_t t)
{
	ON_TYPE(t, h_scale, x, a, n);
	XM_FATAL("xm_scalar_scale", "unexpected scalar type");
}

XM_API void
xm_scalar_axpy(void *x, xm_scalar_t a, const void *y, xm_scalar_t b,
    size_t n, xm_scalar_type_t t)
{
	ON_TYPE(t, h_axpy, x, a, y, b, n);
	XM_FATAL("xm_scalar_axpy", "unexpected scalar type");
}

XM_API void
xm_scalar_vec_mul(void *x, xm_scalar_t a, const void *y, size_t n,
    xm_scalar_type_t t)
{
	ON_TYPE(t, h_vmul, x, a, y, n);
	XM_FATAL("xm_scalar_vec_mul", "unexpected scalar type");
}

XM_API void
xm_scalar_vec_div(void *x, xm_scalar_t a, const void *y, size_t n,
    xm_scalar_type_t t)
{
	ON_TYPE(t, h_vdiv, x, a, y, n);
	XM_FATAL("xm_scalar_vec_div", "unexpected scalar type");
}

XM_API xm_scalar_t
xm_scalar_dot(const void *x, const void *y, size_t n, xm_scalar_type_t t)
{
	switch (t) {
	case XM_SCALAR_FLOAT: {
		const float *a = (const float *)x, *b = (const float *)y;
		float s = 0;
		for (size_t i = 0; i < n; i++)
			s += a[i] * b[i];
		return (double)s;
	}
	case XM_SCALAR_FLOAT_COMPLEX: {
		const std::complex<float> *a = (const std::complex<float> *)x;
		const std::complex<float> *b = (const std::complex<float> *)y;
		std::complex<float> s = 0;
		for (size_t i = 0; i < n; i++)
			s += cmulf(a[i], b[i]);
		return zdouble(s.real(), s.imag());
	}
	case XM_SCALAR_DOUBLE: {
		const double *a = (const double *)x, *b = (const double *)y;
		double s = 0;
		for (size_t i = 0; i < n; i++)
			s += a[i] * b[i];
		return s;
	}
	case XM_SCALAR_DOUBLE_COMPLEX: {
		const zdouble *a = (const zdouble *)x, *b = (const zdouble *)y;
		zdouble s = 0;
		for (size_t i = 0; i < n; i++)
			s += cmul(a[i], b[i]);
		return s;
	}
	}
	XM_FATAL("xm_scalar_dot", "unexpected scalar type");
}

XM_API void
xm_scalar_convert(void *x, const void *y, size_t n, xm_scalar_type_t xt,
    xm_scalar_type_t yt)
{
	if (xt == yt) {
		memcpy(x, y, n * xm_scalar_sizeof(xt));
		return;
	}
	if (xt == XM_SCALAR_DOUBLE && yt == XM_SCALAR_FLOAT) {
		for (size_t i = 0; i < n; i++)
			((double *)x)[i] = ((const float *)y)[i];
		return;
	}
	if (xt == XM_SCALAR_FLOAT && yt == XM_SCALAR_DOUBLE) {
		for (size_t i = 0; i < n; i++)
			((float *)x)[i] = (float)((const double *)y)[i];
		return;
	}
	if (xt == XM_SCALAR_DOUBLE_COMPLEX && yt == XM_SCALAR_FLOAT_COMPLEX) {
		for (size_t i = 0; i < 2 * n; i++)
			((double *)x)[i] = ((const float *)y)[i];
		return;
	}
	if (xt == XM_SCALAR_FLOAT_COMPLEX && yt == XM_SCALAR_DOUBLE_COMPLEX) {
		for (size_t i = 0; i < 2 * n; i++)
			((float *)x)[i] = (float)((const double *)y)[i];
		return;
	}
	XM_FATAL("xm_scalar_convert", "unsupported scalar conversion");
}

/* ------------------------------------------------------------------ */
/* tensor block map (libxm src/tensor.c)                               */
/* ------------------------------------------------------------------ */

static std::atomic<uint64_t> next_tensor_id{1};

size_t
block_flat(const xm_tensor_t *t, const xm_dim_t &idx)
{
	if (idx.n != t->bs->dims.n || !xm_dim_less(&idx, &t->bs->nblocks))
		XM_FATAL("xm_tensor", "block index out of range");
	return xm_dim_offset(&idx, &t->bs->nblocks);
}

xm_block *
tensor_block(const xm_tensor_t *t, const xm_dim_t &idx)
{
	return const_cast<xm_block *>(&t->blocks[block_flat(t, idx)]);
}

xm_dim_t
grid_index(const xm_tensor_t *t, size_t flat)
{
	return xm_dim_from_offset(flat, &t->bs->nblocks);
}

xm_dim_t
block_dims_flat(const xm_tensor_t *t, size_t flat)
{
	return xm_block_space_get_block_dims(t->bs, grid_index(t, flat));
}

size_t
block_elems(const xm_tensor_t *t, const xm_dim_t &idx)
{
	return xm_block_space_get_block_size(t->bs, idx);
}

void
bump_version(xm_tensor_t *t)
{
	t->version++;
}

uint64_t
block_source_ptr(const xm_tensor_t *t, const xm_dim_t &idx)
{
	const xm_block *b = tensor_block(t, idx);
	if (b->type == XM_BLOCK_TYPE_CANONICAL)
		return b->data;
	if (b->type == XM_BLOCK_TYPE_DERIVATIVE)
		return t->blocks[b->data].data;
	return XM_NULL_PTR;
}

XM_API void
xm_tensor_set_zero_block(xm_tensor_t *t, xm_dim_t idx)
{
	xm_block *b = tensor_block(t, idx);
	b->type = XM_BLOCK_TYPE_ZERO;
	b->perm = xm_dim_identity_permutation(idx.n);
	b->scalar = 0;
	b->data = XM_NULL_PTR;
	bump_version(t);
}

/* libxm src/tensor.c:53-81 */
XM_API xm_tensor_t *
xm_tensor_create(const xm_block_space_t *bs, xm_scalar_type_t type,
    xm_allocator_t *alloc)
{
	if (!xm_scalar_check_type(type))
		XM_FATAL("xm_tensor_create", "unexpected scalar type");
	if (bs == nullptr || alloc == nullptr)
		XM_FATAL("xm_tensor_create", "invalid argument");
	xm_tensor_t *t = new (std::nothrow) xm_tensor_t;
	if (t == nullptr || (t->bs = xm_block_space_clone(bs)) == nullptr)
		XM_FATAL("xm_tensor_create", "out of memory");
	t->type = type;
	t->alloc = alloc;
	t->id = next_tensor_id++;
	t->version = 0;
	size_t n = xm_dim_dot(&bs->nblocks);
	xm_block z;
	z.type = XM_BLOCK_TYPE_ZERO;
	z.perm = xm_dim_identity_permutation(bs->dims.n);
	z.scalar = 0;
	z.data = XM_NULL_PTR;
	t->blocks.assign(n, z);
	return t;
}

XM_API xm_tensor_t *
xm_tensor_create_canonical(const xm_block_space_t *bs, xm_scalar_type_t type,
    xm_allocator_t *alloc)
{
	xm_tensor_t *t = xm_tensor_create(bs, type, alloc);
	xm_dim_t nb = bs->nblocks;
	for (xm_dim_t i = xm_dim_zero(nb.n); xm_dim_ne(&i, &nb);
	    xm_dim_inc(&i, &nb))
		xm_tensor_set_canonical_block(t, i);
	return t;
}

/* Structure copy with fresh canonical storage; derivative entries are
 * copied as they are (libxm src/tensor.c:100-122). */
XM_API xm_tensor_t *
xm_tensor_create_structure(const xm_tensor_t *src, xm_scalar_type_t type,
    xm_allocator_t *alloc)
{
	if (alloc == nullptr)
		alloc = src->alloc;
	xm_tensor_t *t = xm_tensor_create(src->bs, type, alloc);
	xm_dim_t nb = src->bs->nblocks;
	for (xm_dim_t i = xm_dim_zero(nb.n); xm_dim_ne(&i, &nb);
	    xm_dim_inc(&i, &nb)) {
		size_t o = xm_dim_offset(&i, &nb);
		t->blocks[o] = src->blocks[o];
		if (t->blocks[o].type == XM_BLOCK_TYPE_CANONICAL) {
			t->blocks[o].type = XM_BLOCK_TYPE_ZERO;
			xm_tensor_set_canonical_block(t, i);
		}
	}
	bump_version(t);
	return t;
}

XM_API const xm_block_space_t *
xm_tensor_get_block_space(const xm_tensor_t *t) { return t->bs; }
XM_API xm_scalar_type_t
xm_tensor_get_scalar_type(const xm_tensor_t *t) { return t->type; }
XM_API xm_allocator_t *
xm_tensor_get_allocator(const xm_tensor_t *t) { return t->alloc; }
XM_API xm_dim_t
xm_tensor_get_abs_dims(const xm_tensor_t *t) { return t->bs->dims; }
XM_API xm_dim_t
xm_tensor_get_nblocks(const xm_tensor_t *t) { return t->bs->nblocks; }

XM_API xm_block_type_t
xm_tensor_get_block_type(const xm_tensor_t *t, xm_dim_t idx)
{
	return tensor_block(t, idx)->type;
}

XM_API xm_dim_t
xm_tensor_get_block_dims(const xm_tensor_t *t, xm_dim_t idx)
{
	return xm_block_space_get_block_dims(t->bs, idx);
}

XM_API size_t
xm_tensor_get_block_size(const xm_tensor_t *t, xm_dim_t idx)
{
	return xm_block_space_get_block_size(t->bs, idx);
}

XM_API size_t
xm_tensor_get_block_bytes(const xm_tensor_t *t, xm_dim_t idx)
{
	return xm_tensor_get_block_size(t, idx) * elem_size(t->type);
}

XM_API size_t
xm_tensor_get_largest_block_size(const xm_tensor_t *t)
{
	return xm_block_space_get_largest_block_size(t->bs);
}

XM_API size_t
xm_tensor_get_largest_block_bytes(const xm_tensor_t *t)
{
	return xm_tensor_get_largest_block_size(t) * elem_size(t->type);
}

XM_API uint64_t
xm_tensor_get_block_data_ptr(const xm_tensor_t *t, xm_dim_t idx)
{
	return block_source_ptr(t, idx);
}

XM_API xm_dim_t
xm_tensor_get_block_permutation(const xm_tensor_t *t, xm_dim_t idx)
{
	return tensor_block(t, idx)->perm;
}

XM_API xm_scalar_t
xm_tensor_get_block_scalar(const xm_tensor_t *t, xm_dim_t idx)
{
	return tensor_block(t, idx)->scalar;
}

XM_API void
xm_tensor_set_canonical_block_raw(xm_tensor_t *t, xm_dim_t idx, uint64_t ptr)
{
	if (ptr == XM_NULL_PTR)
		XM_FATAL("xm_tensor_set_canonical_block_raw",
		    "unexpected null data pointer");
	if (xm_tensor_get_block_type(t, idx) != XM_BLOCK_TYPE_ZERO)
		XM_FATAL("xm_tensor_set_canonical_block_raw",
		    "block must be zero");
	xm_block *b = tensor_block(t, idx);
	b->type = XM_BLOCK_TYPE_CANONICAL;
	b->perm = xm_dim_identity_permutation(idx.n);
	b->scalar = 1;
	b->data = ptr;
	bump_version(t);
}

XM_API void
xm_tensor_set_canonical_block(xm_tensor_t *t, xm_dim_t idx)
{
	if (xm_tensor_get_block_type(t, idx) != XM_BLOCK_TYPE_ZERO)
		XM_FATAL("xm_tensor_set_canonical_block", "block must be zero");
	uint64_t ptr = xm_allocator_allocate(t->alloc,
	    xm_tensor_get_block_bytes(t, idx));
	if (ptr == XM_NULL_PTR)
		XM_FATAL("xm_tensor_set_canonical_block",
		    "unable to allocate block data");
	xm_tensor_set_canonical_block_raw(t, idx, ptr);
}

/* libxm src/tensor.c:287-311 */
XM_API void
xm_tensor_set_derivative_block(xm_tensor_t *t, xm_dim_t idx, xm_dim_t src,
    xm_dim_t perm, xm_scalar_t s)
{
	static const char *fn = "xm_tensor_set_derivative_block";

	if (xm_tensor_get_block_type(t, idx) != XM_BLOCK_TYPE_ZERO)
		XM_FATAL(fn, "block must be zero");
	if (xm_tensor_get_block_type(t, src) != XM_BLOCK_TYPE_CANONICAL)
		XM_FATAL(fn, "derivative blocks must have canonical source blocks");
	if (perm.n != idx.n)
		XM_FATAL(fn, "invalid block permutation");
	{
		bool seen[XM_MAX_DIM] = {false};
		for (size_t k = 0; k < perm.n; k++) {
			if (perm.i[k] >= perm.n || seen[perm.i[k]])
				XM_FATAL(fn, "invalid block permutation");
			seen[perm.i[k]] = true;
		}
	}
	xm_dim_t d1 = xm_block_space_get_block_dims(t->bs, idx);
	d1 = xm_dim_permute(&d1, &perm);
	xm_dim_t d2 = xm_block_space_get_block_dims(t->bs, src);
	if (xm_dim_ne(&d1, &d2))
		XM_FATAL(fn, "invalid block permutation");
	xm_block *b = tensor_block(t, idx);
	b->type = XM_BLOCK_TYPE_DERIVATIVE;
	b->perm = perm;
	b->scalar = s;
	b->data = block_flat(t, src);
	bump_version(t);
}

XM_API void
xm_tensor_get_canonical_block_list(const xm_tensor_t *t, xm_dim_t **list,
    size_t *nlist)
{
	size_t n = 0;
	for (const xm_block &b : t->blocks)
		n += b.type == XM_BLOCK_TYPE_CANONICAL;
	xm_dim_t *l = (xm_dim_t *)malloc((n ? n : 1) * sizeof *l);
	if (l == nullptr)
		XM_FATAL("xm_tensor_get_canonical_block_list", "out of memory");
	size_t k = 0;
	for (size_t f = 0; f < t->blocks.size(); f++)
		if (t->blocks[f].type == XM_BLOCK_TYPE_CANONICAL)
			l[k++] = grid_index(t, f);
	*list = l;
	*nlist = n;
}

XM_API void
xm_tensor_read_block(const xm_tensor_t *t, xm_dim_t idx, void *buf)
{
	if (xm_tensor_get_block_type(t, idx) == XM_BLOCK_TYPE_ZERO)
		XM_FATAL("xm_tensor_read_block",
		    "cannot read data from zero-blocks");
	xm_allocator_read(t->alloc, block_source_ptr(t, idx), buf,
	    xm_tensor_get_block_bytes(t, idx));
}

XM_API void
xm_tensor_write_block(xm_tensor_t *t, xm_dim_t idx, const void *buf)
{
	if (xm_tensor_get_block_type(t, idx) != XM_BLOCK_TYPE_CANONICAL)
		XM_FATAL("xm_tensor_write_block",
		    "can only write to canonical blocks");
	xm_allocator_write(t->alloc, block_source_ptr(t, idx), buf,
	    xm_tensor_get_block_bytes(t, idx));
}

/*
 * Device data epoch: bumped by every entry point that writes block storage
 * (copies into blocks, kernels, deallocation).  Host copies of blocks read
 * by xm_tensor_get_element stay valid while the epoch is unchanged.
 */
static std::atomic<uint64_t> data_epoch{1};

void
bump_data_epoch(void)
{
	data_epoch++;
}

struct element_cache {
	std::mutex mu;
	uint64_t epoch = 0;
	size_t bytes = 0;
	std::map<uint64_t, std::vector<unsigned char>> blocks; /* by address */
};
static element_cache ecache;

/* One element (libxm src/tensor.c:154-178).  The block's storage is copied
 * to the host once per data epoch and kept, so element-by-element loops
 * over a tensor cost one device copy per block, not one per element. */
XM_API xm_scalar_t
xm_tensor_get_element(const xm_tensor_t *t, xm_dim_t idx)
{
	xm_dim_t bi, ei;
	xm_block_space_decompose_index(t->bs, idx, &bi, &ei);
	const xm_block *b = tensor_block(t, bi);
	if (b->type == XM_BLOCK_TYPE_ZERO)
		return 0;
	xm_dim_t bd = xm_tensor_get_block_dims(t, bi);
	xm_dim_t pe = xm_dim_permute(&ei, &b->perm);
	xm_dim_t pd = xm_dim_permute(&bd, &b->perm);
	size_t off = xm_dim_offset(&pe, &pd), es = elem_size(t->type);
	size_t bytes = xm_dim_dot(&bd) * es;
	uint64_t src = block_source_ptr(t, bi);
	unsigned char v[16];
	{
		std::lock_guard<std::mutex> g(ecache.mu);
		uint64_t ep = data_epoch.load();
		if (ecache.epoch != ep || ecache.bytes > (256u << 20)) {
			ecache.blocks.clear();
			ecache.bytes = 0;
			ecache.epoch = ep;
		}
		auto it = ecache.blocks.find(src);
		if (it == ecache.blocks.end() || it->second.size() < bytes) {
			std::vector<unsigned char> h(bytes);
			if (file_ptr(src)) {
				arena_file_read("xm_tensor_get_element", src,
				    h.data(), bytes);
			} else {
				HIP_OK("xm_tensor_get_element", hipMemcpyAsync(
				    h.data(), (const void *)(uintptr_t)src,
				    bytes, hipMemcpyDeviceToHost, xm_stream()));
				HIP_OK("xm_tensor_get_element",
				    hipStreamSynchronize(xm_stream()));
			}
			ecache.bytes += bytes;
			it = ecache.blocks.insert_or_assign(src, std::move(h)).first;
		}
		memcpy(v, it->second.data() + off * es, es);
	}
	return scalar_mul(b->scalar, xm_scalar_get_element(v, 0, t->type),
	    t->type);
}

/*
 * Host unfold / fold (libxm src/tensor.c:368-639): matrix element (i, j)
 * lives at to[j*stride + i]; i runs over the mask_i indices (first
 * fastest), j over mask_j.  The source element is located through the
 * block permutation; the block scalar is not applied.
 */
static void
host_move(const char *fn, const xm_tensor_t *t, xm_dim_t idx, xm_dim_t mi,
    xm_dim_t mj, const void *from, void *to, size_t stride, bool fold)
{
	if (from == nullptr || to == nullptr || from == to)
		XM_FATAL(fn, "invalid argument");
	if (mi.n + mj.n != idx.n)
		XM_FATAL(fn, "invalid mask dimensions");
	if (fold && xm_tensor_get_block_type(t, idx) != XM_BLOCK_TYPE_CANONICAL)
		XM_FATAL(fn, "can only fold canonical blocks");
	xm_dim_t dims = xm_tensor_get_block_dims(t, idx);
	xm_dim_t perm = fold ? xm_dim_identity_permutation(idx.n) :
	    tensor_block(t, idx)->perm;
	xm_dim_t pd = xm_dim_permute(&dims, &perm);
	size_t stride_src[XM_MAX_DIM], s = 1;
	for (size_t k = 0; k < pd.n; k++) {
		stride_src[k] = s;
		s *= pd.i[k];
	}
	/* source stride of each logical block index */
	size_t lstride[XM_MAX_DIM];
	for (size_t k = 0; k < dims.n; k++)
		lstride[k] = stride_src[perm.i[k]];
	size_t ni = xm_dim_dot_mask(&dims, &mi);
	size_t nj = xm_dim_dot_mask(&dims, &mj);
	size_t es = elem_size(t->type);
	xm_dim_t e = xm_dim_zero(dims.n);
	for (size_t j = 0; j < nj; j++) {
		xm_dim_zero_mask(&e, &mi);
		for (size_t i = 0; i < ni; i++) {
			size_t off = 0;
			for (size_t k = 0; k < dims.n; k++)
				off += e.i[k] * lstride[k];
			size_t mat = j * stride + i;
			if (fold)
				memcpy((char *)to + off * es,
				    (const char *)from + mat * es, es);
			else
				memcpy((char *)to + mat * es,
				    (const char *)from + off * es, es);
			xm_dim_inc_mask(&e, &dims, &mi);
		}
		xm_dim_inc_mask(&e, &dims, &mj);
	}
}

XM_API void
xm_tensor_unfold_block(const xm_tensor_t *t, xm_dim_t idx, xm_dim_t mi,
    xm_dim_t mj, const void *from, void *to, size_t stride)
{
	host_move("xm_tensor_unfold_block", t, idx, mi, mj, from, to, stride,
	    false);
}

XM_API void
xm_tensor_fold_block(const xm_tensor_t *t, xm_dim_t idx, xm_dim_t mi,
    xm_dim_t mj, const void *from, void *to, size_t stride)
{
	host_move("xm_tensor_fold_block", t, idx, mi, mj, from, to, stride,
	    true);
}

XM_API void
xm_tensor_free_block_data(xm_tensor_t *t)
{
	for (size_t f = 0; f < t->blocks.size(); f++) {
		xm_block &b = t->blocks[f];
		if (b.type == XM_BLOCK_TYPE_CANONICAL)
			xm_allocator_deallocate(t->alloc, b.data);
		b.type = XM_BLOCK_TYPE_ZERO;
		b.perm = xm_dim_identity_permutation(t->bs->dims.n);
		b.scalar = 0;
		b.data = XM_NULL_PTR;
	}
	bump_version(t);
}

void plan_cache_drop(uint64_t tensor_id);
void op_cache_drop(uint64_t tensor_id);
void bwd_cache_drop(uint64_t tensor_id);

XM_API void
xm_tensor_free(xm_tensor_t *t)
{
	if (t) {
		bwd_cache_drop(t->id);
		plan_cache_drop(t->id);
		op_cache_drop(t->id);
		xm_block_space_free(t->bs);
		delete t;
	}
}

XM_API void
xm_print_banner(void)
{
	printf("libxm API on AMD Instinct MI355X (gfx950): block-sparse "
	    "tensor contractions in HBM\n");
}
