/*
 * dim.h - multi-index algebra of the libxm C API (MI355X build).
 *
 * Same type layout, constants and prototypes as libxm's src/dim.h so that
 * code written against libxm compiles and links unchanged.  xm_dim_t is a
 * 72-byte value type ({size_t n; size_t i[8]}) passed by value.
 * Element layout everywhere in the library is first-index-fastest
 * (column-major): offset(e, d) = sum_k e_k * prod_{l<k} d_l.
 */
#ifndef XM_DIM_H_INCLUDED
#define XM_DIM_H_INCLUDED

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Maximum number of dimensions (libxm src/dim.h:30). */
#define XM_MAX_DIM 8

/* Multi-index / dimension vector (libxm src/dim.h:33-38). */
typedef struct {
	size_t n;		/* number of dimensions, <= XM_MAX_DIM */
	size_t i[XM_MAX_DIM];	/* only the first n entries are used */
} xm_dim_t;

/* Constructors (libxm src/dim.h:43-120). */
xm_dim_t xm_dim_zero(size_t ndim);
xm_dim_t xm_dim_same(size_t ndim, size_t dim);
xm_dim_t xm_dim_1(size_t dim1);
xm_dim_t xm_dim_2(size_t dim1, size_t dim2);
xm_dim_t xm_dim_3(size_t dim1, size_t dim2, size_t dim3);
xm_dim_t xm_dim_4(size_t dim1, size_t dim2, size_t dim3, size_t dim4);
xm_dim_t xm_dim_5(size_t dim1, size_t dim2, size_t dim3, size_t dim4,
    size_t dim5);
xm_dim_t xm_dim_6(size_t dim1, size_t dim2, size_t dim3, size_t dim4,
    size_t dim5, size_t dim6);
xm_dim_t xm_dim_7(size_t dim1, size_t dim2, size_t dim3, size_t dim4,
    size_t dim5, size_t dim6, size_t dim7);
xm_dim_t xm_dim_8(size_t dim1, size_t dim2, size_t dim3, size_t dim4,
    size_t dim5, size_t dim6, size_t dim7, size_t dim8);

/* Comparisons (libxm src/dim.h:127-139). */
int xm_dim_eq(const xm_dim_t *a, const xm_dim_t *b);
int xm_dim_ne(const xm_dim_t *a, const xm_dim_t *b);
int xm_dim_less(const xm_dim_t *idx, const xm_dim_t *dim);

/* Product of all entries (libxm src/dim.h:144). */
size_t xm_dim_dot(const xm_dim_t *dim);

/* Column-major offset and its inverse (libxm src/dim.h:150-156). */
size_t xm_dim_offset(const xm_dim_t *idx, const xm_dim_t *dim);
xm_dim_t xm_dim_from_offset(size_t offset, const xm_dim_t *dim);

/* First-index-fastest increment; past the end idx becomes dim
 * (libxm src/dim.h:162). */
void xm_dim_inc(xm_dim_t *idx, const xm_dim_t *dim);

/* Masked helpers (libxm src/dim.h:167-188). */
void xm_dim_zero_mask(xm_dim_t *dim, const xm_dim_t *mask);
void xm_dim_set_mask(xm_dim_t *a, const xm_dim_t *maska, const xm_dim_t *b,
    const xm_dim_t *maskb);
size_t xm_dim_dot_mask(const xm_dim_t *dim, const xm_dim_t *mask);
void xm_dim_inc_mask(xm_dim_t *idx, const xm_dim_t *dim, const xm_dim_t *mask);

/* Permutations: permute(idx, p).i[p.i[k]] = idx.i[k]
 * (libxm src/dim.h:193-199). */
xm_dim_t xm_dim_identity_permutation(size_t ndim);
xm_dim_t xm_dim_permute(const xm_dim_t *idx, const xm_dim_t *permutation);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* XM_DIM_H_INCLUDED */
