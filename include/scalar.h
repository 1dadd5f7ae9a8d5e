/*
 * scalar.h - scalar types and typed vector helpers of the libxm C API
 * (MI355X build).  Same enum values, scalar type and prototypes as
 * libxm's src/scalar.h.  The vector helpers operate on HOST buffers; the
 * tensor operations run their element work on the GPU.
 */
#ifndef XM_SCALAR_H_INCLUDED
#define XM_SCALAR_H_INCLUDED

#include <stddef.h>

#ifdef __cplusplus
#include <complex>
#else
#include <complex.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Element type of a tensor (libxm src/scalar.h:36-45). */
typedef enum {
	XM_SCALAR_FLOAT = 0,
	XM_SCALAR_FLOAT_COMPLEX,
	XM_SCALAR_DOUBLE,
	XM_SCALAR_DOUBLE_COMPLEX,
} xm_scalar_type_t;

/* Widest scalar, convertible to every element type (libxm
 * src/scalar.h:49-53).  Both spellings travel in two SSE registers under
 * the SysV x86-64 ABI, so C and C++ callers interoperate. */
#ifdef __cplusplus
typedef std::complex<double> xm_scalar_t;
#else
typedef double complex xm_scalar_t;
#endif

/* libxm src/scalar.h:58 */
int xm_scalar_check_type(xm_scalar_type_t type);
/* libxm src/scalar.h:63 */
size_t xm_scalar_sizeof(xm_scalar_type_t type);
/* Arithmetic rounded to the element type (libxm src/scalar.h:70-77). */
xm_scalar_t xm_scalar_add(xm_scalar_t a, xm_scalar_t b, xm_scalar_type_t type);
xm_scalar_t xm_scalar_mul(xm_scalar_t a, xm_scalar_t b, xm_scalar_type_t type);
/* libxm src/scalar.h:84 */
xm_scalar_t xm_scalar_get_element(void *x, size_t idx, xm_scalar_type_t type);
/* x[i] = a (libxm src/scalar.h:91) */
void xm_scalar_set(void *x, xm_scalar_t a, size_t len, xm_scalar_type_t type);
/* x[i] *= a (libxm src/scalar.h:98) */
void xm_scalar_scale(void *x, xm_scalar_t a, size_t len, xm_scalar_type_t type);
/* x[i] = a*x[i] + b*y[i] (libxm src/scalar.h:107) */
void xm_scalar_axpy(void *x, xm_scalar_t a, const void *y, xm_scalar_t b,
    size_t len, xm_scalar_type_t type);
/* x[i] = a*x[i]*y[i] (libxm src/scalar.h:116) */
void xm_scalar_vec_mul(void *x, xm_scalar_t a, const void *y, size_t len,
    xm_scalar_type_t type);
/* x[i] = x[i]/(a*y[i]) (libxm src/scalar.h:125) */
void xm_scalar_vec_div(void *x, xm_scalar_t a, const void *y, size_t len,
    xm_scalar_type_t type);
/* sum_i x[i]*y[i], no conjugation (libxm src/scalar.h:134) */
xm_scalar_t xm_scalar_dot(const void *x, const void *y, size_t len,
    xm_scalar_type_t type);
/* precision conversion between real/real or complex/complex
 * (libxm src/scalar.h:143) */
void xm_scalar_convert(void *x, const void *y, size_t len,
    xm_scalar_type_t xtype, xm_scalar_type_t ytype);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* XM_SCALAR_H_INCLUDED */
