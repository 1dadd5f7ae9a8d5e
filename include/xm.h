/*
 * xm.h - umbrella header and tensor operations of the libxm C API
 * (MI355X build).  Same prototypes as libxm's src/xm.h.
 *
 * Every operation is synchronous: when it returns, its results are visible
 * to xm_tensor_read_block.  Element work runs on the current HIP device.
 */
#ifndef XM_H_INCLUDED
#define XM_H_INCLUDED

#include "tensor.h"

#ifdef __cplusplus
extern "C" {
#endif

/* libxm src/xm.h:30 */
void xm_print_banner(void);

/* Every element of every canonical block of a <- x (libxm src/xm.h:35). */
void xm_set(xm_tensor_t *a, xm_scalar_t x);

/* a = s * b with index permutation and type conversion
 * (libxm src/xm.h:52). */
void xm_copy(xm_tensor_t *a, xm_scalar_t s, const xm_tensor_t *b,
    const char *idxa, const char *idxb);

/* a = alpha * a + beta * b with index permutation (libxm src/xm.h:69). */
void xm_add(xm_scalar_t alpha, xm_tensor_t *a, xm_scalar_t beta,
    const xm_tensor_t *b, const char *idxa, const char *idxb);

/* a = a * b element-wise (libxm src/xm.h:84). */
void xm_mul(xm_tensor_t *a, const xm_tensor_t *b, const char *idxa,
    const char *idxb);

/* a = a / b element-wise (libxm src/xm.h:98). */
void xm_div(xm_tensor_t *a, const xm_tensor_t *b, const char *idxa,
    const char *idxb);

/* sum of a * b over all elements, no conjugation (libxm src/xm.h:112). */
xm_scalar_t xm_dot(const xm_tensor_t *a, const xm_tensor_t *b,
    const char *idxa, const char *idxb);

/* c = alpha * a * b + beta * c over the canonical blocks of c; contracted
 * indices are those shared by idxa and idxb (libxm src/xm.h:135). */
void xm_contract(xm_scalar_t alpha, const xm_tensor_t *a, const xm_tensor_t *b,
    xm_scalar_t beta, xm_tensor_t *c, const char *idxa, const char *idxb,
    const char *idxc);

#ifdef __cplusplus
} /* extern "C" */
#endif

#include "xm_amd.h"

#endif /* XM_H_INCLUDED */
