/*
 * tensor.h - block tensors of the libxm C API (MI355X build).
 *
 * Same enums and prototypes as libxm's src/tensor.h.  A tensor is a grid
 * of blocks; each block is ZERO (no data), CANONICAL (owns data in the
 * allocator's HBM arena) or DERIVATIVE (a canonical source block plus a
 * permutation and a scalar: X(e) = s * S(permute(e, p))).
 * Block metadata lives on the host; block data lives on the GPU.
 * xm_tensor_read_block/write_block copy between a host buffer and the
 * block's device storage; xm_tensor_unfold_block/fold_block operate on
 * host buffers exactly like the reference.
 */
#ifndef TENSOR_H_INCLUDED
#define TENSOR_H_INCLUDED

#include "alloc.h"
#include "blockspace.h"
#include "scalar.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Block kind (libxm src/tensor.h:32-39). */
typedef enum {
	XM_BLOCK_TYPE_ZERO = 0,
	XM_BLOCK_TYPE_CANONICAL,
	XM_BLOCK_TYPE_DERIVATIVE,
} xm_block_type_t;

/* Opaque tensor (libxm src/tensor.h:42). */
typedef struct xm_tensor xm_tensor_t;

/* All blocks ZERO; the block space is cloned (libxm src/tensor.h:49). */
xm_tensor_t *xm_tensor_create(const xm_block_space_t *bs, xm_scalar_type_t type,
    xm_allocator_t *allocator);
/* All blocks CANONICAL, freshly allocated (libxm src/tensor.h:58). */
xm_tensor_t *xm_tensor_create_canonical(const xm_block_space_t *bs,
    xm_scalar_type_t type, xm_allocator_t *allocator);
/* Copy the block structure, allocate new canonical storage; allocator NULL
 * reuses the source's allocator (libxm src/tensor.h:67). */
xm_tensor_t *xm_tensor_create_structure(const xm_tensor_t *tensor,
    xm_scalar_type_t type, xm_allocator_t *allocator);
/* Accessors (libxm src/tensor.h:73-93). */
const xm_block_space_t *xm_tensor_get_block_space(const xm_tensor_t *tensor);
xm_scalar_type_t xm_tensor_get_scalar_type(const xm_tensor_t *tensor);
xm_allocator_t *xm_tensor_get_allocator(const xm_tensor_t *tensor);
xm_dim_t xm_tensor_get_abs_dims(const xm_tensor_t *tensor);
xm_dim_t xm_tensor_get_nblocks(const xm_tensor_t *tensor);
/* One element by absolute index (slow; libxm src/tensor.h:100). */
xm_scalar_t xm_tensor_get_element(const xm_tensor_t *tensor, xm_dim_t idx);
/* Per-block metadata (libxm src/tensor.h:106-157). */
xm_block_type_t xm_tensor_get_block_type(const xm_tensor_t *tensor,
    xm_dim_t blkidx);
xm_dim_t xm_tensor_get_block_dims(const xm_tensor_t *tensor, xm_dim_t blkidx);
size_t xm_tensor_get_block_size(const xm_tensor_t *tensor, xm_dim_t blkidx);
size_t xm_tensor_get_block_bytes(const xm_tensor_t *tensor, xm_dim_t blkidx);
size_t xm_tensor_get_largest_block_size(const xm_tensor_t *tensor);
size_t xm_tensor_get_largest_block_bytes(const xm_tensor_t *tensor);
uint64_t xm_tensor_get_block_data_ptr(const xm_tensor_t *tensor,
    xm_dim_t blkidx);
xm_dim_t xm_tensor_get_block_permutation(const xm_tensor_t *tensor,
    xm_dim_t blkidx);
xm_scalar_t xm_tensor_get_block_scalar(const xm_tensor_t *tensor,
    xm_dim_t blkidx);
/* Structure mutation (libxm src/tensor.h:164-194). */
void xm_tensor_set_zero_block(xm_tensor_t *tensor, xm_dim_t blkidx);
void xm_tensor_set_canonical_block(xm_tensor_t *tensor, xm_dim_t blkidx);
void xm_tensor_set_canonical_block_raw(xm_tensor_t *tensor, xm_dim_t blkidx,
    uint64_t data_ptr);
void xm_tensor_set_derivative_block(xm_tensor_t *tensor, xm_dim_t blkidx,
    xm_dim_t source_blkidx, xm_dim_t permutation, xm_scalar_t scalar);
/* malloc'd list of canonical block indices, first index fastest; the
 * caller frees it (libxm src/tensor.h:204). */
void xm_tensor_get_canonical_block_list(const xm_tensor_t *tensor,
    xm_dim_t **blklist, size_t *nblklist);
/* Block data <-> host buffer (libxm src/tensor.h:211-218). */
void xm_tensor_read_block(const xm_tensor_t *tensor, xm_dim_t blkidx,
    void *buf);
void xm_tensor_write_block(xm_tensor_t *tensor, xm_dim_t blkidx,
    const void *buf);
/* Host-buffer unfold/fold between block layout and matrix form
 * (libxm src/tensor.h:232-247). */
void xm_tensor_unfold_block(const xm_tensor_t *tensor, xm_dim_t blkidx,
    xm_dim_t mask_i, xm_dim_t mask_j, const void *from, void *to,
    size_t stride);
void xm_tensor_fold_block(const xm_tensor_t *tensor, xm_dim_t blkidx,
    xm_dim_t mask_i, xm_dim_t mask_j, const void *from, void *to,
    size_t stride);
/* libxm src/tensor.h:254-259 */
void xm_tensor_free_block_data(xm_tensor_t *tensor);
void xm_tensor_free(xm_tensor_t *tensor);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* TENSOR_H_INCLUDED */
