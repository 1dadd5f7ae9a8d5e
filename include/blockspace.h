/*
 * blockspace.h - block spaces (per-dimension split points) of the libxm C
 * API (MI355X build).  Same prototypes as libxm's src/blockspace.h.
 */
#ifndef XM_BLOCKSPACE_H_INCLUDED
#define XM_BLOCKSPACE_H_INCLUDED

#include "dim.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque block space (libxm src/blockspace.h:30). */
typedef struct xm_block_space xm_block_space_t;

/* libxm src/blockspace.h:35 */
xm_block_space_t *xm_block_space_create(xm_dim_t dims);
/* libxm src/blockspace.h:40 */
xm_block_space_t *xm_block_space_clone(const xm_block_space_t *bs);
/* libxm src/blockspace.h:46 */
xm_block_space_t *xm_block_space_permute_clone(const xm_block_space_t *bs,
    xm_dim_t permutation);
/* libxm src/blockspace.h:52-62 */
size_t xm_block_space_get_ndims(const xm_block_space_t *bs);
xm_dim_t xm_block_space_get_abs_dims(const xm_block_space_t *bs);
xm_dim_t xm_block_space_get_nblocks(const xm_block_space_t *bs);
/* libxm src/blockspace.h:68 */
void xm_block_space_split(xm_block_space_t *bs, size_t dim, size_t x);
/* libxm src/blockspace.h:72 */
void xm_block_space_autosplit(xm_block_space_t *bs);
/* libxm src/blockspace.h:79 */
size_t xm_block_space_get_split(xm_block_space_t *bs, size_t dim, size_t i);
/* libxm src/blockspace.h:85-98 */
xm_dim_t xm_block_space_get_block_dims(const xm_block_space_t *bs,
    xm_dim_t blkidx);
size_t xm_block_space_get_block_size(const xm_block_space_t *bs,
    xm_dim_t blkidx);
size_t xm_block_space_get_largest_block_size(const xm_block_space_t *bs);
/* libxm src/blockspace.h:106 */
void xm_block_space_decompose_index(const xm_block_space_t *bs, xm_dim_t idx,
    xm_dim_t *blkidx, xm_dim_t *elidx);
/* libxm src/blockspace.h:113-121 */
int xm_block_space_eq(const xm_block_space_t *bsa, const xm_block_space_t *bsb);
int xm_block_space_eq1(const xm_block_space_t *bsa, size_t dima,
    const xm_block_space_t *bsb, size_t dimb);
/* libxm src/blockspace.h:126 */
void xm_block_space_free(xm_block_space_t *bs);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* XM_BLOCKSPACE_H_INCLUDED */
