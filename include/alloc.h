/*
 * alloc.h - block storage of the libxm C API (MI355X build).
 *
 * Same prototypes as libxm's src/alloc.h.  Storage is an arena in the
 * GPU's HBM: data_ptr is an opaque 64-bit handle (here the device address
 * of the allocation).  xm_allocator_read/write copy between a host buffer
 * and the device allocation and return when the copy has completed.
 * When HBM is exhausted (or the arena limit set by the environment
 * variable XM_ARENA_BYTES is reached) allocations fall back to pinned,
 * device-mapped host memory so that tensors larger than HBM still work.
 */
#ifndef XM_ALLOC_H
#define XM_ALLOC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Null data pointer (libxm src/alloc.h:31). */
#define XM_NULL_PTR ((uint64_t)(-1))

/* Opaque allocator (libxm src/alloc.h:34). */
typedef struct xm_allocator xm_allocator_t;

/* Create an allocator.  path == NULL or not, data live in HBM; the path is
 * only remembered and returned by xm_allocator_get_path
 * (libxm src/alloc.h:41). */
xm_allocator_t *xm_allocator_create(const char *path);
/* libxm src/alloc.h:46 */
const char *xm_allocator_get_path(xm_allocator_t *allocator);
/* libxm src/alloc.h:53 */
uint64_t xm_allocator_allocate(xm_allocator_t *allocator, size_t size_bytes);
/* Device -> host copy of a whole allocation (libxm src/alloc.h:62). */
void xm_allocator_read(xm_allocator_t *allocator, uint64_t data_ptr,
    void *mem, size_t size_bytes);
/* Host -> device copy of a whole allocation (libxm src/alloc.h:71). */
void xm_allocator_write(xm_allocator_t *allocator, uint64_t data_ptr,
    const void *mem, size_t size_bytes);
/* libxm src/alloc.h:77 */
void xm_allocator_deallocate(xm_allocator_t *allocator, uint64_t data_ptr);
/* libxm src/alloc.h:81 */
void xm_allocator_destroy(xm_allocator_t *allocator);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* XM_ALLOC_H */
