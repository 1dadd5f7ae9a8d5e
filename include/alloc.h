/*
 * alloc.h - block storage of the libxm C API (MI355X build).
 *
 * Same prototypes as libxm's src/alloc.h.  data_ptr is an opaque 64-bit
 * handle.  Allocations come from an arena in the GPU's HBM (data_ptr is
 * then the device address); once HBM is exhausted, or the arena reaches
 * the limit set by the environment variable XM_ARENA_BYTES, from pinned,
 * device-mapped host memory (up to XM_HOST_BYTES); and for an allocator
 * created with a path, once that tier is full too, from the file at that
 * path, as libxm's pagefile allocator.  xm_allocator_read/write copy
 * between a host buffer and the allocation and return when the copy has
 * completed.
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

/* Create an allocator (libxm src/alloc.h:41).  With a path, the file is
 * created (open with O_CREAT, truncated; NULL if that fails, as libxm)
 * and holds the allocations that do not fit HBM and pinned host memory;
 * xm_allocator_destroy removes it. */
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
