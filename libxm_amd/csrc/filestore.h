/*
 * filestore.h - the file tier of allocators created with a path: block
 * storage in a regular file once the HBM and pinned-host tiers are full
 * (libxm keeps all data of such allocators in that file, libxm
 * src/alloc.c:147-347), and the layout of file ranges in a staging
 * buffer.  Plain C++ and POSIX I/O, no device code: the GPU never reads
 * the file, it reads pinned copies that the staging code fills with pread
 * and writes back with pwrite.
 */
#ifndef XM_FILESTORE_H
#define XM_FILESTORE_H

#include <cstddef>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

/*
 * Data pointers of file-tier allocations: bit 62 set, the file's slot in
 * bits 48..61 and the byte offset in the file in bits 0..47.  Device and
 * host addresses on x86-64 stay below 2^47; XM_NULL_PTR (all ones) is
 * excluded explicitly.
 */
#define FILE_PTR_TAG (1ull << 62)
#define FILE_SLOT_SHIFT 48
#define FILE_SLOT_MAX ((1u << 14) - 1)
#define FILE_OFF_MASK ((1ull << FILE_SLOT_SHIFT) - 1)
/* File allocations are rounded to whole pages. */
#define FILE_ALIGN 4096

static inline bool
file_ptr(uint64_t p)
{
	return p != ~0ull && (p & FILE_PTR_TAG) != 0;
}

static inline uint32_t
file_slot(uint64_t p)
{
	return (uint32_t)((p & ~FILE_PTR_TAG) >> FILE_SLOT_SHIFT);
}

static inline uint64_t
file_off(uint64_t p)
{
	return p & FILE_OFF_MASK;
}

static inline uint64_t
file_make_ptr(uint32_t slot, uint64_t off)
{
	return FILE_PTR_TAG | ((uint64_t)slot << FILE_SLOT_SHIFT) | off;
}

/*
 * One backing file: first-fit ranges (whole pages) from a sorted,
 * coalesced free list; the file grows with ftruncate (sparse until
 * written) when no free range fits.
 */
struct file_store {
	int fd = -1;
	std::string path;
	uint64_t length = 0;		/* file size */
	uint64_t used = 0;		/* bytes in live ranges */
	std::vector<std::pair<uint64_t, uint64_t>> free_list; /* off, bytes */

	/* open(path, O_CREAT | O_RDWR, S_IRUSR | S_IWUSR) and truncate to
	 * zero length, as libxm's pagefile; false (errno set, message on
	 * stderr) when either fails. */
	bool create(const char *path);
	/* Offset of a new range of at least `bytes`; UINT64_MAX when the
	 * file cannot grow (message on stderr). */
	uint64_t allocate(uint64_t bytes);
	/* Return the range [off, off + bytes) (bytes as passed to
	 * allocate). */
	void release(uint64_t off, uint64_t bytes);
	/* Whole transfers, in pieces of at most 1 GiB, retried after EINTR
	 * and short counts; false on an I/O error or end of file. */
	bool read(uint64_t off, void *buf, size_t n) const;
	bool write(uint64_t off, const void *buf, size_t n) const;
	/* close and unlink (messages on stderr if they fail). */
	void destroy();
};

/*
 * File ranges of one staging batch laid out in a buffer: requests
 * (pointer, bytes) merged where they overlap or touch within one file,
 * each merged range at a 256-byte aligned offset.  `load`: the range's
 * contents are read before the batch; `store`: written back after it.
 */
struct stage_layout {
	struct range {
		uint64_t p, n;		/* file pointer, bytes */
		uint64_t off;		/* offset in the buffer */
		bool load, store;
	};
	std::vector<range> r;
	uint64_t bytes = 0;		/* buffer bytes the layout needs */
	bool done = false;

	void add(uint64_t p, uint64_t n, bool load, bool store);
	/* Merge and lay out (idempotent; add() afterwards is an error). */
	void finish();
	/* Buffer offset of file pointer p, which must lie in a range
	 * (UINT64_MAX otherwise). */
	uint64_t offset(uint64_t p) const;
	bool empty() const { return r.empty(); }
};

#endif /* XM_FILESTORE_H */
