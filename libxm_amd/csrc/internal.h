/*
 * internal.h - shared declarations of the MI355X libxm implementation.
 *
 * Host side: block-space / tensor metadata (plain C++), the HBM arena and
 * the device context (one HIP stream per process, default priority).
 * Device side: kernels live in *.hip files and are launched through the
 * launch_* functions declared at the bottom.
 */
#ifndef XM_INTERNAL_H
#define XM_INTERNAL_H

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "xm.h"
#include "util.h"
#include "filestore.h"

#define XM_API extern "C" __attribute__((visibility("default")))

/* "libxm: <func>: <msg>" + exit(1), the error convention of libxm
 * (libxm src/util.c:24-35); `fn` is the public function's name. */
#define XM_FATAL(fn, msg) xm_fatal("%s: %s", (fn), (msg))

/* Any HIP failure is fatal and names the public entry point. */
#define HIP_OK(fn, call)                                                  \
	do {                                                              \
		hipError_t e_ = (call);                                   \
		if (e_ != hipSuccess)                                     \
			xm_fatal("%s: HIP error %s (%s) at %s:%d", (fn),  \
			    hipGetErrorName(e_), hipGetErrorString(e_),   \
			    __FILE__, __LINE__);                          \
	} while (0)

typedef std::complex<double> zdouble;

/* ---- metadata ------------------------------------------------------ */

struct xm_block_space {
	xm_dim_t dims, nblocks;
	std::vector<size_t> splits[XM_MAX_DIM];	/* nblocks+1 ascending */
};

struct xm_block {
	xm_block_type_t type;
	xm_dim_t perm;
	xm_scalar_t scalar;
	uint64_t data;	/* canonical: device address, or a file pointer
			   (filestore.h); derivative: flat index of the
			   source block in the block grid */
};

struct xm_tensor {
	xm_scalar_type_t type;
	xm_block_space_t *bs;
	xm_allocator_t *alloc;
	std::vector<xm_block> blocks;
	uint64_t id;		/* process-unique, never reused */
	uint64_t version;	/* bumped on every structure change */
};

/* ---- HBM arena ----------------------------------------------------- */

struct arena_slab {
	char *base;
	size_t bytes;
	bool host;		/* pinned, device-mapped host memory */
	std::vector<std::pair<size_t, size_t>> free_list; /* offset, size */
};

struct xm_allocator {
	std::string path;
	bool has_path;
	std::mutex mu;
	std::vector<arena_slab> slabs;
	std::vector<std::pair<uint64_t, size_t>> live;	/* addr -> bytes */
	size_t device_bytes, host_bytes, limit_bytes;
	size_t host_limit;	/* pinned-host tier cap */
	/* file tier (allocators with a path): the backing file, its slot
	 * in file pointers, bytes of live file ranges */
	file_store *file;
	uint32_t file_slot;
	size_t file_bytes;
	uint64_t id;
};

/* ---- helpers (host_meta.cpp) -------------------------------------- */

size_t elem_size(xm_scalar_type_t t);
bool is_complex(xm_scalar_type_t t);
xm_block *tensor_block(const xm_tensor_t *t, const xm_dim_t &idx);
size_t block_flat(const xm_tensor_t *t, const xm_dim_t &idx);
xm_dim_t block_dims_flat(const xm_tensor_t *t, size_t flat);
xm_dim_t grid_index(const xm_tensor_t *t, size_t flat);
size_t block_elems(const xm_tensor_t *t, const xm_dim_t &idx);
void bump_version(xm_tensor_t *t);
/* Scalar arithmetic in the element precision (libxm src/scalar.c:45-95). */
xm_scalar_t scalar_mul(xm_scalar_t a, xm_scalar_t b, xm_scalar_type_t t);
xm_scalar_t scalar_add(xm_scalar_t a, xm_scalar_t b, xm_scalar_type_t t);

/* Device address of the data of a (canonical or derivative) block. */
uint64_t block_source_ptr(const xm_tensor_t *t, const xm_dim_t &idx);
/* Invalidate host copies of block data (called before any device write to
 * block storage). */
void bump_data_epoch(void);

/* ---- device context (device.cpp) ----------------------------------- */

hipStream_t xm_stream(void);
/* Stream for host -> HBM staging copies (default priority). */
hipStream_t xm_side_stream(void);
/* Base address of the allocator's pinned host slab holding `ptr`, or 0
 * when `ptr` is not in host memory. */
uint64_t arena_host_slab(xm_allocator_t *a, uint64_t ptr);
static inline bool
arena_is_host(xm_allocator_t *a, uint64_t ptr)
{
	return arena_host_slab(a, ptr) != 0;
}
/* pread / pwrite of bytes at a file pointer (fatal on I/O errors). */
void arena_file_read(const char *fn, uint64_t ptr, void *buf, size_t bytes);
void arena_file_write(const char *fn, uint64_t ptr, const void *buf,
    size_t bytes);
/* Does any canonical block of t live in a file? */
bool tensor_in_file(const xm_tensor_t *t);
/* Pinned, device-mapped host scratch of at least `bytes`, grown on demand
 * (slots as for scratch_buffer; the caller has synchronized the work
 * that used the slot before). */
void *pinned_buffer(int slot, size_t bytes);

/*
 * Pinned copies of the file ranges one batch of kernel jobs reads and
 * writes: ranges only read (rd) and ranges written (wr) lie in separate
 * parts of one pinned, device-mapped buffer, so a job reading a block
 * that another job of the batch writes sees the values from before the
 * batch.  load() reads rd and the loaded wr ranges from the files; the
 * kernels use src() / dst() addresses; store() writes wr back.
 */
struct file_stage {
	stage_layout rd, wr;
	char *buf = nullptr;
	int slot;

	explicit file_stage(int pinned_slot) : slot(pinned_slot) {}
	void need_read(uint64_t p, uint64_t bytes)
	{
		if (file_ptr(p))
			rd.add(p, bytes, true, false);
	}
	void need_write(uint64_t p, uint64_t bytes, bool load)
	{
		if (file_ptr(p))
			wr.add(p, bytes, load, true);
	}
	bool empty() const { return rd.empty() && wr.empty(); }
	void load(const char *fn);
	uint64_t src(uint64_t p) const;
	uint64_t dst(uint64_t p) const;
	void store(const char *fn);
};
/* Bytes a batch of file-staged element jobs aims at: XM_STAGE_BYTES, else
 * 1 GiB. */
size_t file_batch_bytes(void);
/* Reusable device scratch of at least `bytes`, grown on demand; slot
 * selects independent buffers.  Valid until the next call with the same
 * slot. */
void *scratch_buffer(int slot, size_t bytes);
/* Upload a host vector into a device scratch slot (async on xm_stream). */
void *upload(int slot, const void *host, size_t bytes);
double elapsed_ms(hipEvent_t a, hipEvent_t b);
#define STAMP_SLOTS 4
uint64_t *stamp_slots(void);
double stamp_ms(uint64_t a, uint64_t b);
/* kernels_elem.hip: one thread stores the device wall clock in *slot */
void launch_stamp(uint64_t *slot, hipStream_t s);

/* ---- kernel launchers (*.hip) -------------------------------------- */

/* A contiguous run of elements in device memory. */
struct seg_desc {
	uint64_t ptr;
	uint64_t n;		/* elements */
	uint64_t first_chunk;	/* prefix of chunk counts */
	uint64_t aux;		/* per-kernel meaning */
};

/* Elements per workgroup of the fill, gather and dot kernels: a list of
 * segments or jobs is cut into chunks of ELEM_CHUNK, first_chunk being a
 * job's first chunk index. */
#define ELEM_CHUNK 4096

void launch_fill_value(const seg_desc *dsegs, size_t nseg, size_t nchunks,
    xm_scalar_type_t type, xm_scalar_t value, hipStream_t s);
void launch_fill_random(const seg_desc *dsegs, size_t nseg, size_t nchunks,
    xm_scalar_type_t type, uint64_t seed, double lo, double hi, double wlo,
    double whi, hipStream_t s);

/*
 * Element-wise map between two blocks: for every element e of the
 * destination block (canonical layout), the source element is at
 * src + sum_d e_d * sstride[d] (strides already include the source
 * block's permutation).  Operation codes are listed in kernels_elem.hip.
 * A pitched destination (drows != 0) is a drows x (n / drows) matrix
 * with leading dimension dld: element e goes to
 * e % drows + (e / drows) * dld.  drows must be the product of the
 * extents of some leading destination dims.
 */
#define XM_ELEM_MAXDIM 8
struct elem_job {
	uint64_t dst, src;
	uint64_t n;			/* elements of the destination */
	uint64_t first_chunk;
	uint32_t ndim;
	uint32_t src_zero;		/* source absent: treat as 0 */
	uint32_t dims[XM_ELEM_MAXDIM];	/* destination extents */
	uint64_t sstride[XM_ELEM_MAXDIM];
	double f_re, f_im;		/* per-job factor */
	uint64_t drows, dld;		/* destination pitch (0: contiguous) */
};

enum elem_op {
	ELEM_COPY_SCALED = 0,	/* dst = convert(f * src)          (xm_copy) */
	ELEM_AXPY = 1,		/* dst = alpha*dst + f*src         (xm_add)  */
	ELEM_MUL = 2,		/* dst = (f*dst)*src               (xm_mul)  */
	ELEM_DIV = 3,		/* dst = dst/(f*src)               (xm_div)  */
	ELEM_GATHER = 4,	/* dst = src (raw permuted copy)   (unfold)  */
	ELEM_ACCUM = 5,		/* dst = dst + f*src               (adjoint fold) */
};

/* A packed job list: gather-form records first, tiled-form records at
 * byte offset toff; jobs and workgroups of each form. */
struct elem_launch {
	size_t ngather = 0, gchunks = 0;
	size_t ntiled = 0, tchunks = 0;
	size_t toff = 0;
};

elem_launch launch_elem(const elem_job *jobs, size_t njobs, int op,
    xm_scalar_type_t dtype, xm_scalar_type_t stype, xm_scalar_t alpha,
    hipStream_t s);
bool elem_tiling_enabled(void);

/* Adjoint fold of one source block: dst = ((dst + f_0 * src_0) + f_1 *
 * src_1) + ..., each src_k a derivative-block gradient read through its
 * permutation (strides in destination dim order), for up to FOLD_MAXSRC
 * gradients in one pass over the destination. */
#define FOLD_MAXSRC 4
struct fold_job {
	uint64_t dst, n;
	uint32_t ndim, nsrc;
	uint32_t dims[XM_ELEM_MAXDIM];		/* destination extents */
	uint64_t src[FOLD_MAXSRC];
	uint64_t sstride[FOLD_MAXSRC][XM_ELEM_MAXDIM];
	double f_re[FOLD_MAXSRC], f_im[FOLD_MAXSRC];
};
std::vector<char> fold_jobs_pack(const std::vector<fold_job> &jobs,
    size_t *nchunks);
void launch_fold_packed(const void *djobs, size_t njobs, size_t nchunks,
    xm_scalar_type_t type, hipStream_t s);

/* Per-block dot products reduced in a fixed order (xm_dot). */
void launch_dot(const elem_job *djobs, size_t njobs, size_t nchunks,
    xm_scalar_type_t type, double *dpartials, hipStream_t s);

/* ---- grouped GEMM --------------------------------------------------- */

/* Operand panel layout: element (p, k) of a P x K panel lives at
 *   PCONTIG: p + k * ld      KCONTIG: k + p * ld                       */
enum { PANEL_PCONTIG = 0, PANEL_KCONTIG = 1 };

struct gemm_pair {
	uint64_t x, y;		/* X panel (P x K), Y panel (Q x K) */
	uint32_t k;		/* contracted extent */
	uint32_t xmode, ymode;	/* PANEL_*; whole dwords, so kernels read
				 * them with scalar loads */
	uint32_t ldx, ldy, pad;
	double a_re, a_im;	/* merged pair factor s_a*s_b */
};

struct gemm_cblock {
	uint64_t c;		/* C panel: P x Q, element (p,q) at p + q*ldc */
	uint32_t p, q, ldc;
	uint32_t pair_begin, pair_end;
	uint32_t flags;		/* bit0: C panel is a copy (not the block) */
	/* K split (FP64): the tiles of a split C block run as two halves,
	 * the first over the 16-k steps before step split_k of pair
	 * split_pair, the second over the rest; KSPLIT_NONE if not split */
	uint32_t split_pair, split_k;
	uint32_t steps, split_steps;	/* 16-k steps: all, first half's */
};
#define KSPLIT_NONE 0xffffffffu
/* doubles of one half's partial sums (a 64 x 128 tile at most) */
#define KSPLIT_SLOT (64 * 128)

struct gemm_tile {
	uint32_t cblock;
	uint32_t p0, q0;
	uint32_t ksplit;	/* 0: all of K; else 1 + 2 * slot + half */
};

/*
 * A 128 x 128 tile of stacked C blocks (FP64): C blocks whose pairs read
 * the same Y panels with the same extents and factors share the tile's
 * columns, their rows stacked in 32-row tiles.  Row tile r of the tile
 * holds rows p0[r] .. p0[r] + 31 of C block cb[r] (cb[r] == TILE_ROWS_NONE:
 * no rows).  head.cblock is the first member (its pairs give Y, k and the
 * factors), head.q0 the first column; head.p0 and head.ksplit are 0.  A
 * candidate's tile list holds one as three consecutive gemm_tile records.
 */
struct gemm_stile {
	gemm_tile head;
	uint32_t cb[4], p0[4];
};
/* The same for the int8 form's 256 x 256 tiles: eight row tiles, five
 * gemm_tile records. */
struct gemm_stile8 {
	gemm_tile head;
	uint32_t cb[8], p0[8];
};
#define TILE_ROWS_NONE 0xffffffffu

/* Scalars of one contraction call, read by the kernels from device
 * memory so a captured launch sequence can be replayed. */
struct gemm_scalars {
	double alpha_re, alpha_im, beta_re, beta_im;
	uint32_t flags;		/* bit0: alpha == 0, bit1: beta == 0 */
	uint32_t pad;
};

/* Pack element jobs into their device form; launch from packed jobs
 * already in device memory. */
std::vector<char> elem_jobs_pack(const std::vector<elem_job> &jobs,
    size_t src_esz, elem_launch *l);
void launch_elem_packed(const void *djobs, const elem_launch &l, int op,
    xm_scalar_type_t dtype, xm_scalar_type_t stype, xm_scalar_t alpha,
    hipStream_t s);

#endif /* XM_INTERNAL_H */
