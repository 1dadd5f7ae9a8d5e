/*
 * xm_amd.h - entry points of the MI355X build that have no counterpart in
 * libxm's C API.  They sit beside the unchanged xm.h prototypes, so code
 * written against libxm is unaffected by them.
 */
#ifndef XM_AMD_H_INCLUDED
#define XM_AMD_H_INCLUDED

#include <stdint.h>
#include "tensor.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Arithmetic used by xm_contract / xm_contract_backward for DOUBLE and
 * DOUBLE_COMPLEX tensors.  No libxm counterpart: libxm always calls the
 * BLAS routine of the element type (libxm src/contract.c:49-84).
 *  NATIVE      - FP64 on the FP64 matrix cores (FP32 for FLOAT types).
 *  SPLIT_BF16  - FP64 emulated by splitting each operand row/column into
 *                BF16 slices that share one FP64 exponent, multiplying the
 *                slices on BF16 matrix cores with FP32 accumulation over
 *                K-chunks short enough to stay exact, and recombining in
 *                FP64.  Exponents, scaling, recombination, alpha/beta and
 *                the C update stay FP64.
 */
typedef enum {
	XM_PRECISION_NATIVE = 0,
	XM_PRECISION_SPLIT_BF16 = 1,
	/*
	 * SPLIT_INT8 - FP64 emulated through the Chinese remainder theorem:
	 *              each operand row is scaled by a power of two (from its
	 *              2-norm) and rounded to an integer, the integer product
	 *              is computed modulo n coprime moduli <= 256 on the int8
	 *              matrix cores (exact int32 sums over the whole K), and
	 *              rebuilt in FP64.  Relative error ~0.6 sqrt(K)
	 *              sqrt(4 W / M) whatever the spread of the entries
	 *              (M the product of the moduli, W the largest merged pair
	 *              factor ratio).  Needs the merged pair factors of every C
	 *              block to be integer multiples of each other (libxm's
	 *              +-1 / +-2 from antisymmetric blocks are); other
	 *              contractions run SPLIT_BF16 (the stats say which).
	 */
	XM_PRECISION_SPLIT_INT8 = 2,
} xm_precision_t;

void xm_set_contract_precision(xm_precision_t precision);
xm_precision_t xm_get_contract_precision(void);
/* Number of BF16 slices per FP64 operand in SPLIT_BF16 mode (3..8;
 * default 6). */
void xm_set_split_slices(int nslices);
int xm_get_split_slices(void);
/* Number of moduli in SPLIT_INT8 mode: 8..13, or 0 (the default) for the
 * fewest whose modelled error is below 1e-13 for the plan's longest K. */
void xm_set_split_moduli(int nmoduli);
int xm_get_split_moduli(void);

/*
 * Graph launch (no libxm counterpart).  Once a contraction plan's GEMM tile
 * shape is decided, its device work is captured as a hipGraph and each
 * later call with that plan is one graph launch; alpha and beta are read
 * from device memory written before the launch.  On by default; 0 turns it
 * off (plans then launch their kernels one by one).  The environment
 * variable XM_GRAPH=0 sets the default to off.  Results are identical.
 */
void xm_set_graph_launch(int on);
int xm_get_graph_launch(void);

/*
 * Figures of the most recent xm_contract / xm_contract_backward call.
 * No libxm counterpart (libxm has no timers or counters).
 * flops counts 2*m*n*k real flops (x4 for complex types) over the block
 * pairs that survive libxm's pair merge with a nonzero factor, i.e. the
 * GEMM work libxm itself executes (libxm src/contract.c:161-218).
 */
typedef struct {
	double flops;		/* after the merge */
	double flops_pre;	/* before the merge */
	size_t npairs;		/* merged pairs with nonzero factor */
	size_t npairs_pre;	/* pairs with two nonzero blocks, before merge */
	size_t nblocks_c;	/* canonical C blocks written */
	size_t nunfold;		/* operand panels built by the permute kernel */
	double unfold_bytes;	/* bytes read + written by permute kernels */
	double gemm_ms;		/* device time of the grouped GEMM launches */
	double prep_ms;		/* device time of unfold / slice launches */
	double post_ms;		/* device time of fold / scale launches */
	double total_ms;	/* host wall time of the call */
	int plan_cached;	/* 1 if the execution plan was reused */
	int precision;		/* xm_precision_t used */
	int nslices;		/* slices per operand (SPLIT_BF16 only) */
	int gemm_launches;	/* number of grouped GEMM launches */
	/* unfold_bytes by stage: operand panels and C-in permutes run inside
	 * prep_ms (the C-in permute only when beta != 0), C-out inside post_ms */
	double panel_bytes;	/* operand panels, read + written */
	double cin_bytes;	/* permuted C in, read + written */
	double cout_bytes;	/* permuted C out, read + written */
	/* output tile of the grouped GEMM launch, and 1 while the plan's tile
	 * shape was still being tuned (the call timed one candidate) */
	int tile_p, tile_q, tile_tuning;
	int graph;		/* 1 if the call replayed the plan's hipGraph */
	/* permute jobs of the call (operand panels, C in, C out) moved by
	 * the LDS-tiled kernel, and by the gather kernel */
	size_t permute_tiled, permute_gather;
	/* contractions touching blocks in pinned host memory (arena overflow)
	 * run in batches of C blocks, each batch's host operand blocks copied
	 * into HBM on a side stream while the previous batch computes: the
	 * batch count (0: nothing staged) and the bytes copied per call */
	int staged_batches;
	double staged_bytes;
	/* FP64 tiles of the GEMM launch run as two halves of their K steps
	 * (C blocks at the end of the launch order), summed by a second
	 * kernel */
	size_t ksplit_tiles;
	/* SPLIT_INT8: the modelled relative error of the result (0 in the
	 * other forms; nslices holds the moduli count) */
	double err_bound;
	/* the grouped GEMM kernel the call ran */
	char gemm_kernel[32];
	/* bytes of blocks in an allocator's file read (pread into pinned
	 * buffers, then staged like host blocks) and written back per call */
	double file_bytes;
	/* dense classes: C blocks whose terms form one GEMM of stacked
	 * blocks (rows R x columns S over the same contracted blocks) run as
	 * one C block each; the count of classes and of C blocks in them */
	size_t classes, class_cblocks;
} xm_contract_stats_t;

void xm_contract_get_stats(xm_contract_stats_t *stats);

/* Jobs (one per canonical block written) of the last xm_add, xm_copy,
 * xm_mul or xm_div moved by the LDS-tiled permute kernel and by the
 * gather kernel.  No libxm counterpart. */
void xm_amd_last_permute(size_t *tiled, size_t *gather);

/*
 * The block pairs xm_contract multiplies for each canonical block of c, in
 * libxm's order (canonical C blocks first-index-fastest, then contracted
 * block combinations first-index-fastest), after libxm's pair merge and
 * with zero-factor pairs dropped (libxm src/contract.c:161-218).
 * m, n, k are the logical GEMM sizes: m over c's indices taken from a, n
 * over c's indices taken from b, k over the contracted indices.
 * Writes at most maxpairs entries; returns the total number of pairs.
 * No libxm counterpart (the list is internal to libxm's compute_block).
 */
typedef struct {
	xm_dim_t blkidxc, blkidxa, blkidxb;
	xm_scalar_t alpha;	/* merged s_a * s_b, in the tensor's precision */
	size_t m, n, k;
} xm_contract_pair_t;

size_t xm_contract_get_pairs(const xm_tensor_t *a, const xm_tensor_t *b,
    const xm_tensor_t *c, const char *idxa, const char *idxb,
    const char *idxc, xm_contract_pair_t *pairs, size_t maxpairs);

/*
 * Adjoint of xm_contract.  No libxm counterpart.
 * For the forward c = alpha * a * b + beta * c and the loss
 * L = sum over the CANONICAL entries of c of dc * c, writes
 *   da = dL/d(canonical data of a),   db = dL/d(canonical data of b).
 * Derivative blocks of dc count as zero.  Every block X of a that is used
 * by the forward contributes through its source canonical block:
 *   da[src(X)](permute(e, p_X)) += s_X * G_X(e),  G = alpha * dc (x) b,
 * and likewise for b.  da / db must have the block structure of a / b
 * (canonical blocks at the same places); either may be NULL.
 */
void xm_contract_backward(xm_scalar_t alpha, const xm_tensor_t *a,
    const xm_tensor_t *b, const xm_tensor_t *dc, xm_tensor_t *da,
    xm_tensor_t *db, const char *idxa, const char *idxb, const char *idxc);

/*
 * xm_contract without any device allocation.  No libxm counterpart: it is
 * for callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, checks and result as
 * xm_contract in native precision (whatever xm_set_contract_precision
 * says), with one tile shape and no K split, so results match
 * xm_contract's to FP64 rounding (FP64: within 1e-13 relative).  Every
 * device the library runs on holds a 64 MiB array for the call's permuted
 * operand and C panels and its descriptors until the process ends; calls
 * on one device share it, so calls that overlap on one device (from
 * several host threads) race.  Returns 0, or -1 after a message on stderr,
 * computing nothing, when the contraction needs more than the array or
 * touches blocks in pinned host memory (those need staging buffers).
 */
int xm_contract_noalloc(xm_scalar_t alpha, const xm_tensor_t *a,
    const xm_tensor_t *b, xm_scalar_t beta, xm_tensor_t *c, const char *idxa,
    const char *idxb, const char *idxc);

/*
 * Fill every canonical block of t with reproducible pseudo-random values
 * generated on the device: x = u * 10^w, u ~ U[lo, hi), w ~ U[wlo, whi)
 * (wlo == whi == 0 gives plain uniform values).  Real and imaginary parts
 * are drawn independently.  Values depend only on (seed, block index,
 * element index).  No libxm counterpart (workload generation).
 */
void xm_amd_fill_random(xm_tensor_t *t, uint64_t seed, double lo, double hi,
    double wlo, double whi);

/* Bytes of device (HBM) and pinned-host memory held by an allocator's
 * arena.  No libxm counterpart. */
void xm_amd_allocator_usage(xm_allocator_t *allocator, size_t *device_bytes,
    size_t *host_bytes);

/* Bytes of live allocations in the file of an allocator created with a
 * path (0 without one).  No libxm counterpart. */
size_t xm_amd_allocator_file_bytes(xm_allocator_t *allocator);

/* Block until all work issued by the library has completed; returns 0.
 * No libxm counterpart (every xm_* operation is already synchronous). */
int xm_amd_synchronize(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* XM_AMD_H_INCLUDED */
