/*
 * contract_plan.h - the execution plan of one contraction (see
 * contract.cpp) and the GEMM / split-precision stages that run it.
 */
#ifndef XM_CONTRACT_PLAN_H
#define XM_CONTRACT_PLAN_H

#include "internal.h"

#include <string>
#include <vector>

#define MAX_TILE_CANDIDATES 2

struct masks_t {
	xm_dim_t cidxa, aidxa, cidxb, aidxb, cidxc, aidxc;
};

/* One contracted block combination of one C block. */
struct block_pair {
	size_t fa, fb;		/* flat block indices in a and b */
	xm_scalar_t alpha;	/* s_a * s_b, merged; 0 if dropped */
	bool nonzero;		/* both blocks nonzero before the merge */
};

/* A surviving pair as libxm reports it (libxm order). */
struct plan_pair {
	size_t fc, fa, fb;
	xm_scalar_t alpha;
	uint64_t m, n, k;
};

/* Split-precision stage state (split.hip, crt.hip). */
struct split_state;
struct crt_state;

struct contract_plan {
	xm_scalar_type_t type;
	int precision = XM_PRECISION_NATIVE;
	int nslices = 0;

	/* work */
	double flops = 0, flops_pre = 0, unfold_bytes = 0;
	double panel_bytes = 0, cin_bytes = 0, cout_bytes = 0;
	size_t npairs = 0, npairs_pre = 0, nblocks_c = 0, nunfold = 0;
	size_t nsubstituted = 0;
	std::vector<plan_pair> pair_list;
	/* dense classes run as one C block each, and the C blocks in them */
	size_t nclasses = 0, class_cblocks = 0;

	/* grouped GEMM descriptors (host copies) */
	std::vector<gemm_cblock> cblocks;
	std::vector<bool> cpanel;
	std::vector<gemm_pair> pairs;
	std::vector<bool> x_in_ws, y_in_ws, pair_swap;
	std::vector<gemm_tile> tiles;	/* the active candidate's */
	int tile_cfg = 0;
	int gemm_launches = 0;

	/* tile-shape candidates (gemm.hip autotuner): candidate c launches
	 * cand_tiles[c] (device copy d_cand_tiles[c]) with cand_bp/bq[c] */
	int ncand = 0;
	std::vector<gemm_tile> cand_tiles[MAX_TILE_CANDIDATES];
	gemm_tile *d_cand_tiles[MAX_TILE_CANDIDATES] = {};
	/* K-split tiles (first halves), one per partial-sum slot: the
	 * sum kernel's work list */
	std::vector<gemm_tile> cand_fix[MAX_TILE_CANDIDATES];
	gemm_tile *d_cand_fix[MAX_TILE_CANDIDATES] = {};
	uint32_t cand_bp[MAX_TILE_CANDIDATES] = {}, cand_bq[MAX_TILE_CANDIDATES] = {};
	/* candidate c is the stacked 128 x 128 form: cand_tiles[c] holds
	 * gemm_stile records (three gemm_tile each) */
	bool cand_stacked[MAX_TILE_CANDIDATES] = {};
	bool stacked = false;		/* the active candidate's */
	float cand_ms[MAX_TILE_CANDIDATES] = {};
	int tile_choice = -1;		/* decided candidate, -1 while tuning */
	int tile_next = 0;		/* calls of the tuning sequence so far */
	int tile_active = 0;		/* candidate of the last launch */
	bool f64_dma = false;		/* k_gemm_f64_dma can run the 64 x 128 tiles */
	bool mfma_dw = false;		/* k_gemm_mfma_dw can run the plan */
	bool mfma_x4 = false;		/* ... with 16-byte DMAs */
	bool even_extents = false;	/* every C block's P and Q even */
	bool split_staged = false;	/* SPLIT_BF16 on k_gemm_split2 */
	/* plans of xm_contract_noalloc: one tile shape (the first
	 * candidate, never tuned) and no K split, whose partial sums would
	 * need device memory of their own */
	bool single_tile = false, no_ksplit = false;
	std::string tile_sig;		/* block-size histogram signature */

	/* permute jobs: operand panels, C in, C out */
	std::vector<elem_job> unfold_jobs, cin_jobs, cout_jobs;
	elem_launch unfold_l, cin_l, cout_l;	/* set by the upload */
	std::vector<bool> unfold_src_ws;	/* source is a staged copy */

	/*
	 * Blocks in pinned host memory (arena overflow).  A plan touching any
	 * runs as batches of C blocks (`batches`, each a plan of its own
	 * whose workspace is region[i % 2]); a batch's host operand blocks are
	 * copied into its workspace on the side stream (`stage`, merged
	 * address ranges) while the previous batch computes, and unfolded
	 * from there, so the GEMM reads HBM only.  Host C blocks go through C
	 * panels.
	 */
	struct stage_copy {
		uint64_t host;		/* host address of a merged range */
		size_t bytes;
		uint64_t off;		/* workspace offset of its copy */
	};
	std::vector<stage_copy> stage;
	bool ws_external = false;	/* ws is the parent's region */
	std::vector<contract_plan *> batches;
	void *region[2] = {nullptr, nullptr};
	hipEvent_t ev_start = nullptr, ev_staged[2] = {}, ev_done[2] = {};
	double staged_bytes = 0;
	/*
	 * Blocks in an allocator's file: a batch reads its file operand
	 * ranges (file_ops, `stage` entries then copy from the pinned
	 * copies) and its file C blocks (file_c, read when beta != 0; the
	 * C-in and C-out permutes use the pinned copies in place) into its
	 * region's pinned, device-mapped buffer bounce[i % 2] of the parent,
	 * and writes file_c back once the batch has run.
	 */
	struct file_copy {
		uint64_t fptr;		/* file pointer */
		size_t bytes;
		size_t off;		/* offset in the bounce buffer */
	};
	std::vector<file_copy> file_ops, file_c;
	void *bounce[2] = {nullptr, nullptr};
	double file_bytes = 0;		/* read + written per call */

	/* device memory */
	size_t ws_bytes = 0;
	void *ws = nullptr;		/* panels */
	void *dev = nullptr;		/* descriptors */
	gemm_scalars *d_scalars = nullptr, *host_scalars = nullptr;
	gemm_cblock *d_cblocks = nullptr;
	gemm_pair *d_pairs = nullptr;
	gemm_tile *d_tiles = nullptr;
	gemm_tile *d_fix = nullptr;	/* the active candidate's */
	size_t nfix = 0;
	double *d_ksw = nullptr;	/* K-split partial sums */
	uint32_t *d_queue = nullptr;	/* persistent GEMM's tile counters */
	void *d_unfold = nullptr, *d_cin = nullptr, *d_cout = nullptr;
	hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};

	split_state *split = nullptr;
	crt_state *crt = nullptr;
	const char *gemm_kernel = "";	/* kernel of the last GEMM launch */

	/* the device launch sequence captured as a graph, one per (alpha == 0,
	 * beta == 0) combination; graph_failed after a capture error */
	hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
	bool graph_failed = false;
	bool last_graph = false;	/* the last call replayed a graph */
};

void contract_masks(const char *fn, const xm_tensor_t *a,
    const xm_tensor_t *b, const xm_tensor_t *c, const char *idxa,
    const char *idxb, const char *idxc, masks_t &mk);
size_t enumerate_pairs(const xm_tensor_t *a, const xm_tensor_t *b,
    const xm_dim_t &blkc, const masks_t &mk, std::vector<block_pair> &all,
    std::vector<std::vector<uint32_t>> *members);
int current_precision(void);
bool graph_launch_enabled(void);
int current_nslices(void);
void plan_cache_clear(void);
void plan_cache_drop(uint64_t tensor_id);
void contract_set_stats(const xm_contract_stats_t &s);
bool images_mergeable(const xm_tensor_t *x, const masks_t &mk);
bool contract_gradient(xm_scalar_t alpha, const xm_tensor_t *p,
    const xm_tensor_t *q, xm_tensor_t *dx, const xm_tensor_t *x,
    const char *ip, const char *iq, const char *ix);

/* gemm.hip */
void build_tiles(contract_plan *pl);
void select_tiles(contract_plan *pl);
void launch_gemm(contract_plan *pl, hipStream_t s);
void gemm_tile_record(contract_plan *pl, float gemm_ms);

/* split.hip */
void split_prepare(contract_plan *pl, const xm_tensor_t *a,
    const xm_tensor_t *b, const xm_tensor_t *c, const masks_t &mk);
void split_slices(contract_plan *pl, hipStream_t s);
void split_gemm(contract_plan *pl, hipStream_t s);
void split_release(contract_plan *pl);
/* crt.hip */
bool crt_eligible(const contract_plan *pl);
void crt_prepare(contract_plan *pl, const masks_t &mk, const xm_tensor_t *c);
void crt_quantize(contract_plan *pl, hipStream_t s);
void crt_gemm(contract_plan *pl, hipStream_t s);
void crt_release(contract_plan *pl);
double crt_error_bound(const contract_plan *pl);
int current_nmoduli(void);

/* Whether a split contraction with this many slices runs the LDS-staged
 * kernel (128 x 128 tiles; slices 3-8) rather than the per-wave one (64 x 64
 * tiles; with XM_SPLIT_KERNEL=1). */
bool split_staged_kernel(int nslices);

#endif /* XM_CONTRACT_PLAN_H */
