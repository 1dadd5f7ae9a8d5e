/*
 * backward.cpp - xm_contract_backward, the adjoint of xm_contract
 * (no libxm counterpart).
 *
 * Forward: c = alpha * a * b + beta * c over the canonical blocks of c
 * (libxm src/contract.c:117-225).  Loss: L = sum over the canonical
 * entries of c of dc * c.  Every nonzero block X of a (canonical or
 * derivative) enters the forward as X(e) = s_X * S(permute(e, p_X)), so
 *
 *   G_a(X) = alpha * sum_{canonical C, k} dc * b     (one GEMM per pair),
 *   da[S](permute(e, p_X)) += s_X * G_a(X)(e)        for every X with
 *                                                     source S,
 *
 * where dc' views the canonical blocks of dc (its derivative blocks count
 * as zero).  When every derivative block of a permutes the dims da takes
 * from dc among themselves and those it takes from b among themselves
 * (antisymmetric index pairs do), da is one contraction into da's
 * canonical blocks whose block S takes the terms of every image X of S
 * read in S's index order, identical terms merged (contract_gradient):
 * on an antisymmetric ladder the images' terms are the canonical block's
 * own, so each GEMM runs once with factor 2.  Otherwise G_a is a
 * contraction xm_contract(alpha, dc', b, 0, G_a, idxc, idxb, idxa) into a
 * tensor with the nonzero blocks of a as independent canonical blocks
 * (canonical ones on da's storage, derivative ones on scratch), and the
 * derivative blocks' gradients are folded into their sources, each source
 * block in one pass that reads the gradients of all its derivative blocks
 * (up to FOLD_MAXSRC per launch) through their permutations.
 * db is computed the same way from the contraction of a with dc' (idxa,
 * idxc -> idxb).
 */
#include "contract_plan.h"

#include <algorithm>
#include <chrono>
#include <list>
#include <map>
#include <mutex>

static const char *FN = "xm_contract_backward";

/* A view of dc holding only its canonical blocks (no data copied). */
static xm_tensor_t *
canonical_view(const xm_tensor_t *dc)
{
	xm_tensor_t *v = xm_tensor_create(dc->bs, dc->type, dc->alloc);
	for (size_t f = 0; f < dc->blocks.size(); f++)
		if (dc->blocks[f].type == XM_BLOCK_TYPE_CANONICAL)
			xm_tensor_set_canonical_block_raw(v, grid_index(dc, f),
			    dc->blocks[f].data);
	return v;
}

/* dx fits x: same scalar type and block space, canonical wherever x is. */
static void
check_gradient(const xm_tensor_t *x, const xm_tensor_t *dx)
{
	if (dx->type != x->type)
		XM_FATAL(FN, "gradient tensors must have the tensor's scalar type");
	if (!xm_block_space_eq(dx->bs, x->bs))
		XM_FATAL(FN, "gradient tensors must have the tensor's block space");
	for (size_t f = 0; f < x->blocks.size(); f++)
		if (x->blocks[f].type == XM_BLOCK_TYPE_CANONICAL &&
		    dx->blocks[f].type != XM_BLOCK_TYPE_CANONICAL)
			XM_FATAL(FN, "gradient tensors must have canonical "
			    "blocks where the tensor has them");
}

/* dx canonical exactly where x is (the merged gradient writes every
 * canonical block of dx). */
static bool
same_canonical(const xm_tensor_t *x, const xm_tensor_t *dx)
{
	for (size_t f = 0; f < x->blocks.size(); f++)
		if ((x->blocks[f].type == XM_BLOCK_TYPE_CANONICAL) !=
		    (dx->blocks[f].type == XM_BLOCK_TYPE_CANONICAL))
			return false;
	return true;
}

/* The gradient tensor of x: every nonzero block of x made canonical;
 * canonical blocks share dx's storage, derivative ones get scratch
 * (returned in scratch). */
static xm_tensor_t *
gradient_tensor(const xm_tensor_t *x, xm_tensor_t *dx,
    std::vector<uint64_t> &scratch)
{
	xm_tensor_t *g = xm_tensor_create(x->bs, x->type, x->alloc);
	for (size_t f = 0; f < x->blocks.size(); f++) {
		const xm_block &bx = x->blocks[f];
		if (bx.type == XM_BLOCK_TYPE_ZERO)
			continue;
		xm_dim_t idx = grid_index(x, f);
		if (bx.type == XM_BLOCK_TYPE_CANONICAL) {
			xm_tensor_set_canonical_block_raw(g, idx,
			    dx->blocks[f].data);
		} else {
			xm_tensor_set_canonical_block(g, idx);
			scratch.push_back(g->blocks[f].data);
		}
	}
	return g;
}

/*
 * The auxiliary tensors of one (a, b, dc, da, db) combination: dc's
 * canonical view and the two gradient tensors, kept across calls while
 * none of the five changes structure, so repeated calls find their
 * contraction plans cached (a new tensor would be a new plan key).
 */
struct fold_wave {
	void *dev;			/* packed jobs, device copy */
	size_t njobs, nchunks;
};

struct bwd_aux {
	uint64_t id[5], ver[5];
	xm_tensor_t *dcv = nullptr, *ga = nullptr, *gb = nullptr;
	std::vector<uint64_t> sa, sb;		/* scratch of ga, gb */
	xm_allocator_t *aa = nullptr, *ab = nullptr;
	/* fold launches of ga, gb (built by their first call) */
	std::vector<fold_wave> fa, fb;
	bool fa_built = false, fb_built = false;
};

#define BWD_CACHE_MAX 4
static std::list<bwd_aux> bwd_cache;
static std::mutex bwd_mu;

static void
aux_release(bwd_aux &x)
{
	for (auto *w : {&x.fa, &x.fb})
		for (fold_wave &f : *w)
			if (f.dev)
				HIP_OK(FN, hipFree(f.dev));
	for (uint64_t p : x.sa)
		xm_allocator_deallocate(x.aa, p);
	for (uint64_t p : x.sb)
		xm_allocator_deallocate(x.ab, p);
	/* metadata only (plans on them are dropped); storage released above
	 * or owned by dc, da, db */
	for (xm_tensor_t *t : {x.ga, x.gb, x.dcv})
		if (t)
			xm_tensor_free(t);
}

/* Forget the auxiliary tensors that involve tensor `id` (being freed). */
void
bwd_cache_drop(uint64_t id)
{
	std::list<bwd_aux> gone;
	{
		std::lock_guard<std::mutex> g(bwd_mu);
		for (auto it = bwd_cache.begin(); it != bwd_cache.end();) {
			bool hit = false;
			for (uint64_t k : it->id)
				hit |= k == id;
			if (hit) {
				auto nx = std::next(it);
				gone.splice(gone.end(), bwd_cache, it);
				it = nx;
			} else {
				++it;
			}
		}
	}
	for (auto &x : gone)
		aux_release(x);
}

/* Forget the auxiliary tensors whose scratch lives in allocator `a` (being
 * destroyed: its storage goes with it). */
void
bwd_cache_drop_alloc(xm_allocator_t *a)
{
	std::list<bwd_aux> gone;
	{
		std::lock_guard<std::mutex> g(bwd_mu);
		for (auto it = bwd_cache.begin(); it != bwd_cache.end();) {
			auto nx = std::next(it);
			if (it->aa == a || it->ab == a || (it->dcv &&
			    it->dcv->alloc == a))
				gone.splice(gone.end(), bwd_cache, it);
			it = nx;
		}
	}
	/* the scratch in the allocator being destroyed goes with it; the
	 * other allocator's is released */
	for (auto &x : gone) {
		if (x.aa == a)
			x.sa.clear();
		if (x.ab == a)
			x.sb.clear();
		aux_release(x);
	}
}

static void accumulate(xm_contract_stats_t *acc);
static void fold_images(const xm_tensor_t *x, xm_tensor_t *dx,
    const xm_tensor_t *g, std::vector<fold_wave> &fold, bool &built);

/* One gradient: g from gradient_tensor; fold: its fold launches, built
 * on the first call (built). */
static void
gradient(xm_scalar_t alpha, const xm_tensor_t *x, xm_tensor_t *dx,
    xm_tensor_t *g, const xm_tensor_t *p, const xm_tensor_t *q,
    const char *ip, const char *iq, const char *ix, xm_contract_stats_t *acc,
    std::vector<fold_wave> &fold, bool &built)
{
	xm_contract(alpha, p, q, 0, g, ip, iq, ix);
	accumulate(acc);
	fold_images(x, dx, g, fold, built);
}

/* Add the last contraction's figures to acc. */
static void
accumulate(xm_contract_stats_t *acc)
{
	xm_contract_stats_t st;
	xm_contract_get_stats(&st);
	acc->flops += st.flops;
	acc->flops_pre += st.flops_pre;
	acc->npairs += st.npairs;
	acc->npairs_pre += st.npairs_pre;
	acc->nblocks_c += st.nblocks_c;
	acc->nunfold += st.nunfold;
	acc->unfold_bytes += st.unfold_bytes;
	acc->gemm_ms += st.gemm_ms;
	acc->prep_ms += st.prep_ms;
	acc->post_ms += st.post_ms;
	acc->gemm_launches += st.gemm_launches;
	acc->panel_bytes += st.panel_bytes;
	acc->cin_bytes += st.cin_bytes;
	acc->cout_bytes += st.cout_bytes;
	acc->permute_tiled += st.permute_tiled;
	acc->permute_gather += st.permute_gather;
	acc->ksplit_tiles += st.ksplit_tiles;
	acc->precision = st.precision;
	acc->nslices = st.nslices;
	acc->plan_cached = st.plan_cached;
	acc->err_bound = std::max(acc->err_bound, st.err_bound);
	memcpy(acc->gemm_kernel, st.gemm_kernel, sizeof st.gemm_kernel);
}

/* One wave of fold jobs over blocks of which some lie in a file: batches
 * within file_batch_bytes(), the destinations and gradients of each read
 * into pinned memory and the destinations written back. */
static void
fold_file_wave(const std::vector<fold_job> &wv, size_t esz,
    xm_scalar_type_t type, hipStream_t s)
{
	const uint64_t budget = file_batch_bytes();
	size_t i = 0;
	while (i < wv.size()) {
		file_stage fs(1);
		uint64_t pend = 0;
		size_t e = i;
		for (; e < wv.size(); e++) {
			const fold_job &j = wv[e];
			const uint64_t one = (j.n * esz + 255) / 256 * 256;
			uint64_t cost = file_ptr(j.dst) ? one : 0;
			for (uint32_t m = 0; m < j.nsrc; m++)
				cost += file_ptr(j.src[m]) ? one : 0;
			if (e > i && pend + cost > budget)
				break;
			pend += cost;
			fs.need_write(j.dst, j.n * esz, true);
			for (uint32_t m = 0; m < j.nsrc; m++)
				fs.need_read(j.src[m], j.n * esz);
		}
		fs.load(FN);
		std::vector<fold_job> b(wv.begin() + i, wv.begin() + e);
		for (fold_job &j : b) {
			j.dst = fs.dst(j.dst);
			for (uint32_t m = 0; m < j.nsrc; m++)
				j.src[m] = fs.src(j.src[m]);
		}
		size_t nchunks = 0;
		std::vector<char> pk = fold_jobs_pack(b, &nchunks);
		void *d = upload(5, pk.data(), pk.size());
		launch_fold_packed(d, b.size(), nchunks, type, s);
		HIP_OK(FN, hipStreamSynchronize(s));
		fs.store(FN);
		i = e;
	}
}

/* Fold the gradients of x's derivative blocks (in g) into their sources
 * in dx; the launches are built on the first call (built). */
static void
fold_images(const xm_tensor_t *x, xm_tensor_t *dx, const xm_tensor_t *g,
    std::vector<fold_wave> &fold, bool &built)
{
	/* fold derivative blocks into their sources:
	 * dx[S](permute(e, p)) += s * g[X](e), the X of one source in block
	 * order; one job per source takes up to FOLD_MAXSRC of them in one
	 * pass over dx[S] (launch g takes each source's g-th group) */
	hipStream_t s = xm_stream();
	/* blocks in a file: staged per call, nothing kept */
	const bool file = tensor_in_file(dx) || tensor_in_file(g);
	if (built && !file) {
		for (const fold_wave &f : fold)
			launch_fold_packed(f.dev, f.njobs, f.nchunks, x->type, s);
		HIP_OK(FN, hipStreamSynchronize(s));
		return;
	}
	std::map<size_t, std::vector<fold_job>> groups;	/* source -> jobs */
	for (size_t f = 0; f < x->blocks.size(); f++) {
		const xm_block &bx = x->blocks[f];
		if (bx.type != XM_BLOCK_TYPE_DERIVATIVE)
			continue;
		size_t src = (size_t)bx.data;
		xm_dim_t sd = block_dims_flat(x, src);
		xm_dim_t xd = block_dims_flat(x, f);
		if (xm_dim_dot(&sd) == 0)
			continue;
		std::vector<fold_job> &gs = groups[src];
		if (gs.empty() || gs.back().nsrc == FOLD_MAXSRC) {
			fold_job j;
			memset(&j, 0, sizeof j);
			j.dst = dx->blocks[src].data;
			j.n = xm_dim_dot(&sd);
			j.ndim = (uint32_t)sd.n;
			for (size_t k = 0; k < sd.n; k++)
				j.dims[k] = (uint32_t)sd.i[k];
			gs.push_back(j);
		}
		fold_job &j = gs.back();
		const uint32_t m = j.nsrc++;
		j.src[m] = g->blocks[f].data;
		uint64_t xs[XM_MAX_DIM], st = 1;
		for (size_t k = 0; k < xd.n; k++) {
			xs[k] = st;
			st *= xd.i[k];
		}
		/* source dim perm[k] is logical dim k of X */
		for (size_t k = 0; k < xd.n; k++)
			j.sstride[m][bx.perm.i[k]] = xs[k];
		j.f_re[m] = bx.scalar.real();
		j.f_im[m] = bx.scalar.imag();
	}
	std::vector<std::vector<fold_job>> waves;
	for (auto &e : groups)
		for (size_t w = 0; w < e.second.size(); w++) {
			if (waves.size() <= w)
				waves.resize(w + 1);
			waves[w].push_back(e.second[w]);
		}
	for (auto &wv : waves) {
		if (file) {
			fold_file_wave(wv, elem_size(x->type), x->type, s);
			continue;
		}
		fold_wave fw{nullptr, wv.size(), 0};
		std::vector<char> pk = fold_jobs_pack(wv, &fw.nchunks);
		HIP_OK(FN, hipMalloc(&fw.dev, pk.size()));
		HIP_OK(FN, hipMemcpy(fw.dev, pk.data(), pk.size(),
		    hipMemcpyHostToDevice));
		fold.push_back(fw);
		launch_fold_packed(fw.dev, fw.njobs, fw.nchunks, x->type, s);
	}
	if (!file)
		built = true;
	HIP_OK(FN, hipStreamSynchronize(s));
}

XM_API void
xm_contract_backward(xm_scalar_t alpha, const xm_tensor_t *a,
    const xm_tensor_t *b, const xm_tensor_t *dc, xm_tensor_t *da,
    xm_tensor_t *db, const char *idxa, const char *idxb, const char *idxc)
{
	auto t0 = std::chrono::steady_clock::now();
	masks_t mk;
	/* the forward's own checks (libxm src/contract.c:237-280) */
	contract_masks(FN, a, b, dc, idxa, idxb, idxc, mk);
	if ((da && da->alloc != a->alloc) || (db && db->alloc != b->alloc))
		XM_FATAL(FN, "tensors must use same allocator");
	xm_contract_stats_t acc;
	memset(&acc, 0, sizeof acc);
	const xm_tensor_t *ts[5] = {a, b, dc, da, db};
	bwd_aux cur;
	for (int k = 0; k < 5; k++) {
		cur.id[k] = ts[k] ? ts[k]->id : 0;
		cur.ver[k] = ts[k] ? ts[k]->version : 0;
	}
	bwd_aux *aux = nullptr;
	{
		std::lock_guard<std::mutex> g(bwd_mu);
		for (auto it = bwd_cache.begin(); it != bwd_cache.end(); ++it)
			if (std::equal(cur.id, cur.id + 5, it->id) &&
			    std::equal(cur.ver, cur.ver + 5, it->ver)) {
				bwd_cache.splice(bwd_cache.begin(), bwd_cache, it);
				aux = &bwd_cache.front();
				break;
			}
	}
	if (da)
		check_gradient(a, da);
	if (db)
		check_gradient(b, db);
	if (aux == nullptr) {
		cur.dcv = canonical_view(dc);
		cur.aa = a->alloc;
		cur.ab = b->alloc;
		std::list<bwd_aux> gone;
		{
			std::lock_guard<std::mutex> g(bwd_mu);
			bwd_cache.push_front(std::move(cur));
			aux = &bwd_cache.front();
			while (bwd_cache.size() > BWD_CACHE_MAX)
				gone.splice(gone.end(), bwd_cache,
				    std::prev(bwd_cache.end()));
		}
		for (auto &x : gone)
			aux_release(x);
	}
	/* the merged gradient where the images allow it, else the
	 * per-image gradient tensor and its folds (built on first use) */
	if (da) {
		if (same_canonical(a, da) && contract_gradient(alpha, aux->dcv,
		    b, da, a, idxc, idxb, idxa)) {
			accumulate(&acc);
		} else {
			if (aux->ga == nullptr)
				aux->ga = gradient_tensor(a, da, aux->sa);
			gradient(alpha, a, da, aux->ga, aux->dcv, b, idxc, idxb,
			    idxa, &acc, aux->fa, aux->fa_built);
		}
	}
	if (db) {
		if (same_canonical(b, db) && contract_gradient(alpha, a,
		    aux->dcv, db, b, idxa, idxc, idxb)) {
			accumulate(&acc);
		} else {
			if (aux->gb == nullptr)
				aux->gb = gradient_tensor(b, db, aux->sb);
			gradient(alpha, b, db, aux->gb, a, aux->dcv, idxa, idxc,
			    idxb, &acc, aux->fb, aux->fb_built);
		}
	}
	auto t1 = std::chrono::steady_clock::now();
	acc.total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
	contract_set_stats(acc);
}
