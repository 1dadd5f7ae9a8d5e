/*
 * contract.cpp - xm_contract on the GPU (libxm src/contract.c:86-314).
 *
 * Host: the same checks and fatal messages as libxm, the same pair
 * enumeration and merge for every canonical C block, then a plan:
 *   - every surviving pair becomes a GEMM term  C(p,q) += al * X(p,k) Y(q,k)
 *     of a grouped GEMM whose K dimension concatenates all pairs of a C
 *     block (one owner per output tile, no atomics);
 *   - operands whose block layout already is a P x K (or K x P) matrix are
 *     read in place; the others are permuted once per contraction into a
 *     panel by the permute kernel and shared by every pair that uses them;
 *   - C blocks whose layout is not a matrix are permuted into a panel and
 *     back.
 * Plans are cached per (tensor structures, index strings, precision) and
 * hold their descriptor lists in device memory; alpha and beta are read by
 * the kernels from device memory, so a plan is reused across calls.
 */
#include "internal.h"
#include "contract_plan.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <list>
#include <map>
#include <set>
#include <unordered_map>

/* ---- precision selection ----------------------------------------- */

static int g_precision = XM_PRECISION_NATIVE;
/* 6 slices keep entries spread over six decades below 1e-12 (5 land at
 * 1.1e-12 on the o=64 v=512 ladder, 1e-13 on uniform entries) */
static int g_nslices = 6;

XM_API void
xm_set_contract_precision(xm_precision_t p)
{
	if (p != XM_PRECISION_NATIVE && p != XM_PRECISION_SPLIT_BF16 &&
	    p != XM_PRECISION_SPLIT_INT8)
		XM_FATAL("xm_set_contract_precision", "unknown precision");
	g_precision = p;
}

XM_API xm_precision_t
xm_get_contract_precision(void)
{
	return (xm_precision_t)g_precision;
}

/* graph launch of decided plans: on unless XM_GRAPH=0 or switched off */
static int g_graph = -1;

bool
graph_launch_enabled(void)
{
	if (g_graph < 0) {
		const char *e = getenv("XM_GRAPH");
		g_graph = !(e && *e == '0');
	}
	return g_graph != 0;
}

XM_API void
xm_set_graph_launch(int on)
{
	g_graph = on != 0;
}

XM_API int
xm_get_graph_launch(void)
{
	return graph_launch_enabled();
}

XM_API void
xm_set_split_slices(int n)
{
	if (n < 3 || n > 8)
		XM_FATAL("xm_set_split_slices", "slices must be in 3..8");
	g_nslices = n;
}

XM_API int
xm_get_split_slices(void)
{
	return g_nslices;
}

int current_precision(void) { return g_precision; }
int current_nslices(void) { return g_nslices; }

/* ---- checks and masks (libxm src/contract.c:237-280) -------------- */

void
contract_masks(const char *fn, const xm_tensor_t *a, const xm_tensor_t *b,
    const xm_tensor_t *c, const char *idxa, const char *idxb,
    const char *idxc, masks_t &mk)
{
	if (a->alloc != c->alloc || b->alloc != c->alloc)
		XM_FATAL(fn, "tensors must use same allocator");
	if (a->type != c->type || b->type != c->type)
		XM_FATAL(fn, "tensors must have same scalar type");
	if (strlen(idxa) != a->bs->dims.n || strlen(idxb) != b->bs->dims.n ||
	    strlen(idxc) != c->bs->dims.n)
		XM_FATAL(fn, "bad contraction indices");
	xm_make_masks(idxa, idxb, &mk.cidxa, &mk.cidxb);
	xm_make_masks(idxc, idxa, &mk.cidxc, &mk.aidxa);
	xm_make_masks(idxc, idxb, &mk.aidxc, &mk.aidxb);
	if (mk.aidxa.n + mk.cidxa.n != a->bs->dims.n ||
	    mk.aidxb.n + mk.cidxb.n != b->bs->dims.n ||
	    mk.aidxc.n + mk.cidxc.n != c->bs->dims.n)
		XM_FATAL(fn, "bad contraction indices");
	if (!(mk.aidxc.n > 0 && mk.aidxc.i[0] == 0) &&
	    !(mk.cidxc.n > 0 && mk.cidxc.i[0] == 0))
		XM_FATAL(fn, "bad contraction indices");
	for (size_t k = 0; k < mk.cidxa.n; k++)
		if (!xm_block_space_eq1(a->bs, mk.cidxa.i[k], b->bs,
		    mk.cidxb.i[k]))
			XM_FATAL(fn, "inconsistent a and b tensor block-spaces");
	for (size_t k = 0; k < mk.cidxc.n; k++)
		if (!xm_block_space_eq1(c->bs, mk.cidxc.i[k], a->bs,
		    mk.aidxa.i[k]))
			XM_FATAL(fn, "inconsistent a and c tensor block-spaces");
	for (size_t k = 0; k < mk.aidxc.n; k++)
		if (!xm_block_space_eq1(c->bs, mk.aidxc.i[k], b->bs,
		    mk.aidxb.i[k]))
			XM_FATAL(fn, "inconsistent b and c tensor block-spaces");
}

/* ---- pair enumeration and merge (libxm src/contract.c:86-192) ----- */

/* libxm src/contract.c:86-115: same source data for a and for b, and the
 * same permutation at the free positions of a and of b. */
static bool
same_contraction(const xm_tensor_t *a, const xm_tensor_t *b,
    const block_pair &x, const block_pair &y, const masks_t &mk)
{
	const xm_block &xa = a->blocks[x.fa], &ya = a->blocks[y.fa];
	const xm_block &xb = b->blocks[x.fb], &yb = b->blocks[y.fb];
	uint64_t pxa = xa.type == XM_BLOCK_TYPE_DERIVATIVE ?
	    a->blocks[xa.data].data : xa.data;
	uint64_t pya = ya.type == XM_BLOCK_TYPE_DERIVATIVE ?
	    a->blocks[ya.data].data : ya.data;
	uint64_t pxb = xb.type == XM_BLOCK_TYPE_DERIVATIVE ?
	    b->blocks[xb.data].data : xb.data;
	uint64_t pyb = yb.type == XM_BLOCK_TYPE_DERIVATIVE ?
	    b->blocks[yb.data].data : yb.data;
	if (pxa != pya || pxb != pyb)
		return false;
	for (size_t k = 0; k < mk.aidxa.n; k++)
		if (xa.perm.i[mk.aidxa.i[k]] != ya.perm.i[mk.aidxa.i[k]])
			return false;
	for (size_t k = 0; k < mk.aidxb.n; k++)
		if (xb.perm.i[mk.aidxb.i[k]] != yb.perm.i[mk.aidxb.i[k]])
			return false;
	return true;
}

/* libxm src/contract.c:161-192 for one C block.  `all` receives every
 * contracted combination (alpha 0 for zero blocks and merged-away pairs),
 * `members[i]` the combinations merged into survivor i.  Returns the
 * number of combinations with two nonzero blocks. */
size_t
enumerate_pairs(const xm_tensor_t *a, const xm_tensor_t *b,
    const xm_dim_t &blkc, const masks_t &mk, std::vector<block_pair> &all,
    std::vector<std::vector<uint32_t>> *members)
{
	xm_dim_t nba = a->bs->nblocks, nbb = b->bs->nblocks;
	size_t n = xm_dim_dot_mask(&nba, &mk.cidxa), npre = 0;
	xm_dim_t ia = xm_dim_zero(nba.n), ib = xm_dim_zero(nbb.n);

	xm_dim_set_mask(&ia, &mk.aidxa, &blkc, &mk.cidxc);
	xm_dim_set_mask(&ib, &mk.aidxb, &blkc, &mk.aidxc);
	all.resize(n);
	for (size_t i = 0; i < n; i++) {
		block_pair &p = all[i];
		p.fa = xm_dim_offset(&ia, &nba);
		p.fb = xm_dim_offset(&ib, &nbb);
		p.alpha = 0;
		const xm_block &ba = a->blocks[p.fa], &bb = b->blocks[p.fb];
		if (ba.type != XM_BLOCK_TYPE_ZERO &&
		    bb.type != XM_BLOCK_TYPE_ZERO) {
			p.alpha = scalar_mul(ba.scalar, bb.scalar, a->type);
			npre++;
		}
		p.nonzero = ba.type != XM_BLOCK_TYPE_ZERO &&
		    bb.type != XM_BLOCK_TYPE_ZERO;
		xm_dim_inc_mask(&ia, &nba, &mk.cidxa);
		xm_dim_inc_mask(&ib, &nbb, &mk.cidxb);
	}
	if (members) {
		members->assign(n, {});
	}
	for (size_t i = 0; i < n; i++) {
		if (all[i].alpha == xm_scalar_t(0))
			continue;
		for (size_t j = i + 1; j < n; j++) {
			if (all[j].alpha == xm_scalar_t(0))
				continue;
			if (same_contraction(a, b, all[i], all[j], mk)) {
				all[i].alpha = scalar_add(all[i].alpha,
				    all[j].alpha, a->type);
				all[j].alpha = 0;
				if (members)
					(*members)[i].push_back((uint32_t)j);
			}
		}
	}
	return npre;
}

/* ---- operand layout analysis -------------------------------------- */

/* Source stride of every logical index of a block (the permutation of a
 * derivative block folded in): X(e) = S(permute(e, p)). */
static void
logical_strides(const xm_tensor_t *t, size_t flat, uint64_t *ls,
    xm_dim_t *dims)
{
	const xm_block &b = t->blocks[flat];
	xm_dim_t d = block_dims_flat(t, flat);
	xm_dim_t sd = xm_dim_permute(&d, &b.perm);
	uint64_t ss[XM_MAX_DIM], s = 1;
	for (size_t k = 0; k < sd.n; k++) {
		ss[k] = s;
		s *= sd.i[k];
	}
	for (size_t k = 0; k < d.n; k++)
		ls[k] = ss[b.perm.i[k]];
	*dims = d;
}

/* Are the dims `order` (in order) laid out contiguously, the first one
 * with stride `start`?  Extent-1 dims are ignored. */
static bool
contiguous_run(const uint64_t *ls, const xm_dim_t &d, const xm_dim_t &order,
    uint64_t start, uint64_t *end)
{
	uint64_t want = start;
	for (size_t k = 0; k < order.n; k++) {
		size_t q = order.i[k];
		if (d.i[q] == 1)
			continue;
		if (ls[q] != want)
			return false;
		want *= d.i[q];
	}
	*end = want;
	return true;
}

/* Layout of a block as a (P x K) operand with free dims `F` (ordered as
 * the matching C dims) and contracted dims `Kd`.  Returns true when the
 * block can be read in place and sets mode/ld. */
static bool
direct_layout(const uint64_t *ls, const xm_dim_t &d, const xm_dim_t &F,
    const xm_dim_t &Kd, uint8_t *mode, uint32_t *ld)
{
	uint64_t P = xm_dim_dot_mask(&d, &F), K = xm_dim_dot_mask(&d, &Kd), e;
	if (contiguous_run(ls, d, F, 1, &e) && contiguous_run(ls, d, Kd, P, &e)) {
		*mode = PANEL_PCONTIG;
		*ld = (uint32_t)std::max<uint64_t>(P, 1);
		return true;
	}
	if (contiguous_run(ls, d, Kd, 1, &e) && contiguous_run(ls, d, F, K, &e)) {
		*mode = PANEL_KCONTIG;
		*ld = (uint32_t)std::max<uint64_t>(K, 1);
		return true;
	}
	return false;
}

/* sigma test for substituting a merged member r for survivor i: the
 * contracted index t' of r maps to the source dim perm_r[cidx[t']], which
 * must be the source dim of the same contracted index of i for both
 * operands, so that both sums run over the same products. */
static bool
same_k_order(const xm_block &xi, const xm_block &xr, const xm_dim_t &cidx,
    int *sigma)
{
	for (size_t t = 0; t < cidx.n; t++) {
		int found = -1;
		for (size_t u = 0; u < cidx.n; u++)
			if (xr.perm.i[cidx.i[t]] == xi.perm.i[cidx.i[u]])
				found = (int)u;
		if (found < 0)
			return false;
		sigma[t] = found;
	}
	return true;
}

static bool
substitutable(const xm_tensor_t *a, const xm_tensor_t *b, const masks_t &mk,
    const block_pair &i, const block_pair &r)
{
	int sa[XM_MAX_DIM], sb[XM_MAX_DIM];
	if (!same_k_order(a->blocks[i.fa], a->blocks[r.fa], mk.cidxa, sa) ||
	    !same_k_order(b->blocks[i.fb], b->blocks[r.fb], mk.cidxb, sb))
		return false;
	for (size_t t = 0; t < mk.cidxa.n; t++)
		if (sa[t] != sb[t])
			return false;
	return true;
}

/* ---- plan building ------------------------------------------------- */

struct panel_key {
	uint64_t src;
	uint32_t n;
	uint64_t ext[XM_MAX_DIM], str[XM_MAX_DIM];
	bool operator<(const panel_key &o) const {
		if (src != o.src) return src < o.src;
		if (n != o.n) return n < o.n;
		for (uint32_t k = 0; k < n; k++) {
			if (ext[k] != o.ext[k]) return ext[k] < o.ext[k];
			if (str[k] != o.str[k]) return str[k] < o.str[k];
		}
		return false;
	}
};

/* Whether a plan's operand panels are laid out P-contiguous, for the
 * DMA-staged GEMM kernels that read only those: FP32 and complex plans,
 * and native FP64 plans whose C blocks all have even extents (the FP64
 * kernel's row-pair DMAs need them; with an odd one the plan runs the
 * register-staged kernel, which reads K-contiguous panels as well and ran
 * 3-4% faster on them).  XM_PANEL_PCONTIG=0: only where the source's
 * stride-1 index is free. */
static bool
pcontig_panels(const contract_plan *pl)
{
	static int on = -1;
	if (on < 0) {
		const char *e = getenv("XM_PANEL_PCONTIG");
		on = !(e && *e == '0');
	}
	/* the CRT form's residue kernels read P-contiguous panels only */
	if (pl->precision == XM_PRECISION_SPLIT_INT8)
		return true;
	return on && pl->precision == XM_PRECISION_NATIVE &&
	    (pl->type != XM_SCALAR_DOUBLE || pl->even_extents);
}

struct plan_builder {
	contract_plan *pl;
	size_t esz;
	std::map<panel_key, uint64_t> panels;	/* key -> workspace offset */
	/* staged plans: host-resident source blocks (address -> bytes, slab) */
	bool stage = false;
	xm_allocator_t *alloc = nullptr;
	std::map<uint64_t, std::pair<size_t, uint64_t>> host_blocks;

	/* Operand of one pair: in place, or a (shared) panel. */
	void operand(const xm_tensor_t *t, size_t flat, const xm_dim_t &F,
	    const xm_dim_t &Kd, uint64_t *ptr, uint8_t *mode, uint32_t *ld,
	    bool *in_ws)
	{
		uint64_t ls[XM_MAX_DIM];
		xm_dim_t d;
		logical_strides(t, flat, ls, &d);
		const xm_block &b = t->blocks[flat];
		uint64_t src = b.type == XM_BLOCK_TYPE_DERIVATIVE ?
		    t->blocks[b.data].data : b.data;
		*in_ws = false;
		/* a host-resident block is never read in place: its staged
		 * copy is unfolded into a panel */
		const uint64_t slab = stage ? arena_host_slab(alloc, src) : 0;
		const bool host = slab != 0;
		/* plans on the staged MFMA kernels (native FP64, FP32,
		 * complex) want P-contiguous panels: those kernels read only
		 * them */
		const bool pfirst = pcontig_panels(pl);
		if (!host && direct_layout(ls, d, F, Kd, mode, ld) &&
		    !(pfirst && *mode == PANEL_KCONTIG)) {
			*ptr = src;
			return;
		}
		/* Permuted panel; P-contiguous unless the source's stride-1
		 * index is a contracted one and the GEMM reads either layout
		 * (the permute then reads and writes in runs). */
		bool kfirst = false;
		for (size_t k = 0; k < Kd.n && !pfirst; k++)
			if (ls[Kd.i[k]] == 1 && d.i[Kd.i[k]] > 1)
				kfirst = true;
		panel_key key;
		key.src = src;
		key.n = 0;
		const xm_dim_t &o1 = kfirst ? Kd : F, &o2 = kfirst ? F : Kd;
		for (size_t k = 0; k < o1.n; k++) {
			key.ext[key.n] = d.i[o1.i[k]];
			key.str[key.n++] = ls[o1.i[k]];
		}
		for (size_t k = 0; k < o2.n; k++) {
			key.ext[key.n] = d.i[o2.i[k]];
			key.str[key.n++] = ls[o2.i[k]];
		}
		uint64_t P = xm_dim_dot_mask(&d, &F), K = xm_dim_dot_mask(&d, &Kd);
		*mode = kfirst ? PANEL_KCONTIG : PANEL_PCONTIG;
		*ld = (uint32_t)std::max<uint64_t>(kfirst ? K : P, 1);
		auto it = panels.find(key);
		if (it != panels.end()) {
			*ptr = it->second;
			*in_ws = true;
			return;
		}
		uint64_t off = pl->ws_bytes;
		pl->ws_bytes += (P * K * esz + 255) / 256 * 256;
		panels[key] = off;
		elem_job j;
		memset(&j, 0, sizeof j);
		j.dst = off;		/* relocated once the workspace exists */
		j.src = src;
		j.n = P * K;
		j.ndim = key.n;
		for (uint32_t k = 0; k < key.n; k++) {
			j.dims[k] = (uint32_t)key.ext[k];
			j.sstride[k] = key.str[k];
		}
		pl->unfold_jobs.push_back(j);
		pl->unfold_src_ws.push_back(host);
		if (host)
			host_blocks[src] = {xm_dim_dot(&d) * esz, slab};
		pl->unfold_bytes += 2.0 * P * K * esz;
		pl->panel_bytes += 2.0 * P * K * esz;
		*ptr = off;
		*in_ws = true;
	}
};

/* Host ranges closer than this are staged as one copy. */
#define STAGE_MERGE_GAP (1u << 20)

/*
 * The plan of canonical C blocks [lo, hi) (libxm's order).  stage: blocks
 * in pinned host memory are staged (see contract_plan) rather than read in
 * place.
 */
/*
 * Derivative blocks of the tensor `images` (block space of c) by source
 * block: the images of every canonical block of c in the gradient plans
 * of xm_contract_backward.
 */
struct image_map {
	const xm_tensor_t *x;
	std::map<size_t, std::vector<size_t>> of;	/* source -> images */
};

/* The free dims of an operand read for image X of a C block: C position
 * pos[j'] takes X's dim pinv[pos[j']], which is pos[j] for some j (an
 * image permutes the C block's dims taken from a among themselves, and
 * those taken from b); the operand dim matched to it is od[j]. */
static bool
image_free_dims(const xm_dim_t &perm, const xm_dim_t &pos,
    const xm_dim_t &od, xm_dim_t *out)
{
	xm_dim_t pinv = perm;
	for (size_t k = 0; k < perm.n; k++)
		pinv.i[perm.i[k]] = k;
	*out = od;
	for (size_t jp = 0; jp < pos.n; jp++) {
		const size_t want = pinv.i[pos.i[jp]];
		size_t j = pos.n;
		for (size_t t = 0; t < pos.n; t++)
			if (pos.i[t] == want)
				j = t;
		if (j == pos.n)
			return false;
		out->i[jp] = od.i[j];
	}
	return true;
}

/* Whether every derivative block of x permutes c's dims from a among
 * themselves and those from b among themselves (x has c's block space). */
bool
images_mergeable(const xm_tensor_t *x, const masks_t &mk)
{
	for (const xm_block &bx : x->blocks) {
		if (bx.type != XM_BLOCK_TYPE_DERIVATIVE)
			continue;
		xm_dim_t f;
		if (!image_free_dims(bx.perm, mk.cidxc, mk.aidxa, &f) ||
		    !image_free_dims(bx.perm, mk.aidxc, mk.aidxb, &f))
			return false;
	}
	return true;
}

/*
 * The GEMM terms of canonical block S (flat fc) of c in a gradient plan:
 * the pairs of S and of every image X = s_X S(permute(., p_X)) of S in
 * the tensor the gradient belongs to, X's pairs read in S's index order
 * (operand free dims reordered by p_X) and scaled by s_X; terms reading
 * the same data in the same order sum the same products and are merged
 * (factors added in the element precision, zero sums dropped), as libxm
 * merges pairs within one C block (libxm src/contract.c:86-115).
 */
template <typename Emit>
static void
image_terms(const xm_tensor_t *a, const xm_tensor_t *b, const xm_tensor_t *c,
    const image_map *images, size_t fc, const masks_t &mk, contract_plan *pl,
    size_t mult, uint64_t m, uint64_t n, Emit &emit)
{
	struct term {
		size_t fa, fb;
		xm_dim_t Fa, Fb;
		xm_scalar_t f;
		uint64_t k;
	};
	std::vector<term> terms;
	std::map<std::vector<uint64_t>, size_t> index;
	std::vector<std::pair<size_t, xm_block>> ims;
	xm_block self;
	memset(&self, 0, sizeof self);
	self.perm = xm_dim_identity_permutation(c->bs->dims.n);
	self.scalar = 1;
	ims.push_back({fc, self});
	auto it = images->of.find(fc);
	if (it != images->of.end())
		for (size_t fx : it->second)
			ims.push_back({fx, images->x->blocks[fx]});
	std::vector<block_pair> all;
	for (auto &im : ims) {
		xm_dim_t Fa, Fb;
		image_free_dims(im.second.perm, mk.cidxc, mk.aidxa, &Fa);
		image_free_dims(im.second.perm, mk.aidxc, mk.aidxb, &Fb);
		const xm_dim_t blkx = grid_index(c, im.first);
		pl->npairs_pre += enumerate_pairs(a, b, blkx, mk, all, nullptr);
		for (const block_pair &bp : all) {
			if (!bp.nonzero)
				continue;
			xm_dim_t da = block_dims_flat(a, bp.fa);
			const uint64_t k = xm_dim_dot_mask(&da, &mk.cidxa);
			pl->flops_pre += 2.0 * mult * m * n * k;
			if (bp.alpha == xm_scalar_t(0))
				continue;
			/* the data and the order each operand is read in */
			std::vector<uint64_t> key;
			for (int side = 0; side < 2; side++) {
				const xm_tensor_t *t = side ? b : a;
				const size_t f = side ? bp.fb : bp.fa;
				const xm_dim_t &F = side ? Fb : Fa;
				const xm_dim_t &K = side ? mk.cidxb : mk.cidxa;
				uint64_t ls[XM_MAX_DIM];
				xm_dim_t d;
				logical_strides(t, f, ls, &d);
				key.push_back(block_source_ptr(t, grid_index(t, f)));
				for (size_t q = 0; q < F.n; q++) {
					key.push_back(d.i[F.i[q]]);
					key.push_back(ls[F.i[q]]);
				}
				for (size_t q = 0; q < K.n; q++) {
					key.push_back(d.i[K.i[q]]);
					key.push_back(ls[K.i[q]]);
				}
			}
			const xm_scalar_t f = scalar_mul(im.second.scalar, bp.alpha,
			    c->type);
			auto ti = index.find(key);
			if (ti == index.end()) {
				index[key] = terms.size();
				terms.push_back({bp.fa, bp.fb, Fa, Fb, f, k});
			} else {
				term &t = terms[ti->second];
				t.f = scalar_add(t.f, f, c->type);
			}
		}
	}
	for (const term &t : terms) {
		if (t.f == xm_scalar_t(0))
			continue;
		pl->flops += 2.0 * mult * m * n * t.k;
		pl->npairs++;
		pl->pair_list.push_back({fc, t.fa, t.fb, t.f, m, n, t.k});
		emit(t.fa, t.fb, t.Fa, t.Fb, t.f, t.k);
	}
}

/* One GEMM term of a C block: block fa of a, block fb of b, merged factor
 * f, contracted extent k. */
struct cterm {
	size_t fa, fb;
	xm_scalar_t f;
	uint64_t k;
};

/* A canonical C block of a plan: extents m (from a) and n (from b),
 * orientation (swap: P = n and X comes from b), whether it is computed in
 * a C panel, and its terms in libxm's order. */
struct cinfo {
	size_t fc;
	uint64_t m, n;
	bool swap, cpanel;
	std::vector<cterm> terms;
};

/*
 * Dense classes.  C blocks (r, s) whose terms read X blocks X(r, t) and
 * Y blocks Y(t, s), t = 0 .. n-1, with the same extents k_t and one
 * factor, for every r of a row set R and s of a column set S, are one GEMM
 * C(R, S) = X(R, T) Y(T, S) of stacked blocks.  (The ring term
 * c(i,j,a,b) = sum t(i,k,a,c) w(k,b,c,j) over tensors that store every
 * symmetry-allowed block is eight such GEMMs, one per irrep of (i,a).)
 * A class runs as one C block with one pair: its X blocks are permuted
 * into one P-contiguous panel of sum_R P rows at their row and column
 * offsets, its Y blocks likewise, C is computed in one panel folded back
 * to its blocks.  Each block's K columns are padded with zeros to a
 * multiple of 4: every 4-k group the GEMM kernels hand to an MFMA then
 * lies within one term, as when each term is a pair of its own, so every
 * output sums the same products in the same groups and order, and the
 * results are the same bits (a zero's sign aside; the K split of the
 * drain, where a plan has one, cuts a class's C block at its own middle
 * step).  A class is formed
 * when the per-C-block form permutes all of its operand blocks and
 * computes its C blocks in panels too (so it adds no permute traffic) and
 * its padded MFMA work (rows in 32s, columns in 128s) is at least 8% below
 * the per-C-block form's (rows in 32s, columns in 64s, each pair's k in
 * 16s).  XM_CLASS_GEMM=0 forms none, =1 every complete one.
 */
struct dense_class {
	std::vector<uint32_t> members;		/* cinfo indices */
	std::vector<uint32_t> rows, cols;	/* a member of each row, column */
	std::vector<uint32_t> mrow, mcol;	/* each member's row, column */
};

static int
class_mode(void)
{
	const char *e = getenv("XM_CLASS_GEMM");
	if (e && *e == '0')
		return 0;
	return e && *e == '1' ? 2 : 1;
}

static uint64_t
round_up(uint64_t x, uint64_t m)
{
	return (x + m - 1) / m * m;
}

/* Whether the per-C-block form reads block `flat` of t in place as an
 * operand with free dims F and contracted dims Kd. */
static bool
read_in_place(const contract_plan *pl, const xm_tensor_t *t, size_t flat,
    const xm_dim_t &F, const xm_dim_t &Kd)
{
	uint64_t ls[XM_MAX_DIM];
	xm_dim_t d;
	uint8_t mode;
	uint32_t ld;
	logical_strides(t, flat, ls, &d);
	return direct_layout(ls, d, F, Kd, &mode, &ld) &&
	    !(pcontig_panels(pl) && mode == PANEL_KCONTIG);
}

static std::vector<dense_class>
find_classes(const xm_tensor_t *a, const xm_tensor_t *b, const masks_t &mk,
    const contract_plan *pl, const std::vector<cinfo> &ci)
{
	std::vector<dense_class> out;
	const int mode = class_mode();
	const size_t N = ci.size();
	if (mode == 0 || N == 0)
		return out;
	/* each C block's X and Y sequences (with swap and the k's) */
	std::map<std::vector<uint64_t>, int64_t> xid, yid;
	std::vector<int64_t> cx(N, -1), cy(N, -1);
	for (size_t i = 0; i < N; i++) {
		const cinfo &q = ci[i];
		if (q.terms.empty())
			continue;
		bool uniform = true;
		std::vector<uint64_t> kx = {q.swap}, ky = {q.swap};
		for (const cterm &t : q.terms) {
			uniform = uniform && t.f == q.terms[0].f;
			kx.push_back(q.swap ? t.fb : t.fa);
			kx.push_back(t.k);
			ky.push_back(q.swap ? t.fa : t.fb);
			ky.push_back(t.k);
		}
		if (!uniform)
			continue;
		cx[i] = xid.emplace(kx, (int64_t)xid.size()).first->second;
		cy[i] = yid.emplace(ky, (int64_t)yid.size()).first->second;
	}
	/* components of the graph joining C blocks that share an X or a Y
	 * sequence (union-find) */
	std::vector<uint32_t> par(N);
	for (size_t i = 0; i < N; i++)
		par[i] = (uint32_t)i;
	auto root = [&](uint32_t i) {
		while (par[i] != i)
			i = par[i] = par[par[i]];
		return i;
	};
	std::vector<int64_t> fx(xid.size(), -1), fy(yid.size(), -1);
	for (size_t i = 0; i < N; i++) {
		if (cx[i] < 0)
			continue;
		for (int side = 0; side < 2; side++) {
			int64_t &f = side ? fy[cy[i]] : fx[cx[i]];
			if (f < 0)
				f = (int64_t)i;
			else
				par[root((uint32_t)i)] = root((uint32_t)f);
		}
	}
	std::map<uint32_t, size_t> cid;
	std::vector<std::vector<uint32_t>> comps;
	for (size_t i = 0; i < N; i++) {
		if (cx[i] < 0)
			continue;
		auto it = cid.emplace(root((uint32_t)i), comps.size()).first;
		if (it->second == comps.size())
			comps.emplace_back();
		comps[it->second].push_back((uint32_t)i);
	}
	const size_t esz = elem_size(pl->type);
	for (auto &mem : comps) {
		dense_class dc;
		std::map<int64_t, uint32_t> rmap, smap;
		std::set<std::pair<uint32_t, uint32_t>> seen;
		bool ok = true;
		for (uint32_t m : mem) {
			auto r = rmap.emplace(cx[m], (uint32_t)rmap.size());
			if (r.second)
				dc.rows.push_back(m);
			auto s = smap.emplace(cy[m], (uint32_t)smap.size());
			if (s.second)
				dc.cols.push_back(m);
			dc.mrow.push_back(r.first->second);
			dc.mcol.push_back(s.first->second);
			ok = ok && seen.insert({r.first->second,
			    s.first->second}).second &&
			    ci[m].terms[0].f == ci[mem[0]].terms[0].f;
		}
		if (!ok || mem.size() != dc.rows.size() * dc.cols.size())
			continue;
		dc.members = mem;
		if (mode == 1) {
			/* padded MFMA work per C block against the class's, and
			 * the bytes the class permutes beyond the per-block
			 * form */
			const cinfo &q0 = ci[mem[0]];
			const bool sw = q0.swap;
			const xm_tensor_t *xt = sw ? b : a, *yt = sw ? a : b;
			const xm_dim_t &XF = sw ? mk.aidxb : mk.aidxa;
			const xm_dim_t &XK = sw ? mk.cidxb : mk.cidxa;
			const xm_dim_t &YF = sw ? mk.aidxa : mk.aidxb;
			const xm_dim_t &YK = sw ? mk.cidxa : mk.cidxb;
			uint64_t K4 = 0, K16 = 0, P = 0, Q = 0;
			for (const cterm &t : q0.terms) {
				K4 += round_up(t.k, 4);
				K16 += round_up(t.k, 16);
			}
			double eb = 0, extra = 0;
			for (uint32_t m : mem) {
				const cinfo &q = ci[m];
				const uint64_t p = sw ? q.n : q.m, qq = sw ? q.m : q.n;
				eb += 2.0 * round_up(p, 32) * round_up(qq, 64) * K16;
				if (!q.cpanel)
					extra += 4.0 * p * qq * esz;
			}
			for (uint32_t r : dc.rows) {
				const cinfo &q = ci[r];
				P += sw ? q.n : q.m;
				for (const cterm &t : q.terms)
					if (read_in_place(pl, xt, sw ? t.fb : t.fa,
					    XF, XK))
						extra += 2.0 * (sw ? q.n : q.m) *
						    t.k * esz;
			}
			for (uint32_t s : dc.cols) {
				const cinfo &q = ci[s];
				Q += sw ? q.m : q.n;
				for (const cterm &t : q.terms)
					if (read_in_place(pl, yt, sw ? t.fa : t.fb,
					    YF, YK))
						extra += 2.0 * (sw ? q.m : q.n) *
						    t.k * esz;
			}
			const double ec = 2.0 * round_up(P, 32) *
			    round_up(Q, 128) * K4;
			if (extra > 0 || ec > 0.92 * eb)
				continue;
		}
		out.push_back(std::move(dc));
	}
	return out;
}

/* The permute job writing block `flat` of t at dst as a P-contiguous
 * (P x K) panel: free dims F, then contracted dims Kd. */
static elem_job
panel_job(const xm_tensor_t *t, size_t flat, const xm_dim_t &F,
    const xm_dim_t &Kd, uint64_t dst)
{
	uint64_t ls[XM_MAX_DIM];
	xm_dim_t d;
	logical_strides(t, flat, ls, &d);
	const xm_block &bl = t->blocks[flat];
	elem_job j;
	memset(&j, 0, sizeof j);
	j.dst = dst;
	j.src = bl.type == XM_BLOCK_TYPE_DERIVATIVE ?
	    t->blocks[bl.data].data : bl.data;
	j.n = xm_dim_dot_mask(&d, &F) * xm_dim_dot_mask(&d, &Kd);
	for (size_t k = 0; k < F.n; k++) {
		j.dims[j.ndim] = (uint32_t)d.i[F.i[k]];
		j.sstride[j.ndim++] = ls[F.i[k]];
	}
	for (size_t k = 0; k < Kd.n; k++) {
		j.dims[j.ndim] = (uint32_t)d.i[Kd.i[k]];
		j.sstride[j.ndim++] = ls[Kd.i[k]];
	}
	return j;
}

/* A job writing zeros to a rows x cols window of a panel with leading
 * dimension ld. */
static elem_job
zero_job(uint64_t dst, uint64_t rows, uint64_t cols, uint64_t ld)
{
	elem_job j;
	memset(&j, 0, sizeof j);
	j.dst = dst;
	j.src_zero = 1;
	j.n = rows * cols;
	j.ndim = 2;
	j.dims[0] = (uint32_t)rows;
	j.dims[1] = (uint32_t)cols;
	j.drows = rows;
	j.dld = ld;
	return j;
}

static contract_plan *
build_plan(const xm_tensor_t *a, const xm_tensor_t *b, const xm_tensor_t *c,
    const masks_t &mk, int precision, size_t lo = 0, size_t hi = SIZE_MAX,
    bool stage = false, const image_map *images = nullptr)
{
	contract_plan *pl = new contract_plan;
	plan_builder pb;
	pb.pl = pl;
	pb.esz = elem_size(c->type);
	pb.stage = stage;
	pb.alloc = c->alloc;
	pl->type = c->type;
	pl->precision = precision;
	size_t mult = is_complex(c->type) ? 4 : 1;

	xm_dim_t *list;
	size_t nlist;
	xm_tensor_get_canonical_block_list(c, &list, &nlist);
	hi = std::min(hi, nlist);
	lo = std::min(lo, hi);
	pl->nblocks_c = hi - lo;
	pl->even_extents = true;
	for (size_t ic = lo; ic < hi; ic++) {
		xm_dim_t dc = block_dims_flat(c, block_flat(c, list[ic]));
		if (xm_dim_dot_mask(&dc, &mk.cidxc) % 2 ||
		    xm_dim_dot_mask(&dc, &mk.aidxc) % 2)
			pl->even_extents = false;
	}

	/* the C blocks: orientation and terms */
	std::vector<cinfo> ci;
	std::vector<block_pair> all;
	std::vector<std::vector<uint32_t>> members;
	for (size_t ic = lo; ic < hi; ic++) {
		cinfo q;
		q.fc = block_flat(c, list[ic]);
		xm_dim_t dc = block_dims_flat(c, q.fc);
		q.m = xm_dim_dot_mask(&dc, &mk.cidxc);
		q.n = xm_dim_dot_mask(&dc, &mk.aidxc);

		/* orientation of the C block */
		uint64_t lsc[XM_MAX_DIM];
		xm_dim_t dd;
		logical_strides(c, q.fc, lsc, &dd);
		uint64_t e;
		q.cpanel = false;	/* swap: P = n (X = B), Q = m (Y = A) */
		if (contiguous_run(lsc, dc, mk.cidxc, 1, &e) &&
		    contiguous_run(lsc, dc, mk.aidxc,
		    std::max<uint64_t>(q.m, 1), &e))
			q.swap = false;
		else if (contiguous_run(lsc, dc, mk.aidxc, 1, &e) &&
		    contiguous_run(lsc, dc, mk.cidxc,
		    std::max<uint64_t>(q.n, 1), &e))
			q.swap = true;
		else {
			q.swap = false;
			q.cpanel = true;
		}
		if (stage && arena_is_host(c->alloc, c->blocks[q.fc].data)) {
			/* a host C block is computed in an HBM panel */
			q.swap = false;
			q.cpanel = true;
		}
		if (images != nullptr) {
			ci.push_back(std::move(q));
			continue;
		}
		size_t npre = enumerate_pairs(a, b, list[ic], mk, all, &members);
		pl->npairs_pre += npre;
		for (size_t i = 0; i < all.size(); i++) {
			if (!all[i].nonzero)
				continue;
			xm_dim_t da = block_dims_flat(a, all[i].fa);
			uint64_t k = xm_dim_dot_mask(&da, &mk.cidxa);
			pl->flops_pre += 2.0 * mult * q.m * q.n * k;
			if (all[i].alpha == xm_scalar_t(0))
				continue;
			pl->flops += 2.0 * mult * q.m * q.n * k;
			pl->npairs++;
			pl->pair_list.push_back({q.fc, all[i].fa, all[i].fb,
			    all[i].alpha, q.m, q.n, k});

			/* pick the cheapest equivalent member of the merge
			 * group (fewest operands needing a permute) */
			block_pair best = all[i];
			int best_cost = 3;
			std::vector<uint32_t> cand = {(uint32_t)i};
			for (uint32_t r : members[i])
				cand.push_back(r);
			for (uint32_t r : cand) {
				if (r != i && !substitutable(a, b, mk, all[i], all[r]))
					continue;
				uint64_t ls[XM_MAX_DIM];
				xm_dim_t d;
				uint8_t md;
				uint32_t ld;
				int cost = 0;
				logical_strides(a, all[r].fa, ls, &d);
				cost += !direct_layout(ls, d, mk.aidxa, mk.cidxa, &md, &ld);
				logical_strides(b, all[r].fb, ls, &d);
				cost += !direct_layout(ls, d, mk.aidxb, mk.cidxb, &md, &ld);
				if (cost < best_cost) {
					best_cost = cost;
					best = all[r];
				}
			}
			if (best.fa != all[i].fa || best.fb != all[i].fb)
				pl->nsubstituted++;
			q.terms.push_back({best.fa, best.fb, all[i].alpha, k});
		}
		ci.push_back(std::move(q));
	}
	free(list);

	/* dense classes (native precision, blocks in HBM) */
	std::vector<dense_class> classes;
	if (images == nullptr && !stage && precision == XM_PRECISION_NATIVE)
		classes = find_classes(a, b, mk, pl, ci);
	std::vector<int64_t> class_of(ci.size(), -1);
	for (size_t k = 0; k < classes.size(); k++)
		for (uint32_t m : classes[k].members)
			class_of[m] = (int64_t)k;

	const size_t esz = pb.esz;
	auto ws_take = [&](uint64_t bytes) {
		uint64_t off = pl->ws_bytes;
		pl->ws_bytes += (bytes + 255) / 256 * 256;
		return off;
	};
	/* C panel jobs of C block q: in (C block -> panel at off, pitch ld)
	 * and out (panel -> C block) */
	auto cpanel_jobs = [&](const cinfo &q, uint64_t off, uint64_t ld) {
		xm_dim_t dc = block_dims_flat(c, q.fc);
		uint64_t lsc[XM_MAX_DIM];
		xm_dim_t dd;
		logical_strides(c, q.fc, lsc, &dd);
		const uint64_t cptr = c->blocks[q.fc].data;
		elem_job in, out;
		memset(&in, 0, sizeof in);
		memset(&out, 0, sizeof out);
		const uint64_t m = q.m, n = q.n;
		if (q.cpanel) {
			/* rows = cidxc dims, columns = aidxc dims */
			in.dst = off;
			in.src = cptr;
			in.n = m * n;
			for (size_t k = 0; k < mk.cidxc.n; k++) {
				in.dims[in.ndim] = (uint32_t)dc.i[mk.cidxc.i[k]];
				in.sstride[in.ndim++] = lsc[mk.cidxc.i[k]];
			}
			for (size_t k = 0; k < mk.aidxc.n; k++) {
				in.dims[in.ndim] = (uint32_t)dc.i[mk.aidxc.i[k]];
				in.sstride[in.ndim++] = lsc[mk.aidxc.i[k]];
			}
			in.drows = m;
			in.dld = ld;
			/* fold: C block element (canonical order) <- panel */
			out.dst = cptr;
			out.src = off;
			out.n = m * n;
			out.ndim = (uint32_t)dc.n;
			uint64_t pstride[XM_MAX_DIM], s = 1;
			for (size_t k = 0; k < mk.cidxc.n; k++) {
				pstride[mk.cidxc.i[k]] = s;
				s *= dc.i[mk.cidxc.i[k]];
			}
			s = ld;
			for (size_t k = 0; k < mk.aidxc.n; k++) {
				pstride[mk.aidxc.i[k]] = s;
				s *= dc.i[mk.aidxc.i[k]];
			}
			for (size_t k = 0; k < dc.n; k++) {
				out.dims[k] = (uint32_t)dc.i[k];
				out.sstride[k] = pstride[k];
			}
		} else {
			/* a C block laid out as its P x Q panel (a class
			 * member): copied in and out of the class panel */
			const uint64_t P = q.swap ? n : m, Q = q.swap ? m : n;
			in.dst = off;
			in.src = cptr;
			in.n = P * Q;
			in.ndim = 2;
			in.dims[0] = (uint32_t)P;
			in.dims[1] = (uint32_t)Q;
			in.sstride[0] = 1;
			in.sstride[1] = P;
			in.drows = P;
			in.dld = ld;
			out.dst = cptr;
			out.src = off;
			out.n = P * Q;
			out.ndim = 2;
			out.dims[0] = (uint32_t)P;
			out.dims[1] = (uint32_t)Q;
			out.sstride[0] = 1;
			out.sstride[1] = ld;
		}
		pl->cin_jobs.push_back(in);
		pl->cout_jobs.push_back(out);
		pl->unfold_bytes += 4.0 * m * n * esz;
		pl->cin_bytes += 2.0 * m * n * esz;
		pl->cout_bytes += 2.0 * m * n * esz;
	};

	/* one class: panels, permutes, one C block with one pair */
	auto emit_class = [&](const dense_class &dc) {
		const cinfo &q0 = ci[dc.members[0]];
		const bool sw = q0.swap;
		const xm_tensor_t *xt = sw ? b : a, *yt = sw ? a : b;
		const xm_dim_t &XF = sw ? mk.aidxb : mk.aidxa;
		const xm_dim_t &XK = sw ? mk.cidxb : mk.cidxa;
		const xm_dim_t &YF = sw ? mk.aidxa : mk.aidxb;
		const xm_dim_t &YK = sw ? mk.cidxa : mk.cidxb;
		auto rows_of = [&](const cinfo &q) { return q.swap ? q.n : q.m; };
		auto cols_of = [&](const cinfo &q) { return q.swap ? q.m : q.n; };
		std::vector<uint64_t> roff, coff, koff;
		uint64_t P = 0, Q = 0, K = 0;
		for (uint32_t r : dc.rows) {
			roff.push_back(P);
			P += rows_of(ci[r]);
		}
		for (uint32_t s : dc.cols) {
			coff.push_back(Q);
			Q += cols_of(ci[s]);
		}
		for (const cterm &t : q0.terms) {
			koff.push_back(K);
			K += round_up(t.k, 4);
		}
		const uint64_t ldx = round_up(P, 4), ldy = round_up(Q, 4);
		if (ldx > 0xffffffffull || ldy > 0xffffffffull ||
		    K > 0xffffffffull)
			XM_FATAL("xm_contract", "class panel too large");
		const uint64_t ox = ws_take(ldx * K * esz);
		const uint64_t oy = ws_take(ldy * K * esz);
		const uint64_t oc = ws_take(ldx * Q * esz);
		for (size_t r = 0; r < dc.rows.size(); r++) {
			const cinfo &q = ci[dc.rows[r]];
			for (size_t t = 0; t < q.terms.size(); t++) {
				elem_job j = panel_job(xt, sw ? q.terms[t].fb :
				    q.terms[t].fa, XF, XK,
				    ox + (roff[r] + koff[t] * ldx) * esz);
				j.drows = rows_of(q);
				j.dld = ldx;
				pl->unfold_jobs.push_back(j);
				pl->unfold_src_ws.push_back(false);
				pl->unfold_bytes += 2.0 * j.n * esz;
				pl->panel_bytes += 2.0 * j.n * esz;
			}
		}
		for (size_t s = 0; s < dc.cols.size(); s++) {
			const cinfo &q = ci[dc.cols[s]];
			for (size_t t = 0; t < q.terms.size(); t++) {
				elem_job j = panel_job(yt, sw ? q.terms[t].fa :
				    q.terms[t].fb, YF, YK,
				    oy + (coff[s] + koff[t] * ldy) * esz);
				j.drows = cols_of(q);
				j.dld = ldy;
				pl->unfold_jobs.push_back(j);
				pl->unfold_src_ws.push_back(false);
				pl->unfold_bytes += 2.0 * j.n * esz;
				pl->panel_bytes += 2.0 * j.n * esz;
			}
		}
		/* zeros: each term's K padding and the rows past P (Q) */
		std::vector<elem_job> zj;
		for (size_t t = 0; t < q0.terms.size(); t++) {
			const uint64_t k = q0.terms[t].k, pad = round_up(k, 4) - k;
			if (!pad)
				continue;
			zj.push_back(zero_job(ox + (koff[t] + k) * ldx * esz,
			    ldx, pad, ldx));
			zj.push_back(zero_job(oy + (koff[t] + k) * ldy * esz,
			    ldy, pad, ldy));
		}
		if (ldx > P)
			zj.push_back(zero_job(ox + P * esz, ldx - P, K, ldx));
		if (ldy > Q)
			zj.push_back(zero_job(oy + Q * esz, ldy - Q, K, ldy));
		for (const elem_job &j : zj) {
			pl->unfold_jobs.push_back(j);
			pl->unfold_src_ws.push_back(false);
			pl->unfold_bytes += 1.0 * j.n * esz;
			pl->panel_bytes += 1.0 * j.n * esz;
		}
		for (size_t i = 0; i < dc.members.size(); i++)
			cpanel_jobs(ci[dc.members[i]], oc + (roff[dc.mrow[i]] +
			    coff[dc.mcol[i]] * ldx) * esz, ldx);

		gemm_cblock cb;
		memset(&cb, 0, sizeof cb);
		cb.split_pair = KSPLIT_NONE;
		cb.c = oc;
		cb.p = (uint32_t)P;
		cb.q = (uint32_t)Q;
		cb.ldc = (uint32_t)ldx;
		cb.pair_begin = (uint32_t)pl->pairs.size();
		cb.pair_end = cb.pair_begin + 1;
		cb.flags = 1;
		gemm_pair gp;
		memset(&gp, 0, sizeof gp);
		gp.x = ox;
		gp.y = oy;
		gp.k = (uint32_t)K;
		gp.xmode = gp.ymode = PANEL_PCONTIG;
		gp.ldx = (uint32_t)ldx;
		gp.ldy = (uint32_t)ldy;
		gp.a_re = q0.terms[0].f.real();
		gp.a_im = q0.terms[0].f.imag();
		pl->pairs.push_back(gp);
		pl->x_in_ws.push_back(true);
		pl->y_in_ws.push_back(true);
		pl->pair_swap.push_back(false);
		pl->cblocks.push_back(cb);
		pl->cpanel.push_back(true);
		pl->nclasses++;
		pl->class_cblocks += dc.members.size();
	};

	for (size_t i = 0; i < ci.size(); i++) {
		const cinfo &q = ci[i];
		if (class_of[i] >= 0) {
			const dense_class &dc = classes[class_of[i]];
			if (dc.members[0] == i)
				emit_class(dc);
			continue;
		}
		const bool swap = q.swap;
		gemm_cblock cb;
		memset(&cb, 0, sizeof cb);
		cb.split_pair = KSPLIT_NONE;
		cb.p = (uint32_t)(swap ? q.n : q.m);
		cb.q = (uint32_t)(swap ? q.m : q.n);
		cb.ldc = std::max<uint32_t>(cb.p, 1);
		cb.pair_begin = (uint32_t)pl->pairs.size();
		if (q.cpanel) {
			uint64_t off = ws_take(q.m * q.n * esz);
			cpanel_jobs(q, off, std::max<uint64_t>(q.m, 1));
			cb.c = off;
			cb.flags |= 1;
		} else {
			cb.c = c->blocks[q.fc].data;
		}

		/* one GEMM term of this C block: operands fa (free dims Fa,
		 * in the order of the C block's dims taken from a) and fb */
		auto emit = [&](size_t fa, size_t fb, const xm_dim_t &Fa,
		    const xm_dim_t &Fb, xm_scalar_t f, uint64_t k) {
			gemm_pair gp;
			memset(&gp, 0, sizeof gp);
			uint64_t pa, pbp;
			uint8_t ma, mb;
			uint32_t lda, ldb;
			bool wa, wb;
			pb.operand(a, fa, Fa, mk.cidxa, &pa, &ma, &lda, &wa);
			pb.operand(b, fb, Fb, mk.cidxb, &pbp, &mb, &ldb, &wb);
			gp.k = (uint32_t)k;
			gp.a_re = f.real();
			gp.a_im = f.imag();
			if (!swap) {
				gp.x = pa; gp.xmode = ma; gp.ldx = lda;
				gp.y = pbp; gp.ymode = mb; gp.ldy = ldb;
				pl->x_in_ws.push_back(wa);
				pl->y_in_ws.push_back(wb);
			} else {
				gp.x = pbp; gp.xmode = mb; gp.ldx = ldb;
				gp.y = pa; gp.ymode = ma; gp.ldy = lda;
				pl->x_in_ws.push_back(wb);
				pl->y_in_ws.push_back(wa);
			}
			pl->pairs.push_back(gp);
			/* split path: which operand is A (kept for the
			 * slicing stage) */
			pl->pair_swap.push_back(swap);
		};
		if (images != nullptr)
			image_terms(a, b, c, images, q.fc, mk, pl, mult, q.m, q.n,
			    emit);
		else
			for (const cterm &t : q.terms)
				emit(t.fa, t.fb, mk.aidxa, mk.aidxb, t.f, t.k);
		cb.pair_end = (uint32_t)pl->pairs.size();
		pl->cblocks.push_back(cb);
		pl->cpanel.push_back(q.cpanel);
	}
	pl->nunfold = pl->unfold_jobs.size();
	if (!pb.host_blocks.empty()) {
		/* merge host blocks of one slab into ranges, one staged copy
		 * each, laid out after the panels; point their unfold jobs at
		 * the copies */
		std::vector<std::pair<uint64_t, uint64_t>> rng;	/* [start, end) */
		uint64_t last_slab = 0;
		for (auto &hb : pb.host_blocks) {
			uint64_t e = hb.first + hb.second.first;
			if (!rng.empty() && hb.second.second == last_slab &&
			    hb.first <= rng.back().second + STAGE_MERGE_GAP)
				rng.back().second = std::max(rng.back().second, e);
			else
				rng.push_back({hb.first, e});
			last_slab = hb.second.second;
		}
		for (auto &r : rng) {
			uint64_t off = pl->ws_bytes;
			pl->ws_bytes += (r.second - r.first + 255) / 256 * 256;
			pl->stage.push_back({r.first, r.second - r.first, off});
			pl->staged_bytes += (double)(r.second - r.first);
		}
		for (size_t i = 0; i < pl->unfold_jobs.size(); i++) {
			if (!pl->unfold_src_ws[i])
				continue;
			elem_job &j = pl->unfold_jobs[i];
			auto it = std::upper_bound(pl->stage.begin(),
			    pl->stage.end(), j.src, [](uint64_t v,
			    const contract_plan::stage_copy &s) {
				return v < s.host;
			});
			--it;
			j.src = it->off + (j.src - it->host);
		}
	}
	return pl;
}

/* ---- plan cache ---------------------------------------------------- */

struct plan_key {
	uint64_t ida, va, idb, vb, idc, vc;
	std::string ia, ib, ic;
	int precision, nslices;
	uint64_t idx = 0, vx = 0;	/* gradient plans: the images' tensor */
	bool operator<(const plan_key &o) const {
		return std::tie(ida, va, idb, vb, idc, vc, ia, ib, ic, precision,
		    nslices, idx, vx) < std::tie(o.ida, o.va, o.idb, o.vb, o.idc,
		    o.vc, o.ia, o.ib, o.ic, o.precision, o.nslices, o.idx, o.vx);
	}
};

static std::mutex cache_mu;
static std::list<std::pair<plan_key, contract_plan *>> cache;
static const size_t CACHE_MAX = 8;

void
plan_release(contract_plan *pl)
{
	if (pl == nullptr)
		return;
	HIP_OK("xm_contract", hipStreamSynchronize(xm_stream()));
	HIP_OK("xm_contract", hipStreamSynchronize(xm_side_stream()));
	for (contract_plan *sub : pl->batches)
		plan_release(sub);
	for (void *r : pl->region)
		if (r)
			HIP_OK("xm_contract", hipFree(r));
	for (void *r : pl->bounce)
		if (r)
			HIP_OK("xm_contract", hipHostFree(r));
	for (hipEvent_t e : {pl->ev_start, pl->ev_staged[0], pl->ev_staged[1],
	    pl->ev_done[0], pl->ev_done[1]})
		if (e)
			HIP_OK("xm_contract", hipEventDestroy(e));
	if (pl->dev)
		HIP_OK("xm_contract", hipFree(pl->dev));
	if (pl->d_ksw)
		HIP_OK("xm_contract", hipFree(pl->d_ksw));
	if (pl->d_queue)
		HIP_OK("xm_contract", hipFree(pl->d_queue));
	if (pl->ws && !pl->ws_external)
		HIP_OK("xm_contract", hipFree(pl->ws));
	if (pl->host_scalars)
		HIP_OK("xm_contract", hipHostFree(pl->host_scalars));
	for (auto &g : pl->graph)
		if (g)
			HIP_OK("xm_contract", hipGraphExecDestroy(g));
	if (pl->ev[0])
		for (auto &e : pl->ev)
			HIP_OK("xm_contract", hipEventDestroy(e));
	split_release(pl);
	crt_release(pl);
	delete pl;
}

/* Turn the plan's workspace offsets into addresses: the workspace starts
 * at wsb. */
static void
plan_relocate(contract_plan *pl, uint64_t wsb)
{
	for (size_t i = 0; i < pl->unfold_jobs.size(); i++) {
		pl->unfold_jobs[i].dst += wsb;
		if (pl->unfold_src_ws[i])
			pl->unfold_jobs[i].src += wsb;
	}
	for (auto &j : pl->cin_jobs)
		j.dst += wsb;
	for (auto &j : pl->cout_jobs)
		j.src += wsb;
	for (size_t i = 0; i < pl->cblocks.size(); i++)
		if (pl->cpanel[i])
			pl->cblocks[i].c += wsb;
	for (size_t i = 0; i < pl->pairs.size(); i++) {
		if (pl->x_in_ws[i])
			pl->pairs[i].x += wsb;
		if (pl->y_in_ws[i])
			pl->pairs[i].y += wsb;
	}
}

/* Relocate workspace offsets to addresses and move every descriptor list
 * to one device allocation. */
static void
plan_upload(contract_plan *pl)
{
	const char *fn = "xm_contract";
	if (pl->ws_bytes && !pl->ws_external)
		HIP_OK(fn, hipMalloc(&pl->ws, pl->ws_bytes));
	plan_relocate(pl, (uint64_t)(uintptr_t)pl->ws);
	build_tiles(pl);

	/* one device allocation: scalars | cblocks | pairs | tiles |
	 * K-split lists | jobs */
	std::vector<std::pair<const void *, size_t>> parts;
	auto add = [&](const void *p, size_t n) { parts.push_back({p, n}); };
	add(nullptr, sizeof(gemm_scalars));
	add(pl->cblocks.data(), pl->cblocks.size() * sizeof(gemm_cblock));
	add(pl->pairs.data(), pl->pairs.size() * sizeof(gemm_pair));
	for (int c = 0; c < pl->ncand; c++)
		add(pl->cand_tiles[c].data(),
		    pl->cand_tiles[c].size() * sizeof(gemm_tile));
	for (int c = 0; c < pl->ncand; c++)
		add(pl->cand_fix[c].data(),
		    pl->cand_fix[c].size() * sizeof(gemm_tile));
	const size_t esz = elem_size(pl->type);
	std::vector<char> ju = elem_jobs_pack(pl->unfold_jobs, esz, &pl->unfold_l);
	std::vector<char> ji = elem_jobs_pack(pl->cin_jobs, esz, &pl->cin_l);
	std::vector<char> jo = elem_jobs_pack(pl->cout_jobs, esz, &pl->cout_l);
	add(ju.data(), ju.size());
	add(ji.data(), ji.size());
	add(jo.data(), jo.size());
	size_t total = 0;
	std::vector<size_t> offs;
	for (auto &p : parts) {
		offs.push_back(total);
		total += (p.second + 255) / 256 * 256;
	}
	HIP_OK(fn, hipMalloc(&pl->dev, std::max<size_t>(total, 256)));
	std::vector<char> host(total, 0);
	for (size_t i = 0; i < parts.size(); i++)
		if (parts[i].first)
			memcpy(host.data() + offs[i], parts[i].first,
			    parts[i].second);
	HIP_OK(fn, hipMemcpy(pl->dev, host.data(), total,
	    hipMemcpyHostToDevice));
	char *base = (char *)pl->dev;
	pl->d_scalars = (gemm_scalars *)(base + offs[0]);
	pl->d_cblocks = (gemm_cblock *)(base + offs[1]);
	pl->d_pairs = (gemm_pair *)(base + offs[2]);
	size_t nslot = 0;
	for (int c = 0; c < pl->ncand; c++) {
		pl->d_cand_tiles[c] = (gemm_tile *)(base + offs[3 + c]);
		pl->d_cand_fix[c] = (gemm_tile *)(base +
		    offs[3 + pl->ncand + c]);
		/* a split stacked tile (three records) takes two slots a
		 * half */
		nslot = std::max(nslot, pl->cand_stacked[c] ?
		    pl->cand_fix[c].size() / 3 * 2 : pl->cand_fix[c].size());
	}
	/* counters of the persistent GEMM's tile queues, beside the K-split
	 * sums */
	HIP_OK(fn, hipMalloc((void **)&pl->d_queue, 64 * sizeof(uint32_t)));
	if (nslot)
		HIP_OK(fn, hipMalloc(&pl->d_ksw, nslot * 2 * KSPLIT_SLOT *
		    sizeof(double)));
	const size_t oj = 3 + 2 * (size_t)pl->ncand;
	pl->d_unfold = base + offs[oj];
	pl->d_cin = base + offs[oj + 1];
	pl->d_cout = base + offs[oj + 2];
	select_tiles(pl);
	HIP_OK(fn, hipHostMalloc(&pl->host_scalars, sizeof(gemm_scalars),
	    hipHostMallocDefault));
	for (auto &e : pl->ev)
		HIP_OK(fn, hipEventCreate(&e));
}

/* Does the contraction read or write any block in pinned host memory or
 * in a file?  (XM_HOST_STAGING=0: treat host blocks like HBM ones, read
 * in place; file blocks are staged regardless.) */
static bool
touches_host(const xm_tensor_t *a, const xm_tensor_t *b,
    const xm_tensor_t *c)
{
	/* blocks in a file are always staged: no kernel reads a file */
	for (const xm_tensor_t *t : {a, b, c})
		if (tensor_in_file(t))
			return true;
	const char *e = getenv("XM_HOST_STAGING");
	if (e && *e == '0')
		return false;
	for (const xm_tensor_t *t : {a, b, c}) {
		if (t->alloc == nullptr || t->alloc->host_bytes == 0)
			continue;
		for (const xm_block &x : t->blocks)
			if (x.type == XM_BLOCK_TYPE_CANONICAL &&
			    arena_is_host(t->alloc, x.data))
				return true;
	}
	return false;
}

/*
 * Workspace bytes a staging batch aims at: XM_STAGE_BYTES, else a sixth of
 * the whole contraction's estimate (six batches let the copies of one
 * overlap the compute of the previous one: on the o=64 v=384 ladder with v
 * in host memory 6 batches ran at the host link's rate, 1 batch 16% slower,
 * 15 and 44 batches re-staged shared blocks), at least 256 MiB and at most
 * a quarter of the free HBM.
 */
static size_t
stage_budget(double total)
{
	if (const char *e = getenv("XM_STAGE_BYTES"))
		return std::max<size_t>(strtoull(e, nullptr, 10), 1);
	size_t fr = 0, tot = 0;
	HIP_OK("xm_contract", hipMemGetInfo(&fr, &tot));
	double b = std::max(total / 6.0, (double)(256ull << 20));
	return (size_t)std::min(b, (double)fr / 4.0);
}

static size_t
round256(size_t n)
{
	return (n + 255) / 256 * 256;
}

/*
 * Lay out a staged batch's file ranges in the pinned buffer at `base`:
 * its file operand ranges (`stage` entries, which then copy from the
 * pinned copies) and its file C blocks (the C-in and C-out permutes then
 * read and write the pinned copies).  base == nullptr: count only.
 * Returns the buffer bytes the batch needs.
 */
static size_t
file_layout(contract_plan *sub, char *base)
{
	const size_t esz = elem_size(sub->type);
	size_t off = 0;
	if (base) {
		sub->file_ops.clear();
		sub->file_c.clear();
		sub->file_bytes = 0;
	}
	for (auto &cp : sub->stage) {
		if (!file_ptr(cp.host))
			continue;
		if (base) {
			sub->file_ops.push_back({cp.host, cp.bytes, off});
			sub->file_bytes += (double)cp.bytes;
			cp.host = (uint64_t)(uintptr_t)(base + off);
		}
		off += round256(cp.bytes);
	}
	std::map<uint64_t, size_t> cofs;
	for (auto &j : sub->cout_jobs) {
		if (!file_ptr(j.dst))
			continue;
		auto it = cofs.find(j.dst);
		if (it == cofs.end()) {
			it = cofs.emplace(j.dst, off).first;
			if (base) {
				sub->file_c.push_back({j.dst, j.n * esz, off});
				sub->file_bytes += 2.0 * (double)(j.n * esz);
			}
			off += round256(j.n * esz);
		}
		if (base)
			j.dst = (uint64_t)(uintptr_t)(base + it->second);
	}
	if (base)
		for (auto &j : sub->cin_jobs) {
			if (!file_ptr(j.src))
				continue;
			auto it = cofs.find(j.src);
			if (it == cofs.end())
				XM_FATAL("xm_contract", "file C block without "
				    "a write-back");
			j.src = (uint64_t)(uintptr_t)(base + it->second);
		}
	return off;
}

/*
 * Plan of a contraction that touches host blocks: C blocks (libxm's
 * order) cut into batches whose estimated workspace (host operand blocks
 * twice: staged copy and panel; other operand blocks once, as an upper
 * bound on their panels; the C block) stays within the budget, each batch
 * a plan of its own on one of two alternating workspace regions.  FP64
 * contractions run in native precision here.
 */
static contract_plan *
build_staged(const xm_tensor_t *a, const xm_tensor_t *b, const xm_tensor_t *c,
    const masks_t &mk)
{
	const char *fn = "xm_contract";
	contract_plan *pl = new contract_plan;
	pl->type = c->type;
	pl->precision = XM_PRECISION_NATIVE;
	const size_t esz = elem_size(c->type);
	xm_dim_t *list;
	size_t nlist;
	xm_tensor_get_canonical_block_list(c, &list, &nlist);
	std::vector<block_pair> all;
	/* batch boundaries for a budget; *total: the estimate summed over
	 * the batches */
	auto partition = [&](double budget, double *total) {
		std::vector<size_t> cuts = {0};
		std::set<uint64_t> seen;
		double est = 0;
		*total = 0;
		for (size_t ic = 0; ic < nlist; ic++) {
			size_t fc = block_flat(c, list[ic]);
			xm_dim_t dc = block_dims_flat(c, fc);
			enumerate_pairs(a, b, list[ic], mk, all, nullptr);
			auto cost = [&](bool fresh) {
				std::set<uint64_t> s2;
				double x = (double)xm_dim_dot(&dc) * esz;
				for (auto &p : all) {
					if (!p.nonzero || p.alpha == xm_scalar_t(0))
						continue;
					for (int w = 0; w < 2; w++) {
						const xm_tensor_t *t = w ? b : a;
						size_t f = w ? p.fb : p.fa;
						const xm_block &x0 = t->blocks[f];
						uint64_t src = x0.type ==
						    XM_BLOCK_TYPE_DERIVATIVE ?
						    t->blocks[x0.data].data : x0.data;
						if ((!fresh && seen.count(src)) ||
						    !s2.insert(src).second)
							continue;
						xm_dim_t d = block_dims_flat(t, f);
						double by = (double)xm_dim_dot(&d) *
						    esz;
						x += arena_is_host(t->alloc, src) ?
						    2 * by : by;
					}
				}
				return std::make_pair(x, s2);
			};
			auto cs = cost(false);
			if (ic > cuts.back() && est + cs.first > budget) {
				cuts.push_back(ic);
				seen.clear();
				cs = cost(true);
				est = 0;
			}
			est += cs.first;
			*total += cs.first;
			seen.insert(cs.second.begin(), cs.second.end());
		}
		cuts.push_back(nlist);
		return cuts;
	};
	double total = 0;
	partition(1e300, &total);
	std::vector<size_t> cuts = partition((double)stage_budget(total),
	    &total);
	free(list);
	size_t rb[2] = {0, 0};
	for (size_t i = 0; i + 1 < cuts.size(); i++) {
		contract_plan *sub = build_plan(a, b, c, mk, XM_PRECISION_NATIVE,
		    cuts[i], cuts[i + 1], true);
		sub->graph_failed = true;	/* direct launches only */
		sub->ws_external = true;
		rb[i % 2] = std::max(rb[i % 2], sub->ws_bytes);
		pl->batches.push_back(sub);
	}
	for (int r = 0; r < 2; r++)
		if (rb[r])
			HIP_OK(fn, hipMalloc(&pl->region[r], rb[r]));
	/* file ranges: in the pinned buffer of the batch's region */
	size_t fb[2] = {0, 0};
	for (size_t i = 0; i < pl->batches.size(); i++)
		fb[i % 2] = std::max(fb[i % 2], file_layout(pl->batches[i],
		    nullptr));
	for (int r = 0; r < 2; r++)
		if (fb[r])
			HIP_OK(fn, hipHostMalloc(&pl->bounce[r], fb[r],
			    hipHostMallocMapped));
	for (size_t i = 0; i < pl->batches.size(); i++) {
		contract_plan *sub = pl->batches[i];
		if (fb[i % 2])
			file_layout(sub, (char *)pl->bounce[i % 2]);
		sub->ws = pl->region[i % 2];
		plan_upload(sub);
	}
	HIP_OK(fn, hipEventCreateWithFlags(&pl->ev_start, hipEventDisableTiming));
	for (int r = 0; r < 2; r++) {
		HIP_OK(fn, hipEventCreateWithFlags(&pl->ev_staged[r],
		    hipEventDisableTiming));
		HIP_OK(fn, hipEventCreateWithFlags(&pl->ev_done[r],
		    hipEventDisableTiming));
	}
	return pl;
}

static contract_plan *
get_plan(const xm_tensor_t *a, const xm_tensor_t *b, const xm_tensor_t *c,
    const char *ia, const char *ib, const char *ic, const masks_t &mk,
    bool *cached, const xm_tensor_t *x = nullptr)
{
	int prec = current_precision();
	if (c->type != XM_SCALAR_DOUBLE)
		prec = XM_PRECISION_NATIVE;
	plan_key key{a->id, a->version, b->id, b->version, c->id, c->version,
	    ia, ib, ic, prec, prec == XM_PRECISION_SPLIT_BF16 ?
	    current_nslices() : prec == XM_PRECISION_SPLIT_INT8 ?
	    100 * current_nslices() + current_nmoduli() : 0};
	image_map images;
	if (x) {
		key.idx = x->id;
		key.vx = x->version;
		images.x = x;
		for (size_t f = 0; f < x->blocks.size(); f++)
			if (x->blocks[f].type == XM_BLOCK_TYPE_DERIVATIVE)
				images.of[(size_t)x->blocks[f].data].push_back(f);
	}
	const image_map *im = x ? &images : nullptr;
	std::lock_guard<std::mutex> g(cache_mu);
	for (auto it = cache.begin(); it != cache.end(); ++it)
		if (!(it->first < key) && !(key < it->first)) {
			cache.splice(cache.begin(), cache, it);
			*cached = true;
			return cache.front().second;
		}
	*cached = false;
	if (x == nullptr && touches_host(a, b, c)) {
		contract_plan *sp = build_staged(a, b, c, mk);
		cache.push_front({key, sp});
		while (cache.size() > CACHE_MAX) {
			plan_release(cache.back().second);
			cache.pop_back();
		}
		return sp;
	}
	contract_plan *pl = build_plan(a, b, c, mk, prec, 0, SIZE_MAX, false,
	    im);
	if (prec == XM_PRECISION_SPLIT_INT8 && !crt_eligible(pl)) {
		/* factors without an integer chain: the BF16 form */
		delete pl;
		prec = XM_PRECISION_SPLIT_BF16;
		pl = build_plan(a, b, c, mk, prec, 0, SIZE_MAX, false, im);
	}
	pl->nslices = prec == XM_PRECISION_SPLIT_BF16 ? current_nslices() : 0;
	plan_upload(pl);
	if (prec == XM_PRECISION_SPLIT_BF16)
		split_prepare(pl, a, b, c, mk);
	if (prec == XM_PRECISION_SPLIT_INT8)
		crt_prepare(pl, mk, c);
	cache.push_front({key, pl});
	while (cache.size() > CACHE_MAX) {
		plan_release(cache.back().second);
		cache.pop_back();
	}
	return pl;
}

void
plan_cache_clear(void)
{
	std::lock_guard<std::mutex> g(cache_mu);
	for (auto &p : cache)
		plan_release(p.second);
	cache.clear();
}

static void noalloc_cache_drop(uint64_t id);

/* Drop the cached plans that reference tensor `id` (the tensor is being
 * freed: its plans can never be reused and hold device memory). */
void
plan_cache_drop(uint64_t id)
{
	noalloc_cache_drop(id);
	std::lock_guard<std::mutex> g(cache_mu);
	for (auto it = cache.begin(); it != cache.end();) {
		const plan_key &k = it->first;
		if (k.ida == id || k.idb == id || k.idc == id) {
			plan_release(it->second);
			it = cache.erase(it);
		} else {
			++it;
		}
	}
}

/* ---- execution ------------------------------------------------------ */

static xm_contract_stats_t last_stats;

static void launch_sequence(contract_plan *pl, uint32_t flags, hipStream_t s,
    bool stamps);

static void
run_plan(contract_plan *pl, xm_scalar_t alpha, xm_scalar_t beta)
{
	const char *fn = "xm_contract";
	hipStream_t s = xm_stream();
	bump_data_epoch();
	gemm_scalars sc;
	sc.alpha_re = alpha.real();
	sc.alpha_im = alpha.imag();
	sc.beta_re = beta.real();
	sc.beta_im = beta.imag();
	/* alpha == 0: C = beta*C only (libxm src/contract.c:159-160) */
	sc.flags = alpha == xm_scalar_t(0) ? 1 : 0;
	sc.flags |= beta == xm_scalar_t(0) ? 2 : 0;
	*pl->host_scalars = sc;
	HIP_OK(fn, hipMemcpyAsync(pl->d_scalars, pl->host_scalars, sizeof sc,
	    hipMemcpyHostToDevice, s));
	/*
	 * Once the tile shape is decided the device sequence is captured as a
	 * graph (per alpha/beta-zero combination, which select the launches)
	 * and replayed: one launch per call.  alpha and beta are read from
	 * d_scalars, written above outside the graph, so a graph serves every
	 * call.  The sequence holds kernels and event records only (no
	 * memset nodes).  A capture that fails leaves the plan on direct
	 * launches.
	 */
	pl->last_graph = false;
	hipGraphExec_t &g = pl->graph[sc.flags & 3];
	if (graph_launch_enabled() && pl->tile_choice >= 0 && !pl->graph_failed &&
	    g == nullptr) {
		hipGraph_t graph = nullptr;
		bool ok = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed) ==
		    hipSuccess;
		if (ok)
			launch_sequence(pl, sc.flags, s, true);
		ok = hipStreamEndCapture(s, &graph) == hipSuccess && ok &&
		    graph != nullptr;
		if (ok)
			ok = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0) ==
			    hipSuccess;
		if (graph)
			(void)hipGraphDestroy(graph);
		if (!ok) {
			(void)hipGetLastError();
			g = nullptr;
			pl->graph_failed = true;
		}
	}
	if (graph_launch_enabled() && g) {
		HIP_OK(fn, hipGraphLaunch(g, s));
		pl->last_graph = true;
	} else {
		launch_sequence(pl, sc.flags, s, false);
	}
	HIP_OK(fn, hipStreamSynchronize(s));
}

/*
 * A staged plan's call: batch i computes on the main stream once its host
 * operand blocks are in region i % 2; the side stream copies batch i + 1's
 * blocks meanwhile, after batch i - 1 (the region's previous user) is done.
 * File ranges: the host reads batch i + 1's into bounce[(i + 1) % 2] while
 * batch i computes, once batch i - 1 (that buffer's previous user) has
 * finished and its C blocks are written back to the file.
 */
static void
run_staged(contract_plan *pl, xm_scalar_t alpha, xm_scalar_t beta)
{
	const char *fn = "xm_contract";
	hipStream_t s = xm_stream(), side = xm_side_stream();
	bump_data_epoch();
	gemm_scalars sc;
	sc.alpha_re = alpha.real();
	sc.alpha_im = alpha.imag();
	sc.beta_re = beta.real();
	sc.beta_im = beta.imag();
	sc.flags = alpha == xm_scalar_t(0) ? 1 : 0;
	sc.flags |= beta == xm_scalar_t(0) ? 2 : 0;
	const bool pairs = !(sc.flags & 1), cin = !(sc.flags & 2);
	const size_t nb = pl->batches.size();
	const bool file = pl->bounce[0] || pl->bounce[1];
	for (contract_plan *sub : pl->batches) {
		*sub->host_scalars = sc;
		HIP_OK(fn, hipMemcpyAsync(sub->d_scalars, sub->host_scalars,
		    sizeof sc, hipMemcpyHostToDevice, s));
		sub->last_graph = false;
	}
	auto stage = [&](size_t i) {
		contract_plan *sub = pl->batches[i];
		char *bb = (char *)pl->bounce[i % 2];
		if (pairs)
			for (auto &f : sub->file_ops)
				arena_file_read(fn, f.fptr, bb + f.off, f.bytes);
		if (cin)
			for (auto &f : sub->file_c)
				arena_file_read(fn, f.fptr, bb + f.off, f.bytes);
		char *base = (char *)sub->ws;
		if (pairs)
			for (auto &cp : sub->stage)
				HIP_OK(fn, hipMemcpyAsync(base + cp.off,
				    (const void *)(uintptr_t)cp.host, cp.bytes,
				    hipMemcpyHostToDevice, side));
		HIP_OK(fn, hipEventRecord(pl->ev_staged[i % 2], side));
	};
	/* batch i's C blocks back to the file (its work has finished) */
	auto write_back = [&](size_t i) {
		char *bb = (char *)pl->bounce[i % 2];
		for (auto &f : pl->batches[i]->file_c)
			arena_file_write(fn, f.fptr, bb + f.off, f.bytes);
	};
	/* the regions may still be read by the previous call */
	HIP_OK(fn, hipEventRecord(pl->ev_start, s));
	HIP_OK(fn, hipStreamWaitEvent(side, pl->ev_start, 0));
	stage(0);
	for (size_t i = 0; i < nb; i++) {
		HIP_OK(fn, hipStreamWaitEvent(s, pl->ev_staged[i % 2], 0));
		launch_sequence(pl->batches[i], sc.flags, s, false);
		HIP_OK(fn, hipEventRecord(pl->ev_done[i % 2], s));
		if (i + 1 < nb) {
			if (i >= 1) {
				if (file) {
					HIP_OK(fn, hipEventSynchronize(
					    pl->ev_done[(i + 1) % 2]));
					write_back(i - 1);
				}
				HIP_OK(fn, hipStreamWaitEvent(side,
				    pl->ev_done[(i + 1) % 2], 0));
			}
			stage(i + 1);
		}
	}
	HIP_OK(fn, hipStreamSynchronize(s));
	HIP_OK(fn, hipStreamSynchronize(side));
	if (file) {
		if (nb >= 2)
			write_back(nb - 2);
		write_back(nb - 1);
	}
}

/* The plan's device work for one call: permutes in, slices, grouped GEMM,
 * permutes out, with four time marks between the stages: HIP events, or
 * (inside a graph, where events cannot be timed) device wall-clock stamps. */
static void
launch_sequence(contract_plan *pl, uint32_t flags, hipStream_t s, bool stamps)
{
	const char *fn = "xm_contract";
	bool do_pairs = !(flags & 1);
	auto mark = [&](int i) {
		if (stamps)
			launch_stamp(stamp_slots() + i, s);
		else
			HIP_OK(fn, hipEventRecord(pl->ev[i], s));
	};
	mark(0);
	if (do_pairs)
		launch_elem_packed(pl->d_unfold, pl->unfold_l, ELEM_GATHER,
		    pl->type, pl->type, 0, s);
	if (!(flags & 2))
		launch_elem_packed(pl->d_cin, pl->cin_l, ELEM_GATHER,
		    pl->type, pl->type, 0, s);
	if (do_pairs && pl->precision == XM_PRECISION_SPLIT_BF16)
		split_slices(pl, s);
	if (do_pairs && pl->precision == XM_PRECISION_SPLIT_INT8)
		crt_quantize(pl, s);
	mark(1);
	if (pl->precision == XM_PRECISION_SPLIT_BF16)
		split_gemm(pl, s);
	else if (pl->precision == XM_PRECISION_SPLIT_INT8)
		crt_gemm(pl, s);
	else
		launch_gemm(pl, s);
	mark(2);
	launch_elem_packed(pl->d_cout, pl->cout_l, ELEM_GATHER, pl->type,
	    pl->type, 0, s);
	mark(3);
}

static void record_stats(contract_plan *pl, bool cached, xm_scalar_t alpha,
    xm_scalar_t beta, double total_ms);

XM_API void
xm_contract(xm_scalar_t alpha, const xm_tensor_t *a, const xm_tensor_t *b,
    xm_scalar_t beta, xm_tensor_t *c, const char *idxa, const char *idxb,
    const char *idxc)
{
	auto t0 = std::chrono::steady_clock::now();
	masks_t mk;
	contract_masks("xm_contract", a, b, c, idxa, idxb, idxc, mk);
	bool cached = false;
	contract_plan *pl = get_plan(a, b, c, idxa, idxb, idxc, mk, &cached);
	if (pl->batches.empty())
		run_plan(pl, alpha, beta);
	else
		run_staged(pl, alpha, beta);
	auto t1 = std::chrono::steady_clock::now();
	record_stats(pl, cached, alpha, beta,
	    std::chrono::duration<double, std::milli>(t1 - t0).count());
}

/*
 * dx = alpha * p (x) q over dx's canonical blocks, each block S also
 * taking the terms of x's derivative blocks whose source is S, read in
 * S's index order and merged with S's own (image_terms): the adjoint's
 * gradient in one grouped GEMM, without per-image outputs or folds.
 * Returns false (nothing done) when an image mixes the dims dx takes
 * from p with those from q, or a block is in host memory or in a file.
 */
bool
contract_gradient(xm_scalar_t alpha, const xm_tensor_t *p,
    const xm_tensor_t *q, xm_tensor_t *dx, const xm_tensor_t *x,
    const char *ip, const char *iq, const char *ix)
{
	auto t0 = std::chrono::steady_clock::now();
	masks_t mk;
	contract_masks("xm_contract_backward", p, q, dx, ip, iq, ix, mk);
	if (!images_mergeable(x, mk) || touches_host(p, q, dx))
		return false;
	bool cached = false;
	contract_plan *pl = get_plan(p, q, dx, ip, iq, ix, mk, &cached, x);
	run_plan(pl, alpha, 0);
	auto t1 = std::chrono::steady_clock::now();
	record_stats(pl, cached, alpha, 0,
	    std::chrono::duration<double, std::milli>(t1 - t0).count());
	return true;
}

static void
record_stats(contract_plan *pl, bool cached, xm_scalar_t alpha,
    xm_scalar_t beta, double total_ms)
{
	xm_contract_stats_t &st = last_stats;
	memset(&st, 0, sizeof st);
	const bool ops = alpha != xm_scalar_t(0), cin = beta != xm_scalar_t(0);
	/* a staged plan reports the sums over its batches */
	std::vector<contract_plan *> parts = pl->batches;
	if (parts.empty())
		parts.push_back(pl);
	std::vector<float> gms;
	for (contract_plan *q : parts) {
		st.flops += q->flops;
		st.flops_pre += q->flops_pre;
		st.npairs += q->npairs;
		st.npairs_pre += q->npairs_pre;
		st.nblocks_c += q->nblocks_c;
		st.nunfold += q->nunfold;
		st.unfold_bytes += q->unfold_bytes;
		double ms[3];
		if (q->last_graph) {
			const uint64_t *t = stamp_slots();
			for (int k = 0; k < 3; k++)
				ms[k] = stamp_ms(t[k], t[k + 1]);
		} else {
			for (int k = 0; k < 3; k++)
				ms[k] = elapsed_ms(q->ev[k], q->ev[k + 1]);
		}
		st.prep_ms += ms[0];
		st.gemm_ms += ms[1];
		st.post_ms += ms[2];
		st.gemm_launches += q->gemm_launches;
		st.ksplit_tiles += q->nfix;
		st.panel_bytes += q->panel_bytes;
		st.cin_bytes += q->cin_bytes;
		st.cout_bytes += q->cout_bytes;
		st.classes += q->nclasses;
		st.class_cblocks += q->class_cblocks;
		st.permute_tiled += (ops ? q->unfold_l.ntiled : 0) +
		    (cin ? q->cin_l.ntiled : 0) + q->cout_l.ntiled;
		st.permute_gather += (ops ? q->unfold_l.ngather : 0) +
		    (cin ? q->cin_l.ngather : 0) + q->cout_l.ngather;
		if (ops)
			st.staged_bytes += q->staged_bytes;
		for (auto &f : q->file_ops)
			st.file_bytes += ops ? (double)f.bytes : 0;
		for (auto &f : q->file_c)
			st.file_bytes += (cin ? 2.0 : 1.0) * (double)f.bytes;
		gms.push_back((float)ms[1]);
	}
	st.total_ms = total_ms;
	st.plan_cached = cached;
	st.precision = pl->precision;
	st.nslices = pl->nslices;
	st.tile_p = (int)parts[0]->cand_bp[parts[0]->tile_active];
	st.tile_q = (int)parts[0]->cand_bq[parts[0]->tile_active];
	st.tile_tuning = parts[0]->tile_choice < 0;
	st.graph = pl->last_graph;
	st.staged_batches = (int)pl->batches.size();
	st.err_bound = crt_error_bound(pl);
	snprintf(st.gemm_kernel, sizeof st.gemm_kernel, "%s",
	    parts[0]->gemm_kernel);
	for (size_t i = 0; i < parts.size(); i++)
		gemm_tile_record(parts[i], gms[i]);
}

XM_API void
xm_contract_get_stats(xm_contract_stats_t *s)
{
	*s = last_stats;
}

void
contract_set_stats(const xm_contract_stats_t &s)
{
	last_stats = s;
}

XM_API size_t
xm_contract_get_pairs(const xm_tensor_t *a, const xm_tensor_t *b,
    const xm_tensor_t *c, const char *idxa, const char *idxb,
    const char *idxc, xm_contract_pair_t *out, size_t maxout)
{
	masks_t mk;
	contract_masks("xm_contract_get_pairs", a, b, c, idxa, idxb, idxc, mk);
	contract_plan *pl = build_plan(a, b, c, mk, XM_PRECISION_NATIVE);
	size_t n = pl->pair_list.size();
	for (size_t i = 0; i < n && i < maxout; i++) {
		const plan_pair &p = pl->pair_list[i];
		out[i].blkidxc = grid_index(c, p.fc);
		out[i].blkidxa = grid_index(a, p.fa);
		out[i].blkidxb = grid_index(b, p.fb);
		out[i].alpha = p.alpha;
		out[i].m = p.m;
		out[i].n = p.n;
		out[i].k = p.k;
	}
	memset(&last_stats, 0, sizeof last_stats);
	last_stats.flops = pl->flops;
	last_stats.flops_pre = pl->flops_pre;
	last_stats.npairs = pl->npairs;
	last_stats.npairs_pre = pl->npairs_pre;
	last_stats.nblocks_c = pl->nblocks_c;
	last_stats.nunfold = pl->nunfold;
	delete pl;
	return n;
}

/* ---- contraction without device allocation ------------------------ */

/*
 * xm_contract_noalloc keeps what a native-precision plan needs on the
 * device beyond the tensors' own blocks (permuted operand and C panels,
 * then the scalars, C-block, pair and tile descriptors and the packed
 * permute jobs) in one file-scope array per device, `scratch`.  In host
 * code the array's name is not a device address: its address on the
 * current device comes from hipGetSymbolAddress, once per device, and
 * only that address is used (kernel arguments, copies).  Every call on a
 * device shares that device's array and leaves its descriptors there;
 * nothing in it needs clearing: each call copies its descriptors in, and
 * the permutes write every panel element before the GEMM reads it.  Calls
 * that overlap on one device (from several host threads) race on it.
 */
#define NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
#define NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

static std::mutex noalloc_mu;
static void *noalloc_ptr[NOALLOC_MAX_DEVICES];

/* Address of `scratch` on the current device (looked up on the first call
 * with that device current). */
static char *
noalloc_scratch(const char *fn, int *dev)
{
	HIP_OK(fn, hipGetDevice(dev));
	if (*dev < 0 || *dev >= NOALLOC_MAX_DEVICES)
		XM_FATAL(fn, "device index out of range");
	std::lock_guard<std::mutex> g(noalloc_mu);
	if (noalloc_ptr[*dev] == nullptr) {
		void *p = nullptr;
		HIP_OK(fn, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
		if (p == nullptr)
			XM_FATAL(fn, "no device address for the scratch array");
		noalloc_ptr[*dev] = p;
	}
	return (char *)noalloc_ptr[*dev];
}

/* A host-only plan laid out for one device's scratch array: panels at
 * offset 0, then the descriptor image (copied in by every call). */
struct noalloc_plan {
	contract_plan *pl = nullptr;
	std::vector<char> image;
	size_t doff = 0;	/* scratch offset of the image */
	size_t need = 0;	/* scratch bytes the call needs */
};

struct noalloc_key {
	plan_key k;
	int dev;
	bool operator<(const noalloc_key &o) const {
		if (dev != o.dev)
			return dev < o.dev;
		return k < o.k;
	}
};

static std::list<std::pair<noalloc_key, noalloc_plan>> noalloc_cache;

static void
noalloc_cache_drop(uint64_t id)
{
	std::lock_guard<std::mutex> g(noalloc_mu);
	for (auto it = noalloc_cache.begin(); it != noalloc_cache.end();) {
		const plan_key &k = it->first.k;
		if (k.ida == id || k.idb == id || k.idc == id) {
			delete it->second.pl;
			it = noalloc_cache.erase(it);
		} else {
			++it;
		}
	}
}

/* The plan of c = alpha a b + beta c with its device state in `base`;
 * np.need says how much of the array it takes (the caller checks it). */
static void
noalloc_build(const xm_tensor_t *a, const xm_tensor_t *b,
    const xm_tensor_t *c, const masks_t &mk, char *base, noalloc_plan &np)
{
	contract_plan *pl = build_plan(a, b, c, mk, XM_PRECISION_NATIVE);
	pl->single_tile = pl->no_ksplit = true;
	const uint64_t wsb = (uint64_t)(uintptr_t)base;
	plan_relocate(pl, wsb);
	build_tiles(pl);
	const size_t esz = elem_size(pl->type);
	std::vector<char> ju = elem_jobs_pack(pl->unfold_jobs, esz, &pl->unfold_l);
	std::vector<char> ji = elem_jobs_pack(pl->cin_jobs, esz, &pl->cin_l);
	std::vector<char> jo = elem_jobs_pack(pl->cout_jobs, esz, &pl->cout_l);
	std::vector<std::pair<const void *, size_t>> parts = {
		{nullptr, sizeof(gemm_scalars)},
		{pl->cblocks.data(), pl->cblocks.size() * sizeof(gemm_cblock)},
		{pl->pairs.data(), pl->pairs.size() * sizeof(gemm_pair)},
		{pl->cand_tiles[0].data(),
		    pl->cand_tiles[0].size() * sizeof(gemm_tile)},
		{ju.data(), ju.size()}, {ji.data(), ji.size()},
		{jo.data(), jo.size()}};
	np.doff = (pl->ws_bytes + 255) / 256 * 256;
	std::vector<size_t> offs;
	size_t total = 0;
	for (auto &p : parts) {
		offs.push_back(total);
		total += (p.second + 255) / 256 * 256;
	}
	np.need = np.doff + total;
	np.pl = pl;
	if (np.need > NOALLOC_SCRATCH_BYTES)
		return;
	np.image.assign(total, 0);
	for (size_t i = 0; i < parts.size(); i++)
		if (parts[i].first && parts[i].second)
			memcpy(np.image.data() + offs[i], parts[i].first,
			    parts[i].second);
	char *d = base + np.doff;
	pl->d_scalars = (gemm_scalars *)(d + offs[0]);
	pl->d_cblocks = (gemm_cblock *)(d + offs[1]);
	pl->d_pairs = (gemm_pair *)(d + offs[2]);
	pl->d_cand_tiles[0] = (gemm_tile *)(d + offs[3]);
	pl->d_cand_fix[0] = nullptr;
	pl->d_unfold = d + offs[4];
	pl->d_cin = d + offs[5];
	pl->d_cout = d + offs[6];
	select_tiles(pl);
}

XM_API int
xm_contract_noalloc(xm_scalar_t alpha, const xm_tensor_t *a,
    const xm_tensor_t *b, xm_scalar_t beta, xm_tensor_t *c, const char *idxa,
    const char *idxb, const char *idxc)
{
	const char *fn = "xm_contract_noalloc";
	auto t0 = std::chrono::steady_clock::now();
	masks_t mk;
	contract_masks(fn, a, b, c, idxa, idxb, idxc, mk);
	if (touches_host(a, b, c)) {
		fprintf(stderr, "libxm: %s: blocks in host memory or in a file "
		    "need staging buffers; use xm_contract\n", fn);
		return -1;
	}
	int dev = 0;
	char *base = noalloc_scratch(fn, &dev);
	noalloc_key key{plan_key{a->id, a->version, b->id, b->version, c->id,
	    c->version, idxa, idxb, idxc, XM_PRECISION_NATIVE, 0}, dev};
	std::unique_lock<std::mutex> g(noalloc_mu);
	auto it = noalloc_cache.begin();
	for (; it != noalloc_cache.end(); ++it)
		if (!(it->first < key) && !(key < it->first))
			break;
	bool cached = it != noalloc_cache.end();
	if (cached) {
		noalloc_cache.splice(noalloc_cache.begin(), noalloc_cache, it);
	} else {
		noalloc_plan np;
		noalloc_build(a, b, c, mk, base, np);
		noalloc_cache.push_front({key, np});
		while (noalloc_cache.size() > CACHE_MAX) {
			delete noalloc_cache.back().second.pl;
			noalloc_cache.pop_back();
		}
	}
	noalloc_plan &np = noalloc_cache.front().second;
	if (np.need > NOALLOC_SCRATCH_BYTES) {
		fprintf(stderr, "libxm: %s: the contraction needs %zu bytes of "
		    "device scratch (panels and descriptors), more than the %zu "
		    "of the no-allocation path; use xm_contract\n", fn, np.need,
		    NOALLOC_SCRATCH_BYTES);
		return -1;
	}
	contract_plan *pl = np.pl;
	hipStream_t s = xm_stream();
	bump_data_epoch();
	gemm_scalars sc;
	memset(&sc, 0, sizeof sc);
	sc.alpha_re = alpha.real();
	sc.alpha_im = alpha.imag();
	sc.beta_re = beta.real();
	sc.beta_im = beta.imag();
	sc.flags = alpha == xm_scalar_t(0) ? 1 : 0;
	sc.flags |= beta == xm_scalar_t(0) ? 2 : 0;
	memcpy(np.image.data(), &sc, sizeof sc);
	HIP_OK(fn, hipMemcpyAsync(base + np.doff, np.image.data(),
	    np.image.size(), hipMemcpyHostToDevice, s));
	if (!(sc.flags & 1))
		launch_elem_packed(pl->d_unfold, pl->unfold_l, ELEM_GATHER,
		    pl->type, pl->type, 0, s);
	if (!(sc.flags & 2))
		launch_elem_packed(pl->d_cin, pl->cin_l, ELEM_GATHER, pl->type,
		    pl->type, 0, s);
	launch_gemm(pl, s);
	launch_elem_packed(pl->d_cout, pl->cout_l, ELEM_GATHER, pl->type,
	    pl->type, 0, s);
	HIP_OK(fn, hipStreamSynchronize(s));
	auto t1 = std::chrono::steady_clock::now();
	xm_contract_stats_t &st = last_stats;
	memset(&st, 0, sizeof st);
	st.flops = pl->flops;
	st.flops_pre = pl->flops_pre;
	st.npairs = pl->npairs;
	st.npairs_pre = pl->npairs_pre;
	st.nblocks_c = pl->nblocks_c;
	st.nunfold = pl->nunfold;
	st.unfold_bytes = pl->unfold_bytes;
	st.panel_bytes = pl->panel_bytes;
	st.cin_bytes = pl->cin_bytes;
	st.cout_bytes = pl->cout_bytes;
	st.total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
	st.plan_cached = cached;
	st.precision = XM_PRECISION_NATIVE;
	st.gemm_launches = pl->gemm_launches;
	st.tile_p = (int)pl->cand_bp[0];
	st.tile_q = (int)pl->cand_bq[0];
	snprintf(st.gemm_kernel, sizeof st.gemm_kernel, "%s", pl->gemm_kernel);
	return 0;
}
