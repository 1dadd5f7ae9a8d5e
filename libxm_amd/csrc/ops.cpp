/*
 * ops.cpp - tensor-wide block operations on the GPU
 * (libxm src/xm.c:28-508).  Each call builds one job per canonical block
 * of the output and runs one kernel over all of them.
 */
#include "contract_plan.h"

#include <algorithm>
#include <list>
#include <map>
#include <mutex>
#include <set>

/* Same checks and messages as libxm src/xm.c:169-189. */
static void
pair_masks(const char *fn, const xm_tensor_t *a, const xm_tensor_t *b,
    const char *idxa, const char *idxb, bool same_type, xm_dim_t *ca,
    xm_dim_t *cb)
{
	if (a->alloc != b->alloc)
		XM_FATAL(fn, "tensors must use same allocator");
	if (same_type && a->type != b->type)
		XM_FATAL(fn, "tensors must have same scalar type");
	if (strlen(idxa) != a->bs->dims.n)
		XM_FATAL(fn, "idxa does not match tensor dimensions");
	if (strlen(idxb) != b->bs->dims.n)
		XM_FATAL(fn, "idxb does not match tensor dimensions");
	xm_make_masks(idxa, idxb, ca, cb);
	if (ca->n != a->bs->dims.n || cb->n != b->bs->dims.n)
		XM_FATAL(fn, "index spaces do not match");
	for (size_t k = 0; k < ca->n; k++)
		if (!xm_block_space_eq1(a->bs, ca->i[k], b->bs, cb->i[k]))
			XM_FATAL(fn, "inconsistent block-spaces");
}

/* Source strides of every logical index of block `flat` of t. */
static void
block_strides(const xm_tensor_t *t, size_t flat, uint64_t *ls)
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
}

/* Workgroups of xm_dot's per-block partial sums (ELEM_CHUNK elements
 * each). */
static size_t
set_chunks(std::vector<elem_job> &jobs)
{
	size_t c = 0;
	for (auto &j : jobs) {
		j.first_chunk = c;
		c += (j.n + ELEM_CHUNK - 1) / ELEM_CHUNK;
	}
	return c;
}

/*
 * One job per canonical block of a, reading the matching block of b
 * through the index pairing (ca, cb) and b's block permutation.
 * `factor` gives the per-block factor from b's block scalar; it returns
 * false to mark the source as absent (treated as zero).
 */
template <typename F>
static std::vector<elem_job>
pair_jobs(const xm_tensor_t *a, const xm_tensor_t *b, const xm_dim_t &ca,
    const xm_dim_t &cb, F factor)
{
	std::vector<elem_job> jobs;
	xm_dim_t ib = xm_dim_zero(cb.n);
	for (size_t fa = 0; fa < a->blocks.size(); fa++) {
		if (a->blocks[fa].type != XM_BLOCK_TYPE_CANONICAL)
			continue;
		xm_dim_t ia = grid_index(a, fa);
		xm_dim_set_mask(&ib, &cb, &ia, &ca);
		size_t fb = block_flat(b, ib);
		const xm_block &bb = b->blocks[fb];
		elem_job j;
		memset(&j, 0, sizeof j);
		xm_dim_t da = block_dims_flat(a, fa);
		j.dst = a->blocks[fa].data;
		j.n = xm_dim_dot(&da);
		j.ndim = (uint32_t)da.n;
		for (size_t k = 0; k < da.n; k++)
			j.dims[k] = (uint32_t)da.i[k];
		xm_scalar_t f = 0;
		bool present = bb.type != XM_BLOCK_TYPE_ZERO && factor(bb, &f);
		if (present) {
			uint64_t ls[XM_MAX_DIM];
			block_strides(b, fb, ls);
			for (size_t k = 0; k < ca.n; k++)
				j.sstride[ca.i[k]] = ls[cb.i[k]];
			j.src = bb.type == XM_BLOCK_TYPE_DERIVATIVE ?
			    b->blocks[bb.data].data : bb.data;
		} else {
			j.src_zero = 1;
		}
		j.f_re = f.real();
		j.f_im = f.imag();
		if (j.n)
			jobs.push_back(j);
	}
	return jobs;
}

/*
 * When a source block is also written by the same call (a and b alias),
 * copy the sources to scratch first so every job reads the values from
 * before the call, as libxm's read-whole-block-then-write order does.
 */
static void
stage_aliased(const xm_tensor_t *a, std::vector<elem_job> &jobs,
    size_t src_esz, hipStream_t s)
{
	std::set<uint64_t> dsts;
	for (auto &j : jobs)
		dsts.insert(j.dst);
	std::map<uint64_t, size_t> need;	/* src ptr -> elements */
	for (auto &j : jobs)
		if (!j.src_zero && dsts.count(j.src))
			need[j.src] = std::max<size_t>(need[j.src], j.n);
	if (need.empty())
		return;
	size_t total = 0;
	std::map<uint64_t, size_t> off;
	for (auto &n : need) {
		off[n.first] = total;
		total += (n.second * src_esz + 255) / 256 * 256;
	}
	char *buf = (char *)scratch_buffer(2, total);
	for (auto &n : need)
		HIP_OK("libxm", hipMemcpyAsync(buf + off[n.first],
		    (const void *)(uintptr_t)n.first, n.second * src_esz,
		    hipMemcpyDeviceToDevice, s));
	for (auto &j : jobs)
		if (!j.src_zero && need.count(j.src))
			j.src = (uint64_t)(uintptr_t)(buf + off[j.src]);
	(void)a;
}

/*
 * Cached job lists.  The jobs of an element operation depend only on the
 * tensors' structures (block kinds, permutations, scalars and data
 * addresses, all covered by tensor id and version), the index strings and
 * the per-call factor that scales b's block scalars; a repeated call with
 * the same key launches the device copy kept here, with no host work
 * beyond validation.  Lists that read a block the same call writes are
 * not cached: they read copies staged in scratch.
 */
struct op_key {
	int op;
	uint64_t ida, va, idb, vb;
	std::string ia, ib;
	double f_re, f_im;
	bool operator==(const op_key &o) const {
		return op == o.op && ida == o.ida && va == o.va &&
		    idb == o.idb && vb == o.vb && ia == o.ia && ib == o.ib &&
		    f_re == o.f_re && f_im == o.f_im;
	}
};

struct op_entry {
	void *dev;		/* packed jobs or segments, device copy */
	elem_launch l;		/* xm_set: ngather segments, gchunks */
};

#define OP_CACHE_MAX 64
static std::list<std::pair<op_key, op_entry>> op_cache;
static std::mutex op_mu;

static bool
op_cache_find(const op_key &k, op_entry *e)
{
	std::lock_guard<std::mutex> g(op_mu);
	for (auto it = op_cache.begin(); it != op_cache.end(); ++it)
		if (it->first == k) {
			op_cache.splice(op_cache.begin(), op_cache, it);
			*e = it->second;
			return true;
		}
	return false;
}

static op_entry
op_cache_insert(const op_key &k, const void *host, size_t bytes,
    const elem_launch &l)
{
	op_entry e{nullptr, l};
	if (bytes) {
		HIP_OK("libxm", hipMalloc(&e.dev, bytes));
		HIP_OK("libxm", hipMemcpy(e.dev, host, bytes,
		    hipMemcpyHostToDevice));
	}
	std::lock_guard<std::mutex> g(op_mu);
	op_cache.push_front({k, e});
	while (op_cache.size() > OP_CACHE_MAX) {
		if (op_cache.back().second.dev)
			HIP_OK("libxm", hipFree(op_cache.back().second.dev));
		op_cache.pop_back();
	}
	return e;
}

/* Forget the job lists of tensor `id` (it is being freed). */
void
op_cache_drop(uint64_t id)
{
	std::lock_guard<std::mutex> g(op_mu);
	for (auto it = op_cache.begin(); it != op_cache.end();) {
		if (it->first.ida == id || it->first.idb == id) {
			if (it->second.dev)
				HIP_OK("libxm", hipFree(it->second.dev));
			it = op_cache.erase(it);
		} else {
			++it;
		}
	}
}

static op_key
make_key(int op, const xm_tensor_t *a, const xm_tensor_t *b,
    const char *ia, const char *ib, xm_scalar_t f)
{
	return op_key{op, a->id, a->version, b ? b->id : 0,
	    b ? b->version : 0, ia ? ia : "", ib ? ib : "", f.real(),
	    f.imag()};
}

static elem_launch last_launch;

/* Jobs of the last xm_add / xm_copy / xm_mul / xm_div in the LDS-tiled and
 * in the gather form (no libxm counterpart). */
XM_API void
xm_amd_last_permute(size_t *tiled, size_t *gather)
{
	if (tiled)
		*tiled = last_launch.ntiled;
	if (gather)
		*gather = last_launch.ngather;
}

/* Run the jobs `build` makes (or their cached device copy) as `op`. */
template <typename B>
static void
run_jobs(const char *fn, const op_key &key, B build, int op,
    xm_scalar_type_t dt, xm_scalar_type_t st, xm_scalar_t alpha,
    const xm_tensor_t *a)
{
	hipStream_t s = xm_stream();
	op_entry e;
	if (!op_cache_find(key, &e)) {
		std::vector<elem_job> jobs = build();
		std::set<uint64_t> dsts;
		bool aliased = false;
		for (auto &j : jobs)
			dsts.insert(j.dst);
		for (auto &j : jobs)
			aliased |= !j.src_zero && dsts.count(j.src) != 0;
		if (aliased) {
			stage_aliased(a, jobs, elem_size(st), s);
			last_launch = launch_elem(jobs.data(), jobs.size(), op,
			    dt, st, alpha, s);
			HIP_OK(fn, hipStreamSynchronize(s));
			return;
		}
		elem_launch l;
		std::vector<char> packed = elem_jobs_pack(jobs, elem_size(st),
		    &l);
		e = op_cache_insert(key, packed.data(), packed.size(), l);
	}
	last_launch = e.l;
	launch_elem_packed(e.dev, e.l, op, dt, st, alpha, s);
	HIP_OK(fn, hipStreamSynchronize(s));
}

/* libxm src/xm.c:28-64 */
XM_API void
xm_set(xm_tensor_t *a, xm_scalar_t x)
{
	const op_key key = make_key(-1, a, nullptr, nullptr, nullptr, 0);
	op_entry e;
	if (!op_cache_find(key, &e)) {
		std::vector<seg_desc> segs;
		size_t c = 0;
		for (size_t f = 0; f < a->blocks.size(); f++) {
			if (a->blocks[f].type != XM_BLOCK_TYPE_CANONICAL)
				continue;
			xm_dim_t d = block_dims_flat(a, f);
			seg_desc sd{a->blocks[f].data, xm_dim_dot(&d), c, f};
			if (sd.n == 0)
				continue;
			c += (sd.n + ELEM_CHUNK - 1) / ELEM_CHUNK;
			segs.push_back(sd);
		}
		elem_launch l;
		l.ngather = segs.size();
		l.gchunks = c;
		e = op_cache_insert(key, segs.data(),
		    segs.size() * sizeof(seg_desc), l);
	}
	if (e.l.ngather == 0)
		return;
	hipStream_t s = xm_stream();
	launch_fill_value((const seg_desc *)e.dev, e.l.ngather, e.l.gchunks,
	    a->type, x, s);
	HIP_OK("xm_set", hipStreamSynchronize(s));
}

XM_API void
xm_amd_fill_random(xm_tensor_t *t, uint64_t seed, double lo, double hi,
    double wlo, double whi)
{
	std::vector<seg_desc> segs;
	size_t c = 0;
	for (size_t f = 0; f < t->blocks.size(); f++) {
		if (t->blocks[f].type != XM_BLOCK_TYPE_CANONICAL)
			continue;
		xm_dim_t d = block_dims_flat(t, f);
		seg_desc sd{t->blocks[f].data, xm_dim_dot(&d), c, f};
		if (sd.n == 0)
			continue;
		c += (sd.n + ELEM_CHUNK - 1) / ELEM_CHUNK;
		segs.push_back(sd);
	}
	if (segs.empty())
		return;
	hipStream_t s = xm_stream();
	void *d = upload(1, segs.data(), segs.size() * sizeof(seg_desc));
	launch_fill_random((const seg_desc *)d, segs.size(), c, t->type, seed,
	    lo, hi, wlo, whi, s);
	HIP_OK("xm_amd_fill_random", hipStreamSynchronize(s));
}

/* libxm src/xm.c:159-249: a = alpha*a + beta*s_b*perm(b). */
XM_API void
xm_add(xm_scalar_t alpha, xm_tensor_t *a, xm_scalar_t beta,
    const xm_tensor_t *b, const char *idxa, const char *idxb)
{
	xm_dim_t ca, cb;
	pair_masks("xm_add", a, b, idxa, idxb, true, &ca, &cb);
	xm_scalar_type_t t = a->type;
	auto build = [&]() {
		return pair_jobs(a, b, ca, cb, [&](const xm_block &bb,
		    xm_scalar_t *f) {
			if (beta == xm_scalar_t(0))
				return false;
			*f = scalar_mul(beta, bb.scalar, t);
			return true;
		});
	};
	run_jobs("xm_add", make_key(ELEM_AXPY, a, b, idxa, idxb, beta), build,
	    ELEM_AXPY, t, t, alpha, a);
}

/* libxm src/xm.c:66-157: a = convert(s*s_b*perm(b)). */
XM_API void
xm_copy(xm_tensor_t *a, xm_scalar_t sc, const xm_tensor_t *b,
    const char *idxa, const char *idxb)
{
	xm_dim_t ca, cb;
	pair_masks("xm_copy", a, b, idxa, idxb, false, &ca, &cb);
	if (is_complex(a->type) != is_complex(b->type))
		XM_FATAL("xm_copy", "unsupported scalar conversion");
	xm_scalar_type_t tb = b->type;
	auto build = [&]() {
		return pair_jobs(a, b, ca, cb, [&](const xm_block &bb,
		    xm_scalar_t *f) {
			if (sc == xm_scalar_t(0))
				return false;
			*f = scalar_mul(sc, bb.scalar, tb);
			return true;
		});
	};
	run_jobs("xm_copy", make_key(ELEM_COPY_SCALED, a, b, idxa, idxb, sc),
	    build, ELEM_COPY_SCALED, a->type, tb, 0, a);
}

/* libxm src/xm.c:251-332: a = (s_b*a)*perm(b). */
XM_API void
xm_mul(xm_tensor_t *a, const xm_tensor_t *b, const char *idxa,
    const char *idxb)
{
	xm_dim_t ca, cb;
	pair_masks("xm_mul", a, b, idxa, idxb, true, &ca, &cb);
	auto build = [&]() {
		return pair_jobs(a, b, ca, cb, [&](const xm_block &bb,
		    xm_scalar_t *f) {
			*f = bb.scalar;
			return true;
		});
	};
	run_jobs("xm_mul", make_key(ELEM_MUL, a, b, idxa, idxb, 0), build,
	    ELEM_MUL, a->type, a->type, 0, a);
}

/* libxm src/xm.c:334-413: a = a/(s_b*perm(b)); zero b blocks are fatal. */
XM_API void
xm_div(xm_tensor_t *a, const xm_tensor_t *b, const char *idxa,
    const char *idxb)
{
	xm_dim_t ca, cb;
	pair_masks("xm_div", a, b, idxa, idxb, true, &ca, &cb);
	auto build = [&]() {
		auto jobs = pair_jobs(a, b, ca, cb, [&](const xm_block &bb,
		    xm_scalar_t *f) {
			*f = bb.scalar;
			return true;
		});
		for (auto &j : jobs)
			if (j.src_zero)
				XM_FATAL("xm_div", "division by zero");
		return jobs;
	};
	run_jobs("xm_div", make_key(ELEM_DIV, a, b, idxa, idxb, 0), build,
	    ELEM_DIV, a->type, a->type, 0, a);
}

/* libxm src/xm.c:415-508: sum over all nonzero block pairs of
 * s_a*s_b*dot(unfold(a), unfold(b)).  Derivative a blocks are first laid
 * out in logical order in scratch; per-chunk partials are summed on the
 * host in a fixed order. */
XM_API xm_scalar_t
xm_dot(const xm_tensor_t *a, const xm_tensor_t *b, const char *idxa,
    const char *idxb)
{
	xm_dim_t ca, cb;
	pair_masks("xm_dot", a, b, idxa, idxb, true, &ca, &cb);
	hipStream_t s = xm_stream();
	size_t esz = elem_size(a->type);
	std::vector<elem_job> gather, dots;
	size_t ws = 0;
	xm_dim_t ib = xm_dim_zero(cb.n);
	for (size_t fa = 0; fa < a->blocks.size(); fa++) {
		const xm_block &ba = a->blocks[fa];
		if (ba.type == XM_BLOCK_TYPE_ZERO)
			continue;
		xm_dim_t ia = grid_index(a, fa);
		xm_dim_set_mask(&ib, &cb, &ia, &ca);
		size_t fb = block_flat(b, ib);
		const xm_block &bb = b->blocks[fb];
		if (bb.type == XM_BLOCK_TYPE_ZERO)
			continue;
		xm_dim_t da = block_dims_flat(a, fa);
		elem_job j;
		memset(&j, 0, sizeof j);
		j.n = xm_dim_dot(&da);
		if (j.n == 0)
			continue;
		j.ndim = (uint32_t)da.n;
		for (size_t k = 0; k < da.n; k++)
			j.dims[k] = (uint32_t)da.i[k];
		uint64_t ls[XM_MAX_DIM];
		block_strides(b, fb, ls);
		for (size_t k = 0; k < ca.n; k++)
			j.sstride[ca.i[k]] = ls[cb.i[k]];
		j.src = bb.type == XM_BLOCK_TYPE_DERIVATIVE ?
		    b->blocks[bb.data].data : bb.data;
		uint64_t asrc = ba.type == XM_BLOCK_TYPE_DERIVATIVE ?
		    a->blocks[ba.data].data : ba.data;
		if (ba.type == XM_BLOCK_TYPE_DERIVATIVE) {
			elem_job g = j;
			g.src = asrc;
			block_strides(a, fa, g.sstride);
			g.dst = ws;		/* relocated below */
			ws += (j.n * esz + 255) / 256 * 256;
			gather.push_back(g);
			j.dst = g.dst;
			j.src_zero = 1;		/* marks: dst is in scratch */
		} else {
			j.dst = asrc;
		}
		xm_scalar_t f = scalar_mul(ba.scalar, bb.scalar, a->type);
		j.f_re = f.real();
		j.f_im = f.imag();
		dots.push_back(j);
	}
	if (dots.empty())
		return 0;
	char *buf = (char *)scratch_buffer(3, ws ? ws : 256);
	for (auto &g : gather)
		g.dst += (uint64_t)(uintptr_t)buf;
	for (auto &j : dots) {
		if (j.src_zero)
			j.dst += (uint64_t)(uintptr_t)buf;
		j.src_zero = 0;
	}
	if (!gather.empty()) {
		launch_elem(gather.data(), gather.size(), ELEM_GATHER,
		    a->type, a->type, 0, s);
		HIP_OK("xm_dot", hipStreamSynchronize(s));
	}
	size_t nc = set_chunks(dots);
	double *dp = (double *)scratch_buffer(4, 2 * nc * sizeof(double));
	launch_dot(dots.data(), dots.size(), nc, a->type, dp, s);
	std::vector<double> hp(2 * nc);
	HIP_OK("xm_dot", hipMemcpyAsync(hp.data(), dp, hp.size() * sizeof(double),
	    hipMemcpyDeviceToHost, s));
	HIP_OK("xm_dot", hipStreamSynchronize(s));
	zdouble sum = 0;
	for (size_t i = 0; i < nc; i++)
		sum += zdouble(hp[2 * i], hp[2 * i + 1]);
	if (a->type == XM_SCALAR_FLOAT || a->type == XM_SCALAR_DOUBLE)
		sum = sum.real();
	if (a->type == XM_SCALAR_FLOAT)
		sum = (double)(float)sum.real();
	if (a->type == XM_SCALAR_FLOAT_COMPLEX)
		sum = zdouble((float)sum.real(), (float)sum.imag());
	return sum;
}
