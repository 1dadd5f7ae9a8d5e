/*
 * kernels_elem.hip - bandwidth-bound block kernels.
 *
 * Work is cut into chunks of CHUNK elements; every workgroup takes one
 * chunk of one job (a block, or a block pair for two-operand ops).  The
 * chunk -> job mapping is a binary search over the jobs' first_chunk
 * prefix, done once per workgroup.
 *
 * Element rounding follows libxm's C loops exactly (libxm
 * src/scalar.c:123-327): every product and sum is rounded separately
 * (explicit __dmul_rn/__dadd_rn style operations, never a fused
 * multiply-add), complex products are (ac - bd) + i(ad + bc).  The file
 * is compiled with -ffp-contract=off as well.
 */
#include "internal.h"

#define CHUNK ELEM_CHUNK
#define WG 256

/* ---- exactly rounded scalar arithmetic ---------------------------- */

struct c32 { float re, im; };
struct c64 { double re, im; };

__device__ __forceinline__ float mulr(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ double mulr(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ float addr(float a, float b) { return __fadd_rn(a, b); }
__device__ __forceinline__ double addr(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ float subr(float a, float b) { return __fsub_rn(a, b); }
__device__ __forceinline__ double subr(double a, double b) { return __dsub_rn(a, b); }
__device__ __forceinline__ float divr(float a, float b) { return __fdiv_rn(a, b); }
__device__ __forceinline__ double divr(double a, double b) { return __ddiv_rn(a, b); }

template <typename R> struct cplx { R re, im; };

template <typename R> __device__ __forceinline__ R emul(R a, R b) { return mulr(a, b); }
template <typename R> __device__ __forceinline__ cplx<R>
emul(cplx<R> a, cplx<R> b)
{
	cplx<R> r;
	r.re = subr(mulr(a.re, b.re), mulr(a.im, b.im));
	r.im = addr(mulr(a.re, b.im), mulr(a.im, b.re));
	return r;
}
template <typename R> __device__ __forceinline__ R eadd(R a, R b) { return addr(a, b); }
template <typename R> __device__ __forceinline__ cplx<R>
eadd(cplx<R> a, cplx<R> b)
{
	cplx<R> r;
	r.re = addr(a.re, b.re);
	r.im = addr(a.im, b.im);
	return r;
}
template <typename R> __device__ __forceinline__ R ediv(R a, R b) { return divr(a, b); }
/* Complex quotient by Smith's method (R. L. Smith, CACM 5(8), 1962). */
template <typename R> __device__ __forceinline__ cplx<R>
ediv(cplx<R> a, cplx<R> b)
{
	cplx<R> r;
	if (fabs(b.re) >= fabs(b.im)) {
		R q = divr(b.im, b.re), d = addr(b.re, mulr(b.im, q));
		r.re = divr(addr(a.re, mulr(a.im, q)), d);
		r.im = divr(subr(a.im, mulr(a.re, q)), d);
	} else {
		R q = divr(b.re, b.im), d = addr(b.im, mulr(b.re, q));
		r.re = divr(addr(mulr(a.re, q), a.im), d);
		r.im = divr(subr(mulr(a.im, q), a.re), d);
	}
	return r;
}

template <typename T> __device__ __forceinline__ T mk(double re, double im);
template <> __device__ __forceinline__ float mk<float>(double re, double) { return (float)re; }
template <> __device__ __forceinline__ double mk<double>(double re, double) { return re; }
template <> __device__ __forceinline__ cplx<float> mk<cplx<float>>(double re, double im)
{ return cplx<float>{(float)re, (float)im}; }
template <> __device__ __forceinline__ cplx<double> mk<cplx<double>>(double re, double im)
{ return cplx<double>{re, im}; }

/* conversions between element types of the same kind (xm_copy) */
template <typename D, typename S> __device__ __forceinline__ D conv(S x) { return (D)x; }
template <> __device__ __forceinline__ cplx<float> conv(cplx<double> x)
{ return cplx<float>{(float)x.re, (float)x.im}; }
template <> __device__ __forceinline__ cplx<double> conv(cplx<float> x)
{ return cplx<double>{(double)x.re, (double)x.im}; }
template <> __device__ __forceinline__ cplx<float> conv(cplx<float> x) { return x; }
template <> __device__ __forceinline__ cplx<double> conv(cplx<double> x) { return x; }

/* ---- chunk -> job lookup ------------------------------------------ */

/*
 * The last job with first_chunk <= chunk (jobs[0].first_chunk is 0).  Each
 * round every thread tests one job of the remaining range at a stride of
 * range / WG and the range shrinks to the bucket found: two rounds of one
 * load per thread for up to WG^2 jobs, where a one-thread binary search
 * would chain log2(njobs) dependent loads.  All threads must call it.
 */
template <typename J>
__device__ __forceinline__ uint32_t
find_job(const J *jobs, uint32_t njobs, uint64_t chunk)
{
	__shared__ uint32_t best;
	uint32_t lo = 0, hi = njobs;	/* answer in [lo, hi) */
	while (hi - lo > 1) {
		const uint32_t stride = (hi - lo + WG - 1) / WG;
		if (threadIdx.x == 0)
			best = 0;
		__syncthreads();
		const uint32_t j = lo + threadIdx.x * stride;
		if (j < hi && jobs[j].first_chunk <= chunk)
			atomicMax(&best, threadIdx.x);
		__syncthreads();
		const uint32_t b = best;
		__syncthreads();
		lo += b * stride;
		hi = min(hi, lo + stride);
	}
	return lo;
}

/* ---- fills -------------------------------------------------------- */

template <typename T>
__global__ void __launch_bounds__(WG)
k_fill_value(const seg_desc *segs, uint32_t nseg, T v)
{
	uint64_t chunk = blockIdx.x;
	uint32_t j = find_job(segs, nseg, chunk);
	const seg_desc s = segs[j];
	uint64_t base = (chunk - s.first_chunk) * CHUNK;
	T *p = (T *)(uintptr_t)s.ptr;
	for (uint64_t i = base + threadIdx.x; i < s.n && i < base + CHUNK;
	    i += WG)
		p[i] = v;
}

__device__ __forceinline__ uint64_t
mix64(uint64_t z)
{
	/* SplitMix64 finalizer (Steele, Lea, Flood, OOPSLA 2014). */
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

__device__ __forceinline__ double
unit(uint64_t h)
{
	return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

template <typename R, int NPART>
__global__ void __launch_bounds__(WG)
k_fill_random(const seg_desc *segs, uint32_t nseg, uint64_t seed, double lo,
    double hi, double wlo, double whi)
{
	uint64_t chunk = blockIdx.x;
	uint32_t j = find_job(segs, nseg, chunk);
	const seg_desc s = segs[j];
	uint64_t base = (chunk - s.first_chunk) * CHUNK;
	R *p = (R *)(uintptr_t)s.ptr;
	uint64_t key = mix64(seed * 0x9e3779b97f4a7c15ull + s.aux);
	for (uint64_t i = base + threadIdx.x; i < s.n && i < base + CHUNK;
	    i += WG) {
		for (int c = 0; c < NPART; c++) {
			uint64_t ctr = 2 * (i * NPART + c);
			double u = lo + (hi - lo) * unit(mix64(key ^ ctr));
			if (whi != wlo || wlo != 0.0) {
				double w = wlo + (whi - wlo) *
				    unit(mix64(key ^ (ctr + 1)));
				u *= exp10(w);
			}
			p[i * NPART + c] = (R)u;
		}
	}
}

/* ---- generic permuted element map -------------------------------- */

/* Division by an invariant 32-bit integer through a multiply-high
 * (T. Granlund, P. L. Montgomery, PLDI 1994, figure 4.1). */
struct fastdiv {
	uint32_t d, m, s1, s2;
};

static fastdiv
make_fdiv(uint32_t d)
{
	fastdiv f;
	uint32_t l = 0;
	while ((1ull << l) < d)
		l++;
	f.d = d;
	f.m = (uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
	f.s1 = l < 1 ? l : 1;
	f.s2 = l > 1 ? l - 1 : 0;
	return f;
}

__device__ __forceinline__ uint32_t
fdiv_q(uint32_t n, const fastdiv &f)
{
	uint32_t t = __umulhi(f.m, n);
	return (t + ((n - t) >> f.s1)) >> f.s2;
}

/*
 * Device forms of one job.  The destination block is walked in its own
 * (canonical) order; the host merges dims the source walks as one run.
 *
 * Gather form (elem_job_dev, kernel k_elem): workgroup w of the job maps
 * destination elements [w*CHUNK, (w+1)*CHUNK); writes are contiguous and
 * reads follow the source strides.  Used where the source's fastest dim
 * is destination dim 0 (reads then come in runs of its extent) and for
 * absent sources.
 *
 * Tiled form (elem_tile_dev, kernel k_permute): destination dim 0 (extent
 * A) against the source's fastest dim t (extent B).  A unit is NS slabs of
 * one TA x TB tile, the slabs being consecutive values of the remaining
 * ("batch") dims: it is read in source order (runs of TB along dim t) into
 * LDS, then written in destination order (runs of TA along dim 0).  A
 * workgroup moves R consecutive units, the next unit's loads in flight in
 * registers while the current one is written out.  Without the tile, the
 * lines of a far-transposed source are fetched once per element by
 * workgroups spread over the XCDs' L2s.
 */
#define TILE_MAX 32			/* TA, TB <= 32 */
#define TILE_BYTES (64 * 33 * 8)	/* tile storage, rows padded to TB + 1 */
#define TILE_NS_MAX 128
#define TILE_PER 9			/* >= TILE_BYTES / 8 / WG elements */
#define TILE_R 8			/* units per workgroup, at most */
/* tile | two (source, destination) slab-base buffers */
#define TILE_LDS (TILE_BYTES + 4 * TILE_NS_MAX * 8)

struct elem_job_dev {
	uint64_t dst, src, n, first_chunk;
	uint32_t ndim, src_zero;
	fastdiv div[XM_ELEM_MAXDIM];	/* merged destination extents */
	uint64_t sstride[XM_ELEM_MAXDIM];	/* and their source strides */
	double f_re, f_im;
	fastdiv drow;			/* pitched destination: rows ... */
	uint64_t dld;			/* ... and leading dimension (0: none) */
};

struct elem_tile_dev {
	uint64_t dst, src, first_chunk;
	double f_re, f_im;
	uint32_t nbd;			/* batch dims */
	uint32_t A, B, TA, TB, NS, nta, ntb;
	uint32_t R, pad;		/* units per workgroup */
	uint64_t nbatch, nunits;
	uint64_t sa, sb, db;		/* source strides of dims 0 and t;
					 * destination stride of t */
	fastdiv f_tatb, f_tb, f_ta;
	fastdiv bdiv[XM_ELEM_MAXDIM];	/* batch extents, destination order */
	uint64_t bs[XM_ELEM_MAXDIM], bd[XM_ELEM_MAXDIM];
};

/* Job record -> LDS, one dword per thread. */
template <typename J>
__device__ __forceinline__ void
load_job(J *dst, const J *src)
{
	constexpr uint32_t W = sizeof(J) / 4;
	static_assert(W <= WG && sizeof(J) % 4 == 0, "job record");
	if (threadIdx.x < W)
		((uint32_t *)dst)[threadIdx.x] =
		    ((const uint32_t *)src)[threadIdx.x];
	__syncthreads();
}

__device__ __forceinline__ uint64_t
src_offset(const elem_job_dev &jb, uint32_t i)
{
	uint64_t off = 0;
	for (uint32_t d = 0; d < jb.ndim; d++) {
		uint32_t q = d + 1 < jb.ndim ? fdiv_q(i, jb.div[d]) : 0;
		uint32_t r = d + 1 < jb.ndim ? i - q * jb.div[d].d : i;
		off += (uint64_t)r * jb.sstride[d];
		i = q;
	}
	return off;
}

/* Destination offset of element i (a pitched destination's rows are
 * runs of drow.d elements, dld apart). */
__device__ __forceinline__ uint64_t
dst_offset(const elem_job_dev &jb, uint32_t i)
{
	if (!jb.dld)
		return i;
	const uint32_t q = fdiv_q(i, jb.drow);
	return (uint64_t)(i - q * jb.drow.d) + (uint64_t)q * jb.dld;
}

template <int OP> struct reads_dst {
	static constexpr bool v = OP == ELEM_AXPY || OP == ELEM_MUL ||
	    OP == ELEM_DIV || OP == ELEM_ACCUM;
};

/* The operation on one destination element: x is the source element (0
 * when the source is absent), d the old destination value. */
template <typename D, typename S, int OP>
__device__ __forceinline__ D
elem_op(D d, S x, bool present, S f, D alpha, int alpha_zero)
{
	if (OP == ELEM_GATHER) {
		return conv<D>(x);
	} else if (OP == ELEM_COPY_SCALED) {
		/* libxm src/xm.c:131-141: scale in b's type, convert */
		return present ? conv<D>(emul(x, f)) : mk<D>(0.0, 0.0);
	} else if (OP == ELEM_AXPY) {
		/* libxm src/xm.c:218-239 with src/scalar.c:193-239 */
		D tmp = present ? conv<D>(emul(x, f)) : mk<D>(0.0, 0.0);
		if (alpha_zero)
			return tmp;
		return eadd(emul(alpha, d), emul(mk<D>(1.0, 0.0), tmp));
	} else if (OP == ELEM_MUL) {
		/* libxm src/xm.c:306-318 with src/scalar.c:241-283 */
		return present ? emul(emul(conv<D>(f), d), conv<D>(x)) :
		    mk<D>(0.0, 0.0);
	} else if (OP == ELEM_DIV) {
		/* libxm src/xm.c:389-399 with src/scalar.c:285-327 */
		return ediv(d, emul(conv<D>(f), conv<D>(x)));
	} else {	/* ELEM_ACCUM */
		return present ? eadd(d, conv<D>(emul(x, f))) : d;
	}
}

template <typename D, typename S, int OP>
__global__ void __launch_bounds__(WG)
k_elem(const elem_job_dev *jobs, uint32_t njobs, double al_re, double al_im,
    int alpha_zero)
{
	const uint64_t chunk = blockIdx.x;
	const uint32_t j = find_job(jobs, njobs, chunk);
	__shared__ elem_job_dev jb;
	load_job(&jb, jobs + j);
	const uint64_t base = (chunk - jb.first_chunk) * CHUNK;
	D *dst = (D *)(uintptr_t)jb.dst;
	const S *src = (const S *)(uintptr_t)jb.src;
	const S f = mk<S>(jb.f_re, jb.f_im);
	const D alpha = mk<D>(al_re, al_im);
	const bool present = !jb.src_zero;
	for (uint64_t i = base + threadIdx.x; i < jb.n && i < base + CHUNK;
	    i += WG) {
		S x = mk<S>(0.0, 0.0);
		if (present)
			x = src[src_offset(jb, (uint32_t)i)];
		const uint64_t di = dst_offset(jb, (uint32_t)i);
		D d = mk<D>(0.0, 0.0);
		if (reads_dst<OP>::v)
			d = dst[di];
		dst[di] = elem_op<D, S, OP>(d, x, present, f, alpha, alpha_zero);
	}
}

/* Unit u of a tiled job: tile origin, extents and first slab. */
struct tile_unit {
	uint32_t a0, b0, ta, tb, ns;
	uint64_t s0;
};

__device__ __forceinline__ tile_unit
unit_of(const elem_tile_dev &jb, uint64_t u)
{
	tile_unit t;
	const uint32_t ia = (uint32_t)(u % jb.nta);
	const uint64_t r1 = u / jb.nta;
	const uint32_t ib = (uint32_t)(r1 % jb.ntb);
	t.s0 = r1 / jb.ntb * jb.NS;
	t.a0 = ia * jb.TA;
	t.b0 = ib * jb.TB;
	t.ta = min(jb.TA, jb.A - t.a0);
	t.tb = min(jb.TB, jb.B - t.b0);
	t.ns = (uint32_t)min((uint64_t)jb.NS, jb.nbatch - t.s0);
	return t;
}

/* Source and destination offsets of the unit's slabs into sb / db. */
__device__ __forceinline__ void
unit_bases(const elem_tile_dev &jb, const tile_unit &t, uint64_t *sb,
    uint64_t *db)
{
	for (uint32_t s = threadIdx.x; s < t.ns; s += WG) {
		uint32_t q = (uint32_t)(t.s0 + s);
		uint64_t so = t.a0 * jb.sa + t.b0 * jb.sb;
		uint64_t dof = t.a0 + t.b0 * jb.db;
		for (uint32_t k = 0; k < jb.nbd; k++) {
			const uint32_t qq = fdiv_q(q, jb.bdiv[k]);
			const uint32_t r = q - qq * jb.bdiv[k].d;
			so += r * jb.bs[k];
			dof += r * jb.bd[k];
			q = qq;
		}
		sb[s] = so;
		db[s] = dof;
	}
}

template <typename D, typename S, int OP>
__global__ void __launch_bounds__(WG)
k_permute(const elem_tile_dev *jobs, uint32_t njobs, double al_re,
    double al_im, int alpha_zero)
{
	extern __shared__ __align__(16) char tile_lds[];
	S *tile = (S *)tile_lds;
	uint64_t *bases = (uint64_t *)(tile_lds + TILE_BYTES);
	const uint64_t chunk = blockIdx.x;
	const uint32_t j = find_job(jobs, njobs, chunk);
	__shared__ elem_tile_dev jb;
	load_job(&jb, jobs + j);
	const uint64_t u0 = (chunk - jb.first_chunk) * jb.R;
	const uint32_t cnt = (uint32_t)min((uint64_t)jb.R, jb.nunits - u0);
	D *dst = (D *)(uintptr_t)jb.dst;
	const S *src = (const S *)(uintptr_t)jb.src;
	const S f = mk<S>(jb.f_re, jb.f_im);
	const D alpha = mk<D>(al_re, al_im);
	const uint32_t TA = jb.TA, TB = jb.TB, TATB = TA * TB;
	const uint32_t tot = jb.NS * TATB, row = TB + 1;

	/* loads of unit t (slab bases in sb) into v, source order */
	S v[TILE_PER];
	auto load = [&](const tile_unit &t, const uint64_t *sb) {
#pragma unroll
		for (int u = 0; u < TILE_PER; u++) {
			const uint32_t e = threadIdx.x + u * WG;
			const uint32_t s = fdiv_q(e, jb.f_tatb);
			const uint32_t r = e - s * TATB;
			const uint32_t a = fdiv_q(r, jb.f_tb);
			const uint32_t b = r - a * TB;
			v[u] = mk<S>(0.0, 0.0);
			if (e < tot && s < t.ns && a < t.ta && b < t.tb)
				v[u] = src[sb[s] + a * jb.sa + b * jb.sb];
		}
	};
	auto to_tile = [&]() {
#pragma unroll
		for (int u = 0; u < TILE_PER; u++) {
			const uint32_t e = threadIdx.x + u * WG;
			const uint32_t s = fdiv_q(e, jb.f_tatb);
			const uint32_t r = e - s * TATB;
			const uint32_t a = fdiv_q(r, jb.f_tb);
			const uint32_t b = r - a * TB;
			if (e < tot)
				tile[(s * TA + a) * row + b] = v[u];
		}
	};

	tile_unit cur = unit_of(jb, u0);
	unit_bases(jb, cur, bases, bases + TILE_NS_MAX);
	__syncthreads();
	load(cur, bases);
	for (uint32_t k = 0; k < cnt; k++) {
		uint64_t *sb = bases + (k & 1) * 2 * TILE_NS_MAX;
		uint64_t *db = sb + TILE_NS_MAX;
		uint64_t *nsb = bases + ((k + 1) & 1) * 2 * TILE_NS_MAX;
		to_tile();
		tile_unit nxt = cur;
		const bool more = k + 1 < cnt;
		if (more) {
			nxt = unit_of(jb, u0 + k + 1);
			unit_bases(jb, nxt, nsb, nsb + TILE_NS_MAX);
		}
		__syncthreads();
		if (more)
			load(nxt, nsb);
		/* destination order (a fastest) out of the tile */
		for (uint32_t e = threadIdx.x; e < tot; e += WG) {
			const uint32_t s = fdiv_q(e, jb.f_tatb);
			const uint32_t r = e - s * TATB;
			const uint32_t b = fdiv_q(r, jb.f_ta);
			const uint32_t a = r - b * TA;
			if (s < cur.ns && a < cur.ta && b < cur.tb) {
				const uint64_t i = db[s] + a + b * jb.db;
				D d = mk<D>(0.0, 0.0);
				if (reads_dst<OP>::v)
					d = dst[i];
				dst[i] = elem_op<D, S, OP>(d,
				    tile[(s * TA + a) * row + b], true, f, alpha,
				    alpha_zero);
			}
		}
		__syncthreads();
		cur = nxt;
	}
}

__device__ __forceinline__ cplx<double> widen(float x, float y)
{ return cplx<double>{(double)x * (double)y, 0.0}; }
__device__ __forceinline__ cplx<double> widen(double x, double y)
{ return cplx<double>{mulr(x, y), 0.0}; }
template <typename R> __device__ __forceinline__ cplx<double>
widen(cplx<R> x, cplx<R> y)
{
	return emul(cplx<double>{(double)x.re, (double)x.im},
	    cplx<double>{(double)y.re, (double)y.im});
}

/* Per-chunk partial dot product of two permuted blocks; partials[2*chunk]
 * holds the real part, partials[2*chunk+1] the imaginary part. */
template <typename S>
__global__ void __launch_bounds__(WG)
k_dot(const elem_job_dev *jobs, uint32_t njobs, double *partials)
{
	uint64_t chunk = blockIdx.x;
	uint32_t j = find_job(jobs, njobs, chunk);
	__shared__ elem_job_dev jb;
	__shared__ double red[2][WG];
	load_job(&jb, jobs + j);
	uint64_t base = (chunk - jb.first_chunk) * CHUNK;
	const S *a = (const S *)(uintptr_t)jb.dst;	/* a, canonical order */
	const S *b = (const S *)(uintptr_t)jb.src;	/* b, permuted */
	double sre = 0, sim = 0;
	for (uint64_t i = base + threadIdx.x; i < jb.n && i < base + CHUNK;
	    i += WG) {
		S x = a[i], y = b[src_offset(jb, (uint32_t)i)];
		cplx<double> p = widen(x, y);
		sre += p.re;
		sim += p.im;
	}
	red[0][threadIdx.x] = sre;
	red[1][threadIdx.x] = sim;
	__syncthreads();
	for (int w = WG / 2; w > 0; w /= 2) {
		if ((int)threadIdx.x < w) {
			red[0][threadIdx.x] += red[0][threadIdx.x + w];
			red[1][threadIdx.x] += red[1][threadIdx.x + w];
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		/* scaled by s_a * s_b (f), rounded per libxm src/xm.c:490-497 */
		cplx<double> s{jb.f_re, jb.f_im}, v{red[0][0], red[1][0]};
		cplx<double> r = emul(s, v);
		partials[2 * chunk] = r.re;
		partials[2 * chunk + 1] = r.im;
	}
}

/* ---- launchers ----------------------------------------------------- */

static void
check_launch(const char *fn)
{
	hipError_t e = hipGetLastError();
	if (e != hipSuccess)
		xm_fatal("%s: kernel launch failed: %s", fn, hipGetErrorString(e));
}

void
launch_fill_value(const seg_desc *dsegs, size_t nseg, size_t nchunks,
    xm_scalar_type_t type, xm_scalar_t v, hipStream_t s)
{
	if (nchunks == 0)
		return;
	bump_data_epoch();
	dim3 g((unsigned)nchunks), b(WG);
	switch (type) {
	case XM_SCALAR_FLOAT:
		k_fill_value<float><<<g, b, 0, s>>>(dsegs, nseg, (float)v.real());
		break;
	case XM_SCALAR_FLOAT_COMPLEX:
		k_fill_value<cplx<float>><<<g, b, 0, s>>>(dsegs, nseg,
		    cplx<float>{(float)v.real(), (float)v.imag()});
		break;
	case XM_SCALAR_DOUBLE:
		k_fill_value<double><<<g, b, 0, s>>>(dsegs, nseg, v.real());
		break;
	case XM_SCALAR_DOUBLE_COMPLEX:
		k_fill_value<cplx<double>><<<g, b, 0, s>>>(dsegs, nseg,
		    cplx<double>{v.real(), v.imag()});
		break;
	}
	check_launch("xm_set");
}

void
launch_fill_random(const seg_desc *dsegs, size_t nseg, size_t nchunks,
    xm_scalar_type_t type, uint64_t seed, double lo, double hi, double wlo,
    double whi, hipStream_t s)
{
	if (nchunks == 0)
		return;
	bump_data_epoch();
	dim3 g((unsigned)nchunks), b(WG);
	switch (type) {
	case XM_SCALAR_FLOAT:
		k_fill_random<float, 1><<<g, b, 0, s>>>(dsegs, nseg, seed, lo, hi, wlo, whi);
		break;
	case XM_SCALAR_FLOAT_COMPLEX:
		k_fill_random<float, 2><<<g, b, 0, s>>>(dsegs, nseg, seed, lo, hi, wlo, whi);
		break;
	case XM_SCALAR_DOUBLE:
		k_fill_random<double, 1><<<g, b, 0, s>>>(dsegs, nseg, seed, lo, hi, wlo, whi);
		break;
	case XM_SCALAR_DOUBLE_COMPLEX:
		k_fill_random<double, 2><<<g, b, 0, s>>>(dsegs, nseg, seed, lo, hi, wlo, whi);
		break;
	}
	check_launch("xm_amd_fill_random");
}

/*
 * Tiled form of a job (dims/str/dstr: merged destination extents, their
 * source and destination strides), or false where the source's fastest
 * dim is destination dim 0 (the gather form reads runs then) or
 * destination dim 0 is not contiguous.  Near transposes (dims 0..t within
 * one 4096-element chunk) ran 4.2-4.3 TB/s gathered and 4.4-4.7 tiled on
 * 4-D FP64 copies; far ones 0.8-2.6 gathered.
 */
static bool
plan_tiled(elem_tile_dev &b, const uint32_t *dims, const uint64_t *str,
    const uint64_t *dstr, uint32_t nd, size_t esz)
{
	uint32_t t = 0;
	for (uint32_t d = 1; d < nd; d++)
		if (str[d] < str[t])
			t = d;
	if (t == 0 || dstr[0] != 1)
		return false;
	b.A = dims[0];
	b.B = dims[t];
	b.nta = (b.A + TILE_MAX - 1) / TILE_MAX;
	b.TA = (b.A + b.nta - 1) / b.nta;
	b.ntb = (b.B + TILE_MAX - 1) / TILE_MAX;
	b.TB = (b.B + b.ntb - 1) / b.ntb;
	b.sa = str[0];
	b.sb = str[t];
	b.db = dstr[t];
	b.nbd = 0;
	b.nbatch = 1;
	for (uint32_t d = 1; d < nd; d++) {
		if (d == t)
			continue;
		b.bdiv[b.nbd] = make_fdiv(dims[d]);
		b.bs[b.nbd] = str[d];
		b.bd[b.nbd] = dstr[d];
		b.nbd++;
		b.nbatch *= dims[d];
	}
	for (uint32_t k = b.nbd; k < XM_ELEM_MAXDIM; k++)
		b.bdiv[k] = make_fdiv(1);
	/* slabs per unit: within the tile storage and the TILE_PER
	 * elements a thread holds in flight */
	uint64_t ns = std::min<uint64_t>(
	    (TILE_BYTES / esz) / ((uint64_t)b.TA * (b.TB + 1)),
	    (uint64_t)TILE_PER * WG / ((uint64_t)b.TA * b.TB));
	ns = std::min<uint64_t>(std::max<uint64_t>(ns, 1), TILE_NS_MAX);
	b.NS = (uint32_t)std::min<uint64_t>(ns, b.nbatch);
	b.f_tatb = make_fdiv(b.TA * b.TB);
	b.f_tb = make_fdiv(b.TB);
	b.f_ta = make_fdiv(b.TA);
	return true;
}

/*
 * Device records of a job list: gather records, then (at l->toff bytes)
 * tiled records, each list with its own workgroup prefix.  esz: source
 * element size (tile storage); tiling false keeps every job a gather.
 */
static std::vector<char>
to_dev(const elem_job *jobs, size_t n, size_t esz, bool tiling,
    elem_launch *l)
{
	std::vector<elem_job_dev> g;
	std::vector<elem_tile_dev> t;
	g.reserve(n);
	uint64_t cg = 0, ct = 0;
	for (size_t i = 0; i < n; i++) {
		const elem_job &a = jobs[i];
		if (a.n > 0xffffffffull)
			XM_FATAL("libxm", "block of more than 2^32 elements");
		/* merge dimension d + 1 into d where the source and the
		 * destination walk them as one run (stride[d+1] = stride[d] *
		 * dims[d] on both sides): an identity permutation becomes one
		 * dimension, one multiply per element.  dstr: destination
		 * strides (a pitched destination's column dims step by dld) */
		const bool pitched = a.drows != 0 && a.drows < a.n;
		if (pitched && (a.drows > 0xffffffffull || a.dld < a.drows))
			XM_FATAL("libxm", "bad destination pitch");
		uint32_t dims[XM_ELEM_MAXDIM];
		uint64_t str[XM_ELEM_MAXDIM], dstr[XM_ELEM_MAXDIM];
		uint32_t nd = 0;
		uint64_t cum = 1;	/* destination elements before dim d */
		for (uint32_t d = 0; d < a.ndim; d++) {
			if (pitched && cum < a.drows && cum * a.dims[d] > a.drows)
				XM_FATAL("libxm", "destination pitch splits a dim");
			const uint64_t ds = !pitched || cum < a.drows ? cum :
			    cum / a.drows * a.dld;
			if (a.dims[d] == 1)
				continue;	/* contributes offset 0 */
			cum *= a.dims[d];
			if (nd > 0 && str[nd - 1] * dims[nd - 1] == a.sstride[d] &&
			    dstr[nd - 1] * dims[nd - 1] == ds &&
			    (uint64_t)dims[nd - 1] * a.dims[d] <= 0xffffffffu) {
				dims[nd - 1] *= a.dims[d];
				continue;
			}
			dims[nd] = a.dims[d];
			dstr[nd] = ds;
			str[nd++] = a.sstride[d];
		}
		elem_tile_dev tb;
		memset(&tb, 0, sizeof tb);
		if (tiling && !a.src_zero && nd >= 2 &&
		    plan_tiled(tb, dims, str, dstr, nd, esz)) {
			tb.dst = a.dst;
			tb.src = a.src;
			tb.f_re = a.f_re;
			tb.f_im = a.f_im;
			tb.nunits = (uint64_t)tb.nta * tb.ntb *
			    ((tb.nbatch + tb.NS - 1) / tb.NS);
			ct += tb.nunits;
			t.push_back(tb);
			continue;
		}
		elem_job_dev b;
		memset(&b, 0, sizeof b);
		b.dst = a.dst;
		b.src = a.src;
		b.n = a.n;
		b.src_zero = a.src_zero;
		b.ndim = nd;
		for (uint32_t d = 0; d < XM_ELEM_MAXDIM; d++) {
			b.div[d] = make_fdiv(d < nd && dims[d] ? dims[d] : 1);
			b.sstride[d] = d < nd ? str[d] : 0;
		}
		b.f_re = a.f_re;
		b.f_im = a.f_im;
		b.drow = make_fdiv(pitched ? (uint32_t)a.drows : 1);
		b.dld = pitched ? a.dld : 0;
		b.first_chunk = cg;
		cg += (a.n + CHUNK - 1) / CHUNK;
		g.push_back(b);
	}
	/* units per workgroup: as many as keep 4096 workgroups (16 per CU)
	 * in the launch, at most TILE_R */
	const uint32_t R = (uint32_t)std::max<uint64_t>(1,
	    std::min<uint64_t>(TILE_R, ct / 4096));
	uint64_t cw = 0;
	for (auto &tb : t) {
		tb.R = R;
		tb.first_chunk = cw;
		cw += (tb.nunits + R - 1) / R;
	}
	l->ngather = g.size();
	l->gchunks = cg;
	l->ntiled = t.size();
	l->tchunks = cw;
	l->toff = (g.size() * sizeof(elem_job_dev) + 255) / 256 * 256;
	std::vector<char> out(l->toff + t.size() * sizeof(elem_tile_dev), 0);
	if (!g.empty())
		memcpy(out.data(), g.data(), g.size() * sizeof(elem_job_dev));
	if (!t.empty())
		memcpy(out.data() + l->toff, t.data(),
		    t.size() * sizeof(elem_tile_dev));
	return out;
}

template <typename D, typename S, int OP>
static void
elem_launch_op(const char *dj, const elem_launch &l, double ar, double ai,
    int az, hipStream_t s)
{
	if (l.gchunks)
		k_elem<D, S, OP><<<dim3((unsigned)l.gchunks), dim3(WG), 0, s>>>(
		    (const elem_job_dev *)dj, (uint32_t)l.ngather, ar, ai, az);
	if (l.tchunks)
		k_permute<D, S, OP><<<dim3((unsigned)l.tchunks), dim3(WG),
		    TILE_LDS, s>>>((const elem_tile_dev *)(dj + l.toff),
		    (uint32_t)l.ntiled, ar, ai, az);
}

template <typename D, typename S>
static void
elem_dispatch(int op, const char *dj, const elem_launch &l,
    xm_scalar_t alpha, hipStream_t s)
{
	double ar = alpha.real(), ai = alpha.imag();
	int az = alpha == xm_scalar_t(0);
	switch (op) {
	case ELEM_COPY_SCALED:
		elem_launch_op<D, S, ELEM_COPY_SCALED>(dj, l, ar, ai, az, s);
		break;
	case ELEM_AXPY:
		elem_launch_op<D, S, ELEM_AXPY>(dj, l, ar, ai, az, s);
		break;
	case ELEM_MUL:
		elem_launch_op<D, S, ELEM_MUL>(dj, l, ar, ai, az, s);
		break;
	case ELEM_DIV:
		elem_launch_op<D, S, ELEM_DIV>(dj, l, ar, ai, az, s);
		break;
	case ELEM_GATHER:
		elem_launch_op<D, S, ELEM_GATHER>(dj, l, ar, ai, az, s);
		break;
	case ELEM_ACCUM:
		elem_launch_op<D, S, ELEM_ACCUM>(dj, l, ar, ai, az, s);
		break;
	}
}

std::vector<char>
elem_jobs_pack(const std::vector<elem_job> &jobs, size_t src_esz,
    elem_launch *l)
{
	return to_dev(jobs.data(), jobs.size(), src_esz,
	    elem_tiling_enabled(), l);
}

void
launch_elem_packed(const void *djobs, const elem_launch &l, int op,
    xm_scalar_type_t dt, xm_scalar_type_t st, xm_scalar_t alpha,
    hipStream_t s)
{
	if (l.gchunks == 0 && l.tchunks == 0)
		return;
	bump_data_epoch();
	const char *dj = (const char *)djobs;
#define PAIR(DT, ST, D, S) \
	if (dt == DT && st == ST) { elem_dispatch<D, S>(op, dj, l, alpha, s); \
		check_launch("libxm element kernel"); return; }
	PAIR(XM_SCALAR_DOUBLE, XM_SCALAR_DOUBLE, double, double)
	PAIR(XM_SCALAR_FLOAT, XM_SCALAR_FLOAT, float, float)
	PAIR(XM_SCALAR_DOUBLE_COMPLEX, XM_SCALAR_DOUBLE_COMPLEX, cplx<double>, cplx<double>)
	PAIR(XM_SCALAR_FLOAT_COMPLEX, XM_SCALAR_FLOAT_COMPLEX, cplx<float>, cplx<float>)
	PAIR(XM_SCALAR_DOUBLE, XM_SCALAR_FLOAT, double, float)
	PAIR(XM_SCALAR_FLOAT, XM_SCALAR_DOUBLE, float, double)
	PAIR(XM_SCALAR_DOUBLE_COMPLEX, XM_SCALAR_FLOAT_COMPLEX, cplx<double>, cplx<float>)
	PAIR(XM_SCALAR_FLOAT_COMPLEX, XM_SCALAR_DOUBLE_COMPLEX, cplx<float>, cplx<double>)
#undef PAIR
	XM_FATAL("libxm", "unsupported scalar conversion");
}

elem_launch
launch_elem(const elem_job *jobs, size_t njobs, int op,
    xm_scalar_type_t dt, xm_scalar_type_t st, xm_scalar_t alpha,
    hipStream_t s)
{
	elem_launch l;
	if (njobs == 0)
		return l;
	std::vector<char> hj = to_dev(jobs, njobs, elem_size(st),
	    elem_tiling_enabled(), &l);
	void *dj = upload(1, hj.data(), hj.size());
	launch_elem_packed(dj, l, op, dt, st, alpha, s);
	return l;
}

/* ---- adjoint folds -------------------------------------------------- */

struct fold_job_dev {
	uint64_t dst, n, first_chunk;
	uint32_t ndim, nsrc;
	fastdiv div[XM_ELEM_MAXDIM];	/* merged destination extents */
	uint64_t src[FOLD_MAXSRC];
	uint64_t sstride[FOLD_MAXSRC][XM_ELEM_MAXDIM];
	double f_re[FOLD_MAXSRC], f_im[FOLD_MAXSRC];
};

/*
 * Gather form over the destination (a canonical block): workgroup w of a
 * job owns destination elements [w*CHUNK, (w+1)*CHUNK); a thread splits
 * its element index once into the merged dims and adds every source's
 * element in the job's order, with the arithmetic of ELEM_ACCUM, so the
 * result equals that of one ELEM_ACCUM launch per source.  The loop over
 * sources is unrolled to FOLD_MAXSRC (uniform predicates) so the offsets
 * stay in registers; the loop over dims is not, so the strides are read
 * from LDS as needed instead of being held in registers (held, they took
 * over 200 VGPRs and left two waves per SIMD).
 */
template <typename T>
__global__ void __launch_bounds__(WG)
k_fold(const fold_job_dev *jobs, uint32_t njobs)
{
	const uint64_t chunk = blockIdx.x;
	const uint32_t j = find_job(jobs, njobs, chunk);
	__shared__ fold_job_dev jb;
	load_job(&jb, jobs + j);
	const uint64_t base = (chunk - jb.first_chunk) * CHUNK;
	T *dst = (T *)(uintptr_t)jb.dst;
	for (uint64_t i = base + threadIdx.x; i < jb.n && i < base + CHUNK;
	    i += WG) {
		uint64_t off[FOLD_MAXSRC] = {};
		uint32_t q = (uint32_t)i;
#pragma unroll 1
		for (int d = 0; d < (int)jb.ndim; d++) {
			const bool last = d + 1 == (int)jb.ndim;
			const uint32_t qn = last ? 0 : fdiv_q(q, jb.div[d]);
			const uint32_t r = last ? q : q - qn * jb.div[d].d;
#pragma unroll
			for (int k = 0; k < FOLD_MAXSRC; k++)
				if (k < (int)jb.nsrc)
					off[k] += (uint64_t)r * jb.sstride[k][d];
			q = qn;
		}
		T v = dst[i];
#pragma unroll
		for (int k = 0; k < FOLD_MAXSRC; k++)
			if (k < (int)jb.nsrc) {
				const T x = ((const T *)(uintptr_t)jb.src[k])[off[k]];
				v = eadd(v, conv<T>(emul(x,
				    mk<T>(jb.f_re[k], jb.f_im[k]))));
			}
		dst[i] = v;
	}
}

/* Device records of fold jobs (destination dims merged where every source
 * walks them as one run) and the workgroup count of their launch. */
std::vector<char>
fold_jobs_pack(const std::vector<fold_job> &jobs, size_t *nchunks)
{
	std::vector<fold_job_dev> out;
	out.reserve(jobs.size());
	uint64_t c = 0;
	for (const fold_job &a : jobs) {
		if (a.n > 0xffffffffull)
			XM_FATAL("libxm", "block of more than 2^32 elements");
		if (a.nsrc == 0 || a.nsrc > FOLD_MAXSRC || a.n == 0)
			XM_FATAL("libxm", "invalid fold job");
		uint32_t dims[XM_ELEM_MAXDIM];
		uint64_t str[FOLD_MAXSRC][XM_ELEM_MAXDIM];
		uint32_t nd = 0;
		for (uint32_t d = 0; d < a.ndim; d++) {
			if (a.dims[d] == 1)
				continue;	/* contributes offset 0 */
			bool run = nd > 0 &&
			    (uint64_t)dims[nd - 1] * a.dims[d] <= 0xffffffffu;
			for (uint32_t k = 0; run && k < a.nsrc; k++)
				run = str[k][nd - 1] * dims[nd - 1] ==
				    a.sstride[k][d];
			if (run) {
				dims[nd - 1] *= a.dims[d];
				continue;
			}
			dims[nd] = a.dims[d];
			for (uint32_t k = 0; k < a.nsrc; k++)
				str[k][nd] = a.sstride[k][d];
			nd++;
		}
		fold_job_dev b;
		memset(&b, 0, sizeof b);
		b.dst = a.dst;
		b.n = a.n;
		b.first_chunk = c;
		b.ndim = nd;
		b.nsrc = a.nsrc;
		for (uint32_t d = 0; d < XM_ELEM_MAXDIM; d++)
			b.div[d] = make_fdiv(d < nd ? dims[d] : 1);
		for (uint32_t k = 0; k < a.nsrc; k++) {
			b.src[k] = a.src[k];
			for (uint32_t d = 0; d < nd; d++)
				b.sstride[k][d] = str[k][d];
			b.f_re[k] = a.f_re[k];
			b.f_im[k] = a.f_im[k];
		}
		c += (a.n + CHUNK - 1) / CHUNK;
		out.push_back(b);
	}
	*nchunks = c;
	std::vector<char> bytes(out.size() * sizeof(fold_job_dev));
	if (!out.empty())
		memcpy(bytes.data(), out.data(), bytes.size());
	return bytes;
}

void
launch_fold_packed(const void *djobs, size_t njobs, size_t nchunks,
    xm_scalar_type_t type, hipStream_t s)
{
	if (njobs == 0 || nchunks == 0)
		return;
	bump_data_epoch();
	const fold_job_dev *dj = (const fold_job_dev *)djobs;
	const dim3 g((unsigned)nchunks), b(WG);
	const uint32_t nj = (uint32_t)njobs;
	switch (type) {
	case XM_SCALAR_FLOAT: k_fold<float><<<g, b, 0, s>>>(dj, nj); break;
	case XM_SCALAR_FLOAT_COMPLEX: k_fold<cplx<float>><<<g, b, 0, s>>>(dj, nj); break;
	case XM_SCALAR_DOUBLE: k_fold<double><<<g, b, 0, s>>>(dj, nj); break;
	case XM_SCALAR_DOUBLE_COMPLEX: k_fold<cplx<double>><<<g, b, 0, s>>>(dj, nj); break;
	}
	check_launch("xm_contract_backward fold");
}

/* XM_ELEM_TILED=0 keeps every permute in the gather form. */
bool
elem_tiling_enabled(void)
{
	static int on = -1;
	if (on < 0) {
		const char *e = getenv("XM_ELEM_TILED");
		on = !(e && *e == '0');
	}
	return on != 0;
}

/* The device wall clock (constant rate, hipDeviceAttributeWallClockRate
 * kHz) when the stream reaches this launch. */
__global__ void
k_stamp(uint64_t *slot)
{
	if (threadIdx.x == 0)
		*slot = wall_clock64();
}

void
launch_stamp(uint64_t *slot, hipStream_t s)
{
	k_stamp<<<1, 64, 0, s>>>(slot);
	check_launch("libxm stamp kernel");
}

void
launch_dot(const elem_job *jobs, size_t njobs, size_t nchunks,
    xm_scalar_type_t type, double *dpartials, hipStream_t s)
{
	if (nchunks == 0 || njobs == 0)
		return;
	elem_launch l;
	std::vector<char> hj = to_dev(jobs, njobs, elem_size(type), false, &l);
	if (l.gchunks != nchunks)
		XM_FATAL("xm_dot", "chunk count mismatch");
	elem_job_dev *dj = (elem_job_dev *)upload(1, hj.data(), hj.size());
	dim3 g((unsigned)nchunks), b(WG);
	switch (type) {
	case XM_SCALAR_FLOAT: k_dot<float><<<g, b, 0, s>>>(dj, njobs, dpartials); break;
	case XM_SCALAR_FLOAT_COMPLEX: k_dot<cplx<float>><<<g, b, 0, s>>>(dj, njobs, dpartials); break;
	case XM_SCALAR_DOUBLE: k_dot<double><<<g, b, 0, s>>>(dj, njobs, dpartials); break;
	case XM_SCALAR_DOUBLE_COMPLEX: k_dot<cplx<double>><<<g, b, 0, s>>>(dj, njobs, dpartials); break;
	}
	check_launch("xm_dot");
}
