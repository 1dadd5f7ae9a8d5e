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

#define CHUNK 4096
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

/* Device form of one job: extents as fast divisors. */
struct elem_job_dev {
	uint64_t dst, src, n, first_chunk;
	uint32_t ndim, src_zero;
	fastdiv div[XM_ELEM_MAXDIM];
	uint64_t sstride[XM_ELEM_MAXDIM];
	double f_re, f_im;
};

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

template <typename D, typename S, int OP>
__global__ void __launch_bounds__(WG)
k_elem(const elem_job_dev *jobs, uint32_t njobs, double al_re, double al_im,
    int alpha_zero)
{
	uint64_t chunk = blockIdx.x;
	uint32_t j = find_job(jobs, njobs, chunk);
	__shared__ elem_job_dev jb;
	if (threadIdx.x == 0)
		jb = jobs[j];
	__syncthreads();
	uint64_t base = (chunk - jb.first_chunk) * CHUNK;
	D *dst = (D *)(uintptr_t)jb.dst;
	const S *src = (const S *)(uintptr_t)jb.src;
	const S f = mk<S>(jb.f_re, jb.f_im);
	const D alpha = mk<D>(al_re, al_im);
	for (uint64_t i = base + threadIdx.x; i < jb.n && i < base + CHUNK;
	    i += WG) {
		S x = mk<S>(0.0, 0.0);
		if (!jb.src_zero)
			x = src[src_offset(jb, (uint32_t)i)];
		if (OP == ELEM_GATHER) {
			dst[i] = conv<D>(x);
		} else if (OP == ELEM_COPY_SCALED) {
			/* libxm src/xm.c:131-141: scale in b's type, convert */
			dst[i] = jb.src_zero ? mk<D>(0.0, 0.0) :
			    conv<D>(emul(x, f));
		} else if (OP == ELEM_AXPY) {
			/* libxm src/xm.c:218-239 with src/scalar.c:193-239 */
			D tmp = jb.src_zero ? mk<D>(0.0, 0.0) :
			    conv<D>(emul(x, f));
			if (alpha_zero)
				dst[i] = tmp;
			else
				dst[i] = eadd(emul(alpha, dst[i]),
				    emul(mk<D>(1.0, 0.0), tmp));
		} else if (OP == ELEM_MUL) {
			/* libxm src/xm.c:306-318 with src/scalar.c:241-283 */
			dst[i] = jb.src_zero ? mk<D>(0.0, 0.0) :
			    emul(emul(conv<D>(f), dst[i]), conv<D>(x));
		} else if (OP == ELEM_DIV) {
			/* libxm src/xm.c:389-399 with src/scalar.c:285-327 */
			dst[i] = ediv(dst[i], emul(conv<D>(f), conv<D>(x)));
		} else if (OP == ELEM_ACCUM) {
			if (!jb.src_zero)
				dst[i] = eadd(dst[i], conv<D>(emul(x, f)));
		}
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
	if (threadIdx.x == 0)
		jb = jobs[j];
	__syncthreads();
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

static std::vector<elem_job_dev>
to_dev(const elem_job *jobs, size_t n)
{
	std::vector<elem_job_dev> out(n);
	for (size_t i = 0; i < n; i++) {
		const elem_job &a = jobs[i];
		elem_job_dev &b = out[i];
		b.dst = a.dst;
		b.src = a.src;
		b.n = a.n;
		b.first_chunk = a.first_chunk;
		b.src_zero = a.src_zero;
		/* merge dimension d + 1 into d where the source walks them as
		 * one run (stride[d+1] = stride[d] * dims[d]): an identity
		 * permutation becomes one dimension, one multiply per element */
		uint32_t dims[XM_ELEM_MAXDIM];
		uint64_t str[XM_ELEM_MAXDIM];
		uint32_t nd = 0;
		for (uint32_t d = 0; d < a.ndim; d++) {
			if (a.dims[d] == 1)
				continue;	/* contributes offset 0 */
			if (nd > 0 && str[nd - 1] * dims[nd - 1] == a.sstride[d] &&
			    (uint64_t)dims[nd - 1] * a.dims[d] <= 0xffffffffu) {
				dims[nd - 1] *= a.dims[d];
				continue;
			}
			dims[nd] = a.dims[d];
			str[nd++] = a.sstride[d];
		}
		b.ndim = nd;
		for (uint32_t d = 0; d < XM_ELEM_MAXDIM; d++) {
			b.div[d] = make_fdiv(d < nd && dims[d] ? dims[d] : 1);
			b.sstride[d] = d < nd ? str[d] : 0;
		}
		b.f_re = a.f_re;
		b.f_im = a.f_im;
	}
	return out;
}

template <typename D, typename S>
static void
elem_dispatch(int op, const elem_job_dev *dj, uint32_t nj, size_t nchunks,
    xm_scalar_t alpha, hipStream_t s)
{
	dim3 g((unsigned)nchunks), b(WG);
	double ar = alpha.real(), ai = alpha.imag();
	int az = alpha == xm_scalar_t(0);
	switch (op) {
	case ELEM_COPY_SCALED:
		k_elem<D, S, ELEM_COPY_SCALED><<<g, b, 0, s>>>(dj, nj, ar, ai, az);
		break;
	case ELEM_AXPY:
		k_elem<D, S, ELEM_AXPY><<<g, b, 0, s>>>(dj, nj, ar, ai, az);
		break;
	case ELEM_MUL:
		k_elem<D, S, ELEM_MUL><<<g, b, 0, s>>>(dj, nj, ar, ai, az);
		break;
	case ELEM_DIV:
		k_elem<D, S, ELEM_DIV><<<g, b, 0, s>>>(dj, nj, ar, ai, az);
		break;
	case ELEM_GATHER:
		k_elem<D, S, ELEM_GATHER><<<g, b, 0, s>>>(dj, nj, ar, ai, az);
		break;
	case ELEM_ACCUM:
		k_elem<D, S, ELEM_ACCUM><<<g, b, 0, s>>>(dj, nj, ar, ai, az);
		break;
	}
}

std::vector<char>
elem_jobs_pack(const std::vector<elem_job> &jobs)
{
	std::vector<elem_job_dev> hj = to_dev(jobs.data(), jobs.size());
	std::vector<char> out(hj.size() * sizeof(elem_job_dev));
	if (!out.empty())
		memcpy(out.data(), hj.data(), out.size());
	return out;
}

void
launch_elem_packed(const void *djobs, size_t njobs, size_t nchunks, int op,
    xm_scalar_type_t dt, xm_scalar_type_t st, xm_scalar_t alpha,
    hipStream_t s)
{
	if (nchunks == 0 || njobs == 0)
		return;
	bump_data_epoch();
	const elem_job_dev *dj = (const elem_job_dev *)djobs;
	uint32_t nj = (uint32_t)njobs;
#define PAIR(DT, ST, D, S) \
	if (dt == DT && st == ST) { elem_dispatch<D, S>(op, dj, nj, nchunks, alpha, s); \
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

void
launch_elem(const elem_job *jobs, size_t njobs, size_t nchunks, int op,
    xm_scalar_type_t dt, xm_scalar_type_t st, xm_scalar_t alpha,
    hipStream_t s)
{
	if (nchunks == 0 || njobs == 0)
		return;
	std::vector<elem_job_dev> hj = to_dev(jobs, njobs);
	void *dj = upload(1, hj.data(), hj.size() * sizeof(elem_job_dev));
	launch_elem_packed(dj, njobs, nchunks, op, dt, st, alpha, s);
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
	std::vector<elem_job_dev> hj = to_dev(jobs, njobs);
	elem_job_dev *dj = (elem_job_dev *)upload(1, hj.data(),
	    hj.size() * sizeof(elem_job_dev));
	dim3 g((unsigned)nchunks), b(WG);
	switch (type) {
	case XM_SCALAR_FLOAT: k_dot<float><<<g, b, 0, s>>>(dj, njobs, dpartials); break;
	case XM_SCALAR_FLOAT_COMPLEX: k_dot<cplx<float>><<<g, b, 0, s>>>(dj, njobs, dpartials); break;
	case XM_SCALAR_DOUBLE: k_dot<double><<<g, b, 0, s>>>(dj, njobs, dpartials); break;
	case XM_SCALAR_DOUBLE_COMPLEX: k_dot<cplx<double>><<<g, b, 0, s>>>(dj, njobs, dpartials); break;
	}
	check_launch("xm_dot");
}
