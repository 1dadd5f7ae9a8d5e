/*
 * split.hip - FP64 contraction emulated on the BF16 matrix cores
 * (Ozaki-style splitting; K. Ozaki, T. Ogita, S. Oishi, S. M. Rump,
 * "Error-free transformations of matrix multiplication by using fast
 * routines of matrix multiplication", Numer. Algorithms 59, 2012).
 *
 * Slicing (k_slice), once per call for every distinct operand panel:
 *   for each row r and K-chunk c of KC = 128 columns, e = exponent of the
 *   chunk's largest |x| (|x| < 2^e), x~ = x * 2^-e in (-1, 1), and
 *     y1 = x~ * 2^8,        d1 = rint(y1),  r1 = y1 - d1  (|r1| <= 1/2)
 *     yt = r(t-1) * 2^9,    dt = rint(yt),  rt = yt - dt
 *   so x~ = sum_t dt * 2^-(8 + 9(t-1)) + O(2^-(8+9s)), every step exact in
 *   FP64 and every digit |dt| <= 256 exact in BF16.  Slice t is stored
 *   pre-scaled, dt * 2^-(8+9(t-1)), in the lane order of the
 *   v_mfma_f32_32x32x16_bf16 operand (32 rows x 16 k per 1 KiB fragment).
 *
 * GEMM (k_gemm_split): per 16-wide k step, the products of slices t, u
 * with t + u <= s + 1 go to one FP32 accumulator per weight class
 * d = t + u: d = 2, 3, 4 and d >= 5.  All products of one class are
 * multiples of the same unit (2^-16, 2^-25, 2^-34), so over one chunk of
 * 128 k the d = 2 and d = 3 sums are exact (partial sums stay below 2^24
 * units) and the d = 4 sum is exact unless its partial sums reach 2^24
 * units, ~40 standard deviations above their typical size for random
 * data.  The d >= 5 class weighs <= 2^-27, so its FP32 rounding stays
 * below 2^-50 of the result.  (One shared accumulator for all d >= 4
 * would round at the unit of the smallest class and lose ~2^-42 sqrt(n).)
 * At every chunk end the classes are recombined in FP64, smallest first:
 * acc += al * 2^(ex + ey) * (((a5 + a4) + a3) + a2).
 * Exponents, scaling, recombination, alpha/beta and the C update are FP64.
 */
#include "contract_plan.h"

#include <algorithm>
#include <map>
#include <tuple>

#define KC 128
#define SPLIT_WG 256
#define MAXS 8

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

struct split_operand {
	uint64_t src;		/* FP64 panel */
	uint32_t mode, ld, rows, k;
	uint32_t rows_pad, kpad, nchunk, pad;
	uint64_t slices;	/* nslices planes of rows_pad * kpad bf16 */
	uint64_t exps;		/* rows_pad * nchunk int32 */
	uint64_t first_wg;	/* prefix of (rows_pad/32) * nchunk */
};

struct split_pair {
	uint64_t xs, ys;	/* slice plane 0 of X and Y */
	uint64_t xe, ye;	/* exponents of X and Y */
	uint64_t xplane, yplane;	/* elements per slice plane */
	uint32_t k, nchunk, xrows_pad, yrows_pad;
	uint32_t nks, pad;	/* 16-wide k steps in kpad */
	double a_re;
};

struct split_state {
	std::vector<split_operand> ops;
	std::vector<split_pair> pairs;
	uint64_t nwg = 0;
	void *mem = nullptr;	/* slices + exponents */
	void *desc = nullptr;	/* operands | pairs */
	split_operand *d_ops = nullptr;
	split_pair *d_pairs = nullptr;
	int nslices = 5;
};

/* ---- slicing ------------------------------------------------------ */

__device__ __forceinline__ uint16_t
bf16_bits_exact(double v)
{
	/* v has at most 9 significant bits: exact in FP32 and in BF16 */
	float f = (float)v;
	return (uint16_t)(__float_as_uint(f) >> 16);
}

template <int S>
__global__ void __launch_bounds__(SPLIT_WG)
k_slice(const split_operand *__restrict__ ops, uint32_t nops)
{
	__shared__ double tile[32][KC + 1];
	__shared__ double rmax[32][8];
	__shared__ uint32_t found;
	const uint64_t wg = blockIdx.x;
	if (threadIdx.x == 0) {
		uint32_t lo = 0, hi = nops - 1;
		while (lo < hi) {
			uint32_t mid = (lo + hi + 1) / 2;
			if (ops[mid].first_wg <= wg)
				lo = mid;
			else
				hi = mid - 1;
		}
		found = lo;
	}
	__syncthreads();
	const split_operand op = ops[found];
	const uint64_t local = wg - op.first_wg;
	const uint32_t rt = (uint32_t)(local / op.nchunk);
	const uint32_t c = (uint32_t)(local % op.nchunk);
	const double *X = (const double *)(uintptr_t)op.src;
	const int tid = threadIdx.x;

	/* load 32 rows x KC columns (zero outside the panel) */
	for (int e = tid; e < 32 * KC; e += SPLIT_WG) {
		int r, k;
		if (op.mode == PANEL_PCONTIG) {
			r = e % 32; k = e / 32;
		} else {
			k = e % KC; r = e / KC;
		}
		uint32_t gr = rt * 32 + r, gk = c * KC + k;
		double v = 0.0;
		if (gr < op.rows && gk < op.k)
			v = op.mode == PANEL_PCONTIG ?
			    X[(size_t)gr + (size_t)gk * op.ld] :
			    X[(size_t)gk + (size_t)gr * op.ld];
		tile[r][k] = v;
	}
	__syncthreads();
	/* row maxima: 8 threads per row */
	const int r = tid / 8, part = tid % 8;
	double m = 0.0;
	for (int k = part * 16; k < part * 16 + 16; k++)
		m = fmax(m, fabs(tile[r][k]));
	rmax[r][part] = m;
	__syncthreads();
	m = 0.0;
	for (int q = 0; q < 8; q++)
		m = fmax(m, rmax[r][q]);
	int ex = 0;
	if (m > 0.0)
		(void)frexp(m, &ex);
	const uint32_t grow = rt * 32 + r;
	if (part == 0)
		((int *)(uintptr_t)op.exps)[(size_t)grow * op.nchunk + c] = ex;

	/* this thread: row r, k in [c*KC + 16*part, +16) */
	const uint32_t k0 = c * KC + 16 * part;
	if (k0 >= op.kpad)
		return;
	const uint32_t nks = op.kpad / 16, ks = k0 / 16;
	const size_t plane = (size_t)op.rows_pad * op.kpad;
	uint16_t dig[S][16];
#pragma unroll
	for (int j = 0; j < 16; j++) {
		double y = ldexp(tile[r][16 * part + j], 8 - ex);
		double d = rint(y);
		dig[0][j] = bf16_bits_exact(ldexp(d, -8));
		double rem = y - d;
#pragma unroll
		for (int t = 1; t < S; t++) {
			y = rem * 512.0;
			d = rint(y);
			rem = y - d;
			dig[t][j] = bf16_bits_exact(ldexp(d, -8 - 9 * t));
		}
	}
	/* fragment order: ((rt * nks + ks) * 64 + lane) * 8 + j, lane = row +
	 * 32 * (k % 16 >= 8) */
	uint16_t *base = (uint16_t *)(uintptr_t)op.slices;
	const size_t f = ((size_t)rt * nks + ks) * 64;
#pragma unroll
	for (int t = 0; t < S; t++) {
		uint4 lo, hi;
		lo.x = dig[t][0] | ((uint32_t)dig[t][1] << 16);
		lo.y = dig[t][2] | ((uint32_t)dig[t][3] << 16);
		lo.z = dig[t][4] | ((uint32_t)dig[t][5] << 16);
		lo.w = dig[t][6] | ((uint32_t)dig[t][7] << 16);
		hi.x = dig[t][8] | ((uint32_t)dig[t][9] << 16);
		hi.y = dig[t][10] | ((uint32_t)dig[t][11] << 16);
		hi.z = dig[t][12] | ((uint32_t)dig[t][13] << 16);
		hi.w = dig[t][14] | ((uint32_t)dig[t][15] << 16);
		uint16_t *p = base + t * plane;
		*(uint4 *)(p + (f + r) * 8) = lo;
		*(uint4 *)(p + (f + r + 32) * 8) = hi;
	}
}

/* ---- GEMM ------------------------------------------------------------ */

__device__ __forceinline__ bf8
ld_frag(const uint16_t *p)
{
	uint4 v = *(const uint4 *)p;
	bf8 r;
	__builtin_memcpy(&r, &v, 16);
	return r;
}

/* 64 x 64 outputs per workgroup; wave w computes the 32 x 32 block
 * (p: (w & 1) * 32, q: (w >> 1) * 32) with one 32x32 MFMA tile. */
template <int S>
__global__ void __launch_bounds__(SPLIT_WG)
k_gemm_split(const gemm_tile *__restrict__ tiles,
    const gemm_cblock *__restrict__ cbs, const split_pair *__restrict__ pairs,
    const gemm_scalars *__restrict__ scp)
{
	const gemm_tile tl = tiles[blockIdx.x];
	if (tl.cblock == 0xffffffffu)
		return;
	const gemm_cblock cb = cbs[tl.cblock];
	const gemm_scalars sc = *scp;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const uint32_t pw = tl.p0 + (wave & 1) * 32, qw = tl.q0 + (wave >> 1) * 32;

	double acc[16];
#pragma unroll
	for (int i = 0; i < 16; i++)
		acc[i] = 0.0;
	const uint32_t pbeg = (sc.flags & 1) ? cb.pair_end : cb.pair_begin;
	if (pw < cb.p && qw < cb.q) {
		for (uint32_t ip = pbeg; ip < cb.pair_end; ip++) {
			const split_pair pr = pairs[ip];
			const double al = __dmul_rn(sc.alpha_re, pr.a_re);
			const uint16_t *xs = (const uint16_t *)(uintptr_t)pr.xs;
			const uint16_t *ys = (const uint16_t *)(uintptr_t)pr.ys;
			const int *xe = (const int *)(uintptr_t)pr.xe;
			const int *ye = (const int *)(uintptr_t)pr.ye;
			const size_t xf = (size_t)(pw / 32) * pr.nks * 64 + lane;
			const size_t yf = (size_t)(qw / 32) * pr.nks * 64 + lane;
			for (uint32_t c = 0; c < pr.nchunk; c++) {
				/* one FP32 accumulator per weight class d = t + u:
				 * d = 2, 3, 4 and d >= 5 */
				f16v a2 = {}, a3 = {}, a4 = {}, a5 = {};
				const uint32_t ks0 = c * (KC / 16);
				const uint32_t ks1 = min(ks0 + KC / 16, pr.nks);
				for (uint32_t ks = ks0; ks < ks1; ks++) {
					bf8 fx[S], fy[S];
#pragma unroll
					for (int t = 0; t < S; t++) {
						fx[t] = ld_frag(xs + t * pr.xplane +
						    (xf + (size_t)ks * 64) * 8);
						fy[t] = ld_frag(ys + t * pr.yplane +
						    (yf + (size_t)ks * 64) * 8);
					}
					/* D[i = q][j = p]: A = Y slice, B = X slice */
#pragma unroll
					for (int d = 2; d <= S + 1; d++)
#pragma unroll
						for (int t = 1; t < d; t++) {
							const int u = d - t;
							if (t > S || u > S)
								continue;
							if (d == 2)
								a2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
								    fy[u - 1], fx[t - 1], a2, 0, 0, 0);
							else if (d == 3)
								a3 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
								    fy[u - 1], fx[t - 1], a3, 0, 0, 0);
							else if (d == 4)
								a4 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
								    fy[u - 1], fx[t - 1], a4, 0, 0, 0);
							else
								a5 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
								    fy[u - 1], fx[t - 1], a5, 0, 0, 0);
						}
				}
				/* recombine, smallest weights first: lane -> p row,
				 * register -> q row */
				const int exp_p = xe[(size_t)(pw + (lane & 31)) *
				    pr.nchunk + c];
#pragma unroll
				for (int i = 0; i < 16; i++) {
					uint32_t qq = qw + 8 * (i / 4) + 4 * (lane >> 5) + (i % 4);
					int exp_q = ye[(size_t)qq * pr.nchunk + c];
					double v = (((double)a5[i] + (double)a4[i]) +
					    (double)a3[i]) + (double)a2[i];
					acc[i] = fma(ldexp(al, exp_p + exp_q), v, acc[i]);
				}
			}
		}
	}
	double *C = (double *)(uintptr_t)cb.c;
	const bool bz = sc.flags & 2, none = pbeg == cb.pair_end;
#pragma unroll
	for (int i = 0; i < 16; i++) {
		uint32_t p = pw + (lane & 31);
		uint32_t q = qw + 8 * (i / 4) + 4 * (lane >> 5) + (i % 4);
		if (p < cb.p && q < cb.q) {
			size_t o = (size_t)p + (size_t)q * cb.ldc;
			double v;
			if (bz)
				v = none ? 0.0 : acc[i];
			else if (none)
				v = __dmul_rn(sc.beta_re, C[o]);
			else
				v = __dadd_rn(__dmul_rn(sc.beta_re, C[o]), acc[i]);
			C[o] = v;
		}
	}
}

/* ---- host ---------------------------------------------------------- */

void
split_prepare(contract_plan *pl, const xm_tensor_t *, const xm_tensor_t *,
    const xm_tensor_t *, const masks_t &)
{
	const char *fn = "xm_contract";
	split_state *ss = new split_state;
	ss->nslices = pl->nslices;
	pl->split = ss;
	std::map<std::tuple<uint64_t, uint32_t, uint32_t, uint32_t, uint32_t>,
	    uint32_t> seen;
	auto opnd = [&](uint64_t src, uint32_t mode, uint32_t ld, uint32_t rows,
	    uint32_t k) -> uint32_t {
		auto key = std::make_tuple(src, mode, ld, rows, k);
		auto it = seen.find(key);
		if (it != seen.end())
			return it->second;
		split_operand o;
		memset(&o, 0, sizeof o);
		o.src = src;
		o.mode = mode;
		o.ld = ld;
		o.rows = rows;
		o.k = k;
		o.rows_pad = (rows + 31) / 32 * 32;
		o.kpad = (k + 15) / 16 * 16;
		o.nchunk = (o.kpad + KC - 1) / KC;
		uint32_t id = (uint32_t)ss->ops.size();
		ss->ops.push_back(o);
		seen[key] = id;
		return id;
	};
	std::vector<std::pair<uint32_t, uint32_t>> use(pl->pairs.size());
	for (size_t cbi = 0; cbi < pl->cblocks.size(); cbi++) {
		const gemm_cblock &cb = pl->cblocks[cbi];
		for (uint32_t i = cb.pair_begin; i < cb.pair_end; i++) {
			const gemm_pair &g = pl->pairs[i];
			use[i].first = opnd(g.x, g.xmode, g.ldx, cb.p, g.k);
			use[i].second = opnd(g.y, g.ymode, g.ldy, cb.q, g.k);
		}
	}
	/* storage: slices (16-byte aligned) then exponents */
	size_t total = 0;
	for (auto &o : ss->ops) {
		o.slices = total;
		total += (size_t)ss->nslices * o.rows_pad * o.kpad * 2;
		total = (total + 255) / 256 * 256;
	}
	for (auto &o : ss->ops) {
		o.exps = total;
		total += (size_t)o.rows_pad * o.nchunk * 4;
		total = (total + 255) / 256 * 256;
	}
	if (total)
		HIP_OK(fn, hipMalloc(&ss->mem, total));
	uint64_t base = (uint64_t)(uintptr_t)ss->mem, wgs = 0;
	for (auto &o : ss->ops) {
		o.slices += base;
		o.exps += base;
		o.first_wg = wgs;
		wgs += (uint64_t)(o.rows_pad / 32) * o.nchunk;
	}
	ss->nwg = wgs;
	for (size_t i = 0; i < pl->pairs.size(); i++) {
		const split_operand &x = ss->ops[use[i].first];
		const split_operand &y = ss->ops[use[i].second];
		split_pair sp;
		memset(&sp, 0, sizeof sp);
		sp.xs = x.slices;
		sp.ys = y.slices;
		sp.xe = x.exps;
		sp.ye = y.exps;
		sp.xplane = (uint64_t)x.rows_pad * x.kpad;
		sp.yplane = (uint64_t)y.rows_pad * y.kpad;
		sp.k = x.k;
		sp.nchunk = x.nchunk;
		sp.xrows_pad = x.rows_pad;
		sp.yrows_pad = y.rows_pad;
		sp.nks = x.kpad / 16;
		sp.a_re = pl->pairs[i].a_re;
		ss->pairs.push_back(sp);
	}
	size_t ob = ss->ops.size() * sizeof(split_operand);
	size_t pb = ss->pairs.size() * sizeof(split_pair);
	size_t ob_al = (ob + 255) / 256 * 256;
	HIP_OK(fn, hipMalloc(&ss->desc, std::max<size_t>(ob_al + pb, 256)));
	if (ob)
		HIP_OK(fn, hipMemcpy(ss->desc, ss->ops.data(), ob,
		    hipMemcpyHostToDevice));
	if (pb)
		HIP_OK(fn, hipMemcpy((char *)ss->desc + ob_al, ss->pairs.data(),
		    pb, hipMemcpyHostToDevice));
	ss->d_ops = (split_operand *)ss->desc;
	ss->d_pairs = (split_pair *)((char *)ss->desc + ob_al);
	pl->unfold_bytes += 0;
}

#define SDISPATCH(KERNEL, S, ...)					\
	switch (S) {							\
	case 3: KERNEL<3><<<__VA_ARGS__>>>; break;			\
	case 4: KERNEL<4><<<__VA_ARGS__>>>; break;			\
	case 5: KERNEL<5><<<__VA_ARGS__>>>; break;			\
	case 6: KERNEL<6><<<__VA_ARGS__>>>; break;			\
	case 7: KERNEL<7><<<__VA_ARGS__>>>; break;			\
	default: KERNEL<8><<<__VA_ARGS__>>>; break;			\
	}

void
split_slices(contract_plan *pl, hipStream_t s)
{
	split_state *ss = pl->split;
	if (ss == nullptr || ss->nwg == 0)
		return;
	uint32_t nops = (uint32_t)ss->ops.size();
	switch (ss->nslices) {
	case 3: k_slice<3><<<dim3((unsigned)ss->nwg), dim3(SPLIT_WG), 0, s>>>(ss->d_ops, nops); break;
	case 4: k_slice<4><<<dim3((unsigned)ss->nwg), dim3(SPLIT_WG), 0, s>>>(ss->d_ops, nops); break;
	case 5: k_slice<5><<<dim3((unsigned)ss->nwg), dim3(SPLIT_WG), 0, s>>>(ss->d_ops, nops); break;
	case 6: k_slice<6><<<dim3((unsigned)ss->nwg), dim3(SPLIT_WG), 0, s>>>(ss->d_ops, nops); break;
	case 7: k_slice<7><<<dim3((unsigned)ss->nwg), dim3(SPLIT_WG), 0, s>>>(ss->d_ops, nops); break;
	default: k_slice<8><<<dim3((unsigned)ss->nwg), dim3(SPLIT_WG), 0, s>>>(ss->d_ops, nops); break;
	}
	hipError_t e = hipGetLastError();
	if (e != hipSuccess)
		xm_fatal("xm_contract: slice launch failed: %s",
		    hipGetErrorString(e));
}

void
split_gemm(contract_plan *pl, hipStream_t s)
{
	split_state *ss = pl->split;
	size_t nt = pl->tiles.size();
	pl->gemm_launches = 0;
	if (ss == nullptr || nt == 0)
		return;
	dim3 g((unsigned)nt), b(SPLIT_WG);
	switch (ss->nslices) {
	case 3: k_gemm_split<3><<<g, b, 0, s>>>(pl->d_tiles, pl->d_cblocks, ss->d_pairs, pl->d_scalars); break;
	case 4: k_gemm_split<4><<<g, b, 0, s>>>(pl->d_tiles, pl->d_cblocks, ss->d_pairs, pl->d_scalars); break;
	case 5: k_gemm_split<5><<<g, b, 0, s>>>(pl->d_tiles, pl->d_cblocks, ss->d_pairs, pl->d_scalars); break;
	case 6: k_gemm_split<6><<<g, b, 0, s>>>(pl->d_tiles, pl->d_cblocks, ss->d_pairs, pl->d_scalars); break;
	case 7: k_gemm_split<7><<<g, b, 0, s>>>(pl->d_tiles, pl->d_cblocks, ss->d_pairs, pl->d_scalars); break;
	default: k_gemm_split<8><<<g, b, 0, s>>>(pl->d_tiles, pl->d_cblocks, ss->d_pairs, pl->d_scalars); break;
	}
	hipError_t e = hipGetLastError();
	if (e != hipSuccess)
		xm_fatal("xm_contract: split GEMM launch failed: %s",
		    hipGetErrorString(e));
	pl->gemm_launches = 1;
}

void
split_release(contract_plan *pl)
{
	split_state *ss = pl->split;
	if (ss == nullptr)
		return;
	if (ss->mem)
		HIP_OK("xm_contract", hipFree(ss->mem));
	if (ss->desc)
		HIP_OK("xm_contract", hipFree(ss->desc));
	delete ss;
	pl->split = nullptr;
}
