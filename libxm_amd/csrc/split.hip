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
typedef const __attribute__((address_space(1))) double gdouble_t;
typedef float f16v __attribute__((ext_vector_type(16)));

struct split_operand {
	uint64_t src;		/* FP64 panel */
	uint32_t mode, ld, rows, k;
	uint32_t rows_pad, kpad, nchunk, pad;
	uint64_t slices;	/* nslices * rows_pad * kpad bf16, fragment order */
	uint64_t exps;		/* int32 [nchunk][rows_pad] */
	uint64_t first_wg;	/* prefix of (rows_pad/32) * nchunk */
};

struct split_pair {
	uint64_t xs, ys;	/* slices of X and Y */
	uint64_t xe, ye;	/* exponents of X and Y */
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

/*
 * One workgroup per (32-row tile, 128-k chunk) of an operand panel; thread
 * (r, part) = (tid % 32, tid / 32) owns row r and k in [16 part, +16) of the
 * chunk and loads them straight from the panel: for a P-contiguous panel
 * the 32 lanes of a row group read 256 consecutive bytes per k, for a
 * K-contiguous one a lane reads a contiguous 128-byte run.  Only the row
 * maximum goes through LDS.
 */
template <int S>
__global__ void __launch_bounds__(SPLIT_WG)
k_slice(const split_operand *__restrict__ ops, uint32_t nops)
{
	__shared__ double rmax[8][32];
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
	const gdouble_t *X = (const gdouble_t *)(uintptr_t)op.src;
	const int tid = threadIdx.x, r = tid % 32, part = tid / 32;
	const uint32_t grow = rt * 32 + r, k0 = c * KC + 16 * part;

	/* this thread's 16 values (zero outside the panel) */
	double x[16];
	const bool rin = grow < op.rows;
	if (op.mode == PANEL_PCONTIG) {
#pragma unroll
		for (int j = 0; j < 16; j++) {
			const bool in = rin && k0 + j < op.k;
			const double v = X[in ? (size_t)grow +
			    (size_t)(k0 + j) * op.ld : 0];
			x[j] = in ? v : 0.0;
		}
	} else {
		const gdouble_t *row = X + (size_t)grow * op.ld;
#pragma unroll
		for (int j = 0; j < 16; j++) {
			const bool in = rin && k0 + j < op.k;
			const double v = (in ? row : X)[in ? k0 + j : 0];
			x[j] = in ? v : 0.0;
		}
	}
	/* row maxima over the chunk: 8 parts per row */
	double m = 0.0;
#pragma unroll
	for (int j = 0; j < 16; j++)
		m = fmax(m, fabs(x[j]));
	rmax[part][r] = m;
	__syncthreads();
	m = 0.0;
#pragma unroll
	for (int q = 0; q < 8; q++)
		m = fmax(m, rmax[q][r]);
	int ex = 0;
	if (m > 0.0)
		(void)frexp(m, &ex);
	if (part == 0)
		((int *)(uintptr_t)op.exps)[(size_t)c * op.rows_pad + grow] = ex;

	if (k0 >= op.kpad)
		return;
	const uint32_t nks = op.kpad / 16, ks = k0 / 16;
	uint16_t dig[S][16];
#pragma unroll
	for (int j = 0; j < 16; j++) {
		double y = ldexp(x[j], 8 - ex);
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
	/* fragment order: (((rt * nks + ks) * S + t) * 64 + lane) * 8 + j,
	 * lane = row + 32 * (k % 16 >= 8): the S slices of one row tile and
	 * k step are one contiguous S KiB run */
	uint16_t *base = (uint16_t *)(uintptr_t)op.slices;
#pragma unroll
	for (int t = 0; t < S; t++) {
		const size_t f = (((size_t)rt * nks + ks) * S + t) * 64;
		uint4 lo, hi;
		lo.x = dig[t][0] | ((uint32_t)dig[t][1] << 16);
		lo.y = dig[t][2] | ((uint32_t)dig[t][3] << 16);
		lo.z = dig[t][4] | ((uint32_t)dig[t][5] << 16);
		lo.w = dig[t][6] | ((uint32_t)dig[t][7] << 16);
		hi.x = dig[t][8] | ((uint32_t)dig[t][9] << 16);
		hi.y = dig[t][10] | ((uint32_t)dig[t][11] << 16);
		hi.z = dig[t][12] | ((uint32_t)dig[t][13] << 16);
		hi.w = dig[t][14] | ((uint32_t)dig[t][15] << 16);
		*(uint4 *)(base + (f + r) * 8) = lo;
		*(uint4 *)(base + (f + r + 32) * 8) = hi;
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

/*
 * Workgroup tile 128 (P) x 128 (Q), 8 waves, one workgroup per CU (two
 * waves per SIMD).  Wave w owns P rows (w & 3) * 32 + [0,32) and Q columns
 * (w >> 2) * 64 + [0,64) as two 32 x 32 MFMA tiles.  A tile record
 * (gemm_stile) names the C block and first row of each 32-row tile: the
 * C blocks of a group that read the same Y slices with the same factors
 * stack their rows (XM_SPLIT_STACK=0: one C block a tile), so the ladder's
 * C blocks of 144-400 rows are not padded to multiples of 128 (its GEMM
 * went from 15.0 to 13.5 ms at S = 5); an X fragment or exponent row comes
 * from its row tile's C block's pair.  For every 16-wide k
 * step the workgroup stages the S slices of its 4 X and 4 Y row tiles (8S
 * fragments of 1 KiB) and the exponents of the step's chunk in LDS.  A
 * fragment is a contiguous 1 KiB run, so it is one wave-wide
 * global_load_lds_dwordx4 (lane l's 16 bytes land at the LDS base + 16 l):
 * no staging registers.  NST stages rotate with the DMA running NST - 1
 * steps ahead (three stages up to S = 6; two at S = 7 and 8, whose 56 and
 * 64 KiB stages do not fit three times: on the o=64 v=384 ladder their
 * GEMMs then ran 21.8 and 26.5 ms against 35.5 and 44.3 on the per-wave
 * kernel, at 255 and 256 VGPRs, S = 8 spilling one); every wave issues
 * exactly S + 1 DMAs per step (row tiles
 * past an operand's padding read a zero page), so the end of a step waits
 * with a constant count for the next step's data and leaves the later
 * steps' loads in flight; one barrier per step.  The (pair, k step)
 * sequence of the C block is walked flat, so the prefetch crosses pair
 * boundaries.  At each 128-k chunk end the weight classes are recombined
 * into FP64 register sums; C is read and written once, at the end.
 */
#define SP_BP 128
#define SP_BQ 128
#define SP_WG 512

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void gbl_void_t;

/* source of the DMAs of row tiles past an operand's padding */
static __device__ uint4 sp_zero_page[64];

__device__ __forceinline__ bf8
frag_of(uint4 v)
{
	bf8 r;
	__builtin_memcpy(&r, &v, 16);
	return r;
}

/* One LDS stage: the 8S fragments of a k step, then the exponents of the
 * step's chunk for the tile's 128 X rows and 128 Y rows. */
template <int S>
struct sp_stage {
	uint4 frag[8 * S * 64];
	int ex[SP_BP];
	int ey[SP_BQ];
};

/* stages that fit the 160 KiB of LDS */
template <int S>
constexpr int
sp_nstages(void)
{
	return sizeof(sp_stage<S>) * 3 <= 160 * 1024 ? 3 : 2;
}

template <int S>
__global__ void __launch_bounds__(SP_WG)
k_gemm_split2(const gemm_stile *__restrict__ tiles,
    const gemm_cblock *__restrict__ cbs, const split_pair *__restrict__ pairs,
    const gemm_scalars *__restrict__ scp)
{
	constexpr int FX = 4 * S, NST = sp_nstages<S>();
	/* s_waitcnt: vmcnt = the DMAs of the NST - 2 later steps, expcnt and
	 * lgkmcnt not waited on */
	constexpr int WAIT = ((NST - 2) * (S + 1)) | (0x7 << 4) | (0xf << 8);
	__shared__ sp_stage<S> st[NST];

	const gemm_stile *tp = tiles + blockIdx.x;
	const gemm_tile tl = tp->head;
	if (tl.cblock == 0xffffffffu)
		return;
	/* the first C block of the tile's stack: its pairs give Y, k steps
	 * and factors, shared by every C block of the stack */
	const gemm_cblock cb = cbs[tl.cblock];
	const gemm_scalars sc = *scp;
	const int tid = threadIdx.x, lane = tid & 63;
	const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	const int wp = wave & 3, wq = wave >> 2;
	const uint32_t qw = tl.q0 + wq * 64;
	const uint32_t pend = cb.pair_end;
	const uint32_t pbeg = (sc.flags & 1) ? pend : cb.pair_begin;
	const uint32_t ry0 = tl.q0 / 32;
	/* row tile r of the tile: rows rp0[r] .. + 31 of C block tp->cb[r]
	 * (none: TILE_ROWS_NONE), whose pairs start at rb[r] (the first
	 * C block's for none, so loads stay in range) */
	uint32_t rb[4], rp0[4];
	bool rhas[4];
#pragma unroll
	for (int r = 0; r < 4; r++) {
		const uint32_t c = tp->cb[r];
		rhas[r] = c != TILE_ROWS_NONE;
		rb[r] = cbs[rhas[r] ? c : tl.cblock].pair_begin;
		rp0[r] = rhas[r] ? tp->p0[r] : 0;
	}
	/* this wave's rows: row tile wp (selects, not indexed loads: an array
	 * indexed at run time would go to scratch memory) */
	uint32_t wb = rb[0], wr0 = rp0[0];
	bool whas = rhas[0];
#pragma unroll
	for (int r = 1; r < 4; r++) {
		wb = wp == r ? rb[r] : wb;
		wr0 = wp == r ? rp0[r] : wr0;
		whas = wp == r ? rhas[r] : whas;
	}
	const uint32_t mcb = tp->cb[wp];
	const uint32_t prow = wr0 + (lane & 31);	/* row of C block mcb */
	const uint32_t wrows = pairs[wb].xrows_pad;

	f16v a2[2] = {}, a3[2] = {}, a4[2] = {}, a5[2] = {};
	double acc[2][16];
#pragma unroll
	for (int j = 0; j < 2; j++)
#pragma unroll
		for (int i = 0; i < 16; i++)
			acc[j][i] = 0.0;

	/*
	 * k step kx of the fetch pair into stage nx: S + 1 DMAs per wave.
	 * Fragment f = wave + 8 i is wave's i-th (1 KiB); a row tile past the
	 * operand's padded rows reads the zero page.  Then exponent quarter
	 * wave % 4 (64 rows: X rows 0-63, 64-127, Y rows 0-63, 64-127 of the
	 * tile) of the step's chunk, 4 bytes a lane; waves 4-7 repeat waves
	 * 0-3's, writing the same values.  Rows past the padding read row 0
	 * and are never used.  The address arithmetic is done once per pair
	 * (setup): a step adds kx times a per-DMA stride (0 for the zero
	 * page) to a base.  Recomputed per DMA it took ~20 scalar
	 * instructions each, issued by both waves of a SIMD at the start of
	 * every step while the MFMA pipe waited; hoisted, the o=64 v=384
	 * ladder's GEMM went from 17.2 to 16.3 ms at S = 5.
	 */
	const uint4 *dsrc[S];
	uint32_t dstp[S];
	const int *esrc;
	uint32_t erows;
	const int quarter = wave & 3;
	const bool eisx = quarter < 2;
	/* pair ip of the stack (px: the first C block's): Y from px, X row
	 * tile r from C block tp->cb[r]'s pair ip */
	auto setup = [&](const split_pair px, uint32_t ip) {
		const uint32_t ri = ip - pbeg;
#pragma unroll
		for (int i = 0; i < S; i++) {
			const int f = wave + 8 * i;
			const bool isx = f < FX;
			const int g = isx ? f : f - FX;
			const int t = g % S;
			dsrc[i] = sp_zero_page;
			dstp[i] = 0;
			if (isx) {
				const int r = g / S;
				uint32_t b = rb[0], p0 = rp0[0];
				bool has = rhas[0];
#pragma unroll
				for (int q = 1; q < 4; q++) {
					b = r == q ? rb[q] : b;
					p0 = r == q ? rp0[q] : p0;
					has = r == q ? rhas[q] : has;
				}
				if (has) {
					const split_pair xp = pairs[b + ri];
					dsrc[i] = (const uint4 *)(uintptr_t)xp.xs +
					    ((size_t)(p0 / 32) * xp.nks * S + t) * 64;
					dstp[i] = S * 64;
				}
			} else {
				const uint32_t rt = ry0 + g / S;
				if (rt * 32 < px.yrows_pad) {
					dsrc[i] = (const uint4 *)(uintptr_t)px.ys +
					    ((size_t)rt * px.nks * S + t) * 64;
					dstp[i] = S * 64;
				}
			}
		}
		if (eisx) {
			/* lanes 0-31: row tile 2 (quarter & 1), 32-63 the next */
			const int r0 = 2 * (quarter & 1);
			const split_pair e0 = pairs[(r0 == 0 ? rb[0] : rb[2]) + ri];
			const split_pair e1 = pairs[(r0 == 0 ? rb[1] : rb[3]) + ri];
			const bool h0 = r0 == 0 ? rhas[0] : rhas[2];
			const bool h1 = r0 == 0 ? rhas[1] : rhas[3];
			const uint32_t p00 = r0 == 0 ? rp0[0] : rp0[2];
			const uint32_t p01 = r0 == 0 ? rp0[1] : rp0[3];
			const bool hi = lane >= 32;
			erows = hi ? e1.xrows_pad : e0.xrows_pad;
			const uint64_t xe = hi ? e1.xe : e0.xe;
			uint32_t row = (hi ? p01 : p00) + (lane & 31);
			if (!(hi ? h1 : h0) || row >= erows)
				row = 0;
			esrc = (const int *)(uintptr_t)xe + row;
		} else {
			erows = px.yrows_pad;
			uint32_t row = tl.q0 + 64 * (quarter & 1) + lane;
			if (row >= erows)
				row = 0;
			esrc = (const int *)(uintptr_t)px.ye + row;
		}
	};
	auto fetch = [&](sp_stage<S> *__restrict__ nx, uint32_t kx) {
#pragma unroll
		for (int i = 0; i < S; i++) {
			const int f = wave + 8 * i;
			const uint4 *src = dsrc[i] + (size_t)kx * dstp[i] + lane;
			__builtin_amdgcn_global_load_lds((gbl_void_t *)src,
			    (lds_void_t *)&nx->frag[f * 64], 16, 0, 0);
		}
		const int *src = esrc + (size_t)(kx / (KC / 16)) * erows;
		int *dst = (eisx ? nx->ex : nx->ey) + 64 * (quarter & 1);
		__builtin_amdgcn_global_load_lds((gbl_void_t *)src,
		    (lds_void_t *)dst, 4, 0, 0);
	};
	/* chunk done: acc += al * 2^(ex + ey) * (((a5 + a4) + a3) + a2); a
	 * lane's 4 consecutive q rows are one int4 of the stage's exponents */
	auto flush = [&](const sp_stage<S> *__restrict__ cur,
	    const split_pair px) {
		const double al = __dmul_rn(sc.alpha_re, px.a_re);
		const int ep = whas && prow < wrows ?
		    cur->ex[wp * 32 + (lane & 31)] : 0;
#pragma unroll
		for (int j = 0; j < 2; j++)
#pragma unroll
			for (int g = 0; g < 4; g++) {
				const uint32_t ql = wq * 64 + 32 * j + 8 * g +
				    4 * (lane >> 5);
				const int4 e4 = tl.q0 + ql < px.yrows_pad ?
				    *(const int4 *)(cur->ey + ql) :
				    make_int4(0, 0, 0, 0);
				const int eq[4] = {e4.x, e4.y, e4.z, e4.w};
#pragma unroll
				for (int r = 0; r < 4; r++) {
					const int i = 4 * g + r;
					const double v = (((double)a5[j][i] +
					    (double)a4[j][i]) + (double)a3[j][i]) +
					    (double)a2[j][i];
					acc[j][i] = fma(ldexp(al, ep + eq[r]), v,
					    acc[j][i]);
				}
			}
#pragma unroll
		for (int j = 0; j < 2; j++) {
			a2[j] = f16v{};
			a3[j] = f16v{};
			a4[j] = f16v{};
			a5[j] = f16v{};
		}
	};

	/*
	 * Descriptors are values, never references into two address spaces,
	 * so they stay in scalar registers and the loop stays uniform.  The
	 * stages reach step() as restrict pointers: the alias scopes that
	 * inlining gives them let the compiler read this stage while the DMA
	 * into the other is in flight (without them it waits for the DMA
	 * before the first read, serialising every step).
	 */
	auto advance = [&](uint32_t &i, uint32_t &k, split_pair &p) {
		if (++k >= p.nks) {
			k = 0;
			if (++i < pend)
				p = pairs[i];
		}
	};
	/* the fetch cursor also sets up its new pair's DMAs */
	auto advance_fetch = [&](uint32_t &i, uint32_t &k, split_pair &p) {
		if (++k >= p.nks) {
			k = 0;
			if (++i < pend) {
				p = pairs[i];
				setup(p, i);
			}
		}
	};
	/* compute cursor (ip, ks, pr); fetch cursor NST - 1 steps ahead */
	uint32_t ip = pbeg, ks = 0, fip = pbeg, fks = 0;
	split_pair pr, fpr;
	if (ip < pend) {
		pr = pairs[ip];
		fpr = pr;
		setup(fpr, ip);
	}
#pragma unroll
	for (int s = 0; s < NST - 1; s++)
		if (fip < pend) {
			fetch(&st[s], fks);
			advance_fetch(fip, fks, fpr);
		}
	__builtin_amdgcn_s_waitcnt(0);
	__syncthreads();
	auto step = [&](const sp_stage<S> *__restrict__ cur,
	    sp_stage<S> *__restrict__ nxt) {
		const bool more = fip < pend;
		if (more) {
			fetch(nxt, fks);
			advance_fetch(fip, fks, fpr);
		}
		/* MFMAs of this k step: D[q][p] += Y_u * X_t by weight class */
		bf8 fx[S];
#pragma unroll
		for (int t = 0; t < S; t++)
			fx[t] = frag_of(cur->frag[(wp * S + t) * 64 + lane]);
#pragma unroll
		for (int u = 1; u <= S; u++) {
			const bf8 fy0 = frag_of(cur->frag[(FX + (2 * wq) * S + u - 1) * 64 + lane]);
			const bf8 fy1 = frag_of(cur->frag[(FX + (2 * wq + 1) * S + u - 1) * 64 + lane]);
#pragma unroll
			for (int t = 1; t + u <= S + 1; t++) {
				const int d = t + u;
				if (d == 2) {
					a2[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fy0, fx[t - 1], a2[0], 0, 0, 0);
					a2[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fy1, fx[t - 1], a2[1], 0, 0, 0);
				} else if (d == 3) {
					a3[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fy0, fx[t - 1], a3[0], 0, 0, 0);
					a3[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fy1, fx[t - 1], a3[1], 0, 0, 0);
				} else if (d == 4) {
					a4[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fy0, fx[t - 1], a4[0], 0, 0, 0);
					a4[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fy1, fx[t - 1], a4[1], 0, 0, 0);
				} else {
					a5[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fy0, fx[t - 1], a5[0], 0, 0, 0);
					a5[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fy1, fx[t - 1], a5[1], 0, 0, 0);
				}
			}
		}
		if ((ks % (KC / 16)) == KC / 16 - 1 || ks + 1 == pr.nks)
			flush(cur, pr);
		/*
		 * This step's stage is read (every LDS read's value is used by
		 * now) and this wave's DMAs of the next step have landed; those
		 * just issued stay in flight.  A bare s_barrier then publishes
		 * the stage: __syncthreads()' release fence would also wait for
		 * the DMAs in flight.  The signal fences only stop the compiler
		 * from moving LDS accesses across the barrier.
		 */
		if (more)
			__builtin_amdgcn_s_waitcnt(WAIT);
		else
			__builtin_amdgcn_s_waitcnt(0);
		__atomic_signal_fence(__ATOMIC_SEQ_CST);
		__builtin_amdgcn_s_barrier();
		__atomic_signal_fence(__ATOMIC_SEQ_CST);
		advance(ip, ks, pr);
	};
	int cs = 0;
	while (ip < pend) {
		const int ns = cs + NST - 1 < NST ? cs + NST - 1 : cs - 1;
		step(&st[cs], &st[ns]);
		cs = cs + 1 < NST ? cs + 1 : 0;
	}
	if (!whas)
		return;
	const gemm_cblock mc = cbs[mcb];
	double *C = (double *)(uintptr_t)mc.c;
	const bool bz = sc.flags & 2, none = pbeg == pend;
#pragma unroll
	for (int j = 0; j < 2; j++)
#pragma unroll
		for (int i = 0; i < 16; i++) {
			const uint32_t q = qw + 32 * j + 8 * (i / 4) +
			    4 * (lane >> 5) + (i % 4);
			if (prow < mc.p && q < mc.q) {
				const size_t o = (size_t)prow + (size_t)q * mc.ldc;
				double v;
				if (bz)
					v = none ? 0.0 : acc[j][i];
				else if (none)
					v = __dmul_rn(sc.beta_re, C[o]);
				else
					v = __dadd_rn(__dmul_rn(sc.beta_re, C[o]),
					    acc[j][i]);
				C[o] = v;
			}
		}
}

/* 64 x 64 outputs per workgroup; wave w computes the 32 x 32 block
 * (p: (w & 1) * 32, q: (w >> 1) * 32) with one 32x32 MFMA tile.  Kept as
 * the reference form of k_gemm_split2 (XM_SPLIT_KERNEL=1). */
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
			const size_t xf = (size_t)(pw / 32) * pr.nks;
			const size_t yf = (size_t)(qw / 32) * pr.nks;
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
						fx[t] = ld_frag(xs + (((xf + ks) * S + t) *
						    64 + lane) * 8);
						fy[t] = ld_frag(ys + (((yf + ks) * S + t) *
						    64 + lane) * 8);
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
				const int exp_p = xe[(size_t)c * pr.xrows_pad +
				    pw + (lane & 31)];
#pragma unroll
				for (int i = 0; i < 16; i++) {
					uint32_t qq = qw + 8 * (i / 4) + 4 * (lane >> 5) + (i % 4);
					int exp_q = ye[(size_t)c * pr.yrows_pad + qq];
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

/* KERNEL<S><<<launch>>>(args) for the runtime slice count S. */
#define SDISPATCH(KERNEL, S, ...) SDISPATCH_ARGS(KERNEL, S, (__VA_ARGS__))
#define SDISPATCH_ARGS(KERNEL, S, LAUNCH) [&](auto... args) {		\
	switch (S) {							\
	case 3: KERNEL<3><<<SD_UNPACK LAUNCH>>>(args...); break;	\
	case 4: KERNEL<4><<<SD_UNPACK LAUNCH>>>(args...); break;	\
	case 5: KERNEL<5><<<SD_UNPACK LAUNCH>>>(args...); break;	\
	case 6: KERNEL<6><<<SD_UNPACK LAUNCH>>>(args...); break;	\
	case 7: KERNEL<7><<<SD_UNPACK LAUNCH>>>(args...); break;	\
	default: KERNEL<8><<<SD_UNPACK LAUNCH>>>(args...); break;	\
	}								\
}
#define SD_UNPACK(...) __VA_ARGS__

bool
split_staged_kernel(int nslices)
{
	const char *e = getenv("XM_SPLIT_KERNEL");
	return !(e && *e == '1') && nslices >= 3 && nslices <= 8;
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
	const dim3 g((unsigned)nt);
	const gemm_tile *t = pl->d_tiles;
	const gemm_cblock *c = pl->d_cblocks;
	const gemm_scalars *sc = pl->d_scalars;
	if (!pl->split_staged) {
		SDISPATCH(k_gemm_split, ss->nslices, g, dim3(SPLIT_WG), 0, s)(t,
		    c, ss->d_pairs, sc);
	} else {
		/* stacked tile records, three gemm_tile each */
		const dim3 b(SP_WG), g3((unsigned)(nt / 3));
		const gemm_stile *ts = (const gemm_stile *)t;
		switch (ss->nslices) {
		case 3: k_gemm_split2<3><<<g3, b, 0, s>>>(ts, c, ss->d_pairs, sc); break;
		case 4: k_gemm_split2<4><<<g3, b, 0, s>>>(ts, c, ss->d_pairs, sc); break;
		case 5: k_gemm_split2<5><<<g3, b, 0, s>>>(ts, c, ss->d_pairs, sc); break;
		case 6: k_gemm_split2<6><<<g3, b, 0, s>>>(ts, c, ss->d_pairs, sc); break;
		case 7: k_gemm_split2<7><<<g3, b, 0, s>>>(ts, c, ss->d_pairs, sc); break;
		default: k_gemm_split2<8><<<g3, b, 0, s>>>(ts, c, ss->d_pairs, sc); break;
		}
	}
	hipError_t e = hipGetLastError();
	if (e != hipSuccess)
		xm_fatal("xm_contract: split GEMM launch failed: %s",
		    hipGetErrorString(e));
	pl->gemm_launches = 1;
	pl->gemm_kernel = pl->split_staged ? "k_gemm_split2" :
	    "k_gemm_split";
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
