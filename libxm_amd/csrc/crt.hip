/*
 * crt.hip - FP64 contraction emulated on the int8 matrix cores through the
 * Chinese remainder theorem (an Ozaki scheme II form: K. Ozaki, Y. Uchino,
 * T. Imamura, "Ozaki Scheme II: A GEMM-oriented emulation of floating-point
 * matrix multiplication using an integer modular technique", 2025).
 *
 * For every C block the GEMM is C(p,q) += sum_i f_i X_i(p,k) Y_i(q,k) over
 * its merged pairs i (K concatenated).  Each operand row is scaled by a
 * power of two and rounded to an integer, x'(p,k) = rint(x(p,k) 2^e_p),
 * e_p chosen from the 2-norm of the row over every operand the row meets in
 * any C block (its "slab": the C blocks with the same block indices at the
 * row's dims), so that the integer result
 *   C'(p,q) = sum_i (f_i / f_n) x'_i(p,.) . y'_i(q,.)
 * obeys |C'| <= sqrt(W_x W_y) R^2 <= M / 4 by Cauchy-Schwarz, with
 * M = m_0 ... m_{n-1} the product of the moduli, R >= the scaled row norms
 * and W the largest total factor weight an operand gets in a C block.
 * C' mod m_l is then computed exactly for every modulus on the int8 matrix
 * cores (v_mfma_i32_32x32x32_i8: residues fit a byte, products 2^14, int32
 * sums exact for K < 2^17), C' is rebuilt by Garner's algorithm in FP64,
 * and C = beta C + alpha f_n 2^-(e_p + e_q) C'.  The merged pair factors
 * must form a chain f_1 / f_2, f_2 / f_3, ... of integers once sorted by
 * magnitude (the ladder's are 2 and 1); the accumulator is multiplied by
 * each ratio where the factor changes (Horner).  The only rounding is that
 * of x' and y' (half a unit of a row whose norm is about R) and of the
 * final FP64 steps: the relative error of C is about
 * 0.6 sqrt(K) sqrt(4W / M), independent of how the entries are spread.
 *
 * Per call: k_crt_norm (row sums of squares of every operand instance),
 * k_crt_scale (slab row exponents), k_crt_quant (integer rows -> n residue
 * planes of bytes in the MFMA operand order), k_gemm_crt (n passes of an
 * int8 GEMM over each tile, residues of C' to a byte buffer), k_crt_recon
 * (Garner, scaling, alpha/beta update of C).
 */
#include "contract_plan.h"
#include "crt_math.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>

/* the tables as compile-time constants (folded into immediates where the
 * index is known after unrolling) and in constant memory (runtime index) */
static constexpr crt_tables CT = crt_make_tables();
__constant__ crt_tables d_ct = crt_make_tables();

/* exponent marking a row whose norm is not finite: its outputs are NaN */
#define CRT_POISON (-0x40000000)

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

/* An operand as one slab sees it: a P-contiguous panel, element (p, k) at
 * src[p + k * ld], rows x k, padded to 32-row tiles and 32-k steps. */
struct crt_inst {
	uint64_t src;
	uint32_t ld, rows, k, nrt, nks;	/* row tiles, k steps */
	uint32_t slab;
	uint64_t planes;		/* n x nrt x nks fragments of 1 KiB */
	uint64_t norm;			/* double [rows] */
	uint64_t first_norm_wg;		/* prefix of ceil(rows / 64) */
	uint64_t first_quant_wg;	/* prefix of nrt * ceil(nks / 4) */
};

struct crt_slab {
	uint32_t rows, ninst;
	uint32_t inst0, pad;		/* first entry in the slab's instance list */
	uint64_t exps;			/* int32 [rows] */
	uint64_t first_row;		/* prefix of rows */
};

/* A pair as the GEMM walks it (whole dwords: read with scalar loads). */
struct crt_pair {
	uint64_t xs, ys;		/* planes of the X and Y instances */
	uint32_t nsteps;		/* GEMM k steps (CR_KQ fragments each) */
	uint32_t rtfr;			/* fragments of a row tile and modulus */
	uint32_t xrt, yrt;
	int32_t ratio;			/* multiply the sum by this first (1: no) */
	uint32_t pad;
};

struct crt_cblock {
	uint64_t cres;			/* int8 [n][q][p] */
	uint64_t c;			/* C block or panel, (p, q) at p + q ldc */
	uint64_t xexp, yexp;		/* slab row exponents */
	uint32_t p, q, ldc, npairs;
	double scale;			/* f_n: the smallest factor */
	uint64_t first_wg;		/* prefix of ceil(p q / RECON_CHUNK) */
};

struct crt_state {
	int nmod = 13;
	double log2m = 0, radius = 0, err_bound = 0;
	double kmax = 0, wx = 1, wy = 1;
	std::vector<crt_inst> insts;
	std::vector<crt_slab> slabs;
	std::vector<uint32_t> slab_list;
	std::vector<crt_pair> pairs;
	std::vector<crt_cblock> cbs;
	uint64_t norm_wgs = 0, quant_wgs = 0, slab_rows = 0, recon_wgs = 0;
	void *mem = nullptr;		/* planes | norms | exponents | residues */
	void *desc = nullptr;
	crt_inst *d_insts = nullptr;
	crt_slab *d_slabs = nullptr;
	uint32_t *d_list = nullptr;
	crt_pair *d_pairs = nullptr;
	crt_cblock *d_cbs = nullptr;
	double plane_bytes = 0;
};

/* ---- device helpers ------------------------------------------------ */


template <typename T>
__device__ __forceinline__ uint32_t
find_first(const T *__restrict__ v, uint32_t n, uint64_t x,
    uint64_t T::*field)
{
	uint32_t lo = 0, hi = n - 1;
	while (lo < hi) {
		uint32_t mid = (lo + hi + 1) / 2;
		if (v[mid].*field <= x)
			lo = mid;
		else
			hi = mid - 1;
	}
	return lo;
}

/* ---- norms, scales, residues --------------------------------------- */

#define NORM_WG 256

/* One workgroup per 64 rows of an instance: thread (r, part) sums the
 * squares of row r over k = part, part + 4, ...; the four parts are added
 * in a fixed order. */
__global__ void __launch_bounds__(NORM_WG)
k_crt_norm(const crt_inst *__restrict__ insts, uint32_t ninst)
{
	__shared__ uint32_t found;
	__shared__ double part_sum[4][64];
	const uint64_t wg = blockIdx.x;
	if (threadIdx.x == 0)
		found = find_first(insts, ninst, wg, &crt_inst::first_norm_wg);
	__syncthreads();
	/* uniform: the descriptor comes in scalar loads */
	const crt_inst in = insts[__builtin_amdgcn_readfirstlane(found)];
	const uint32_t r = threadIdx.x % 64, part = threadIdx.x / 64;
	const uint32_t row = (uint32_t)(wg - in.first_norm_wg) * 64 + r;
	typedef const __attribute__((address_space(1))) double gdouble;
	double s = 0.0;
	if (row < in.rows) {
		/* eight loads in flight, the sum in the order of k */
		const gdouble *Xr = (const gdouble *)(uintptr_t)in.src + row;
		uint32_t k = part;
		for (; k + 28 < in.k; k += 32) {
			double x[8];
#pragma unroll
			for (int j = 0; j < 8; j++)
				x[j] = Xr[(size_t)(k + 4 * j) * in.ld];
#pragma unroll
			for (int j = 0; j < 8; j++)
				s = fma(x[j], x[j], s);
		}
		for (; k < in.k; k += 4) {
			const double x = Xr[(size_t)k * in.ld];
			s = fma(x, x, s);
		}
	}
	part_sum[part][r] = s;
	__syncthreads();
	if (part == 0 && row < in.rows)
		((double *)(uintptr_t)in.norm)[row] = ((part_sum[0][r] +
		    part_sum[1][r]) + part_sum[2][r]) + part_sum[3][r];
}

/* One thread per slab row: the row's squared norm over the slab's
 * instances (in list order), then the largest e with 2^e n <= radius. */
__global__ void __launch_bounds__(256)
k_crt_scale(const crt_slab *__restrict__ slabs, uint32_t nslab,
    const crt_inst *__restrict__ insts, const uint32_t *__restrict__ list,
    uint64_t nrows, double radius)
{
	const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
	if (t >= nrows)
		return;
	const uint32_t si = find_first(slabs, nslab, t, &crt_slab::first_row);
	const crt_slab sl = slabs[si];
	const uint32_t row = (uint32_t)(t - sl.first_row);
	double n2 = 0.0;
	for (uint32_t j = 0; j < sl.ninst; j++)
		n2 += ((const double *)(uintptr_t)insts[list[sl.inst0 + j]].norm)[row];
	int e = 0;
	if (!isfinite(n2)) {
		e = CRT_POISON;
	} else if (n2 > 0.0) {
		/* n rounded up by a few ulps so that 2^e n <= radius holds for
		 * the true norm */
		const double n = sqrt(n2) * (1.0 + 0x1p-48);
		int x;
		(void)frexp(radius / n, &x);
		e = x - 1;
		while (ldexp(n, e) > radius)
			e--;
	}
	((int *)(uintptr_t)sl.exps)[row] = e;
}

/* The 16 residue bytes of a lane for moduli I .. N - 1 (I a template
 * parameter, so the table entries are immediates). */
template <int N, int I = 0>
__device__ __forceinline__ void
quant_planes(const float (&lm)[16][4], const int (&l0)[16], uint8_t *dst,
    size_t plane)
{
	if constexpr (I < N) {
		uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
		for (int j = 0; j < 16; j++) {
			const int r = crt_residue13<I>(lm[j], l0[j], CT);
			w[j / 4] |= (uint32_t)(r & 0xff) << (8 * (j % 4));
		}
		*(uint4 *)(dst + (size_t)I * plane) = make_uint4(w[0], w[1],
		    w[2], w[3]);
		quant_planes<N, I + 1>(lm, l0, dst, plane);
	}
}

/*
 * One workgroup per (32-row tile, 4 k steps of 32) of an instance.  Thread
 * (L, part) = (tid % 64, tid / 64) owns lane L of the step's fragments:
 * row L % 32 of the tile, k = 32 ks + 16 (L / 32) + [0, 16).  It rounds the
 * 16 scaled values to integers, splits them into limbs and writes 16 bytes
 * of residues to each modulus' plane: plane l holds fragment (rt, ks) at
 * ((l nrt + rt) nks + ks) KiB, lane L's bytes at 16 L (the order both MFMA
 * operands of v_mfma_i32_32x32x32_i8 take, pairing byte b of lane L of X
 * and Y alike).
 */
template <int N>
__global__ void __launch_bounds__(256)
k_crt_quant(const crt_inst *__restrict__ insts, uint32_t ninst,
    const crt_slab *__restrict__ slabs)
{
	__shared__ uint32_t found;
	const uint64_t wg = blockIdx.x;
	if (threadIdx.x == 0)
		found = find_first(insts, ninst, wg, &crt_inst::first_quant_wg);
	__syncthreads();
	const crt_inst in = insts[__builtin_amdgcn_readfirstlane(found)];
	const uint32_t local = (uint32_t)(wg - in.first_quant_wg);
	const uint32_t nkq = (in.nks + 3) / 4;
	const uint32_t rt = local / nkq, ks = (local % nkq) * 4 + threadIdx.x / 64;
	if (ks >= in.nks)
		return;
	const uint32_t L = threadIdx.x % 64, row = rt * 32 + L % 32;
	const uint32_t k0 = ks * 32 + 16 * (L / 32);
	const bool rin = row < in.rows;
	const int e = rin ? ((const int *)(uintptr_t)slabs[in.slab].exps)[row] : 0;
	/* all 16 loads issued before the first is used: entries past the
	 * instance read its first element instead and are dropped */
	typedef const __attribute__((address_space(1))) double gdouble;
	const gdouble *X0 = (const gdouble *)(uintptr_t)in.src;
	const gdouble *Xr = X0 + row + (size_t)k0 * in.ld;
	double xv[16];
#pragma unroll
	for (int j = 0; j < 16; j++) {
		const bool ok = rin && k0 + j < in.k;
		xv[j] = *(ok ? Xr + (size_t)j * in.ld : X0);
	}
	float lm[16][4];
	int l0[16];
#pragma unroll
	for (int j = 0; j < 16; j++) {
		const bool ok = rin && k0 + j < in.k && e != CRT_POISON;
		int li[4];
		crt_limbs13(rint(ldexp(ok ? xv[j] : 0.0, e)), li);
#pragma unroll
		for (int t = 0; t < 4; t++)
			lm[j][t] = (float)li[t];
		l0[j] = li[0];
	}
	uint8_t *base = (uint8_t *)(uintptr_t)in.planes;
	const size_t plane = (size_t)in.nrt * in.nks * 1024;
	const size_t off = ((size_t)rt * in.nks + ks) * 1024 + 16 * L;
	quant_planes<N>(lm, l0, base + off, plane);
}

/* ---- GEMM ------------------------------------------------------------ */

/*
 * Workgroup tile 256 (P) x 256 (Q), one part of the moduli (head.ksplit:
 * parts << 8 | part + 1, 0 for all), 8 waves in 4 x 2 (two a SIMD), each
 * 64 x 128 as 2 x 4 MFMA tiles of 32 x 32 (128 int32 accumulators a
 * lane); one workgroup per CU.  A tile record (gemm_stile8) names the
 * C block and first row of each of its eight 32-row tiles: the rows of one
 * C block (by default), or of the C blocks of a group that read the same
 * Y fragments, stacked (XM_CRT_STACK=1).  With P up to 256 a tile covers a
 * C block's rows, so every Y fragment is read once per C block and
 * modulus, and the C block's tiles (one per 256 columns, placed on one
 * XCD) share the X fragments in L2.  A step covers CR_KQ 32-k fragments: the workgroup
 * stages 8 X and 8 Y row tiles of CR_KQ 1 KiB fragments in LDS, every
 * fragment one wave-wide global_load_lds_dwordx4 (wave w moves X row tile
 * w and Y row tile w), and each wave runs 8 CR_KQ MFMAs on it, reading the
 * next fragment's operands while the current one's MFMAs run.  CR_NST
 * stages rotate with the DMA CR_NST - 1 steps ahead of the MFMAs, one
 * barrier per step.  The fetch walks the tile's (modulus, pair, k step)
 * sequence flat, so the prefetch crosses pair and modulus boundaries; the
 * MFMAs walk it as nested loops (the accumulators stay put in registers):
 * where the factor changes between pairs they are multiplied by the ratio
 * (mod m), at the end of each modulus reduced to residues and stored as
 * bytes.
 *
 * What bounds it is the operand stream into LDS, not the matrix cores: a
 * model kernel with the same per-step work measured 11.3 TB/s of
 * global_load_lds traffic (all CUs) from L2-resident data and 6.2 from
 * HBM, while the ladder's 256 x 128 tiles need 29 TB/s at the int8 peak
 * (170 int8 ops per staged byte).  256 x 256 tiles stage 1.5 times fewer
 * bytes per op (256): on the o=64 v=384 ladder the GEMM and reconstruction
 * went from 10.7 to 9.5 ms, although they pad Q = 576 to 768.  Tried and
 * dropped on that ladder: 128-k steps in two stages (9.6 ms) and 32-k steps
 * in eight (9.9); fragments of the next step read into a second register
 * set during the current step's MFMAs (11.1: spills); waves whose columns
 * all lie past the C block skipping their MFMAs (9.7); not fetching row
 * tiles past the C block (10.6).  The earlier 256 x 128 form ran 10.7 ms,
 * and 128 x 128 wave tiles (one wave a SIMD) 12.7.
 */
#define CR_BP 256
#define CR_BQ 256
#define CR_WG 512
#define CR_NST 4
#define CR_KQ 2			/* 32-k fragments a step */

typedef __attribute__((address_space(3))) void cr_lds_t;
typedef __attribute__((address_space(1))) void cr_gbl_t;

/* source of the DMAs of row tiles past an instance's rows (a step) */
static __device__ uint4 cr_zero_page[CR_KQ * 64];

struct cr_stage {
	/* X row tile j, fragment h at (j CR_KQ + h) 64; Y row tile j at
	 * ((8 + j) CR_KQ + h) 64 */
	uint4 frag[16 * CR_KQ * 64];
};

__device__ __forceinline__ v4i
frag4(uint4 v)
{
	v4i r;
	__builtin_memcpy(&r, &v, 16);
	return r;
}

__global__ void __launch_bounds__(CR_WG, 1)
k_gemm_crt(const gemm_stile8 *__restrict__ tiles,
    const gemm_cblock *__restrict__ cbs, const crt_pair *__restrict__ pairs,
    const crt_cblock *__restrict__ ccbs, const gemm_scalars *__restrict__ scp,
    int nmod)
{
	/* vmcnt = the DMAs of the CR_NST - 2 later steps (2 CR_KQ a step) */
	constexpr int WAIT = ((CR_NST - 2) * 2 * CR_KQ) | (0x7 << 4) | (0xf << 8);
	__shared__ cr_stage st[CR_NST];

	const gemm_stile8 *tp = tiles + blockIdx.x;
	const gemm_tile tl = tp->head;
	if (tl.cblock == 0xffffffffu)
		return;
	const uint32_t ns = tl.ksplit ? tl.ksplit >> 8 : 1;
	const uint32_t pj = tl.ksplit ? (tl.ksplit & 0xff) - 1 : 0;
	const uint32_t m0 = pj * (uint32_t)nmod / ns;
	const uint32_t m1 = (pj + 1) * (uint32_t)nmod / ns;
	if (m0 >= m1)
		return;		/* a part with no modulus */
	const gemm_scalars sc = *scp;
	/* the first C block of the tile's stack: its pairs give Y, k steps
	 * and factor ratios, shared by every C block of the stack */
	const gemm_cblock cb = cbs[tl.cblock];
	const uint32_t pbeg = cb.pair_begin, pend = cb.pair_end;
	if ((sc.flags & 1) || pbeg == pend)
		return;		/* C = beta C: the reconstruction alone */
	const int tid = threadIdx.x, lane = tid & 63;
	const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	const int wp = wave & 3, wq = wave >> 2;
	const uint32_t ry0 = tl.q0 / 32;
	/* X row tile `wave` of the tile: row tile xrt0 of C block xcb (its
	 * pairs from xpb on), none if xcb is TILE_ROWS_NONE (loads at
	 * wave-uniform addresses: scalar) */
	const uint32_t xcb = tp->cb[wave];
	const bool xhas = xcb != TILE_ROWS_NONE;
	const uint32_t xrt0 = tp->p0[wave] / 32;
	const uint32_t xpb = xhas ? cbs[xcb].pair_begin : pbeg;

	v16i acc[4][2];
#pragma unroll
	for (int a = 0; a < 4; a++)
#pragma unroll
		for (int b = 0; b < 2; b++)
			acc[a][b] = v16i{};

	/* this wave's DMA sources for the fetch cursor's (modulus, pair):
	 * step ks reads base + ks * stride + h * 64 + lane (uint4), X row
	 * tile `wave` and Y row tile `wave` */
	const uint4 *bx = cr_zero_page, *by = cr_zero_page;
	uint32_t sx = 0, sy = 0;
	/* pair i of the stack: Y from the first C block's pair fp, X from
	 * this wave's C block's pair i */
	auto setup = [&](const crt_pair fp, uint32_t l, uint32_t i) {
		const uint32_t rty = ry0 + wave;
		bx = cr_zero_page;
		sx = 0;
		if (xhas) {
			const crt_pair xp = pairs[xpb + (i - pbeg)];
			bx = (const uint4 *)(uintptr_t)xp.xs +
			    ((size_t)l * xp.xrt + xrt0) * xp.rtfr * 64;
			sx = CR_KQ * 64;
		}
		by = cr_zero_page;
		sy = 0;
		if (rty < fp.yrt) {
			by = (const uint4 *)(uintptr_t)fp.ys +
			    ((size_t)l * fp.yrt + rty) * fp.rtfr * 64;
			sy = CR_KQ * 64;
		}
	};
	auto fetch = [&](cr_stage *__restrict__ nx, uint32_t ks) {
		const uint4 *a = bx + ks * sx + lane;
		const uint4 *b = by + ks * sy + lane;
#pragma unroll
		for (int h = 0; h < CR_KQ; h++)
			__builtin_amdgcn_global_load_lds((cr_gbl_t *)(a + 64 * h),
			    (cr_lds_t *)&nx->frag[(wave * CR_KQ + h) * 64], 16, 0, 0);
#pragma unroll
		for (int h = 0; h < CR_KQ; h++)
			__builtin_amdgcn_global_load_lds((cr_gbl_t *)(b + 64 * h),
			    (cr_lds_t *)&nx->frag[((8 + wave) * CR_KQ + h) * 64],
			    16, 0, 0);
	};
	uint32_t fl = m0, fi = pbeg, fk = 0;
	crt_pair fp = pairs[pbeg];
	setup(fp, m0, pbeg);
	bool fmore = true;
	auto fadvance = [&]() {
		if (++fk < fp.nsteps)
			return;
		fk = 0;
		if (++fi == pend) {
			fi = pbeg;
			if (++fl == m1) {
				fmore = false;
				return;
			}
		}
		fp = pairs[fi];
		setup(fp, fl, fi);
	};
#pragma unroll
	for (int s = 0; s < CR_NST - 1; s++)
		if (fmore) {
			fetch(&st[s], fk);
			fadvance();
		}
	__builtin_amdgcn_s_waitcnt(0);
	__syncthreads();

	/* residues of the finished modulus l: row tile 2 wp + b of the tile
	 * is rows p0 .. p0 + 31 of its own C block */
	auto store = [&](uint32_t l) {
		const int m = CRT_MOD[l];
		const uint32_t mg = d_ct.magic[l];
		const int c31 = d_ct.c31[l], hi = d_ct.hi[l];
#pragma unroll
		for (int b = 0; b < 2; b++) {
			const uint32_t mcb = tp->cb[2 * wp + b];
			if (mcb != TILE_ROWS_NONE) {
				const crt_cblock mc = ccbs[mcb];
				/* the extents made opaque here, so the store
				 * offsets are computed at each store, not kept
				 * across the k loop */
				uint32_t P = mc.p, Q = mc.q;
				asm volatile("" : "+s"(P), "+s"(Q));
				uint8_t *out = (uint8_t *)(uintptr_t)mc.cres +
				    l * (size_t)P * Q;
				const uint32_t p = tp->p0[2 * wp + b] +
				    (lane & 31);
#pragma unroll
				for (int a = 0; a < 4; a++)
#pragma unroll
					for (int i = 0; i < 16; i++) {
						const uint32_t q = tl.q0 + wq * 128 +
						    32 * a + 8 * (i / 4) +
						    4 * (lane >> 5) + (i % 4);
						const int r = crt_smod(acc[a][b][i], m,
						    mg, c31, hi);
						if (p < P && q < Q)
							out[(size_t)q * P + p] =
							    (uint8_t)(r & 0xff);
					}
			}
#pragma unroll
			for (int a = 0; a < 4; a++)
				acc[a][b] = v16i{};
		}
	};
	auto rescale = [&](uint32_t l, int ratio) {
		const int m = CRT_MOD[l];
		const uint32_t mg = d_ct.magic[l];
		const int c31 = d_ct.c31[l], hi = d_ct.hi[l];
#pragma unroll
		for (int a = 0; a < 4; a++)
#pragma unroll
			for (int b = 0; b < 2; b++)
#pragma unroll
				for (int i = 0; i < 16; i++)
					acc[a][b][i] = crt_smod(acc[a][b][i], m, mg,
					    c31, hi) * ratio;
	};
	auto step = [&](const cr_stage *__restrict__ cur,
	    cr_stage *__restrict__ nxt) {
		const bool more = fmore;
		if (more) {
			fetch(nxt, fk);
			fadvance();
		}
		/* D[q][p] += Y(q, k) X(p, k): A = Y fragment, B = X fragment;
		 * fragment h + 1's operands are read while h's MFMAs run */
		v4i f[2][6];
		auto rd = [&](v4i (&g)[6], int h) {
			g[0] = frag4(cur->frag[((2 * wp) * CR_KQ + h) * 64 + lane]);
			g[1] = frag4(cur->frag[((2 * wp + 1) * CR_KQ + h) * 64 + lane]);
#pragma unroll
			for (int a = 0; a < 4; a++)
				g[2 + a] = frag4(cur->frag[((8 + 4 * wq + a) *
				    CR_KQ + h) * 64 + lane]);
		};
		rd(f[0], 0);
#pragma unroll
		for (int h = 0; h < CR_KQ; h++) {
			if (h + 1 < CR_KQ)
				rd(f[(h + 1) & 1], h + 1);
			const v4i (&g)[6] = f[h & 1];
#pragma unroll
			for (int a = 0; a < 4; a++)
#pragma unroll
				for (int b = 0; b < 2; b++)
					acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
					    g[2 + a], g[b], acc[a][b], 0, 0, 0);
		}
		__builtin_amdgcn_sched_group_barrier(0x100, 6, 0);
#pragma unroll
		for (int h = 0; h < CR_KQ; h++) {
			if (h + 1 < CR_KQ)
				__builtin_amdgcn_sched_group_barrier(0x100, 6, 0);
			__builtin_amdgcn_sched_group_barrier(0x008, 8, 0);
		}
		if (more)
			__builtin_amdgcn_s_waitcnt(WAIT);
		else
			__builtin_amdgcn_s_waitcnt(0);
		__atomic_signal_fence(__ATOMIC_SEQ_CST);
		__builtin_amdgcn_s_barrier();
		__atomic_signal_fence(__ATOMIC_SEQ_CST);
	};
	int cs = 0;
	for (uint32_t l = m0; l < m1; l++) {
		for (uint32_t i = pbeg; i < pend; i++) {
			const crt_pair cp = pairs[i];
			if (i != pbeg && cp.ratio != 1)
				rescale(l, cp.ratio);
			for (uint32_t ks = 0; ks < cp.nsteps; ks++) {
				const int ns = cs + CR_NST - 1 < CR_NST ?
				    cs + CR_NST - 1 : cs - 1;
				step(&st[cs], &st[ns]);
				cs = cs + 1 < CR_NST ? cs + 1 : 0;
			}
		}
		store(l);
		__builtin_amdgcn_s_waitcnt(0);
	}
}

/* ---- reconstruction ------------------------------------------------ */

#define RECON_CHUNK 1024

/* 256 threads per 1024 elements of a C block (4 each, p fastest). */
template <int N>
__global__ void __launch_bounds__(256)
k_crt_recon(const crt_cblock *__restrict__ ccbs, uint32_t ncb,
    const gemm_scalars *__restrict__ scp)
{
	__shared__ uint32_t found;
	const uint64_t wg = blockIdx.x;
	if (threadIdx.x == 0)
		found = find_first(ccbs, ncb, wg, &crt_cblock::first_wg);
	__syncthreads();
	const crt_cblock cb = ccbs[__builtin_amdgcn_readfirstlane(found)];
	const gemm_scalars sc = *scp;
	const size_t pq = (size_t)cb.p * cb.q;
	const size_t e0 = (size_t)(wg - cb.first_wg) * RECON_CHUNK;
	const bool bz = sc.flags & 2;
	const bool none = (sc.flags & 1) || cb.npairs == 0;
	const double al = __dmul_rn(sc.alpha_re, cb.scale);
	double *C = (double *)(uintptr_t)cb.c;
	const int8_t *R = (const int8_t *)(uintptr_t)cb.cres;
	const int *xe = (const int *)(uintptr_t)cb.xexp;
	const int *ye = (const int *)(uintptr_t)cb.yexp;
	for (uint32_t j = 0; j < RECON_CHUNK / 256; j++) {
		const size_t e = e0 + j * 256 + threadIdx.x;
		if (e >= pq)
			return;
		const uint32_t p = (uint32_t)(e % cb.p), q = (uint32_t)(e / cb.p);
		const size_t o = (size_t)p + (size_t)q * cb.ldc;
		double v = 0.0;
		if (!none) {
			int r[N];
#pragma unroll
			for (int l = 0; l < N; l++)
				r[l] = R[(size_t)l * pq + e];
			const int ep = xe[p], eq = ye[q];
			const double x = crt_reconstruct<N>(r, CT);
			v = (ep == CRT_POISON || eq == CRT_POISON) ? NAN :
			    __dmul_rn(ldexp(x, -(ep + eq)), al);
		}
		double out;
		if (bz)
			out = v;
		else if (none)
			out = __dmul_rn(sc.beta_re, C[o]);
		else
			out = __dadd_rn(__dmul_rn(sc.beta_re, C[o]), v);
		C[o] = out;
	}
}

/* ---- host ------------------------------------------------------------ */

static int g_nmod = 0;		/* 0: chosen per plan */

XM_API void
xm_set_split_moduli(int n)
{
	if (n != 0 && (n < 8 || n > CRT_MAXN))
		XM_FATAL("xm_set_split_moduli", "moduli must be 0 or in 8..13");
	g_nmod = n;
}

XM_API int
xm_get_split_moduli(void)
{
	return g_nmod;
}

int current_nmoduli(void) { return g_nmod; }

/* The modelled relative error of C for n moduli (fit of the emulation
 * against exact products: 0.62 sqrt(K) sqrt(4 W / M), K the longest
 * concatenated K of a C block). */
static double
crt_err_model(int n, double kmax, double w)
{
	return 0.62 * std::sqrt(std::max(kmax, 1.0)) *
	    std::sqrt(4.0 * w * std::exp2(-crt_log2_modulus(n)));
}

/*
 * Whether the plan can run in the CRT form: FP64, every merged factor
 * chain of a C block (sorted by magnitude, largest first) made of integer
 * ratios of at most 2^7, and every C block's K below 2^17 (int32 sums of
 * residue products stay exact).
 */
bool
crt_eligible(const contract_plan *pl)
{
	if (pl->type != XM_SCALAR_DOUBLE)
		return false;
	for (const gemm_cblock &cb : pl->cblocks) {
		std::vector<double> f;
		double kt = 0;
		for (uint32_t i = cb.pair_begin; i < cb.pair_end; i++) {
			f.push_back(std::fabs(pl->pairs[i].a_re));
			kt += (pl->pairs[i].k + 127) / 128 * 128;
		}
		if (kt >= 131072)
			return false;
		std::sort(f.begin(), f.end(), std::greater<double>());
		for (size_t i = 1; i < f.size(); i++) {
			const double r = f[i - 1] / f[i];
			if (r != std::rint(r) || r > 128 || r * f[i] != f[i - 1])
				return false;
		}
	}
	return true;
}

void
crt_prepare(contract_plan *pl, const masks_t &mk, const xm_tensor_t *c)
{
	const char *fn = "xm_contract";
	crt_state *cs = new crt_state;
	pl->crt = cs;
	/* slabs: (X or Y, swap, C block indices at the row dims) */
	xm_dim_t *list;
	size_t nlist;
	xm_tensor_get_canonical_block_list(c, &list, &nlist);
	std::map<std::vector<size_t>, uint32_t> slab_of;
	std::map<std::tuple<uint32_t, uint64_t, uint32_t, uint32_t, uint32_t>,
	    uint32_t> inst_of;
	auto slab = [&](int role, bool swap, const xm_dim_t &blk,
	    uint32_t rows) -> uint32_t {
		const xm_dim_t &dims = (role == 0) != swap ? mk.cidxc : mk.aidxc;
		std::vector<size_t> key = {(size_t)role, (size_t)swap};
		for (size_t k = 0; k < dims.n; k++)
			key.push_back(blk.i[dims.i[k]]);
		auto it = slab_of.find(key);
		if (it != slab_of.end())
			return it->second;
		crt_slab s;
		memset(&s, 0, sizeof s);
		s.rows = rows;
		uint32_t id = (uint32_t)cs->slabs.size();
		cs->slabs.push_back(s);
		slab_of[key] = id;
		return id;
	};
	std::vector<std::vector<uint32_t>> members;
	auto inst = [&](uint32_t sl, uint64_t src, uint32_t ld, uint32_t rows,
	    uint32_t k) -> uint32_t {
		auto key = std::make_tuple(sl, src, ld, rows, k);
		auto it = inst_of.find(key);
		if (it != inst_of.end())
			return it->second;
		crt_inst in;
		memset(&in, 0, sizeof in);
		in.src = src;
		in.ld = ld;
		in.rows = rows;
		in.k = k;
		in.nrt = (rows + 31) / 32;
		in.nks = (k + 127) / 128 * 4;	/* 32-k steps, 4 a GEMM step */
		in.slab = sl;
		uint32_t id = (uint32_t)cs->insts.size();
		cs->insts.push_back(in);
		if (members.size() <= sl)
			members.resize(sl + 1);
		members[sl].push_back(id);
		inst_of[key] = id;
		return id;
	};
	cs->pairs.resize(pl->pairs.size());
	cs->cbs.resize(pl->cblocks.size());
	for (size_t bi = 0; bi < pl->cblocks.size(); bi++) {
		const gemm_cblock &cb = pl->cblocks[bi];
		const bool swap = cb.pair_begin < cb.pair_end ?
		    (bool)pl->pair_swap[cb.pair_begin] : false;
		const uint32_t sx = slab(0, swap, list[bi], cb.p);
		const uint32_t sy = slab(1, swap, list[bi], cb.q);
		/* pairs sorted by factor magnitude, largest first */
		std::vector<uint32_t> ord;
		for (uint32_t i = cb.pair_begin; i < cb.pair_end; i++)
			ord.push_back(i);
		std::stable_sort(ord.begin(), ord.end(), [&](uint32_t u,
		    uint32_t v) {
			return std::fabs(pl->pairs[u].a_re) >
			    std::fabs(pl->pairs[v].a_re);
		});
		std::map<uint32_t, double> wx, wy;
		const double fn = ord.empty() ? 1.0 : pl->pairs[ord.back()].a_re;
		double kt = 0;
		for (size_t j = 0; j < ord.size(); j++) {
			const gemm_pair &g = pl->pairs[ord[j]];
			if (g.xmode != PANEL_PCONTIG || g.ymode != PANEL_PCONTIG)
				XM_FATAL(fn, "CRT form needs P-contiguous panels");
			const uint32_t ix = inst(sx, g.x, g.ldx, cb.p, g.k);
			const uint32_t iy = inst(sy, g.y, g.ldy, cb.q, g.k);
			crt_pair cp;
			memset(&cp, 0, sizeof cp);
			cp.xs = ix;	/* instance ids until the planes exist */
			cp.ys = iy;
			cp.nsteps = (g.k + 32 * CR_KQ - 1) / (32 * CR_KQ);
			cp.rtfr = (g.k + 127) / 128 * 4;
			cp.xrt = (cb.p + 31) / 32;
			cp.yrt = (cb.q + 31) / 32;
			cp.ratio = j == 0 ? 1 : (int32_t)std::lrint(
			    pl->pairs[ord[j - 1]].a_re / g.a_re);
			cs->pairs[cb.pair_begin + j] = cp;
			const double w = std::fabs(g.a_re / fn);
			wx[ix] += w;
			wy[iy] += w;
			kt += (g.k + 127) / 128 * 128.0;
		}
		for (auto &e : wx)
			cs->wx = std::max(cs->wx, e.second);
		for (auto &e : wy)
			cs->wy = std::max(cs->wy, e.second);
		cs->kmax = std::max(cs->kmax, kt);
		crt_cblock cc;
		memset(&cc, 0, sizeof cc);
		cc.c = cb.c;
		cc.p = cb.p;
		cc.q = cb.q;
		cc.ldc = cb.ldc;
		cc.npairs = cb.pair_end - cb.pair_begin;
		cc.scale = fn;
		cc.xexp = sx;	/* slab ids until the exponents exist */
		cc.yexp = sy;
		cs->cbs[bi] = cc;
	}
	free(list);
	/* moduli: the fewest whose modelled error is below 1e-13, or the
	 * pinned count; the radius keeps |x'| below 2^51 */
	const double w = std::sqrt(cs->wx * cs->wy);
	int n = current_nmoduli();
	if (n == 0) {
		n = CRT_MAXN;
		for (int t = 8; t <= CRT_MAXN; t++)
			if (crt_err_model(t, cs->kmax, w) <= 1e-13) {
				n = t;
				break;
			}
	}
	cs->nmod = n;
	cs->log2m = crt_log2_modulus(n);
	cs->radius = std::min(std::exp2(0.5 * (cs->log2m - 2.0) -
	    0.5 * std::log2(w)), 0x1p51);
	cs->err_bound = crt_err_model(n, cs->kmax, w);
	pl->nslices = n;

	/* storage: planes | norms | exponents | residues */
	size_t total = 0;
	auto take = [&](size_t b) {
		size_t o = total;
		total += (b + 255) / 256 * 256;
		return o;
	};
	for (auto &in : cs->insts) {
		in.planes = take((size_t)n * in.nrt * in.nks * 1024);
		cs->plane_bytes += (double)n * in.nrt * in.nks * 1024;
	}
	for (auto &in : cs->insts)
		in.norm = take((size_t)in.rows * 8);
	for (auto &s : cs->slabs)
		s.exps = take((size_t)s.rows * 4);
	for (auto &cc : cs->cbs)
		cc.cres = take((size_t)n * cc.p * cc.q);
	HIP_OK(fn, hipMalloc(&cs->mem, std::max<size_t>(total, 256)));
	const uint64_t base = (uint64_t)(uintptr_t)cs->mem;
	uint64_t nw = 0, qw = 0;
	for (auto &in : cs->insts) {
		in.planes += base;
		in.norm += base;
		in.first_norm_wg = nw;
		nw += (in.rows + 63) / 64;
		in.first_quant_wg = qw;
		qw += (uint64_t)in.nrt * ((in.nks + 3) / 4);
	}
	cs->norm_wgs = nw;
	cs->quant_wgs = qw;
	uint64_t rows = 0;
	for (uint32_t si = 0; si < cs->slabs.size(); si++) {
		crt_slab &s = cs->slabs[si];
		s.exps += base;
		s.first_row = rows;
		rows += s.rows;
		s.inst0 = (uint32_t)cs->slab_list.size();
		s.ninst = si < members.size() ? (uint32_t)members[si].size() : 0;
		if (si < members.size())
			for (uint32_t id : members[si])
				cs->slab_list.push_back(id);
	}
	cs->slab_rows = rows;
	for (auto &cp : cs->pairs) {
		cp.xs = cs->insts[cp.xs].planes;
		cp.ys = cs->insts[cp.ys].planes;
	}
	uint64_t rw = 0;
	for (auto &cc : cs->cbs) {
		cc.cres += base;
		cc.xexp = cs->slabs[cc.xexp].exps;
		cc.yexp = cs->slabs[cc.yexp].exps;
		cc.first_wg = rw;
		rw += ((uint64_t)cc.p * cc.q + RECON_CHUNK - 1) / RECON_CHUNK;
	}
	cs->recon_wgs = rw;

	/* descriptors */
	std::vector<std::pair<const void *, size_t>> parts = {
		{cs->insts.data(), cs->insts.size() * sizeof(crt_inst)},
		{cs->slabs.data(), cs->slabs.size() * sizeof(crt_slab)},
		{cs->slab_list.data(), cs->slab_list.size() * 4},
		{cs->pairs.data(), cs->pairs.size() * sizeof(crt_pair)},
		{cs->cbs.data(), cs->cbs.size() * sizeof(crt_cblock)}};
	std::vector<size_t> offs;
	size_t dt = 0;
	for (auto &p : parts) {
		offs.push_back(dt);
		dt += (p.second + 255) / 256 * 256;
	}
	HIP_OK(fn, hipMalloc(&cs->desc, std::max<size_t>(dt, 256)));
	std::vector<char> host(std::max<size_t>(dt, 256), 0);
	for (size_t i = 0; i < parts.size(); i++)
		if (parts[i].second)
			memcpy(host.data() + offs[i], parts[i].first,
			    parts[i].second);
	HIP_OK(fn, hipMemcpy(cs->desc, host.data(), host.size(),
	    hipMemcpyHostToDevice));
	char *d = (char *)cs->desc;
	cs->d_insts = (crt_inst *)(d + offs[0]);
	cs->d_slabs = (crt_slab *)(d + offs[1]);
	cs->d_list = (uint32_t *)(d + offs[2]);
	cs->d_pairs = (crt_pair *)(d + offs[3]);
	cs->d_cbs = (crt_cblock *)(d + offs[4]);
	/* the traffic of the per-call residue stage (bytes read + written):
	 * two reads of every instance, the planes, the C residues twice */
	double opnd = 0;
	for (auto &in : cs->insts)
		opnd += (double)in.rows * in.k * 8;
	pl->unfold_bytes += 2 * opnd + cs->plane_bytes;
}

static void
check(const char *what)
{
	hipError_t e = hipGetLastError();
	if (e != hipSuccess)
		xm_fatal("xm_contract: %s launch failed: %s", what,
		    hipGetErrorString(e));
}

void
crt_quantize(contract_plan *pl, hipStream_t s)
{
	crt_state *cs = pl->crt;
	if (cs == nullptr || cs->insts.empty())
		return;
	const uint32_t ni = (uint32_t)cs->insts.size();
	if (cs->norm_wgs == 0 || cs->slab_rows == 0 || cs->quant_wgs == 0)
		return;		/* every operand has no rows or no k */
	k_crt_norm<<<dim3((unsigned)cs->norm_wgs), dim3(NORM_WG), 0, s>>>(
	    cs->d_insts, ni);
	check("CRT norm");
	k_crt_scale<<<dim3((unsigned)((cs->slab_rows + 255) / 256)), dim3(256),
	    0, s>>>(cs->d_slabs, (uint32_t)cs->slabs.size(), cs->d_insts,
	    cs->d_list, cs->slab_rows, cs->radius);
	check("CRT scale");
	const dim3 g((unsigned)cs->quant_wgs), b(256);
	switch (cs->nmod) {
	case 8: k_crt_quant<8><<<g, b, 0, s>>>(cs->d_insts, ni, cs->d_slabs); break;
	case 9: k_crt_quant<9><<<g, b, 0, s>>>(cs->d_insts, ni, cs->d_slabs); break;
	case 10: k_crt_quant<10><<<g, b, 0, s>>>(cs->d_insts, ni, cs->d_slabs); break;
	case 11: k_crt_quant<11><<<g, b, 0, s>>>(cs->d_insts, ni, cs->d_slabs); break;
	case 12: k_crt_quant<12><<<g, b, 0, s>>>(cs->d_insts, ni, cs->d_slabs); break;
	default: k_crt_quant<13><<<g, b, 0, s>>>(cs->d_insts, ni, cs->d_slabs); break;
	}
	check("CRT residue");
}

void
crt_gemm(contract_plan *pl, hipStream_t s)
{
	crt_state *cs = pl->crt;
	pl->gemm_launches = 0;
	if (cs == nullptr || cs->cbs.empty())
		return;
	const size_t nt = pl->tiles.size();
	if (nt) {
		k_gemm_crt<<<dim3((unsigned)(nt / 5)), dim3(CR_WG), 0, s>>>(
		    (const gemm_stile8 *)pl->d_tiles, pl->d_cblocks,
		    cs->d_pairs, cs->d_cbs, pl->d_scalars, cs->nmod);
		check("CRT GEMM");
		pl->gemm_launches = 1;
		pl->gemm_kernel = "k_gemm_crt";
	}
	if (cs->recon_wgs == 0)
		return;
	const dim3 g((unsigned)cs->recon_wgs), b(256);
	const uint32_t nc = (uint32_t)cs->cbs.size();
	switch (cs->nmod) {
	case 8: k_crt_recon<8><<<g, b, 0, s>>>(cs->d_cbs, nc, pl->d_scalars); break;
	case 9: k_crt_recon<9><<<g, b, 0, s>>>(cs->d_cbs, nc, pl->d_scalars); break;
	case 10: k_crt_recon<10><<<g, b, 0, s>>>(cs->d_cbs, nc, pl->d_scalars); break;
	case 11: k_crt_recon<11><<<g, b, 0, s>>>(cs->d_cbs, nc, pl->d_scalars); break;
	case 12: k_crt_recon<12><<<g, b, 0, s>>>(cs->d_cbs, nc, pl->d_scalars); break;
	default: k_crt_recon<13><<<g, b, 0, s>>>(cs->d_cbs, nc, pl->d_scalars); break;
	}
	check("CRT reconstruction");
}

double
crt_error_bound(const contract_plan *pl)
{
	return pl->crt ? pl->crt->err_bound : 0.0;
}

void
crt_release(contract_plan *pl)
{
	crt_state *cs = pl->crt;
	if (cs == nullptr)
		return;
	if (cs->mem)
		HIP_OK("xm_contract", hipFree(cs->mem));
	if (cs->desc)
		HIP_OK("xm_contract", hipFree(cs->desc));
	delete cs;
	pl->crt = nullptr;
}
