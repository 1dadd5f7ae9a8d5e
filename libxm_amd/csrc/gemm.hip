/*
 * gemm.hip - grouped GEMM over the block pairs of a contraction
 * (replaces the per-pair ?gemm_ loop of libxm src/contract.c:193-216).
 *
 * One launch covers every canonical C block.  A work tile is
 * (C block, P-range, Q-range); its K loop walks all surviving pairs of
 * that C block back to back (K concatenation), so each output element has
 * one owner and no atomics are needed.  Per pair the X tile is scaled by
 * al = alpha * (s_a * s_b) while it is staged into LDS.  The epilogue
 * applies beta: C = acc (beta == 0, C not read) or beta*C + acc.
 *
 * FP64: v_mfma_f64_16x16x4_f64.  The MFMA's A operand is the Y tile and
 * its B operand the X tile, so D is laid out with P along the lanes and
 * the stores to the P-contiguous C panel are coalesced.
 * FP32 / complex: a register-tiled FMA kernel of the same structure.
 */
#include "contract_plan.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <type_traits>

typedef double d4 __attribute__((ext_vector_type(4)));

#define GEMM_WG 256

typedef const __attribute__((address_space(1))) double gdouble;

/* Element (p, k) of a panel if inside P x K, else element 0 (to be
 * discarded by the caller).  A global (not flat) load, so it counts only
 * against vmcnt and LDS waits do not wait for it; no branch. */
__device__ __forceinline__ double
ld_raw(const double *x, uint32_t mode, uint32_t ld, uint32_t p, uint32_t k,
    uint32_t P, uint32_t K)
{
	const bool in = p < P && k < K;
	const size_t off = mode == PANEL_PCONTIG ?
	    (size_t)p + (size_t)k * ld : (size_t)k + (size_t)p * ld;
	return ((gdouble *)x)[in ? off : 0];
}

/*
 * K split.  The tiles of a split C block run as two halves over the two
 * parts of its (pair, 16-k step) sequence, cut at (split_pair, split_k);
 * each half writes its sums, without beta, to its slot of partial sums
 * (element (p, q) of the tile at p + q * BP), and k_ksplit_sum then writes
 * C = beta * C + (half 0 + half 1).  Every kernel that runs FP64 tiles cuts
 * a C block at the same step, so all tile shapes give identical results.
 * With alpha == 0 both halves are empty.
 */
__device__ __forceinline__ void
ksplit_range(const gemm_tile &tl, const gemm_cblock &cb, uint32_t pbeg,
    uint32_t &sp, uint32_t &sk, uint32_t &ep, uint32_t &ek)
{
	sp = pbeg;
	sk = 0;
	ep = cb.pair_end;
	ek = 0;
	if (tl.ksplit && pbeg != cb.pair_end) {
		if ((tl.ksplit - 1) & 1) {
			sp = cb.split_pair;
			sk = cb.split_k;
		} else {
			ep = cb.split_pair;
			ek = cb.split_k;
		}
	}
}

/* The same as a start cursor and a count of 16-k steps. */
__device__ __forceinline__ void
ksplit_steps(const gemm_tile &tl, const gemm_cblock &cb, uint32_t pbeg,
    uint32_t &sp, uint32_t &sk, uint32_t &n)
{
	sp = pbeg;
	sk = 0;
	n = pbeg == cb.pair_end ? 0 : cb.steps;
	if (tl.ksplit && n) {
		if ((tl.ksplit - 1) & 1) {
			sp = cb.split_pair;
			sk = cb.split_k;
			n = cb.steps - cb.split_steps;
		} else {
			n = cb.split_steps;
		}
	}
}

template <int MT, int NT, int BP>
__device__ __forceinline__ void
ksplit_store(double *__restrict__ ksw, const gemm_tile &tl, int wp, int wq,
    int lane, const d4 (&acc)[MT][NT])
{
	static_assert(BP == 32 * MT && BP * 32 * NT <= KSPLIT_SLOT,
	    "a tile's partial sums fit one slot");
	double *W = ksw + (size_t)(tl.ksplit - 1) * KSPLIT_SLOT;
#pragma unroll
	for (int i = 0; i < MT; i++)
#pragma unroll
		for (int j = 0; j < NT; j++)
#pragma unroll
			for (int r = 0; r < 4; r++) {
				const int p = wp + i * 16 + (lane & 15);
				const int q = wq + j * 16 + 4 * r + (lane >> 4);
				W[p + q * BP] = acc[i][j][r];
			}
}

/* C = beta * C + (half 0 + half 1) over one split tile per workgroup
 * (fix[] holds the first halves). */
template <int BP, int BQ>
__global__ void __launch_bounds__(GEMM_WG)
k_ksplit_sum(const gemm_tile *__restrict__ fix,
    const gemm_cblock *__restrict__ cbs, const double *__restrict__ ksw,
    const gemm_scalars *__restrict__ scp)
{
	static_assert(BP * BQ <= KSPLIT_SLOT, "a tile fits one slot");
	const gemm_tile tl = fix[blockIdx.x];
	const gemm_cblock cb = cbs[tl.cblock];
	const gemm_scalars sc = *scp;
	const double *W0 = ksw + (size_t)(tl.ksplit - 1) * KSPLIT_SLOT;
	const double *W1 = W0 + KSPLIT_SLOT;
	double *C = (double *)(uintptr_t)cb.c;
	const bool bz = sc.flags & 2, none = sc.flags & 1;
	for (int e = threadIdx.x; e < BP * BQ; e += GEMM_WG) {
		const uint32_t p = tl.p0 + e % BP, q = tl.q0 + e / BP;
		if (p < cb.p && q < cb.q) {
			const size_t o = (size_t)p + (size_t)q * cb.ldc;
			double v;
			if (bz)
				v = none ? 0.0 : __dadd_rn(W0[e], W1[e]);
			else if (none)
				v = __dmul_rn(sc.beta_re, C[o]);
			else
				v = __dadd_rn(__dmul_rn(sc.beta_re, C[o]),
				    __dadd_rn(W0[e], W1[e]));
			C[o] = v;
		}
	}
}

/*
 * FP64 tile kernel.  BP x BQ outputs per workgroup, 4 waves as 2 x 2, each
 * wave (BP/2) x (BQ/2) = MT x NT MFMA tiles of 16 x 16, BK k per LDS step.
 * One LDS buffer and a small register footprint keep several waves per
 * SIMD resident (3 at 64 x 128 x 16): with 64-cycle MFMAs it is the other
 * waves that hide a wave's loads, stores and barriers.  Measured on the
 * o=64 v=384 ladder, 64 x 128 x 16 beats 64 x 64 x {8,16,32}, 128 x 64,
 * 64 x 192, 64 x 256, 64 x 128 x {8,32} and 128 x 128 tiles with a
 * register-prefetched double buffer (2 waves per SIMD).  A wave whose rows
 * or columns all lie past the C block skips its MFMAs.
 */
template <int BP, int BQ, int BK>
__global__ void __launch_bounds__(GEMM_WG)
k_gemm_f64(const gemm_tile *__restrict__ tiles,
    const gemm_cblock *__restrict__ cbs, const gemm_pair *__restrict__ pairs,
    const gemm_scalars *__restrict__ scp, double *__restrict__ ksw)
{
	constexpr int MT = BP / 2 / 16, NT = BQ / 2 / 16;
	constexpr int EX = BP * BK / GEMM_WG, EY = BQ * BK / GEMM_WG;
	constexpr int LP = BP + 1, LQ = BQ + 1;
	__shared__ double Xs[BK * LP];
	__shared__ double Ys[BK * LQ];

	const gemm_tile tl = tiles[blockIdx.x];
	if (tl.cblock == 0xffffffffu)
		return;
	const gemm_cblock cb = cbs[tl.cblock];
	const gemm_scalars sc = *scp;
	const int tid = threadIdx.x, lane = tid & 63;
	const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	const int wp = (wave & 1) * (BP / 2), wq = (wave >> 1) * (BQ / 2);
	const bool live = tl.p0 + wp < cb.p && tl.q0 + wq < cb.q;

	d4 acc[MT][NT];
	for (int i = 0; i < MT; i++)
		for (int j = 0; j < NT; j++)
			acc[i][j] = d4{0, 0, 0, 0};

	/* step k0 of pair pr: load() fetches the raw values into registers
	 * rx / ry; store() zeroes those outside the panels, scales X by al
	 * and writes the LDS tiles. */
	double rx[EX], ry[EY];
	auto load = [&](const gemm_pair pr, uint32_t k0) {
		const double *X = (const double *)(uintptr_t)pr.x;
		const double *Y = (const double *)(uintptr_t)pr.y;
#pragma unroll
		for (int e = 0; e < EX; e++) {
			int idx = tid + GEMM_WG * e, p, k;
			if (pr.xmode == PANEL_PCONTIG) {
				p = idx % BP; k = idx / BP;
			} else {
				k = idx % BK; p = idx / BK;
			}
			rx[e] = ld_raw(X, pr.xmode, pr.ldx, tl.p0 + p, k0 + k,
			    cb.p, pr.k);
		}
#pragma unroll
		for (int e = 0; e < EY; e++) {
			int idx = tid + GEMM_WG * e, q, k;
			if (pr.ymode == PANEL_PCONTIG) {
				q = idx % BQ; k = idx / BQ;
			} else {
				k = idx % BK; q = idx / BK;
			}
			ry[e] = ld_raw(Y, pr.ymode, pr.ldy, tl.q0 + q, k0 + k,
			    cb.q, pr.k);
		}
	};
	auto store = [&](const gemm_pair pr, uint32_t k0) {
		const double al = __dmul_rn(sc.alpha_re, pr.a_re);
#pragma unroll
		for (int e = 0; e < EX; e++) {
			int idx = tid + GEMM_WG * e, p, k;
			if (pr.xmode == PANEL_PCONTIG) {
				p = idx % BP; k = idx / BP;
			} else {
				k = idx % BK; p = idx / BK;
			}
			const bool in = tl.p0 + p < cb.p && k0 + k < pr.k;
			Xs[k * LP + p] = in ? rx[e] * al : 0.0;
		}
#pragma unroll
		for (int e = 0; e < EY; e++) {
			int idx = tid + GEMM_WG * e, q, k;
			if (pr.ymode == PANEL_PCONTIG) {
				q = idx % BQ; k = idx / BQ;
			} else {
				k = idx % BK; q = idx / BK;
			}
			const bool in = tl.q0 + q < cb.q && k0 + k < pr.k;
			Ys[k * LQ + q] = in ? ry[e] : 0.0;
		}
	};

	/* descriptors are values (scalar registers), never references */
	const uint32_t pbeg = (sc.flags & 1) ? cb.pair_end : cb.pair_begin;
	uint32_t sp, sk, ep, ek;
	ksplit_range(tl, cb, pbeg, sp, sk, ep, ek);
	const uint32_t elim = ek ? ep + 1 : ep;
	for (uint32_t ip = sp; ip < elim; ip++) {
		const gemm_pair pr = pairs[ip];
		const uint32_t kb = ip == sp ? sk : 0;
		const uint32_t ke = ip == ep ? ek : pr.k;
		for (uint32_t k0 = kb; k0 < ke; k0 += BK) {
			load(pr, k0);
			__syncthreads();
			store(pr, k0);
			__syncthreads();
			if (!live)
				continue;
#pragma unroll
			for (int kk = 0; kk < BK; kk += 4) {
				const int kr = kk + (lane >> 4), c16 = lane & 15;
				double xa[MT], yb[NT];
#pragma unroll
				for (int i = 0; i < MT; i++)
					xa[i] = Xs[kr * LP + wp + i * 16 + c16];
#pragma unroll
				for (int j = 0; j < NT; j++)
					yb[j] = Ys[kr * LQ + wq + j * 16 + c16];
#pragma unroll
				for (int i = 0; i < MT; i++)
#pragma unroll
					for (int j = 0; j < NT; j++)
						acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
						    yb[j], xa[i], acc[i][j], 0, 0, 0);
			}
		}
	}

	/* D[q][p] of v_mfma_f64_16x16x4_f64: lane -> p = lane % 16,
	 * register r -> q = 4 * r + lane / 16 */
	if (tl.ksplit) {
		ksplit_store<MT, NT, BP>(ksw, tl, wp, wq, lane, acc);
		return;
	}
	double *C = (double *)(uintptr_t)cb.c;
	const bool bz = sc.flags & 2;
	const bool none = pbeg == cb.pair_end;
#pragma unroll
	for (int i = 0; i < MT; i++)
#pragma unroll
		for (int j = 0; j < NT; j++)
#pragma unroll
			for (int r = 0; r < 4; r++) {
				uint32_t p = tl.p0 + wp + i * 16 + (lane & 15);
				uint32_t q = tl.q0 + wq + j * 16 + 4 * r + (lane >> 4);
				if (p < cb.p && q < cb.q) {
					size_t o = (size_t)p + (size_t)q * cb.ldc;
					double v;
					if (bz)
						v = none ? 0.0 : acc[i][j][r];
					else if (none)
						v = __dmul_rn(sc.beta_re, C[o]);
					else
						v = __dadd_rn(__dmul_rn(sc.beta_re, C[o]),
						    acc[i][j][r]);
					C[o] = v;
				}
			}
}

/* ---- persistent tile queues ---------------------------------------- */

#define TILE_NONE 0xffffffffu
#define NXCD_Q 8

/*
 * A launch's tiles as NXCD_Q lists (list x holds tiles[s * 8 + x], s <
 * per: the order the launch-order kernels run them in) with one counter
 * each, zeroed before the launch.  A workgroup takes the next tile of its
 * own XCD's list (the XCC_ID hardware register) and, once that is empty,
 * of the others in turn, so every list drains whatever XCD a workgroup
 * runs on; each counter passes `per` at most once per workgroup.
 */
struct tile_queue {
	uint32_t *ctr;
	uint32_t per, x0, done;

	__device__ tile_queue(uint32_t *c, uint32_t n) : ctr(c), per(n), done(0)
	{
		/* s_getreg_b32 hwreg(HW_REG_XCC_ID, 0, 4) */
		x0 = (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | 20) % NXCD_Q;
	}
	/* called by one thread of the workgroup */
	__device__ uint32_t take()
	{
		for (uint32_t d = 0; d < NXCD_Q; d++) {
			const uint32_t x = (x0 + d) % NXCD_Q;
			if (done & (1u << x))
				continue;
			const uint32_t s = atomicAdd(&ctr[x], 1u);
			if (s < per)
				return s * NXCD_Q + x;
			done |= 1u << x;
		}
		return TILE_NONE;
	}
};

__global__ void
k_queue_reset(uint32_t *__restrict__ ctr)
{
	if (threadIdx.x < NXCD_Q)
		ctr[threadIdx.x] = 0;
}

/* ---- FP64, operands staged by DMA -------------------------------- */

/*
 * FP64 tile kernel for plans whose every X and Y panel is P-contiguous
 * with an even row count and leading dimension and a 16-byte aligned base
 * (the planner checks): 64 x 128 outputs, 4 waves as 2 x 2 of 32 x 64.
 * Per 16-k step the tiles move global -> LDS by global_load_lds_dwordx4,
 * lane l carrying rows 2(l % 32), +1 of one column: an X DMA holds two
 * 64-row columns, a Y DMA one 128-row column, laid out [k][row]; columns
 * past K and row pairs past the panel read a zero page.  1.5 BK DMAs per
 * step over the 4 waves; NST rotating stages with the DMA NST - 1 steps
 * ahead, a counted vmcnt wait and a bare s_barrier per step, and the
 * stages reach the step as restrict pointers (see k_gemm_split2 for why
 * each matters).  MINW (waves per SIMD the compiler must allow for) keeps
 * the accumulators in VGPRs: without it they go to AGPRs and the kernel
 * ran 13% slower.  On the o=64 v=384 ladder BK 16, 2 stages, 3 waves per
 * SIMD (18.4 ms) edged out BK 16 / 3 stages / 2 waves (18.7), BK 8 with
 * 2-4 stages at 3-5 waves (18.7-19.0).  The pair
 * factor al scales X values in registers, after the LDS read.  Same
 * summation order per output as k_gemm_f64: results are identical.
 */
#define FD_BP 64
#define FD_BQ 128

typedef __attribute__((address_space(3))) void fd_lds_t;
typedef __attribute__((address_space(1))) void fd_gbl_t;

static __device__ uint4 fd_zero_page[64];

template <int FD_BK>
struct fd_stage {
	double x[FD_BK * FD_BP];	/* [k][p] */
	double y[FD_BK * FD_BQ];	/* [k][q] */
};

/* One 64 x 128 tile on the workgroup's stages st. */
template <int FD_BK, int FD_NST>
__device__ __forceinline__ void
f64_dma_tile(const gemm_tile tl, const gemm_cblock *__restrict__ cbs,
    const gemm_pair *__restrict__ pairs, const gemm_scalars *__restrict__ scp,
    double *__restrict__ ksw, fd_stage<FD_BK> *__restrict__ st)
{
	constexpr int MT = 2, NT = 4;
	constexpr int NXD = FD_BK / 2, NYD = FD_BK;	/* X / Y DMAs per step */
	constexpr int DPW = (NXD + NYD) / 4;		/* per wave */
	static_assert(NXD % 4 == 0 || NXD == 4, "X DMAs per wave");
	constexpr int WAIT = (DPW * (FD_NST - 2)) | (0x7 << 4) | (0xf << 8);
	typedef fd_stage<FD_BK> stage_t;

	if (tl.cblock == 0xffffffffu)
		return;
	const gemm_cblock cb = cbs[tl.cblock];
	const gemm_scalars sc = *scp;
	const int tid = threadIdx.x, lane = tid & 63;
	const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	const int wp = (wave & 1) * 32, wq = (wave >> 1) * 64;
	const bool live = tl.p0 + wp < cb.p && tl.q0 + wq < cb.q;
	const uint32_t pend = cb.pair_end;
	const uint32_t pbeg = (sc.flags & 1) ? pend : cb.pair_begin;

	d4 acc[MT][NT];
#pragma unroll
	for (int i = 0; i < MT; i++)
#pragma unroll
		for (int j = 0; j < NT; j++)
			acc[i][j] = d4{0, 0, 0, 0};

	/* step at column k0 of pair px into stage nx: wave w issues X DMAs
	 * w, w + 4, ... (two columns each) and Y DMAs w, w + 4, ... */
	auto fetch = [&](stage_t *__restrict__ nx, const gemm_pair px,
	    uint32_t k0) {
		const double *X = (const double *)(uintptr_t)px.x;
		const double *Y = (const double *)(uintptr_t)px.y;
		const uint32_t r2 = 2 * (lane & 31);
#pragma unroll
		for (int d = 0; d < NXD / 4; d++) {
			const int xd = wave + 4 * d;
			const int col = 2 * xd + (lane >> 5);
			const uint32_t k = k0 + col, p = tl.p0 + r2;
			const void *src = &fd_zero_page[lane];
			if (k < px.k && p < cb.p)
				src = X + (size_t)p + (size_t)k * px.ldx;
			__builtin_amdgcn_global_load_lds((fd_gbl_t *)src,
			    (fd_lds_t *)&nx->x[2 * xd * FD_BP], 16, 0, 0);
		}
#pragma unroll
		for (int d = 0; d < NYD / 4; d++) {
			const int col = wave + 4 * d;
			const uint32_t k = k0 + col, q = tl.q0 + 2 * lane;
			const void *src = &fd_zero_page[lane];
			if (k < px.k && q < cb.q)
				src = Y + (size_t)q + (size_t)k * px.ldy;
			__builtin_amdgcn_global_load_lds((fd_gbl_t *)src,
			    (fd_lds_t *)&nx->y[col * FD_BQ], 16, 0, 0);
		}
	};
	auto advance = [&](uint32_t &i, uint32_t &k, gemm_pair &p) {
		k += FD_BK;
		if (k >= p.k) {
			k = 0;
			if (++i < pend)
				p = pairs[i];
		}
	};

	/* compute cursor (ip, kc, pr) with n steps left; fetch cursor
	 * FD_NST - 1 steps ahead with fn left */
	uint32_t ip, kc, n;
	ksplit_steps(tl, cb, pbeg, ip, kc, n);
	uint32_t fip = ip, fk = kc, fn = n;
	gemm_pair pr, fpr;
	if (n) {
		pr = pairs[ip];
		fpr = pr;
	}
#pragma unroll
	for (int s = 0; s < FD_NST - 1; s++)
		if (fn) {
			fetch(&st[s], fpr, fk);
			advance(fip, fk, fpr);
			fn--;
		}
	__builtin_amdgcn_s_waitcnt(0);
	__syncthreads();
	auto step = [&](const stage_t *__restrict__ cur,
	    stage_t *__restrict__ nxt) {
		const bool more = fn != 0;
		if (more) {
			fetch(nxt, fpr, fk);
			advance(fip, fk, fpr);
			fn--;
		}
		if (live) {
			const double al = __dmul_rn(sc.alpha_re, pr.a_re);
#pragma unroll
			for (int kk = 0; kk < FD_BK; kk += 4) {
				const int kr = kk + (lane >> 4), c16 = lane & 15;
				double xa[MT], yb[NT];
#pragma unroll
				for (int i = 0; i < MT; i++)
					xa[i] = cur->x[kr * FD_BP + wp + i * 16 + c16] * al;
#pragma unroll
				for (int j = 0; j < NT; j++)
					yb[j] = cur->y[kr * FD_BQ + wq + j * 16 + c16];
#pragma unroll
				for (int i = 0; i < MT; i++)
#pragma unroll
					for (int j = 0; j < NT; j++)
						acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
						    yb[j], xa[i], acc[i][j], 0, 0, 0);
			}
		}
		/* this step's stage is read (every LDS value is used by now) and
		 * this wave's DMAs of the next step have landed */
		if (more)
			__builtin_amdgcn_s_waitcnt(WAIT);
		else
			__builtin_amdgcn_s_waitcnt(0);
		__atomic_signal_fence(__ATOMIC_SEQ_CST);
		__builtin_amdgcn_s_barrier();
		__atomic_signal_fence(__ATOMIC_SEQ_CST);
		advance(ip, kc, pr);
	};
	int cs = 0;
	for (; n; n--) {
		const int ns = cs + FD_NST - 1 < FD_NST ? cs + FD_NST - 1 : cs - 1;
		step(&st[cs], &st[ns]);
		cs = cs + 1 < FD_NST ? cs + 1 : 0;
	}

	/* D[q][p] of v_mfma_f64_16x16x4_f64: lane -> p = lane % 16,
	 * register r -> q = 4 * r + lane / 16 */
	if (tl.ksplit) {
		ksplit_store<MT, NT, FD_BP>(ksw, tl, wp, wq, lane, acc);
		return;
	}
	double *C = (double *)(uintptr_t)cb.c;
	const bool bz = sc.flags & 2;
	const bool none = pbeg == pend;
#pragma unroll
	for (int i = 0; i < MT; i++)
#pragma unroll
		for (int j = 0; j < NT; j++)
#pragma unroll
			for (int r = 0; r < 4; r++) {
				uint32_t p = tl.p0 + wp + i * 16 + (lane & 15);
				uint32_t q = tl.q0 + wq + j * 16 + 4 * r + (lane >> 4);
				if (p < cb.p && q < cb.q) {
					size_t o = (size_t)p + (size_t)q * cb.ldc;
					double v;
					if (bz)
						v = none ? 0.0 : acc[i][j][r];
					else if (none)
						v = __dmul_rn(sc.beta_re, C[o]);
					else
						v = __dadd_rn(__dmul_rn(sc.beta_re, C[o]),
						    acc[i][j][r]);
					C[o] = v;
				}
			}
}

/* One tile per workgroup, in launch order. */
template <int FD_BK, int FD_NST, int MINW>
__global__ void __launch_bounds__(GEMM_WG, MINW)
k_gemm_f64_dma(const gemm_tile *__restrict__ tiles,
    const gemm_cblock *__restrict__ cbs, const gemm_pair *__restrict__ pairs,
    const gemm_scalars *__restrict__ scp, double *__restrict__ ksw)
{
	__shared__ fd_stage<FD_BK> st[FD_NST];
	f64_dma_tile<FD_BK, FD_NST>(tiles[blockIdx.x], cbs, pairs, scp, ksw, st);
}

/* Persistent form: a grid of resident workgroups takes tiles from the
 * XCD lists until all are empty (tile_next). */
template <int FD_BK, int FD_NST, int MINW>
__global__ void __launch_bounds__(GEMM_WG, MINW)
k_gemm_f64_dma_q(const gemm_tile *__restrict__ tiles, uint32_t per,
    uint32_t *__restrict__ queue, const gemm_cblock *__restrict__ cbs,
    const gemm_pair *__restrict__ pairs, const gemm_scalars *__restrict__ scp,
    double *__restrict__ ksw)
{
	__shared__ fd_stage<FD_BK> st[FD_NST];
	__shared__ uint32_t next;
	tile_queue q(queue, per);
	for (;;) {
		if (threadIdx.x == 0)
			next = q.take();
		__syncthreads();
		const uint32_t t = next;
		__syncthreads();
		if (t == TILE_NONE)
			return;
		f64_dma_tile<FD_BK, FD_NST>(tiles[t], cbs, pairs, scp, ksw, st);
	}
}

/*
 * Stacked form of the DMA kernel: 128 x 128 tiles (gemm_stile), 8 waves as
 * 4 x 2 of 32 x 64, the same 16-k steps, DMAs and per-output k order as
 * k_gemm_f64_dma, so results are identical to it without a K split.  The
 * tile's rows are four 32-row tiles, each of its own C block: C blocks
 * whose pairs read the same Y panels in the same order with the same k
 * and factors (the ladder's C blocks of one v(a,b) block and one irrep of
 * (i,j)) stack their rows, so a tile runs 128 rows without padding a
 * C block of 144 or 400 rows to a multiple of 128, and stages 1.5 times
 * fewer operand bytes per flop than 64 x 128.  Lane l moves rows 2l, 2l+1
 * of an X column (row tile l / 16, whose C block's X panel for the
 * current pair comes from four scalar loads per pair); a Y DMA moves a
 * 128-row column as before.  Two workgroups a CU (64 KiB of stages each).
 */
#define FS_BP 128
#define FS_WG 512
/* doubles of one half's partial sums of a split stacked tile */
#define FS_KSLOT (2 * KSPLIT_SLOT)
static_assert(FS_BP * FD_BQ <= FS_KSLOT, "a stacked tile fits its slot");

/* C = beta * C + (half 0 + half 1) over one split stacked tile per
 * workgroup (fix[] holds the first halves); row tile r of the tile holds
 * rows p0[r] .. of C block cb[r]. */
__global__ void __launch_bounds__(GEMM_WG)
k_ksplit_sum_s(const gemm_stile *__restrict__ fix,
    const gemm_cblock *__restrict__ cbs, const double *__restrict__ ksw,
    const gemm_scalars *__restrict__ scp)
{
	const gemm_stile tl = fix[blockIdx.x];
	const gemm_scalars sc = *scp;
	const double *W0 = ksw + (size_t)(tl.head.ksplit - 1) * FS_KSLOT;
	const double *W1 = W0 + FS_KSLOT;
	const bool bz = sc.flags & 2, none = sc.flags & 1;
	for (int e = threadIdx.x; e < FS_BP * FD_BQ; e += GEMM_WG) {
		const int pt = e % FS_BP, r = pt / 32;
		const uint32_t m = r == 0 ? tl.cb[0] : r == 1 ? tl.cb[1] :
		    r == 2 ? tl.cb[2] : tl.cb[3];
		if (m == TILE_ROWS_NONE)
			continue;
		const uint32_t p0 = r == 0 ? tl.p0[0] : r == 1 ? tl.p0[1] :
		    r == 2 ? tl.p0[2] : tl.p0[3];
		const gemm_cblock cb = cbs[m];
		const uint32_t p = p0 + pt % 32, q = tl.head.q0 + e / FS_BP;
		if (p < cb.p && q < cb.q) {
			double *C = (double *)(uintptr_t)cb.c;
			const size_t o = (size_t)p + (size_t)q * cb.ldc;
			double v;
			if (bz)
				v = none ? 0.0 : __dadd_rn(W0[e], W1[e]);
			else if (none)
				v = __dmul_rn(sc.beta_re, C[o]);
			else
				v = __dadd_rn(__dmul_rn(sc.beta_re, C[o]),
				    __dadd_rn(W0[e], W1[e]));
			C[o] = v;
		}
	}
}

template <int FD_BK>
struct fs_stage {
	double x[FD_BK * FS_BP];	/* [k][p] */
	double y[FD_BK * FD_BQ];	/* [k][q] */
};

template <int FD_BK, int FD_NST, int MINW>
__global__ void __launch_bounds__(FS_WG, MINW)
k_gemm_f64_dma_s(const gemm_stile *__restrict__ tiles,
    const gemm_cblock *__restrict__ cbs, const gemm_pair *__restrict__ pairs,
    const gemm_scalars *__restrict__ scp, double *__restrict__ ksw)
{
	constexpr int MT = 2, NT = 4;
	constexpr int NXD = FD_BK, NYD = FD_BK;	/* one column per DMA */
	constexpr int DPW = (NXD + NYD) / 8;	/* per wave */
	static_assert(NXD % 8 == 0, "X DMAs per wave");
	constexpr int WAIT = (DPW * (FD_NST - 2)) | (0x7 << 4) | (0xf << 8);
	typedef fs_stage<FD_BK> stage_t;
	__shared__ stage_t st[FD_NST];

	const gemm_stile tl = tiles[blockIdx.x];
	if (tl.head.cblock == 0xffffffffu)
		return;
	const gemm_cblock cb = cbs[tl.head.cblock];
	const gemm_scalars sc = *scp;
	const int tid = threadIdx.x, lane = tid & 63;
	const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	const int wr = wave & 3;			/* this wave's row tile */
	const int wp = wr * 32, wq = (wave >> 2) * 64;
	const uint32_t q0 = tl.head.q0;
	const uint32_t pend = cb.pair_end;
	const uint32_t pbeg = (sc.flags & 1) ? pend : cb.pair_begin;

	/* the row tiles' C blocks: first pair and row count (0: none; the
	 * first pair is then the first member's, so loads stay in range) */
	uint32_t mb[4], mp[4];
#pragma unroll
	for (int r = 0; r < 4; r++) {
		const bool has = tl.cb[r] != TILE_ROWS_NONE;
		const gemm_cblock m = cbs[has ? tl.cb[r] : tl.head.cblock];
		mb[r] = m.pair_begin;
		mp[r] = has ? m.p : 0;
	}
	/* (selects, not an indexed load: an array indexed at run time
	 * would go to scratch memory) */
	const uint32_t mcb = wr == 0 ? tl.cb[0] : wr == 1 ? tl.cb[1] :
	    wr == 2 ? tl.cb[2] : tl.cb[3];
	const uint32_t pr0 = wr == 0 ? tl.p0[0] : wr == 1 ? tl.p0[1] :
	    wr == 2 ? tl.p0[2] : tl.p0[3];
	const bool live = mcb != TILE_ROWS_NONE && q0 + wq < cb.q;
	/* this lane's X rows in the DMAs: row tile lr, rows lp, lp + 1 of its
	 * C block, which has lP rows */
	const int lr = lane >> 4;
	uint32_t lp = tl.p0[0], lP = mp[0];
#pragma unroll
	for (int r = 1; r < 4; r++) {
		lp = lr == r ? tl.p0[r] : lp;
		lP = lr == r ? mp[r] : lP;
	}
	lp += 2 * (lane & 15);

	d4 acc[MT][NT];
#pragma unroll
	for (int i = 0; i < MT; i++)
#pragma unroll
		for (int j = 0; j < NT; j++)
			acc[i][j] = d4{0, 0, 0, 0};

	/* the lane's X panel and leading dimension for pair ri of the
	 * C blocks' lists */
	/* (four scalar loads and per-lane selects: a per-lane load would
	 * count in vmcnt, and its wait would drain the DMAs in flight) */
	auto xpanel = [&](uint32_t ri, uint64_t &x, uint32_t &ldx) {
		uint32_t xl[4], xh[4], ld[4];
#pragma unroll
		for (int r = 0; r < 4; r++) {
			const gemm_pair *g = pairs + mb[r] + ri;
			xl[r] = __builtin_amdgcn_readfirstlane((uint32_t)g->x);
			xh[r] = __builtin_amdgcn_readfirstlane(
			    (uint32_t)(g->x >> 32));
			ld[r] = __builtin_amdgcn_readfirstlane(g->ldx);
		}
		uint32_t lo = xl[0], hi = xh[0];
		ldx = ld[0];
#pragma unroll
		for (int r = 1; r < 4; r++) {
			lo = lr == r ? xl[r] : lo;
			hi = lr == r ? xh[r] : hi;
			ldx = lr == r ? ld[r] : ldx;
		}
		x = (uint64_t)hi << 32 | lo;
	};
	/* step at column k0 of pair px (X of this lane: fx, fldx) into
	 * stage nx: wave w issues X and Y DMAs for columns w, w + 8 */
	auto fetch = [&](stage_t *__restrict__ nx, const gemm_pair px,
	    uint64_t fx, uint32_t fldx, uint32_t k0) {
		const double *X = (const double *)(uintptr_t)fx;
		const double *Y = (const double *)(uintptr_t)px.y;
#pragma unroll
		for (int d = 0; d < NXD / 8; d++) {
			const int col = wave + 8 * d;
			const uint32_t k = k0 + col;
			const void *src = &fd_zero_page[lane];
			if (k < px.k && lp < lP)
				src = X + (size_t)lp + (size_t)k * fldx;
			__builtin_amdgcn_global_load_lds((fd_gbl_t *)src,
			    (fd_lds_t *)&nx->x[col * FS_BP], 16, 0, 0);
		}
#pragma unroll
		for (int d = 0; d < NYD / 8; d++) {
			const int col = wave + 8 * d;
			const uint32_t k = k0 + col, q = q0 + 2 * lane;
			const void *src = &fd_zero_page[lane];
			if (k < px.k && q < cb.q)
				src = Y + (size_t)q + (size_t)k * px.ldy;
			__builtin_amdgcn_global_load_lds((fd_gbl_t *)src,
			    (fd_lds_t *)&nx->y[col * FD_BQ], 16, 0, 0);
		}
	};

	/* compute cursor (ip, kc, pr) with n steps left; fetch cursor
	 * FD_NST - 1 steps ahead with fn left, its lane X in (fx, fldx) */
	/* (a split tile runs one half of the steps; the stack's C blocks
	 * have the same pair sequence, so the head's cut is everyone's) */
	uint32_t ip, kc, n;
	ksplit_steps(tl.head, cb, pbeg, ip, kc, n);
	uint32_t fip = ip, fk = kc, fn = n;
	gemm_pair pr, fpr;
	uint64_t fx = 0;
	uint32_t fldx = 0;
	if (n) {
		pr = pairs[ip];
		fpr = pr;
		xpanel(ip - cb.pair_begin, fx, fldx);
	}
	auto fadvance = [&]() {
		fk += FD_BK;
		if (fk >= fpr.k) {
			fk = 0;
			if (++fip < pend) {
				fpr = pairs[fip];
				xpanel(fip - cb.pair_begin, fx, fldx);
			}
		}
	};
#pragma unroll
	for (int s = 0; s < FD_NST - 1; s++)
		if (fn) {
			fetch(&st[s], fpr, fx, fldx, fk);
			fadvance();
			fn--;
		}
	__builtin_amdgcn_s_waitcnt(0);
	__syncthreads();
	auto step = [&](const stage_t *__restrict__ cur,
	    stage_t *__restrict__ nxt) {
		const bool more = fn != 0;
		if (more) {
			fetch(nxt, fpr, fx, fldx, fk);
			fadvance();
			fn--;
		}
		if (live) {
			const double al = __dmul_rn(sc.alpha_re, pr.a_re);
#pragma unroll
			for (int kk = 0; kk < FD_BK; kk += 4) {
				const int kr = kk + (lane >> 4), c16 = lane & 15;
				double xa[MT], yb[NT];
#pragma unroll
				for (int i = 0; i < MT; i++)
					xa[i] = cur->x[kr * FS_BP + wp + i * 16 + c16] * al;
#pragma unroll
				for (int j = 0; j < NT; j++)
					yb[j] = cur->y[kr * FD_BQ + wq + j * 16 + c16];
#pragma unroll
				for (int i = 0; i < MT; i++)
#pragma unroll
					for (int j = 0; j < NT; j++)
						acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
						    yb[j], xa[i], acc[i][j], 0, 0, 0);
			}
		}
		if (more)
			__builtin_amdgcn_s_waitcnt(WAIT);
		else
			__builtin_amdgcn_s_waitcnt(0);
		__atomic_signal_fence(__ATOMIC_SEQ_CST);
		__builtin_amdgcn_s_barrier();
		__atomic_signal_fence(__ATOMIC_SEQ_CST);
		kc += FD_BK;
		if (kc >= pr.k) {
			kc = 0;
			if (++ip < pend)
				pr = pairs[ip];
		}
	};
	int cs = 0;
	for (; n; n--) {
		const int ns = cs + FD_NST - 1 < FD_NST ? cs + FD_NST - 1 : cs - 1;
		step(&st[cs], &st[ns]);
		cs = cs + 1 < FD_NST ? cs + 1 : 0;
	}

	/* this wave's rows: rows p0[wr] + wp' of C block mcb */
	if (mcb == TILE_ROWS_NONE)
		return;
	if (tl.head.ksplit) {
		/* the half's sums, element (p, q) of the tile at p + 128 q */
		double *W = ksw + (size_t)(tl.head.ksplit - 1) * FS_KSLOT;
#pragma unroll
		for (int i = 0; i < MT; i++)
#pragma unroll
			for (int j = 0; j < NT; j++)
#pragma unroll
				for (int r = 0; r < 4; r++) {
					const int p = wp + i * 16 + (lane & 15);
					const int q = wq + j * 16 + 4 * r + (lane >> 4);
					W[p + q * FS_BP] = acc[i][j][r];
				}
		return;
	}
	const gemm_cblock mc = cbs[mcb];
	double *C = (double *)(uintptr_t)mc.c;
	const bool bz = sc.flags & 2;
	const bool none = pbeg == pend;
#pragma unroll
	for (int i = 0; i < MT; i++)
#pragma unroll
		for (int j = 0; j < NT; j++)
#pragma unroll
			for (int r = 0; r < 4; r++) {
				uint32_t p = pr0 + i * 16 + (lane & 15);
				uint32_t q = q0 + wq + j * 16 + 4 * r + (lane >> 4);
				if (p < mc.p && q < mc.q) {
					size_t o = (size_t)p + (size_t)q * mc.ldc;
					double v;
					if (bz)
						v = none ? 0.0 : acc[i][j][r];
					else if (none)
						v = __dmul_rn(sc.beta_re, C[o]);
					else
						v = __dadd_rn(__dmul_rn(sc.beta_re, C[o]),
						    acc[i][j][r]);
					C[o] = v;
				}
			}
}

/* ---- FP32 / complex / reference kernel ---------------------------- */

template <typename R> struct cx { R re, im; };

template <typename T> struct ops;
template <> struct ops<float> {
	typedef float T;
	static __device__ T zero() { return 0.f; }
	static __device__ T mk(double re, double) { return (float)re; }
	static __device__ T mul(T a, T b) { return a * b; }
	static __device__ T fma(T a, T b, T c) { return __fmaf_rn(a, b, c); }
	static __device__ T add(T a, T b) { return a + b; }
};
template <> struct ops<double> {
	typedef double T;
	static __device__ T zero() { return 0.0; }
	static __device__ T mk(double re, double) { return re; }
	static __device__ T mul(T a, T b) { return a * b; }
	static __device__ T fma(T a, T b, T c) { return __fma_rn(a, b, c); }
	static __device__ T add(T a, T b) { return a + b; }
};
template <typename R> struct ops<cx<R>> {
	typedef cx<R> T;
	static __device__ T zero() { return T{0, 0}; }
	static __device__ T mk(double re, double im) { return T{(R)re, (R)im}; }
	static __device__ T mul(T a, T b) {
		return T{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
	}
	static __device__ T fma(T a, T b, T c) {
		c.re = ops<R>::fma(a.re, b.re, c.re);
		c.re = ops<R>::fma(-a.im, b.im, c.re);
		c.im = ops<R>::fma(a.re, b.im, c.im);
		c.im = ops<R>::fma(a.im, b.re, c.im);
		return c;
	}
	static __device__ T add(T a, T b) { return T{a.re + b.re, a.im + b.im}; }
};

/* 64 x 64 outputs per workgroup, 16 x 16 threads, 4 x 4 per thread. */
template <typename T>
__global__ void __launch_bounds__(GEMM_WG)
k_gemm_simple(const gemm_tile *__restrict__ tiles,
    const gemm_cblock *__restrict__ cbs, const gemm_pair *__restrict__ pairs,
    const gemm_scalars *__restrict__ scp)
{
	typedef ops<T> O;
	constexpr int B = 64, BK = 16;
	__shared__ T Xs[BK][B + 1];
	__shared__ T Ys[BK][B + 1];
	const gemm_tile tl = tiles[blockIdx.x];
	if (tl.cblock == 0xffffffffu)
		return;
	const gemm_cblock cb = cbs[tl.cblock];
	const gemm_scalars sc = *scp;
	const int tid = threadIdx.x, tp = tid % 16, tq = tid / 16;
	const T alpha = O::mk(sc.alpha_re, sc.alpha_im);
	const T beta = O::mk(sc.beta_re, sc.beta_im);
	T acc[4][4];
	for (int i = 0; i < 4; i++)
		for (int j = 0; j < 4; j++)
			acc[i][j] = O::zero();
	const uint32_t pbeg = (sc.flags & 1) ? cb.pair_end : cb.pair_begin;
	for (uint32_t ip = pbeg; ip < cb.pair_end; ip++) {
		const gemm_pair pr = pairs[ip];
		const T al = O::mul(alpha, O::mk(pr.a_re, pr.a_im));
		const T *X = (const T *)(uintptr_t)pr.x;
		const T *Y = (const T *)(uintptr_t)pr.y;
		for (uint32_t k0 = 0; k0 < pr.k; k0 += BK) {
			__syncthreads();
			for (int e = 0; e < B * BK / GEMM_WG; e++) {
				int idx = tid + GEMM_WG * e, p, k;
				if (pr.xmode == PANEL_PCONTIG) { p = idx % B; k = idx / B; }
				else { k = idx % BK; p = idx / BK; }
				uint32_t gp = tl.p0 + p, gk = k0 + k;
				T v = O::zero();
				if (gp < cb.p && gk < pr.k)
					v = O::mul(X[pr.xmode == PANEL_PCONTIG ?
					    (size_t)gp + (size_t)gk * pr.ldx :
					    (size_t)gk + (size_t)gp * pr.ldx], al);
				Xs[k][p] = v;
				if (pr.ymode == PANEL_PCONTIG) { p = idx % B; k = idx / B; }
				else { k = idx % BK; p = idx / BK; }
				uint32_t gq = tl.q0 + p;
				gk = k0 + k;
				v = O::zero();
				if (gq < cb.q && gk < pr.k)
					v = Y[pr.ymode == PANEL_PCONTIG ?
					    (size_t)gq + (size_t)gk * pr.ldy :
					    (size_t)gk + (size_t)gq * pr.ldy];
				Ys[k][p] = v;
			}
			__syncthreads();
			for (int k = 0; k < BK; k++) {
				T xa[4], yb[4];
				for (int i = 0; i < 4; i++)
					xa[i] = Xs[k][tp + 16 * i];
				for (int j = 0; j < 4; j++)
					yb[j] = Ys[k][tq + 16 * j];
				for (int i = 0; i < 4; i++)
					for (int j = 0; j < 4; j++)
						acc[i][j] = O::fma(xa[i], yb[j], acc[i][j]);
			}
		}
	}
	T *C = (T *)(uintptr_t)cb.c;
	const bool bz = sc.flags & 2, none = pbeg == cb.pair_end;
	for (int i = 0; i < 4; i++)
		for (int j = 0; j < 4; j++) {
			uint32_t p = tl.p0 + tp + 16 * i, q = tl.q0 + tq + 16 * j;
			if (p < cb.p && q < cb.q) {
				size_t o = (size_t)p + (size_t)q * cb.ldc;
				T v;
				if (bz)
					v = none ? O::zero() : acc[i][j];
				else if (none)
					v = O::mul(beta, C[o]);
				else
					v = O::add(O::mul(beta, C[o]), acc[i][j]);
				C[o] = v;
			}
		}
}

/* ---- FP32, complex64 and complex128 on MFMA ------------------------ */

/*
 * FP32 / complex64 / complex128 tile kernel for plans whose every X and Y
 * panel is P-contiguous: 64 x BQ outputs, 4 waves as 2 x 2 of 32 x BQ/2,
 * each a set of 16 x 16 x 4 MFMA blocks (v_mfma_f32_16x16x4_f32, or
 * v_mfma_f64_16x16x4_f64 for complex128).  Operands move global -> LDS by
 * dword global_load_lds (lane l moves word l of a 64-word run of the stage,
 * laid out [k][row][re, im]; any extents and alignment), two stages with
 * the DMA one M32_BK-k step ahead (16, 8 for complex128), a counted wait and a bare s_barrier per step
 * (the FP64 kernel's scheme).  A complex product takes four real MFMAs into
 * a real and an imaginary accumulator: re += yr xr - yi xi, im += yr xi +
 * yi xr.  The pair factor multiplies X in registers.  With Y as the A
 * operand P runs along the lanes: the FP32 MFMA leaves row 4 (lane / 16) +
 * r of column lane % 16 in register r, the FP64 one row 4 r + lane / 16
 * (both checked on the device).
 */
typedef float f4 __attribute__((ext_vector_type(4)));

/* element -> real type R and components C */
template <typename T> struct mdw { typedef float R; static constexpr int C = 1; };
template <> struct mdw<cx<float>> { typedef float R; static constexpr int C = 2; };
template <> struct mdw<cx<double>> { typedef double R; static constexpr int C = 2; };

__device__ __forceinline__ f4
mfma16(float a, float b, f4 c)
{
	return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ d4
mfma16(double a, double b, d4 c)
{
	return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

/* One wave-wide global -> LDS DMA of DW words a lane (the size must be a
 * literal). */
template <int DW>
__device__ __forceinline__ void
lds_dma(const void *src, void *dst)
{
	if constexpr (DW == 4)
		__builtin_amdgcn_global_load_lds((fd_gbl_t *)src,
		    (fd_lds_t *)dst, 16, 0, 0);
	else
		__builtin_amdgcn_global_load_lds((fd_gbl_t *)src,
		    (fd_lds_t *)dst, 4, 0, 0);
}

/* DW = 4 (plans whose panels allow it): a lane's DMA moves 16 bytes of
 * one column, 4 consecutive FP32 rows, 2 complex64 rows or one complex128
 * row, a quarter of the DMAs (and of their address arithmetic) of the
 * dword form. */
template <typename T, int BQ, int M32_BK, int DW = 1>
__global__ void __launch_bounds__(GEMM_WG,
    sizeof(typename mdw<T>::R) == 8 ? 2 : 4)
k_gemm_mfma_dw(const gemm_tile *__restrict__ tiles,
    const gemm_cblock *__restrict__ cbs, const gemm_pair *__restrict__ pairs,
    const gemm_scalars *__restrict__ scp)
{
	typedef typename mdw<T>::R R;
	typedef R acc_t __attribute__((ext_vector_type(4)));
	constexpr int NC = mdw<T>::C;		/* components per element */
	constexpr bool CX = NC == 2;
	constexpr int W = NC * (int)sizeof(R) / 4;	/* words per element */
	constexpr int NT = BQ / 32;		/* 16-column blocks per wave */
	constexpr int XW = 64 * W, YW = BQ * W;	/* words of one column */
	constexpr int SX = M32_BK * XW, SY = M32_BK * YW;	/* of a stage */
	static_assert(SX % (256 * DW) == 0 && SY % (256 * DW) == 0,
	    "DMAs per wave");
	static_assert(DW == 1 || 4 % W == 0, "16-byte DMAs: whole elements");
	__shared__ __align__(16) float st[2][SX + SY];	/* [stage][X, Y][k][row] */

	const gemm_tile tl = tiles[blockIdx.x];
	if (tl.cblock == 0xffffffffu)
		return;
	const gemm_cblock cb = cbs[tl.cblock];
	const gemm_scalars sc = *scp;
	const int tid = threadIdx.x, lane = tid & 63;
	const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	const int wp = (wave & 1) * 32, wq = (wave >> 1) * (BQ / 2);
	const bool live = tl.p0 + wp < cb.p && tl.q0 + wq < cb.q;
	const uint32_t pend = cb.pair_end;
	const uint32_t pbeg = (sc.flags & 1) ? pend : cb.pair_begin;

	acc_t acc[2][NT][NC];
#pragma unroll
	for (int i = 0; i < 2; i++)
#pragma unroll
		for (int j = 0; j < NT; j++)
#pragma unroll
			for (int c = 0; c < NC; c++)
				acc[i][j][c] = acc_t{0, 0, 0, 0};

	/* step at column k0 of pair px into stage nx (X words, then Y words
	 * at SX): DMA d of an operand moves its words [64 DW d, 64 DW (d + 1));
	 * wave w issues d = w, w + 4, ... */
	auto fetch_op = [&](float *__restrict__ nz, const float *Z, uint32_t ld,
	    uint32_t z0, uint32_t zn, uint32_t kn, uint32_t k0, auto cw,
	    auto sw) {
		constexpr int CW = decltype(cw)::value, SWD = decltype(sw)::value;
#pragma unroll
		for (int d = 0; d < SWD / (256 * DW); d++) {
			const int dd = wave + 4 * d, e = (dd * 64 + lane) * DW;
			const uint32_t k = k0 + e / CW;
			const uint32_t r = z0 + (e % CW) / W;
			const void *src = (const float *)fd_zero_page + lane * DW;
			if (k < kn && r < zn)
				src = Z + ((size_t)r + (size_t)k * ld) * W + e % W;
			lds_dma<DW>(src, &nz[dd * 64 * DW]);
		}
	};
	auto fetch = [&](float *__restrict__ nx, const gemm_pair px,
	    uint32_t k0) {
		fetch_op(nx, (const float *)(uintptr_t)px.x, px.ldx, tl.p0, cb.p,
		    px.k, k0, std::integral_constant<int, XW>(),
		    std::integral_constant<int, SX>());
		fetch_op(nx + SX, (const float *)(uintptr_t)px.y, px.ldy, tl.q0,
		    cb.q, px.k, k0, std::integral_constant<int, YW>(),
		    std::integral_constant<int, SY>());
	};
	auto advance = [&](uint32_t &i, uint32_t &k, gemm_pair &p) {
		k += M32_BK;
		if (k >= p.k) {
			k = 0;
			if (++i < pend)
				p = pairs[i];
		}
	};
	uint32_t ip = pbeg, kc = 0, fip = pbeg, fk = 0;
	gemm_pair pr, fpr;
	if (ip < pend) {
		pr = pairs[ip];
		fpr = pr;
		fetch(st[0], fpr, fk);
		advance(fip, fk, fpr);
	}
	__builtin_amdgcn_s_waitcnt(0);
	__syncthreads();
	/* the stages reach step() as restrict pointers, so the LDS reads of
	 * this stage need not wait for the DMA into the other */
	auto step = [&](const float *__restrict__ cur, float *__restrict__ nxt) {
		const bool more = fip < pend;
		if (more) {
			fetch(nxt, fpr, fk);
			advance(fip, fk, fpr);
		}
		if (live) {
			const R ar = (R)(sc.alpha_re * pr.a_re -
			    sc.alpha_im * pr.a_im);
			const R ai = (R)(sc.alpha_re * pr.a_im +
			    sc.alpha_im * pr.a_re);
#pragma unroll
			for (int kk = 0; kk < M32_BK; kk += 4) {
				const int kr = kk + (lane >> 4), c16 = lane & 15;
				R xa[2][NC], yb[NT][NC];
#pragma unroll
				for (int i = 0; i < 2; i++) {
					const R *x = (const R *)(cur + kr * XW +
					    (wp + i * 16 + c16) * W);
					if (CX) {
						const R xr = x[0], xi = x[NC - 1];
						xa[i][0] = xr * ar - xi * ai;
						xa[i][NC - 1] = xr * ai + xi * ar;
					} else {
						xa[i][0] = x[0] * ar;
					}
				}
#pragma unroll
				for (int j = 0; j < NT; j++) {
					const R *y = (const R *)(cur + SX + kr * YW +
					    (wq + j * 16 + c16) * W);
#pragma unroll
					for (int c = 0; c < NC; c++)
						yb[j][c] = y[c];
				}
#pragma unroll
				for (int i = 0; i < 2; i++)
#pragma unroll
					for (int j = 0; j < NT; j++) {
						acc[i][j][0] = mfma16(yb[j][0], xa[i][0],
						    acc[i][j][0]);
						if (CX) {
							acc[i][j][0] = mfma16(-yb[j][NC - 1],
							    xa[i][NC - 1], acc[i][j][0]);
							acc[i][j][NC - 1] = mfma16(yb[j][0],
							    xa[i][NC - 1], acc[i][j][NC - 1]);
							acc[i][j][NC - 1] = mfma16(yb[j][NC - 1],
							    xa[i][0], acc[i][j][NC - 1]);
						}
					}
			}
		}
		/* this stage is read and this wave's DMAs of the next step have
		 * landed; a bare s_barrier then publishes the next stage */
		__builtin_amdgcn_s_waitcnt(0);
		__atomic_signal_fence(__ATOMIC_SEQ_CST);
		__builtin_amdgcn_s_barrier();
		__atomic_signal_fence(__ATOMIC_SEQ_CST);
		advance(ip, kc, pr);
	};
	int cs = 0;
	while (ip < pend) {
		step(st[cs], st[cs ^ 1]);
		cs ^= 1;
	}

	typedef ops<T> O;
	T *C = (T *)(uintptr_t)cb.c;
	const T beta = O::mk(sc.beta_re, sc.beta_im);
	const bool bz = sc.flags & 2, none = pbeg == pend;
#pragma unroll
	for (int i = 0; i < 2; i++)
#pragma unroll
		for (int j = 0; j < NT; j++)
#pragma unroll
			for (int r = 0; r < 4; r++) {
				const uint32_t p = tl.p0 + wp + i * 16 + (lane & 15);
				const uint32_t q = tl.q0 + wq + j * 16 + (sizeof(R) == 4 ?
				    4 * (lane >> 4) + r : 4 * r + (lane >> 4));
				if (p < cb.p && q < cb.q) {
					const size_t o = (size_t)p + (size_t)q * cb.ldc;
					const T a = O::mk(acc[i][j][0][r],
					    acc[i][j][NC - 1][r]);
					T v;
					if (bz)
						v = none ? O::zero() : a;
					else if (none)
						v = O::mul(beta, C[o]);
					else
						v = O::add(O::mul(beta, C[o]), a);
					C[o] = v;
				}
			}
}

/* ---- tiles and launch --------------------------------------------- */

static bool
use_simple(void)
{
	static int v = -1;
	if (v < 0) {
		const char *e = getenv("XM_GEMM_SIMPLE");
		v = e && *e == '1';
	}
	return v;
}

/*
 * Tile-shape candidates of a plan.  The FP64 kernel is compiled for
 * 64 x 128 and 64 x 64 output tiles: 64 x 128 moves fewer operand bytes
 * per FLOP, 64 x 64 wastes less on small C blocks (P, Q of a few 16s).  The
 * first calls of a plan time each candidate once (the GEMM's own HIP
 * events); later calls launch the faster one.  Decisions are also kept per
 * block-size histogram signature, so a new plan whose C blocks have the
 * same shape distribution starts with the decision.  Every candidate gives
 * bit-identical results: each output element accumulates over the same k
 * sequence in the same order.  XM_GEMM_TILE=64x128 or 64x64 pins a shape.
 * The split form has one shape per slice count; FP32 and complex one.
 */
static void
tile_candidates(const contract_plan *pl, std::vector<std::pair<uint32_t,
    uint32_t>> &out)
{
	if (pl->precision == XM_PRECISION_SPLIT_BF16) {
		out.push_back(pl->split_staged ?
		    std::make_pair(128u, 128u) : std::make_pair(64u, 64u));
	} else if (pl->precision == XM_PRECISION_SPLIT_INT8) {
		out.push_back({256, 256});
	} else if (pl->type == XM_SCALAR_DOUBLE && !use_simple()) {
		/* the stacked 128 x 128 form where the DMA kernel can run,
		 * then 64 x 128 (the no-allocation entry point's plans take the
		 * first candidate and have no storage for stacked tiles) */
		const char *e = getenv("XM_GEMM_TILE");
		const bool st = pl->f64_dma && !pl->single_tile;
		if (e && strcmp(e, "64x64") == 0)
			out.push_back({64, 64});
		else if (e && strcmp(e, "64x128") == 0)
			out.push_back({64, 128});
		else if (e && strcmp(e, "128x128") == 0 && st)
			out.push_back({128, 128});
		else if (st) {
			out.push_back({128, 128});
			out.push_back({64, 128});
		} else {
			out.push_back({64, 128});
			out.push_back({64, 64});
		}
	} else if (pl->mfma_dw && pl->type == XM_SCALAR_FLOAT) {
		/* (complex64 at 64 x 128 would hold 259 registers: 1 wave) */
		const char *e = getenv("XM_GEMM_TILE");
		if (e && strcmp(e, "64x64") == 0)
			out.push_back({64, 64});
		else if (e && strcmp(e, "64x128") == 0)
			out.push_back({64, 128});
		else {
			out.push_back({64, 128});
			out.push_back({64, 64});
		}
	} else {
		out.push_back({64, 64});
	}
	if (pl->single_tile)
		out.resize(1);
}

/* decided candidate per block-size histogram signature */
static std::map<std::string, int> g_tile_choice;

/*
 * Signature: element type, precision, and the count of C blocks (rounded
 * to a power of two) per (log2 P, log2 Q, log2 K) bucket.
 */
static std::string
tile_signature(const contract_plan *pl)
{
	std::map<std::tuple<int, int, int>, size_t> h;
	auto lg = [](double x) { return x < 1 ? 0 : (int)std::log2(x); };
	for (const gemm_cblock &cb : pl->cblocks) {
		double kt = 0;
		for (uint32_t p = cb.pair_begin; p < cb.pair_end; p++)
			kt += pl->pairs[p].k;
		h[std::make_tuple(lg(cb.p), lg(cb.q), lg(kt))]++;
	}
	std::string sig = std::to_string((int)pl->type) + "/" +
	    std::to_string(pl->precision) + "/" + std::to_string(pl->nslices);
	for (auto &e : h)
		sig += ":" + std::to_string(std::get<0>(e.first)) + "," +
		    std::to_string(std::get<1>(e.first)) + "," +
		    std::to_string(std::get<2>(e.first)) + "x" +
		    std::to_string(lg((double)e.second));
	return sig;
}

#define NXCD 8

/*
 * C blocks per XCD list.  Workgroups w and w + 8 of a launch run on the
 * same XCD (round-robin dispatch) and each XCD has its own L2, so the
 * tiles of one C block are kept on one XCD: C blocks are spread over the 8
 * lists by a greedy longest-first balance of their work.
 */
static void
xcd_lists(const contract_plan *pl, uint32_t BP, uint32_t BQ,
    std::vector<uint32_t> (&lists)[NXCD])
{
	std::vector<std::pair<double, uint32_t>> order;
	for (uint32_t i = 0; i < pl->cblocks.size(); i++) {
		const gemm_cblock &cb = pl->cblocks[i];
		double kt = 1;
		for (uint32_t p = cb.pair_begin; p < cb.pair_end; p++)
			kt += pl->pairs[p].k;
		double nt = (double)((cb.p + BP - 1) / BP) * ((cb.q + BQ - 1) / BQ);
		order.push_back({kt * nt, i});
	}
	std::stable_sort(order.begin(), order.end(),
	    [](const std::pair<double, uint32_t> &a,
	    const std::pair<double, uint32_t> &b) { return a.first > b.first; });
	double load[NXCD] = {0};
	for (auto &o : order) {
		int x = (int)(std::min_element(load, load + NXCD) - load);
		load[x] += o.first;
		lists[x].push_back(o.second);
	}
}

/*
 * Grouped order (XM_TILE_GROUPS=0 turns it off): C blocks whose pairs read
 * the same Y panels in the same
 * order form a group, and the tiles of a group at one Q offset read the
 * same Y fragments step for step.  Items of a group's tiles at QCHUNK
 * consecutive Q offsets are dealt to the XCD lists longest first (each
 * list keeps that order, so its short items come last), and an item lists
 * its tiles C block by C block, P tile by P tile: the workgroups an XCD
 * runs together share Y fragments across C blocks in its L2.  The one
 * workgroup a CU of the int8 kernel starts its tiles in step, and the
 * tiles of an item have equal K.  On the o=64 v=384 ladder the FP64 GEMM
 * went from 18.0 to 17.1-17.5 ms, the BF16 split GEMM from 16.5 to 14.9-15.6
 * and the int8 one from 11.1 to 10.4-10.6; lists ordered by group instead
 * of by length ran 12.1-12.5 ms (int8: their tails drained late), and
 * items of 2, 4 or 8 Q offsets ran no faster than 1.
 */
#define QCHUNK 1

/*
 * A group of one C block with at least 2 TILE_RECT tiles along P and along
 * Q (a dense class: thousands of rows and columns) is cut into items of
 * TILE_RECT x TILE_RECT tiles instead of columns: the tiles an XCD runs
 * together then read TILE_RECT X and TILE_RECT Y panels per k step rather
 * than every X panel of the C block.  TILE_RECT is 8 (XM_TILE_RECT sets
 * it, 0 keeps columns).  On the o=96 v=480 ring (eight classes of about
 * 5800 x 5800 x 5800) the stacked FP64 GEMM ran 46.7 ms against 47.7 with
 * columns (rectangles of 4 and 6: 46.8, of 16: 47.8); its MFMA pipe was
 * busy 85% of the time already with columns, so the HBM reads they cost
 * (106 GB a launch, 2.2 TB/s) were mostly hidden.
 */
static uint32_t
tile_rect(void)
{
	static int v = -1;
	if (v < 0) {
		const char *e = getenv("XM_TILE_RECT");
		v = e && *e ? atoi(e) : 8;
		if (v < 0)
			v = 0;
	}
	return (uint32_t)v;
}

static bool
grouped_order(const contract_plan *pl)
{
	const char *e = getenv("XM_TILE_GROUPS");
	(void)pl;
	return !(e && *e == '0');
}

struct tile_item {
	double work;
	uint32_t g, q0, q1;
	uint32_t p0, p1;	/* rows of a one-C-block group's item */
};

/* The groups, their items and the items of each XCD list (longest
 * first) for tiles of BP x BQ. */
static void
group_items(const contract_plan *pl, uint32_t BP, uint32_t BQ,
    std::vector<std::vector<uint32_t>> &groups, std::vector<tile_item> &items,
    std::vector<size_t> (&lists)[NXCD])
{
	std::map<std::vector<uint64_t>, uint32_t> gid;
	groups.clear();
	items.clear();
	for (uint32_t i = 0; i < pl->cblocks.size(); i++) {
		const gemm_cblock &cb = pl->cblocks[i];
		std::vector<uint64_t> key;
		for (uint32_t p = cb.pair_begin; p < cb.pair_end; p++) {
			key.push_back(pl->pairs[p].y);
			key.push_back(pl->pairs[p].k);
		}
		key.push_back((uint64_t)cb.q);
		if (cb.pair_begin == cb.pair_end)
			key.push_back(0x8000000000000000ull | i);	/* alone */
		auto it = gid.find(key);
		if (it == gid.end()) {
			it = gid.emplace(key, (uint32_t)groups.size()).first;
			groups.emplace_back();
		}
		groups[it->second].push_back(i);
	}
	for (uint32_t g = 0; g < groups.size(); g++) {
		const gemm_cblock &c0 = pl->cblocks[groups[g][0]];
		double kt = 1;
		for (uint32_t p = c0.pair_begin; p < c0.pair_end; p++)
			kt += pl->pairs[p].k;
		double np = 0;
		for (uint32_t i : groups[g])
			np += (pl->cblocks[i].p + BP - 1) / BP;
		const uint32_t TR = tile_rect();
		if (TR && groups[g].size() == 1 && np >= 2 * TR &&
		    c0.q >= 2 * TR * BQ) {
			for (uint32_t p0 = 0; p0 < c0.p; p0 += TR * BP)
				for (uint32_t q0 = 0; q0 < c0.q; q0 += TR * BQ) {
					const uint32_t p1 = std::min(c0.p,
					    p0 + TR * BP);
					const uint32_t q1 = std::min(c0.q,
					    q0 + TR * BQ);
					items.push_back({kt * ((p1 - p0 + BP - 1) /
					    BP) * ((q1 - q0 + BQ - 1) / BQ), g, q0,
					    q1, p0, p1});
				}
			continue;
		}
		for (uint32_t q0 = 0; q0 < c0.q; q0 += QCHUNK * BQ) {
			const uint32_t q1 = std::min(c0.q, q0 + QCHUNK * BQ);
			items.push_back({kt * np * ((q1 - q0 + BQ - 1) / BQ), g, q0,
			    q1, 0, 0xffffffffu});
		}
	}
	std::vector<size_t> order(items.size());
	for (size_t i = 0; i < order.size(); i++)
		order[i] = i;
	std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
		return items[a].work > items[b].work;
	});
	double load[NXCD] = {0};
	for (auto &l : lists)
		l.clear();
	for (size_t o : order) {
		int x = (int)(std::min_element(load, load + NXCD) - load);
		load[x] += items[o].work;
		lists[x].push_back(o);
	}
}

/* Tiles in the grouped order; split C blocks (marked on the grouped lists
 * of 64 x 128 tiles) come as adjacent halves, as in order_tiles. */
static void
order_tiles_grouped(const contract_plan *pl, uint32_t BP, uint32_t BQ,
    std::vector<gemm_tile> &tiles, std::vector<gemm_tile> &fix)
{
	std::vector<std::vector<uint32_t>> groups;
	std::vector<tile_item> items;
	std::vector<size_t> lists[NXCD];
	group_items(pl, BP, BQ, groups, items, lists);
	std::vector<gemm_tile> bins[NXCD];
	fix.clear();
	for (int x = 0; x < NXCD; x++)
		for (size_t o : lists[x]) {
			const tile_item &it = items[o];
			for (uint32_t q0 = it.q0; q0 < it.q1; q0 += BQ)
				for (uint32_t i : groups[it.g]) {
					const gemm_cblock &cb = pl->cblocks[i];
					const uint32_t pe = std::min(cb.p, it.p1);
					for (uint32_t p0 = it.p0; p0 < pe; p0 += BP) {
						if (cb.split_pair == KSPLIT_NONE) {
							bins[x].push_back({i, p0, q0, 0});
							continue;
						}
						const uint32_t k = 1 + 2 *
						    (uint32_t)fix.size();
						bins[x].push_back({i, p0, q0, k});
						bins[x].push_back({i, p0, q0, k + 1});
						fix.push_back({i, p0, q0, k});
					}
				}
		}
	size_t per = 0;
	for (auto &b2 : bins)
		per = std::max(per, b2.size());
	tiles.assign(per * NXCD, gemm_tile{0xffffffffu, 0, 0, 0});
	for (int x = 0; x < NXCD; x++)
		for (size_t s2 = 0; s2 < bins[x].size(); s2++)
			tiles[s2 * NXCD + x] = bins[x][s2];
}

/*
 * Tiles of the stacked forms (records of a head and R row tiles: R = 4,
 * gemm_stile, three gemm_tile records; R = 8, gemm_stile8, five): C blocks
 * with the same Q and the same pair sequence on the Y side (Y panel, its
 * leading dimension, k and factor at every position) form a group, which
 * is one stack (merge) or one stack per C block; a stack's rows are its
 * C blocks' rows in 32-row tiles (a C block of P rows takes ceil(P / 32)),
 * R row tiles to a tile.  Items (a group's tiles at one Q offset) are
 * dealt to the XCD lists longest first, as in the grouped order, so the
 * workgroups an XCD runs together read the same Y fragments.
 */
static uint32_t crt_moduli_parts(void);

template <int R>
struct stacked_order {
	struct rec {
		gemm_tile head;
		uint32_t cb[R], p0[R];
	};
	struct item {
		double work;
		uint32_t g, q0, q1, s0, s1;	/* Q range, record range */
	};
	std::vector<std::vector<uint32_t>> groups;
	std::vector<std::vector<rec>> recs;	/* a group's records */
	std::vector<item> items;
	std::vector<size_t> lists[NXCD];	/* items per XCD, in order */
};

template <int R>
static void
stacked_items(const contract_plan *pl, uint32_t BQ, bool merge,
    stacked_order<R> &so)
{
	typedef typename stacked_order<R>::rec rec;
	static_assert(sizeof(rec) % sizeof(gemm_tile) == 0, "tile records");
	/* groups: C blocks with the same Q and Y-side pair sequence */
	std::map<std::vector<uint64_t>, uint32_t> gid;
	std::vector<std::vector<uint32_t>> &groups = so.groups;
	for (uint32_t i = 0; i < pl->cblocks.size(); i++) {
		const gemm_cblock &cb = pl->cblocks[i];
		std::vector<uint64_t> key = {cb.q, cb.pair_end - cb.pair_begin};
		for (uint32_t p = cb.pair_begin; p < cb.pair_end; p++) {
			const gemm_pair &g = pl->pairs[p];
			uint64_t re, im;
			memcpy(&re, &g.a_re, 8);
			memcpy(&im, &g.a_im, 8);
			key.insert(key.end(), {g.y, g.ldy, g.k, re, im});
		}
		if (cb.pair_begin == cb.pair_end)
			key.push_back(0x8000000000000000ull | i);	/* alone */
		auto it = gid.find(key);
		if (it == gid.end()) {
			it = gid.emplace(key, (uint32_t)groups.size()).first;
			groups.emplace_back();
		}
		groups[it->second].push_back(i);
	}
	/* a group's stacks (all its C blocks, or one each without merge),
	 * their rows in 32-row tiles, R to a tile record */
	so.recs.assign(groups.size(), {});
	for (uint32_t g = 0; g < groups.size(); g++) {
		std::vector<std::vector<std::pair<uint32_t, uint32_t>>> rows;
		for (uint32_t i : groups[g]) {
			if (!merge || rows.empty())
				rows.emplace_back();
			for (uint32_t p0 = 0; p0 < pl->cblocks[i].p; p0 += 32)
				rows.back().push_back({i, p0});
		}
		for (const auto &rw : rows)
			for (size_t r0 = 0; r0 < rw.size(); r0 += R) {
				rec t;
				t.head = {rw[0].first, 0, 0, 0};
				for (size_t r = 0; r < R; r++) {
					const bool in = r0 + r < rw.size();
					t.cb[r] = in ? rw[r0 + r].first :
					    TILE_ROWS_NONE;
					t.p0[r] = in ? rw[r0 + r].second : 0;
				}
				so.recs[g].push_back(t);
			}
		const gemm_cblock &c0 = pl->cblocks[groups[g][0]];
		double kt = 1;
		for (uint32_t p = c0.pair_begin; p < c0.pair_end; p++)
			kt += pl->pairs[p].k;
		const uint32_t nt = (uint32_t)so.recs[g].size(), TR = tile_rect();
		if (TR && nt >= 2 * TR && c0.q >= 2 * TR * BQ) {
			for (uint32_t s0 = 0; s0 < nt; s0 += TR)
				for (uint32_t q0 = 0; q0 < c0.q; q0 += TR * BQ) {
					const uint32_t s1 = std::min(nt, s0 + TR);
					const uint32_t q1 = std::min(c0.q,
					    q0 + TR * BQ);
					so.items.push_back({kt * (s1 - s0) *
					    ((q1 - q0 + BQ - 1) / BQ), g, q0, q1, s0,
					    s1});
				}
			continue;
		}
		for (uint32_t q0 = 0; q0 < c0.q; q0 += BQ)
			so.items.push_back({kt * (double)nt, g, q0,
			    std::min(c0.q, q0 + BQ), 0, nt});
	}
	std::vector<size_t> order(so.items.size());
	for (size_t i = 0; i < order.size(); i++)
		order[i] = i;
	std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
		return so.items[a].work > so.items[b].work;
	});
	double load[NXCD] = {0};
	for (size_t o : order) {
		const int x = (int)(std::min_element(load, load + NXCD) - load);
		load[x] += so.items[o].work;
		so.lists[x].push_back(o);
	}
}

/* The stacked tiles in launch order (slot s of XCD x is workgroup
 * 8 s + x); a tile of a split C block comes as adjacent halves, and `fix`
 * gets the first halves (records of the same size). */
template <int R>
static void
order_tiles_stacked(const contract_plan *pl, uint32_t BQ, bool merge,
    std::vector<gemm_tile> &out, std::vector<gemm_tile> &fix)
{
	typedef typename stacked_order<R>::rec rec;
	stacked_order<R> so;
	stacked_items<R>(pl, BQ, merge, so);
	const uint32_t nsp = R == 8 ? crt_moduli_parts() : 1;
	std::vector<rec> bins[NXCD], fixes;
	for (int x = 0; x < NXCD; x++)
		for (size_t o : so.lists[x]) {
			const typename stacked_order<R>::item &it = so.items[o];
			if (R == 8 && nsp > 1) {
				/* int8: the item's tiles part by part, so the
				 * tiles running together read the same
				 * moduli's planes */
				for (uint32_t j = 0; j < nsp; j++)
					for (uint32_t q0 = it.q0; q0 < it.q1; q0 += BQ)
						for (uint32_t s = it.s0; s < it.s1; s++) {
							rec t = so.recs[it.g][s];
							t.head.q0 = q0;
							t.head.ksplit = nsp << 8 | (j + 1);
							bins[x].push_back(t);
						}
				continue;
			}
			for (uint32_t q0 = it.q0; q0 < it.q1; q0 += BQ)
				for (uint32_t s = it.s0; s < it.s1; s++) {
					rec t = so.recs[it.g][s];
					t.head.q0 = q0;
					if (pl->cblocks[t.head.cblock].split_pair ==
					    KSPLIT_NONE) {
						bins[x].push_back(t);
						continue;
					}
					const uint32_t k = 1 + 2 *
					    (uint32_t)fixes.size();
					t.head.ksplit = k;
					bins[x].push_back(t);
					fixes.push_back(t);
					t.head.ksplit = k + 1;
					bins[x].push_back(t);
				}
		}
	size_t per = 0;
	for (auto &b : bins)
		per = std::max(per, b.size());
	rec none;
	memset(&none, 0, sizeof none);
	none.head.cblock = 0xffffffffu;
	std::vector<rec> flat(per * NXCD, none);
	for (int x = 0; x < NXCD; x++)
		for (size_t s = 0; s < bins[x].size(); s++)
			flat[s * NXCD + x] = bins[x][s];
	out.resize(flat.size() * (sizeof(rec) / sizeof(gemm_tile)));
	if (!flat.empty())
		memcpy(out.data(), flat.data(), flat.size() * sizeof(rec));
	fix.resize(fixes.size() * (sizeof(rec) / sizeof(gemm_tile)));
	if (!fixes.empty())
		memcpy(fix.data(), fixes.data(), fixes.size() * sizeof(rec));
}
static_assert(sizeof(gemm_stile) == 3 * sizeof(gemm_tile),
    "a stacked FP64 tile is three tile records");
static_assert(sizeof(gemm_stile8) == 5 * sizeof(gemm_tile),
    "a stacked int8 tile is five tile records");

/*
 * K split of FP64 tiles (not with XM_GEMM_KSPLIT=0 or the FMA kernel).
 * An XCD's list ends when its last tiles finish, and with 64 x 128 tiles
 * of up to 18432 k (the o=64 v=384 ladder: 4470 tiles on 768 slots) the
 * slots idled 8.6% of the GEMM's time waiting for the last ones (per-tile
 * start and end stamps).  The C blocks holding the last KSPLIT_TAIL tiles
 * of each list are cut in two halves of their 16-k steps, which halves the
 * granularity where the lists drain; C blocks of fewer than
 * KSPLIT_MIN_STEPS steps are not cut.  The lists are those of the stacked
 * 128 x 128 tiles where a candidate is stacked (the last KSPLIT_TAIL_S
 * tiles of each: the stacked tiles of the ladder, 4.4 a slot, left the
 * slots idle 8.3% of the time, mostly in the last 3 ms; cutting the
 * groups of the last 32, 64, 128 or 256 tiles of each list ran 16.22,
 * 16.19, 16.23 and 16.49 ms against 16.29 uncut), else those of
 * 64 x 128 tiles, whatever the candidates, so every tile shape cuts the
 * same C blocks at the same step and gives the same results (a stacked
 * tile cuts where its first C block does: the C blocks of a stack have the
 * same pair sequence, and a group's C blocks are cut together).  Launches
 * of more than KSPLIT_MAX_TILES 64 x 128 tiles (over 12 per slot of the
 * 768 that 3 workgroups on each of 256 CUs give) drain in a small share of
 * their time and are not split.
 */
#define KSPLIT_TAIL 96
#define KSPLIT_TAIL_S 64
#define KSPLIT_MIN_STEPS 64
#define KSPLIT_MAX_TILES (12 * 768)

static bool
ksplit_enabled(const contract_plan *pl)
{
	const char *e = getenv("XM_GEMM_KSPLIT");
	return !(e && *e == '0') && !pl->no_ksplit &&
	    pl->type == XM_SCALAR_DOUBLE &&
	    pl->precision == XM_PRECISION_NATIVE && !use_simple();
}

static void
mark_ksplit(contract_plan *pl)
{
	const uint32_t BP = 64, BQ = 128;
	for (gemm_cblock &cb : pl->cblocks) {
		cb.split_pair = KSPLIT_NONE;
		cb.split_k = cb.split_steps = 0;
		uint64_t steps = 0;
		for (uint32_t i = cb.pair_begin; i < cb.pair_end; i++)
			steps += (pl->pairs[i].k + 15) / 16;
		if (steps > 0xffffffffu)
			xm_fatal("xm_contract: C block of more than 2^32 k steps");
		cb.steps = (uint32_t)steps;
	}
	if (!ksplit_enabled(pl))
		return;
	bool stacked = false;
	for (int c = 0; c < pl->ncand; c++)
		stacked = stacked || pl->cand_stacked[c];
	size_t ntiles = 0;
	for (const gemm_cblock &cb : pl->cblocks)
		ntiles += (size_t)((cb.p + BP - 1) / BP) * ((cb.q + BQ - 1) / BQ);
	if (ntiles > KSPLIT_MAX_TILES)
		return;
	auto split = [&](gemm_cblock &cb) {
		if (cb.steps < KSPLIT_MIN_STEPS || cb.split_pair != KSPLIT_NONE)
			return;
		/* first pair whose steps reach past the middle */
		const uint32_t mid = cb.steps / 2;
		uint32_t before = 0;
		for (uint32_t i = cb.pair_begin; i < cb.pair_end; i++) {
			const uint32_t ns = (pl->pairs[i].k + 15) / 16;
			if (before + ns > mid) {
				cb.split_pair = i;
				cb.split_k = (mid - before) * 16;
				cb.split_steps = mid;
				break;
			}
			before += ns;
		}
	};
	/* tiles at the end of each list whose C blocks are cut */
	const char *e = getenv("XM_KSPLIT_TAIL");
	const uint32_t want = e && *e ? (uint32_t)atoi(e) : stacked ?
	    KSPLIT_TAIL_S : KSPLIT_TAIL;
	if (stacked) {
		stacked_order<4> so;
		stacked_items<4>(pl, 128, true, so);
		for (auto &l : so.lists) {
			uint32_t tail = 0;
			for (size_t n = l.size(); n-- > 0 && tail < want;) {
				const auto &it = so.items[l[n]];
				tail += (it.s1 - it.s0) *
				    ((it.q1 - it.q0 + 127) / 128);
				for (uint32_t i : so.groups[it.g])
					split(pl->cblocks[i]);
			}
		}
		return;
	}
	if (grouped_order(pl)) {
		std::vector<std::vector<uint32_t>> groups;
		std::vector<tile_item> items;
		std::vector<size_t> lists[NXCD];
		group_items(pl, BP, BQ, groups, items, lists);
		for (auto &l : lists) {
			uint32_t tail = 0;
			for (size_t n = l.size(); n-- > 0 && tail < want;) {
				const tile_item &it = items[l[n]];
				for (uint32_t i : groups[it.g]) {
					gemm_cblock &cb = pl->cblocks[i];
					tail += ((cb.p + BP - 1) / BP) *
					    ((it.q1 - it.q0 + BQ - 1) / BQ);
					split(cb);
				}
			}
		}
		return;
	}
	std::vector<uint32_t> lists[NXCD];
	xcd_lists(pl, BP, BQ, lists);
	for (auto &l : lists) {
		uint32_t tail = 0;
		for (size_t n = l.size(); n-- > 0 && tail < want;) {
			gemm_cblock &cb = pl->cblocks[l[n]];
			tail += ((cb.p + BP - 1) / BP) * ((cb.q + BQ - 1) / BQ);
			split(cb);
		}
	}
}

/*
 * Work tiles of one shape in launch order: every XCD list walks its C
 * blocks with the P-tiles innermost (tiles running together share the Y
 * panel columns and the X panel), and slot s of list x becomes workgroup
 * 8 * s + x.  Shorter lists are padded with empty tiles.  A split C
 * block's tiles come as adjacent halves, each pair with a slot of partial
 * sums; `fix` gets the first halves.
 */
static void
order_tiles(const contract_plan *pl, uint32_t BP, uint32_t BQ,
    std::vector<gemm_tile> &tiles, std::vector<gemm_tile> &fix)
{
	std::vector<uint32_t> lists[NXCD];
	xcd_lists(pl, BP, BQ, lists);
	std::vector<gemm_tile> bins[NXCD];
	fix.clear();
	for (int x = 0; x < NXCD; x++)
		for (uint32_t i : lists[x]) {
			const gemm_cblock &cb = pl->cblocks[i];
			for (uint32_t q0 = 0; q0 < cb.q; q0 += BQ)
				for (uint32_t p0 = 0; p0 < cb.p; p0 += BP) {
					if (cb.split_pair == KSPLIT_NONE) {
						bins[x].push_back({i, p0, q0, 0});
						continue;
					}
					const uint32_t k = 1 + 2 * (uint32_t)fix.size();
					bins[x].push_back({i, p0, q0, k});
					bins[x].push_back({i, p0, q0, k + 1});
					fix.push_back({i, p0, q0, k});
				}
		}
	size_t per = 0;
	for (auto &b : bins)
		per = std::max(per, b.size());
	tiles.assign(per * NXCD, gemm_tile{0xffffffffu, 0, 0, 0});
	for (int x = 0; x < NXCD; x++)
		for (size_t s = 0; s < bins[x].size(); s++)
			tiles[s * NXCD + x] = bins[x][s];
}

/* Every X and Y panel P-contiguous with an even leading dimension that
 * holds the C block's rows (columns) rounded up to even, base 16-byte
 * aligned: what k_gemm_f64_dma's 16-byte row-pair DMAs need (a pair past
 * an odd row count reads the padding row, whose products reach only
 * outputs past the C block).  XM_GEMM_DMA=0 turns the DMA kernel off. */
static bool
f64_dma_eligible(const contract_plan *pl)
{
	const char *e = getenv("XM_GEMM_DMA");
	if ((e && *e == '0') || pl->type != XM_SCALAR_DOUBLE ||
	    pl->precision != XM_PRECISION_NATIVE || use_simple())
		return false;
	for (const gemm_cblock &cb : pl->cblocks) {
		if (cb.pair_begin == cb.pair_end)
			continue;
		for (uint32_t i = cb.pair_begin; i < cb.pair_end; i++) {
			const gemm_pair &g = pl->pairs[i];
			if (g.xmode != PANEL_PCONTIG || g.ymode != PANEL_PCONTIG ||
			    g.ldx % 2 || g.ldy % 2 || g.x % 16 || g.y % 16 ||
			    g.ldx < cb.p + cb.p % 2 || g.ldy < cb.q + cb.q % 2)
				return false;
		}
	}
	return true;
}

/* Plans whose panels have 16-byte aligned bases and a leading dimension
 * divisible by the elements of 16 bytes, E (4 FP32, 2 complex64, 1
 * complex128), that holds the C block's rows (columns) rounded up to E:
 * what k_gemm_mfma_dw's 16-byte DMAs need.  XM_GEMM_DMA=0 keeps the dword
 * form. */
static bool
mfma_x4_eligible(const contract_plan *pl)
{
	const char *e = getenv("XM_GEMM_DMA");
	if (!pl->mfma_dw || (e && *e == '0'))
		return false;
	const uint32_t E = 16 / (uint32_t)elem_size(pl->type);
	for (const gemm_cblock &cb : pl->cblocks) {
		if (cb.pair_begin == cb.pair_end)
			continue;
		for (uint32_t i = cb.pair_begin; i < cb.pair_end; i++) {
			const gemm_pair &g = pl->pairs[i];
			if (g.ldx % E || g.ldy % E || g.x % 16 || g.y % 16 ||
			    g.ldx < (cb.p + E - 1) / E * E ||
			    g.ldy < (cb.q + E - 1) / E * E)
				return false;
		}
	}
	return true;
}

/* Every X and Y panel P-contiguous in an FP32, complex64 or complex128
 * plan: what
 * k_gemm_mfma_dw reads.  XM_GEMM_SIMPLE=1 keeps the FMA kernel. */
static bool
mfma_dw_eligible(const contract_plan *pl)
{
	if ((pl->type != XM_SCALAR_FLOAT && pl->type != XM_SCALAR_FLOAT_COMPLEX &&
	    pl->type != XM_SCALAR_DOUBLE_COMPLEX) || use_simple())
		return false;
	for (const gemm_pair &g : pl->pairs)
		if (g.xmode != PANEL_PCONTIG || g.ymode != PANEL_PCONTIG)
			return false;
	return true;
}

/*
 * The int8 form's tiles hold one C block each unless XM_CRT_STACK=1
 * merges a group's C blocks into one stack.  On the o=64 v=384 ladder the
 * merged stacks stage 8% fewer fragments but ran 9.75 ms against 9.50
 * (a group's C blocks in their own tiles, run together); the FP64 form,
 * whose 128-row tiles pad the ladder's C blocks far more, gains from it.
 */
static bool
stacked_crt(void)
{
	const char *e = getenv("XM_CRT_STACK");
	return e && *e == '1';
}

/*
 * The int8 form's tiles each run a part of the moduli (their residue
 * planes are independent; the reconstruction reads them all): tile t of
 * a launch computes moduli [j n / parts, (j + 1) n / parts) for part j in
 * head.ksplit (parts << 8 | j + 1; 0: all), an item's tiles listed part
 * by part, so the tiles running together read the same moduli's planes.
 * Twice the tiles of half the length let the XCD lists drain more evenly
 * (each plane is still staged once a tile): on the o=64 v=384 ladder 2
 * parts ran the GEMM 2.2% faster than 1 (the parts of a tile adjacent
 * instead: 1.6%, with twice the HBM reads), 3 parts as fast as 1, 4 and 6
 * slower.  XM_CRT_MSPLIT sets the parts (default 2, 1 to 8).
 */
static uint32_t
crt_moduli_parts(void)
{
	const char *e = getenv("XM_CRT_MSPLIT");
	const int v = e && *e ? atoi(e) : 2;
	return (uint32_t)std::min(8, std::max(1, v));
}

/* The BF16 split form's 128 x 128 tiles stack a group's C blocks like the
 * FP64 form (XM_SPLIT_STACK=0: one C block a stack). */
static bool
stacked_split(void)
{
	const char *e = getenv("XM_SPLIT_STACK");
	return !(e && *e == '0');
}

void
build_tiles(contract_plan *pl)
{
	/* the split kernel is the plan's for good: its tile records depend
	 * on it */
	pl->split_staged = pl->precision == XM_PRECISION_SPLIT_BF16 &&
	    split_staged_kernel(pl->nslices);
	pl->f64_dma = f64_dma_eligible(pl);
	pl->mfma_dw = mfma_dw_eligible(pl);
	pl->mfma_x4 = mfma_x4_eligible(pl);
	std::vector<std::pair<uint32_t, uint32_t>> cand;
	tile_candidates(pl, cand);
	pl->ncand = (int)std::min<size_t>(cand.size(), MAX_TILE_CANDIDATES);
	/* the int8 kernel reads stacked tile records only (one C block a
	 * stack with XM_CRT_STACK=0) */
	for (int c = 0; c < pl->ncand; c++)
		pl->cand_stacked[c] = pl->type == XM_SCALAR_DOUBLE &&
		    ((pl->precision == XM_PRECISION_NATIVE &&
		    cand[c].first == 128) ||
		    pl->precision == XM_PRECISION_SPLIT_INT8 ||
		    (pl->precision == XM_PRECISION_SPLIT_BF16 &&
		    cand[c].first == 128));
	mark_ksplit(pl);
	for (int c = 0; c < pl->ncand; c++) {
		pl->cand_bp[c] = cand[c].first;
		pl->cand_bq[c] = cand[c].second;
		if (pl->cand_stacked[c]) {
			if (pl->precision == XM_PRECISION_SPLIT_INT8)
				order_tiles_stacked<8>(pl, cand[c].second,
				    stacked_crt(), pl->cand_tiles[c],
				    pl->cand_fix[c]);
			else
				order_tiles_stacked<4>(pl, cand[c].second,
				    pl->precision == XM_PRECISION_NATIVE ||
				    stacked_split(), pl->cand_tiles[c],
				    pl->cand_fix[c]);
		} else if (grouped_order(pl))
			order_tiles_grouped(pl, cand[c].first, cand[c].second,
			    pl->cand_tiles[c], pl->cand_fix[c]);
		else
			order_tiles(pl, cand[c].first, cand[c].second,
			    pl->cand_tiles[c], pl->cand_fix[c]);
		pl->cand_ms[c] = -1.0f;
	}
	pl->tile_sig = tile_signature(pl);
	pl->tile_next = 0;
	pl->tile_choice = pl->ncand == 1 ? 0 : -1;
	auto it = g_tile_choice.find(pl->tile_sig);
	if (pl->tile_choice < 0 && it != g_tile_choice.end())
		pl->tile_choice = it->second;
}

/* Make the candidate to launch next the plan's active tile list (the
 * decided one, or candidate 0 for the plan's first, untimed call). */
void
select_tiles(contract_plan *pl)
{
	int c = pl->tile_choice >= 0 ? pl->tile_choice : 0;
	pl->tile_active = c;
	pl->stacked = pl->cand_stacked[c];
	pl->tiles = pl->cand_tiles[c];
	pl->d_tiles = pl->d_cand_tiles[c];
	pl->d_fix = pl->d_cand_fix[c];
	pl->nfix = pl->cand_fix[c].size() / (pl->cand_stacked[c] ?
	    sizeof(gemm_stile) / sizeof(gemm_tile) : 1);
	pl->tile_cfg = (int)(pl->cand_bp[c] * 1000 + pl->cand_bq[c]);
}

/* Time of the last launch (candidate tile_active); decide once every
 * candidate has one.  The plan's first call runs candidate 0 untimed: it
 * pays for cold caches and first launches, which would bias the choice.
 * tile_next counts calls: call 0 warms, call c + 1 times candidate c. */
void
gemm_tile_record(contract_plan *pl, float gemm_ms)
{
	if (pl->tile_choice >= 0)
		return;
	if (pl->tile_next > 0)
		pl->cand_ms[pl->tile_active] = gemm_ms;
	if (++pl->tile_next <= pl->ncand) {
		const int c = pl->tile_next - 1;
		pl->tile_active = c;
		pl->stacked = pl->cand_stacked[c];
		pl->tiles = pl->cand_tiles[c];
		pl->d_tiles = pl->d_cand_tiles[c];
		pl->d_fix = pl->d_cand_fix[c];
		pl->nfix = pl->cand_fix[c].size() / (pl->cand_stacked[c] ?
		    sizeof(gemm_stile) / sizeof(gemm_tile) : 1);
		pl->tile_cfg = (int)(pl->cand_bp[c] * 1000 + pl->cand_bq[c]);
		return;
	}
	int best = 0;
	for (int c = 1; c < pl->ncand; c++)
		if (pl->cand_ms[c] < pl->cand_ms[best])
			best = c;
	pl->tile_choice = best;
	g_tile_choice[pl->tile_sig] = best;
	select_tiles(pl);
}

/* The kernel launch_gemm runs for the plan's active tile shape (the same
 * choices as below). */
/*
 * The persistent form of the FP64 DMA kernel (XM_GEMM_PERSIST=1; plans
 * without queue counters, the no-allocation entry point's, never).  On the
 * o=64 v=384 ladder it ran 17.9-18.0 ms against 17.2 for one workgroup per
 * tile in launch order (same tiles, same occupancy of 3 workgroups a CU,
 * whether a workgroup starts at its XCC_ID's list or at blockIdx % 8), so
 * launch order stays the default.  Results are identical.
 */
static bool
persistent(const contract_plan *pl)
{
	const char *e = getenv("XM_GEMM_PERSIST");
	return e && *e == '1' && pl->d_queue != nullptr;
}

/* Workgroups of `kernel` (256 threads, no dynamic LDS) resident on the
 * whole device at once. */
static unsigned
resident_wgs(const void *kernel)
{
	int per_cu = 0, ncu = 0, dev = 0;
	HIP_OK("xm_contract", hipGetDevice(&dev));
	HIP_OK("xm_contract", hipOccupancyMaxActiveBlocksPerMultiprocessor(
	    &per_cu, kernel, GEMM_WG, 0));
	HIP_OK("xm_contract", hipDeviceGetAttribute(&ncu,
	    hipDeviceAttributeMultiprocessorCount, dev));
	return (unsigned)std::max(1, per_cu * ncu);
}

static const char *
gemm_kernel_name(const contract_plan *pl)
{
	const bool q128 = pl->cand_bq[pl->tile_active] == 128;
	if (use_simple())
		return "k_gemm_simple";
	switch (pl->type) {
	case XM_SCALAR_DOUBLE:
		if (pl->stacked)
			return "k_gemm_f64_dma_s";
		if (q128 && pl->f64_dma)
			return persistent(pl) ? "k_gemm_f64_dma_q" :
			    "k_gemm_f64_dma";
		return "k_gemm_f64";
	default:
		if (pl->mfma_x4)
			return "k_gemm_mfma_dw<x4>";
		return pl->mfma_dw ? "k_gemm_mfma_dw" : "k_gemm_simple";
	}
}

void
launch_gemm(contract_plan *pl, hipStream_t s)
{
	size_t nt = pl->tiles.size();
	pl->gemm_launches = 0;
	if (nt == 0)
		return;
	dim3 g((unsigned)nt), b(GEMM_WG);
	switch (pl->type) {
	case XM_SCALAR_DOUBLE: {
		const bool q128 = pl->cand_bq[pl->tile_active] == 128;
		if (use_simple())
			k_gemm_simple<double><<<g, b, 0, s>>>(pl->d_tiles,
			    pl->d_cblocks, pl->d_pairs, pl->d_scalars);
		else if (pl->stacked) {
			k_gemm_f64_dma_s<16, 2, 2><<<dim3((unsigned)(nt / 3)),
			    dim3(FS_WG), 0, s>>>((const gemm_stile *)pl->d_tiles,
			    pl->d_cblocks, pl->d_pairs, pl->d_scalars, pl->d_ksw);
			if (pl->nfix)
				k_ksplit_sum_s<<<dim3((unsigned)pl->nfix), b, 0,
				    s>>>((const gemm_stile *)pl->d_fix,
				    pl->d_cblocks, pl->d_ksw, pl->d_scalars);
		}
		else if (q128 && pl->f64_dma && persistent(pl)) {
			static const unsigned nwg = resident_wgs(
			    (const void *)k_gemm_f64_dma_q<16, 2, 3>);
			k_queue_reset<<<1, 64, 0, s>>>(pl->d_queue);
			k_gemm_f64_dma_q<16, 2, 3><<<dim3(std::min<unsigned>(nwg,
			    (unsigned)nt)), b, 0, s>>>(pl->d_tiles,
			    (uint32_t)(nt / NXCD), pl->d_queue, pl->d_cblocks,
			    pl->d_pairs, pl->d_scalars, pl->d_ksw);
		} else if (q128 && pl->f64_dma)
			k_gemm_f64_dma<16, 2, 3><<<g, b, 0, s>>>(pl->d_tiles,
			    pl->d_cblocks, pl->d_pairs, pl->d_scalars, pl->d_ksw);
		else if (q128)
			k_gemm_f64<64, 128, 16><<<g, b, 0, s>>>(pl->d_tiles,
			    pl->d_cblocks, pl->d_pairs, pl->d_scalars, pl->d_ksw);
		else
			k_gemm_f64<64, 64, 16><<<g, b, 0, s>>>(pl->d_tiles,
			    pl->d_cblocks, pl->d_pairs, pl->d_scalars, pl->d_ksw);
		if (pl->nfix && pl->stacked)
			;	/* summed above */
		else if (pl->nfix && q128)
			k_ksplit_sum<64, 128><<<dim3((unsigned)pl->nfix), b, 0,
			    s>>>(pl->d_fix, pl->d_cblocks, pl->d_ksw,
			    pl->d_scalars);
		else if (pl->nfix)
			k_ksplit_sum<64, 64><<<dim3((unsigned)pl->nfix), b, 0,
			    s>>>(pl->d_fix, pl->d_cblocks, pl->d_ksw,
			    pl->d_scalars);
		break;
	}
	case XM_SCALAR_FLOAT:
		if (pl->mfma_x4 && pl->cand_bq[pl->tile_active] == 128)
			k_gemm_mfma_dw<float, 128, 16, 4><<<g, b, 0, s>>>(
			    pl->d_tiles, pl->d_cblocks, pl->d_pairs,
			    pl->d_scalars);
		else if (pl->mfma_x4)
			k_gemm_mfma_dw<float, 64, 16, 4><<<g, b, 0, s>>>(
			    pl->d_tiles, pl->d_cblocks, pl->d_pairs,
			    pl->d_scalars);
		else if (pl->mfma_dw && pl->cand_bq[pl->tile_active] == 128)
			k_gemm_mfma_dw<float, 128, 16><<<g, b, 0, s>>>(pl->d_tiles,
			    pl->d_cblocks, pl->d_pairs, pl->d_scalars);
		else if (pl->mfma_dw)
			k_gemm_mfma_dw<float, 64, 16><<<g, b, 0, s>>>(pl->d_tiles,
			    pl->d_cblocks, pl->d_pairs, pl->d_scalars);
		else
			k_gemm_simple<float><<<g, b, 0, s>>>(pl->d_tiles,
			    pl->d_cblocks, pl->d_pairs, pl->d_scalars);
		break;
	case XM_SCALAR_FLOAT_COMPLEX:
		if (pl->mfma_x4)
			k_gemm_mfma_dw<cx<float>, 64, 16, 4><<<g, b, 0, s>>>(
			    pl->d_tiles, pl->d_cblocks, pl->d_pairs,
			    pl->d_scalars);
		else if (pl->mfma_dw)
			k_gemm_mfma_dw<cx<float>, 64, 16><<<g, b, 0, s>>>(pl->d_tiles,
			    pl->d_cblocks, pl->d_pairs, pl->d_scalars);
		else
			k_gemm_simple<cx<float>><<<g, b, 0, s>>>(pl->d_tiles,
			    pl->d_cblocks, pl->d_pairs, pl->d_scalars);
		break;
	case XM_SCALAR_DOUBLE_COMPLEX:
		if (pl->mfma_x4)
			k_gemm_mfma_dw<cx<double>, 64, 8, 4><<<g, b, 0, s>>>(
			    pl->d_tiles, pl->d_cblocks, pl->d_pairs,
			    pl->d_scalars);
		else if (pl->mfma_dw)
			k_gemm_mfma_dw<cx<double>, 64, 8><<<g, b, 0, s>>>(pl->d_tiles,
			    pl->d_cblocks, pl->d_pairs, pl->d_scalars);
		else
			k_gemm_simple<cx<double>><<<g, b, 0, s>>>(pl->d_tiles,
			    pl->d_cblocks, pl->d_pairs, pl->d_scalars);
		break;
	}
	hipError_t e = hipGetLastError();
	if (e != hipSuccess)
		xm_fatal("xm_contract: GEMM launch failed: %s",
		    hipGetErrorString(e));
	pl->gemm_launches = 1;
	pl->gemm_kernel = gemm_kernel_name(pl);
}
