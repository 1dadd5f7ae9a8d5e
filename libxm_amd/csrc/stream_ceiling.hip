/*
 * stream_ceiling.hip - how fast operand fragments can stream into LDS on
 * this part, measured with a model of the int8 CRT GEMM's inner step (no
 * contraction involved; the numbers behind the tile-shape choice of
 * k_gemm_crt in crt.hip).
 *
 * One workgroup of 8 waves per CU (256 workgroups).  A step of each wave:
 * NDMA wave-wide global_load_lds_dwordx4 (1 KiB each) into the stage
 * CR_NST - 1 steps ahead, NRD ds_read_b128 of the current stage and NMF
 * v_mfma_i32_32x32x32_i8, then a counted vmcnt wait and s_barrier: the
 * 256 x 128 tile's 128-k step is (6, 16, 16), the 256 x 256 tile's 64-k
 * step (4, 12, 16).  The DMA sources cycle through a region per workgroup:
 * 64 KiB (L2-resident: 2 MiB per XCD) or 32 MiB (streamed from HBM).
 * Prints one JSON line per configuration: int8 TOP/s and the DMA rate in
 * TB/s over all CUs.  Usage: stream_ceiling (no arguments).
 */
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(3))) void lds_t;
typedef __attribute__((address_space(1))) void gbl_t;

#define CHECK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
	fprintf(stderr, "stream_ceiling: %s: %s\n", #x, \
	    hipGetErrorString(e_)); exit(1); } } while (0)

template <int NDMA, int NRD, int NMF, int NST, int NACC, bool DMA>
__global__ void __launch_bounds__(512, 1)
k_model(int iters, const uint4 *src, size_t region, int *out)
{
	constexpr int SF = 8 * NDMA;		/* fragments per stage */
	__shared__ uint4 st[NST][SF * 64];
	const int lane = threadIdx.x & 63;
	const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	for (int i = threadIdx.x; i < NST * SF * 64; i += 512)
		(&st[0][0])[i] = make_uint4(i, i * 3, i * 5, i * 7);
	__syncthreads();
	v16i acc[NACC];
	for (int a = 0; a < NACC; a++)
		acc[a] = v16i{};
	const size_t nfr = region / 1024;
	const uint4 *base = src + (size_t)blockIdx.x * (region / 16);
	size_t fr = (size_t)w * NDMA;
	int cs = 0;
	for (int it = 0; it < iters; it++) {
		if (DMA) {
			const int ns = cs == 0 ? NST - 1 : cs - 1;
#pragma unroll
			for (int d = 0; d < NDMA; d++) {
				const size_t f = (fr + d) % nfr;
				__builtin_amdgcn_global_load_lds(
				    (gbl_t *)(base + f * 64 + lane),
				    (lds_t *)&st[ns][(w * NDMA + d) * 64], 16, 0, 0);
			}
			fr += 8 * NDMA;
		}
		const uint4 *cur = st[cs];
		v4i g[NRD];
#pragma unroll
		for (int j = 0; j < NRD; j++) {
			const uint4 u = cur[((j * 5 + w) % SF) * 64 + lane];
			g[j] = v4i{(int)u.x, (int)u.y, (int)u.z, (int)u.w};
		}
#pragma unroll
		for (int j = 0; j < NMF; j++)
			acc[j % NACC] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
			    g[j % NRD], g[(j * 3 + 1) % NRD], acc[j % NACC], 0, 0, 0);
		__builtin_amdgcn_s_waitcnt(((NST - 2) * NDMA) | (0x7 << 4) |
		    (0xf << 8));
		__atomic_signal_fence(__ATOMIC_SEQ_CST);
		__builtin_amdgcn_s_barrier();
		__atomic_signal_fence(__ATOMIC_SEQ_CST);
		cs = cs + 1 < NST ? cs + 1 : 0;
	}
	__builtin_amdgcn_s_waitcnt(0);
	int s = 0;
	for (int a = 0; a < NACC; a++)
		for (int i = 0; i < 16; i++)
			s += acc[a][i];
	if (s == 0x12345678)
		out[threadIdx.x] = s;
}

template <int NDMA, int NRD, int NMF, int NST, int NACC, bool DMA>
static void
run(const char *tile, const char *source, const uint4 *src, size_t region,
    int *out)
{
	const long mfmas = 256L * 8 * 16 * 4000;
	const int iters = (int)(mfmas / (256L * 8 * NMF));
	hipEvent_t a, b;
	CHECK(hipEventCreate(&a));
	CHECK(hipEventCreate(&b));
	k_model<NDMA, NRD, NMF, NST, NACC, DMA><<<256, 512>>>(iters, src,
	    region, out);
	CHECK(hipGetLastError());
	CHECK(hipEventRecord(a));
	k_model<NDMA, NRD, NMF, NST, NACC, DMA><<<256, 512>>>(iters, src,
	    region, out);
	CHECK(hipEventRecord(b));
	CHECK(hipEventSynchronize(b));
	float ms;
	CHECK(hipEventElapsedTime(&ms, a, b));
	const double nw = 256.0 * 8 * iters;
	printf("{\"tile\": \"%s\", \"source\": \"%s\", \"ms\": %.3f, "
	    "\"int8_tops\": %.1f, \"dma_tbs\": %.2f}\n", tile, source, ms,
	    nw * NMF * 65536 / ms / 1e9,
	    DMA ? nw * NDMA * 1024 / ms / 1e9 : 0.0);
	fflush(stdout);
	CHECK(hipEventDestroy(a));
	CHECK(hipEventDestroy(b));
}

int
main()
{
	const size_t big = (size_t)8 << 30;	/* 256 x 32 MiB */
	uint4 *src;
	int *out;
	CHECK(hipMalloc(&src, big));
	CHECK(hipMalloc(&out, 4096));
	CHECK(hipMemset(src, 1, big));
	const size_t l2 = 64 << 10, hbm = (size_t)32 << 20;
	run<6, 16, 16, 3, 4, false>("256x128 k128", "none", src, l2, out);
	run<6, 16, 16, 3, 4, true>("256x128 k128", "L2", src, l2, out);
	run<6, 16, 16, 3, 4, true>("256x128 k128", "HBM", src, hbm, out);
	run<4, 12, 16, 4, 8, false>("256x256 k64", "none", src, l2, out);
	run<4, 12, 16, 4, 8, true>("256x256 k64", "L2", src, l2, out);
	run<4, 12, 16, 4, 8, true>("256x256 k64", "HBM", src, hbm, out);
	CHECK(hipFree(src));
	CHECK(hipFree(out));
	return 0;
}
