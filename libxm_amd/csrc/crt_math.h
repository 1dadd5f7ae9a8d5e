/*
 * crt_math.h - integer arithmetic of the int8 CRT (Chinese remainder
 * theorem) form of FP64 contraction, shared by the GPU kernels (crt.hip)
 * and a CPU check program.  Plain C++ (host and device).
 *
 * The moduli m_0..m_12 are pairwise coprime and at most 256, so every
 * residue in its symmetric range [-floor(m/2), ceil(m/2) - 1] fits a signed
 * byte and a product of two residues is at most 2^14 in magnitude: int32
 * sums of up to 2^17 such products are exact.
 */
#ifndef XM_CRT_MATH_H
#define XM_CRT_MATH_H

#include <cmath>
#include <cstdint>

#ifdef __HIPCC__
#define CRT_HD __host__ __device__
#else
#define CRT_HD
#endif

#define CRT_MAXN 13

/* Products of operands below 2^23 in magnitude: the full-rate 24-bit
 * multiply on the device (32-bit ones run at a quarter of the rate). */
#ifdef __HIP_DEVICE_COMPILE__
#define CRT_MUL24(a, b) __mul24((a), (b))
#else
#define CRT_MUL24(a, b) ((a) * (b))
#endif

/* 256 = 2^8, 255 = 3*5*17, 253 = 11*23, 247 = 13*19, 217 = 7*31; the others
 * are primes */
static constexpr int CRT_MOD[CRT_MAXN] = {256, 255, 253, 251, 247, 241, 239,
    233, 229, 227, 223, 217, 211};

struct crt_tables {
	int lo[CRT_MAXN], hi[CRT_MAXN];	/* symmetric residue range */
	uint32_t magic[CRT_MAXN];	/* ceil(2^32 / m) */
	float finv[CRT_MAXN];		/* 1 / m, rounded to FP32 */
	int c31[CRT_MAXN];		/* 2^31 mod m, in [0, m) */
	int p17[CRT_MAXN], p34[CRT_MAXN];	/* 2^17, 2^34 mod m (symmetric) */
	int p13[CRT_MAXN], p26[CRT_MAXN], p39[CRT_MAXN]; /* 2^13, 2^26, 2^39 */
	/* Garner: ginv[i] = (m_0 ... m_{i-1})^-1 mod m_i and
	 * pm[i][j] = m_0 ... m_{j-1} mod m_i (j < i), symmetric */
	int ginv[CRT_MAXN];
	int pm[CRT_MAXN][CRT_MAXN];
};

static constexpr int
crt_sym(long long v, int m)
{
	long long r = v % m;
	if (r < 0)
		r += m;
	/* [0, m) -> [-floor(m/2), ceil(m/2) - 1] */
	return (int)(r >= (m + 1) / 2 ? r - m : r);
}

static constexpr long long
crt_pow2_mod(int e, int m)
{
	long long r = 1;
	for (int k = 0; k < e; k++)
		r = (r * 2) % m;
	return r;
}

static constexpr int
crt_inverse(long long a, int m)
{
	a %= m;
	if (a < 0)
		a += m;
	for (int x = 1; x < m; x++)
		if ((a * x) % m == 1)
			return x;
	return 0;
}

static constexpr crt_tables
crt_make_tables(void)
{
	crt_tables t{};
	for (int i = 0; i < CRT_MAXN; i++) {
		const int m = CRT_MOD[i];
		t.lo[i] = -(m / 2);
		t.hi[i] = (m + 1) / 2 - 1;
		t.magic[i] = (uint32_t)((0x100000000ull + m - 1) / m);
		t.finv[i] = (float)(1.0 / m);
		t.c31[i] = (int)(0x80000000ull % m);
		t.p17[i] = crt_sym(crt_pow2_mod(17, m), m);
		t.p34[i] = crt_sym(crt_pow2_mod(34, m), m);
		t.p13[i] = crt_sym(crt_pow2_mod(13, m), m);
		t.p26[i] = crt_sym(crt_pow2_mod(26, m), m);
		t.p39[i] = crt_sym(crt_pow2_mod(39, m), m);
		long long prod = 1;	/* m_0 ... m_{j-1} mod m_i */
		for (int j = 0; j < CRT_MAXN; j++) {
			t.pm[i][j] = j < i ? crt_sym(prod, m) : 0;
			prod = (prod * (CRT_MOD[j] % m)) % m;
		}
		long long pi = 1;
		for (int j = 0; j < i; j++)
			pi = (pi * CRT_MOD[j]) % m;
		t.ginv[i] = i == 0 ? 1 : crt_sym(crt_inverse(pi, m), m);
	}
	return t;
}

/* log2 of m_0 ... m_{n-1} */
static inline double
crt_log2_modulus(int n)
{
	double s = 0;
	for (int i = 0; i < n; i++)
		s += std::log2((double)CRT_MOD[i]);
	return s;
}

/*
 * v mod m in the symmetric range, for any int32 v, in integer arithmetic:
 * u = v + 2^31 in [0, 2^32), q = floor(u magic / 2^32) with
 * magic = ceil(2^32 / m) is floor(u / m) or one more (u (magic - 2^32/m)
 * < 2^32), so u - q m lies in [-m, m); then 2^31 mod m is taken off
 * again.
 */
CRT_HD static inline int
crt_smod(int v, int m, uint32_t magic, int c31, int hi)
{
	const uint32_t u = (uint32_t)v + 0x80000000u;
	const uint32_t q = (uint32_t)(((uint64_t)u * magic) >> 32);
	int r = (int)(u - q * (uint32_t)m);
	r += r < 0 ? m : 0;
	r -= c31;
	r += r < 0 ? m : 0;
	return r > hi ? r - m : r;
}

/*
 * v mod m in the symmetric range for |v| < 2^25, at full rate: (float)v is
 * within 1 of v, so the FP32 quotient rint(v / m) is within 0.03 of the
 * true one and v - q m lies within 0.53 m of 0; one correction either way.
 * (The integer form above needs quarter-rate 32-bit multiplies.)
 */
CRT_HD static inline int
crt_smod25(int v, int m, float finv, int lo, int hi)
{
	const int q = (int)rintf((float)v * finv);
	int r = v - CRT_MUL24(q, m);
	r = r > hi ? r - m : r;
	return r < lo ? r + m : r;
}

/*
 * An integer-valued double x (|x| < 2^51) as three int32 limbs:
 * x = h * 2^34 + md * 2^17 + lo with |h| <= 2^17, |md|, |lo| <= 2^16.
 * Every step is exact in FP64.
 */
CRT_HD static inline void
crt_limbs(double x, int *h, int *md, int *lo)
{
	const double th = rint(x * 0x1p-34);
	const double r1 = x - th * 0x1p34;
	const double tm = rint(r1 * 0x1p-17);
	const double tl = r1 - tm * 0x1p17;
	*h = (int)th;
	*md = (int)tm;
	*lo = (int)tl;
}

/* Residue of limbs (h, md, lo) modulo m_i: |h p34 + md p17 + lo| < 2^25. */
CRT_HD static inline int
crt_residue(int h, int md, int lo, int p34, int p17, int m, float finv,
    int rlo, int rhi)
{
	return crt_smod25(CRT_MUL24(h, p34) + CRT_MUL24(md, p17) + lo, m, finv,
	    rlo, rhi);
}

/*
 * An integer-valued double x (|x| < 2^51) as four limbs of 13 bits:
 * x = l[3] 2^39 + l[2] 2^26 + l[1] 2^13 + l[0] with |l[i]| <= 2^12.  Every
 * step is exact in FP64.
 */
CRT_HD static inline void
crt_limbs13(double x, int *l)
{
	const double t3 = rint(x * 0x1p-39);
	const double r2 = x - t3 * 0x1p39;
	const double t2 = rint(r2 * 0x1p-26);
	const double r1 = r2 - t2 * 0x1p26;
	const double t1 = rint(r1 * 0x1p-13);
	l[3] = (int)t3;
	l[2] = (int)t2;
	l[1] = (int)t1;
	l[0] = (int)(r1 - t1 * 0x1p13);
}

/*
 * Residue modulo m_I of the integer with limbs l (crt_limbs13), in the
 * symmetric range of m_I, in FP32 arithmetic that is exact: the limbs are
 * given as floats (|l| <= 2^12), v = l0 + l1 c13 + l2 c26 + l3 c39 (c the
 * symmetric 2^13k mod m, |c| <= 128) is an integer below 2^12 (1 + 3 * 128)
 * < 2^21 at every step of its fused multiply-adds, and fl(v / m) is within
 * |v| / m 2^-23 < 0.0009 of v / m, less than v / m's distance to a
 * half-integer for odd m (at least 1 / (2m) > 0.0019): its rounding q is
 * the nearest integer to v / m, and v - q m (exact in one fused
 * multiply-add) lies in [-(m - 1) / 2, (m - 1) / 2] with no correction.
 * m_0 = 256 divides 2^13: the residue is the low byte of l0, sign extended
 * (l0i: l0 as an integer).  No integer multiply is left for the compiler to
 * widen.
 */
template <int I>
CRT_HD static inline int
crt_residue13(const float *l, int l0i, const crt_tables &t)
{
	if (I == 0)
		return (int)(int8_t)(uint8_t)(l0i & 0xff);
	const float m = (float)CRT_MOD[I];
	const float v = fmaf(l[3], (float)t.p39[I], fmaf(l[2], (float)t.p26[I],
	    fmaf(l[1], (float)t.p13[I], l[0])));
	const float q = rintf(v * t.finv[I]);
	return (int)fmaf(-q, m, v);
}

/*
 * The integer X with X = r_i (mod m_i) for i < N and X inside the balanced
 * range of m_0 ... m_{N-1} (which holds every |X| < m_0 ... m_{N-1} / 4), as
 * a double: Garner's algorithm with balanced mixed-radix digits d_i,
 * X = d_0 + m_0 (d_1 + m_1 (d_2 + ...)), evaluated from the top digit in
 * FP64 (relative error at most N 2^-53).  Residues r_i in the symmetric
 * range of m_i.  N is a template parameter so the loops unroll and the
 * table entries become immediates.
 */
template <int N>
CRT_HD static inline double
crt_reconstruct(const int *r, const crt_tables &t)
{
	int d[N];
#pragma unroll
	for (int i = 0; i < N; i++) {
		const int m = CRT_MOD[i];
		int s = 0;	/* (d_0 + m_0 d_1 + ...) mod m_i, unreduced */
#pragma unroll
		for (int j = 0; j < i; j++)
			s += CRT_MUL24(d[j], t.pm[i][j]);
		/* |s| < 12 * 128 * 128, |(r - s) ginv| < 2^16 */
		s = crt_smod25(s, m, t.finv[i], t.lo[i], t.hi[i]);
		d[i] = crt_smod25(CRT_MUL24(r[i] - s, t.ginv[i]), m, t.finv[i],
		    t.lo[i], t.hi[i]);
	}
	double x = d[N - 1];
#pragma unroll
	for (int i = N - 2; i >= 0; i--)
		x = fma(x, (double)CRT_MOD[i], (double)d[i]);
	return x;
}

#endif /* XM_CRT_MATH_H */
