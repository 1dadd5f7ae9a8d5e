// CPU check of the CRT arithmetic in libxm_amd/csrc/crt_math.h against
// exact 128-bit integer arithmetic: limb splits and residues of integers
// up to 2^50 (17-bit limbs) and 2^51 (13-bit limbs, every quotient of the
// range), and int32 residue-product sums (what the int8 matrix cores
// compute per modulus) reconstructed by Garner's algorithm.
#include "crt_math.h"

#include <cstdio>
#include <cstdlib>
#include <random>

static constexpr crt_tables T = crt_make_tables();

static int
exact_mod(__int128 v, int m)
{
	__int128 r = v % m;
	if (r < 0)
		r += m;
	return (int)(r >= (m + 1) / 2 ? r - m : r);
}

/* crt_residue13<I> on every v of its range, through limbs (l0 = v, the
 * rest zero, covers v exactly when |v| <= 4096; larger v through l1), and
 * on random integers of up to 51 bits split by crt_limbs13. */
template <int I>
static void
check13(long &bad, std::mt19937_64 &rng)
{
	const int m = CRT_MOD[I];
	const int V = 4096 * 385;
	for (int v = -V; v <= V; v++) {
		/* every v the limbs can give, as l0 = v (the function's FP32
		 * steps on v itself; the I = 0 path reads l0 as an integer) */
		const float lf[4] = {(float)v, 0.0f, 0.0f, 0.0f};
		const int r = crt_residue13<I>(lf, v, T);
		if (I != 0 && r != exact_mod(v, m))
			bad++;
	}
	for (int it = 0; it < 300000; it++) {
		long long x = (long long)(rng() >> 12) - (1ll << 51) + 1;
		if (it < 8)
			x = it < 4 ? (1ll << 51) - 1 - it : -(1ll << 51) + 1 + it;
		int l[4];
		crt_limbs13((double)x, l);
		if ((__int128)l[3] * (1ll << 39) + (__int128)l[2] * (1ll << 26) +
		    (__int128)l[1] * (1 << 13) + l[0] != x)
			bad++;
		for (int k = 0; k < 4; k++)
			if (l[k] < -4096 || l[k] > 4096)
				bad++;
		const float lf[4] = {(float)l[0], (float)l[1], (float)l[2],
		    (float)l[3]};
		const int r = crt_residue13<I>(lf, l[0], T);
		if (r != exact_mod(x, m))
			bad++;
	}
}

template <int I = 0>
static void
check13_all(long &bad, std::mt19937_64 &rng)
{
	if constexpr (I < CRT_MAXN) {
		check13<I>(bad, rng);
		check13_all<I + 1>(bad, rng);
	}
}

int
main(void)
{
	std::mt19937_64 rng(12345);
	long bad = 0;
	// tables
	for (int i = 0; i < CRT_MAXN; i++) {
		__int128 p = 1;
		for (int j = 0; j < i; j++)
			p *= CRT_MOD[j];
		if (i && exact_mod(p * T.ginv[i], CRT_MOD[i]) != 1)
			bad++, printf("ginv %d\n", i);
		for (int j = 0; j < i; j++) {
			__int128 q = 1;
			for (int k = 0; k < j; k++)
				q *= CRT_MOD[k];
			if (exact_mod(q, CRT_MOD[i]) != T.pm[i][j])
				bad++, printf("pm %d %d\n", i, j);
		}
	}
	// the modular reduction on the whole int32 range: both ends, a dense
	// window around 0 and random values
	for (int i = 0; i < CRT_MAXN; i++) {
		const int m = CRT_MOD[i];
		auto one = [&](int v) {
			int r = crt_smod(v, m, T.magic[i], T.c31[i], T.hi[i]);
			if (r != exact_mod(v, m))
				bad++;
		};
		for (long long v = -(1ll << 31); v < -(1ll << 31) + 100000; v++)
			one((int)v);
		for (long long v = (1ll << 31) - 100000; v < (1ll << 31); v++)
			one((int)v);
		for (int v = -(1 << 20); v <= (1 << 20); v++)
			one(v);
		for (int it = 0; it < 2000000; it++)
			one((int)(uint32_t)rng());
		/* the FP32 form on its range |v| < 2^25: every value below 2^21,
		 * both ends and random values */
		auto small = [&](int v) {
			int r = crt_smod25(v, m, T.finv[i], T.lo[i], T.hi[i]);
			if (r != exact_mod(v, m))
				bad++;
		};
		for (int v = -(1 << 21); v <= (1 << 21); v++)
			small(v);
		for (int v = (1 << 25) - 200000; v < (1 << 25); v++)
			small(v), small(-v);
		for (int it = 0; it < 2000000; it++)
			small((int)(rng() % (1u << 26)) - (1 << 25) + 1);
	}
	// residues of single integers, |x| < 2^50
	for (int it = 0; it < 200000; it++) {
		long long x = (long long)(rng() >> 13) - (1ll << 50);
		if (it < 8)
			x = it < 4 ? (1ll << 50) - it : -(1ll << 50) + it;
		int h, md, lo;
		crt_limbs((double)x, &h, &md, &lo);
		if ((__int128)h * (1ll << 34) + (__int128)md * (1 << 17) + lo != x)
			bad++;
		for (int i = 0; i < CRT_MAXN; i++) {
			int r = crt_residue(h, md, lo, T.p34[i], T.p17[i],
			    CRT_MOD[i], T.finv[i], T.lo[i], T.hi[i]);
			if (r != exact_mod(x, CRT_MOD[i]) || r < -128 || r > 127)
				bad++;
		}
	}
	// the 13-bit limb form: the FP32 quotient of every v the limbs can
	// give (|v| <= 2^12 * 385) for every modulus, and residues of single
	// integers up to 2^51, both ends included
	check13_all(bad, rng);
	// dot products of K integers of up to 2^49.5: int32 sums of residue
	// products per modulus, then Garner
	for (int n = 11; n <= CRT_MAXN; n++) {
		double maxrel = 0;
		for (int it = 0; it < 300; it++) {
			const int K = 1 + (int)(rng() % 4000);
			const double lim = std::ldexp(1.0, (int)(crt_log2_modulus(n) / 2) - 1) /
			    std::sqrt((double)K);
			__int128 exact = 0;
			int acc[CRT_MAXN] = {0};
			for (int k = 0; k < K; k++) {
				std::uniform_real_distribution<double> u(-lim, lim);
				long long a = (long long)std::rint(u(rng));
				long long b = (long long)std::rint(u(rng));
				exact += (__int128)a * b;
				for (int i = 0; i < n; i++)
					acc[i] += exact_mod(a, CRT_MOD[i]) *
					    exact_mod(b, CRT_MOD[i]);
			}
			int r[CRT_MAXN];
			for (int i = 0; i < n; i++)
				r[i] = crt_smod(acc[i], CRT_MOD[i], T.magic[i],
				    T.c31[i], T.hi[i]);
			double got = n == 11 ? crt_reconstruct<11>(r, T) :
			    n == 12 ? crt_reconstruct<12>(r, T) :
			    crt_reconstruct<13>(r, T);
			double want = (double)exact;
			double rel = want != 0 ? std::fabs(got - want) / std::fabs(want) : std::fabs(got);
			if (rel > maxrel)
				maxrel = rel;
			if (rel > 1e-15)
				bad++;
		}
		printf("n=%d log2M=%.1f max rel %.2e\n", n, crt_log2_modulus(n), maxrel);
	}
	printf(bad ? "FAIL %ld\n" : "OK\n", bad);
	return bad != 0;
}
