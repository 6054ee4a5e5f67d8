// Counter-based random numbers on the device.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11) keyed by the 64-bit seed, with the counter
// (site, direction, draw | ctr2 << 5, stream tag).  Normals come from a
// Box-Muller transform whose log and sin/cos are evaluated with +, *, fma,
// division and sqrt only (argument reduction plus series), so the same
// counter gives the same bits on the host and on the GPU.
#pragma once
#include "su3_dev.h"

namespace mesh_dev {

#define MD __device__ __forceinline__
#define HD __host__ __device__ __forceinline__

enum : uint32_t
{
	TAG_HAAR = 0x48414152u,
	TAG_MOM = 0x4d4f4d30u,
	TAG_METRO = 0x4d455452u,
	TAG_GAUGE = 0x47415547u,
};

struct u4
{
	uint32_t x, y, z, w;
};

HD u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
	for (int r = 0; r < 10; ++r)
	{
		uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
		uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
		u4 n;
		n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
		n.y = (uint32_t)p1;
		n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
		n.w = (uint32_t)p0;
		c = n;
		k0 += 0x9E3779B9u;
		k1 += 0xBB67AE85u;
	}
	return c;
}

HD double u01_from_bits(uint32_t hi, uint32_t lo)
{
	uint64_t w = ((((uint64_t)hi) << 32) | lo) >> 11;
	return (double)(w + 1) * (1.0 / 9007199254740992.0);
}

MD double det_log(double x)
{
	uint64_t bits = (uint64_t)__double_as_longlong(x);
	int e = (int)((bits >> 52) & 0x7ff) - 1023;
	bits = (bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
	double m = __longlong_as_double((long long)bits);
	if (m > 1.4142135623730951)
	{
		m *= 0.5;
		e += 1;
	}
	double s = (m - 1.0) / (m + 1.0);
	double s2 = s * s;
	double p = 1.0 / 23.0;
	p = fma(p, s2, 1.0 / 21.0);
	p = fma(p, s2, 1.0 / 19.0);
	p = fma(p, s2, 1.0 / 17.0);
	p = fma(p, s2, 1.0 / 15.0);
	p = fma(p, s2, 1.0 / 13.0);
	p = fma(p, s2, 1.0 / 11.0);
	p = fma(p, s2, 1.0 / 9.0);
	p = fma(p, s2, 1.0 / 7.0);
	p = fma(p, s2, 1.0 / 5.0);
	p = fma(p, s2, 1.0 / 3.0);
	p = fma(p, s2, 1.0);
	double lm = 2.0 * s * p;
	return fma((double)e, 0.6931471805599453094, lm);
}

MD void det_sincos2pi(double u, double *sn, double *cs)
{
	double q = floor(4.0 * u + 0.5);
	double f = u - q * 0.25;
	double th = f * 6.283185307179586477;
	double t2 = th * th;
	double ps = -1.0 / 121645100408832000.0;
	ps = fma(ps, t2, 1.0 / 355687428096000.0);
	ps = fma(ps, t2, -1.0 / 1307674368000.0);
	ps = fma(ps, t2, 1.0 / 6227020800.0);
	ps = fma(ps, t2, -1.0 / 39916800.0);
	ps = fma(ps, t2, 1.0 / 362880.0);
	ps = fma(ps, t2, -1.0 / 5040.0);
	ps = fma(ps, t2, 1.0 / 120.0);
	ps = fma(ps, t2, -1.0 / 6.0);
	ps = fma(ps, t2, 1.0);
	double S = th * ps;
	double pc = 1.0 / 2432902008176640000.0;
	pc = fma(pc, t2, -1.0 / 6402373705728000.0);
	pc = fma(pc, t2, 1.0 / 20922789888000.0);
	pc = fma(pc, t2, -1.0 / 87178291200.0);
	pc = fma(pc, t2, 1.0 / 479001600.0);
	pc = fma(pc, t2, -1.0 / 3628800.0);
	pc = fma(pc, t2, 1.0 / 40320.0);
	pc = fma(pc, t2, -1.0 / 720.0);
	pc = fma(pc, t2, 1.0 / 24.0);
	pc = fma(pc, t2, -0.5);
	double C = fma(pc, t2, 1.0);
	int qi = ((int)q) & 3;
	double s_ = qi == 0 ? S : (qi == 1 ? C : (qi == 2 ? -S : -C));
	double c_ = qi == 0 ? C : (qi == 1 ? -S : (qi == 2 ? -C : S));
	*sn = s_;
	*cs = c_;
}

MD void normal_pair(u4 r, double *z0, double *z1)
{
	double u1 = u01_from_bits(r.x, r.y);
	double u2 = u01_from_bits(r.z, r.w);
	double R = sqrt(-2.0 * det_log(u1));
	double sn, cs;
	det_sincos2pi(u2, &sn, &cs);
	*z0 = R * cs;
	*z1 = R * sn;
}

// n (even) normals for (site, mu) of a stream
template <int N>
MD void link_normals(uint64_t seed, uint32_t tag, uint32_t ctr2, uint32_t site, uint32_t mu, double *z)
{
	uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
	for (int d = 0; d < N / 2; ++d)
	{
		u4 c = {site, mu, (uint32_t)d | (ctr2 << 5), tag};
		u4 r = philox4x32_10(c, k0, k1);
		normal_pair(r, &z[2 * d], &z[2 * d + 1]);
	}
}

// Haar SU(3): Gram-Schmidt on the rows of a complex Gaussian matrix, row 0
// divided by the determinant (src/gauge/groups.h:255-261)
MD m3<double> haar_from_normals(double const *z)
{
	m3<double> a;
#pragma unroll
	for (int k = 0; k < 9; ++k)
		a.e[k] = {z[2 * k], z[2 * k + 1]};
#pragma unroll
	for (int i = 0; i < 3; ++i)
	{
		// projection done twice (re-orthogonalisation) for full accuracy
#pragma unroll
		for (int pass = 0; pass < 2; ++pass)
#pragma unroll
			for (int k = 0; k < i; ++k)
			{
				cpx<double> c = c0<double>();
#pragma unroll
				for (int j = 0; j < 3; ++j)
					cmadd_ca(c, a.e[3 * k + j], a.e[3 * i + j]);
#pragma unroll
				for (int j = 0; j < 3; ++j)
					cmsub(a.e[3 * i + j], c, a.e[3 * k + j]);
			}
		double n2 = 0.0;
#pragma unroll
		for (int j = 0; j < 3; ++j)
		{
			n2 = fma(a.e[3 * i + j].re, a.e[3 * i + j].re, n2);
			n2 = fma(a.e[3 * i + j].im, a.e[3 * i + j].im, n2);
		}
		double inv = 1.0 / sqrt(n2);
#pragma unroll
		for (int j = 0; j < 3; ++j)
		{
			a.e[3 * i + j].re *= inv;
			a.e[3 * i + j].im *= inv;
		}
	}
	cpx<double> d = det(a);
	double dn = fma(d.re, d.re, d.im * d.im);
	double id = 1.0 / dn;
#pragma unroll
	for (int j = 0; j < 3; ++j)
	{
		cpx<double> t = c0<double>();
		cmadd_cb(t, a.e[j], d);
		a.e[j] = {t.re * id, t.im * id};
	}
	return a;
}

// sum_a p_a i T^a with T^a = lambda^a / 2
MD m3<double> algebra_from_normals(double const *p)
{
	m3<double> m = zero_m3<double>();
	const double h = 0.5;
	const double r3 = 0.28867513459481288225;
	m.e[1] = {h * p[1], h * p[0]};
	m.e[3] = {-(h * p[1]), h * p[0]};
	m.e[2] = {h * p[4], h * p[3]};
	m.e[6] = {-(h * p[4]), h * p[3]};
	m.e[5] = {h * p[6], h * p[5]};
	m.e[7] = {-(h * p[6]), h * p[5]};
	m.e[0] = {0.0, fma(h, p[2], r3 * p[7])};
	m.e[4] = {0.0, fma(-h, p[2], r3 * p[7])};
	m.e[8] = {0.0, -2.0 * (r3 * p[7])};
	return m;
}

#undef MD
#undef HD

} // namespace mesh_dev
