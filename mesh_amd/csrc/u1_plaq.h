// Plaquette-space pieces shared by the U(1) 2-d flows that propose
// plaquettes (pflow.hip, pcflow.hip): geometry, Philox angles, block sums,
// and the integration of V - 1 free plaquettes into links.
//
// On an L0 x L1 torus the links are fixed by the V - 1 plaquettes other than
// the last, P(L0-1, L1-1), two holonomies and a gauge transformation: in the
// maximal-tree gauge theta_1(x0, x1) = 0 for x1 < L1 - 1 and
// theta_0(x0, L1 - 1) = 0 for x0 < L0 - 1,
//   theta_0(x0, x1) = sum_{k = x1}^{L1 - 2} P(x0, k)   (+ h0 on row L0 - 1)
//   theta_1(0, L1 - 1) = h1,
//   theta_1(x0 + 1, L1 - 1) = theta_1(x0, L1 - 1) + P(x0, L1 - 1)
//                             - theta_0(x0, L1 - 1) + theta_0(x0, 0),
// followed by the gauge transformation theta_mu(x) += a(x) - a(x + mu).  The
// map is integer-linear and unimodular, so a density over the free
// plaquettes with uniform holonomies and gauge angles is that density times
// (2 pi)^-(V + 1) over the links.
#pragma once

#include "mesh_internal.h"
#include "rng_dev.h"

namespace u1_plaq {

struct PGeo
{
	int L0, L1, V;
};

__device__ __forceinline__ double wrapd(double a)
{
	const double pi = 3.141592653589793238, tpi = 6.283185307179586477;
	a = a - tpi * floor((a + pi) * (1.0 / tpi));
	return a >= pi ? a - tpi : (a < -pi ? a + tpi : a);
}

// site (x0, x1) for 0 <= x0 <= L0, 0 <= x1 <= L1 (one periodic step)
__device__ __forceinline__ int site(PGeo const &g, int x0, int x1)
{
	x0 = x0 == g.L0 ? 0 : x0;
	x1 = x1 == g.L1 ? 0 : x1;
	return x0 * g.L1 + x1;
}

// plaquette at (x0, x1) of an LDS configuration t[2V], unwrapped
__device__ __forceinline__ float plaqp(PGeo const &g, const float *t, int x0, int x1)
{
	return t[site(g, x0, x1)] + t[g.V + site(g, x0 + 1, x1)] - t[site(g, x0, x1 + 1)] - t[g.V + site(g, x0, x1)];
}

// uniform on [-pi, pi) from Philox(seed; a, b, step << 5, tag)
__device__ __forceinline__ float uniform_angle(uint64_t seed, uint32_t step, uint32_t a, uint32_t b, uint32_t tag)
{
	mesh_dev::u4 ctr = {a, b, step << 5, tag};
	mesh_dev::u4 r = mesh_dev::philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
	double u = mesh_dev::u01_from_bits(r.x, r.y) - 0x1p-53;
	return (float)(6.283185307179586477 * u - 3.141592653589793238);
}

// sums over the 256 threads of a workgroup, in a fixed order
template <int NQ> __device__ __forceinline__ void block_sum(double (&v)[NQ], double (*red)[4])
{
	int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	for (int q = 0; q < NQ; ++q)
	{
		double s = v[q];
		for (int o = 32; o >= 1; o >>= 1)
			s += __shfl_xor(s, o, 64);
		if (lane == 0)
			red[q][w] = s;
	}
	__syncthreads();
	for (int q = 0; q < NQ; ++q)
		v[q] = ((red[q][0] + red[q][1]) + red[q][2]) + red[q][3];
	__syncthreads();
}

__device__ __forceinline__ float wrapf_p(float a)
{
	const float PI = 3.14159265358979323846f, TPI = 6.28318530717958647692f;
	a = a - TPI * floorf((a + PI) * (1.0f / TPI));
	return a >= PI ? a - TPI : (a < -PI ? a + TPI : a);
}

// links out[2V] of the free plaquettes P[V] (the last unused), holonomies
// h0, h1 and gauge angles al[V]; t[2V] is LDS scratch.  Called by every
// thread of the workgroup; P and al must be visible to all of them.
__device__ void links_from_plaquettes(PGeo const &g, const float *P, float h0, float h1, const float *al, float *t,
                                      float *__restrict__ out, int nthreads)
{
	// direction 0: suffix sums along each row (one thread per row)
	for (int x0 = threadIdx.x; x0 < g.L0; x0 += nthreads)
	{
		double acc = x0 == g.L0 - 1 ? (double)h0 : 0.0;
		t[x0 * g.L1 + g.L1 - 1] = (float)wrapd(acc);
		for (int x1 = g.L1 - 2; x1 >= 0; --x1)
		{
			acc += (double)P[x0 * g.L1 + x1];
			t[x0 * g.L1 + x1] = (float)wrapd(acc);
		}
		for (int x1 = 0; x1 < g.L1 - 1; ++x1)
			t[g.V + x0 * g.L1 + x1] = 0.f;
	}
	__syncthreads();
	// direction 1, last column: prefix sum over rows (one thread)
	if (threadIdx.x == 0)
	{
		double acc = (double)h1;
		t[g.V + g.L1 - 1] = (float)wrapd(acc);
		for (int x0 = 0; x0 + 1 < g.L0; ++x0)
		{
			int r = x0 * g.L1;
			acc += (double)P[r + g.L1 - 1] - (double)t[r + g.L1 - 1] + (double)t[r];
			t[g.V + r + g.L1 + g.L1 - 1] = (float)wrapd(acc);
		}
	}
	__syncthreads();
	// gauge transformation into the output
	for (int x = threadIdx.x; x < g.V; x += nthreads)
	{
		int x0 = x / g.L1, x1 = x - x0 * g.L1;
		out[x] = wrapf_p(t[x] + al[x] - al[site(g, x0 + 1, x1)]);
		out[g.V + x] = wrapf_p(t[g.V + x] + al[x] - al[site(g, x0, x1 + 1)]);
	}
}

} // namespace u1_plaq
