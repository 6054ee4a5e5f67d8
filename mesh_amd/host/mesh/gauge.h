// C++ host layer: the gauge call surface of krox/mesh (src/gauge/utils.h,
// hmc.h, hmc.cpp, landau.h) executed by libmesh_hip.so on the GPU.
// Same names and argument meaning as the reference; each function cites the
// reference function it stands for.
#pragma once

#include "mesh/lattice.h"

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <numeric>
#include <string_view>

namespace mesh {

inline Context &ctx() { return Context::global(); }

// ---- observables ------------------------------------------------------------

// normalised to [0,1] (src/gauge/utils.h:93-108)
template <class G> double plaquette(GaugeField<G> const &U)
{
	double r;
	ctx().check(mesh_su3_plaquette(ctx().get(), U.grid().c(), G::prec, U.data(), &r));
	return r;
}

// (1 - plaquette) * beta * Nd(Nd-1)/2 * V (src/gauge/utils.h:141-146)
template <class G> double wilsonAction(GaugeField<G> const &U, double beta)
{
	double r;
	ctx().check(mesh_su3_wilson_action(ctx().get(), U.grid().c(), G::prec, U.data(), beta, &r));
	return r;
}

// sum ||U||^2 (src/gauge/utils.h:28, 35-41)
template <class G> double norm2(Field<G> const &P)
{
	double r;
	ctx().check(mesh_su3_norm2(ctx().get(), P.grid().c(), P.nfields(), G::prec, P.data(), &r));
	return r;
}

// sum_real_trace(A * B) (src/gauge/utils.h:29)
template <class G> double sum_real_trace_prod(Field<G> const &A, Field<G> const &B)
{
	double r;
	ctx().check(mesh_su3_sum_real_trace_prod(ctx().get(), A.grid().c(), A.nfields(), G::prec, A.data(), B.data(),
	                                         &r));
	return r;
}

// staple sum of direction mu (src/gauge/utils.h:110-139)
template <class G> Lattice<G> stapleSum(GaugeField<G> const &U, int mu)
{
	Lattice<G> S(U.grid());
	ctx().check(mesh_su3_staple(ctx().get(), U.grid().c(), G::prec, U.data(), mu, S.data()));
	return S;
}

// (beta / 2 Nc) TA(U_mu stapleSum(U, mu)) (src/gauge/utils.h:148-153)
template <class G> Lattice<G> wilsonDeriv(GaugeField<G> const &U, int mu, double beta)
{
	Lattice<G> F(U.grid());
	ctx().check(mesh_su3_force(ctx().get(), U.grid().c(), G::prec, U.data(), mu, beta, F.data()));
	return F;
}

// circular shift: b[x] = a[x + offset e_dir] (src/lattice/lattice.h:366-389)
template <class G> Lattice<G> cshift(Lattice<G> const &a, int dir, int offset)
{
	Lattice<G> r(a.grid());
	ctx().check(mesh_su3_cshift(ctx().get(), a.grid().c(), 1, G::prec, r.data(), a.data(), dir, offset));
	return r;
}

// sum of every matrix of the field, in FP64 (lattice_sum, src/lattice/lattice.h:270-273)
template <class G> std::array<std::complex<double>, 9> lattice_sum(Field<G> const &a)
{
	double o[18];
	ctx().check(mesh_su3_sum(ctx().get(), a.grid().c(), a.nfields(), G::prec, a.data(), o));
	std::array<std::complex<double>, 9> r;
	for (int k = 0; k < 9; ++k)
		r[k] = {o[2 * k], o[2 * k + 1]};
	return r;
}

// lhs += a * rhs, the product rounded before the sum (add_mul,
// src/lattice/lattice.h:358-363; `lhs += rhs * a`, lattice.h:302-309, 349-355)
template <class G> void add_mul(Field<G> &lhs, double a, Field<G> const &rhs)
{
	if (lhs.nfields() != rhs.nfields())
		throw std::runtime_error("add_mul: fields differ in the number of matrices per site");
	ctx().check(mesh_su3_axpy(ctx().get(), lhs.grid().c(), lhs.nfields(), G::prec, lhs.data(), a, rhs.data()));
}

// project_on_group_fast on every link (src/gauge/utils.h:80-91)
template <class G> void reunitize(Field<G> &U)
{
	ctx().check(mesh_su3_reunit(ctx().get(), U.grid().c(), U.nfields(), G::prec, U.data()));
}

// ---- random fields (Philox-keyed; the reference draws from a host xoshiro256)

template <class G> void random_gauge_field(Field<G> &U, uint64_t seed, uint32_t stream = 0)
{
	ctx().check(mesh_su3_random_gauge(ctx().get(), U.grid().c(), U.nfields(), G::prec, U.data(), seed, stream));
}

template <class G> void random_algebra_field(GaugeField<G> &P, uint64_t seed, uint32_t traj)
{
	ctx().check(mesh_su3_momenta(ctx().get(), P.grid().c(), G::prec, P.data(), nullptr, seed, traj));
}

// Philox key of a command-line seed: a decimal number is the key itself, any
// other string is hashed (FNV-1a; the reference seeds xoshiro256 with
// sha3<256>(seed), src/apps/hmc.cpp:93)
inline uint64_t seed_key(std::string const &s)
{
	if (!s.empty() && s.size() <= 20 && s.find_first_not_of("0123456789") == std::string::npos)
		return std::strtoull(s.c_str(), nullptr, 10);
	uint64_t h = 1469598103934665603ull;
	for (unsigned char c : s)
	{
		h ^= c;
		h *= 1099511628211ull;
	}
	return h;
}

// ---- molecular dynamics -------------------------------------------------------

// Integrator step table (src/gauge/hmc.cpp:7-51)
inline std::vector<double> makeDeltas(std::string_view scheme, double epsilon, int substeps)
{
	std::vector<double> d(9 * (size_t)std::max(substeps, 1) + 1);
	std::string s(scheme);
	int n = mesh_make_deltas(s.c_str(), epsilon, substeps, d.data(), (int)d.size());
	if (n < 0)
		throw std::runtime_error("unknown scheme '" + s + "'");
	d.resize(n);
	return d;
}

// U = exp(TA(P) t) U (src/gauge/hmc.h:26-36), per direction or whole stack
template <class G> void evolve(Field<G> &U, Field<G> const &P, double t)
{
	if (U.nfields() != P.nfields())
		throw std::runtime_error("evolve: U and P differ in the number of directions");
	ctx().check(mesh_su3_exp_mul(ctx().get(), U.grid().c(), U.nfields(), G::prec, U.data(), P.data(), t));
}

// Hamiltonian dynamics over deltas (src/gauge/hmc.h:38-59); even index: all
// directions' link step, odd index: fused momentum step
template <class G>
void runHmd(GaugeField<G> &U, GaugeField<G> &P, double beta, std::vector<double> const &deltas)
{
	if (deltas.size() % 2 != 1)
		throw std::runtime_error("runHmd: deltas must have odd length");
	for (size_t i = 0; i < deltas.size(); ++i)
		if (i % 2 == 0)
			ctx().check(mesh_su3_link_update(ctx().get(), U.grid().c(), G::prec, U.data(), U.data(), P.data(),
			                                 deltas[i], 0));
		else
			ctx().check(mesh_su3_force_update(ctx().get(), U.grid().c(), G::prec, U.data(), P.data(), beta,
			                                  deltas[i]));
}

namespace detail {
// same quantities as Hmc::print_summary (src/gauge/hmc.h:88-100)
inline void print_hmc_summary(std::vector<double> const &plaq, std::vector<double> const &dH,
                              std::vector<double> const &acc, std::vector<double> const &time)
{
	auto mean = [](std::vector<double> const &v) {
		return v.empty() ? 0.0 : std::accumulate(v.begin(), v.end(), 0.0) / v.size();
	};
	double mp = mean(plaq), var = 0;
	for (double p : plaq)
		var += (p - mp) * (p - mp);
	size_t n = plaq.size();
	var = n > 1 ? var / (n - 1) : 0;
	double emdh = 0;
	for (double d : dH)
		emdh += std::exp(-d);
	std::printf("========== HMC summary ==========\n");
	std::printf("time per step = %.3f s (min = %.3f, max = %.3f)\n", mean(time),
	            n ? *std::min_element(time.begin(), time.end()) : 0.0,
	            n ? *std::max_element(time.begin(), time.end()) : 0.0);
	std::printf("plaquette = %.4f +- %.4f\n", mp, n ? std::sqrt(var / n) : 0.0);
	std::printf("acceptance = %.2f\n", mean(acc));
	std::printf("<exp(-dH)> = %.4f\n", dH.empty() ? 0.0 : emdh / dH.size());
}
} // namespace detail

// HMC state and observables (src/gauge/hmc.h:61-115)
template <class G> class Hmc
{
  public:
	Grid const &g;
	// P is the trajectory's workspace: after runHmcUpdate it holds the final
	// momenta in the compact antihermitian layout (mesh_su3_hmc_trajectory),
	// after randomizeMomenta the matrix layout
	GaugeField<G> U, P, U_new;
	uint64_t seed = 0; // keys the Philox streams (the reference seeds xoshiro256)
	uint32_t traj = 0;
	std::vector<double> plaq_history, deltaH_history, accept_history, time_history;

  private:
	void *K_ = nullptr; // per-site kinetic energy, V doubles

  public:
	explicit Hmc(Grid const &g_) : g(g_), U(g_), P(g_), U_new(g_)
	{
		ctx().check(mesh_malloc(ctx().get(), &K_, g_.size() * sizeof(double)));
	}
	~Hmc()
	{
		if (K_)
			mesh_free(ctx().get(), K_);
	}
	Hmc(Hmc const &) = delete;
	Hmc &operator=(Hmc const &) = delete;

	void reset_observables()
	{
		plaq_history.clear();
		deltaH_history.clear();
		accept_history.clear();
		time_history.clear();
	}

	void randomizeGaugeField() { random_gauge_field(U, seed); }
	void randomizeMomenta()
	{
		ctx().check(mesh_su3_momenta(ctx().get(), g.c(), G::prec, P.data(), (double *)K_, seed, traj));
	}

	// momenta -> trajectory -> accept/reject -> measure (src/gauge/hmc.cpp:75-107);
	// metropolis == false accepts unconditionally (thermalisation)
	void runHmcUpdate(double beta, std::vector<double> const &deltas, bool metropolis = true)
	{
		auto t0 = std::chrono::steady_clock::now();
		void *pu = U.data(), *pw = U_new.data();
		double out[6];
		ctx().check(mesh_su3_hmc_trajectory(ctx().get(), g.c(), G::prec, &pu, &pw, P.data(), (double *)K_, beta,
		                                    deltas.data(), (int)deltas.size(), seed, traj, metropolis ? 1 : 0,
		                                    out));
		if (pu != U.data())
			swap(U, U_new);
		++traj;
		double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
		accept_history.push_back(out[1]);
		plaq_history.push_back(out[2]);
		deltaH_history.push_back(out[0]);
		time_history.push_back(dt);
	}

	void print_summary() const
	{
		detail::print_hmc_summary(plaq_history, deltaH_history, accept_history, time_history);
	}
};

// U(1) links as angles on a 2- to 4-d lattice: theta[f][x] for nf
// directions (nf = Nd: the gauge field, nf = 1: a gauge transformation
// g = exp(i phi)), FP32 (the reference's GaugeField<U1<T>> of complex<T>
// links exp(i theta))
class U1Field
{
	Grid const *g_;
	int nf_;
	void *p_ = nullptr;

  public:
	explicit U1Field(Grid const &g, int nf = -1) : g_(&g), nf_(nf < 0 ? g.ndim() : nf)
	{
		if (g.ndim() < 2 || g.ndim() > 4)
			throw std::runtime_error("U(1) fields are implemented on lattices of 2 to 4 dimensions");
		ctx().check(mesh_malloc(ctx().get(), &p_, bytes()));
	}
	~U1Field()
	{
		if (p_)
			mesh_free(ctx().get(), p_);
	}
	U1Field(U1Field const &) = delete;
	U1Field &operator=(U1Field const &) = delete;
	U1Field(U1Field &&o) noexcept : g_(o.g_), nf_(o.nf_), p_(o.p_) { o.p_ = nullptr; }
	Grid const &grid() const { return *g_; }
	int nf() const { return nf_; }
	size_t bytes() const { return (size_t)nf_ * g_->size() * sizeof(float); }
	float *data() const { return (float *)p_; }
	// theta = 0: the unit configuration
	void fill_unit() { ctx().check(mesh_memclear(ctx().get(), p_, bytes())); }
	void copy_from(U1Field const &o)
	{
		if (o.bytes() != bytes())
			throw std::runtime_error("U1Field::copy_from: size mismatch");
		ctx().check(mesh_memcpy(ctx().get(), p_, o.p_, bytes()));
	}
	std::vector<float> download() const
	{
		std::vector<float> h((size_t)nf_ * g_->size());
		ctx().check(mesh_memcpy_to_host(ctx().get(), h.data(), p_, bytes()));
		return h;
	}
	void upload(std::span<const float> h)
	{
		if (h.size() * sizeof(float) != bytes())
			throw std::runtime_error("U1Field::upload: size mismatch");
		ctx().check(mesh_memcpy_from_host(ctx().get(), p_, h.data(), bytes()));
	}
};

// mean cos theta_P (= Re tr P / Nc; plaquette, src/gauge/utils.h:93-108)
inline double plaquette(U1Field const &U)
{
	double *d = nullptr, h[4];
	ctx().check(mesh_malloc(ctx().get(), (void **)&d, sizeof h));
	int rc = mesh_u1_observables(ctx().get(), U.grid().c(), 1, U.data(), d);
	if (rc == MESH_OK)
		rc = mesh_memcpy_to_host(ctx().get(), h, d, sizeof h);
	mesh_free(ctx().get(), d);
	ctx().check(rc);
	const int nd = U.grid().ndim();
	return h[2] / ((double)U.grid().size() * nd * (nd - 1) / 2);
}

// HMC for U(1) (Hmc<U1<T>>, src/gauge/hmc.cpp:109-110) through mesh_u1_hmc
// with a batch of one: on small 2-d lattices the whole trajectory is one
// kernel, otherwise the site-parallel kernels
template <class T> class Hmc<U1<T>>
{
  public:
	Grid const &g;
	U1Field U;
	uint64_t seed = 0;
	uint32_t traj = 0;
	std::vector<double> plaq_history, deltaH_history, accept_history, time_history;

  private:
	double *out_ = nullptr; // device: 4 HMC results + 4 observables

  public:
	explicit Hmc(Grid const &g_) : g(g_), U(g_)
	{
		ctx().check(mesh_malloc(ctx().get(), (void **)&out_, 8 * sizeof(double)));
	}
	~Hmc()
	{
		if (out_)
			mesh_free(ctx().get(), out_);
	}
	Hmc(Hmc const &) = delete;
	Hmc &operator=(Hmc const &) = delete;

	void reset_observables()
	{
		plaq_history.clear();
		deltaH_history.clear();
		accept_history.clear();
		time_history.clear();
	}
	void randomizeGaugeField() { ctx().check(mesh_u1_random(ctx().get(), g.c(), 1, U.data(), seed)); }

	void runHmcUpdate(double beta, std::vector<double> const &deltas, bool metropolis = true)
	{
		auto t0 = std::chrono::steady_clock::now();
		Context &c = ctx();
		c.check(mesh_u1_hmc(c.get(), g.c(), 1, U.data(), deltas.data(), (int)deltas.size(), beta, seed, traj,
		                    metropolis ? 1 : 0, out_));
		c.check(mesh_u1_observables(c.get(), g.c(), 1, U.data(), out_ + 4));
		double h[8];
		c.check(mesh_memcpy_to_host(c.get(), h, out_, sizeof h));
		++traj;
		double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
		accept_history.push_back(h[1]);
		const int np = g.ndim() * (g.ndim() - 1) / 2;
		plaq_history.push_back(h[6] / ((double)g.size() * np)); // mean cos theta_P = Re tr P / Nc
		deltaH_history.push_back(h[0]);
		time_history.push_back(dt);
	}

	void print_summary() const
	{
		detail::print_hmc_summary(plaq_history, deltaH_history, accept_history, time_history);
	}
};

// SU(2) links as quaternions in the library's SoA layout (4 planes per
// direction); FP64 or FP32 after T
template <class T> class SU2Field
{
	Grid const *g_;
	int nf_;
	void *p_ = nullptr;

  public:
	static constexpr int prec = sizeof(T) == 8 ? MESH_F64 : MESH_F32;
	explicit SU2Field(Grid const &g, int nf = -1) : g_(&g), nf_(nf < 0 ? g.ndim() : nf)
	{
		if (g.ndim() < 2 || g.ndim() > 4)
			throw std::runtime_error("SU(2) fields are implemented on lattices of 2 to 4 dimensions");
		ctx().check(mesh_malloc(ctx().get(), &p_, bytes()));
	}
	~SU2Field()
	{
		if (p_)
			mesh_free(ctx().get(), p_);
	}
	SU2Field(SU2Field const &) = delete;
	SU2Field &operator=(SU2Field const &) = delete;
	SU2Field(SU2Field &&o) noexcept : g_(o.g_), nf_(o.nf_), p_(o.p_) { o.p_ = nullptr; }
	void copy_from(SU2Field const &o)
	{
		if (o.bytes() != bytes())
			throw std::runtime_error("SU2Field::copy_from: size mismatch");
		ctx().check(mesh_memcpy(ctx().get(), p_, o.p_, bytes()));
	}
	friend void swap(SU2Field &a, SU2Field &b) noexcept
	{
		std::swap(a.g_, b.g_);
		std::swap(a.nf_, b.nf_);
		std::swap(a.p_, b.p_);
	}
	Grid const &grid() const { return *g_; }
	int nf() const { return nf_; }
	size_t bytes() const { return mesh_su2_field_bytes(g_->c(), nf_, prec); }
	void *data() const { return p_; }
	void fill_unit() { ctx().check(mesh_su2_unit(ctx().get(), g_->c(), nf_, prec, p_)); }
	// host layout [nf][V][4] (w, x, y, z)
	std::vector<T> download() const
	{
		std::vector<T> h((size_t)nf_ * g_->size() * 4);
		ctx().check(mesh_su2_transfer(ctx().get(), g_->c(), nf_, prec, h.data(), p_, 0));
		return h;
	}
	void upload(std::span<const T> h)
	{
		if (h.size() != (size_t)nf_ * g_->size() * 4)
			throw std::runtime_error("SU2Field::upload: size mismatch");
		ctx().check(mesh_su2_transfer(ctx().get(), g_->c(), nf_, prec, (void *)h.data(), p_, 1));
	}
};

// mean Re tr P / 2 (plaquette, src/gauge/utils.h:93-108)
template <class T> double plaquette(SU2Field<T> const &U)
{
	double r;
	ctx().check(mesh_su2_plaquette(ctx().get(), U.grid().c(), SU2Field<T>::prec, U.data(), &r));
	return r;
}

// HMC for SU(2) (Hmc<SU2<T>>, src/gauge/hmc.cpp:111-112) through
// mesh_su2_hmc_trajectory.  The kinetic energy is the 2x2 matrix norm of the
// momenta, 2 (x^2 + y^2 + z^2): the norm that the force
// (beta / 4) project_on_algebra(U Sigma) conserves.
template <class T> class Hmc<SU2<T>>
{
  public:
	Grid const &g;
	SU2Field<T> U, P, U_new;
	uint64_t seed = 0;
	uint32_t traj = 0;
	std::vector<double> plaq_history, deltaH_history, accept_history, time_history;

  private:
	void *K_ = nullptr;

  public:
	explicit Hmc(Grid const &g_) : g(g_), U(g_), P(g_), U_new(g_)
	{
		ctx().check(mesh_malloc(ctx().get(), &K_, g_.size() * sizeof(double)));
	}
	~Hmc()
	{
		if (K_)
			mesh_free(ctx().get(), K_);
	}
	Hmc(Hmc const &) = delete;
	Hmc &operator=(Hmc const &) = delete;

	void reset_observables()
	{
		plaq_history.clear();
		deltaH_history.clear();
		accept_history.clear();
		time_history.clear();
	}
	void randomizeGaugeField()
	{
		ctx().check(mesh_su2_random(ctx().get(), g.c(), g.ndim(), SU2Field<T>::prec, U.data(), seed));
	}

	void runHmcUpdate(double beta, std::vector<double> const &deltas, bool metropolis = true)
	{
		auto t0 = std::chrono::steady_clock::now();
		void *pu = U.data(), *pw = U_new.data();
		double out[6];
		ctx().check(mesh_su2_hmc_trajectory(ctx().get(), g.c(), SU2Field<T>::prec, &pu, &pw, P.data(), (double *)K_,
		                                    beta, deltas.data(), (int)deltas.size(), seed, traj,
		                                    metropolis ? 1 : 0, out));
		if (pu != U.data())
			swap(U, U_new);
		++traj;
		double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
		accept_history.push_back(out[1]);
		plaq_history.push_back(out[2]);
		deltaH_history.push_back(out[0]);
		time_history.push_back(dt);
	}

	void print_summary() const
	{
		detail::print_hmc_summary(plaq_history, deltaH_history, accept_history, time_history);
	}
};

// ---- Landau gauge (src/gauge/landau.h) ----------------------------------------

namespace detail {
// The reference's solver (landau.h:67-136) over the field operations of a
// group, supplied by D: the gauge field `g` (type D::gauge_type), the
// functional F[h] as operator()(h), condition(), deriv() (D::deriv_type),
// dot(A, B) = sum Re tr (A B), axpy(Y, a, X), assign(dst, src),
// step(h, F, e): h = exp(e F) h, and reunit(h).
template <class D> struct LandauAscent
{
	// abscissa of the vertex of the parabola through three points (landau.h:68-75)
	static double solve_quadratic(double x1, double x2, double x3, double y1, double y2, double y3)
	{
		return 0.5 * ((x2 * x2 - x3 * x3) * y1 - (x1 * x1 - x3 * x3) * y2 + (x1 * x1 - x2 * x2) * y3) /
		       ((x2 - x3) * y1 - (x1 - x3) * y2 + (x1 - x2) * y3);
	}

	// step length along F (landau.h:77-98)
	template <class Fd> double line_search(Fd const &F, double eps) const
	{
		D const &d = static_cast<D const &>(*this);
		double base = d();
		typename D::gauge_type h = d.new_gauge();
		auto trial = [&](double e) {
			d.assign(h, d.g);
			d.step(h, F, e);
			return d(h);
		};
		double y1 = trial(eps);
		while (y1 <= base)
		{
			eps *= 0.5;
			y1 = trial(eps);
			if (eps < 1e-10)
				return 0.0;
		}
		double y2 = trial(eps * 1.2);
		if (y2 > y1)
			return 1.2 * eps;
		double y3 = trial(eps * 0.8);
		return solve_quadratic(eps, 1.2 * eps, 0.8 * eps, y1, y2, y3);
	}

	// ascent along deriv() with the line search, Polak-Ribiere steps after
	// iteration 20 (as written in landau.h:126-129, `last = F` precedes
	// `F += last * beta`, so the step is (1 + beta) F), then
	// g = reunitize(exp(eps F) g), until condition() < tol; returns iterations
	int run_ascent(int max_iter, double tol = 1e-10)
	{
		D &d = static_cast<D &>(*this);
		double eps = 3000.0, lastNorm = 0.0;
		typename D::deriv_type last = d.new_deriv();
		for (int iter = 0; iter < max_iter; ++iter)
		{
			double cond = d.condition();
			if (d.verbose)
				std::printf("iter = %d, cond = %.6e, term = %.12f, eps = %.6e\n", iter + 1, cond, d(), eps);
			if (cond < tol)
				return iter;
			typename D::deriv_type F = d.deriv();
			double norm = d.dot(F, F);
			double beta = iter > 20 ? (norm - d.dot(F, last)) / lastNorm : 0.0;
			d.assign(last, F);
			if (iter > 20 && beta > 0)
				d.axpy(F, beta, last);
			eps = line_search(F, eps);
			d.step(d.g, F, eps);
			d.reunit(d.g);
			lastNorm = norm;
		}
		return max_iter;
	}

	// Even/odd overrelaxation until theta = Nd * condition() < tol
	// (replaces the line-search ascent of landau.h:100-136); returns sweeps.
	// The sweeps update the transformed links U^g and g together; every
	// reunit_every sweeps g is reunitarised and U^g rebuilt from it, which
	// bounds the drift of g from the group that FP32 products accumulate.
	int run(int max_iter, double omega = 1.7, double tol = 1e-12, int check_every = 10, int reunit_every = 100)
	{
		D &d = static_cast<D &>(*this);
		auto Ug = d.gauge_fixed();
		int it = 0;
		while (it < max_iter)
		{
			for (int k = 0; k < check_every && it < max_iter; ++k)
			{
				for (int parity = 0; parity < 2; ++parity)
					d.or_halfsweep(Ug, parity, omega);
				++it;
				if (reunit_every > 0 && it % reunit_every == 0)
				{
					d.reunit(d.g);
					d.transform_into(Ug);
				}
			}
			last_theta = d.divsum_of(Ug) / (double)d.U.grid().size();
			if (d.verbose)
				std::printf("iter = %d, theta = %.3e\n", it, last_theta);
			if (last_theta < tol)
				break;
		}
		return it;
	}

	// theta of the transformed links at the last check of run()
	double last_theta = 0;
};
} // namespace detail

template <class G> struct Landau : detail::LandauAscent<Landau<G>>
{
	using gauge_type = Lattice<G>;
	using deriv_type = Lattice<G>;
	GaugeField<G> const &U;
	Lattice<G> g;
	bool verbose = false;

	void reset() { g.fill_unit(); }

	explicit Landau(GaugeField<G> const &U_) : U(U_), g(U_.grid()) { reset(); }

	// F[h] = sum Re tr h U h(x+mu)^+ / (Nc Nd V) (landau.h:23-34)
	double operator()(Lattice<G> const &h) const
	{
		double r;
		ctx().check(mesh_landau_functional(ctx().get(), U.grid().c(), G::prec, U.data(), h.data(), &r));
		return r;
	}
	double operator()() const { return (*this)(g); }

	// (1/(Nd V)) sum ||TA(g d)||^2 (landau.h:37-49)
	double condition() const
	{
		double r;
		ctx().check(mesh_landau_divsum(ctx().get(), U.grid().c(), G::prec, U.data(), g.data(), &r));
		return r / (U.grid().ndim() * (double)U.grid().size());
	}

	// -0.5/(Nc Nd V) TA(g d) (landau.h:52-65)
	Lattice<G> deriv() const
	{
		Lattice<G> F(U.grid());
		ctx().check(mesh_landau_deriv(ctx().get(), U.grid().c(), G::prec, U.data(), g.data(), F.data()));
		return F;
	}

	// the gauge-transformed field U^g_mu(x) = g(x) U_mu(x) g(x+mu)^+
	GaugeField<G> gauge_fixed() const
	{
		GaugeField<G> Ug(U.grid());
		transform_into(Ug);
		return Ug;
	}

	// field operations of the solvers
	Lattice<G> new_gauge() const { return Lattice<G>(U.grid()); }
	Lattice<G> new_deriv() const { return Lattice<G>(U.grid()); }
	void assign(Lattice<G> &dst, Lattice<G> const &src) const { dst = src; }
	void step(Lattice<G> &h, Lattice<G> const &F, double e) const { evolve(h, F, e); }
	void reunit(Lattice<G> &h) const { reunitize(h); }
	double dot(Lattice<G> const &A, Lattice<G> const &B) const { return sum_real_trace_prod(A, B); }
	void axpy(Lattice<G> &Y, double a, Lattice<G> const &X) const { add_mul(Y, a, X); }
	void transform_into(GaugeField<G> &Ug) const
	{
		ctx().check(mesh_su3_gauge_transform(ctx().get(), U.grid().c(), G::prec, Ug.data(), U.data(), g.data()));
	}
	void or_halfsweep(GaugeField<G> &Ug, int parity, double omega)
	{
		ctx().check(mesh_landau_or_halfsweep(ctx().get(), U.grid().c(), G::prec, Ug.data(), g.data(), parity, omega));
	}
	double divsum_of(GaugeField<G> const &Ug) const
	{
		double d;
		ctx().check(mesh_landau_divsum(ctx().get(), U.grid().c(), G::prec, Ug.data(), nullptr, &d));
		return d;
	}
};

// Landau<SU2<T>>: g one quaternion per site
template <class T> struct Landau<SU2<T>> : detail::LandauAscent<Landau<SU2<T>>>
{
	using F = SU2Field<T>;
	using gauge_type = F;
	using deriv_type = F;
	static constexpr int prec = F::prec;
	F const &U;
	F g;
	bool verbose = false;

	void reset() { g.fill_unit(); }
	explicit Landau(F const &U_) : U(U_), g(U_.grid(), 1) { reset(); }

	double operator()(F const &h) const
	{
		double r;
		ctx().check(mesh_su2_landau_functional(ctx().get(), U.grid().c(), prec, U.data(), h.data(), &r));
		return r;
	}
	double operator()() const { return (*this)(g); }
	double condition() const
	{
		double r;
		ctx().check(mesh_su2_landau_divsum(ctx().get(), U.grid().c(), prec, U.data(), g.data(), &r));
		return r / (U.grid().ndim() * (double)U.grid().size());
	}
	F deriv() const
	{
		F D(U.grid(), 1);
		ctx().check(mesh_su2_landau_deriv(ctx().get(), U.grid().c(), prec, U.data(), g.data(), D.data()));
		return D;
	}
	F gauge_fixed() const
	{
		F Ug(U.grid());
		transform_into(Ug);
		return Ug;
	}

	F new_gauge() const { return F(U.grid(), 1); }
	F new_deriv() const { return F(U.grid(), 1); }
	void assign(F &dst, F const &src) const { dst.copy_from(src); }
	void step(F &h, F const &D, double e) const
	{
		ctx().check(mesh_su2_link_update(ctx().get(), U.grid().c(), 1, prec, h.data(), h.data(), D.data(), e, 0));
	}
	void reunit(F &h) const { ctx().check(mesh_su2_reunit(ctx().get(), U.grid().c(), h.nf(), prec, h.data())); }
	double dot(F const &A, F const &B) const
	{
		double r;
		ctx().check(mesh_su2_dot(ctx().get(), U.grid().c(), A.nf(), prec, A.data(), B.data(), &r));
		return r;
	}
	void axpy(F &Y, double a, F const &X) const
	{
		ctx().check(mesh_su2_axpy(ctx().get(), U.grid().c(), Y.nf(), prec, Y.data(), a, X.data()));
	}
	void transform_into(F &Ug) const
	{
		ctx().check(mesh_su2_gauge_transform(ctx().get(), U.grid().c(), prec, Ug.data(), U.data(), g.data()));
	}
	void or_halfsweep(F &Ug, int parity, double omega)
	{
		ctx().check(
		    mesh_su2_landau_or_halfsweep(ctx().get(), U.grid().c(), prec, Ug.data(), g.data(), parity, omega));
	}
	double divsum_of(F const &Ug) const
	{
		double d;
		ctx().check(mesh_su2_landau_divsum(ctx().get(), U.grid().c(), prec, Ug.data(), nullptr, &d));
		return d;
	}
};

// Landau<U1<T>>: g = exp(i phi), phi one FP32 angle per site
template <class T> struct Landau<U1<T>> : detail::LandauAscent<Landau<U1<T>>>
{
	using gauge_type = U1Field;
	using deriv_type = U1Field;
	U1Field const &U;
	U1Field g;
	bool verbose = false;

	void reset() { g.fill_unit(); }
	explicit Landau(U1Field const &U_) : U(U_), g(U_.grid(), 1) { reset(); }

	double operator()(U1Field const &h) const
	{
		double r;
		ctx().check(mesh_u1_landau_functional(ctx().get(), U.grid().c(), U.data(), h.data(), &r));
		return r;
	}
	double operator()() const { return (*this)(g); }
	double condition() const
	{
		double r;
		ctx().check(mesh_u1_landau_divsum(ctx().get(), U.grid().c(), U.data(), g.data(), &r));
		return r / (U.grid().ndim() * (double)U.grid().size());
	}
	U1Field deriv() const
	{
		U1Field D(U.grid(), 1);
		ctx().check(mesh_u1_landau_deriv(ctx().get(), U.grid().c(), U.data(), g.data(), D.data()));
		return D;
	}
	U1Field gauge_fixed() const
	{
		U1Field Ug(U.grid());
		transform_into(Ug);
		return Ug;
	}

	U1Field new_gauge() const { return U1Field(U.grid(), 1); }
	U1Field new_deriv() const { return U1Field(U.grid(), 1); }
	void assign(U1Field &dst, U1Field const &src) const { dst.copy_from(src); }
	void step(U1Field &h, U1Field const &D, double e) const
	{
		ctx().check(mesh_u1_gauge_step(ctx().get(), U.grid().c(), h.data(), h.data(), D.data(), e));
	}
	void reunit(U1Field &) const {} // angles stay on the group
	double dot(U1Field const &A, U1Field const &B) const
	{
		double r;
		ctx().check(mesh_u1_dot(ctx().get(), (long)(A.bytes() / sizeof(float)), A.data(), B.data(), &r));
		return r;
	}
	void axpy(U1Field &Y, double a, U1Field const &X) const
	{
		ctx().check(mesh_u1_axpy(ctx().get(), (long)(Y.bytes() / sizeof(float)), Y.data(), a, X.data()));
	}
	void transform_into(U1Field &Ug) const
	{
		ctx().check(mesh_u1_gauge_transform(ctx().get(), U.grid().c(), Ug.data(), U.data(), g.data()));
	}
	void or_halfsweep(U1Field &Ug, int parity, double omega)
	{
		ctx().check(mesh_u1_landau_or_halfsweep(ctx().get(), U.grid().c(), Ug.data(), g.data(), parity, omega));
	}
	double divsum_of(U1Field const &Ug) const
	{
		double d;
		ctx().check(mesh_u1_landau_divsum(ctx().get(), U.grid().c(), Ug.data(), nullptr, &d));
		return d;
	}
};

// the field type of a group's gauge configuration
template <class G> struct FieldOf
{
	using type = GaugeField<G>;
};
template <class T> struct FieldOf<SU2<T>>
{
	using type = SU2Field<T>;
};
template <class T> struct FieldOf<U1<T>>
{
	using type = U1Field;
};

// ---- runtime dispatch (src/gauge/utils.h:155-186) ------------------------------

template <typename F> void dispatchByGroup(F f, std::string const &group, int precision)
{
	if (precision != 1 && precision != 2)
		throw std::runtime_error("invlid precision level '" + std::to_string(precision) + "'");
	if (group == "su3")
	{
		if (precision == 1)
			f.template operator()<SU3<float>>();
		else
			f.template operator()<SU3<double>>();
	}
	else if (group == "su2")
	{
		if (precision == 1)
			f.template operator()<SU2<float>>();
		else
			f.template operator()<SU2<double>>();
	}
	else if (group == "u1")
	{
		if (precision == 1)
			f.template operator()<U1<float>>();
		else
			f.template operator()<U1<double>>();
	}
	else
		throw std::runtime_error("unknown gauge group '" + group + "'");
}

} // namespace mesh
