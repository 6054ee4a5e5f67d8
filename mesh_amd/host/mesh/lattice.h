// C++ host layer over libmesh_hip.so: geometry, device context and lattice
// containers with the call surface of krox/mesh's src/lattice/{grid.h,
// lattice.h, device.h}.  Header-only; needs only include/mesh_hip.h (no HIP
// headers) and links against libmesh_hip.so.
//
// Differences from the reference, by design:
//  * storage is the library's blocked structure-of-arrays layout, so element
//    access goes through copy_from_host / copy_to_host (AoS order as in the
//    reference, src/lattice/lattice.h:102-115);
//  * arbitrary device lambdas (lattice_apply, src/lattice/lattice.h:256-261)
//    are replaced by the named fused operations of mesh/gauge.h;
//  * errors from the C ABI are rethrown as std::runtime_error
//    ("hip error: ...", as src/lattice/device.cpp:10-19).
#pragma once

#include "mesh_hip.h"

#include <algorithm>
#include <array>
#include <complex>
#include <cstdint>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace mesh {

// ---------------------------------------------------------------- context

class Context
{
	mesh_ctx *ctx_ = nullptr;

  public:
	explicit Context(int device = 0)
	{
		int rc = mesh_ctx_create(&ctx_, device);
		if (rc != MESH_OK)
			throw std::runtime_error("hip error: cannot create a context on device " + std::to_string(device) +
			                         " (code " + std::to_string(rc) + ")");
	}
	~Context()
	{
		if (ctx_)
			mesh_ctx_destroy(ctx_);
	}
	Context(Context const &) = delete;
	Context &operator=(Context const &) = delete;

	mesh_ctx *get() const { return ctx_; }

	// throws on a non-zero status of the last call on this context
	void check(int rc) const
	{
		if (rc != MESH_OK)
			throw std::runtime_error(std::string(mesh_last_error(ctx_)) + " (code " + std::to_string(rc) + ")");
	}
	void synchronize() const { check(mesh_synchronize(ctx_)); }

	// process-wide default context on device 0 (the reference uses the
	// implicit CUDA context)
	static Context &global()
	{
		static Context c(0);
		return c;
	}
};

// untyped backend, same names as src/lattice/device.h:14-24
inline void *device_malloc(size_t n)
{
	void *p = nullptr;
	Context::global().check(mesh_malloc(Context::global().get(), &p, n));
	return p;
}
inline void device_free(void *p) { Context::global().check(mesh_free(Context::global().get(), p)); }
inline void device_memcpy(void *dst, void const *src, size_t n)
{
	Context::global().check(mesh_memcpy(Context::global().get(), dst, src, n));
}
inline void device_memcpy_2d(void *dst, size_t dpitch, void const *src, size_t spitch, size_t width, size_t height)
{
	Context::global().check(mesh_memcpy_2d(Context::global().get(), dst, dpitch, src, spitch, width, height));
}
inline void device_memclear(void *ptr, size_t n)
{
	Context::global().check(mesh_memclear(Context::global().get(), ptr, n));
}
inline void device_memcpy_to_host(void *dst, void const *src, size_t n)
{
	Context::global().check(mesh_memcpy_to_host(Context::global().get(), dst, src, n));
}
inline void device_memcpy_from_host(void *dst, void const *src, size_t n)
{
	Context::global().check(mesh_memcpy_from_host(Context::global().get(), dst, src, n));
}
inline void device_synchronize() { Context::global().synchronize(); }

// ---------------------------------------------------------------- grid

using Coordinate = std::vector<int32_t>;

// lattice geometry, row-major, dimension 0 slowest (src/lattice/grid.h:13-58)
class Grid
{
	mesh_grid g_ = {1, {0, 0, 0, 0}};
	size_t size_ = 0;

  public:
	Grid() = default;
	explicit Grid(Coordinate const &shape)
	{
		if (shape.size() < 2 || shape.size() > 4)
			throw std::runtime_error("grid must have 2 to 4 dimensions");
		g_.nd = (int)shape.size();
		size_ = 1;
		for (size_t i = 0; i < shape.size(); ++i)
		{
			g_.dims[i] = shape[i];
			size_ *= shape[i];
		}
	}
	int ndim() const { return g_.nd; }
	size_t size() const { return size_; }
	Coordinate shape() const { return Coordinate(g_.dims, g_.dims + g_.nd); }
	int32_t shape(size_t i) const { return g_.dims[i]; }
	size_t flat_index(Coordinate const &index) const
	{
		size_t r = 0;
		for (int i = 0; i < ndim(); ++i)
		{
			if (index[i] < 0 || index[i] >= shape(i))
				throw std::out_of_range("coordinate out of range");
			r = r * shape(i) + index[i];
		}
		return r;
	}
	std::string to_string() const
	{
		std::string s = "[";
		for (int i = 0; i < ndim(); ++i)
			s += (i ? ", " : "") + std::to_string(shape(i));
		return s + "]";
	}
	mesh_grid const *c() const { return &g_; }
	bool operator==(Grid const &o) const
	{
		return g_.nd == o.g_.nd && std::equal(g_.dims, g_.dims + g_.nd, o.g_.dims);
	}
};

// ---------------------------------------------------------------- groups

// element types: the reference's U1<T> = complex<T>, SU3<T> = Matrix<complex<T>,3>
//
// SU3<T> is one 3x3 complex matrix, row-major (re, im) pairs -- the value a
// device lambda of lattice_apply (mesh/lattice_apply.h) sees for one site --
// and carries the group's traits.
template <class T> struct SU3
{
	using real = T;
	using element = std::array<std::complex<T>, 9>; // row-major (host transfers)
	static constexpr int prec = sizeof(T) == 8 ? MESH_F64 : MESH_F32;
	static std::string name() { return "su3"; }
	static constexpr int Nc() { return 3; }
	static constexpr int dim() { return 8; }
	T v[18];
};

// U(1): the reference's complex<T> link exp(i theta) is stored as its angle,
// in FP32 whatever T is (energies, sums and the Metropolis test are FP64);
// lattices of 2 to 4 dimensions
template <class T> struct U1
{
	using real = T;
	using element = float; // the angle theta in [-pi, pi)
	static constexpr int prec = MESH_F32;
	static std::string name() { return "u1"; }
	static constexpr int Nc() { return 1; }
	static constexpr int dim() { return 1; }
};

// SU(2): unit quaternions (w, x, y, z) = w - i (x, y, z).sigma, the
// reference's util::quaternion<T> (re, im1, im2, im3)
template <class T> struct SU2
{
	using real = T;
	using element = std::array<T, 4>;
	static constexpr int prec = sizeof(T) == 8 ? MESH_F64 : MESH_F32;
	static std::string name() { return "su2"; }
	static constexpr int Nc() { return 2; }
	static constexpr int dim() { return 3; }
};

template <class G> struct GaugeTraits
{
	static std::string name() { return G::name(); }
	static constexpr int Nc() { return G::Nc(); }
	static constexpr int dim() { return G::dim(); }
};

// ---------------------------------------------------------------- containers

namespace detail {
struct DeviceDeleter
{
	mesh_ctx *ctx;
	void operator()(void *p) const noexcept
	{
		if (p)
			mesh_free(ctx, p);
	}
};
} // namespace detail

// nf matrices per site on the device: nf == 1 is Lattice<G>, nf == Nd the
// gauge-field stack (LatticeStack<G>, src/lattice/lattice.h:159-192)
template <class G> class Field
{
  protected:
	Grid grid_;
	int nf_ = 0;
	size_t bytes_ = 0;
	std::shared_ptr<void> owner_; // null for a view
	void *ptr_ = nullptr;

	void allocate()
	{
		bytes_ = mesh_su3_field_bytes(grid_.c(), nf_, G::prec);
		if (bytes_ == 0)
			throw std::runtime_error("invalid lattice geometry " + grid_.to_string());
		void *p = nullptr;
		Context &c = Context::global();
		c.check(mesh_malloc(c.get(), &p, bytes_));
		owner_ = std::shared_ptr<void>(p, detail::DeviceDeleter{c.get()});
		ptr_ = p;
	}

  public:
	using value_type = typename G::element;

	Field() = default;
	Field(Grid const &g, int nf) : grid_(g), nf_(nf) { allocate(); }

	// deep copy (src/lattice/lattice.h:71-76, 83-93)
	Field(Field const &o) : grid_(o.grid_), nf_(o.nf_)
	{
		allocate();
		device_memcpy(ptr_, o.ptr_, bytes_);
	}
	// copy assignment reallocates only if the size differs
	// (src/lattice/lattice.h:83-93); a view of one direction of a stack keeps
	// its memory, as the reference's U[mu] = ... writes into the stack
	Field &operator=(Field const &o)
	{
		if (this == &o)
			return *this;
		if (!ptr_ || bytes_ != o.bytes_)
		{
			if (!owner_ && ptr_)
				throw std::runtime_error("assignment of a differently sized lattice to a view");
			grid_ = o.grid_;
			nf_ = o.nf_;
			allocate();
		}
		device_memcpy(ptr_, o.ptr_, bytes_);
		return *this;
	}
	Field(Field &&o) noexcept = default;
	// move assignment steals the buffer (src/lattice/lattice.h:95-100), except
	// into a view, whose memory belongs to its stack: there it copies
	Field &operator=(Field &&o)
	{
		if (this == &o)
			return *this;
		if (!owner_ && ptr_)
			return *this = static_cast<Field const &>(o);
		grid_ = std::move(o.grid_);
		nf_ = std::exchange(o.nf_, 0);
		bytes_ = std::exchange(o.bytes_, 0);
		owner_ = std::move(o.owner_);
		ptr_ = std::exchange(o.ptr_, nullptr);
		return *this;
	}

	Grid const &grid() const { return grid_; }
	int nfields() const { return nf_; }
	size_t bytes() const { return bytes_; }
	void *data() { return ptr_; }
	void const *data() const { return ptr_; }
	size_t size() const { return grid_.size() * nf_; }

	void fill_zeros() { device_memclear(ptr_, bytes_); }
	void fill_unit()
	{
		Context &c = Context::global();
		c.check(mesh_su3_unit(c.get(), grid_.c(), nf_, G::prec, ptr_));
	}
	// host transfer in the reference's AoS order
	void copy_from_host(std::span<const value_type> data)
	{
		if (data.size() != size())
			throw std::runtime_error("copy_from_host: size mismatch");
		Context &c = Context::global();
		c.check(mesh_su3_upload(c.get(), grid_.c(), nf_, G::prec, data.data(), ptr_));
	}
	std::vector<value_type> copy_to_host() const
	{
		std::vector<value_type> r(size());
		Context &c = Context::global();
		c.check(mesh_su3_download(c.get(), grid_.c(), nf_, G::prec, ptr_, r.data()));
		return r;
	}
	friend void swap(Field &a, Field &b) noexcept
	{
		std::swap(a.grid_, b.grid_);
		std::swap(a.nf_, b.nf_);
		std::swap(a.bytes_, b.bytes_);
		std::swap(a.owner_, b.owner_);
		std::swap(a.ptr_, b.ptr_);
	}
};

template <class G> class Lattice : public Field<G>
{
  public:
	Lattice() = default;
	explicit Lattice(Grid const &g) : Field<G>(g, 1) {}
	static Lattice zeros(Grid const &g)
	{
		Lattice a(g);
		a.fill_zeros();
		return a;
	}
	// view of one direction of a stack (no ownership of its own)
	static Lattice view(Field<G> &stack, int mu)
	{
		Lattice v;
		v.grid_ = stack.grid();
		v.nf_ = 1;
		v.bytes_ = stack.bytes() / stack.nfields();
		v.ptr_ = static_cast<char *>(stack.data()) + mu * v.bytes_;
		return v;
	}
};

template <class G> class LatticeStack : public Field<G>
{
	// one non-owning Lattice view per direction, rebuilt when the stack's
	// memory changes (copy, move, swap)
	mutable std::vector<Lattice<G>> views_;
	mutable void const *views_of_ = nullptr;

	std::vector<Lattice<G>> &views() const
	{
		if (views_of_ != this->ptr_ || (int)views_.size() != this->nf_)
		{
			views_.clear();
			for (int mu = 0; mu < this->nf_; ++mu)
				views_.push_back(Lattice<G>::view(const_cast<LatticeStack &>(*this), mu));
			views_of_ = this->ptr_;
		}
		return views_;
	}

  public:
	LatticeStack() = default;
	explicit LatticeStack(Grid const &g) : Field<G>(g, g.ndim()) {}
	LatticeStack(LatticeStack const &o) : Field<G>(o) {}
	LatticeStack(LatticeStack &&o) noexcept : Field<G>(std::move(o)) {}
	LatticeStack &operator=(LatticeStack const &o)
	{
		Field<G>::operator=(o);
		return *this;
	}
	LatticeStack &operator=(LatticeStack &&o)
	{
		Field<G>::operator=(std::move(o));
		return *this;
	}
	size_t size() const { return (size_t)this->nf_; }
	// direction mu: a Lattice that views the stack's memory
	// (src/lattice/lattice.h:190-191 returns a reference to it as well)
	Lattice<G> &operator[](size_t mu)
	{
		if ((int)mu >= this->nf_)
			throw std::out_of_range("direction out of range");
		return views()[mu];
	}
	Lattice<G> const &operator[](size_t mu) const
	{
		if ((int)mu >= this->nf_)
			throw std::out_of_range("direction out of range");
		return views()[mu];
	}
};

template <class G> using GaugeField = LatticeStack<G>;

} // namespace mesh
