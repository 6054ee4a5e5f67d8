// Ensemble and configuration files.
//
// The reference keeps an ensemble in one HDF5 file: the attributes group,
// beta, geometry, hmc_{scheme,epsilon,substeps,metropolis} and
// markov_{seed,start,discard,count,spacing} (src/apps/hmc.cpp:65-86), one
// dataset "/configs/<n>" per configuration of shape [Nd, L0, ..., L(nd-1),
// ncomp] (writeToFile, src/lattice/lattice.h:205-233) and the four histories
// (hmc.cpp:119-124); readConfig opens "file.h5/configs/<n>" and checks the
// group and the geometry (src/gauge/utils.h:188-219).  HDF5 is not part of
// this build, so an ensemble is a directory that holds the same things:
//
//     <dir>/attributes.json       one flat JSON object, the reference's
//                                 attribute names plus precision, philox_key
//                                 and philox_traj0 (below)
//     <dir>/configs/<n>.npy       NumPy format 1.0, little endian, C order:
//                                 SU(3) [Nd, L0.., 18] reals (each link's 3x3
//                                 complex matrix row-major, re/im interleaved),
//                                 U(1) [Nd, L0.., 1] FP32 angles
//     <dir>/<history>.npy         one-dimensional FP64 arrays
//
// and a configuration is named "<dir>/configs/<n>" as in the reference.
// philox_key is the 64-bit key of the Philox streams and philox_traj0 the
// counter of the chain's first trajectory, so configuration n of a chain
// continues bit for bit from counter philox_traj0 + n.
#pragma once
#include "mesh/gauge.h"

#include <cstdio>
#include <cstring>
#include <filesystem>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace mesh {
namespace io {

// ---- NumPy .npy files (format 1.0; 2.0 and 3.0 are read) -------------------

struct Npy
{
	std::string descr; // "<f8", "<f4", ...
	std::vector<size_t> shape;
	std::vector<char> data;
	size_t count() const
	{
		size_t n = 1;
		for (size_t s : shape)
			n *= s;
		return n;
	}
};

inline size_t descr_size(std::string const &d)
{
	if (d == "<f8" || d == "<i8" || d == "<u8")
		return 8;
	if (d == "<f4" || d == "<i4" || d == "<u4")
		return 4;
	throw std::runtime_error("npy: unsupported dtype '" + d + "'");
}

// written to path.tmp and renamed, so a reader never sees a partial file
inline void write_npy(std::string const &path, std::string const &descr, std::vector<size_t> const &shape,
                      void const *data)
{
	size_t n = descr_size(descr);
	std::string sh = "(";
	for (size_t i = 0; i < shape.size(); ++i)
	{
		n *= shape[i];
		sh += std::to_string(shape[i]) + (shape.size() == 1 ? "," : (i + 1 < shape.size() ? ", " : ""));
	}
	sh += ")";
	std::string hdr = "{'descr': '" + descr + "', 'fortran_order': False, 'shape': " + sh + ", }";
	// magic (6) + version (2) + length (2) + header + '\n', padded to 64
	size_t total = 10 + hdr.size() + 1;
	hdr.append((64 - total % 64) % 64, ' ');
	hdr += '\n';
	if (hdr.size() > 65535)
		throw std::runtime_error("npy: header too long");
	std::string tmp = path + ".tmp";
	FILE *f = std::fopen(tmp.c_str(), "wb");
	if (!f)
		throw std::runtime_error("cannot create '" + tmp + "'");
	unsigned char pre[10] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0, (unsigned char)(hdr.size() & 255),
	                         (unsigned char)(hdr.size() >> 8)};
	bool ok = std::fwrite(pre, 1, 10, f) == 10 && std::fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size() &&
	          (n == 0 || std::fwrite(data, 1, n, f) == n);
	ok = (std::fclose(f) == 0) && ok;
	if (!ok)
		throw std::runtime_error("write error on '" + tmp + "'");
	std::filesystem::rename(tmp, path);
}

inline Npy read_npy(std::string const &path)
{
	FILE *f = std::fopen(path.c_str(), "rb");
	if (!f)
		throw std::runtime_error("cannot open '" + path + "'");
	auto fail = [&](std::string const &m) {
		std::fclose(f);
		throw std::runtime_error("npy '" + path + "': " + m);
	};
	unsigned char pre[8];
	if (std::fread(pre, 1, 8, f) != 8 || std::memcmp(pre, "\x93NUMPY", 6) != 0)
		fail("not a .npy file");
	size_t hlen = 0;
	if (pre[6] == 1)
	{
		unsigned char l[2];
		if (std::fread(l, 1, 2, f) != 2)
			fail("truncated header");
		hlen = l[0] | (size_t)l[1] << 8;
	}
	else if (pre[6] == 2 || pre[6] == 3)
	{
		unsigned char l[4];
		if (std::fread(l, 1, 4, f) != 4)
			fail("truncated header");
		hlen = l[0] | (size_t)l[1] << 8 | (size_t)l[2] << 16 | (size_t)l[3] << 24;
	}
	else
		fail("unknown format version");
	std::string hdr(hlen, '\0');
	if (std::fread(hdr.data(), 1, hlen, f) != hlen)
		fail("truncated header");
	Npy r;
	auto value_of = [&](std::string const &key) -> size_t {
		size_t p = hdr.find("'" + key + "'");
		if (p == std::string::npos)
			fail("header has no '" + key + "'");
		p = hdr.find(':', p);
		if (p == std::string::npos)
			fail("malformed header");
		return p + 1;
	};
	size_t p = value_of("descr");
	size_t q0 = hdr.find('\'', p), q1 = q0 == std::string::npos ? q0 : hdr.find('\'', q0 + 1);
	if (q1 == std::string::npos)
		fail("malformed descr");
	r.descr = hdr.substr(q0 + 1, q1 - q0 - 1);
	if (r.descr.size() == 3 && r.descr[0] == '|')
		r.descr[0] = '<';
	p = value_of("fortran_order");
	while (p < hdr.size() && hdr[p] == ' ')
		++p;
	if (hdr.compare(p, 5, "False") != 0)
		fail("Fortran-ordered arrays are not supported");
	p = value_of("shape");
	size_t a = hdr.find('(', p), b = a == std::string::npos ? a : hdr.find(')', a);
	if (b == std::string::npos)
		fail("malformed shape");
	for (size_t i = a + 1; i < b;)
	{
		while (i < b && (hdr[i] == ' ' || hdr[i] == ','))
			++i;
		if (i >= b)
			break;
		size_t j = i;
		while (j < b && hdr[j] >= '0' && hdr[j] <= '9')
			++j;
		if (j == i)
			fail("malformed shape");
		r.shape.push_back(std::stoull(hdr.substr(i, j - i)));
		i = j;
	}
	size_t n = descr_size(r.descr) * r.count();
	r.data.resize(n);
	if (n && std::fread(r.data.data(), 1, n, f) != n)
		fail("truncated data");
	std::fclose(f);
	return r;
}

// ---- flat JSON object of attributes ---------------------------------------

// Values are kept as their JSON text: strings, numbers, arrays of numbers,
// true / false / null.  Nested objects are not used by ensemble files.
class Attributes
{
	std::map<std::string, std::string> v_;

	static std::string quote(std::string const &s)
	{
		std::string r = "\"";
		for (char c : s)
		{
			if (c == '"' || c == '\\')
				r += '\\', r += c;
			else if (c == '\n')
				r += "\\n";
			else if ((unsigned char)c < 0x20)
			{
				char b[8];
				std::snprintf(b, sizeof b, "\\u%04x", c);
				r += b;
			}
			else
				r += c;
		}
		return r + "\"";
	}

  public:
	void set(std::string const &k, std::string const &s) { v_[k] = quote(s); }
	void set(std::string const &k, char const *s) { v_[k] = quote(s); }
	void set(std::string const &k, double x)
	{
		char b[40];
		std::snprintf(b, sizeof b, "%.17g", x);
		v_[k] = b;
	}
	void set(std::string const &k, long long x) { v_[k] = std::to_string(x); }
	void set(std::string const &k, int x) { v_[k] = std::to_string(x); }
	void set(std::string const &k, std::vector<int> const &a)
	{
		std::string r = "[";
		for (size_t i = 0; i < a.size(); ++i)
			r += (i ? ", " : "") + std::to_string(a[i]);
		v_[k] = r + "]";
	}
	bool has(std::string const &k) const { return v_.count(k) != 0; }
	std::string const &raw(std::string const &k) const
	{
		auto it = v_.find(k);
		if (it == v_.end())
			throw std::runtime_error("attribute '" + k + "' not found");
		return it->second;
	}
	std::string get_string(std::string const &k) const
	{
		std::string const &t = raw(k);
		if (t.size() < 2 || t.front() != '"')
			throw std::runtime_error("attribute '" + k + "' is not a string");
		std::string r;
		for (size_t i = 1; i + 1 < t.size(); ++i)
		{
			if (t[i] != '\\')
			{
				r += t[i];
				continue;
			}
			char c = t[++i];
			if (c == 'n')
				r += '\n';
			else if (c == 't')
				r += '\t';
			else if (c == 'u' && i + 4 < t.size())
			{
				r += (char)std::stoi(t.substr(i + 1, 4), nullptr, 16);
				i += 4;
			}
			else
				r += c;
		}
		return r;
	}
	double get_double(std::string const &k) const { return std::stod(raw(k)); }
	long long get_int(std::string const &k) const { return std::stoll(raw(k)); }
	std::vector<int> get_ints(std::string const &k) const
	{
		std::string const &t = raw(k);
		if (t.empty() || t.front() != '[')
			throw std::runtime_error("attribute '" + k + "' is not an array");
		std::vector<int> r;
		for (size_t i = 1; i < t.size();)
		{
			while (i < t.size() && (t[i] == ' ' || t[i] == ',' || t[i] == '\n'))
				++i;
			if (i >= t.size() || t[i] == ']')
				break;
			size_t used = 0;
			r.push_back(std::stoi(t.substr(i), &used));
			i += used;
		}
		return r;
	}

	std::string to_json() const
	{
		std::string r = "{";
		bool first = true;
		for (auto const &[k, v] : v_)
		{
			r += (first ? "\n  " : ",\n  ") + quote(k) + ": " + v;
			first = false;
		}
		return r + "\n}\n";
	}

	static Attributes parse(std::string const &s)
	{
		Attributes a;
		size_t i = 0;
		auto ws = [&]() {
			while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\r' || s[i] == '\t'))
				++i;
		};
		auto fail = [&](char const *m) { throw std::runtime_error(std::string("attributes: ") + m); };
		auto string_end = [&](size_t j) {
			for (++j; j < s.size(); ++j)
				if (s[j] == '\\')
					++j;
				else if (s[j] == '"')
					return j + 1;
			fail("unterminated string");
			return j;
		};
		ws();
		if (i >= s.size() || s[i] != '{')
			fail("expected an object");
		++i;
		ws();
		if (i < s.size() && s[i] == '}')
			return a;
		for (;;)
		{
			ws();
			if (i >= s.size() || s[i] != '"')
				fail("expected a key");
			size_t e = string_end(i);
			Attributes key;
			key.v_["k"] = s.substr(i, e - i);
			std::string k = key.get_string("k");
			i = e;
			ws();
			if (i >= s.size() || s[i] != ':')
				fail("expected ':'");
			++i;
			ws();
			size_t b = i;
			if (i < s.size() && s[i] == '"')
				i = string_end(i);
			else if (i < s.size() && s[i] == '[')
			{
				while (i < s.size() && s[i] != ']')
					++i;
				if (i >= s.size())
					fail("unterminated array");
				++i;
			}
			else
				while (i < s.size() && s[i] != ',' && s[i] != '}' && s[i] != ' ' && s[i] != '\n')
					++i;
			if (i == b)
				fail("expected a value");
			a.v_[k] = s.substr(b, i - b);
			ws();
			if (i < s.size() && s[i] == ',')
			{
				++i;
				continue;
			}
			if (i < s.size() && s[i] == '}')
				return a;
			fail("expected ',' or '}'");
		}
	}
};

// ---- ensemble directory ----------------------------------------------------

class Ensemble
{
	std::filesystem::path dir_;
	Attributes attr_;

	void flush() const
	{
		std::string tmp = (dir_ / "attributes.json.tmp").string();
		FILE *f = std::fopen(tmp.c_str(), "wb");
		if (!f)
			throw std::runtime_error("cannot create '" + tmp + "'");
		std::string j = attr_.to_json();
		bool ok = std::fwrite(j.data(), 1, j.size(), f) == j.size();
		ok = (std::fclose(f) == 0) && ok;
		if (!ok)
			throw std::runtime_error("write error on '" + tmp + "'");
		std::filesystem::rename(tmp, dir_ / "attributes.json");
	}

  public:
	Ensemble() = default;
	explicit operator bool() const { return !dir_.empty(); }
	std::filesystem::path const &dir() const { return dir_; }
	Attributes const &attributes() const { return attr_; }

	// a new ensemble directory; an existing non-empty one only with
	// allow_overwrite, which drops its configurations (the reference's
	// Hdf5File::create(filename, allow_overwrite))
	// A new ensemble in `dir`: a missing or empty directory, or -- with
	// allow_overwrite -- an existing ensemble (a directory holding
	// attributes.json), whose configurations and data files are removed first.
	// Anything else is refused, so that a stray path cannot be emptied.
	static Ensemble create(std::string const &dir, bool allow_overwrite)
	{
		namespace fs = std::filesystem;
		const fs::path d(dir);
		if (fs::exists(d) && !(fs::is_directory(d) && fs::is_empty(d)))
		{
			if (!allow_overwrite)
				throw std::runtime_error("ensemble '" + dir + "' exists (allow overwriting to replace it)");
			if (!fs::is_directory(d) || !fs::is_regular_file(d / "attributes.json"))
				throw std::runtime_error("'" + dir + "' is not an ensemble (no attributes.json); not overwritten");
			fs::remove_all(d / "configs");
			std::vector<fs::path> data;
			for (auto const &f : fs::directory_iterator(d))
				if (f.is_regular_file() && f.path().extension() == ".npy")
					data.push_back(f.path());
			for (auto const &f : data)
				fs::remove(f);
		}
		Ensemble e;
		e.dir_ = dir;
		fs::create_directories(e.dir_ / "configs");
		e.flush();
		return e;
	}
	static Ensemble open(std::string const &dir)
	{
		Ensemble e;
		e.dir_ = dir;
		std::string path = (e.dir_ / "attributes.json").string();
		FILE *f = std::fopen(path.c_str(), "rb");
		if (!f)
			throw std::runtime_error("cannot open '" + path + "'");
		std::string s;
		char buf[4096];
		size_t n;
		while ((n = std::fread(buf, 1, sizeof buf, f)) > 0)
			s.append(buf, n);
		std::fclose(f);
		e.attr_ = Attributes::parse(s);
		return e;
	}

	template <class T> void set_attribute(std::string const &k, T const &v)
	{
		attr_.set(k, v);
		flush();
	}

	void write_data(std::string const &name, std::vector<double> const &v) const
	{
		write_npy((dir_ / (name + ".npy")).string(), "<f8", {v.size()}, v.data());
	}
	std::vector<double> read_data(std::string const &name) const
	{
		Npy a = read_npy((dir_ / (name + ".npy")).string());
		if (a.descr != "<f8" || a.shape.size() != 1)
			throw std::runtime_error("'" + name + "' is not a one-dimensional FP64 array");
		std::vector<double> r(a.count());
		std::memcpy(r.data(), a.data.data(), a.data.size());
		return r;
	}
	std::string config_path(std::string const &name) const { return (dir_ / "configs" / (name + ".npy")).string(); }
};

// "dir/configs/<n>" -> (ensemble directory, configuration name)
inline std::pair<std::string, std::string> split_config_name(std::string const &name)
{
	auto p = std::filesystem::path(name);
	if (p.parent_path().filename() != "configs" || p.filename().empty())
		throw std::runtime_error("unknown config name '" + name + "' (expected <ensemble>/configs/<n>)");
	return {p.parent_path().parent_path().string(), p.filename().string()};
}

inline std::vector<size_t> config_shape(Grid const &g, size_t ncomp)
{
	std::vector<size_t> s{(size_t)g.ndim()};
	for (int d = 0; d < g.ndim(); ++d)
		s.push_back((size_t)g.shape(d));
	s.push_back(ncomp);
	return s;
}

} // namespace io

// store a gauge field as configuration `name` of the ensemble
// (writeToFile, src/lattice/lattice.h:205-233)
template <class G> void writeConfig(io::Ensemble const &e, std::string const &name, GaugeField<G> const &U)
{
	auto h = U.copy_to_host();
	static_assert(sizeof(h[0]) == 18 * sizeof(typename G::real));
	io::write_npy(e.config_path(name), sizeof(typename G::real) == 8 ? "<f8" : "<f4",
	              io::config_shape(U.grid(), 18), h.data());
}
inline void writeConfig(io::Ensemble const &e, std::string const &name, U1Field const &U)
{
	auto h = U.download();
	io::write_npy(e.config_path(name), "<f4", io::config_shape(U.grid(), 1), h.data());
}

template <class T> void writeConfig(io::Ensemble const &e, std::string const &name, SU2Field<T> const &U)
{
	auto h = U.download();
	io::write_npy(e.config_path(name), sizeof(T) == 8 ? "<f8" : "<f4", io::config_shape(U.grid(), 4), h.data());
}

namespace io {
// opens "dir/configs/<n>", checks the group and the geometry against the
// ensemble's attributes and the array's shape (readConfig,
// src/gauge/utils.h:188-219)
inline Npy open_config(std::string const &configName, std::string const &group, Grid const &g, size_t ncomp)
{
	auto [dir, name] = split_config_name(configName);
	Ensemble e = Ensemble::open(dir);
	auto const &at = e.attributes();
	if (at.has("group") && at.get_string("group") != group)
		throw std::runtime_error("group mismatch on load. Expected " + group + ", got " + at.get_string("group"));
	auto shape = g.shape();
	std::vector<int> want(shape.begin(), shape.end());
	if (at.has("geometry") && at.get_ints("geometry") != want)
		throw std::runtime_error("grid mismatch on load ('" + configName + "')");
	Npy a = read_npy(e.config_path(name));
	if (a.shape != config_shape(g, ncomp))
		throw std::runtime_error("shape mismatch on load ('" + configName + "')");
	return a;
}
} // namespace io

// loads configuration "dir/configs/<n>" into U, whose grid must be the
// ensemble's geometry.  An FP64 file loads into an FP32 field rounded to
// nearest, an FP32 file into an FP64 field exactly.
template <class G> void readConfig(std::string const &configName, GaugeField<G> &U)
{
	using R = typename G::real;
	io::Npy a = io::open_config(configName, G::name(), U.grid(), 18);
	size_t n = a.count();
	std::vector<R> v(n);
	if (a.descr == "<f8")
	{
		double const *s = (double const *)a.data.data();
		for (size_t i = 0; i < n; ++i)
			v[i] = (R)s[i];
	}
	else if (a.descr == "<f4")
	{
		float const *s = (float const *)a.data.data();
		for (size_t i = 0; i < n; ++i)
			v[i] = (R)s[i];
	}
	else
		throw std::runtime_error("configuration '" + configName + "' has dtype " + a.descr);
	using V = typename G::element;
	static_assert(sizeof(V) == 18 * sizeof(R));
	U.copy_from_host(std::span<const V>((V const *)v.data(), n / 18));
}
inline void readConfig(std::string const &configName, U1Field &U)
{
	io::Npy a = io::open_config(configName, "u1", U.grid(), 1);
	if (a.descr != "<f4")
		throw std::runtime_error("U(1) configuration '" + configName + "' is not FP32 angles");
	ctx().check(mesh_memcpy_from_host(ctx().get(), U.data(), a.data.data(), a.data.size()));
}

template <class T> void readConfig(std::string const &configName, SU2Field<T> &U)
{
	io::Npy a = io::open_config(configName, "su2", U.grid(), 4);
	const size_t n = a.count();
	std::vector<T> h(n);
	if (a.descr == "<f8")
		for (size_t i = 0; i < n; ++i)
			h[i] = (T)((double const *)a.data.data())[i];
	else if (a.descr == "<f4")
		for (size_t i = 0; i < n; ++i)
			h[i] = (T)((float const *)a.data.data())[i];
	else
		throw std::runtime_error("SU(2) configuration '" + configName + "' is neither <f8 nor <f4");
	U.upload(h);
}

// geometry stored with a configuration ("dir/configs/<n>")
inline std::vector<int> configGeometry(std::string const &configName)
{
	return io::Ensemble::open(io::split_config_name(configName).first).attributes().get_ints("geometry");
}

} // namespace mesh
