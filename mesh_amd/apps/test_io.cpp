// Ensemble file format of mesh/io.h, host only (no GPU call):
//   test_io write <dir>   writes an ensemble with known attributes and arrays
//   test_io read <dir>    prints what an ensemble holds, one "key = value" a line
//   test_io overwrite <dir>  checks what create(..., allow_overwrite) replaces
#include "mesh/io.h"

#include <cstdio>
#include <filesystem>
#include <string>

using namespace mesh;

int main(int argc, char **argv)
{
	if (argc != 3)
	{
		std::fprintf(stderr, "usage: test_io write|read|overwrite <dir>\n");
		return 2;
	}
	std::string mode = argv[1], dir = argv[2];
	try
	{
		if (mode == "overwrite")
		{
			// allow_overwrite replaces an ensemble (its configurations and
			// histories go) and refuses a directory that is not one
			namespace fs = std::filesystem;
			std::string ens = dir + "/ens", other = dir + "/other";
			auto e = io::Ensemble::create(ens, false);
			e.write_data("plaq_history", {1.0});
			std::vector<double> u(2 * 2 * 2 * 18, 1.0);
			io::write_npy(e.config_path("1"), "<f8", io::config_shape(Grid({2, 2}), 18), u.data());
			io::Ensemble::create(ens, true);
			if (fs::exists(ens + "/plaq_history.npy") || !fs::is_empty(ens + "/configs"))
			{
				std::printf("old ensemble files kept\n");
				return 1;
			}
			fs::create_directories(other);
			std::FILE *f = std::fopen((other + "/notes.txt").c_str(), "w");
			std::fputs("x", f);
			std::fclose(f);
			try
			{
				io::Ensemble::create(other, true);
				std::printf("overwrite of a non-ensemble not refused\n");
				return 1;
			}
			catch (std::runtime_error const &)
			{
			}
			if (!fs::exists(other + "/notes.txt"))
			{
				std::printf("non-ensemble touched\n");
				return 1;
			}
			std::printf("overwrite ok\n");
			return 0;
		}
		if (mode == "write")
		{
			auto e = io::Ensemble::create(dir, false);
			e.set_attribute("group", "su3");
			e.set_attribute("beta", 5.75);
			e.set_attribute("geometry", std::vector<int>{2, 3, 4});
			e.set_attribute("markov_seed", "a \"quoted\" seed\\path");
			e.set_attribute("markov_count", 7);
			e.set_attribute("philox_key", std::to_string(18446744073709551615ull));
			e.set_attribute("philox_traj0", 4294967295ll);
			e.write_data("plaq_history", {0.5, 0.25, 1.0 / 3.0});
			std::vector<double> u(3 * 2 * 3 * 4 * 18);
			for (size_t i = 0; i < u.size(); ++i)
				u[i] = 0.5 * (double)i - 1.0 / 7.0;
			io::write_npy(e.config_path("1"), "<f8", io::config_shape(Grid({2, 3, 4}), 18), u.data());
			std::vector<float> t(2 * 5 * 6);
			for (size_t i = 0; i < t.size(); ++i)
				t[i] = (float)i / 3.0f;
			io::write_npy(e.config_path("2"), "<f4", io::config_shape(Grid({5, 6}), 1), t.data());
			// a second create of the same directory must refuse, and so must a
			// create over a non-empty directory that is not an ensemble
			for (std::string d : {dir, dir + "/configs"})
				try
				{
					io::Ensemble::create(d, false);
					std::printf("overwrite of %s not refused\n", d.c_str());
					return 1;
				}
				catch (std::runtime_error const &)
				{
				}
			std::printf("written\n");
			return 0;
		}
		auto e = io::Ensemble::open(dir);
		auto const &at = e.attributes();
		std::printf("group = %s\n", at.get_string("group").c_str());
		std::printf("beta = %.17g\n", at.get_double("beta"));
		auto geom = at.get_ints("geometry");
		std::printf("geometry =");
		for (int x : geom)
			std::printf(" %d", x);
		std::printf("\n");
		std::printf("markov_seed = %s\n", at.get_string("markov_seed").c_str());
		std::printf("markov_count = %lld\n", at.get_int("markov_count"));
		auto h = e.read_data("plaq_history");
		double s = 0;
		for (double x : h)
			s += x;
		std::printf("plaq_history = %zu %.17g\n", h.size(), s);
		for (char const *name : {"1", "2"})
		{
			io::Npy a = io::read_npy(e.config_path(name));
			double sum = 0;
			for (size_t i = 0; i < a.count(); ++i)
				sum += a.descr == "<f8" ? ((double const *)a.data.data())[i] : ((float const *)a.data.data())[i];
			std::printf("config %s = %s", name, a.descr.c_str());
			for (size_t x : a.shape)
				std::printf(" %zu", x);
			std::printf(" sum %.17g\n", sum);
		}
		std::printf("config name = %s | %s\n", io::split_config_name(dir + "/configs/12").first.c_str(),
		            io::split_config_name(dir + "/configs/12").second.c_str());
		return 0;
	}
	catch (std::exception const &ex)
	{
		std::fprintf(stderr, "error: %s\n", ex.what());
		return 1;
	}
}
