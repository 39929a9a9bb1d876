// mirror_driver.cpp — drives the MARS:: C++ interface (mars-md_amd/host/MARS.hpp) the
// way SimManager drives the reference's Patch, for the GPU test
// tests/test_gpu_cpp_mirror.py.
//
// Input (raw little-endian, written by the test):
//   int64 n, int64 num_knots, int64 steps, uint64 seed,
//   float L, cutoff, pairlist_cutoff, dt, kT, diffusion, table_start, table_step_inv,
//   float pos4[n][4], int32 types[n], float knots[num_knots]
// Output files (raw): <out>.force (n x 4 float, after the step-1 force pass),
// <out>.pairs (num_pairs x 2 int32, canonical, Patch pairlist), <out>.pairs_standalone
// (the same from a separate MARS::Pairlist built on the input positions),
// <out>.pos (n x 4 float after `steps` BD steps), and a summary line on stdout.
#include "MARS.hpp"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

namespace {

template<class T>
void read_into(std::ifstream& f, T* dst, size_t count) {
	f.read(reinterpret_cast<char*>(dst), std::streamsize(sizeof(T) * count));
	if (!f)
		throw MARS::Exception(MARS::ExceptionType::FileIoError, "short input file");
}

template<class T>
void write_file(const std::string& path, const T* src, size_t count) {
	std::ofstream f(path, std::ios::binary);
	f.write(reinterpret_cast<const char*>(src), std::streamsize(sizeof(T) * count));
	if (!f)
		throw MARS::Exception(MARS::ExceptionType::FileIoError, "cannot write " + path);
}

std::vector<MARS::int2> pairs_to_host(MARS::Pairlist& pl) {
	const size_t np = pl.get_num_pairs();
	std::vector<MARS::int2> host(np);
	const MARS::int2* dev = pl.get_neighbor_pairs();
	if (np && hipMemcpy(host.data(), dev, np * sizeof(MARS::int2), hipMemcpyDeviceToHost) != hipSuccess)
		throw MARS::Exception(MARS::ExceptionType::CUDARuntimeError, "hipMemcpy pairs");
	return host;
}

} // namespace

int main(int argc, char** argv) {
	if (argc != 3) {
		std::fprintf(stderr, "usage: %s <input> <output prefix>\n", argv[0]);
		return 2;
	}
	try {
		std::ifstream in(argv[1], std::ios::binary);
		int64_t n = 0, num_knots = 0, steps = 0;
		uint64_t seed = 0;
		float L, cutoff, plcut, dt, kT, D, tstart, tstep_inv;
		read_into(in, &n, 1);
		read_into(in, &num_knots, 1);
		read_into(in, &steps, 1);
		read_into(in, &seed, 1);
		float hdr[8];
		read_into(in, hdr, 8);
		L = hdr[0];
		cutoff = hdr[1];
		plcut = hdr[2];
		dt = hdr[3];
		kT = hdr[4];
		D = hdr[5];
		tstart = hdr[6];
		tstep_inv = hdr[7];
		std::vector<MARS::Vector3> pos(static_cast<size_t>(n));
		std::vector<int> types(static_cast<size_t>(n));
		std::vector<float> knots(static_cast<size_t>(num_knots));
		read_into(in, reinterpret_cast<float*>(pos.data()), size_t(4 * n));
		read_into(in, types.data(), size_t(n));
		read_into(in, knots.data(), size_t(num_knots));

		const MARS::PeriodicBox box(MARS::Vector3{L, L, L});
		MARS::PatchSetup setup;
		setup.types.num_types = 1;
		setup.types.mass[0] = 1.0f;
		for (int a = 0; a < 3; ++a)
			setup.types.diffusion[0][a] = D;
		setup.term_matrix = {uint32_t(MARSMD_PAIR_TERM_TABULATED)}; // table 0 for the (0, 0) pair
		MARS::TabulatedPotential t{};
		t.pot = knots.data();
		t.size = uint32_t(num_knots);
		t.start = tstart;
		t.step_inv = tstep_inv;
		t.is_periodic = 0;
		setup.nb_tables.push_back(t);

		MARS::Patch patch(0, size_t(n), setup, box);
		patch.set_base_seed(seed);
		patch.copy_particles_from_host(pos.data(), nullptr, types.data(), size_t(n));

		// step 1: the force pass alone, as SimManager::execute_force_calculation does
		patch.calculate_nonbonded_forces(plcut, cutoff, 1, 20, MARS::Vector3{}, MARS::InterpolationOrder::Linear, true)
			.wait();
		std::vector<MARS::Vector3> fe(static_cast<size_t>(n));
		patch.copy_particles_to_host(nullptr, nullptr, fe.data(), nullptr, nullptr, size_t(n));
		write_file(std::string(argv[2]) + ".force", reinterpret_cast<const float*>(fe.data()), size_t(4 * n));
		const std::vector<MARS::int2> pp = pairs_to_host(patch.get_pairlist());
		write_file(std::string(argv[2]) + ".pairs", reinterpret_cast<const int*>(pp.data()), 2 * pp.size());
		const MARS::PairlistStatistics st = patch.get_pairlist().get_statistics();

		// the rest of step 1 and steps 2..steps: integrate, then force + integrate
		patch.integrate_motion(dt, kT, MARS::IntegratorType::Brownian, 1).wait();
		for (int64_t s = 2; s <= steps; ++s) {
			patch.calculate_nonbonded_forces(plcut, cutoff, size_t(s), 20);
			patch.integrate_motion(dt, kT, MARS::IntegratorType::Brownian, size_t(s)).wait();
		}
		std::vector<MARS::Vector3> pos_out(static_cast<size_t>(n));
		patch.copy_particles_to_host(pos_out.data(), nullptr, nullptr, nullptr, nullptr, size_t(n));
		write_file(std::string(argv[2]) + ".pos", reinterpret_cast<const float*>(pos_out.data()), size_t(4 * n));

		// a stand-alone pairlist on the input positions (device copy), as the unit tests use it
		MARS::Vector3* dpos = nullptr;
		if (hipMalloc(reinterpret_cast<void**>(&dpos), sizeof(MARS::Vector3) * size_t(n)) != hipSuccess)
			throw MARS::Exception(MARS::ExceptionType::CUDARuntimeError, "hipMalloc");
		if (hipMemcpy(dpos, pos.data(), sizeof(MARS::Vector3) * size_t(n), hipMemcpyHostToDevice) != hipSuccess)
			throw MARS::Exception(MARS::ExceptionType::CUDARuntimeError, "hipMemcpy");
		{
			// max_pairs is the regrow ceiling (the reference passes half the device memory)
			auto pl = MARS::create_pairlist(MARS::PairlistBuilderType::ZOrder, size_t(n), size_t(1) << 32);
			pl->set_periodic_box(box);
			pl->build_pairlist(dpos, size_t(n), plcut).wait();
			const std::vector<MARS::int2> ps = pairs_to_host(*pl);
			write_file(std::string(argv[2]) + ".pairs_standalone", reinterpret_cast<const int*>(ps.data()),
					   2 * ps.size());
			const bool stale = pl->needs_update(dpos, size_t(n), plcut - cutoff);
			std::printf("standalone_pairs %zu needs_update %d\n", ps.size(), stale ? 1 : 0);
		}
		(void)hipFree(dpos);

		// error behaviour: an unsupported builder and an out-of-range type are exceptions
		int caught = 0;
		try {
			(void)MARS::create_pairlist(MARS::PairlistBuilderType::CellList, 16, 16);
		} catch (const MARS::Exception& e) {
			caught += e.type() == MARS::ExceptionType::NotImplementedError ? 1 : 0;
		}
		try {
			std::vector<int> bad(types);
			bad[0] = 7;
			patch.copy_particles_from_host(pos.data(), nullptr, bad.data(), size_t(n));
		} catch (const MARS::Exception& e) {
			caught += e.type() == MARS::ExceptionType::ValueError ? 1 : 0;
		}
		std::printf("pairs %zu max_neighbors %u grid_bits %d exceptions %d\n", pp.size(),
					st.max_neighbors_per_particle, st.cell_grid_bits, caught);
	} catch (const MARS::Exception& e) {
		std::fprintf(stderr, "MARS::Exception (%d): %s\n", int(e.type()), e.what());
		return 1;
	}
	return 0;
}
