// mirror_driver.cpp — drives the MARS:: C++ interface (mars-md_amd/host/MARS.hpp) the
// way SimManager drives the reference's Patch: the reference's Patch constructor, its
// containers (TablesRegistry, NonBondedInteractions, BondedInteractions,
// DeviceParticleTypes, grid views, Temperature, HostParticleData) and the per-step calls
// of SimManager::execute_force_calculation / execute_integration with their argument
// lists unchanged (SimManager.cpp:343-349 reorder cadence, 401-425 forces, 485
// integration). For the GPU test tests/test_gpu_cpp_mirror.py.
//
// Input (raw little-endian, written by the test):
//   int64 n, num_types, num_pair_entries, nb_knots, num_nb_tables, num_bonds, num_angles,
//         num_excl, bond_knots, angle_knots, steps, rebuild_period, reorder_period,
//         energy_period
//   uint64 seed
//   float L, cutoff, pairlist_cutoff, timestep, temperature (K), nb_step
//   float types[num_types][3]            mass, charge, diffusion
//   int32 pairs[num_pair_entries][3]     type a, type b, nonbonded table index
//   float nb_y[num_nb_tables][nb_knots]  (abscissa k * nb_step)
//   float bond_x[bond_knots], bond_y[bond_knots]
//   float angle_x[angle_knots] (degrees), angle_y[angle_knots]
//   float pos[n][4], int32 type_ids[n], int32 bonds[num_bonds][2],
//   int32 angles[num_angles][3], int32 excl[num_excl][2]
// Output files (raw): <out>.force (n x 4 float by global id, after the step-1 force
// pass, energy in the 4th), <out>.pairs (step-1 pair set, global ids),
// <out>.pos (n x 4 float by global id after `steps` steps), <out>.pairs_standalone
// (a separate MARS::Pairlist on the input positions), and summary lines on stdout.
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

// The patch's particles as n x 4 floats in global-id order (force + energy, or pos).
std::vector<float> by_global_id(const MARS::Patch& patch, bool force, size_t n) {
	MARS::HostParticleData h;
	patch.copy_particles_to_host(h, 0, n, true);
	std::vector<float> out(4 * n, 0.0f);
	for (size_t i = 0; i < n; ++i) {
		const size_t g = size_t(h.global_id[i]);
		const MARS::Vector3& v = force ? h.force[i] : h.pos[i];
		out[4 * g + 0] = v.x;
		out[4 * g + 1] = v.y;
		out[4 * g + 2] = v.z;
		out[4 * g + 3] = force ? h.energy[i] : 0.0f;
	}
	return out;
}

} // namespace

int main(int argc, char** argv) {
	using namespace MARS;
	if (argc != 3) {
		std::fprintf(stderr, "usage: %s <input> <output prefix>\n", argv[0]);
		return 2;
	}
	try {
		std::ifstream in(argv[1], std::ios::binary);
		int64_t H[14];
		read_into(in, H, 14);
		const size_t n = size_t(H[0]), num_types = size_t(H[1]), num_pair_entries = size_t(H[2]),
					 nb_knots = size_t(H[3]), num_nb_tables = size_t(H[4]), num_bonds = size_t(H[5]),
					 num_angles = size_t(H[6]), num_excl = size_t(H[7]), bond_knots = size_t(H[8]),
					 angle_knots = size_t(H[9]);
		const int64_t steps = H[10], rebuild_period = H[11], reorder_period = H[12], energy_period = H[13];
		uint64_t seed = 0;
		read_into(in, &seed, 1);
		float prm[6];
		read_into(in, prm, 6);
		const float L = prm[0], cutoff = prm[1], plcut = prm[2], timestep = prm[3], temp_K = prm[4],
					nb_step = prm[5];
		std::vector<float> tprm(3 * num_types);
		read_into(in, tprm.data(), tprm.size());
		std::vector<int32_t> pair_entries(3 * num_pair_entries);
		read_into(in, pair_entries.data(), pair_entries.size());
		std::vector<float> nb_y(num_nb_tables * nb_knots), bond_x(bond_knots), bond_y(bond_knots),
			angle_x(angle_knots), angle_y(angle_knots);
		read_into(in, nb_y.data(), nb_y.size());
		read_into(in, bond_x.data(), bond_knots);
		read_into(in, bond_y.data(), bond_knots);
		read_into(in, angle_x.data(), angle_knots);
		read_into(in, angle_y.data(), angle_knots);
		std::vector<Vector3> pos(n);
		std::vector<int> type_ids(n);
		std::vector<int32_t> bonds(2 * num_bonds), angles(3 * num_angles), excl(2 * num_excl);
		read_into(in, reinterpret_cast<float*>(pos.data()), 4 * n);
		read_into(in, type_ids.data(), n);
		read_into(in, bonds.data(), bonds.size());
		read_into(in, angles.data(), angles.size());
		read_into(in, excl.data(), excl.size());

		// ---- the system, as SimSystem / SystemState / TablesRegistry hold it
		const Resource resource(0);
		const PeriodicBox box(Vector3{L, L, L});
		TablesRegistry tables;
		for (size_t t = 0; t < num_nb_tables; ++t) {
			Table tab(TabulatedType::NonBondedPair);
			std::vector<mars_real> x(nb_knots), y(nb_y.begin() + long(t * nb_knots),
												  nb_y.begin() + long((t + 1) * nb_knots));
			for (size_t k = 0; k < nb_knots; ++k)
				x[k] = mars_real(k) * nb_step;
			tab.set_xy(std::move(x), std::move(y));
			tables.add_nonbonded(std::move(tab));
		}
		Table bond_tab(TabulatedType::Bond);
		bond_tab.set_xy(bond_x, bond_y);
		const int bond_idx = tables.add_bond(std::move(bond_tab));
		Table angle_tab(TabulatedType::Angle);
		angle_tab.set_xy(angle_x, angle_y); // degrees in, radians kept
		const int angle_idx = tables.add_angle(std::move(angle_tab));

		std::vector<PairNonBonded> pn;
		for (size_t k = 0; k < num_pair_entries; ++k)
			pn.emplace_back(pair_entries[3 * k], pair_entries[3 * k + 1], "tabulated", pair_entries[3 * k + 2]);
		const NonBondedInteractions nonbonded(pn);

		std::vector<Bond> bl;
		for (size_t k = 0; k < num_bonds; ++k)
			bl.emplace_back(bonds[2 * k], bonds[2 * k + 1], bond_idx);
		std::vector<Angle> al(num_angles);
		for (size_t k = 0; k < num_angles; ++k) {
			al[k].ind1 = angles[3 * k];
			al[k].ind2 = angles[3 * k + 1];
			al[k].ind3 = angles[3 * k + 2];
			al[k].function_index = angle_idx;
		}
		std::vector<Exclude> el;
		for (size_t k = 0; k < num_excl; ++k)
			el.emplace_back(excl[2 * k], excl[2 * k + 1]);
		const BondedInteractions bonded(bl, al, {}, el, {});

		std::vector<ParticleType> ptypes(num_types);
		for (size_t t = 0; t < num_types; ++t) {
			ptypes[t].name = "T" + std::to_string(t);
			ptypes[t].mass = tprm[3 * t];
			ptypes[t].charge = tprm[3 * t + 1];
			ptypes[t].diffusion = Vector3{tprm[3 * t + 2], tprm[3 * t + 2], tprm[3 * t + 2]};
		}
		const DeviceParticleTypes particle_types(ptypes, resource);
		const DeviceBuffer<BaseGridView<mars_real>> grid_views; // no grids
		const Temperature temperature(temp_K);

		HostParticleData host;
		host.resize(n);
		for (size_t i = 0; i < n; ++i) {
			host.global_id[i] = int(i);
			host.type_id[i] = type_ids[i];
			host.pos[i] = pos[i];
		}

		// ---- Patch (Patch.h:57-69), particles in (Patch.cpp:507-554)
		Patch patch(0, n, resource, box);
		patch.set_base_seed(seed);
		patch.copy_particles_from_host(host, 0, n);
		const size_t resource_idx = 0;

		std::vector<MARS::int2> step1_pairs;
		size_t energy_lines = 0;
		for (int64_t si = 1; si <= steps; ++si) {
			const size_t step = size_t(si);
			// ParticleReorderManager::maybe_reorder (ReorderManager.cpp:8-25)
			const int64_t rp = reorder_period > 0 ? reorder_period : rebuild_period;
			if (rp > 0 && (step - 1) % size_t(rp) == 0 && patch.is_bonded_prepared())
				patch.reorder_particles();
			// SimManager::execute_force_calculation (SimManager.cpp:401-425)
			const size_t energy_output_period = static_cast<size_t>(energy_period);
			const bool compute_energy = energy_output_period > 0 && step % energy_output_period == 0;
			Event evt = patch.calculate_nonbonded_forces(nonbonded,
														 bonded,
														 particle_types,
														 grid_views,
														 tables,
														 resource_idx,
														 static_cast<float>(plcut),
														 static_cast<float>(cutoff),
														 step,
														 static_cast<size_t>(rebuild_period),
														 Vector3{0.0, 0.0, 0.0},
														 InterpolationOrder::Linear,
														 compute_energy);

			Event bonded_evt = patch.calculate_bonded_forces(bonded,
															 particle_types,
															 tables,
															 resource_idx,
															 compute_energy);

			bonded_evt.wait();
			evt.wait();
			if (step == 1) {
				const std::vector<float> fe = by_global_id(patch, true, n);
				write_file(std::string(argv[2]) + ".force", fe.data(), fe.size());
				step1_pairs = pairs_to_host(patch.get_pairlist());
			}
			energy_lines += compute_energy ? 1 : 0;
			// SimManager::execute_integration (SimManager.cpp:485)
			Event ievt = patch.integrate_motion(timestep, temperature, IntegratorType::Brownian, step);
			ievt.wait();
		}
		write_file(std::string(argv[2]) + ".pairs", reinterpret_cast<const int*>(step1_pairs.data()),
				   2 * step1_pairs.size());
		const std::vector<float> pout = by_global_id(patch, false, n);
		write_file(std::string(argv[2]) + ".pos", pout.data(), pout.size());
		const PairlistStatistics st = patch.get_pairlist().get_statistics();

		// a stand-alone pairlist on the input positions (device copy), as the unit tests use it
		Vector3* dpos = nullptr;
		if (hipMalloc(reinterpret_cast<void**>(&dpos), sizeof(Vector3) * n) != hipSuccess)
			throw Exception(ExceptionType::CUDARuntimeError, "hipMalloc");
		if (hipMemcpy(dpos, pos.data(), sizeof(Vector3) * n, hipMemcpyHostToDevice) != hipSuccess)
			throw Exception(ExceptionType::CUDARuntimeError, "hipMemcpy");
		{
			// max_pairs is the regrow ceiling (the reference passes half the device memory)
			auto pl = create_pairlist(PairlistBuilderType::ZOrder, n, size_t(1) << 32);
			pl->set_periodic_box(box);
			pl->build_pairlist(dpos, n, plcut).wait();
			const std::vector<MARS::int2> ps = pairs_to_host(*pl);
			write_file(std::string(argv[2]) + ".pairs_standalone", reinterpret_cast<const int*>(ps.data()),
					   2 * ps.size());
			const bool stale = pl->needs_update(dpos, n, plcut - cutoff);
			// ZOrderSort displacement tracking on the same positions: none moved
			ZOrderSort& sorter = pl->get_sorter();
			sorter.update_positions_incremental(dpos, n);
			const float moved = sorter.compute_max_displacement(dpos, n, Vector3{L, L, L});
			std::printf("standalone_pairs %zu needs_update %d max_displacement %g\n", ps.size(), stale ? 1 : 0,
						double(moved));
		}
		(void)hipFree(dpos);

		// error behaviour: an unsupported builder, an out-of-range type, a host range past
		// the data, and containers swapped after the first force call are exceptions
		int caught = 0;
		try {
			(void)create_pairlist(PairlistBuilderType::CellList, 16, 16);
		} catch (const Exception& e) {
			caught += e.type() == ExceptionType::NotImplementedError ? 1 : 0;
		}
		try {
			HostParticleData bad = host;
			bad.type_id[0] = 7;
			patch.copy_particles_from_host(bad, 0, n);
		} catch (const Exception& e) {
			caught += e.type() == ExceptionType::ValueError ? 1 : 0;
		}
		try {
			patch.copy_particles_from_host(host, 1, n);
		} catch (const Exception& e) {
			caught += e.type() == ExceptionType::ValueError ? 1 : 0;
		}
		try {
			const NonBondedInteractions other(pn);
			patch.calculate_nonbonded_forces(other, bonded, particle_types, grid_views, tables, resource_idx, plcut,
											 cutoff, size_t(steps + 1), size_t(rebuild_period));
		} catch (const Exception& e) {
			caught += e.type() == ExceptionType::ValueError ? 1 : 0;
		}
		std::printf("pairs %zu max_neighbors %u grid_bits %d energy_steps %zu exceptions %d\n", step1_pairs.size(),
					st.max_neighbors_per_particle, st.cell_grid_bits, energy_lines, caught);
	} catch (const MARS::Exception& e) {
		std::fprintf(stderr, "MARS::Exception (%d): %s\n", int(e.type()), e.what());
		return 1;
	}
	return 0;
}
