// MARS.hpp — the reference's C++ interface for the per-step particle path, over the
// marsmd C ABI (include/marsmd.h) and the gfx950 library behind it.
//
// Header-only, C++17, host compiler only (g++ or clang++; link -lmarsmd_hip
// -lamdhip64). Class, function and enum names, argument meaning and error behaviour
// follow the reference:
//   MARS::Exception / ExceptionType      src/MARSException.h:103-130
//   MARS::ZOrderSort                     src/PatchOperation/ZOrderKernels/ZOrderSort.h:43-313
//   MARS::Pairlist, create_pairlist      src/PatchOperation/Pairlist.h:56-252, Pairlist.cpp:7-31
//   MARS::launch_*                       src/Interactions/NonBondedInteraction.h:205-296,
//                                        DeviceBondedInteraction.h:495-546,
//                                        src/PatchOperation/Integrator.h:25-144
//   MARS::Patch                          src/PatchOperation/Patch.h:48-645
// Every call takes (or defaults to) one stream and never synchronises implicitly
// except where the reference reads a count back to the host (pairlist build,
// needs_update). Calls return an Event whose wait() synchronises that stream, as the
// reference's Event::wait does for the caller in SimManager.cpp:424-425.
//
// Where the HIP library owns state the reference passes per call (interaction
// containers, types, tables, grids), the Patch takes it once at construction
// (PatchSetup); the per-step signatures keep the reference's remaining arguments.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <exception>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/marsmd.h"

namespace MARS {

// ------------------------------------------------------------------ errors

enum class ExceptionType {
	UnspecifiedError = 0,
	NotImplementedError = 1,
	ValueError = 2,
	DivideByZeroError = 3,
	CUDARuntimeError = 4,
	SYCLRuntimeError = 5,
	MetalRuntimeError = 6,
	FileIoError = 7,
	FileOpenError = 8,
	RuntimeError = 9
};

class Exception : public std::exception {
  public:
	Exception(ExceptionType type, std::string message) : type_(type), message_(std::move(message)) {}
	const char* what() const noexcept override { return message_.c_str(); }
	ExceptionType type() const noexcept { return type_; }

  private:
	ExceptionType type_;
	std::string message_;
};

namespace detail {
// Status codes of the C ABI -> the reference's exception categories. A failed HIP
// runtime call is the HIP counterpart of the reference's CUDA_CHECK failure.
inline ExceptionType exception_type(marsmd_status s) {
	switch (s) {
	case MARSMD_ERR_VALUE:
		return ExceptionType::ValueError;
	case MARSMD_ERR_DEVICE:
		return ExceptionType::CUDARuntimeError;
	case MARSMD_ERR_RUNTIME:
	case MARSMD_ERR_CAPACITY:
		return ExceptionType::RuntimeError;
	default:
		return ExceptionType::UnspecifiedError;
	}
}
inline void check(marsmd_status s) {
	if (s != MARSMD_OK) {
		const char* msg = marsmd_last_error();
		throw Exception(exception_type(s), msg ? msg : "marsmd error");
	}
}
inline marsmd_stream as_stream(hipStream_t s) { return reinterpret_cast<marsmd_stream>(s); }
} // namespace detail

// ------------------------------------------------------------------ value types

// Vector3_t<float> (src/Types/Vector3.h:41-132): 16-byte aligned {x, y, z, t}.
struct alignas(16) Vector3 {
	float x = 0.0f, y = 0.0f, z = 0.0f, t = 0.0f;
	Vector3() = default;
	Vector3(float x_, float y_, float z_, float t_ = 0.0f) : x(x_), y(y_), z(z_), t(t_) {}
};
static_assert(sizeof(Vector3) == 16, "Vector3 is 16 bytes");

// Vec2<int> (src/Types/Types.h:44-46): the pair record {x < y}.
struct int2 {
	int x, y;
};

enum class IntegratorType { Langevin, Brownian, VelocityVerlet };     // src/SimParam.h:160
enum class InterpolationOrder : uint8_t { Linear = 1, Cubic = 3 };     // src/Types/GridTerm.h:36
enum class PairlistBuilderType { ZOrder, CellList, VerletList, Hierarchical }; // Pairlist.h:40-45

// An event on the stream a call was enqueued on.
class Event {
  public:
	explicit Event(hipStream_t s = nullptr) : stream_(s) {}
	void wait() const {
		if (hipStreamSynchronize(stream_) != hipSuccess)
			throw Exception(ExceptionType::CUDARuntimeError, "hipStreamSynchronize failed");
	}

  private:
	hipStream_t stream_;
};

// Orthorhombic PeriodicBox (src/System/PeriodicBox.h:16-483) as the ABI's POD box.
class PeriodicBox {
  public:
	PeriodicBox() = default;
	PeriodicBox(const Vector3& size, const Vector3& origin = Vector3{}, bool px = true, bool py = true,
				bool pz = true) {
		box_.size[0] = size.x;
		box_.size[1] = size.y;
		box_.size[2] = size.z;
		box_.origin[0] = origin.x;
		box_.origin[1] = origin.y;
		box_.origin[2] = origin.z;
		box_.periodic[0] = px;
		box_.periodic[1] = py;
		box_.periodic[2] = pz;
	}
	const marsmd_box& c() const { return box_; }

  private:
	marsmd_box box_{};
};

using TabulatedPotential = marsmd_table;     // TabulatedPotential.h:8-13
using ExclusionView = marsmd_exclusion_view; // DeviceExclusions.h:22-75
using ParticleView = marsmd_particle_view;   // DeviceParticle.h:18-28
using ParticleTypeView = marsmd_particle_types;
using GridView = marsmd_grid_view;           // BaseGridView<float>, BaseGridDevice.h:24-30
using PairlistStatistics = marsmd_pairlist_stats;

// ------------------------------------------------------------------ ZOrderSort

class ZOrderSort {
  public:
	explicit ZOrderSort(size_t max_particles) { detail::check(marsmd_zorder_create(max_particles, &z_)); }
	~ZOrderSort() {
		if (owned_ && z_)
			marsmd_zorder_destroy(z_);
	}
	ZOrderSort(const ZOrderSort&) = delete;
	ZOrderSort& operator=(const ZOrderSort&) = delete;

	// sort_particles(positions, n, box_min, box_max): Morton keys + stable sort.
	Event sort_particles(const Vector3* positions, size_t n, const Vector3& box_min, const Vector3& box_max,
						 hipStream_t stream = nullptr) {
		const float lo[3] = {box_min.x, box_min.y, box_min.z}, hi[3] = {box_max.x, box_max.y, box_max.z};
		detail::check(marsmd_zorder_sort_particles(z_, reinterpret_cast<const float*>(positions), n, lo, hi,
												   detail::as_stream(stream)));
		return Event(stream);
	}
	const uint32_t* get_morton_codes() const { return buffers().codes; }
	const uint32_t* get_sorted_indices() const { return buffers().sorted; }
	const uint32_t* get_inverse_indices() const { return buffers().inverse; }
	// reorder_data<T>: out[i] = in[sorted_indices[i]] (T of 4, 8 or 16 bytes).
	template<class T>
	Event reorder_data(const T* in, T* out, size_t n, hipStream_t stream = nullptr) {
		static_assert(sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16, "element must be 4, 8 or 16 bytes");
		detail::check(marsmd_zorder_reorder_data(z_, in, out, n, uint32_t(sizeof(T)), detail::as_stream(stream)));
		return Event(stream);
	}
	Event remap_indices(int* indices, size_t n, hipStream_t stream = nullptr) {
		detail::check(marsmd_zorder_remap_indices(z_, indices, n, detail::as_stream(stream)));
		return Event(stream);
	}

  private:
	friend class Pairlist;
	explicit ZOrderSort(marsmd_zorder* borrowed) : z_(borrowed), owned_(false) {}
	struct Buffers {
		const uint32_t *codes, *sorted, *inverse;
	};
	Buffers buffers() const {
		Buffers b{nullptr, nullptr, nullptr};
		detail::check(marsmd_zorder_get_buffers(z_, &b.codes, &b.sorted, &b.inverse));
		return b;
	}
	marsmd_zorder* z_ = nullptr;
	bool owned_ = true;
};

// ------------------------------------------------------------------ Pairlist

class Pairlist {
  public:
	Pairlist(size_t max_particles, size_t max_pairs) {
		detail::check(marsmd_pairlist_create(max_particles, max_pairs, &pl_));
	}
	~Pairlist() {
		if (owned_ && pl_)
			marsmd_pairlist_destroy(pl_);
	}
	Pairlist(const Pairlist&) = delete;
	Pairlist& operator=(const Pairlist&) = delete;

	void set_periodic_box(const PeriodicBox& box) { detail::check(marsmd_pairlist_set_periodic_box(pl_, &box.c())); }
	void set_exclusions(const ExclusionView& view) { detail::check(marsmd_pairlist_set_exclusions(pl_, &view)); }
	void set_max_pairs_ceiling(size_t ceiling) { detail::check(marsmd_pairlist_set_max_pairs_ceiling(pl_, ceiling)); }
	// build_pairlist(positions, n, cutoff); type_ids optional (read as 0 when null).
	Event build_pairlist(const Vector3* positions, size_t n, float pairlist_cutoff, const int* type_ids = nullptr,
						 hipStream_t stream = nullptr) {
		detail::check(marsmd_pairlist_build(pl_, reinterpret_cast<const float*>(positions), type_ids, n,
											pairlist_cutoff, detail::as_stream(stream)));
		return Event(stream);
	}
	// needs_update: max displacement since the last build > skin / 2.
	bool needs_update(const Vector3* positions, size_t n, float skin, hipStream_t stream = nullptr) {
		int32_t needs = 0;
		float max_disp = 0.0f;
		detail::check(marsmd_pairlist_needs_update(pl_, reinterpret_cast<const float*>(positions), n, skin, &needs,
												   &max_disp, detail::as_stream(stream)));
		return needs != 0;
	}
	size_t get_num_pairs() const {
		uint64_t np = 0;
		detail::check(marsmd_pairlist_get_num_pairs(pl_, &np));
		return size_t(np);
	}
	// Device pointer to get_num_pairs() int2 records {x < y}.
	const int2* get_neighbor_pairs(hipStream_t stream = nullptr) {
		const int32_t* p = nullptr;
		detail::check(marsmd_pairlist_get_neighbor_pairs(pl_, &p, detail::as_stream(stream)));
		return reinterpret_cast<const int2*>(p);
	}
	PairlistStatistics get_statistics() const {
		PairlistStatistics st{};
		detail::check(marsmd_pairlist_get_statistics(pl_, &st));
		return st;
	}
	// The sorter of the last build (borrowed; valid while this pairlist lives).
	ZOrderSort& get_sorter() {
		if (!sorter_) {
			marsmd_zorder* z = nullptr;
			detail::check(marsmd_pairlist_get_sorter(pl_, &z));
			sorter_.reset(new ZOrderSort(z));
		}
		return *sorter_;
	}

  private:
	friend class Patch;
	explicit Pairlist(marsmd_pairlist* borrowed) : pl_(borrowed), owned_(false) {}
	marsmd_pairlist* pl_ = nullptr;
	bool owned_ = true;
	std::unique_ptr<ZOrderSort> sorter_;
};

// create_pairlist (Pairlist.cpp:7-31): the Z-order builder is the one implemented.
inline std::unique_ptr<Pairlist> create_pairlist(PairlistBuilderType type, size_t max_particles,
												 size_t max_pairs) {
	if (type != PairlistBuilderType::ZOrder)
		throw Exception(ExceptionType::NotImplementedError, "Only the ZOrder pairlist builder is implemented");
	return std::unique_ptr<Pairlist>(new Pairlist(max_particles, max_pairs));
}

// ------------------------------------------------------------------ launchers

inline Event launch_resolve_pair_tables(const int2* pairs, const int* type_ids, const uint32_t* term_matrix,
										size_t num_types, uint32_t* pair_tag, size_t num_pairs,
										hipStream_t stream = nullptr) {
	detail::check(marsmd_launch_resolve_pair_tables(reinterpret_cast<const int32_t*>(pairs), type_ids, term_matrix,
													num_types, pair_tag, num_pairs, detail::as_stream(stream)));
	return Event(stream);
}

// SolventParams (NonBondedInteraction.h:121-127) with the Gaussian term's constants;
// the defaults are the reference's.
struct SolventParams : marsmd_pair_params {
	SolventParams() : marsmd_pair_params{10.0f, 80.0f, 0.1f, 80.0f, 6.86f, 1.0f, 1.0f, 0.0f} {}
};

inline Event launch_pairwise_nonbonded(const int2* pairs, const ParticleView& particles, const uint32_t* pair_tag,
									   const TabulatedPotential* tables, const ParticleTypeView& types,
									   const PeriodicBox& box, bool get_energy, size_t num_pairs, float cutoff_squared,
									   uint32_t enabled_terms, const SolventParams& solvent = SolventParams(),
									   hipStream_t stream = nullptr) {
	detail::check(marsmd_launch_pairwise_nonbonded(reinterpret_cast<const int32_t*>(pairs), &particles, pair_tag,
												   tables, &types, &box.c(), get_energy ? 1 : 0, num_pairs,
												   cutoff_squared, enabled_terms, &solvent,
												   detail::as_stream(stream)));
	return Event(stream);
}

// launch_analytical_pair_nonbonded (AnalyticalPairKernels.h:110-118) over caller arrays.
inline Event launch_analytical_pair_nonbonded(const int2* pairs, const Vector3* pos, Vector3* force_energy,
											  const int* type_ids, const uint32_t* pair_tag,
											  const ParticleTypeView& types, const PeriodicBox& box,
											  uint32_t enabled_terms, const SolventParams& params, bool get_energy,
											  size_t num_pairs, float cutoff_squared, hipStream_t stream = nullptr) {
	detail::check(marsmd_launch_analytical_pair_nonbonded(
		reinterpret_cast<const int32_t*>(pairs), reinterpret_cast<const float*>(pos),
		reinterpret_cast<float*>(force_energy), type_ids, pair_tag, &types, &box.c(), enabled_terms, &params,
		get_energy ? 1 : 0, num_pairs, cutoff_squared, detail::as_stream(stream)));
	return Event(stream);
}

inline Event launch_tabulated_bonds(const int2* bonds, const Vector3* pos, Vector3* force_energy,
									const TabulatedPotential* tables, const int* table_indices, const int* forms,
									const PeriodicBox& box, bool get_energy, size_t num_bonds,
									hipStream_t stream = nullptr) {
	detail::check(marsmd_launch_tabulated_bonds(reinterpret_cast<const int32_t*>(bonds),
												reinterpret_cast<const float*>(pos),
												reinterpret_cast<float*>(force_energy), tables, table_indices, forms,
												&box.c(), get_energy ? 1 : 0, num_bonds, detail::as_stream(stream)));
	return Event(stream);
}

// angles: 16-byte int records {i, j, k, unused} (int3 of the reference is 16 bytes).
inline Event launch_tabulated_angles(const int* angles, const Vector3* pos, Vector3* force_energy,
									 const TabulatedPotential* tables, const int* table_indices, const int* forms,
									 const PeriodicBox& box, bool get_energy, size_t num_angles,
									 hipStream_t stream = nullptr) {
	detail::check(marsmd_launch_tabulated_angles(angles, reinterpret_cast<const float*>(pos),
												 reinterpret_cast<float*>(force_energy), tables, table_indices, forms,
												 &box.c(), get_energy ? 1 : 0, num_angles, detail::as_stream(stream)));
	return Event(stream);
}

inline Event launch_BD(const ParticleView& particles, const ParticleTypeView& types, float timestep,
					   size_t current_step, float kT, size_t num_particles, const PeriodicBox& box,
					   uint64_t base_seed, uint32_t base_ctr, const GridView* grids = nullptr,
					   const Vector3& electric_field = Vector3{},
					   InterpolationOrder interpolation = InterpolationOrder::Linear, hipStream_t stream = nullptr) {
	const float ef[3] = {electric_field.x, electric_field.y, electric_field.z};
	detail::check(marsmd_launch_bd(&particles, &types, timestep, current_step, kT, num_particles, &box.c(), base_seed,
								   base_ctr, grids, ef, int32_t(interpolation), detail::as_stream(stream)));
	return Event(stream);
}

inline Event launch_BAOAB(const ParticleView& particles, const ParticleTypeView& types, const PeriodicBox& box,
						  float timestep, size_t current_step, float kT, size_t num_particles, uint64_t base_seed,
						  uint32_t base_ctr, hipStream_t stream = nullptr) {
	detail::check(marsmd_launch_baoab(&particles, &types, &box.c(), timestep, current_step, kT, num_particles,
									  base_seed, base_ctr, detail::as_stream(stream)));
	return Event(stream);
}

inline Event launch_BAOAB_LastUpdate(const ParticleView& particles, const ParticleTypeView& types, float timestep,
									 size_t num_particles, hipStream_t stream = nullptr) {
	detail::check(
		marsmd_launch_baoab_last_update(&particles, &types, timestep, num_particles, detail::as_stream(stream)));
	return Event(stream);
}

inline Event launch_PMF(const ParticleView& particles, const ParticleTypeView& types, size_t num_particles,
						const GridView* grids, const Vector3& electric_field = Vector3{},
						InterpolationOrder interpolation = InterpolationOrder::Linear, hipStream_t stream = nullptr) {
	const float ef[3] = {electric_field.x, electric_field.y, electric_field.z};
	detail::check(marsmd_launch_pmf(&particles, &types, num_particles, grids, ef, int32_t(interpolation),
									detail::as_stream(stream)));
	return Event(stream);
}

// ------------------------------------------------------------------ Patch

// What the reference passes to each Patch call as NonBondedInteractions,
// BondedInteractions, DeviceParticleTypes, grid views and TablesRegistry; host data,
// copied to HBM when the Patch is made.
struct PatchSetup {
	ParticleTypeView types{};
	std::vector<uint32_t> term_matrix;        // num_types^2 pair tags (make_pair_tag)
	uint32_t enabled_terms = MARSMD_PAIR_TERM_TABULATED;
	std::vector<TabulatedPotential> nb_tables; // .pot: host knots
	std::vector<int32_t> bonds, bond_table;    // 2 indices per bond
	std::vector<int32_t> angles, angle_table;  // 3 indices per angle
	std::vector<TabulatedPotential> bond_tables, angle_tables;
	std::vector<int32_t> exclusions;           // 2 indices per exclusion
	std::vector<GridView> grids;               // .data: host values
	std::vector<int32_t> dihedrals, dihedral_table; // 4 indices per dihedral
	std::vector<TabulatedPotential> dihedral_tables; // periodic, radians
	std::vector<int32_t> restraint_ids;        // particle per harmonic restraint
	std::vector<float> restraint_anchors;      // 4 floats per restraint
	std::vector<float> restraint_k;
	SolventParams solvent;                     // analytical pair terms
};

class Patch {
  public:
	Patch(uint64_t patch_id, size_t capacity, const PatchSetup& setup, const PeriodicBox& periodic_box,
		  PairlistBuilderType pairlist_type = PairlistBuilderType::ZOrder) {
		if (pairlist_type != PairlistBuilderType::ZOrder)
			throw Exception(ExceptionType::NotImplementedError, "Only the ZOrder pairlist builder is implemented");
		marsmd_patch_desc d{};
		d.capacity = capacity;
		d.patch_id = patch_id;
		d.box = periodic_box.c();
		d.types = setup.types;
		d.term_matrix = setup.term_matrix.empty() ? nullptr : setup.term_matrix.data();
		d.enabled_terms = setup.enabled_terms;
		d.num_nb_tables = int32_t(setup.nb_tables.size());
		d.nb_tables = setup.nb_tables.data();
		d.num_bonds = setup.bonds.size() / 2;
		d.bonds = setup.bonds.data();
		d.bond_table = setup.bond_table.empty() ? nullptr : setup.bond_table.data();
		d.num_angles = setup.angles.size() / 3;
		d.angles = setup.angles.data();
		d.angle_table = setup.angle_table.empty() ? nullptr : setup.angle_table.data();
		d.num_bond_tables = int32_t(setup.bond_tables.size());
		d.bond_tables = setup.bond_tables.data();
		d.num_angle_tables = int32_t(setup.angle_tables.size());
		d.angle_tables = setup.angle_tables.data();
		d.num_exclusions = setup.exclusions.size() / 2;
		d.exclusions = setup.exclusions.data();
		d.num_grids = int32_t(setup.grids.size());
		d.grids = setup.grids.data();
		d.num_dihedrals = setup.dihedrals.size() / 4;
		d.dihedrals = setup.dihedrals.data();
		d.dihedral_table = setup.dihedral_table.empty() ? nullptr : setup.dihedral_table.data();
		d.num_dihedral_tables = int32_t(setup.dihedral_tables.size());
		d.dihedral_tables = setup.dihedral_tables.data();
		d.num_restraints = setup.restraint_ids.size();
		d.restraint_ids = setup.restraint_ids.data();
		d.restraint_anchors = setup.restraint_anchors.data();
		d.restraint_k = setup.restraint_k.data();
		d.pair_params = &setup.solvent;
		detail::check(marsmd_patch_create(&d, &p_));
	}
	~Patch() {
		if (p_)
			marsmd_patch_destroy(p_);
	}
	Patch(const Patch&) = delete;
	Patch& operator=(const Patch&) = delete;

	void set_base_seed(uint64_t seed) { detail::check(marsmd_patch_set_base_seed(p_, seed)); }
	// copy_particles_from_host: host arrays in the caller's order; mom may be null.
	void copy_particles_from_host(const Vector3* pos, const Vector3* mom, const int* type_ids, size_t n) {
		detail::check(marsmd_patch_copy_particles_from_host(p_, reinterpret_cast<const float*>(pos),
															reinterpret_cast<const float*>(mom), type_ids, n));
	}
	// copy_particles_to_host: any output may be null; ids are the particles' global ids.
	void copy_particles_to_host(Vector3* pos, Vector3* mom, Vector3* force_energy, int* type_ids, int* ids,
								size_t n) {
		detail::check(marsmd_patch_copy_particles_to_host(p_, reinterpret_cast<float*>(pos),
														  reinterpret_cast<float*>(mom),
														  reinterpret_cast<float*>(force_energy), type_ids, ids, n));
	}
	Event calculate_nonbonded_forces(float pairlist_cutoff, float interaction_cutoff, size_t step,
									 size_t rebuild_period, const Vector3& electric_field = Vector3{},
									 InterpolationOrder interpolation_scheme = InterpolationOrder::Linear,
									 bool compute_energy = false, hipStream_t stream = nullptr) {
		const float ef[3] = {electric_field.x, electric_field.y, electric_field.z};
		detail::check(marsmd_patch_calculate_nonbonded_forces(p_, pairlist_cutoff, interaction_cutoff, step,
															  rebuild_period, ef, int32_t(interpolation_scheme),
															  compute_energy ? 1 : 0, detail::as_stream(stream)));
		return Event(stream);
	}
	Event calculate_bonded_forces(bool compute_energy = false, hipStream_t stream = nullptr) {
		detail::check(marsmd_patch_calculate_bonded_forces(p_, compute_energy ? 1 : 0, detail::as_stream(stream)));
		return Event(stream);
	}
	// integrate_motion(dt, temperature (kT, kcal/mol), integrator, step).
	Event integrate_motion(float dt, float kT, IntegratorType integrator_type, size_t step = 0,
						   hipStream_t stream = nullptr) {
		int32_t integ = -1;
		if (integrator_type == IntegratorType::Brownian)
			integ = MARSMD_INTEGRATOR_BROWNIAN;
		else if (integrator_type == IntegratorType::Langevin)
			integ = MARSMD_INTEGRATOR_LANGEVIN;
		else
			throw Exception(ExceptionType::ValueError, "Unsupported integrator type");
		detail::check(marsmd_patch_integrate_motion(p_, dt, kT, integ, step, detail::as_stream(stream)));
		return Event(stream);
	}
	Event finish_deferred_kick(float dt, hipStream_t stream = nullptr) {
		detail::check(marsmd_patch_finish_deferred_kick(p_, dt, detail::as_stream(stream)));
		return Event(stream);
	}
	Event reorder_particles(hipStream_t stream = nullptr) {
		detail::check(marsmd_patch_reorder_particles(p_, detail::as_stream(stream)));
		return Event(stream);
	}
	// The per-step loop (reorder cadence, rebuild policy, forces, integration).
	Event run(const marsmd_run_params& params, size_t first_step, size_t num_steps, hipStream_t stream = nullptr) {
		detail::check(marsmd_patch_run(p_, &params, first_step, num_steps, detail::as_stream(stream)));
		return Event(stream);
	}
	Pairlist& get_pairlist() {
		if (!pairlist_) {
			marsmd_pairlist* pl = nullptr;
			detail::check(marsmd_patch_get_pairlist(p_, &pl));
			pairlist_.reset(new Pairlist(pl));
		}
		return *pairlist_;
	}

  private:
	marsmd_patch* p_ = nullptr;
	std::unique_ptr<Pairlist> pairlist_;
};

} // namespace MARS
