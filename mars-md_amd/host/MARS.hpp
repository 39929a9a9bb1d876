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
// The Patch keeps the reference's constructor (Patch.h:57-69) and per-step signatures
// (calculate_nonbonded_forces / calculate_bonded_forces / integrate_motion, Patch.h:
// 222-260, 284-287), taking the reference's containers as the host-side mirror types
// below (Resource, Table / TablesRegistry, PairNonBonded / NonBondedInteractions,
// Bond / Angle / Dihedral / Exclude / Restraint / BondedInteractions, ParticleType /
// DeviceParticleTypes, BaseGridView / DeviceBuffer, Temperature, HostParticleData):
// what the Patch reads from them is uploaded to HBM at the first force call and kept
// (the HIP library then owns it, reordering the topology with the particles as the
// reference's device copies are). A Patch may instead be given everything at
// construction (PatchSetup) and called with the shorter signatures.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
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
using mars_real = float;  // src/Header.h (single precision build)
using idx_t = size_t;     // src/Header.h:258
using patch_t = size_t;   // src/Header.h:259

namespace constants {
constexpr float BOLTZMANN = 0.001987191f;        // src/Constants.h:25
constexpr float SQRT_CAL_TO_JOULE = 2.046167337f; // src/Constants.h:34
constexpr double PI = 3.14159265358979323846;
} // namespace constants

// Resource (src/Backend/Resource.h:79): the GPU a patch runs on (its HIP device id).
class Resource {
  public:
	explicit Resource(int id = 0) : id_(id) {}
	int id() const { return id_; }
	bool operator==(const Resource& o) const { return id_ == o.id_; }

  private:
	int id_;
};

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
	// update_positions_incremental (ZOrderSort.h:172-184): the displacement reference.
	void update_positions_incremental(const Vector3* positions, size_t n, hipStream_t stream = nullptr) {
		detail::check(marsmd_zorder_update_positions_incremental(z_, reinterpret_cast<const float*>(positions), n,
																 detail::as_stream(stream)));
	}
	// compute_max_displacement (ZOrderSort.h:160-170): box_len <= 0 marks an open axis.
	float compute_max_displacement(const Vector3* positions, size_t n, const Vector3& box_len = Vector3{},
								   hipStream_t stream = nullptr) const {
		const float bl[3] = {box_len.x, box_len.y, box_len.z};
		float d = 0.0f;
		detail::check(marsmd_zorder_compute_max_displacement(z_, reinterpret_cast<const float*>(positions), n, bl, &d,
															 detail::as_stream(stream)));
		return d;
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

// launch_tabulated_dihedrals (DeviceBondedInteraction.h:551-572): 16-byte int records
// {i, j, k, l}.
inline Event launch_tabulated_dihedrals(const int* dihedrals, const Vector3* pos, Vector3* force_energy,
										const TabulatedPotential* tables, const int* table_indices, const int* forms,
										const PeriodicBox& box, bool get_energy, size_t num_dihedrals,
										hipStream_t stream = nullptr) {
	detail::check(marsmd_launch_tabulated_dihedrals(dihedrals, reinterpret_cast<const float*>(pos),
													reinterpret_cast<float*>(force_energy), tables, table_indices,
													forms, &box.c(), get_energy ? 1 : 0, num_dihedrals,
													detail::as_stream(stream)));
	return Event(stream);
}

// launch_harmonic_restraints (DeviceBondedInteraction.h:599-620): particle ids below 0
// are skipped; anchors are Vector3 records.
inline Event launch_harmonic_restraints(const int* particle_ids, const Vector3* pos, Vector3* force_energy,
										const Vector3* anchors, const float* spring_constants, const PeriodicBox& box,
										bool get_energy, size_t num_restraints, hipStream_t stream = nullptr) {
	detail::check(marsmd_launch_harmonic_restraints(particle_ids, reinterpret_cast<const float*>(pos),
													reinterpret_cast<float*>(force_energy),
													reinterpret_cast<const float*>(anchors), spring_constants,
													&box.c(), get_energy ? 1 : 0, num_restraints,
													detail::as_stream(stream)));
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

// ------------------------------------------------------------------ the reference's containers
// Host-side mirrors of what SimManager hands the Patch each step. They carry the
// fields the Patch reads, under the reference's names.

enum class TabulatedType { NonBondedPair, Bond, Angle, Dihedral, Default }; // src/Objects/Tables.h:21
enum class InteractionForm { Grid, Tabulated, Analytical };                  // src/SimParam.h:174

// Table (src/Objects/Tables.h:25-125): abscissa X (uniform) and values Y.
struct Table {
	TabulatedType type;
	std::string name;
	mars_real step_size{0.0f};
	mars_real start{0.0f};
	std::vector<mars_real> X;
	std::vector<mars_real> Y;
	explicit Table(TabulatedType t = TabulatedType::Default) : type(t) {
		if (type == TabulatedType::Dihedral)
			start = -mars_real(constants::PI);
	}
	// step and start from X (Tables.h:40-56)
	bool check_same_step_size() {
		if (X.size() < 2)
			return false;
		step_size = X[1] - X[0];
		start = X[0];
		for (size_t i = 2; i < X.size(); ++i)
			if (std::abs(X[i] - X[i - 1]) - step_size > 1e-4f)
				return false;
		return true;
	}
	// degrees -> radians for angle and dihedral abscissae (Tables.h:109-119)
	void convert_angular_abscissa_to_radians() {
		if (type != TabulatedType::Angle && type != TabulatedType::Dihedral)
			return;
		constexpr mars_real deg_to_rad = static_cast<mars_real>(constants::PI / 180.0);
		for (auto& x : X)
			x *= deg_to_rad;
		start *= deg_to_rad;
		step_size *= deg_to_rad;
	}
	// X and Y given, as read_file leaves a table (Tables.h:59-100)
	void set_xy(std::vector<mars_real> x, std::vector<mars_real> y) {
		X = std::move(x);
		Y = std::move(y);
		check_same_step_size();
		convert_angular_abscissa_to_radians();
	}
};

// TablesRegistry (src/Objects/Tables.h:155-330): tables by kind; add_* returns the
// index a PairNonBonded / Bond / Angle / Dihedral carries as function_index.
class TablesRegistry {
  public:
	int add_nonbonded(Table t) { return push(nonbonded_, std::move(t)); }
	int add_bond(Table t) { return push(bond_, std::move(t)); }
	int add_angle(Table t) { return push(angle_, std::move(t)); }
	int add_dihedral(Table t) { return push(dihedral_, std::move(t)); }
	const std::vector<Table>& get_nonbonded() const { return nonbonded_; }
	const std::vector<Table>& get_bond() const { return bond_; }
	const std::vector<Table>& get_angle() const { return angle_; }
	const std::vector<Table>& get_dihedral() const { return dihedral_; }

  private:
	static int push(std::vector<Table>& v, Table t) {
		v.push_back(std::move(t));
		return int(v.size()) - 1;
	}
	std::vector<Table> nonbonded_, bond_, angle_, dihedral_;
};

// pair_term_from_name (src/Interactions/NonBondedInteraction.h:16-34)
inline uint32_t pair_term_from_name(const std::string& f) {
	if (f == "coulomb" || f == "columb")
		return MARSMD_PAIR_TERM_COULOMB;
	if (f == "debye_huckel")
		return MARSMD_PAIR_TERM_DEBYE_HUCKEL;
	if (f == "onck")
		return MARSMD_PAIR_TERM_ONCK;
	if (f == "gaussian")
		return MARSMD_PAIR_TERM_GAUSSIAN;
	if (f == "softcore")
		return MARSMD_PAIR_TERM_SOFTCORE;
	return MARSMD_PAIR_TERM_NONE;
}

// PairNonBonded (NonBondedInteraction.h:38-110): one type pair, tabulated by index into
// TablesRegistry::get_nonbonded() or analytical by name.
struct PairNonBonded {
	int type_id_1{-1};
	int type_id_2{-1};
	std::string function_name{};
	InteractionForm form{InteractionForm::Tabulated};
	int function_index{-1};
	PairNonBonded(int a, int b, const std::string& fn) : PairNonBonded(a, b, fn, -1) {}
	PairNonBonded(int a, int b, const std::string& fn, int table_index)
		: type_id_1(std::min(a, b)), type_id_2(std::max(a, b)), function_name(fn) {
		form = table_index >= 0 ? InteractionForm::Tabulated : InteractionForm::Analytical;
		function_index = table_index;
	}
	bool is_analytical() const { return form == InteractionForm::Analytical; }
	uint32_t term_bits() const { return is_analytical() ? pair_term_from_name(function_name) : MARSMD_PAIR_TERM_TABULATED; }
};

class NonBondedInteractions {
  public:
	NonBondedInteractions() = default;
	explicit NonBondedInteractions(std::vector<PairNonBonded> pairs) : pairs_(std::move(pairs)) {}
	void set_solvent_params(const SolventParams& p) { solvent_ = p; }
	const SolventParams& get_solvent_params() const { return solvent_; }
	// a later entry for the same type pair replaces the earlier one
	void add_pair_nonbonded(const PairNonBonded& p) {
		for (auto& e : pairs_)
			if (e.type_id_1 == p.type_id_1 && e.type_id_2 == p.type_id_2) {
				e = p;
				return;
			}
		pairs_.push_back(p);
	}
	const std::vector<PairNonBonded>& get_pair_nonbonded() const { return pairs_; }
	size_t get_num_pair_nonbonded() const { return pairs_.size(); }

  private:
	std::vector<PairNonBonded> pairs_;
	SolventParams solvent_{};
};

// Host-side bonded terms (src/Interactions/BondedInteraction.h:78-173); indices are
// the particles' global ids. Only tabulated forms reach the Patch's kernels, as in
// the reference's Tabulated*Computer.
struct Exclude {
	int ind1{-1}, ind2{-1};
	Exclude() = default;
	Exclude(int a, int b) : ind1(a), ind2(b) {}
};
struct Bond {
	int ind1{-1}, ind2{-1};
	std::string function_name{};
	InteractionForm form{InteractionForm::Tabulated};
	int function_index{-1};
	Bond() = default;
	Bond(int a, int b, int table_index) : ind1(a), ind2(b), function_index(table_index) {}
};
struct Angle {
	int ind1{-1}, ind2{-1}, ind3{-1};
	std::string function_name{};
	InteractionForm form{InteractionForm::Tabulated};
	int function_index{-1};
};
struct Dihedral {
	int ind1{-1}, ind2{-1}, ind3{-1}, ind4{-1};
	std::string function_name{};
	InteractionForm form{InteractionForm::Tabulated};
	int function_index{-1};
};
struct Restraint {
	int ind{-1};
	Vector3 r0{};
	float k{0.0f};
	Restraint() = default;
	Restraint(int id, Vector3 r, float kk) : ind(id), r0(r), k(kk) {}
};

class BondedInteractions {
  public:
	BondedInteractions() = default;
	BondedInteractions(std::vector<Bond> b, std::vector<Angle> a, std::vector<Dihedral> d, std::vector<Exclude> e,
					   std::vector<Restraint> r)
		: bonds_(std::move(b)), angles_(std::move(a)), dihedrals_(std::move(d)), exclusions_(std::move(e)),
		  restraints_(std::move(r)) {}
	void add_bond(const Bond& b) { bonds_.push_back(b); }
	void add_angle(const Angle& a) { angles_.push_back(a); }
	void add_dihedral(const Dihedral& d) { dihedrals_.push_back(d); }
	void add_exclude(const Exclude& e) { exclusions_.push_back(e); }
	void add_restraint(const Restraint& r) { restraints_.push_back(r); }
	size_t get_num_bonds() const { return bonds_.size(); }
	size_t get_num_angles() const { return angles_.size(); }
	size_t get_num_dihedrals() const { return dihedrals_.size(); }
	const std::vector<Bond>& get_bonds() const { return bonds_; }
	const std::vector<Angle>& get_angles() const { return angles_; }
	const std::vector<Dihedral>& get_dihedrals() const { return dihedrals_; }
	const std::vector<Exclude>& get_exclusions() const { return exclusions_; }
	const std::vector<Restraint>& get_restraints() const { return restraints_; }

  private:
	std::vector<Bond> bonds_;
	std::vector<Angle> angles_;
	std::vector<Dihedral> dihedrals_;
	std::vector<Exclude> exclusions_;
	std::vector<Restraint> restraints_;
};

// GridTerm (src/Types/GridTerm.h:55-69) and ParticleType (Objects/ParticleProperties.h:
// 27-60): the fields the per-step path reads.
struct GridTerm {
	int grid_id = -1;
	float scale = 1.0f;
	float scale_slope = 0.0f;
	int boundary_condition = -1;
};
struct ParticleType {
	std::string name;
	float mass = 1.0f;
	float charge = 0.0f;
	float radius = 0.0f;
	float eps = 0.0f;
	Vector3 diffusion{};
	Vector3 trans_damping{};
	std::vector<GridTerm> pmf_grids;
	ParticleType() = default;
	explicit ParticleType(std::string n) : name(std::move(n)) {}
};

// DeviceParticleTypes (src/Objects/DeviceParticleManager.h:284-360): the type table the
// kernels read (held by value here: it travels as a kernel argument).
class DeviceParticleTypes {
  public:
	DeviceParticleTypes(const std::vector<ParticleType>& types, const Resource& res) : resource_(res) {
		if (types.empty() || types.size() > MARSMD_MAX_TYPES)
			throw Exception(ExceptionType::ValueError, "DeviceParticleTypes: 1.." +
														   std::to_string(MARSMD_MAX_TYPES) + " types");
		view_.num_types = int32_t(types.size());
		int terms = 0;
		for (size_t t = 0; t < types.size(); ++t) {
			const ParticleType& p = types[t];
			view_.mass[t] = p.mass;
			view_.charge[t] = p.charge;
			view_.radius[t] = p.radius;
			view_.eps[t] = p.eps;
			view_.diffusion[t][0] = p.diffusion.x;
			view_.diffusion[t][1] = p.diffusion.y;
			view_.diffusion[t][2] = p.diffusion.z;
			view_.trans_damping[t][0] = p.trans_damping.x;
			view_.trans_damping[t][1] = p.trans_damping.y;
			view_.trans_damping[t][2] = p.trans_damping.z;
			view_.pmf_grid_offset[t] = terms;
			view_.pmf_grid_count[t] = int32_t(p.pmf_grids.size());
			for (const GridTerm& g : p.pmf_grids) {
				if (terms >= MARSMD_MAX_PMF_TERMS)
					throw Exception(ExceptionType::ValueError, "DeviceParticleTypes: too many PMF terms");
				view_.pmf_grid_terms[terms++] = marsmd_grid_term{g.grid_id, g.scale, g.scale_slope,
																  g.boundary_condition};
			}
			for (int a = 0; a < 3; ++a) {
				view_.force_grid_id[t][a] = -1;
				view_.force_grid_scale[t][a] = 0.0f;
			}
		}
	}
	const ParticleTypeView& view() const { return view_; }
	const Resource& get_resource() const { return resource_; }
	size_t size() const { return size_t(view_.num_types); }

  private:
	ParticleTypeView view_{};
	Resource resource_;
};

// BaseGridView<T> (src/Types/BaseGridDevice.h:24-30) with host values in .data, and a
// DeviceBuffer<T> standing for the per-resource array GridManager hands the Patch.
template<class T>
struct BaseGridView;
template<>
struct BaseGridView<float> : GridView {};
template<class T>
class DeviceBuffer {
  public:
	DeviceBuffer() = default;
	explicit DeviceBuffer(std::vector<T> host) : host_(std::move(host)) {}
	size_t size() const { return host_.size(); }
	bool empty() const { return host_.empty(); }
	const T* data() const { return host_.data(); }
	void copy_from_host(const std::vector<T>& v) { host_ = v; }

  private:
	std::vector<T> host_;
};

// Temperature (src/SimParam.h:25-45), scalar form: kT = T * BOLTZMANN.
struct Temperature {
	float value;
	float kT;
	Temperature(float v = 298.15f) : value(v), kT(v * constants::BOLTZMANN) {}
};

// HostParticleData (src/Objects/ParticleProperties.h:129-175): the fields the path uses.
struct HostParticleData {
	std::vector<int> global_id;
	std::vector<int> type_id;
	std::vector<Vector3> pos;
	std::vector<Vector3> mom;
	std::vector<Vector3> force;
	std::vector<float> energy;
	size_t size() const { return global_id.size(); }
	void resize(size_t n) {
		global_id.resize(n);
		type_id.resize(n);
		pos.resize(n);
		mom.resize(n);
		force.resize(n);
		energy.resize(n);
	}
};

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
	// The reference's constructor (Patch.h:57-69). What the kernels read (types, tables,
	// topology, grids) arrives with the first calculate_nonbonded_forces call; particles
	// copied in before it wait on the host until then.
	Patch(patch_t patch_id, idx_t capacity, const Resource& resource, const PeriodicBox& periodic_box,
		  PairlistBuilderType pairlist_type = PairlistBuilderType::ZOrder)
		: id_(patch_id), capacity_(capacity), resource_(resource), box_(periodic_box) {
		check_builder(pairlist_type);
	}
	// Everything at construction; the shorter per-step signatures below then apply.
	Patch(uint64_t patch_id, size_t capacity, const PatchSetup& setup, const PeriodicBox& periodic_box,
		  PairlistBuilderType pairlist_type = PairlistBuilderType::ZOrder)
		: id_(patch_id), capacity_(capacity), box_(periodic_box) {
		check_builder(pairlist_type);
		create(setup);
	}
	~Patch() {
		if (p_)
			marsmd_patch_destroy(p_);
	}
	Patch(const Patch&) = delete;
	Patch& operator=(const Patch&) = delete;

	patch_t get_patch_id() const { return id_; }
	idx_t get_capacity() const { return capacity_; }
	idx_t get_particle_count() const { return n_; }
	const Resource& get_resource() const { return resource_; }
	const PeriodicBox& get_periodic_box() const { return box_; }
	void set_periodic_box(const PeriodicBox& box) {
		if (p_)
			throw Exception(ExceptionType::NotImplementedError, "Patch: the box is fixed after the first force call");
		box_ = box;
	}
	// Patch::set_base_seed (Patch.h:122-130)
	void set_base_seed(uint64_t seed) {
		seed_ = seed;
		if (p_)
			detail::check(marsmd_patch_set_base_seed(p_, seed));
	}

	// copy_particles_from_host (Patch.cpp:507-554): host entries [start_idx, start_idx +
	// count) become the patch's particles, in that order.
	void copy_particles_from_host(const HostParticleData& h, idx_t start_idx, idx_t count) {
		if (start_idx + count > h.size())
			throw Exception(ExceptionType::ValueError, "Host data range exceeds available data: start=" +
														   std::to_string(start_idx) + ", count=" +
														   std::to_string(count) + ", available=" +
														   std::to_string(h.size()));
		if (count > capacity_)
			throw Exception(ExceptionType::RuntimeError, "Cannot copy " + std::to_string(count) +
															 " particles to patch with capacity " +
															 std::to_string(capacity_));
		staged_pos_.assign(h.pos.begin() + long(start_idx), h.pos.begin() + long(start_idx + count));
		staged_type_.assign(h.type_id.begin() + long(start_idx), h.type_id.begin() + long(start_idx + count));
		staged_mom_.clear();
		if (h.mom.size() >= start_idx + count)
			staged_mom_.assign(h.mom.begin() + long(start_idx), h.mom.begin() + long(start_idx + count));
		global_id_.assign(h.global_id.begin() + long(start_idx), h.global_id.begin() + long(start_idx + count));
		n_ = count;
		if (p_)
			upload_staged();
	}
	// copy_particles_from_host over plain arrays (mom may be null); global ids 0..n-1.
	void copy_particles_from_host(const Vector3* pos, const Vector3* mom, const int* type_ids, size_t n) {
		staged_pos_.assign(pos, pos + n);
		staged_type_.assign(type_ids, type_ids + n);
		staged_mom_.clear();
		if (mom)
			staged_mom_.assign(mom, mom + n);
		global_id_.resize(n);
		for (size_t i = 0; i < n; ++i)
			global_id_[i] = int(i);
		n_ = n;
		if (p_)
			upload_staged();
	}
	// copy_particles_to_host (Patch.cpp:556-600): the patch's particles in its current
	// order with their global ids into host entries [start_idx, start_idx + count).
	void copy_particles_to_host(HostParticleData& h, idx_t start_idx, idx_t count, bool need_energy = false) const {
		if (count > n_)
			throw Exception(ExceptionType::RuntimeError, "Cannot copy " + std::to_string(count) +
															 " particles from patch containing only " +
															 std::to_string(n_));
		if (h.size() < start_idx + count)
			h.resize(start_idx + count);
		std::vector<Vector3> pos(n_), mom(n_), fe(n_);
		std::vector<int> type(n_), ids(n_);
		if (p_) {
			detail::check(marsmd_patch_copy_particles_to_host(p_, reinterpret_cast<float*>(pos.data()),
															  reinterpret_cast<float*>(mom.data()),
															  reinterpret_cast<float*>(fe.data()), type.data(),
															  ids.data(), n_));
		} else {
			for (size_t i = 0; i < n_; ++i) {
				pos[i] = staged_pos_[i];
				mom[i] = staged_mom_.empty() ? Vector3{} : staged_mom_[i];
				type[i] = staged_type_[i];
				ids[i] = int(i);
			}
		}
		for (idx_t i = 0; i < count; ++i) {
			const size_t k = start_idx + i;
			h.global_id[k] = global_id_[size_t(ids[i])];
			h.type_id[k] = type[i];
			h.pos[k] = Vector3{pos[i].x, pos[i].y, pos[i].z};
			h.mom[k] = Vector3{mom[i].x, mom[i].y, mom[i].z};
			if (need_energy) {
				h.force[k] = Vector3{fe[i].x, fe[i].y, fe[i].z};
				h.energy[k] = fe[i].t;
			}
		}
	}
	// over plain arrays: any output may be null; ids are the particles' global ids.
	void copy_particles_to_host(Vector3* pos, Vector3* mom, Vector3* force_energy, int* type_ids, int* ids,
								size_t n) {
		need_created("copy_particles_to_host");
		detail::check(marsmd_patch_copy_particles_to_host(p_, reinterpret_cast<float*>(pos),
														  reinterpret_cast<float*>(mom),
														  reinterpret_cast<float*>(force_energy), type_ids, ids, n));
		if (ids)
			for (size_t i = 0; i < n; ++i)
				ids[i] = global_id_.empty() ? ids[i] : global_id_[size_t(ids[i])];
	}

	// Patch::calculate_nonbonded_forces as SimManager calls it (Patch.h:222-236,
	// SimManager.cpp:403-416). resource_idx selects nothing here (one resource per
	// library instance); the containers must stay the same objects from call to call.
	Event calculate_nonbonded_forces(const NonBondedInteractions& interactions,
									 const BondedInteractions& bonded_interactions,
									 const DeviceParticleTypes& particle_types,
									 const DeviceBuffer<BaseGridView<mars_real>>& grid_views,
									 const TablesRegistry& tables_registry, size_t resource_idx, float pairlist_cutoff,
									 float interaction_cutoff, size_t step, size_t rebuild_period,
									 const Vector3& electric_field = Vector3{0.0f, 0.0f, 0.0f},
									 InterpolationOrder interpolation_scheme = InterpolationOrder::Linear,
									 bool compute_energy = false, hipStream_t stream = nullptr) {
		(void)resource_idx;
		prepare(&interactions, &bonded_interactions, &particle_types, &grid_views, &tables_registry);
		return calculate_nonbonded_forces(pairlist_cutoff, interaction_cutoff, step, rebuild_period, electric_field,
										  interpolation_scheme, compute_energy, stream);
	}
	Event calculate_nonbonded_forces(float pairlist_cutoff, float interaction_cutoff, size_t step,
									 size_t rebuild_period, const Vector3& electric_field = Vector3{},
									 InterpolationOrder interpolation_scheme = InterpolationOrder::Linear,
									 bool compute_energy = false, hipStream_t stream = nullptr) {
		need_created("calculate_nonbonded_forces");
		const float ef[3] = {electric_field.x, electric_field.y, electric_field.z};
		detail::check(marsmd_patch_calculate_nonbonded_forces(p_, pairlist_cutoff, interaction_cutoff, step,
															  rebuild_period, ef, int32_t(interpolation_scheme),
															  compute_energy ? 1 : 0, detail::as_stream(stream)));
		bonded_prepared_ = true;
		return Event(stream);
	}
	// Patch::calculate_bonded_forces as SimManager calls it (Patch.h:256-260,
	// SimManager.cpp:418-422).
	Event calculate_bonded_forces(const BondedInteractions& interactions, const DeviceParticleTypes& particle_types,
								  const TablesRegistry& tables_registry, size_t resource_idx,
								  bool compute_energy = false, hipStream_t stream = nullptr) {
		(void)resource_idx;
		if (!p_)
			throw Exception(ExceptionType::RuntimeError,
							"Patch: calculate_nonbonded_forces must run before calculate_bonded_forces");
		if (!from_setup_ &&
			(&interactions != src_.bonded || &particle_types != src_.types || &tables_registry != src_.tables))
			throw Exception(ExceptionType::ValueError,
							"Patch: bonded interactions, types and tables must be the objects of the first force call");
		return calculate_bonded_forces(compute_energy, stream);
	}
	Event calculate_bonded_forces(bool compute_energy = false, hipStream_t stream = nullptr) {
		need_created("calculate_bonded_forces");
		detail::check(marsmd_patch_calculate_bonded_forces(p_, compute_energy ? 1 : 0, detail::as_stream(stream)));
		return Event(stream);
	}
	// Patch::integrate_motion (Patch.h:284-287): the scalar temperature's kT.
	Event integrate_motion(float dt, const Temperature& temperature, IntegratorType integrator_type, size_t step = 0,
						   hipStream_t stream = nullptr) {
		return integrate_motion(dt, temperature.kT, integrator_type, step, stream);
	}
	// integrate_motion(dt, kT (kcal/mol), integrator, step).
	Event integrate_motion(float dt, float kT, IntegratorType integrator_type, size_t step = 0,
						   hipStream_t stream = nullptr) {
		need_created("integrate_motion");
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
		if (!p_)
			return Event(stream);
		detail::check(marsmd_patch_finish_deferred_kick(p_, dt, detail::as_stream(stream)));
		return Event(stream);
	}
	// Patch::has_deferred_kick (Patch.h:300-304)
	bool has_deferred_kick() const {
		if (!p_)
			return false;
		int32_t pending = 0;
		detail::check(marsmd_patch_has_deferred_kick(p_, &pending));
		return pending != 0;
	}
	// Patch::is_bonded_prepared: the first force call has run (ReorderManager.cpp:19-22).
	bool is_bonded_prepared() const { return bonded_prepared_; }
	void reorder_particles(hipStream_t stream = nullptr) {
		need_created("reorder_particles");
		detail::check(marsmd_patch_reorder_particles(p_, detail::as_stream(stream)));
	}
	// The per-step loop (reorder cadence, rebuild policy, forces, integration).
	Event run(const marsmd_run_params& params, size_t first_step, size_t num_steps, hipStream_t stream = nullptr) {
		need_created("run");
		detail::check(marsmd_patch_run(p_, &params, first_step, num_steps, detail::as_stream(stream)));
		bonded_prepared_ = bonded_prepared_ || num_steps > 0;
		return Event(stream);
	}
	Pairlist& get_pairlist() {
		need_created("get_pairlist");
		if (!pairlist_) {
			marsmd_pairlist* pl = nullptr;
			detail::check(marsmd_patch_get_pairlist(p_, &pl));
			pairlist_.reset(new Pairlist(pl));
		}
		return *pairlist_;
	}

  private:
	struct Sources {
		const NonBondedInteractions* nb = nullptr;
		const BondedInteractions* bonded = nullptr;
		const DeviceParticleTypes* types = nullptr;
		const DeviceBuffer<BaseGridView<mars_real>>* grids = nullptr;
		const TablesRegistry* tables = nullptr;
	};

	static void check_builder(PairlistBuilderType t) {
		if (t != PairlistBuilderType::ZOrder)
			throw Exception(ExceptionType::NotImplementedError, "Only the ZOrder pairlist builder is implemented");
	}
	void need_created(const char* what) const {
		if (!p_)
			throw Exception(ExceptionType::RuntimeError,
							std::string("Patch::") + what + ": no interactions yet (calculate_nonbonded_forces with "
															 "the system's containers comes first)");
	}
	static TabulatedPotential table_view(const Table& t, bool nonbonded, bool periodic) {
		TabulatedPotential v{};
		v.pot = t.Y.data();
		v.size = uint32_t(t.Y.size());
		v.start = t.start;
		// DevicePairNonBondedInteraction.h:183 (float reciprocal); DeviceBondedInteraction.h:
		// 213, 230, 246 (double reciprocal, 0 for a zero step)
		if (nonbonded)
			v.step_inv = 1.0f / static_cast<mars_real>(t.step_size);
		else
			v.step_inv = t.step_size > 0.0f ? static_cast<mars_real>(1.0 / t.step_size) : mars_real(0);
		v.is_periodic = periodic ? 1 : 0;
		return v;
	}
	// The Patch's view of the reference's containers: the type-pair tag matrix and enabled
	// term mask from the pair list, the tables by kind, tabulated bonded terms only.
	static PatchSetup setup_from(const Sources& src) {
		PatchSetup s;
		s.types = src.types->view();
		const int nt = s.types.num_types;
		const auto& nbt = src.tables->get_nonbonded();
		s.term_matrix.assign(size_t(nt) * size_t(nt), 0u);
		s.enabled_terms = 0u;
		for (const PairNonBonded& p : src.nb->get_pair_nonbonded()) {
			if (p.type_id_1 < 0 || p.type_id_2 < 0 || p.type_id_1 >= nt || p.type_id_2 >= nt)
				throw Exception(ExceptionType::ValueError, "PairNonBonded: type id out of range");
			uint32_t tag = p.term_bits();
			if (!p.is_analytical()) {
				if (p.function_index < 0 || size_t(p.function_index) >= nbt.size())
					throw Exception(ExceptionType::ValueError, "PairNonBonded: table index out of range");
				tag |= uint32_t(p.function_index) << MARSMD_PAIR_TABLE_SHIFT;
			}
			s.term_matrix[size_t(p.type_id_1) * size_t(nt) + size_t(p.type_id_2)] = tag;
			s.term_matrix[size_t(p.type_id_2) * size_t(nt) + size_t(p.type_id_1)] = tag;
			s.enabled_terms |= p.term_bits();
		}
		for (const Table& t : nbt)
			s.nb_tables.push_back(table_view(t, true, false));
		for (const Table& t : src.tables->get_bond())
			s.bond_tables.push_back(table_view(t, false, false));
		for (const Table& t : src.tables->get_angle())
			s.angle_tables.push_back(table_view(t, false, true));
		for (const Table& t : src.tables->get_dihedral())
			s.dihedral_tables.push_back(table_view(t, false, true));
		const BondedInteractions& b = *src.bonded;
		for (const Bond& x : b.get_bonds())
			if (x.form == InteractionForm::Tabulated) {
				s.bonds.insert(s.bonds.end(), {x.ind1, x.ind2});
				s.bond_table.push_back(x.function_index);
			}
		for (const Angle& x : b.get_angles())
			if (x.form == InteractionForm::Tabulated) {
				s.angles.insert(s.angles.end(), {x.ind1, x.ind2, x.ind3});
				s.angle_table.push_back(x.function_index);
			}
		for (const Dihedral& x : b.get_dihedrals())
			if (x.form == InteractionForm::Tabulated) {
				s.dihedrals.insert(s.dihedrals.end(), {x.ind1, x.ind2, x.ind3, x.ind4});
				s.dihedral_table.push_back(x.function_index);
			}
		for (const Exclude& x : b.get_exclusions())
			s.exclusions.insert(s.exclusions.end(), {x.ind1, x.ind2});
		for (const Restraint& x : b.get_restraints()) {
			s.restraint_ids.push_back(x.ind);
			s.restraint_anchors.insert(s.restraint_anchors.end(), {x.r0.x, x.r0.y, x.r0.z, 0.0f});
			s.restraint_k.push_back(x.k);
		}
		for (size_t k = 0; k < src.grids->size(); ++k)
			s.grids.push_back(src.grids->data()[k]);
		static_cast<marsmd_pair_params&>(s.solvent) = src.nb->get_solvent_params();
		return s;
	}
	void prepare(const NonBondedInteractions* nb, const BondedInteractions* bonded, const DeviceParticleTypes* types,
				 const DeviceBuffer<BaseGridView<mars_real>>* grids, const TablesRegistry* tables) {
		const Sources now{nb, bonded, types, grids, tables};
		if (p_) {
			if (from_setup_)
				return; // built from a PatchSetup: the containers only stand in for it
			if (now.nb != src_.nb || now.bonded != src_.bonded || now.types != src_.types ||
				now.grids != src_.grids || now.tables != src_.tables)
				throw Exception(ExceptionType::ValueError,
								"Patch: interactions, types, grids and tables must be the objects of the first "
								"force call");
			return;
		}
		src_ = now;
		create(setup_from(now));
		from_setup_ = false;
	}
	void create(const PatchSetup& setup) {
		marsmd_patch_desc d{};
		d.capacity = capacity_;
		d.patch_id = id_;
		d.box = box_.c();
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
		from_setup_ = true;
		detail::check(marsmd_patch_set_base_seed(p_, seed_));
		if (n_)
			upload_staged();
	}
	void upload_staged() {
		detail::check(marsmd_patch_copy_particles_from_host(
			p_, reinterpret_cast<const float*>(staged_pos_.data()),
			staged_mom_.empty() ? nullptr : reinterpret_cast<const float*>(staged_mom_.data()), staged_type_.data(),
			n_));
		// the library holds them now
		staged_pos_.clear();
		staged_mom_.clear();
		staged_type_.clear();
	}

	marsmd_patch* p_ = nullptr;
	patch_t id_ = 0;
	idx_t capacity_ = 0;
	Resource resource_{};
	PeriodicBox box_{};
	uint64_t seed_ = 0;
	size_t n_ = 0;
	bool bonded_prepared_ = false;
	bool from_setup_ = false;
	Sources src_{};
	std::vector<Vector3> staged_pos_, staged_mom_;
	std::vector<int> staged_type_, global_id_;
	std::unique_ptr<Pairlist> pairlist_;
};

} // namespace MARS
