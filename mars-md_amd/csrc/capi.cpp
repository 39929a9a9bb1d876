// capi.cpp — the extern "C" boundary declared in include/marsmd.h. Every entry
// point converts exceptions into a marsmd_status and a thread-local message.
#include "batch.hpp"
#include "fm.hpp"
#include "launchers.hpp"
#include "pairlist.hpp"
#include "patch.hpp"

#include <cmath>
#include <cstring>
#include <string>

struct marsmd_zorder {
	mmd::ZSort impl;
	explicit marsmd_zorder(uint64_t n) : impl(n) {}
};
struct marsmd_pairlist {
	mmd::PairlistImpl* impl;
	bool owned;
};
struct marsmd_patch {
	mmd::PatchImpl impl;
	marsmd_pairlist pl_view{};
	explicit marsmd_patch(const marsmd_patch_desc& d) : impl(d) {}
};

namespace {

thread_local std::string g_last_error;

template<class F>
marsmd_status guard(F&& f) {
	try {
		f();
		return MARSMD_OK;
	} catch (const mmd::Error& e) {
		g_last_error = e.what();
		return e.code;
	} catch (const std::exception& e) {
		g_last_error = e.what();
		return MARSMD_ERR_RUNTIME;
	} catch (...) {
		g_last_error = "unknown error";
		return MARSMD_ERR_RUNTIME;
	}
}

hipStream_t st(marsmd_stream s) { return static_cast<hipStream_t>(s); }

void need(const void* p, const char* what) {
	if (!p)
		mmd::raise(MARSMD_ERR_VALUE, std::string(what) + " is NULL");
}

} // namespace

extern "C" {

const char* marsmd_last_error(void) { return g_last_error.c_str(); }
const char* marsmd_version(void) { return "marsmd-hip 0.1.0 gfx950"; }

// ---------------------------------------------------------------- Z-order sort
marsmd_status marsmd_zorder_create(uint64_t max_particles, marsmd_zorder** out) {
	return guard([&] {
		need(out, "out");
		*out = new marsmd_zorder(max_particles);
	});
}
marsmd_status marsmd_zorder_destroy(marsmd_zorder* z) {
	return guard([&] { delete z; });
}
marsmd_status marsmd_zorder_sort_particles(marsmd_zorder* z, const float* pos4, uint64_t n, const float box_min[3],
										   const float box_max[3], marsmd_stream stream) {
	return guard([&] {
		need(z, "sorter");
		need(box_min, "box_min");
		need(box_max, "box_max");
		if (n)
			need(pos4, "positions");
		z->impl.sort(reinterpret_cast<const float4*>(pos4), n, box_min, box_max, nullptr, false, st(stream));
	});
}
marsmd_status marsmd_zorder_get_buffers(marsmd_zorder* z, const uint32_t** codes, const uint32_t** sorted,
										const uint32_t** inverse) {
	return guard([&] {
		need(z, "sorter");
		if (codes)
			*codes = z->impl.codes.data();
		if (sorted)
			*sorted = z->impl.sorted_idx.data();
		if (inverse)
			*inverse = z->impl.inverse.data();
	});
}
marsmd_status marsmd_zorder_reorder_data(marsmd_zorder* z, const void* in, void* out, uint64_t n, uint32_t elem,
										 marsmd_stream stream) {
	return guard([&] {
		need(z, "sorter");
		z->impl.reorder_data(in, out, n, elem, st(stream));
	});
}
marsmd_status marsmd_zorder_remap_indices(marsmd_zorder* z, int32_t* idx, uint64_t n, marsmd_stream stream) {
	return guard([&] {
		need(z, "sorter");
		z->impl.remap_indices(idx, n, st(stream));
	});
}
marsmd_status marsmd_zorder_update_positions_incremental(marsmd_zorder* z, const float* pos4, uint64_t n,
														 marsmd_stream stream) {
	return guard([&] {
		need(z, "sorter");
		if (n)
			need(pos4, "positions");
		z->impl.update_positions_incremental(reinterpret_cast<const float4*>(pos4), n, st(stream));
	});
}
marsmd_status marsmd_zorder_compute_max_displacement(marsmd_zorder* z, const float* pos4, uint64_t n,
													 const float box_len[3], float* max_displacement,
													 marsmd_stream stream) {
	return guard([&] {
		need(z, "sorter");
		need(max_displacement, "max_displacement");
		if (n)
			need(pos4, "positions");
		*max_displacement =
			z->impl.compute_max_displacement(reinterpret_cast<const float4*>(pos4), n, box_len, st(stream));
	});
}
marsmd_status marsmd_radix_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
									  uint32_t* vals_out, uint64_t n, int32_t key_bits, void* scratch,
									  uint64_t scratch_bytes, marsmd_stream stream) {
	return guard([&] {
		mmd::radix_sort_pairs(keys_in, vals_in, keys_out, vals_out, n, key_bits, scratch, scratch_bytes, st(stream));
	});
}
uint64_t marsmd_radix_sort_scratch_bytes(uint64_t n) { return mmd::radix_sort_scratch_bytes(n); }
marsmd_status marsmd_radix_sort_pairs_static(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
											 uint32_t* vals_out, uint64_t n, int32_t key_bits, marsmd_stream stream) {
	static_assert(MARSMD_STATIC_SORT_MAX_KEYS == mmd::kStaticSortMaxKeys, "static sort limit");
	return guard([&] {
		mmd::radix_sort_pairs_static(keys_in, vals_in, keys_out, vals_out, n, key_bits, st(stream));
	});
}
uint64_t marsmd_radix_sort_static_scratch_bytes(void) { return mmd::static_sort_scratch_bytes(); }

// ---------------------------------------------------------------- pairlist
marsmd_status marsmd_pairlist_create(uint64_t max_particles, uint64_t max_pairs, marsmd_pairlist** out) {
	return guard([&] {
		need(out, "out");
		auto* p = new marsmd_pairlist;
		p->impl = new mmd::PairlistImpl(max_particles, max_pairs);
		p->owned = true;
		*out = p;
	});
}
marsmd_status marsmd_pairlist_destroy(marsmd_pairlist* pl) {
	return guard([&] {
		if (!pl)
			return;
		if (pl->owned) {
			delete pl->impl;
			delete pl;
		}
	});
}
marsmd_status marsmd_pairlist_set_periodic_box(marsmd_pairlist* pl, const marsmd_box* box) {
	return guard([&] {
		need(pl, "pairlist");
		need(box, "box");
		pl->impl->set_box(*box);
	});
}
marsmd_status marsmd_pairlist_set_exclusions(marsmd_pairlist* pl, const marsmd_exclusion_view* view) {
	return guard([&] {
		need(pl, "pairlist");
		pl->impl->excl = view ? *view : marsmd_exclusion_view{};
	});
}
marsmd_status marsmd_pairlist_set_max_pairs_ceiling(marsmd_pairlist* pl, uint64_t ceiling) {
	return guard([&] {
		need(pl, "pairlist");
		pl->impl->max_pairs_ceiling = ceiling;
		pl->impl->ref_max_pairs = std::min<uint64_t>(pl->impl->ref_max_pairs, ceiling);
	});
}
marsmd_status marsmd_pairlist_build(marsmd_pairlist* pl, const float* pos4, const int32_t* type_ids, uint64_t n,
									float pairlist_cutoff, marsmd_stream stream) {
	return guard([&] {
		need(pl, "pairlist");
		if (n)
			need(pos4, "positions");
		pl->impl->build(reinterpret_cast<const float4*>(pos4), nullptr, false, type_ids, n, pairlist_cutoff,
						st(stream));
	});
}
marsmd_status marsmd_pairlist_needs_update(marsmd_pairlist* pl, const float* pos4, uint64_t n, float skin,
										   int32_t* needs, float* max_disp, marsmd_stream stream) {
	return guard([&] {
		need(pl, "pairlist");
		if (!pl->impl->built || n != pl->impl->n)
			mmd::raise(MARSMD_ERR_RUNTIME, "needs_update: no pairlist built for this particle count");
		const float d = pl->impl->max_displacement(reinterpret_cast<const float4*>(pos4),
												   pl->impl->sorter.inverse.data(), n, st(stream));
		if (needs)
			*needs = d > skin * 0.5f ? 1 : 0;
		if (max_disp)
			*max_disp = d;
	});
}
marsmd_status marsmd_pairlist_get_num_pairs(marsmd_pairlist* pl, uint64_t* num_pairs) {
	return guard([&] {
		need(pl, "pairlist");
		need(num_pairs, "num_pairs");
		*num_pairs = pl->impl->num_pairs();
	});
}
marsmd_status marsmd_pairlist_get_neighbor_pairs(marsmd_pairlist* pl, const int32_t** pairs, marsmd_stream stream) {
	return guard([&] {
		need(pl, "pairlist");
		need(pairs, "pairs");
		*pairs = pl->impl->neighbor_pairs(st(stream));
	});
}
marsmd_status marsmd_pairlist_get_statistics(marsmd_pairlist* pl, marsmd_pairlist_stats* out) {
	return guard([&] {
		need(pl, "pairlist");
		need(out, "out");
		*out = pl->impl->stats();
	});
}
marsmd_status marsmd_pairlist_get_sorter(marsmd_pairlist* pl, marsmd_zorder** out) {
	return guard([&] {
		need(pl, "pairlist");
		need(out, "out");
		// A non-owning view: the sorter object lives inside the pairlist.
		*out = reinterpret_cast<marsmd_zorder*>(&pl->impl->sorter);
	});
}
marsmd_status marsmd_pairlist_get_reference_cells(marsmd_pairlist* pl, const uint32_t** begin, const uint32_t** end,
												  const uint32_t** neighbors, uint64_t* num_cells, int32_t* slots,
												  marsmd_stream stream) {
	return guard([&] {
		need(pl, "pairlist");
		uint64_t nc = 0;
		pl->impl->reference_cells(st(stream), &nc);
		if (begin)
			*begin = pl->impl->ref_begin.data();
		if (end)
			*end = pl->impl->ref_end.data();
		if (neighbors)
			*neighbors = pl->impl->ref_neighbors.data();
		if (num_cells)
			*num_cells = nc;
		if (slots)
			*slots = int32_t(pl->impl->ref_grid.slots);
	});
}

// ---------------------------------------------------------------- launchers
marsmd_status marsmd_launch_resolve_pair_tables(const int32_t* pairs, const int32_t* type_ids,
												const uint32_t* term_matrix, uint64_t num_types, uint32_t* pair_tag,
												uint64_t num_pairs, marsmd_stream stream) {
	return guard([&] {
		mmd::launch_resolve_pair_tables(pairs, type_ids, term_matrix, num_types, pair_tag, num_pairs, st(stream));
	});
}
marsmd_status marsmd_launch_pairwise_nonbonded(const int32_t* pairs, const marsmd_particle_view* view,
											   const uint32_t* pair_tag, const marsmd_table* tables,
											   const marsmd_particle_types* types, const marsmd_box* box,
											   int32_t get_energy, uint64_t num_pairs, float cutoff_squared,
											   uint32_t enabled_terms, const marsmd_pair_params* params,
											   marsmd_stream stream) {
	return guard([&] {
		need(view, "view");
		need(box, "box");
		mmd::launch_pairwise_nonbonded(pairs, *view, pair_tag, tables, types, *box, get_energy != 0, num_pairs,
									   cutoff_squared, enabled_terms, params, st(stream));
	});
}
marsmd_status marsmd_launch_analytical_pair_nonbonded(const int32_t* pairs, const float* pos4, float* force_energy4,
													  const int32_t* type_ids, const uint32_t* pair_tag,
													  const marsmd_particle_types* types, const marsmd_box* box,
													  uint32_t enabled_terms, const marsmd_pair_params* params,
													  int32_t get_energy, uint64_t num_pairs, float cutoff_squared,
													  marsmd_stream stream) {
	return guard([&] {
		need(types, "types");
		need(box, "box");
		if (num_pairs) {
			need(pairs, "pairs");
			need(pos4, "pos4");
			need(force_energy4, "force_energy4");
			need(type_ids, "type_ids");
		}
		mmd::launch_analytical_pair_nonbonded(pairs, pos4, force_energy4, type_ids, pair_tag, *types, *box,
											  enabled_terms, params, get_energy != 0, num_pairs, cutoff_squared,
											  st(stream));
	});
}
marsmd_status marsmd_launch_tabulated_bonds(const int32_t* bonds, const float* pos4, float* fe,
											const marsmd_table* tables, const int32_t* table_indices,
											const int32_t* forms, const marsmd_box* box, int32_t get_energy,
											uint64_t num_bonds, marsmd_stream stream) {
	return guard([&] {
		need(box, "box");
		mmd::launch_tabulated_bonds(bonds, pos4, fe, tables, table_indices, forms, *box, get_energy != 0, num_bonds,
									st(stream));
	});
}
marsmd_status marsmd_launch_tabulated_angles(const int32_t* angles, const float* pos4, float* fe,
											 const marsmd_table* tables, const int32_t* table_indices,
											 const int32_t* forms, const marsmd_box* box, int32_t get_energy,
											 uint64_t num_angles, marsmd_stream stream) {
	return guard([&] {
		need(box, "box");
		mmd::launch_tabulated_angles(angles, pos4, fe, tables, table_indices, forms, *box, get_energy != 0,
									 num_angles, st(stream));
	});
}

marsmd_status marsmd_launch_tabulated_dihedrals(const int32_t* dihedrals, const float* pos4, float* fe,
												const marsmd_table* tables, const int32_t* table_indices,
												const int32_t* forms, const marsmd_box* box, int32_t get_energy,
												uint64_t num_dihedrals, marsmd_stream stream) {
	return guard([&] {
		need(box, "box");
		mmd::launch_tabulated_dihedrals(dihedrals, pos4, fe, tables, table_indices, forms, *box, get_energy != 0,
										num_dihedrals, st(stream));
	});
}

marsmd_status marsmd_launch_harmonic_restraints(const int32_t* particle_ids, const float* pos4, float* fe,
												const float* anchors4, const float* spring_constants,
												const marsmd_box* box, int32_t get_energy, uint64_t num_restraints,
												marsmd_stream stream) {
	return guard([&] {
		need(box, "box");
		if (num_restraints) {
			need(particle_ids, "particle_ids");
			need(anchors4, "anchors");
			need(spring_constants, "spring_constants");
		}
		mmd::launch_harmonic_restraints(particle_ids, pos4, fe, anchors4, spring_constants, *box, get_energy != 0,
										num_restraints, st(stream));
	});
}
marsmd_status marsmd_launch_bd(const marsmd_particle_view* view, const marsmd_particle_types* types, float timestep,
							   uint64_t current_step, float kT, uint64_t num_particles, const marsmd_box* box,
							   uint64_t base_seed, uint32_t base_ctr, const marsmd_grid_view* grids,
							   const float electric_field[3], int32_t interpolation, marsmd_stream stream) {
	return guard([&] {
		need(view, "view");
		need(types, "types");
		need(box, "box");
		const float zero[3] = {0, 0, 0};
		mmd::launch_bd(*view, *types, timestep, current_step, kT, num_particles, *box, base_seed, base_ctr, grids,
					   electric_field ? electric_field : zero, interpolation, st(stream));
	});
}
marsmd_status marsmd_launch_baoab(const marsmd_particle_view* view, const marsmd_particle_types* types,
								  const marsmd_box* box, float timestep, uint64_t current_step, float kT,
								  uint64_t num_particles, uint64_t base_seed, uint32_t base_ctr, marsmd_stream stream) {
	return guard([&] {
		need(view, "view");
		need(types, "types");
		need(box, "box");
		mmd::launch_baoab(*view, *types, *box, timestep, current_step, kT, num_particles, base_seed, base_ctr,
						  st(stream));
	});
}
marsmd_status marsmd_launch_baoab_last_update(const marsmd_particle_view* view, const marsmd_particle_types* types,
											  float timestep, uint64_t num_particles, marsmd_stream stream) {
	return guard([&] {
		(void)types;
		need(view, "view");
		mmd::launch_baoab_last_update(*view, timestep, num_particles, st(stream));
	});
}
marsmd_status marsmd_launch_pmf(const marsmd_particle_view* view, const marsmd_particle_types* types,
								uint64_t num_particles, const marsmd_grid_view* grids, const float electric_field[3],
								int32_t interpolation, marsmd_stream stream) {
	return guard([&] {
		need(view, "view");
		need(types, "types");
		const float zero[3] = {0, 0, 0};
		mmd::launch_pmf(*view, *types, num_particles, grids, electric_field ? electric_field : zero, interpolation,
						st(stream));
	});
}
marsmd_status marsmd_philox_draw(uint64_t seed, uint32_t base_ctr, uint32_t global_seed, uint32_t ctr1, uint32_t draws,
								 uint64_t n, uint32_t* out, marsmd_stream stream) {
	return guard([&] { mmd::philox_draw(seed, base_ctr, global_seed, ctr1, draws, n, out, st(stream)); });
}

marsmd_status marsmd_selftest_pair_sqrt(uint32_t lo_bits, uint32_t hi_bits, uint64_t* mismatches,
										marsmd_stream stream) {
	return guard([&] {
		need(mismatches, "mismatches");
		*mismatches = mmd::selftest_pair_sqrt(lo_bits, hi_bits, st(stream));
	});
}

// ---------------------------------------------------------------- Patch
marsmd_status marsmd_patch_create(const marsmd_patch_desc* desc, marsmd_patch** out) {
	return guard([&] {
		need(desc, "desc");
		need(out, "out");
		auto* p = new marsmd_patch(*desc);
		p->pl_view.impl = p->impl.pl.get();
		p->pl_view.owned = false;
		*out = p;
	});
}
marsmd_status marsmd_patch_destroy(marsmd_patch* p) {
	return guard([&] { delete p; });
}
marsmd_status marsmd_patch_set_base_seed(marsmd_patch* p, uint64_t seed) {
	return guard([&] {
		need(p, "patch");
		p->impl.base_seed = seed;
	});
}
marsmd_status marsmd_patch_copy_particles_from_host(marsmd_patch* p, const float* pos4, const float* mom4,
													const int32_t* type_ids, uint64_t n) {
	return guard([&] {
		need(p, "patch");
		if (n) {
			need(pos4, "positions");
			need(type_ids, "type_ids");
		}
		p->impl.copy_from_host(pos4, mom4, type_ids, n);
	});
}
marsmd_status marsmd_patch_copy_particles_to_host(marsmd_patch* p, float* pos4, float* mom4, float* force_energy4,
												  int32_t* type_ids, int32_t* ids, uint64_t n) {
	return guard([&] {
		need(p, "patch");
		p->impl.copy_to_host(pos4, mom4, force_energy4, type_ids, ids, n);
	});
}
marsmd_status marsmd_patch_calculate_nonbonded_forces(marsmd_patch* p, float pairlist_cutoff, float interaction_cutoff,
													  uint64_t step, uint64_t rebuild_period,
													  const float electric_field[3], int32_t interpolation,
													  int32_t compute_energy, marsmd_stream stream) {
	return guard([&] {
		need(p, "patch");
		const float zero[3] = {0, 0, 0};
		p->impl.calculate_nonbonded(pairlist_cutoff, interaction_cutoff, step, rebuild_period,
									electric_field ? electric_field : zero, interpolation, compute_energy != 0,
									st(stream));
	});
}
marsmd_status marsmd_patch_calculate_bonded_forces(marsmd_patch* p, int32_t compute_energy, marsmd_stream stream) {
	return guard([&] {
		need(p, "patch");
		p->impl.calculate_bonded(compute_energy != 0, st(stream));
	});
}
marsmd_status marsmd_patch_integrate_motion(marsmd_patch* p, float dt, float kT, int32_t integrator, uint64_t step,
											marsmd_stream stream) {
	return guard([&] {
		need(p, "patch");
		p->impl.integrate(dt, kT, integrator, step, st(stream));
	});
}
marsmd_status marsmd_patch_finish_deferred_kick(marsmd_patch* p, float dt, marsmd_stream stream) {
	return guard([&] {
		need(p, "patch");
		p->impl.finish_deferred_kick(dt, st(stream));
	});
}
marsmd_status marsmd_patch_has_deferred_kick(marsmd_patch* p, int32_t* pending) {
	return guard([&] {
		need(p, "patch");
		need(pending, "pending");
		*pending = p->impl.deferred_kick ? 1 : 0;
	});
}
marsmd_status marsmd_patch_reorder_particles(marsmd_patch* p, marsmd_stream stream) {
	return guard([&] {
		need(p, "patch");
		p->impl.reorder(st(stream));
	});
}
marsmd_status marsmd_patch_get_pairlist(marsmd_patch* p, marsmd_pairlist** out) {
	return guard([&] {
		need(p, "patch");
		need(out, "out");
		*out = &p->pl_view;
	});
}
marsmd_status marsmd_patch_run(marsmd_patch* p, const marsmd_run_params* params, uint64_t first_step,
							   uint64_t num_steps, marsmd_stream stream) {
	return guard([&] {
		need(p, "patch");
		need(params, "params");
		p->impl.run(*params, first_step, num_steps, st(stream));
	});
}

marsmd_status marsmd_patch_set_kernel_timing(marsmd_patch* p, int32_t enable) {
	return guard([&] {
		need(p, "patch");
		p->impl.set_kernel_timing(enable != 0);
	});
}

marsmd_status marsmd_patch_get_kernel_timing(marsmd_patch* p, double* total_ms, uint64_t* launches) {
	return guard([&] {
		need(p, "patch");
		need(total_ms, "total_ms");
		need(launches, "launches");
		p->impl.kernel_timing(total_ms, launches);
	});
}

marsmd_status marsmd_patch_set_launch(marsmd_patch* p, const marsmd_patch_launch* launch) {
	return guard([&] {
		need(p, "patch");
		need(launch, "launch");
		p->impl.set_launch(launch->blocks_per_cu, launch->split_pmf != 0);
	});
}

marsmd_status marsmd_patch_get_launch(marsmd_patch* p, marsmd_patch_launch* launch) {
	return guard([&] {
		need(p, "patch");
		need(launch, "launch");
		launch->blocks_per_cu = p->impl.blocks_per_cu;
		launch->split_pmf = p->impl.split_pmf ? 1 : 0;
	});
}

marsmd_status marsmd_patch_set_cells(marsmd_patch* p, int32_t home_target, int32_t fine_offset) {
	return guard([&] {
		need(p, "patch");
		p->impl.set_cells(home_target, fine_offset);
	});
}
marsmd_status marsmd_patch_get_cells(marsmd_patch* p, int32_t* home_target, int32_t* fine_offset) {
	return guard([&] {
		need(p, "patch");
		need(home_target, "home_target");
		need(fine_offset, "fine_offset");
		*home_target = p->impl.pl->home_target;
		*fine_offset = p->impl.pl->fine_offset;
	});
}
marsmd_status marsmd_patch_set_pruning(marsmd_patch* p, int32_t mode, float margin) {
	return guard([&] {
		need(p, "patch");
		if (mode < 0 || mode > 3)
			mmd::raise(MARSMD_ERR_VALUE, "Patch: pruning mode must be 0, 1, 2 or 3");
		if (!(margin < 0.0f || std::isfinite(margin)))
			mmd::raise(MARSMD_ERR_VALUE, "Patch: pruning margin must be finite");
		p->impl.prune_mode = mode;
		p->impl.prune_margin_fixed = margin < 0.0f ? -1.0f : margin;
		p->impl.prune_valid = false;
	});
}
marsmd_status marsmd_patch_get_pruning(marsmd_patch* p, int32_t* mode, float* margin, int32_t* keep_all,
									   float* displacement) {
	return guard([&] {
		need(p, "patch");
		need(mode, "mode");
		need(margin, "margin");
		need(keep_all, "keep_all");
		need(displacement, "displacement");
		mmd::PatchImpl& P = p->impl;
		*mode = P.prune_mode;
		*margin = P.prune_valid ? P.prune_margin : 0.0f;
		*keep_all = P.prune_valid && P.prune_keep_all ? 1 : 0;
		*displacement = 0.0f;
		if (P.prune_recording && P.prune_disp.count >= mmd::kPruneDispWords) {
			uint32_t bits = 0;
			MMD_HIP(hipMemcpy(&bits, P.prune_disp.data() + 4, 4, hipMemcpyDeviceToHost));
			float d2;
			std::memcpy(&d2, &bits, 4);
			*displacement = std::sqrt(d2);
		}
	});
}
marsmd_status marsmd_patch_autotune_cells(marsmd_patch* p, const marsmd_run_params* params, uint64_t first_step,
										  uint64_t num_steps, double* walk_ms, marsmd_stream stream) {
	return guard([&] {
		need(p, "patch");
		need(params, "params");
		need(walk_ms, "walk_ms");
		p->impl.autotune_cells(*params, first_step, num_steps, st(stream), walk_ms);
	});
}
marsmd_status marsmd_patch_autotune(marsmd_patch* p, const marsmd_run_params* params, uint64_t first_step,
									uint64_t num_steps, marsmd_patch_launch* best, double* best_ms_per_step,
									marsmd_stream stream) {
	return guard([&] {
		need(p, "patch");
		need(params, "params");
		need(best, "best");
		need(best_ms_per_step, "best_ms_per_step");
		p->impl.autotune(*params, first_step, num_steps, st(stream), best_ms_per_step);
		best->blocks_per_cu = p->impl.blocks_per_cu;
		best->split_pmf = p->impl.split_pmf ? 1 : 0;
	});
}

// ------------------------------------------------------------------ replica batch
struct marsmd_batch {
	mmd::BatchImpl impl;
	explicit marsmd_batch(const marsmd_batch_desc& d) : impl(d) {}
};

marsmd_status marsmd_batch_create(const marsmd_batch_desc* desc, marsmd_batch** out) {
	return guard([&] {
		need(desc, "desc");
		need(out, "out");
		*out = nullptr;
		*out = new marsmd_batch(*desc);
	});
}

marsmd_status marsmd_batch_destroy(marsmd_batch* b) {
	return guard([&] { delete b; });
}

marsmd_status marsmd_batch_set_base_seed(marsmd_batch* b, uint64_t seed) {
	return guard([&] {
		need(b, "batch");
		b->impl.base_seed = seed;
	});
}

marsmd_status marsmd_batch_copy_particles_from_host(marsmd_batch* b, const float* pos4, const float* mom4,
													const int32_t* type_ids, uint64_t total) {
	return guard([&] {
		need(b, "batch");
		need(pos4, "pos4");
		need(type_ids, "type_ids");
		b->impl.copy_from_host(pos4, mom4, type_ids, total);
	});
}

marsmd_status marsmd_batch_copy_particles_to_host(marsmd_batch* b, float* pos4, float* mom4, float* force_energy4,
												  int32_t* type_ids, int32_t* ids, uint64_t total) {
	return guard([&] {
		need(b, "batch");
		b->impl.copy_to_host(pos4, mom4, force_energy4, type_ids, ids, total);
	});
}

marsmd_status marsmd_batch_run(marsmd_batch* b, const marsmd_run_params* params, uint64_t first_step,
							   uint64_t num_steps, marsmd_stream stream) {
	return guard([&] {
		need(b, "batch");
		need(params, "params");
		b->impl.run(*params, first_step, num_steps, st(stream));
	});
}

marsmd_status marsmd_batch_finish_deferred_kick(marsmd_batch* b, float dt, marsmd_stream stream) {
	return guard([&] {
		need(b, "batch");
		b->impl.finish_deferred_kick(dt, st(stream));
	});
}

marsmd_status marsmd_batch_get_neighbor_pairs(marsmd_batch* b, uint64_t replica, const int32_t** pairs,
											  uint64_t* num_pairs, marsmd_stream stream) {
	return guard([&] {
		need(b, "batch");
		need(pairs, "pairs");
		need(num_pairs, "num_pairs");
		*pairs = b->impl.neighbor_pairs(replica, num_pairs, st(stream));
	});
}

marsmd_status marsmd_batch_set_kernel_timing(marsmd_batch* b, int32_t enable) {
	return guard([&] {
		need(b, "batch");
		b->impl.set_kernel_timing(enable != 0);
	});
}

marsmd_status marsmd_batch_get_kernel_timing(marsmd_batch* b, double* total_ms, uint64_t* steps) {
	return guard([&] {
		need(b, "batch");
		need(total_ms, "total_ms");
		need(steps, "steps");
		b->impl.kernel_timing(total_ms, steps);
	});
}

marsmd_status marsmd_batch_set_launch(marsmd_batch* b, const marsmd_batch_launch* launch) {
	return guard([&] {
		need(b, "batch");
		need(launch, "launch");
		b->impl.set_launch(launch->blocks_per_cu, launch->use_graph != 0);
	});
}

marsmd_status marsmd_batch_get_launch(marsmd_batch* b, marsmd_batch_launch* launch) {
	return guard([&] {
		need(b, "batch");
		need(launch, "launch");
		launch->blocks_per_cu = b->impl.blocks_per_cu;
		launch->use_graph = b->impl.use_graph ? 1 : 0;
	});
}

marsmd_status marsmd_batch_autotune(marsmd_batch* b, const marsmd_run_params* params, uint64_t first_step,
									uint64_t num_steps, marsmd_batch_launch* best, double* best_ms_per_step,
									marsmd_stream stream) {
	return guard([&] {
		need(b, "batch");
		need(params, "params");
		need(best, "best");
		need(best_ms_per_step, "best_ms_per_step");
		b->impl.autotune(*params, first_step, num_steps, st(stream), best_ms_per_step);
		best->blocks_per_cu = b->impl.blocks_per_cu;
		best->use_graph = b->impl.use_graph ? 1 : 0;
	});
}
marsmd_status marsmd_batch_set_cells(marsmd_batch* b, const marsmd_batch_cells* cells) {
	return guard([&] {
		need(b, "batch");
		need(cells, "cells");
		b->impl.set_cells(*cells);
	});
}
marsmd_status marsmd_batch_get_cells(marsmd_batch* b, marsmd_batch_cells* cells) {
	return guard([&] {
		need(b, "batch");
		need(cells, "cells");
		b->impl.get_cells(cells);
	});
}
marsmd_status marsmd_batch_autotune_cells(marsmd_batch* b, const marsmd_run_params* params, uint64_t first_step,
										  uint64_t num_steps, double* walk_ms, marsmd_stream stream) {
	return guard([&] {
		need(b, "batch");
		need(params, "params");
		need(walk_ms, "walk_ms");
		b->impl.autotune_cells(*params, first_step, num_steps, st(stream), walk_ms);
	});
}

// ------------------------------------------------------------------ force matching
struct marsmd_fm {
	mmd::FMImpl impl;
	explicit marsmd_fm(const marsmd_fm_desc& d) : impl(d) {}
};

marsmd_status marsmd_fm_create(const marsmd_fm_desc* desc, marsmd_fm** out) {
	return guard([&] {
		need(desc, "desc");
		need(out, "out");
		*out = nullptr;
		*out = new marsmd_fm(*desc);
	});
}

marsmd_status marsmd_fm_destroy(marsmd_fm* fm) {
	return guard([&] { delete fm; });
}

marsmd_status marsmd_fm_num_pairs(marsmd_fm* fm, uint64_t* num_pairs) {
	return guard([&] {
		need(fm, "fm");
		need(num_pairs, "num_pairs");
		*num_pairs = fm->impl.num_pairs;
	});
}

marsmd_status marsmd_fm_forces(marsmd_fm* fm, const float* knots, float* force_energy4, marsmd_stream stream) {
	return guard([&] {
		need(fm, "fm");
		need(knots, "knots");
		need(force_energy4, "force_energy4");
		fm->impl.forces(knots, force_energy4, st(stream));
	});
}

marsmd_status marsmd_fm_mse(marsmd_fm* fm, const float* force_energy4, const float* target3, float* grad_force4,
							double* loss, marsmd_stream stream) {
	return guard([&] {
		need(fm, "force matching");
		need(force_energy4, "forces");
		need(target3, "target");
		need(grad_force4, "grad_force4");
		need(loss, "loss");
		fm->impl.mse(force_energy4, target3, grad_force4, loss, st(stream));
	});
}
marsmd_status marsmd_fm_adjoint(marsmd_fm* fm, const float* knots, const float* grad_force4, double* grad_knots,
								marsmd_stream stream) {
	return guard([&] {
		need(fm, "fm");
		need(knots, "knots");
		need(grad_force4, "grad_force4");
		need(grad_knots, "grad_knots");
		fm->impl.adjoint(knots, grad_force4, grad_knots, st(stream));
	});
}
marsmd_status marsmd_fm_loss_sorted(marsmd_fm* fm, const float* knots, const float* target3, float* grad_force4,
									double* loss, marsmd_stream stream) {
	return guard([&] {
		need(fm, "force matching");
		need(knots, "knots");
		need(target3, "target");
		need(grad_force4, "grad_force4");
		need(loss, "loss");
		fm->impl.loss_sorted(knots, target3, grad_force4, loss, st(stream));
	});
}
marsmd_status marsmd_fm_forces_sorted(marsmd_fm* fm, const float* knots, float* force_energy4,
									  marsmd_stream stream) {
	return guard([&] {
		need(fm, "fm");
		need(knots, "knots");
		need(force_energy4, "force_energy4");
		fm->impl.forces(knots, force_energy4, st(stream), true);
	});
}
marsmd_status marsmd_fm_adjoint_sorted(marsmd_fm* fm, const float* knots, const float* grad_force4,
									   double* grad_knots, marsmd_stream stream) {
	return guard([&] {
		need(fm, "fm");
		need(knots, "knots");
		need(grad_force4, "grad_force4");
		need(grad_knots, "grad_knots");
		fm->impl.adjoint(knots, grad_force4, grad_knots, st(stream), true);
	});
}
marsmd_status marsmd_fm_to_sorted(marsmd_fm* fm, const float* in, int32_t width, float* out, marsmd_stream stream) {
	return guard([&] {
		need(fm, "fm");
		need(in, "in");
		need(out, "out");
		if (in == out)
			mmd::raise(MARSMD_ERR_VALUE, "force matching: to_sorted needs distinct in and out");
		fm->impl.to_sorted(in, width, out, st(stream));
	});
}

} // extern "C"
