// patch.hpp — host class behind marsmd_patch (see patch.hip).
#pragma once

#include <map>
#include <memory>
#include <utility>
#include <vector>

#include "pairlist.hpp"

namespace mmd {

// words of the pruned rows' displacement record: four steps' maxima (ring), the life's
constexpr uint64_t kPruneDispWords = 5;

struct PatchImpl {
	struct Tables;
	uint64_t capacity = 0, patch_id = 0, base_seed = 0;
	marsmd_box box{};
	marsmd_particle_types types{};
	Tables* tables = nullptr;
	// topology (canonical) and its slot-order copy
	uint64_t num_bonds = 0, num_angles = 0, num_excl = 0;
	int64_t topo_max_index = -1;
	DevArray<int32_t> bonds_canon, bonds_slot, bond_table;
	DevArray<int32_t> angles_canon, angles_slot, angle_table; // 4 ints per angle
	uint64_t num_dihedrals = 0, num_restraints = 0;
	DevArray<int32_t> dihedrals_canon, dihedrals_slot, dihedral_table; // 4 ints per dihedral
	DevArray<int32_t> restraint_canon, restraint_slot;
	DevArray<float4> restraint_anchor;
	DevArray<float> restraint_k;
	DevArray<int32_t> excl_pairs;
	DevArray<uint32_t> excl_offsets, excl_deg, excl_nbrs;
	DevArray<uint8_t> excl_scan;
	// grids
	int num_grids = 0;
	std::vector<DevArray<float>> grid_data;
	DevArray<marsmd_grid_view> grids_dev;
	// particle state, slot order
	uint64_t n = 0;
	DevArray<float4> pos[2];
	int cur = 0;
	DevArray<float4> mom, mom_tmp, force;
	DevArray<int32_t> ids, ids_tmp, remap_tmp;
	DevArray<uint32_t> slot_of_canon, slot_tmp, relabel;
	std::unique_ptr<PairlistImpl> pl;
	bool built = false, force_rebuild = false, deferred_kick = false, bonded_prepared = false;
	float efield[3] = {0, 0, 0};
	int interpolation = MARSMD_INTERP_LINEAR;
	// persistent launches: kernel -> (dynamic LDS bytes, resident workgroups per CU)
	std::map<const void*, std::pair<unsigned, int>> occupancy;
	int num_cus = 0;
	// launch timing of the fused step kernel (events recorded around each launch)
	bool timing = false;
	std::vector<hipEvent_t> timing_events; // pairs: start, stop
	size_t timing_used = 0;
	// launch configuration: persistent step-kernel workgroups per CU (0: as many as
	// fit) and whether a kernel of its own precomputes the position-dependent force
	int blocks_per_cu = 0;
	bool split_pmf = false;
	DevArray<float4> pmf_buf;
	// Pruned rows of the fused step (k_prune): the neighbour list's rows holding the
	// entries within cutoff + 2 margin of the positions at the prune (all entries in
	// sort-only mode), placed in each 256-slot block by ascending length (modes 1, 2)
	// or at their own slots (mode 3). mode 0 (default): off, 1: the rows sorted only,
	// 2: sorted and pruned, 3: pruned; the margin adapts to the displacement of the
	// previous list's life (fixed when margin_fixed >= 0).
	int prune_mode = 0;
	float prune_margin_fixed = -1.0f;
	float prune_margin = 0.0f;     // margin of the live pruned rows (0: sort only)
	bool prune_keep_all = true;    // the live pruned rows keep every entry
	bool prune_valid = false;      // the pruned rows match the list and the positions
	uint64_t prune_list_version = 0;
	uint32_t prune_since = 0;      // fused steps since the prune
	bool prune_recording = false;    // the fused steps since the last prune record displacements
	bool prune_life_tracked = false; // the previous life's record was copied to prune_host
	uint64_t prune_fallback_steps = 0, prune_steps = 0;
	DevArray<uint32_t> prune_lane_slot, prune_count, prune_tile_base, prune_nbr, prune_disp;
	DevArray<float4> prune_ref;
	// bonded shares (share r of term i of a kind with R particles at kind base + R i + r)
	// and, per slot, its shares in share-index order: share_vals[q + begin .. q + end)
	DevArray<float4> shares;
	DevArray<uint32_t> share_keys, share_vals, share_begin, share_end;
	DevArray<uint8_t> share_scratch;
	PinnedWords prune_host{2};

	explicit PatchImpl(const marsmd_patch_desc& d);
	~PatchImpl();
	PatchImpl(const PatchImpl&) = delete;
	PatchImpl& operator=(const PatchImpl&) = delete;

	void copy_from_host(const float* pos4, const float* mom4, const int32_t* type_ids, uint64_t n);
	void copy_to_host(float* pos4, float* mom4, float* force4, int32_t* type_ids, int32_t* ids, uint64_t n);
	void calculate_nonbonded(float pairlist_cutoff, float cutoff, uint64_t step, uint64_t period, const float ef[3],
							 int interp, bool energy, hipStream_t s);
	// Bonded forces into `force`: added to it (after calculate_nonbonded) or written.
	void calculate_bonded(bool energy, hipStream_t s, bool add = true);
	void integrate(float dt, float kT, int integrator, uint64_t step, hipStream_t s);
	void finish_deferred_kick(float dt, hipStream_t s);
	void reorder(hipStream_t s);
	// store_last: the last step keeps its forces in `force` (the caller reads them)
	void run(const marsmd_run_params& p, uint64_t first_step, uint64_t num_steps, hipStream_t s,
			 bool store_last = true);
	void set_kernel_timing(bool enable);
	void kernel_timing(double* total_ms, uint64_t* launches);
	void set_launch(int blocks_per_cu, bool split_pmf);
	// Steps num_steps real steps while timing every candidate launch configuration on
	// the step kernels, and keeps the fastest (its ms per step in *best_ms).
	void autotune(const marsmd_run_params& p, uint64_t first_step, uint64_t num_steps, hipStream_t s,
				  double* best_ms);
	void set_cells(int home_target, int fine_offset);
	// Steps num_steps real steps after choosing the pairlist walk's cell configuration
	// (each candidate timed re-walking the first build's positions; the live list is
	// left as the choice walks it); *walk_ms: the choice's walk time.
	void autotune_cells(const marsmd_run_params& p, uint64_t first_step, uint64_t num_steps, hipStream_t s,
						double* walk_ms);

  private:
	void build(float pairlist_cutoff, hipStream_t s);
	// (Re)build the pruned rows when the list or the positions changed since the last
	// prune (after a build: margin from the previous life's displacement).
	void prepare_prune(float cutoff, float pairlist_cutoff, hipStream_t s);
	bool rebuild_policy(float pairlist_cutoff, float cutoff, uint64_t step, uint64_t period, hipStream_t s);
	void rebuild_exclusions(hipStream_t s);
	void k_remap_canon(int32_t* idx, uint64_t count, hipStream_t s);
	uint64_t num_bonded_terms() const { return num_bonds + num_angles + num_dihedrals + num_restraints; }
	uint64_t num_shares() const;
	void build_share_index(hipStream_t s);
	unsigned lds_bytes() const;
	unsigned persistent_grid(const void* kernel, uint64_t n);
};

} // namespace mmd
