// batch.hpp — host class behind marsmd_batch (see batch.hip).
#pragma once

#include <map>
#include <vector>

#include "pairlist.hpp"

namespace mmd {

struct NbTables;
struct StepParams;

// Size buckets of the per-bucket cell configuration: bucket b holds the replicas with
// 2^b <= n < 2^(b+1) particles.
constexpr int kBatchBuckets = MARSMD_BATCH_BUCKETS;
int batch_size_bucket(uint64_t n);

struct BatchImpl {
	uint64_t R = 0, total = 0, base_seed = 0;
	std::vector<uint64_t> off, count, pid;
	std::vector<marsmd_box> boxes;
	marsmd_particle_types types{};
	NbTables* tables = nullptr;
	DevArray<float> knots;
	// state in slot order: replica r owns slots [off_r, off_r + n_r), Morton order within
	DevArray<float4> pos[2], mom, force, snap;
	int cur = 0;
	DevArray<int32_t> ids, ids_tmp;
	DevArray<float4> mom_tmp;
	DevArray<uint32_t> rep_of_slot, slot_of_canon, slot_of_canon_tmp;
	DevArray<uint32_t> sort_order; // replicas by size, largest first
	// rebuild
	DevArray<uint32_t> keys, vals, keys_s, vals_s, rkeys, rvals, rkeys_s, rvals_s, codes, perm;
	DevArray<uint8_t> sort_scratch, geom;
	DevArray<uint32_t> fine_begin, fine_end, home_begin, home_end, item_off;
	DevArray<uint32_t> counts, tile_base, nbr, scalars;
	DevArray<int32_t> pairs;
	DevArray<uint32_t> half_counts, half_offsets;
	DevArray<uint8_t> scan_scratch;
	uint64_t num_fine = 0, num_home = 0, num_items = 0, nbr_entries = 0;
	uint32_t cap = 0, max_neighbors = 0;
	std::vector<uint8_t> geom_host;     // host copies behind the asynchronous uploads
	// fine-cell stencils of the home walk, one per distinct replica geometry (key: the
	// walk parameters it depends on; value: offset and count in stencils)
	std::map<std::vector<float>, std::pair<uint32_t, uint32_t>> stencil_index;
	std::vector<int4> stencils_host;
	DevArray<int4> stencils;
	size_t stencils_uploaded = 0;
	std::vector<uint32_t> items_host;
	struct Plan { // what geom_host and items_host were planned from
		bool valid = false;
		float rc = 0.0f;
		std::vector<int> target, fine;
		std::vector<uint64_t> ref;
		uint64_t fine_total = 0, home_total = 0;
	} plan;
	std::vector<uint64_t> ref_max_pairs; // per replica: Pairlist::max_pairs_ bookkeeping
	std::vector<uint64_t> rep_pairs_host;
	uint8_t* staging = nullptr;   // pinned: the geometry and item offsets of a build's uploads
	uint64_t* readback = nullptr; // pinned: the walk's statistics (3 words), then the replicas' pair counts
	DevArray<unsigned long long> rep_pairs;
	int rep_bits = 1;
	float built_cutoff = 0.0f;
	bool built = false, prepared = false, deferred_kick = false;
	// launch timing of the step kernel: event pairs around each launch or replayed
	// segment, with the number of steps each pair spans
	bool timing = false;
	std::vector<hipEvent_t> timing_events;
	std::vector<uint64_t> timing_steps;
	size_t timing_used = 0;
	int occupancy = 0, num_cus = 0;
	// launch configuration: persistent step-kernel workgroups per CU (0: as many as
	// fit) and whether runs of steps that need no host decision replay as hipGraphs
	int blocks_per_cu = 0;
	bool use_graph = false;
	struct SegmentGraph {
		std::vector<uint8_t> key;
		hipGraphExec_t exec = nullptr;
	};
	std::vector<SegmentGraph> graphs;
	hipStream_t capture_stream = nullptr;
	DevArray<uint32_t> step_word;
	// cell configuration per size bucket: the walk's home cells hold at least
	// cell_target particles on average; its fine cells lie cell_fine levels below
	std::vector<int> cell_target = std::vector<int>(size_t(kBatchBuckets), 24);
	std::vector<int> cell_fine = std::vector<int>(size_t(kBatchBuckets), 2);
	// scratch of the walk timing (the live list is untouched)
	DevArray<uint8_t> t_geom;
	DevArray<uint32_t> t_items, t_fine_begin, t_fine_end, t_home_begin, t_home_end, t_counts, t_nbr;
	hipEvent_t t_events[2] = {nullptr, nullptr};

	explicit BatchImpl(const marsmd_batch_desc& d);
	~BatchImpl();
	BatchImpl(const BatchImpl&) = delete;
	BatchImpl& operator=(const BatchImpl&) = delete;

	void copy_from_host(const float* pos4, const float* mom4, const int32_t* type_ids, uint64_t n);
	void copy_to_host(float* pos4, float* mom4, float* force4, int32_t* type_ids, int32_t* ids, uint64_t n);
	void run(const marsmd_run_params& p, uint64_t first_step, uint64_t num_steps, hipStream_t s);
	void finish_deferred_kick(float dt, hipStream_t s);
	// Canonical (x < y, replica-local indices) pairs of one replica's current list.
	const int32_t* neighbor_pairs(uint64_t replica, uint64_t* num_pairs, hipStream_t s);
	void set_kernel_timing(bool enable);
	void kernel_timing(double* total_ms, uint64_t* steps);
	void set_launch(int blocks_per_cu, bool graph);
	// Steps num_steps real steps, timing every candidate launch configuration on runs
	// of steps, and keeps the fastest (its ms per step in *best_ms).
	void autotune(const marsmd_run_params& p, uint64_t first_step, uint64_t num_steps, hipStream_t s,
				  double* best_ms);
	void set_cells(const marsmd_batch_cells& c);
	void get_cells(marsmd_batch_cells* c) const;
	// Steps num_steps real steps after choosing, bucket by bucket, the cell
	// configuration whose cell tables and walk over the last build's positions take
	// least time (*walk_ms: that time); the choice applies from the next build on.
	void autotune_cells(const marsmd_run_params& p, uint64_t first_step, uint64_t num_steps, hipStream_t s,
						double* walk_ms);

  private:
	void plan_geometry(float rc, std::vector<uint8_t>& gh, std::vector<uint32_t>& items, uint64_t& fine_total,
					   uint64_t& home_total);
	double walk_ms(hipStream_t s);
	void build(float pairlist_cutoff, bool relabel, hipStream_t s);
	bool needs_rebuild(float skin, hipStream_t s);
	unsigned step_grid(const void* kernel);
	// count consecutive steps from `first` with the current list; `store` keeps the
	// forces of the last one
	void launch_steps(const marsmd_run_params& p, uint64_t first, uint64_t count, bool store, hipStream_t s);
	hipGraphExec_t segment_graph(const void* kernel, unsigned grid, unsigned lds, const StepParams& S,
								 uint64_t count);
	void clear_graphs();
};

} // namespace mmd
