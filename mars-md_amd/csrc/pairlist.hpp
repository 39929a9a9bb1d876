// pairlist.hpp — host classes behind marsmd_zorder / marsmd_pairlist.
#pragma once

#include "internal.hpp"

#include <vector>

namespace mmd {

// Reference cell grid of ZOrderPairlist::select_cell_grid (ZOrderPairlist.cpp:24-81).
struct CellGrid {
	int m = 0;
	int r[3] = {1, 1, 1};
	long long slots = 1;
};
CellGrid select_cell_grid(float cutoff, uint64_t n, const float extent[3], uint64_t max_pairs);

// ZOrderSort (src/PatchOperation/ZOrderKernels/ZOrderSort.h:43-313).
struct ZSort {
	uint64_t max_particles = 0;
	uint64_t n = 0;
	DevArray<uint32_t> keys_unsorted, vals_unsorted; // canonical-order keys and iota
	DevArray<uint32_t> codes, sorted_idx, inverse;    // results
	DevArray<uint8_t> scratch;
	DevArray<uint32_t> outside_flag; // bit 0: a coordinate was clamped by the encode
	explicit ZSort(uint64_t max_n);
	void ensure(uint64_t n);
	// Encode positions (slot order) into canonical-order keys and sort them.
	// canon_from_meta: the canonical index of slot s is meta_canon(pos[s].w);
	// otherwise it is s. domain_dev (6 floats: min xyz, max xyz) overrides bmin/bmax
	// when non-null.
	// presorted: the canonical-order keys are known to be nondecreasing (the particles
	// were just relabelled by a sort of the same keys), so the stable sort is the
	// identity and is replaced by a copy that checks the order (bit 1 of outside_flag
	// when violated).
	void sort(const float4* pos, uint64_t n, const float bmin[3], const float bmax[3],
			  const float* domain_dev, bool canon_from_meta, hipStream_t s, bool presorted = false);
	// ReorderDataKernel<T> (ZOrderKernels.h:48-63).
	void reorder_data(const void* in, void* out, uint64_t n, uint32_t elem, hipStream_t s) const;
	// RemapIndicesKernel (ZOrderKernels.h:87-101).
	void remap_indices(int32_t* idx, uint64_t count, hipStream_t s) const;
	// ZOrderSort::update_positions_incremental (ZOrderSort.h:172-184): keep a copy of
	// the positions as the reference of compute_max_displacement.
	void update_positions_incremental(const float4* pos, uint64_t n, hipStream_t s);
	// ZOrderSort::compute_max_displacement (ZOrderSort.h:160-170): the largest distance
	// any particle moved since that copy, minimum image on axes with box_len > 0.
	// Reads the result back (synchronises the stream).
	float compute_max_displacement(const float4* pos, uint64_t n, const float box_len[3], hipStream_t s);
	DevArray<float4> disp_ref;
	DevArray<uint32_t> disp_bits;
	uint64_t disp_n = 0;
	bool disp_valid = false;
};

// Full neighbour list in wave tiles. Slots 64w .. 64w+63 form tile w, whose rows hold
// R_w = 4 * ceil(max count in the tile / 4) entries per slot, stored as R_w / 4 groups
// of 64 slots x 4 indices: entry k of slot t = 64w + l is
//   nbr[tile_base[w] + (k / 4) * 256 + 4 l + k % 4].
// One 16-byte load per lane therefore brings 4 neighbours and a wave's loads are
// contiguous. A slot's last group is padded with the slot itself (a zero-distance
// pair, which the force kernels skip like the reference's d < 1e-6 test).
struct NbrList {
	const uint32_t* count;
	const uint32_t* tile_base;
	const uint32_t* nbr;
};
__host__ __device__ inline uint64_t nbr_slot_index(uint64_t tile_base, uint32_t t, uint32_t k) {
	return tile_base + uint64_t(k >> 2) * 256u + uint64_t(t & 63u) * 4u + (k & 3u);
}

struct PairlistImpl {
	uint64_t max_particles = 0;
	uint64_t ref_max_pairs = 0;     // Pairlist::max_pairs_ bookkeeping (grows x1.25)
	bool presorted_next = false;    // set by a reorder: the next build's sort is the identity
	uint64_t max_pairs_ceiling = ~uint64_t(0);
	marsmd_box box{};
	marsmd_exclusion_view excl{};
	ZSort sorter;
	uint64_t n = 0;
	float cutoff = 0.0f, cutoff_squared = 0.0f;
	bool built = false;
	// sorted positions (meta in .w) of the last build = the displacement snapshot
	DevArray<float4> spos;
	// build-grid cell ranges
	DevArray<uint32_t> cell_begin, cell_end;
	// full neighbour list over sorted indices in wave tiles (see NbrList)
	DevArray<uint32_t> counts, tile_rows, tile_base, nbr;
	// home-cell walk: home-level cell ranges and the per-slot capacity of its rows
	DevArray<uint32_t> home_begin, home_end;
	uint32_t home_cap = 0;
	// cell configuration of the home walk: home cells of the finest level holding at
	// least home_target particles on average, candidate cells fine_offset levels below
	int home_target = 24, fine_offset = 2;
	std::vector<int4> stencil_host; // the home walk's fine-cell stencil
	DevArray<int4> stencil_dev;
	std::vector<float> stencil_key; // the parameters stencil_dev was made for
	PinnedWords readback{8};
	int home_level = -1; // level of the last home-cell walk (-1: count / scan / fill path)
	uint64_t list_version = 0; // counts the builds and re-walks that rewrote the rows
	uint64_t nbr_entries = 0; // = 2 * num_pairs
	uint64_t nbr_storage = 0; // padded entries held by nbr
	DevArray<uint8_t> scan_scratch;
	// lazily materialised int2 pairs (canonical, x < y)
	DevArray<int32_t> pairs;
	DevArray<uint32_t> half_counts, half_offsets;
	bool pairs_valid = false;
	// bookkeeping
	CellGrid ref_grid, build_grid;
	bool ref_complete = false;
	int build_m_override = -1;
	float domain_min[3]{}, domain_max[3]{};
	double last_build_ms = 0.0;
	uint32_t max_neighbors = 0;
	// misc device scalars: [1] max displacement^2 bits, [2] max count, [4..5] entries (u64)
	DevArray<uint32_t> dev_scalars;
	DevArray<float> dev_domain;
	DevArray<uint32_t> ref_begin, ref_end, ref_neighbors;

	PairlistImpl(uint64_t max_particles_, uint64_t max_pairs_ceiling_);
	void set_box(const marsmd_box& b) {
		box = b;
		presorted_next = false; // the keys of a new box are not the reorder's
	}
	// Build from positions in slot order. slot_of_canon maps canonical index -> slot
	// (NULL = identity); canon_from_meta takes canonical indices (and types) from
	// pos[s].w, else from the slot index and `types`.
	void build(const float4* pos, const uint32_t* slot_of_canon, bool canon_from_meta,
			   const int32_t* types, uint64_t n, float pairlist_cutoff, hipStream_t s);
	// max |min_image(pos - snapshot)|^2 over particles; pos in slot order, mapped to
	// sorted slots by `sorted_of_slot` (NULL = pos already in sorted order).
	float max_displacement(const float4* pos, const uint32_t* sorted_of_slot, uint64_t n,
						   hipStream_t s);
	uint64_t num_pairs() const { return nbr_entries / 2; }
	NbrList list() const { return NbrList{counts.data(), tile_base.data(), nbr.data()}; }
	const int32_t* neighbor_pairs(hipStream_t s);
	// The home walk again over the last build's sorted positions with the current cell
	// configuration (the same pair set; rows may list it in another order): its time in
	// ms, or a negative value when the home walk does not apply.
	double rewalk_ms(hipStream_t s);
	void reference_cells(hipStream_t s, uint64_t* num_cells);
	marsmd_pairlist_stats stats() const;

  private:
	void choose_grids(uint64_t n);
	void walk(bool fill, bool use_ref_grid, hipStream_t s);
	bool walk_home(hipStream_t s, bool& outside);
	void count_stats(uint64_t n, hipStream_t s);
	void ensure(uint64_t n);
};

} // namespace mmd
