// batch.hip — many independent replicas of a periodic nonbonded system, each one the
// reference's Patch, stepped together on one GPU: one persistent launch per step over
// every replica's particles and one batched rebuild (sort, cell tables, home-cell walk,
// one capacity read-back) for all replicas.
//
// Each replica follows the Patch / SimManager sequence (Patch.cpp:85-121, 238-352,
// 456-480; ReorderManager.cpp:8-25) exactly: its id is the patch id of the seed mix
// (Patch.cpp:252-254); its canonical index (Philox counter, exported order) is
// relabelled at every reorder as Patch::reorder_particles does, by a sort whose ties
// fall to the canonical index as the reference's stable sort; its pair set is its own
// ZOrderPairlist's. Only the scheduling is shared: replica r owns slots
// [off_r, off_r + n_r), every sort is stable within replicas (each replica's range
// sorted by Morton key in one workgroup; above 20480 particles a global sort by key,
// then by replica id), and each kernel serves all replicas.
//
// Cadence: at a reorder step every replica reorders and rebuilds, as each Patch does.
// At a rebuild-period step that is not a reorder step the displacements of all
// replicas are reduced in one kernel and, if any replica has moved more than skin/2,
// every replica gets a new list (forces are those of the same pairs within the cutoff;
// only the FP32 summation order of a replica that would have kept its list changes).
// With the reference's default reorderPeriod (= decompPeriod) the two coincide.
//
// Scope: the replicas share particle types and pair tables and carry no bonded
// topology, exclusions or grids (the workload of many independent nonbonded systems);
// every box is periodic on all axes and particles must lie inside their box.
#include "batch.hpp"

#include "homewalk.hpp"
#include "nbforce.hpp"
#include "step.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace mmd {

namespace {

constexpr int kBatchBlock = 256;

// Per-replica geometry of the batched build and step.
struct RepGeom {
	DevBox box;
	float lo[3], hi[3]; // Morton domain = the periodic box
	float cw[3], margin[3];
	uint32_t slot0, n;
	int h, mf, reach[3], full[3];
	uint32_t home_off, fine_off;
	uint64_t pid;
	// the reference grid's stencil, when it misses seam pairs (see HomeParams)
	int ref_filter, ref_shift, ref_cells, ref_r[3], ref_full[3];
	// the home-frame prefilter of the walk (see HomeParams)
	int pre;
	float pre_lo2, pre_hi2;
	// the replica's fine-cell stencil in the batch's table (count 0: none)
	uint32_t stencil_off, stencil_n;
};

// A wrapped coordinate lies in [lo, hi] up to the rounding of PeriodicBox::wrap_scalar
// (x - floor((x - o) inv_L) L can end an ulp or two below lo or above hi; the Morton
// quantisation clamps such a coordinate, as the reference's does). One more than 2^-20
// of the box length outside, or NaN, has left the box.
__device__ __forceinline__ bool beyond_box(float x, float lo, float hi) {
	const float tol = (hi - lo) * 0x1p-20f;
	return !(x >= lo - tol && x <= hi + tol);
}

// keys / vals in canonical position: global canonical g = off_r + canon.
__global__ void k_b_encode(const float4* __restrict__ pos, uint64_t total, const uint32_t* __restrict__ rep_of_slot,
						   const RepGeom* __restrict__ geom, uint32_t* __restrict__ keys,
						   uint32_t* __restrict__ vals, uint32_t* __restrict__ outside) {
	const uint64_t s = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (s >= total)
		return;
	const float4 p = pos[s];
	const RepGeom& G = geom[rep_of_slot[s]];
	const uint32_t g = G.slot0 + meta_canon(__float_as_uint(p.w));
	bool out = false;
	const uint32_t qx = morton_quantize(p.x, G.lo[0], G.hi[0], out);
	const uint32_t qy = morton_quantize(p.y, G.lo[1], G.hi[1], out);
	const uint32_t qz = morton_quantize(p.z, G.lo[2], G.hi[2], out);
	// stored at the canonical index, so the stable sort breaks ties by it; the value is
	// the particle's current slot, which the gather reads from
	keys[g] = split_by3(qz) | (split_by3(qy) << 1) | (split_by3(qx) << 2);
	vals[g] = uint32_t(s);
	if (beyond_box(p.x, G.lo[0], G.hi[0]) || beyond_box(p.y, G.lo[1], G.hi[1]) || beyond_box(p.z, G.lo[2], G.hi[2]))
		atomicOr(outside, 1u);
}

// Second pass key of the global sort (replicas above the per-replica sort's size): the
// replica of each Morton-sorted entry (a slot of that replica), carrying the entry's
// current slot along.
__global__ void k_b_rep_keys(const uint32_t* __restrict__ vals_s, uint64_t total,
							 const uint32_t* __restrict__ rep_of_slot, uint32_t* __restrict__ rkeys,
							 uint32_t* __restrict__ rvals) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= total)
		return;
	const uint32_t old = vals_s[i];
	rkeys[i] = rep_of_slot[old];
	rvals[i] = old;
}

// The same order in one launch when every replica has at most kSegSortMax particles:
// a stable LSD radix sort of each replica's Morton keys (three 10-bit digits) by one
// workgroup of kSegWaves waves, the replica resident in LDS. The encode leaves replica
// r's entries at [slot0, slot0 + n) in canonical order, so sorting each such range
// stably by key gives the order of the global stable sort by key followed by the stable
// sort by replica. Wave w takes the run [w C, w C + C) of the range (C = ceil(n /
// kSegWaves)) in batches of 64 held in registers, counts its digits, and after the
// block's exclusive scan (digit-major, wave-minor) scatters each batch in order: the
// entries of one batch with one digit rank by lane (their peer mask from 10 ballots) and
// the lowest of them moves the wave's offset for that digit on. Equal digits therefore
// keep wave, batch and lane order. A pass reads the whole range into registers before
// its first barrier and scatters it back after the scan, so one LDS copy suffices; the
// entries carry their canonical position (16 bits), and the old slots are read once,
// in order, at the end. Counts and offsets stay below 2^16, so two waves share a
// histogram word (wave 2k in the low half, 2k + 1 in the high half).
constexpr int kSegWaves = 16;
constexpr int kSegThreads = kSegWaves * 64;
constexpr int kSegPerLane = 20;
constexpr uint32_t kSegSortMax = uint32_t(kSegThreads) * kSegPerLane; // 20480: 159.8 KB of LDS
constexpr int kSegStride = kSegWaves / 2 + 1;                          // words per digit, padded
constexpr uint32_t kSegSpare = 1024u * kSegStride;                      // the bin of idle lanes

__global__ __launch_bounds__(kSegThreads) void k_b_seg_sort(const RepGeom* __restrict__ geom,
															const uint32_t* __restrict__ order,
															const uint32_t* __restrict__ keys,
															const uint32_t* __restrict__ vals,
															uint32_t* __restrict__ out_vals) {
	__shared__ uint32_t s_key[kSegSortMax];
	__shared__ uint16_t s_pos[kSegSortMax];
	__shared__ uint32_t hist[1024 * kSegStride + 1];
	__shared__ uint32_t wsum[kSegWaves];
	const uint32_t r = order[blockIdx.x];
	const uint32_t base = geom[r].slot0, n = geom[r].n;
	const int wv = int(threadIdx.x >> 6), lane = int(threadIdx.x & 63);
	const uint32_t half = 16u * uint32_t(wv & 1), col = uint32_t(wv >> 1);
	const uint32_t C = (n + kSegWaves - 1) / kSegWaves; // <= 64 * kSegPerLane
	const uint32_t w0_ = min(n, uint32_t(wv) * C), w1_ = min(n, w0_ + C);
	const uint64_t lt = (uint64_t(1) << lane) - 1;
	uint32_t key[kSegPerLane], at[kSegPerLane];
	if (w1_ > w0_) { // the first pass's entries, from global memory (reads clamped into the run)
		const uint32_t nb = (w1_ - w0_ + 63u) / 64u;
#pragma unroll
		for (int j = 0; j < kSegPerLane; ++j) {
			if (uint32_t(j) < nb) { // wave-uniform
				const uint32_t e = min(w0_ + uint32_t(j) * 64u + uint32_t(lane), w1_ - 1u);
				key[j] = keys[base + e];
				at[j] = e;
			}
		}
	}
#pragma unroll 1
	for (int pass = 0; pass < 3; ++pass) {
		const int shift = 10 * pass;
		// the run's bounds, opaque to the compiler each pass: otherwise it hoists every
		// batch's address and mask out of the pass loop (64 SGPRs and 32 VGPRs)
		uint32_t w0 = w0_, w1 = w1_;
		asm volatile("" : "+v"(w0), "+v"(w1));
		const uint32_t nb = (w1 - w0 + 63u) / 64u; // batches of this wave
		for (uint32_t i = threadIdx.x; i <= kSegSpare; i += kSegThreads)
			hist[i] = 0u;
		__syncthreads();
		if (w1 > w0) { // reads clamped into the run, a lane past its end counted in the spare bin
			if (pass > 0) {
#pragma unroll
				for (int j = 0; j < kSegPerLane; ++j) {
					if (uint32_t(j) < nb) { // wave-uniform
						const uint32_t e = min(w0 + uint32_t(j) * 64u + uint32_t(lane), w1 - 1u);
						key[j] = s_key[e];
						at[j] = s_pos[e];
					}
				}
			}
#pragma unroll
			for (int j = 0; j < kSegPerLane; ++j) {
				if (uint32_t(j) < nb) {
					const bool live = w0 + uint32_t(j) * 64u + uint32_t(lane) < w1;
					atomicAdd(&hist[live ? ((key[j] >> shift) & 1023u) * kSegStride + col : kSegSpare],
							  live ? 1u << half : 1u);
				}
			}
		}
		__syncthreads();
		{ // thread t owns digit t: its waves' counts become their first destinations
			const uint32_t row = threadIdx.x * kSegStride;
			uint32_t c[kSegWaves / 2], s = 0;
#pragma unroll
			for (int k = 0; k < kSegWaves / 2; ++k) {
				c[k] = hist[row + uint32_t(k)];
				s += (c[k] & 0xFFFFu) + (c[k] >> 16);
			}
			const uint32_t incl = wave_inclusive_scan(s);
			if (lane == 63)
				wsum[wv] = incl;
			__syncthreads();
			uint32_t run = incl - s;
			for (int w = 0; w < wv; ++w)
				run += wsum[w];
#pragma unroll
			for (int k = 0; k < kSegWaves / 2; ++k) {
				const uint32_t lo = run;
				run += c[k] & 0xFFFFu;
				hist[row + uint32_t(k)] = lo | (run << 16);
				run += c[k] >> 16;
			}
		}
		__syncthreads();
#pragma unroll
		for (int j = 0; j < kSegPerLane; ++j) {
			if (uint32_t(j) >= nb) // wave-uniform
				continue;
			const bool live = w0 + uint32_t(j) * 64u + uint32_t(lane) < w1;
			const uint32_t d = (key[j] >> shift) & 1023u;
			// the peer mask on 32-bit halves, sx all ones where the lane's bit is set: the
			// lanes that agree on every bit stay (4 VALU a bit instead of 10 on 64-bit masks)
			const uint64_t lv = __ballot(live);
			uint32_t plo = uint32_t(lv), phi = uint32_t(lv >> 32);
#pragma unroll
			for (int b = 0; b < 10; ++b) {
				const int32_t sx = int32_t(d << (31 - b)) >> 31;
				const uint64_t ones = __ballot(sx != 0);
				plo &= ~(uint32_t(ones) ^ uint32_t(sx));
				phi &= ~(uint32_t(ones >> 32) ^ uint32_t(sx));
			}
			const uint64_t peers = uint64_t(plo) | (uint64_t(phi) << 32);
			if (live) {
				// the other half of the word is the sibling wave's: it moves by atomic adds
				uint32_t* h = &hist[d * kSegStride + col];
				const uint32_t first = (*h >> half) & 0xFFFFu;
				const uint32_t rank = uint32_t(__popcll(peers & lt));
				if (rank == 0u)
					atomicAdd(h, uint32_t(__popcll(peers)) << half);
				const uint32_t dst = first + rank;
				if (pass < 2)
					s_key[dst] = key[j];
				s_pos[dst] = uint16_t(at[j]);
			}
		}
		__syncthreads();
	}
#pragma unroll
	for (int k = 0; k < kSegPerLane; ++k) { // every load in flight at once
		const uint32_t i = threadIdx.x + uint32_t(k) * kSegThreads;
		if (i < n)
			out_vals[base + i] = vals[base + s_pos[i]];
	}
}

// New slot i holds the particle of current slot old_slot[i], its Morton code computed
// again from the same position with the same operations as k_b_encode.
__global__ void k_b_gather(const uint32_t* __restrict__ old_slot, uint64_t total,
						   const uint32_t* __restrict__ rep_of_slot, const RepGeom* __restrict__ geom,
						   const float4* __restrict__ pos_in, const float4* __restrict__ mom_in,
						   const int32_t* __restrict__ ids_in, bool relabel, float4* __restrict__ pos_out,
						   float4* __restrict__ snap, float4* __restrict__ mom_out, int32_t* __restrict__ ids_out,
						   uint32_t* __restrict__ codes, uint32_t* __restrict__ slot_of_canon_out) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= total)
		return;
	const uint32_t old = old_slot[i];
	const RepGeom& G = geom[rep_of_slot[i]];
	float4 p = pos_in[old];
	bool out = false;
	const uint32_t qx = morton_quantize(p.x, G.lo[0], G.hi[0], out);
	const uint32_t qy = morton_quantize(p.y, G.lo[1], G.hi[1], out);
	const uint32_t qz = morton_quantize(p.z, G.lo[2], G.hi[2], out);
	const uint32_t m = __float_as_uint(p.w);
	const uint32_t canon = relabel ? uint32_t(i) - G.slot0 : meta_canon(m);
	p.w = __uint_as_float(make_meta(meta_type(m), canon));
	pos_out[i] = p;
	snap[i] = p;
	mom_out[i] = mom_in[old];
	ids_out[i] = ids_in[old];
	codes[i] = split_by3(qz) | (split_by3(qy) << 1) | (split_by3(qx) << 2);
	slot_of_canon_out[G.slot0 + canon] = uint32_t(i);
}

// Fine and home cell ranges of every replica (BuildCellRangesKernel, ZOrderNeighbor.h:27-43,
// per replica), in global slots.
__global__ void k_b_cell_ranges(const uint32_t* __restrict__ codes, uint64_t total,
								const uint32_t* __restrict__ rep_of_slot, const RepGeom* __restrict__ geom,
								uint32_t* __restrict__ fine_begin, uint32_t* __restrict__ fine_end,
								uint32_t* __restrict__ home_begin, uint32_t* __restrict__ home_end,
								uint32_t* __restrict__ flags) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= total)
		return;
	const RepGeom& G = geom[rep_of_slot[i]];
	const bool first = i == G.slot0;
	const bool last = i + 1 == uint64_t(G.slot0) + G.n;
	const uint32_t c = codes[i];
	// the gather recomputes each code from the position it moves: codes out of order
	// within a replica (a mismatch with the sort's keys) would give wrong cell ranges
	if (!first && codes[i - 1] > c && flags)
		atomicOr(flags, 2u);
	const int fs = 3 * (10 - G.mf), hs = 3 * (10 - G.h);
	const uint32_t fc = c >> fs, hc = c >> hs;
	if (first || (codes[i - 1] >> fs) != fc)
		fine_begin[G.fine_off + fc] = uint32_t(i);
	if (last || (codes[i + 1] >> fs) != fc)
		fine_end[G.fine_off + fc] = uint32_t(i + 1);
	if (first || (codes[i - 1] >> hs) != hc)
		home_begin[G.home_off + hc] = uint32_t(i);
	if (last || (codes[i + 1] >> hs) != hc)
		home_end[G.home_off + hc] = uint32_t(i + 1);
}

// The home-cell walk over the work items of all replicas (item_off[r] = first item of
// replica r, one item per home cell), at 6 waves per SIMD (80 VGPRs) as the Patch's walk:
// more waves hide the per-home-particle LDS reads.
__global__ __launch_bounds__(256, 6) void k_b_walk(const RepGeom* __restrict__ geom, const uint32_t* __restrict__ item_off,
												uint32_t R, uint32_t num_items, const float4* __restrict__ spos,
												const uint32_t* __restrict__ home_begin,
												const uint32_t* __restrict__ home_end,
												const uint32_t* __restrict__ fine_begin,
												const uint32_t* __restrict__ fine_end, float rc2, uint32_t cap,
												const int4* __restrict__ stencils, uint32_t* __restrict__ counts,
												uint32_t* __restrict__ nbr) {
	__shared__ uint32_t seg_b[kHomeWaves][kSegMax];
	__shared__ uint32_t seg_pre[kHomeWaves][kSegMax + 1];
	__shared__ float4 home_pos[kHomeWaves][kHomeLds];
	__shared__ uint32_t keep[kHomeWaves][kHomeKeep];
	const int wv = threadIdx.x >> 6;
	const TileRange TR = xcd_tiles((num_items + kHomeWaves - 1) / kHomeWaves);
	for (uint32_t gidx = TR.tile; gidx < TR.end; gidx += TR.stride) {
		const uint32_t item = gidx * kHomeWaves + uint32_t(wv);
		if (item >= num_items)
			continue;
		uint32_t lo = 0, hi = R; // item_off[lo] <= item < item_off[hi]
		while (hi - lo > 1) {
			const uint32_t mid = (lo + hi) >> 1;
			if (item_off[mid] <= item)
				lo = mid;
			else
				hi = mid;
		}
		const uint32_t r = uint32_t(__builtin_amdgcn_readfirstlane(int(lo)));
		const RepGeom& G = geom[r];
		HomeParams P{};
		P.spos = spos;
		P.home_begin = home_begin + G.home_off;
		P.home_end = home_end + G.home_off;
		P.fine_begin = fine_begin + G.fine_off;
		P.fine_end = fine_end + G.fine_off;
		P.num_home = uint32_t(1) << (3 * G.h);
		P.h = G.h;
		P.s = G.mf - G.h;
		for (int a = 0; a < 3; ++a) {
			P.reach[a] = G.reach[a];
			P.full[a] = G.full[a];
			P.cw[a] = G.cw[a];
			P.margin[a] = G.margin[a];
		}
		P.rc2 = rc2;
		P.box = G.box;
		P.ex = marsmd_exclusion_view{};
		P.cap = cap;
		P.ref_filter = G.ref_filter;
		P.ref_shift = G.ref_shift;
		P.ref_cells = G.ref_cells;
		for (int a = 0; a < 3; ++a) {
			P.ref_r[a] = G.ref_r[a];
			P.ref_full[a] = G.ref_full[a];
			P.dom_lo[a] = G.lo[a];
			P.dom_hi[a] = G.hi[a];
		}
		P.pre = G.pre;
		P.pre_lo2 = G.pre_lo2;
		P.pre_hi2 = G.pre_hi2;
		P.stencil = G.stencil_n ? stencils + G.stencil_off : nullptr;
		P.stencil_n = G.stencil_n;
		walk_home_cell<true>(P, item - item_off[r], seg_b[wv], seg_pre[wv], home_pos[wv], keep[wv], counts, nbr);
	}
}

// max over particles of |min_image(pos - snapshot)|^2, per batch (the largest replica
// displacement decides): DisplacementKernel (AdaptiveKernels.h:48-101).
__global__ __launch_bounds__(256) void k_b_max_disp(const float4* __restrict__ cur, const float4* __restrict__ snap,
													uint64_t total, const uint32_t* __restrict__ rep_of_slot,
													const RepGeom* __restrict__ geom,
													uint32_t* __restrict__ out_bits) {
	float m = 0.0f;
	for (uint64_t s = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; s < total;
		 s += uint64_t(gridDim.x) * blockDim.x) {
		const float4 a = cur[s];
		const float4 b = snap[s];
		const RepGeom& G = geom[rep_of_slot[s]];
		float d[3] = {a.x - b.x, a.y - b.y, a.z - b.z};
#pragma unroll
		for (int k = 0; k < 3; ++k) {
			const float L = G.box.L[k];
			int guard = 0;
			while (d[k] > 0.5f * L && guard++ < 8)
				d[k] -= L;
			while (d[k] < -0.5f * L && guard++ < 16)
				d[k] += L;
		}
		m = fmaxf(m, d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
	}
	for (int o = 32; o > 0; o >>= 1)
		m = fmaxf(m, __shfl_xor(m, o, 64));
	if ((threadIdx.x & 63) == 0)
		atomicMax(out_bits, __float_as_uint(m));
}

// The fused step over every replica's particles, persistent over XCD-local tiles: the
// replica of a slot gives its box and its seed mix. Bounded to 96 VGPRs (5 waves per
// SIMD; no spills in the BD case, 2 dwords in the others): 160 -> 156.5 us at 512
// replicas against 4 waves with 123 VGPRs.
template<int INTEG, bool ENERGY, bool ANALYTIC>
__global__ __launch_bounds__(kBatchBlock, 5) void k_b_step(const float4* __restrict__ pos, float4* __restrict__ pos_next,
														  float4* __restrict__ mom, float4* __restrict__ force,
														  uint64_t total, const uint32_t* __restrict__ rep_of_slot,
														  const RepGeom* __restrict__ geom, NbrList L, NbTables T,
														  const float* __restrict__ knots_g,
														  marsmd_particle_types types_arg, StepParams S,
														  uint64_t base_seed, int store_force,
														  const uint32_t* __restrict__ step_base) {
	extern __shared__ float lds_knots[];
	__shared__ NbLds lds_nb;
	__shared__ marsmd_particle_types lds_types;
	__shared__ TypeStep lds_steps[MARSMD_MAX_TYPES]; // the integrator's per-type constants
	stage_types(types_arg, lds_types);
	if (threadIdx.x < uint32_t(types_arg.num_types))
		lds_steps[threadIdx.x] = INTEG == MARSMD_INTEGRATOR_BROWNIAN ? bd_type_step(types_arg, int(threadIdx.x), S)
																	 : baoab_type_step(types_arg, int(threadIdx.x), S);
	stage_tables(T, knots_g, lds_knots, lds_nb);
	const marsmd_particle_types& types = lds_types;
	// a replayed segment carries the step's offset in S.step and its first step in memory
	const uint32_t step = S.step + (step_base ? step_base[0] : 0u);
	const TileRange R = xcd_tiles(uint32_t((total + kBatchBlock - 1) / kBatchBlock));
	for (uint32_t tile = R.tile; tile < R.end; tile += R.stride) {
		const uint64_t t = uint64_t(tile) * kBatchBlock + threadIdx.x;
		if (t >= total)
			continue;
		const RepGeom& G = geom[rep_of_slot[t]];
		StepParams P = S;
		P.box = G.box;
		P.step = step;
		P.seed = mix_seed(base_seed, G.pid, uint64_t(step));
		const float4 p = pos[t];
		float4 f = nb_force<ANALYTIC>(uint32_t(t), p, pos, L, T, lds_knots, knots_g, lds_nb, P.box, P.cut_d, ENERGY);
		float4 pn;
		if (INTEG == MARSMD_INTEGRATOR_BROWNIAN) {
			const float4 pmf = position_dependent_force(p.x, p.y, p.z, int(meta_type(fb(p.w))), types, nullptr, P.ef,
														P.scheme);
			pn = bd_update(p, f, lds_steps[meta_type(fb(p.w))], pmf, P);
		} else {
			float4 mo = mom[t];
			pn = baoab_update(p, mo, f, lds_steps[meta_type(fb(p.w))], P);
			mom[t] = mo;
		}
		pos_next[t] = pn;
		if (ENERGY || store_force)
			force[t] = f;
	}
}

__global__ void k_b_import(const float4* __restrict__ pos_h, const float4* __restrict__ mom_h,
						   const int32_t* __restrict__ types, uint64_t total, const uint32_t* __restrict__ rep_of_slot,
						   const RepGeom* __restrict__ geom, float4* __restrict__ pos, float4* __restrict__ mom,
						   int32_t* __restrict__ ids, uint32_t* __restrict__ slot_of_canon,
						   float4* __restrict__ force) {
	const uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (g >= total)
		return;
	const uint32_t c = uint32_t(g) - geom[rep_of_slot[g]].slot0;
	const float4 p = pos_h[g];
	pos[g] = make_float4(p.x, p.y, p.z, __uint_as_float(make_meta(uint32_t(types[g]), c)));
	mom[g] = mom_h ? make_float4(mom_h[g].x, mom_h[g].y, mom_h[g].z, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	ids[g] = int32_t(c);
	slot_of_canon[g] = uint32_t(g);
	force[g] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__global__ void k_b_export(const float4* __restrict__ pos, const float4* __restrict__ mom,
						   const float4* __restrict__ force, const int32_t* __restrict__ ids,
						   const uint32_t* __restrict__ slot_of_canon, uint64_t total, float4* __restrict__ pos_o,
						   float4* __restrict__ mom_o, float4* __restrict__ force_o, int32_t* __restrict__ type_o,
						   int32_t* __restrict__ ids_o) {
	const uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (g >= total)
		return;
	const uint32_t s = slot_of_canon[g];
	const float4 p = pos[s];
	pos_o[g] = make_float4(p.x, p.y, p.z, 0.0f);
	mom_o[g] = mom[s];
	force_o[g] = force[s];
	type_o[g] = int32_t(meta_type(__float_as_uint(p.w)));
	ids_o[g] = ids[s];
}

__global__ void k_b_half_counts(NbrList L, uint64_t slot0, uint64_t n, uint32_t* __restrict__ half) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const uint32_t t = uint32_t(slot0 + i);
	const uint64_t base = L.tile_base[t >> 6];
	uint32_t h = 0;
	for (uint32_t k = 0; k < L.count[t]; ++k)
		h += L.nbr[nbr_slot_index(base, t, k)] > t ? 1u : 0u;
	half[i] = h;
}

__global__ void k_b_write_pairs(NbrList L, const float4* __restrict__ pos, uint64_t slot0, uint64_t n,
								const uint32_t* __restrict__ half_off, int32_t* __restrict__ pairs) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const uint32_t t = uint32_t(slot0 + i);
	uint64_t w = half_off[i];
	const uint64_t base = L.tile_base[t >> 6];
	const int32_t a = int32_t(meta_canon(__float_as_uint(pos[t].w)));
	for (uint32_t k = 0; k < L.count[t]; ++k) {
		const uint32_t j = L.nbr[nbr_slot_index(base, t, k)];
		if (j > t) {
			const int32_t b = int32_t(meta_canon(__float_as_uint(pos[j].w)));
			pairs[2 * w] = a < b ? a : b;
			pairs[2 * w + 1] = a < b ? b : a;
			++w;
		}
	}
}

// Full-list entries of each replica (twice its pair count): one workgroup per replica
// sums the counts of its slots, which are one contiguous range.
__global__ __launch_bounds__(256) void k_b_rep_pairs(const uint32_t* __restrict__ counts,
													 const RepGeom* __restrict__ geom,
													 unsigned long long* __restrict__ out) {
	__shared__ unsigned long long wsum[4];
	const uint32_t b = geom[blockIdx.x].slot0;
	const uint32_t e = b + geom[blockIdx.x].n;
	unsigned long long sum = 0;
	for (uint32_t i = b + threadIdx.x; i < e; i += 1024u) { // four loads in flight
		const uint32_t c0 = counts[i];
		const uint32_t c1 = i + 256u < e ? counts[i + 256u] : 0u;
		const uint32_t c2 = i + 512u < e ? counts[i + 512u] : 0u;
		const uint32_t c3 = i + 768u < e ? counts[i + 768u] : 0u;
		sum += (unsigned long long)c0 + c1 + c2 + c3;
	}
	for (int o = 32; o > 0; o >>= 1)
		sum += __shfl_xor(sum, o, 64);
	if ((threadIdx.x & 63) == 0)
		wsum[threadIdx.x >> 6] = sum;
	__syncthreads();
	if (threadIdx.x == 0)
		out[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// BAOAB_LastUpdate (BAOAB.h:151-172) on every replica's particles.
__global__ void k_b_last_update(float4* __restrict__ mom, const float4* __restrict__ force, uint64_t total, float dt) {
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= total)
		return;
	const float4 f = force[t];
	float4 m = mom[t];
	const float hdt = 0.5f * dt;
	const float FCF = 4.18679994e4f;
	m.x += hdt * f.x * FCF;
	m.y += hdt * f.y * FCF;
	m.z += hdt * f.z * FCF;
	mom[t] = m;
}

// The first step of a replayed segment, written by the stream before the replay.
__global__ void k_b_set_word(uint32_t* __restrict__ word, uint32_t value) {
	if (threadIdx.x == 0)
		word[0] = value;
}

using StepKernel = void (*)(const float4*, float4*, float4*, float4*, uint64_t, const uint32_t*, const RepGeom*,
							NbrList, NbTables, const float*, marsmd_particle_types, StepParams, uint64_t, int,
							const uint32_t*);

// What a captured segment bakes in besides the step offsets: the graph cache's key.
struct SegmentKey {
	const void* kernel;
	unsigned grid, lds;
	uint64_t count, total, base_seed;
	const void* ptr[9];
	StepParams S;
	NbTables T;
	marsmd_particle_types types;
};

__global__ void k_b_uniform_tile_base(uint32_t* __restrict__ tile_base, uint64_t tiles, uint32_t row) {
	const uint64_t w = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (w <= tiles)
		tile_base[w] = uint32_t(w * row);
}

__global__ __launch_bounds__(256) void k_b_count_stats(const uint32_t* __restrict__ counts, uint64_t n,
													   unsigned long long* __restrict__ total,
													   uint32_t* __restrict__ max_count) {
	__shared__ uint32_t wmax[4];
	__shared__ unsigned long long wsum[4];
	uint32_t m = 0;
	unsigned long long sum = 0;
	const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
	for (uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += 4 * stride) {
		// four loads in flight per thread (0 is neutral for both the sum and the maximum)
		const uint32_t c0 = counts[t];
		const uint32_t c1 = t + stride < n ? counts[t + stride] : 0u;
		const uint32_t c2 = t + 2 * stride < n ? counts[t + 2 * stride] : 0u;
		const uint32_t c3 = t + 3 * stride < n ? counts[t + 3 * stride] : 0u;
		m = max(max(m, c0), max(c1, max(c2, c3)));
		sum += (unsigned long long)c0 + c1 + c2 + c3;
	}
	for (int o = 32; o > 0; o >>= 1) {
		m = max(m, uint32_t(__shfl_xor(int(m), o, 64)));
		sum += __shfl_xor(sum, o, 64);
	}
	const int w = threadIdx.x >> 6;
	if ((threadIdx.x & 63) == 0) {
		wmax[w] = m;
		wsum[w] = sum;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		uint32_t bm = 0;
		unsigned long long bs = 0;
		for (int k = 0; k < int(blockDim.x >> 6); ++k) {
			bm = max(bm, wmax[k]);
			bs += wsum[k];
		}
		atomicAdd(total, bs);
		atomicMax(max_count, bm);
	}
}

} // namespace

BatchImpl::BatchImpl(const marsmd_batch_desc& d) : tables(new NbTables{}) {
	R = d.num_replicas;
	if (R == 0)
		raise(MARSMD_ERR_VALUE, "batch: no replicas");
	if (!d.num_particles || !d.boxes)
		raise(MARSMD_ERR_VALUE, "batch: num_particles and boxes are required");
	types = d.types;
	if (types.num_types < 1 || types.num_types > MARSMD_MAX_TYPES)
		raise(MARSMD_ERR_VALUE, "batch: num_types must be in [1, 16]");
	for (int t = 0; t < types.num_types; ++t)
		if (types.pmf_grid_count[t] != 0)
			raise(MARSMD_ERR_VALUE, "batch: grids are not on the batched path");
	off.assign(R + 1, 0);
	for (uint64_t r = 0; r < R; ++r) {
		const marsmd_box& b = d.boxes[r];
		for (int a = 0; a < 3; ++a)
			if (!(b.periodic[a] && b.size[a] > 0.0f))
				raise(MARSMD_ERR_VALUE, "batch: every replica box must be periodic with a positive size");
		if (d.num_particles[r] >= kMaxParticles)
			raise(MARSMD_ERR_VALUE, "batch: a replica holds at most 2^24 particles");
		boxes.push_back(b);
		count.push_back(d.num_particles[r]);
		pid.push_back(d.patch_ids ? d.patch_ids[r] : r);
		off[r + 1] = off[r] + d.num_particles[r];
	}
	total = off[R];
	if (total == 0 || total >= (uint64_t(1) << 28)) // position records addressed by 32-bit byte offsets
		raise(MARSMD_ERR_VALUE, "batch: total particle count must be in [1, 2^28)");
	while ((uint64_t(1) << rep_bits) < R)
		++rep_bits;
	// pair tables (as PatchImpl)
	NbTables& T = *tables;
	T.num_types = types.num_types;
	T.enabled = d.enabled_terms;
	T.coef = make_pair_coeffs(types, d.pair_params);
	for (int i = 0; i < types.num_types * types.num_types; ++i)
		T.tag[i] = d.term_matrix ? d.term_matrix[i] : 0u;
	if (d.num_nb_tables < 0 || d.num_nb_tables > kMaxTables)
		raise(MARSMD_ERR_VALUE, "batch: too many nonbonded tables");
	T.num_tables = d.num_nb_tables;
	std::vector<float> kn;
	for (int k = 0; k < d.num_nb_tables; ++k) {
		const marsmd_table& t = d.nb_tables[k];
		if (t.size < 2 || t.is_periodic)
			raise(MARSMD_ERR_VALUE, "batch: nonbonded tables need >= 2 knots and are not periodic");
		T.start[k] = t.start;
		T.step_inv[k] = t.step_inv;
		T.size[k] = int(t.size);
		T.offset[k] = int(kn.size());
		kn.insert(kn.end(), t.pot, t.pot + t.size);
	}
	for (int i = 0; i < types.num_types * types.num_types; ++i)
		if ((T.tag[i] & MARSMD_PAIR_TERM_TABULATED) && int(T.tag[i] >> MARSMD_PAIR_TABLE_SHIFT) >= d.num_nb_tables)
			raise(MARSMD_ERR_VALUE, "batch: a pair tag names a missing table");
	T.total_knots = int(kn.size());
	T.lds = T.total_knots <= kLdsKnots ? 1 : 0;
	knots.resize(std::max<uint64_t>(kn.size(), 1));
	if (!kn.empty())
		MMD_HIP(hipMemcpy(knots.data(), kn.data(), kn.size() * 4, hipMemcpyHostToDevice));
	// slot -> replica
	std::vector<uint32_t> rep(total);
	for (uint64_t r = 0; r < R; ++r)
		std::fill(rep.begin() + long(off[r]), rep.begin() + long(off[r + 1]), uint32_t(r));
	rep_of_slot.resize(total);
	MMD_HIP(hipMemcpy(rep_of_slot.data(), rep.data(), 4 * total, hipMemcpyHostToDevice));
	// the per-replica sort's workgroups, largest replica first (the dispatcher hands the
	// later, smaller ones to the CUs that free up first)
	std::vector<uint32_t> order(R);
	for (uint64_t r = 0; r < R; ++r)
		order[r] = uint32_t(r);
	std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return count[a] > count[b]; });
	sort_order.resize(R);
	MMD_HIP(hipMemcpy(sort_order.data(), order.data(), 4 * R, hipMemcpyHostToDevice));
	for (int b = 0; b < 2; ++b)
		pos[b].resize(total);
	mom.resize(total);
	mom_tmp.resize(total);
	force.resize(total);
	snap.resize(total);
	ids.resize(total);
	ids_tmp.resize(total);
	slot_of_canon.resize(total);
	slot_of_canon_tmp.resize(total);
	scalars.resize(8);
	ref_max_pairs.resize(R);
	for (uint64_t r = 0; r < R; ++r)
		ref_max_pairs[r] = std::max<uint64_t>(count[r], 1);
	rep_pairs.resize(R);
	rep_pairs_host.resize(R);
}

BatchImpl::~BatchImpl() {
	clear_graphs();
	if (readback)
		(void)hipHostFree(readback);
	if (staging)
		(void)hipHostFree(staging);
	if (capture_stream)
		(void)hipStreamDestroy(capture_stream);
	for (hipEvent_t e : timing_events)
		(void)hipEventDestroy(e);
	for (hipEvent_t e : t_events)
		if (e)
			(void)hipEventDestroy(e);
	delete tables;
}

void BatchImpl::clear_graphs() {
	for (SegmentGraph& g : graphs)
		if (g.exec)
			(void)hipGraphExecDestroy(g.exec);
	graphs.clear();
}

void BatchImpl::copy_from_host(const float* pos4, const float* mom4, const int32_t* type_ids, uint64_t n) {
	if (n != total)
		raise(MARSMD_ERR_VALUE, "batch: copy_particles_from_host takes every replica's particles (" +
									std::to_string(total) + ")");
	for (uint64_t r = 0; r < R; ++r)
		for (uint64_t g = off[r]; g < off[r + 1]; ++g) {
			if (type_ids[g] < 0 || type_ids[g] >= types.num_types)
				raise(MARSMD_ERR_VALUE, "batch: particle type out of range");
			for (int a = 0; a < 3; ++a) {
				const float x = pos4[4 * g + uint64_t(a)];
				const float lo = boxes[r].origin[a];
				if (!(x >= lo && x <= lo + boxes[r].size[a]))
					raise(MARSMD_ERR_VALUE, "batch: particle " + std::to_string(g - off[r]) + " of replica " +
												std::to_string(r) + " lies outside its box");
			}
		}
	DevArray<float4> hp(total), hm(mom4 ? total : 0);
	DevArray<int32_t> ht(total);
	MMD_HIP(hipMemcpy(hp.data(), pos4, 16 * total, hipMemcpyHostToDevice));
	if (mom4)
		MMD_HIP(hipMemcpy(hm.data(), mom4, 16 * total, hipMemcpyHostToDevice));
	MMD_HIP(hipMemcpy(ht.data(), type_ids, 4 * total, hipMemcpyHostToDevice));
	// geometry needs only the boxes here; the build fills the rest
	std::vector<RepGeom> G(R);
	for (uint64_t r = 0; r < R; ++r) {
		G[r].box = make_devbox(boxes[r]);
		G[r].slot0 = uint32_t(off[r]);
		G[r].n = uint32_t(count[r]);
		G[r].pid = pid[r];
	}
	geom.resize(R * sizeof(RepGeom));
	MMD_HIP(hipMemcpy(geom.data(), G.data(), R * sizeof(RepGeom), hipMemcpyHostToDevice));
	k_b_import<<<grid_for(total, 256), 256>>>(hp.data(), mom4 ? hm.data() : nullptr, ht.data(), total,
											  rep_of_slot.data(), reinterpret_cast<const RepGeom*>(geom.data()),
											  pos[0].data(), mom.data(), ids.data(), slot_of_canon.data(),
											  force.data());
	MMD_LAUNCH_CHECK();
	MMD_HIP(hipDeviceSynchronize());
	cur = 0;
	built = false;
	prepared = false;
	deferred_kick = false;
}

void BatchImpl::copy_to_host(float* pos4, float* mom4, float* force4, int32_t* type_ids, int32_t* id_out,
							 uint64_t n) {
	if (n != total)
		raise(MARSMD_ERR_VALUE, "batch: copy_particles_to_host returns every replica's particles");
	DevArray<float4> op(total), om(total), of(total);
	DevArray<int32_t> ot(total), oi(total);
	MMD_HIP(hipDeviceSynchronize());
	k_b_export<<<grid_for(total, 256), 256>>>(pos[cur].data(), mom.data(), force.data(), ids.data(),
											  slot_of_canon.data(), total, op.data(), om.data(), of.data(), ot.data(),
											  oi.data());
	MMD_LAUNCH_CHECK();
	MMD_HIP(hipDeviceSynchronize());
	if (pos4)
		MMD_HIP(hipMemcpy(pos4, op.data(), 16 * total, hipMemcpyDeviceToHost));
	if (mom4)
		MMD_HIP(hipMemcpy(mom4, om.data(), 16 * total, hipMemcpyDeviceToHost));
	if (force4)
		MMD_HIP(hipMemcpy(force4, of.data(), 16 * total, hipMemcpyDeviceToHost));
	if (type_ids)
		MMD_HIP(hipMemcpy(type_ids, ot.data(), 4 * total, hipMemcpyDeviceToHost));
	if (id_out)
		MMD_HIP(hipMemcpy(id_out, oi.data(), 4 * total, hipMemcpyDeviceToHost));
}

// Sort every replica (stable by Morton key of its box, ties by canonical index), move
// the state into the new slot order (relabelling canonical indices when reordering),
// then cell tables and the home-cell walk of all replicas, with one read-back.
int batch_size_bucket(uint64_t n) {
	int b = 0;
	while (b + 1 < kBatchBuckets && (uint64_t(1) << (b + 1)) <= n)
		++b;
	return b;
}

// Per-replica geometry of a build with the cell configuration of each replica's size
// bucket: home cells of the largest level h <= m - 1 holding at least `home_target`
// particles on average, fine cells `fine_offset` levels below (at most level 10).
void BatchImpl::plan_geometry(float rc, std::vector<uint8_t>& gh, std::vector<uint32_t>& items,
							  uint64_t& fine_total, uint64_t& home_total) {
	const float rc2 = rc * rc;
	gh.assign(R * sizeof(RepGeom), 0);
	RepGeom* G = reinterpret_cast<RepGeom*>(gh.data());
	items.assign(R + 1, 0);
	fine_total = 0;
	home_total = 0;
	for (uint64_t r = 0; r < R; ++r) {
		RepGeom& g = G[r];
		const uint64_t n = count[r];
		g.box = make_devbox(boxes[r]);
		g.slot0 = uint32_t(off[r]);
		g.n = uint32_t(n);
		g.pid = pid[r];
		const int bucket = batch_size_bucket(n);
		const uint64_t target = uint64_t(std::max(1, cell_target[size_t(bucket)]));
		int m = 0;
		while (m < 10 && (uint64_t(1) << (3 * (m + 1))) <= 2 * std::max<uint64_t>(n, 1))
			++m;
		int h = 0;
		while (h + 1 <= m - 1 && (n >> (3 * (h + 1))) >= target)
			++h;
		g.h = h;
		g.mf = std::min(10, h + std::max(1, cell_fine[size_t(bucket)]));
		const int fine_cells = 1 << g.mf;
		const int span = 1 << (g.mf - h);
		const int k = 1024 / fine_cells;
		for (int a = 0; a < 3; ++a) {
			g.lo[a] = boxes[r].origin[a];
			g.hi[a] = boxes[r].origin[a] + boxes[r].size[a];
			const float extent = std::fmax(g.hi[a] - g.lo[a], 1e-6f);
			const double W = double(extent) / 1023.0;
			const double need = (double(rc) * (1.0 + 1e-6) + 1e-3) / W + 1.0;
			g.reach[a] = std::max(1, int(std::min(std::ceil(need / double(k)), double(fine_cells))));
			g.full[a] = span + 2 * g.reach[a] >= fine_cells ? 1 : 0;
			g.cw[a] = float(double(extent) * k / 1023.0);
			g.margin[a] = float(W * 1.001 + 1e-5 * double(extent) + 1e-3);
		}
		// the reference grid of this replica: when it misses seam pairs, the walk keeps
		// exactly the pairs its stencil reaches
		const float ext[3] = {boxes[r].size[0], boxes[r].size[1], boxes[r].size[2]};
		const CellGrid ref = select_cell_grid(rc, n, ext, ref_max_pairs[r]);
		const int rcells = 1 << ref.m;
		bool complete = true;
		for (int a = 0; a < 3; ++a) {
			g.ref_r[a] = ref.r[a];
			g.ref_full[a] = 2 * ref.r[a] + 1 >= rcells ? 1 : 0;
			if (g.ref_full[a])
				continue;
			const double W = double(ext[a]) / 1023.0;
			if ((double(ref.r[a]) * (1024 / rcells) - 1.0) * W < double(rc) * (1.0 + 1e-6) + 1e-3)
				complete = false;
		}
		g.ref_filter = complete ? 0 : 1;
		g.pre = home_prefilter_band(boxes[r], g.lo, g.hi, h, rc, &g.pre_lo2, &g.pre_hi2) ? 1 : 0;
		// the walk's stencil, shared by the replicas of one geometry
		const std::vector<float> key = {float(h),		   float(g.mf - h),	   float(g.reach[0]),
										float(g.reach[1]), float(g.reach[2]),  float(g.full[0]),
										float(g.full[1]),  float(g.full[2]),   g.cw[0],
										g.cw[1],		   g.cw[2],			   g.margin[0],
										g.margin[1],	   g.margin[2],		   rc2};
		auto it = stencil_index.find(key);
		if (it == stencil_index.end()) {
			const std::vector<int4> st = home_stencil(h, g.mf - h, g.reach, g.full, g.cw, g.margin, rc2, 8192);
			it = stencil_index.emplace(key, std::make_pair(uint32_t(stencils_host.size()), uint32_t(st.size()))).first;
			stencils_host.insert(stencils_host.end(), st.begin(), st.end());
		}
		g.stencil_off = it->second.first;
		g.stencil_n = it->second.second;
		g.ref_shift = 10 - ref.m;
		g.ref_cells = rcells;
		g.fine_off = uint32_t(fine_total);
		g.home_off = uint32_t(home_total);
		fine_total += uint64_t(1) << (3 * g.mf);
		home_total += uint64_t(1) << (3 * h);
		items[r + 1] = items[r] + (uint32_t(1) << (3 * h));
	}
	if (fine_total >= (uint64_t(1) << 32))
		raise(MARSMD_ERR_VALUE, "batch: cell tables exceed 2^32 entries");
	if (stencils_host.size() > stencils_uploaded) { // new geometries (a synchronous upload, rare)
		stencils.resize(stencils_host.size());
		MMD_HIP(hipMemcpy(stencils.data(), stencils_host.data(), sizeof(int4) * stencils_host.size(),
						  hipMemcpyHostToDevice));
		stencils_uploaded = stencils_host.size();
	}
}

void BatchImpl::build(float pairlist_cutoff, bool relabel, hipStream_t s) {
	const float rc = pairlist_cutoff;
	const float rc2 = rc * rc;
	std::vector<uint32_t>& items = items_host;
	// the plan depends on the cutoff, the cell configuration and the pair bookkeeping
	// (boxes and sizes are fixed): planned again only when one of them changed, which
	// keeps it off the GPU's idle time when a run starts with a rebuild
	if (!(plan.valid && plan.rc == rc && plan.target == cell_target && plan.fine == cell_fine &&
		  plan.ref == ref_max_pairs)) {
		plan.valid = false; // until planned whole
		plan_geometry(rc, geom_host, items, plan.fine_total, plan.home_total);
		plan.valid = true;
		plan.rc = rc;
		plan.target = cell_target;
		plan.fine = cell_fine;
		plan.ref = ref_max_pairs;
	}
	const uint64_t fine_total = plan.fine_total, home_total = plan.home_total;
	num_fine = fine_total;
	num_home = home_total;
	num_items = items[R];
	// both uploads from pinned memory, so that they queue behind the running steps
	// without the host waiting for them (the last build's copies ended before its wait)
	const size_t gbytes = R * sizeof(RepGeom);
	if (!staging)
		MMD_HIP(hipHostMalloc(reinterpret_cast<void**>(&staging), gbytes + 4 * (R + 1), hipHostMallocDefault));
	std::memcpy(staging, geom_host.data(), gbytes);
	std::memcpy(staging + gbytes, items.data(), 4 * (R + 1));
	geom.resize(gbytes);
	MMD_HIP(hipMemcpyAsync(geom.data(), staging, gbytes, hipMemcpyHostToDevice, s));
	item_off.resize(R + 1);
	MMD_HIP(hipMemcpyAsync(item_off.data(), staging + gbytes, 4 * (R + 1), hipMemcpyHostToDevice, s));
	const RepGeom* gd = reinterpret_cast<const RepGeom*>(geom.data());
	// sort: Morton key (stable; ties by canonical index), then replica (stable)
	keys.resize(total);
	vals.resize(total);
	keys_s.resize(total);
	vals_s.resize(total);
	rkeys.resize(total);
	rvals.resize(total);
	rkeys_s.resize(total);
	rvals_s.resize(total);
	codes.resize(total);
	fill_u32(scalars.data(), 0u, 8, s);
	k_b_encode<<<grid_for(total, 256), 256, 0, s>>>(pos[cur].data(), total, rep_of_slot.data(), gd, keys.data(),
													vals.data(), scalars.data());
	MMD_LAUNCH_CHECK();
	if (*std::max_element(count.begin(), count.end()) <= kSegSortMax) {
		k_b_seg_sort<<<unsigned(R), kSegThreads, 0, s>>>(gd, sort_order.data(), keys.data(), vals.data(),
														 rvals_s.data());
		MMD_LAUNCH_CHECK();
	} else {
		sort_scratch.resize(radix_sort_scratch_bytes(total));
		radix_sort_pairs(keys.data(), vals.data(), keys_s.data(), vals_s.data(), total, 30, sort_scratch.data(),
						 sort_scratch.count, s);
		k_b_rep_keys<<<grid_for(total, 256), 256, 0, s>>>(vals_s.data(), total, rep_of_slot.data(), rkeys.data(),
														  rvals.data());
		MMD_LAUNCH_CHECK();
		radix_sort_pairs(rkeys.data(), rvals.data(), rkeys_s.data(), rvals_s.data(), total, rep_bits,
						 sort_scratch.data(), sort_scratch.count, s);
	}
	const int nxt = 1 - cur;
	k_b_gather<<<grid_for(total, 256), 256, 0, s>>>(rvals_s.data(), total, rep_of_slot.data(), gd, pos[cur].data(),
													mom.data(), ids.data(), relabel, pos[nxt].data(), snap.data(),
													mom_tmp.data(), ids_tmp.data(), codes.data(),
													slot_of_canon_tmp.data());
	MMD_LAUNCH_CHECK();
	cur = nxt;
	mom.swap(mom_tmp);
	ids.swap(ids_tmp);
	slot_of_canon.swap(slot_of_canon_tmp);
	// cell tables
	fine_begin.resize(num_fine);
	fine_end.resize(num_fine);
	home_begin.resize(num_home);
	home_end.resize(num_home);
	fill_u32(fine_begin.data(), 0u, num_fine, s);
	fill_u32(fine_end.data(), 0u, num_fine, s);
	fill_u32(home_begin.data(), 0u, num_home, s);
	fill_u32(home_end.data(), 0u, num_home, s);
	k_b_cell_ranges<<<grid_for(total, 256), 256, 0, s>>>(codes.data(), total, rep_of_slot.data(), gd,
														 fine_begin.data(), fine_end.data(), home_begin.data(),
														 home_end.data(), scalars.data());
	MMD_LAUNCH_CHECK();
	// walk with uniform rows of `cap` entries; once more with larger rows if needed
	const uint64_t tiles = (total + 63) / 64;
	counts.resize(total);
	tile_base.resize(tiles + 1);
	if (cap == 0) {
		double vol = 0.0;
		for (uint64_t r = 0; r < R; ++r)
			vol += double(boxes[r].size[0]) * boxes[r].size[1] * boxes[r].size[2];
		const double mean = double(total) / vol * 4.0 / 3.0 * 3.141592653589793 * double(rc) * rc * rc;
		cap = uint32_t(std::min(1.6 * mean + 16.0, 4096.0) + 3) & ~3u;
	}
	const unsigned groups = unsigned((num_items + kHomeWaves - 1) / kHomeWaves);
	const unsigned grid = std::max(8u, (groups + 7u) / 8u * 8u);
	for (int attempt = 0; attempt < 2; ++attempt) {
		const uint64_t storage = tiles * 64 * uint64_t(cap);
		if (storage >= (uint64_t(1) << 32))
			raise(MARSMD_ERR_CAPACITY, "batch: neighbour storage exceeds 2^32 entries");
		if (storage > nbr.count)
			nbr.resize(storage + storage / 8 + 1024);
		k_b_uniform_tile_base<<<grid_for(tiles + 1, 256), 256, 0, s>>>(tile_base.data(), tiles, 64u * cap);
		MMD_LAUNCH_CHECK();
		k_b_walk<<<grid, 256, 0, s>>>(gd, item_off.data(), uint32_t(R), uint32_t(num_items), pos[cur].data(),
									  home_begin.data(), home_end.data(), fine_begin.data(), fine_end.data(), rc2,
									  cap, stencils.data(), counts.data(), nbr.data());
		MMD_LAUNCH_CHECK();
		fill_u32(scalars.data() + 2, 0u, 4, s);
		k_b_count_stats<<<std::min<unsigned>(grid_for(total, 256), 240), 256, 0, s>>>(
			counts.data(), total, reinterpret_cast<unsigned long long*>(scalars.data() + 4), scalars.data() + 2);
		MMD_LAUNCH_CHECK();
		// Pairlist::grow_pair_capacity (Pairlist.h:244-251) bookkeeping per replica, read
		// back with the walk's statistics: one wait for both (into pinned memory)
		k_b_rep_pairs<<<unsigned(R), 256, 0, s>>>(counts.data(), gd, rep_pairs.data());
		MMD_LAUNCH_CHECK();
		if (!readback)
			MMD_HIP(hipHostMalloc(reinterpret_cast<void**>(&readback), 32 + 8 * R, hipHostMallocDefault));
		MMD_HIP(hipMemcpyAsync(readback, scalars.data(), 24, hipMemcpyDeviceToHost, s));
		MMD_HIP(hipMemcpyAsync(readback + 4, rep_pairs.data(), 8 * R, hipMemcpyDeviceToHost, s));
		MMD_HIP(hipStreamSynchronize(s));
		uint32_t host[6];
		std::memcpy(host, readback, 24);
		if (host[0] & 1u)
			raise(MARSMD_ERR_VALUE, "batch: a particle left its replica's box (positions must stay wrapped)");
		if (host[0] & 2u)
			raise(MARSMD_ERR_RUNTIME, "batch: Morton codes after the sort were not in order");
		max_neighbors = host[2];
		std::memcpy(&nbr_entries, &host[4], 8);
		if (max_neighbors <= cap)
			break;
		cap = (max_neighbors + max_neighbors / 8 + 4 + 3) & ~3u;
	}
	// the pair counts set the reference grid of the next build
	std::memcpy(rep_pairs_host.data(), readback + 4, 8 * R);
	for (uint64_t r = 0; r < R; ++r) {
		const uint64_t np = rep_pairs_host[r] / 2;
		if (np > ref_max_pairs[r])
			ref_max_pairs[r] = np + np / 4;
	}
	built = true;
	built_cutoff = pairlist_cutoff;
}

bool BatchImpl::needs_rebuild(float skin, hipStream_t s) {
	fill_u32(scalars.data() + 1, 0u, 1, s);
	k_b_max_disp<<<std::min<unsigned>(grid_for(total, 256), 1024), 256, 0, s>>>(
		pos[cur].data(), snap.data(), total, rep_of_slot.data(), reinterpret_cast<const RepGeom*>(geom.data()),
		scalars.data() + 1);
	MMD_LAUNCH_CHECK();
	// into the pinned word after the walk's statistics (a build has allocated it)
	MMD_HIP(hipMemcpyAsync(readback + 3, scalars.data() + 1, 4, hipMemcpyDeviceToHost, s));
	MMD_HIP(hipStreamSynchronize(s));
	float d2;
	std::memcpy(&d2, readback + 3, 4);
	return std::sqrt(d2) > skin * 0.5f;
}

unsigned BatchImpl::step_grid(const void* kernel) {
	const unsigned lds = tables->lds ? unsigned(std::max(tables->total_knots, 2)) * 4u : 0u;
	if (occupancy == 0) {
		MMD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occupancy, kernel, kBatchBlock, lds));
		occupancy = std::max(occupancy, 1);
		int dev = 0;
		MMD_HIP(hipGetDevice(&dev));
		MMD_HIP(hipDeviceGetAttribute(&num_cus, hipDeviceAttributeMultiprocessorCount, dev));
		num_cus = std::max(num_cus, 1);
	}
	const int per_cu = blocks_per_cu > 0 ? std::min(blocks_per_cu, occupancy) : occupancy;
	const uint64_t tiles = (total + kBatchBlock - 1) / kBatchBlock;
	uint64_t g = std::min<uint64_t>(uint64_t(per_cu) * uint64_t(num_cus), (tiles + 7) / 8 * 8);
	return unsigned(std::max<uint64_t>((g + 7) / 8 * 8, 8));
}

void BatchImpl::set_launch(int bpc, bool graph) {
	if (bpc < 0)
		raise(MARSMD_ERR_VALUE, "batch: blocks_per_cu must be >= 0");
	blocks_per_cu = bpc;
	use_graph = graph;
}

hipGraphExec_t BatchImpl::segment_graph(const void* kernel, unsigned grid, unsigned lds, const StepParams& S,
										uint64_t count) {
	SegmentKey k;
	std::memset(&k, 0, sizeof(k));
	k.kernel = kernel;
	k.grid = grid;
	k.lds = lds;
	k.count = count;
	k.total = total;
	k.base_seed = base_seed;
	const void* ptrs[9] = {pos[cur].data(), pos[1 - cur].data(), mom.data(), force.data(), rep_of_slot.data(),
						   geom.data(), counts.data(), tile_base.data(), nbr.data()};
	std::memcpy(k.ptr, ptrs, sizeof(ptrs));
	k.S = S;
	k.S.step = 0;
	k.T = *tables;
	k.types = types;
	const uint8_t* kb = reinterpret_cast<const uint8_t*>(&k);
	std::vector<uint8_t> key(kb, kb + sizeof(k));
	for (const SegmentGraph& g : graphs)
		if (g.key == key)
			return g.exec;
	if (!capture_stream)
		MMD_HIP(hipStreamCreateWithFlags(&capture_stream, hipStreamNonBlocking));
	step_word.resize(1);
	const StepKernel kern = reinterpret_cast<StepKernel>(const_cast<void*>(kernel));
	const NbrList L{counts.data(), tile_base.data(), nbr.data()};
	MMD_HIP(hipStreamBeginCapture(capture_stream, hipStreamCaptureModeThreadLocal));
	hipError_t err = hipSuccess;
	for (uint64_t j = 0; j < count && err == hipSuccess; ++j) {
		StepParams Sj = S;
		Sj.step = uint32_t(j);
		const int c = (cur + int(j & 1)) & 1;
		kern<<<grid, kBatchBlock, lds, capture_stream>>>(pos[c].data(), pos[1 - c].data(), mom.data(), force.data(),
														  total, rep_of_slot.data(),
														  reinterpret_cast<const RepGeom*>(geom.data()), L, *tables,
														  knots.data(), types, Sj, base_seed, 0, step_word.data());
		err = hipGetLastError();
	}
	hipGraph_t graph = nullptr;
	const hipError_t end = hipStreamEndCapture(capture_stream, &graph);
	if (err != hipSuccess || end != hipSuccess) {
		if (graph)
			(void)hipGraphDestroy(graph);
		raise(MARSMD_ERR_DEVICE, std::string("batch: capturing a step segment failed: ") +
								  hipGetErrorString(err != hipSuccess ? err : end));
	}
	hipGraphExec_t exec = nullptr;
	const hipError_t inst = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
	(void)hipGraphDestroy(graph);
	if (inst != hipSuccess)
		raise(MARSMD_ERR_DEVICE, std::string("batch: instantiating a step segment failed: ") + hipGetErrorString(inst));
	if (graphs.size() >= 8) { // segments of other lengths or of an older list
		(void)hipGraphExecDestroy(graphs.front().exec);
		graphs.erase(graphs.begin());
	}
	graphs.push_back(SegmentGraph{std::move(key), exec});
	return exec;
}

void BatchImpl::launch_steps(const marsmd_run_params& P, uint64_t first, uint64_t count, bool store,
							 hipStream_t s) {
	const bool bd = P.integrator == MARSMD_INTEGRATOR_BROWNIAN;
	const bool energy = count == 1 && P.energy_period > 0 && first % P.energy_period == 0;
	const bool an = (tables->enabled & kAnalyticTerms) != 0;
	const StepKernel kern =
		bd ? (energy ? (an ? k_b_step<MARSMD_INTEGRATOR_BROWNIAN, true, true> : k_b_step<MARSMD_INTEGRATOR_BROWNIAN, true, false>)
					 : (an ? k_b_step<MARSMD_INTEGRATOR_BROWNIAN, false, true>
						   : k_b_step<MARSMD_INTEGRATOR_BROWNIAN, false, false>))
		   : (energy ? (an ? k_b_step<MARSMD_INTEGRATOR_LANGEVIN, true, true> : k_b_step<MARSMD_INTEGRATOR_LANGEVIN, true, false>)
					 : (an ? k_b_step<MARSMD_INTEGRATOR_LANGEVIN, false, true>
						   : k_b_step<MARSMD_INTEGRATOR_LANGEVIN, false, false>));
	const unsigned grid = step_grid(reinterpret_cast<const void*>(kern));
	const unsigned lds = tables->lds ? unsigned(std::max(tables->total_knots, 2)) * 4u : 0u;
	StepParams S{};
	S.cut_d = cutoff_threshold(P.cutoff > 0.0f ? P.cutoff * P.cutoff : 0.0f);
	S.dt = P.timestep;
	S.kT = P.kT;
	S.ctr = 0;
	S.scheme = MARSMD_INTERP_LINEAR;
	S.has_grids = 0;
	S.deferred = deferred_kick ? 1 : 0;
	// a segment replays when the launches differ only in their step: no forces kept,
	// and the deferred-kick state settled (Langevin defers from its first step on)
	const bool replay = use_graph && count >= 2 && !store && (bd || deferred_kick);
	hipGraphExec_t exec = replay ? segment_graph(reinterpret_cast<const void*>(kern), grid, lds, S, count) : nullptr;
	if (timing) {
		while (timing_events.size() < timing_used + 2) {
			hipEvent_t e;
			MMD_HIP(hipEventCreate(&e));
			timing_events.push_back(e);
		}
		MMD_HIP(hipEventRecord(timing_events[timing_used], s));
	}
	if (exec) {
		k_b_set_word<<<1, 64, 0, s>>>(step_word.data(), uint32_t(first));
		MMD_LAUNCH_CHECK();
		MMD_HIP(hipGraphLaunch(exec, s));
		if (count & 1)
			cur = 1 - cur;
	} else {
		const NbrList L{counts.data(), tile_base.data(), nbr.data()};
		for (uint64_t j = 0; j < count; ++j) {
			S.step = uint32_t(first + j);
			S.deferred = deferred_kick ? 1 : 0;
			const int keep = store && j + 1 == count ? 1 : 0;
			kern<<<grid, kBatchBlock, lds, s>>>(pos[cur].data(), pos[1 - cur].data(), mom.data(), force.data(), total,
												 rep_of_slot.data(), reinterpret_cast<const RepGeom*>(geom.data()), L,
												 *tables, knots.data(), types, S, base_seed, keep, nullptr);
			MMD_LAUNCH_CHECK();
			if (!bd)
				deferred_kick = true;
			cur = 1 - cur;
		}
	}
	if (timing) {
		MMD_HIP(hipEventRecord(timing_events[timing_used + 1], s));
		timing_used += 2;
		timing_steps.push_back(count);
	}
}

void BatchImpl::run(const marsmd_run_params& P, uint64_t first_step, uint64_t num_steps, hipStream_t s) {
	if (P.integrator != MARSMD_INTEGRATOR_BROWNIAN && P.integrator != MARSMD_INTEGRATOR_LANGEVIN)
		raise(MARSMD_ERR_VALUE, "Unsupported integrator type: " + std::to_string(P.integrator));
	const uint64_t rp = P.reorder_period > 0 ? uint64_t(P.reorder_period) : P.rebuild_period;
	const uint64_t last = first_step + num_steps;
	// a step needs the host when it may reorder or rebuild or sums energies
	auto host_step = [&](uint64_t step) {
		return (rp > 0 && (step - 1) % rp == 0) || P.rebuild_period == 0 ||
			   (step - 1) % P.rebuild_period == 0 || (P.energy_period > 0 && step % P.energy_period == 0);
	};
	uint64_t step = first_step + 1;
	while (step <= last) {
		// maybe_reorder (ReorderManager.cpp:8-25): the reorder forces a rebuild
		const bool reorder = rp > 0 && (step - 1) % rp == 0 && prepared;
		prepared = true;
		// rebuild policy (Patch.cpp:85-121), over all replicas
		bool rebuild = !built || reorder || built_cutoff != P.pairlist_cutoff;
		if (!rebuild && (P.rebuild_period == 0 || (step - 1) % P.rebuild_period == 0)) {
			const float skin = P.pairlist_cutoff - P.cutoff;
			rebuild = P.rebuild_period == 0 || skin <= 0.0f || needs_rebuild(skin, s);
		}
		if (rebuild)
			build(P.pairlist_cutoff, reorder, s);
		launch_steps(P, step, 1, step == last, s);
		++step;
		// then every step up to the next one the host decides on, and not the run's last
		// (which keeps its forces): the same list, no reorder, no energies
		uint64_t end = step;
		while (end < last && !host_step(end))
			++end;
		if (end > step) {
			launch_steps(P, step, end - step, false, s);
			step = end;
		}
	}
}

void BatchImpl::autotune(const marsmd_run_params& P, uint64_t first_step, uint64_t num_steps, hipStream_t s,
						 double* best_ms) {
	// candidates: workgroups per CU from the occupancy down, each with and without
	// segment replay; results do not depend on them (each particle's step is one lane's
	// work, and a replay launches the same kernels with the same arguments)
	(void)step_grid(reinterpret_cast<const void*>(k_b_step<MARSMD_INTEGRATOR_BROWNIAN, false, false>));
	std::vector<int> per_cu = {occupancy};
	for (int v : {occupancy - 1, occupancy / 2, 1})
		if (v >= 1 && std::find(per_cu.begin(), per_cu.end(), v) == per_cu.end())
			per_cu.push_back(v);
	struct Cand {
		int bpc;
		bool graph;
		double ms = 1e300;
	};
	std::vector<Cand> cands;
	for (int v : per_cu)
		for (bool g : {false, true})
			cands.push_back(Cand{v, g});
	const bool was_timing = timing;
	const uint64_t passes = 2;
	const uint64_t chunk = num_steps / (passes * cands.size());
	uint64_t step = first_step;
	if (chunk >= 2) {
		for (uint64_t pass = 0; pass < passes; ++pass)
			for (Cand& c : cands) {
				set_launch(c.bpc, c.graph);
				set_kernel_timing(true);
				run(P, step, chunk, s);
				step += chunk;
				// the runs of steps without host decisions; every launch if there are none
				double ms = 0.0, ms_all = 0.0;
				uint64_t n = 0, n_all = 0;
				for (size_t k = 0; k + 1 < timing_used; k += 2) {
					MMD_HIP(hipEventSynchronize(timing_events[k + 1]));
					float t = 0.0f;
					MMD_HIP(hipEventElapsedTime(&t, timing_events[k], timing_events[k + 1]));
					const uint64_t steps = timing_steps[k / 2];
					ms_all += t;
					n_all += steps;
					if (steps >= 2) {
						ms += t;
						n += steps;
					}
				}
				const double per = n ? ms / double(n) : (n_all ? ms_all / double(n_all) : 1e300);
				c.ms = std::min(c.ms, per);
			}
	}
	const Cand* best = &cands.front();
	for (const Cand& c : cands)
		if (c.ms < best->ms)
			best = &c;
	set_launch(best->bpc, best->graph);
	set_kernel_timing(was_timing);
	if (step < first_step + num_steps)
		run(P, step, first_step + num_steps - step, s);
	*best_ms = best->ms < 1e300 ? best->ms : 0.0;
}

void BatchImpl::set_cells(const marsmd_batch_cells& c) {
	for (int b = 0; b < kBatchBuckets; ++b) {
		if (c.home_target[b] < 1 || c.home_target[b] > 4096)
			raise(MARSMD_ERR_VALUE, "batch: home_target must be in [1, 4096]");
		if (c.fine_offset[b] < 1 || c.fine_offset[b] > 4)
			raise(MARSMD_ERR_VALUE, "batch: fine_offset must be in [1, 4]");
	}
	for (int b = 0; b < kBatchBuckets; ++b) {
		cell_target[size_t(b)] = c.home_target[b];
		cell_fine[size_t(b)] = c.fine_offset[b];
	}
}

void BatchImpl::get_cells(marsmd_batch_cells* c) const {
	for (int b = 0; b < kBatchBuckets; ++b) {
		c->home_target[b] = cell_target[size_t(b)];
		c->fine_offset[b] = cell_fine[size_t(b)];
	}
}

double BatchImpl::walk_ms(hipStream_t s) {
	std::vector<uint8_t> gh;
	std::vector<uint32_t> items;
	uint64_t ft = 0, ht = 0;
	plan_geometry(built_cutoff, gh, items, ft, ht);
	t_geom.resize(R * sizeof(RepGeom));
	MMD_HIP(hipMemcpyAsync(t_geom.data(), gh.data(), R * sizeof(RepGeom), hipMemcpyHostToDevice, s));
	t_items.resize(R + 1);
	MMD_HIP(hipMemcpyAsync(t_items.data(), items.data(), 4 * (R + 1), hipMemcpyHostToDevice, s));
	t_fine_begin.resize(ft);
	t_fine_end.resize(ft);
	t_home_begin.resize(ht);
	t_home_end.resize(ht);
	t_counts.resize(total);
	const uint64_t storage = (total + 63) / 64 * 64 * uint64_t(cap);
	t_nbr.resize(std::max<uint64_t>(storage, 1));
	if (!t_events[0]) {
		MMD_HIP(hipEventCreate(&t_events[0]));
		MMD_HIP(hipEventCreate(&t_events[1]));
	}
	const RepGeom* gd = reinterpret_cast<const RepGeom*>(t_geom.data());
	const uint64_t nitems = items[R];
	const unsigned groups = unsigned((nitems + kHomeWaves - 1) / kHomeWaves);
	const unsigned grid = std::max(8u, (groups + 7u) / 8u * 8u);
	MMD_HIP(hipEventRecord(t_events[0], s));
	fill_u32(t_fine_begin.data(), 0u, ft, s);
	fill_u32(t_fine_end.data(), 0u, ft, s);
	fill_u32(t_home_begin.data(), 0u, ht, s);
	fill_u32(t_home_end.data(), 0u, ht, s);
	k_b_cell_ranges<<<grid_for(total, 256), 256, 0, s>>>(codes.data(), total, rep_of_slot.data(), gd,
														 t_fine_begin.data(), t_fine_end.data(), t_home_begin.data(),
														 t_home_end.data(), nullptr);
	MMD_LAUNCH_CHECK();
	k_b_walk<<<grid, 256, 0, s>>>(gd, t_items.data(), uint32_t(R), uint32_t(nitems), snap.data(), t_home_begin.data(),
								  t_home_end.data(), t_fine_begin.data(), t_fine_end.data(),
								  built_cutoff * built_cutoff, cap, stencils.data(), t_counts.data(), t_nbr.data());
	MMD_LAUNCH_CHECK();
	MMD_HIP(hipEventRecord(t_events[1], s));
	MMD_HIP(hipEventSynchronize(t_events[1]));
	float ms = 0.0f;
	MMD_HIP(hipEventElapsedTime(&ms, t_events[0], t_events[1]));
	return double(ms);
}

void BatchImpl::autotune_cells(const marsmd_run_params& P, uint64_t first_step, uint64_t num_steps, hipStream_t s,
							   double* walk_out) {
	uint64_t step = first_step;
	if (!built && num_steps > 0) { // the first build
		run(P, step, 1, s);
		++step;
	}
	if (!built)
		raise(MARSMD_ERR_RUNTIME, "batch: the cell autotuner needs a first build (num_steps >= 1)");
	std::vector<int> present;
	for (uint64_t r = 0; r < R; ++r) {
		const int b = batch_size_bucket(count[r]);
		if (std::find(present.begin(), present.end(), b) == present.end())
			present.push_back(b);
	}
	std::sort(present.begin(), present.end());
	// the fastest of three timings of one configuration
	const auto measure = [&]() {
		double best = 1e300;
		for (int k = 0; k < 3; ++k)
			best = std::min(best, walk_ms(s));
		return best;
	};
	(void)walk_ms(s); // warm the scratch buffers and stencils
	double best = measure();
	const int targets[] = {8, 16, 24, 48, 96};
	const int fines[] = {1, 2, 3};
	// one bucket at a time, the others held; a candidate must beat the incumbent by 1%
	for (int b : present) {
		int bt = cell_target[size_t(b)], bf = cell_fine[size_t(b)];
		for (int t : targets)
			for (int f : fines) {
				if (t == bt && f == bf)
					continue;
				cell_target[size_t(b)] = t;
				cell_fine[size_t(b)] = f;
				const double ms = measure();
				if (ms < 0.99 * best) {
					best = ms;
					bt = t;
					bf = f;
				}
			}
		cell_target[size_t(b)] = bt;
		cell_fine[size_t(b)] = bf;
	}
	*walk_out = best < 1e300 ? best : 0.0;
	if (step < first_step + num_steps)
		run(P, step, first_step + num_steps - step, s);
}

void BatchImpl::finish_deferred_kick(float dt, hipStream_t s) {
	if (!deferred_kick)
		return;
	deferred_kick = false;
	k_b_last_update<<<grid_for(total, 256), 256, 0, s>>>(mom.data(), force.data(), total, dt);
	MMD_LAUNCH_CHECK();
}

const int32_t* BatchImpl::neighbor_pairs(uint64_t replica, uint64_t* np_out, hipStream_t s) {
	if (replica >= R)
		raise(MARSMD_ERR_VALUE, "batch: replica index out of range");
	if (!built)
		raise(MARSMD_ERR_RUNTIME, "batch: no pairlist has been built yet");
	const uint64_t n = count[replica], slot0 = off[replica];
	half_counts.resize(n + 1);
	half_offsets.resize(n + 1);
	scan_scratch.resize(scan_scratch_bytes(n + 1));
	const NbrList L{counts.data(), tile_base.data(), nbr.data()};
	uint32_t np = 0;
	if (n) {
		k_b_half_counts<<<grid_for(n, 256), 256, 0, s>>>(L, slot0, n, half_counts.data());
		MMD_LAUNCH_CHECK();
		exclusive_scan_u32(half_counts.data(), half_offsets.data(), n, scan_scratch.data(), s);
		MMD_HIP(hipMemcpyAsync(&np, half_offsets.data() + n, 4, hipMemcpyDeviceToHost, s));
		MMD_HIP(hipStreamSynchronize(s));
		pairs.resize(2 * uint64_t(np) + 2);
		k_b_write_pairs<<<grid_for(n, 256), 256, 0, s>>>(L, pos[cur].data(), slot0, n, half_offsets.data(),
														 pairs.data());
		MMD_LAUNCH_CHECK();
	} else {
		pairs.resize(2);
	}
	*np_out = np;
	return pairs.data();
}

void BatchImpl::set_kernel_timing(bool enable) {
	timing = enable;
	timing_used = 0;
	timing_steps.clear();
}

void BatchImpl::kernel_timing(double* total_ms, uint64_t* steps) {
	double sum = 0.0;
	uint64_t n = 0;
	for (size_t k = 0; k + 1 < timing_used; k += 2) {
		MMD_HIP(hipEventSynchronize(timing_events[k + 1]));
		float ms = 0.0f;
		MMD_HIP(hipEventElapsedTime(&ms, timing_events[k], timing_events[k + 1]));
		sum += double(ms);
		n += timing_steps[k / 2];
	}
	*total_ms = sum;
	*steps = n;
}

} // namespace mmd
