// patch.hip — the Patch of the reference (src/PatchOperation/Patch.cpp:43-480) on one
// MI355X: particle state resident in HBM in Morton (sorted) slot order, the
// rebuild / reorder policy on the host, and per-particle kernels.
//
// Layout in HBM (slot order = sorted order of the last pairlist build):
//   pos[2]  float4 {x, y, z, meta}; meta = canonical index << 8 | type. Two
//           buffers so the fused step reads step-n positions while writing n+1.
//   mom     float4 {px, py, pz, 0}            (Langevin)
//   force   float4 {fx, fy, fz, energy}       (ForceEnergy; bonded scatter target)
//   ids     int32 global id per slot; slot_of_canon: canonical index -> slot.
// The canonical index is the reference's slot index: it is the Philox counter of a
// particle and the index space of bonds, exclusions and exported arrays, and it is
// relabelled at each reorder exactly as Patch::reorder_particles permutes the SoA.
//
// Kernels:
//   k_nb_forces      nonbonded force from the CSR full list, written (not added)
//                    to `force` — calculate_nonbonded_forces.
//   k_integrate_*    BD / PMF / BAOAB last-update / BAOAB on slot order, Philox
//                    counter = canonical index — integrate_motion.
//   k_step<...>      nonbonded force + bonded force (read from `force`) + PMF +
//                    integrator in one pass; positions double-buffered. Used by run().
// The per-pair term is evaluated with the reference's operation order (canonical
// orientation: r = wrap_diff(pos[max canon] - pos[min canon])), so it is bitwise
// the term the reference's PairNonbondedComputer adds; only the summation order of
// a particle's terms differs (fixed here: its neighbour-list order).
#include "launchers.hpp"
#include "nbforce.hpp"
#include "step.hpp"
#include "pairlist.hpp"
#include "patch.hpp"

#include <algorithm>
#include <cstring>

namespace mmd {

namespace {

constexpr int kBlock = 256;


template<bool ANALYTIC>
__global__ __launch_bounds__(kBlock) void k_nb_forces(const float4* __restrict__ pos, uint64_t n, NbrList L,
													  NbTables T, const float* __restrict__ knots_g, DevBox box,
													  float cut_d, int energy, float4* __restrict__ force) {
	extern __shared__ float lds_knots[];
	__shared__ NbLds lds_nb;
	stage_tables(T, knots_g, lds_knots, lds_nb);
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	force[t] = nb_force<ANALYTIC>(uint32_t(t), pos[t], pos, L, T, lds_knots, knots_g, lds_nb, box, cut_d, energy != 0);
}

__global__ void k_zero_force(float4* __restrict__ f, uint64_t n) {
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t < n)
		f[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__global__ __launch_bounds__(kBlock) void k_integrate_bd(float4* __restrict__ pos, float4* __restrict__ force,
														 uint64_t n, marsmd_particle_types types,
														 const marsmd_grid_view* __restrict__ grids, StepParams S) {
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	float4 f = force[t];
	const float4 p0 = pos[t];
	const float4 pmf = position_dependent_force(p0.x, p0.y, p0.z, int(meta_type(fb(p0.w))), types,
												S.has_grids ? grids : nullptr, S.ef, S.scheme);
	const float4 p = bd_update(p0, f, types, pmf, S);
	pos[t] = p;
	force[t].w = f.w;
}

// The position-dependent force of every particle (compute_position_dependent_force,
// Pmf.h:68-124: grid terms and the electric field) for a step kernel that reads it
// instead of sampling the grids itself: the same function of the same inputs, so the
// step's results are the same bits.
__global__ __launch_bounds__(kBlock) void k_pmf_force(const float4* __restrict__ pos, uint64_t n,
													  marsmd_particle_types types_arg,
													  const marsmd_grid_view* __restrict__ grids, StepParams S,
													  float4* __restrict__ out) {
	__shared__ marsmd_particle_types lds_types;
	stage_types(types_arg, lds_types);
	__syncthreads();
	const marsmd_particle_types& types = lds_types;
	// XCD-local tiles (a contiguous eighth of the Morton order per XCD): each XCD's L2
	// then holds one region's grid lines, not lines from all over the grid
	const TileRange R = xcd_tiles(uint32_t((n + kBlock - 1) / kBlock));
	for (uint32_t tile = R.tile; tile < R.end; tile += R.stride) {
		const uint64_t t = uint64_t(tile) * kBlock + threadIdx.x;
		if (t >= n)
			continue;
		const float4 p = pos[t];
		out[t] = position_dependent_force(p.x, p.y, p.z, int(meta_type(fb(p.w))), types,
										  S.has_grids ? grids : nullptr, S.ef, S.scheme);
	}
}

__global__ __launch_bounds__(kBlock) void k_integrate_pmf(const float4* __restrict__ pos, float4* __restrict__ force,
														  uint64_t n, marsmd_particle_types types,
														  const marsmd_grid_view* __restrict__ grids, StepParams S) {
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	const float4 p = pos[t];
	const float4 g = position_dependent_force(p.x, p.y, p.z, int(meta_type(fb(p.w))), types, grids, S.ef, S.scheme);
	float4 f = force[t];
	f.x += g.x;
	f.y += g.y;
	f.z += g.z;
	f.w += g.w;
	force[t] = f;
}

__global__ __launch_bounds__(kBlock) void k_integrate_last(float4* __restrict__ mom, const float4* __restrict__ force,
														   uint64_t n, float dt) {
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= n)
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

__global__ __launch_bounds__(kBlock) void k_integrate_baoab(float4* __restrict__ pos, float4* __restrict__ mom,
															const float4* __restrict__ force, uint64_t n,
															marsmd_particle_types types, StepParams S) {
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	float4 mo = mom[t];
	const float4 p = baoab_update(pos[t], mo, force[t], types, S);
	pos[t] = p;
	mom[t] = mo;
}

// ------------------------------------------------------------------ pruned rows
// The fused step's view of the pruned rows (k_prune). Lane position t of a step tile
// takes slot lane_slot[t] and that slot's pruned row list.{count, tile_base, nbr}[t].
// A pruned row keeps, in list order, the entries that lay within cutoff + 2 margin at
// the prune; an entry it drops is farther than the cutoff at every step at which no
// particle has moved a margin from its position at the prune, and such an entry adds
// exactly +-0 to the sums (the step kernels' tabulated path; the analytical path skips
// it), so a pruned row gives the full row's bits. disp[k & 3] records the largest
// squared displacement from the prune after fused step k of the prune's life (k = 0:
// the first step after it), disp[4] the largest over the life; a step whose previous
// record reaches limit2 = (margin - slack)^2 walks the full rows instead. limit2 = 0:
// rows sorted only (every entry kept, no record).
struct PruneView {
	NbrList list;
	const uint32_t* lane_slot; // NULL: no pruned rows (lane position = slot)
	const float4* ref;         // positions at the prune
	uint32_t* disp;
	uint32_t since;            // fused steps since the prune before this one
	float limit2;
};
constexpr uint32_t kPruneBins = 64; // rows by group count (0..62), dead lanes last

// Whether entry q of the row of p is kept (squared distance within keep2).
__device__ __forceinline__ bool prune_keep(const DevBox& box, const float4& p, const float4& q, float keep2) {
	return len2(wrap_diff_uniform(box, sub3(q, p))) <= keep2;
}

// Pruned rows of the 256-slot block blockIdx.x (see PruneView): each slot counts the
// entries of its row to keep (squared distance <= keep2, or all when keep_all), reading
// the row four entries (one 16-byte load) and four positions at a time and keeping a
// bit per entry (the first 128; later ones are tested again when written). With sort,
// the slots take lane positions by ascending group count (ties: lower wavefront, then
// lower lane first), else their own; each writes its kept entries in row order at its
// lane position, padding the last group with itself (a zero distance, skipped by the
// force kernels). The positions become the displacement reference; block 0 clears the
// record.
__global__ __launch_bounds__(256) void k_prune(const float4* __restrict__ pos, uint32_t n, NbrList L, DevBox box,
											   float keep2, int keep_all, int sort, uint32_t cap,
											   uint32_t* __restrict__ lane_slot, uint32_t* __restrict__ p_count,
											   uint32_t* __restrict__ p_tile_base, uint32_t* __restrict__ p_nbr,
											   float4* __restrict__ ref, uint32_t* __restrict__ disp) {
	__shared__ uint32_t hist[4][kPruneBins];
	__shared__ uint32_t base[kPruneBins];
	const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
	const uint32_t s = blockIdx.x * 256u + tid;
	const bool live = s < n;
	if (blockIdx.x == 0 && tid < kPruneDispWords)
		disp[tid] = 0u;
	if (tid < 4u)
		p_tile_base[blockIdx.x * 4u + tid] = (blockIdx.x * 4u + tid) * 64u * cap;
	for (uint32_t i = tid; i < 4u * kPruneBins; i += 256u)
		(&hist[0][0])[i] = 0u;
	float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	uint32_t cnt = 0, kept = 0;
	uint64_t keep_lo = 0, keep_hi = 0; // entries 0..63, 64..127
	const uint4* __restrict__ row = nullptr;
	if (live) {
		p = pos[s];
		ref[s] = p;
		cnt = L.count[s];
		row = reinterpret_cast<const uint4*>(L.nbr + L.tile_base[s >> 6]) + (s & 63u);
		if (keep_all) {
			kept = cnt;
		} else {
			for (uint32_t g = 0; 4u * g < cnt; ++g) {
				const uint4 j = row[uint64_t(g) * 64u];
				const float4 q0 = pos[j.x], q1 = pos[j.y], q2 = pos[j.z], q3 = pos[j.w];
				// the pads past the count are the slot itself: never kept
				uint32_t m4 = (prune_keep(box, p, q0, keep2) ? 1u : 0u) | (prune_keep(box, p, q1, keep2) ? 2u : 0u) |
							  (prune_keep(box, p, q2, keep2) ? 4u : 0u) | (prune_keep(box, p, q3, keep2) ? 8u : 0u);
				const uint32_t valid = cnt - 4u * g >= 4u ? 15u : (1u << (cnt - 4u * g)) - 1u;
				m4 &= valid;
				kept += __popc(m4);
				if (g < 16u)
					keep_lo |= uint64_t(m4) << (4u * g);
				else if (g < 32u)
					keep_hi |= uint64_t(m4) << (4u * (g - 16u));
			}
		}
	}
	uint32_t at = tid;
	if (sort) {
		const uint32_t bin = live ? min((kept + 3u) >> 2, kPruneBins - 2u) : kPruneBins - 1u;
		__syncthreads();
		// rank among the wavefront's lanes of the same bin, lower lanes first
		uint32_t rank = 0;
		uint64_t todo = __ballot(1);
		while (todo) {
			const int leader = __ffsll(static_cast<unsigned long long>(todo)) - 1;
			const uint32_t b = __shfl(bin, leader);
			const uint64_t same = __ballot(bin == b);
			if (bin == b) {
				rank = __popcll(same & ((1ull << lane) - 1ull));
				if (int(lane) == leader)
					hist[wave][b] = __popcll(same);
			}
			todo &= ~same;
		}
		__syncthreads();
		if (tid == 0) {
			uint32_t a = 0;
			for (uint32_t b = 0; b < kPruneBins; ++b) {
				base[b] = a;
				a += hist[0][b] + hist[1][b] + hist[2][b] + hist[3][b];
			}
		}
		__syncthreads();
		at = base[bin] + rank;
		for (uint32_t w = 0; w < wave; ++w)
			at += hist[w][bin];
	}
	if (!live)
		return; // dead lanes sort last: every live slot's lane position is below n
	const uint32_t t = blockIdx.x * 256u + at;
	lane_slot[t] = s;
	p_count[t] = kept;
	const uint64_t pb = uint64_t(t >> 6) * 64u * cap;
	uint32_t m = 0;
	for (uint32_t g = 0; 4u * g < cnt; ++g) {
		const uint4 j4 = row[uint64_t(g) * 64u];
		const uint32_t jv[4] = {j4.x, j4.y, j4.z, j4.w};
#pragma unroll
		for (int e = 0; e < 4; ++e) {
			const uint32_t k = 4u * g + uint32_t(e);
			if (k >= cnt)
				break;
			const bool keep = keep_all || (k < 64u	 ? ((keep_lo >> k) & 1u) != 0u
											: k < 128u ? ((keep_hi >> (k - 64u)) & 1u) != 0u
													   : prune_keep(box, p, pos[jv[e]], keep2));
			if (keep)
				p_nbr[nbr_slot_index(pb, t, m++)] = jv[e];
		}
	}
	for (; (m & 3u) != 0u; ++m)
		p_nbr[nbr_slot_index(pb, t, m)] = s;
}

// The fused step: nonbonded + bonded (from `force` when has_bonded) + PMF + integrator,
// persistent over XCD-local tiles so the table knots are staged once per workgroup.
// Bounded to 128 VGPRs (4 waves per SIMD; the tricubic sampler alone would take ~140
// and leave 3): measured faster despite a few dwords of spill.
template<int INTEG, bool ENERGY, bool ANALYTIC, bool PMF_IN>
__global__ __launch_bounds__(kBlock, 4) void k_step(const float4* __restrict__ pos, float4* __restrict__ pos_next,
												 float4* __restrict__ mom, float4* __restrict__ force, uint64_t n,
												 NbrList L, NbTables T, const float* __restrict__ knots_g,
												 marsmd_particle_types types_arg,
												 const marsmd_grid_view* __restrict__ grids, StepParams S,
												 int has_bonded, int store_force, const float4* __restrict__ pmf_in,
												 PruneView PR) {
	extern __shared__ float lds_knots[];
	__shared__ NbLds lds_nb;
	__shared__ marsmd_particle_types lds_types;
	__shared__ float lds_red[kBlock / 64];
	__shared__ TypeStep lds_steps[MARSMD_MAX_TYPES]; // the integrator's per-type constants
	stage_types(types_arg, lds_types);
	if (threadIdx.x < uint32_t(types_arg.num_types))
		lds_steps[threadIdx.x] = INTEG == MARSMD_INTEGRATOR_BROWNIAN ? bd_type_step(types_arg, int(threadIdx.x), S)
																	 : baoab_type_step(types_arg, int(threadIdx.x), S);
	stage_tables(T, knots_g, lds_knots, lds_nb);
	const marsmd_particle_types& types = lds_types;
	// the pruned rows hold while the previous step's displacement record is below the
	// limit; else this step walks the full rows (the same bits, more entries)
	const bool track = PR.lane_slot && PR.limit2 > 0.0f;
	bool pruned = PR.lane_slot != nullptr;
	if (track && PR.since > 0)
		pruned = __uint_as_float(PR.disp[(PR.since - 1u) & 3u]) < PR.limit2;
	const NbrList LL = pruned ? PR.list : L;
	float disp2 = 0.0f;
	// the rest of slot t's step once its nonbonded force f is summed (kernel arguments
	// captured by value: by reference the register allocation around the grid sampler
	// spills)
	const auto finish = [&, pos, pos_next, mom, force, grids, S, has_bonded, store_force, pmf_in, PR](uint32_t t,
																										 float4 f) {
		const float4 p = pos[t];
		if (has_bonded) {
			const float4 b = force[t];
			f.x += b.x;
			f.y += b.y;
			f.z += b.z;
			f.w += b.w;
		}
		float4 pn;
		if (INTEG == MARSMD_INTEGRATOR_BROWNIAN) {
			const float4 pmf = PMF_IN ? pmf_in[t]
									  : position_dependent_force(p.x, p.y, p.z, int(meta_type(fb(p.w))), types,
																 S.has_grids ? grids : nullptr, S.ef, S.scheme);
			pn = bd_update(p, f, lds_steps[meta_type(fb(p.w))], pmf, S);
		} else {
			if (S.has_grids) {
				const float4 g = PMF_IN ? pmf_in[t]
										: position_dependent_force(p.x, p.y, p.z, int(meta_type(fb(p.w))), types,
																   grids, S.ef, S.scheme);
				f.x += g.x;
				f.y += g.y;
				f.z += g.z;
				f.w += g.w;
			}
			float4 mo = mom[t];
			pn = baoab_update(p, mo, f, lds_steps[meta_type(fb(p.w))], S);
			mom[t] = mo;
		}
		pos_next[t] = pn;
		if (ENERGY || store_force)
			force[t] = f;
		if (track)
			disp2 = fmaxf(disp2, len2(wrap_diff_uniform(S.box, sub3(pn, PR.ref[t]))));
	};
	const TileRange R = xcd_tiles(uint32_t((n + kBlock - 1) / kBlock));
	for (uint32_t tile = R.tile; tile < R.end; tile += R.stride) {
		const uint32_t lt = tile * uint32_t(kBlock) + threadIdx.x; // lane position
		if (lt >= n)
			continue;
		const uint32_t t = PR.lane_slot ? PR.lane_slot[lt] : lt; // slot
		finish(t, nb_force<ANALYTIC>(pruned ? lt : t, pos[t], pos, LL, T, lds_knots, knots_g, lds_nb, S.box, S.cut_d,
									 ENERGY));
	}
	if (track) {
		// the workgroup's largest squared displacement (non-negative floats order as
		// their bits) into this step's record and the life's
		for (int o = 32; o > 0; o >>= 1)
			disp2 = fmaxf(disp2, __shfl_xor(disp2, o));
		if ((threadIdx.x & 63u) == 0u)
			lds_red[threadIdx.x >> 6] = disp2;
		__syncthreads();
		if (threadIdx.x == 0) {
			float m = lds_red[0];
			for (int w = 1; w < kBlock / 64; ++w)
				m = fmaxf(m, lds_red[w]);
			atomicMax(&PR.disp[PR.since & 3u], __float_as_uint(m));
			atomicMax(&PR.disp[4], __float_as_uint(m));
			if (blockIdx.x == 0)
				PR.disp[(PR.since + 1u) & 3u] = 0u; // the next step's record starts empty
		}
	}
}

using StepFn = decltype(&k_step<MARSMD_INTEGRATOR_BROWNIAN, false, false, false>);
template<int I, bool E, bool A>
StepFn pick_step_pmf(bool pmf_in) {
	return pmf_in ? k_step<I, E, A, true> : k_step<I, E, A, false>;
}
template<int I, bool E>
StepFn pick_step_an(bool an, bool pmf_in) {
	return an ? pick_step_pmf<I, E, true>(pmf_in) : pick_step_pmf<I, E, false>(pmf_in);
}
template<int I>
StepFn pick_step_e(bool energy, bool an, bool pmf_in) {
	return energy ? pick_step_an<I, true>(an, pmf_in) : pick_step_an<I, false>(an, pmf_in);
}
// The fused step kernel for an integrator, energy step, analytical terms and whether
// the position-dependent force comes precomputed.
StepFn pick_step(bool bd, bool energy, bool an, bool pmf_in) {
	return bd ? pick_step_e<MARSMD_INTEGRATOR_BROWNIAN>(energy, an, pmf_in)
			  : pick_step_e<MARSMD_INTEGRATOR_LANGEVIN>(energy, an, pmf_in);
}

// Permute slot-indexed state into the new sorted order:
// out[t] = in[old_slot_of_canon[sorted_canon[t]]].
// The slot-order state after a build: sorted positions, momenta and ids gathered from
// the old slots, and the new canonical -> slot map (into slot_out: old_slot is read).
__global__ void k_gather_state(const uint32_t* __restrict__ sidx, const uint32_t* __restrict__ old_slot, uint64_t n,
							   const float4* __restrict__ spos, float4* __restrict__ pos_out,
							   const uint32_t* __restrict__ inverse, uint32_t* __restrict__ slot_out,
							   const float4* __restrict__ mom_in, float4* __restrict__ mom_out,
							   const int32_t* __restrict__ ids_in, int32_t* __restrict__ ids_out) {
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	pos_out[t] = spos[t];
	slot_out[t] = inverse[t];
	const uint32_t s = old_slot[sidx[t]];
	if (s >= n)
		return;
	mom_out[t] = mom_in[s];
	ids_out[t] = ids_in[s];
}

// canonical -> slot remap of bonded index tuples (stride ints per term, first k used).
// ------------------------------------------------------------------ bonded gather
// The bonded terms' shares (the launchers' contribution slots: bonds 2 per term, then
// angles 3, dihedrals 4, restraints 1) as (slot of the particle, share index) pairs,
// for a stable sort by slot: each particle's shares then sit in share-index order.
__global__ void k_share_keys(const int32_t* __restrict__ bonds, uint64_t nb, const int32_t* __restrict__ angles,
							 uint64_t na, const int32_t* __restrict__ dih, uint64_t nd,
							 const int32_t* __restrict__ restr, uint64_t nr, uint32_t* __restrict__ keys,
							 uint32_t* __restrict__ vals) {
	const uint64_t q = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	const uint64_t a0 = 2 * nb, d0 = a0 + 3 * na, r0 = d0 + 4 * nd, total = r0 + nr;
	if (q >= total)
		return;
	int32_t slot;
	if (q < a0)
		slot = bonds[q];
	else if (q < d0)
		slot = angles[4 * ((q - a0) / 3) + (q - a0) % 3];
	else if (q < r0)
		slot = dih[q - d0];
	else
		slot = restr[q - r0];
	// a restraint without a particle (id < 0) has a zero share: sort it last
	keys[q] = slot < 0 ? 0xffffffffu : uint32_t(slot);
	vals[q] = uint32_t(q);
}

// [begin, end) of each slot's run of shares in the sorted keys (begin = end = 0 when
// the slot has none: the arrays are cleared first).
__global__ void k_share_ranges(const uint32_t* __restrict__ keys, uint64_t total, uint64_t n,
							   uint32_t* __restrict__ begin, uint32_t* __restrict__ end) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= total)
		return;
	const uint32_t k = keys[i];
	if (k >= n)
		return;
	if (i == 0 || keys[i - 1] != k)
		begin[k] = uint32_t(i);
	if (i + 1 == total || keys[i + 1] != k)
		end[k] = uint32_t(i + 1);
}

// force[t] = (add ? force[t] : 0) + each of slot t's shares in share-index order, one
// FP32 addition at a time (the reference adds them with atomics, in no fixed order).
__global__ void k_bonded_gather(uint64_t n, const uint32_t* __restrict__ begin, const uint32_t* __restrict__ end,
								const uint32_t* __restrict__ share_of, const float4* __restrict__ shares, int add,
								float4* __restrict__ force) {
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	float4 f = add ? force[t] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	for (uint32_t e = begin[t]; e < end[t]; ++e) {
		const float4 c = shares[share_of[e]];
		f.x += c.x;
		f.y += c.y;
		f.z += c.z;
		f.w += c.w;
	}
	force[t] = f;
}

__global__ void k_topo_to_slots(const int32_t* __restrict__ canon, int32_t* __restrict__ slots, uint64_t count,
								const uint32_t* __restrict__ slot_of_canon, uint64_t n) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= count)
		return;
	const int32_t c = canon[i];
	slots[i] = (c >= 0 && uint64_t(c) < n) ? int32_t(slot_of_canon[c]) : c;
}

// Reorder relabel: canonical c -> relabel[c] in every slot's meta.
__global__ void k_relabel(float4* __restrict__ pos, uint64_t n, const uint32_t* __restrict__ relabel,
						  uint32_t* __restrict__ slot_of_canon) {
	const uint64_t s = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (s >= n)
		return;
	const uint32_t m = fb(pos[s].w);
	const uint32_t c = meta_canon(m);
	if (c >= n)
		return;
	const uint32_t cn = relabel[c];
	pos[s].w = __uint_as_float(make_meta(meta_type(m), cn));
	slot_of_canon[cn] = uint32_t(s);
}

__global__ void k_relabel_pairs(int32_t* __restrict__ pairs, uint64_t ne, const uint32_t* __restrict__ relabel,
								uint64_t n) {
	const uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (k >= ne)
		return;
	int32_t a = pairs[2 * k], b = pairs[2 * k + 1];
	if (a >= 0 && uint64_t(a) < n)
		a = int32_t(relabel[a]);
	if (b >= 0 && uint64_t(b) < n)
		b = int32_t(relabel[b]);
	pairs[2 * k] = a < b ? a : b;
	pairs[2 * k + 1] = a < b ? b : a;
}

// Exclusion CSR on the device (DeviceBondedInteraction.h:111-167 and 386-437):
// degree count, scan, fill. Rows are sets; their internal order is not used.
__global__ void k_excl_degree(const int32_t* __restrict__ pairs, uint64_t ne, uint32_t* __restrict__ deg) {
	const uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (k >= ne)
		return;
	atomicAdd(&deg[pairs[2 * k]], 1u);
	atomicAdd(&deg[pairs[2 * k + 1]], 1u);
}
__global__ void k_excl_fill(const int32_t* __restrict__ pairs, uint64_t ne, uint32_t* __restrict__ cursor,
							int32_t* __restrict__ excluded) {
	const uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (k >= ne)
		return;
	const int32_t a = pairs[2 * k], b = pairs[2 * k + 1];
	excluded[atomicAdd(&cursor[a], 1u)] = b;
	excluded[atomicAdd(&cursor[b], 1u)] = a;
}

__global__ void k_import(const float4* __restrict__ pos_h, const float4* __restrict__ mom_h,
						 const int32_t* __restrict__ types, uint64_t n, float4* __restrict__ pos,
						 float4* __restrict__ mom, int32_t* __restrict__ ids, uint32_t* __restrict__ slot_of_canon) {
	const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (c >= n)
		return;
	const float4 p = pos_h[c];
	pos[c] = make_float4(p.x, p.y, p.z, __uint_as_float(make_meta(uint32_t(types[c]), uint32_t(c))));
	mom[c] = mom_h ? make_float4(mom_h[c].x, mom_h[c].y, mom_h[c].z, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	ids[c] = int32_t(c);
	slot_of_canon[c] = uint32_t(c);
}

__global__ void k_export(const float4* __restrict__ pos, const float4* __restrict__ mom,
						 const float4* __restrict__ force, const int32_t* __restrict__ ids,
						 const uint32_t* __restrict__ slot_of_canon, uint64_t n, float4* __restrict__ pos_o,
						 float4* __restrict__ mom_o, float4* __restrict__ force_o, int32_t* __restrict__ type_o,
						 int32_t* __restrict__ ids_o) {
	const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (c >= n)
		return;
	const uint32_t s = slot_of_canon[c];
	const float4 p = pos[s];
	pos_o[c] = make_float4(p.x, p.y, p.z, 0.0f);
	mom_o[c] = mom[s];
	force_o[c] = force[s];
	type_o[c] = int32_t(meta_type(fb(p.w)));
	ids_o[c] = ids[s];
}

uint64_t align_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

} // namespace

// ------------------------------------------------------------------ PatchImpl

struct PatchImpl::Tables {
	NbTables nb{};
	DevArray<float> nb_knots;
	std::vector<marsmd_table> bond_h, angle_h, dihedral_h;
	DevArray<float> bond_knots, angle_knots, dihedral_knots;
	DevArray<marsmd_table> bond_dev, angle_dev, dihedral_dev;
};

PatchImpl::PatchImpl(const marsmd_patch_desc& d) : tables(new Tables) {
	capacity = d.capacity;
	patch_id = d.patch_id;
	box = d.box;
	types = d.types;
	if (types.num_types < 1 || types.num_types > MARSMD_MAX_TYPES)
		raise(MARSMD_ERR_VALUE, "Patch: num_types must be in [1, " + std::to_string(MARSMD_MAX_TYPES) + "]");
	if (capacity >= kMaxParticles)
		raise(MARSMD_ERR_VALUE, "Patch: capacity must be below 2^24");
	NbTables& T = tables->nb;
	T.num_types = types.num_types;
	T.enabled = d.enabled_terms;
	T.coef = make_pair_coeffs(types, d.pair_params);
	const int nt2 = types.num_types * types.num_types;
	for (int i = 0; i < nt2; ++i)
		T.tag[i] = d.term_matrix ? d.term_matrix[i] : 0u;
	if (d.num_nb_tables > kMaxTables)
		raise(MARSMD_ERR_VALUE, "Patch: too many nonbonded tables");
	T.num_tables = d.num_nb_tables;
	std::vector<float> knots;
	for (int k = 0; k < d.num_nb_tables; ++k) {
		const marsmd_table& t = d.nb_tables[k];
		if (t.size < 2)
			raise(MARSMD_ERR_VALUE, "Patch: tables need at least 2 knots");
		if (t.is_periodic)
			raise(MARSMD_ERR_VALUE, "Patch: nonbonded tables are not periodic");
		T.start[k] = t.start;
		T.step_inv[k] = t.step_inv;
		T.size[k] = int(t.size);
		T.offset[k] = int(knots.size());
		knots.insert(knots.end(), t.pot, t.pot + t.size);
	}
	for (int i = 0; i < nt2; ++i) {
		const uint32_t tag = T.tag[i];
		if ((tag & MARSMD_PAIR_TERM_TABULATED) && int(tag >> MARSMD_PAIR_TABLE_SHIFT) >= d.num_nb_tables)
			raise(MARSMD_ERR_VALUE, "Patch: a pair tag names a missing table");
	}
	T.total_knots = int(knots.size());
	T.lds = T.total_knots <= kLdsKnots ? 1 : 0;
	tables->nb_knots.resize(std::max<uint64_t>(knots.size(), 1));
	if (!knots.empty())
		MMD_HIP(hipMemcpy(tables->nb_knots.data(), knots.data(), knots.size() * 4, hipMemcpyHostToDevice));
	auto upload_tables = [&](int num, const marsmd_table* src, std::vector<marsmd_table>& host, DevArray<float>& kd,
							 DevArray<marsmd_table>& td) {
		std::vector<float> kn;
		std::vector<size_t> off;
		for (int k = 0; k < num; ++k) {
			off.push_back(kn.size());
			kn.insert(kn.end(), src[k].pot, src[k].pot + src[k].size);
		}
		kd.resize(std::max<uint64_t>(kn.size(), 1));
		if (!kn.empty())
			MMD_HIP(hipMemcpy(kd.data(), kn.data(), kn.size() * 4, hipMemcpyHostToDevice));
		host.clear();
		for (int k = 0; k < num; ++k) {
			marsmd_table t = src[k];
			t.pot = kd.data() + off[size_t(k)];
			host.push_back(t);
		}
		td.resize(std::max<uint64_t>(host.size(), 1));
		if (!host.empty())
			MMD_HIP(hipMemcpy(td.data(), host.data(), host.size() * sizeof(marsmd_table), hipMemcpyHostToDevice));
	};
	upload_tables(d.num_bond_tables, d.bond_tables, tables->bond_h, tables->bond_knots, tables->bond_dev);
	upload_tables(d.num_angle_tables, d.angle_tables, tables->angle_h, tables->angle_knots, tables->angle_dev);
	upload_tables(d.num_dihedral_tables, d.dihedral_tables, tables->dihedral_h, tables->dihedral_knots,
				  tables->dihedral_dev);
	// topology
	num_bonds = d.num_bonds;
	num_angles = d.num_angles;
	for (uint64_t k = 0; k < 2 * num_bonds; ++k)
		topo_max_index = std::max<int64_t>(topo_max_index, d.bonds[k]);
	for (uint64_t k = 0; k < 3 * num_angles; ++k)
		topo_max_index = std::max<int64_t>(topo_max_index, d.angles[k]);
	for (uint64_t k = 0; k < 2 * d.num_exclusions; ++k)
		topo_max_index = std::max<int64_t>(topo_max_index, d.exclusions[k]);
	num_dihedrals = d.num_dihedrals;
	num_restraints = d.num_restraints;
	if (num_dihedrals && !d.dihedrals)
		raise(MARSMD_ERR_VALUE, "Patch: num_dihedrals > 0 without dihedral indices");
	if (num_restraints && (!d.restraint_ids || !d.restraint_anchors || !d.restraint_k))
		raise(MARSMD_ERR_VALUE, "Patch: num_restraints > 0 without ids, anchors and spring constants");
	for (uint64_t k = 0; k < 4 * num_dihedrals; ++k) {
		if (d.dihedrals[k] < 0)
			raise(MARSMD_ERR_VALUE, "Patch: negative dihedral index");
		topo_max_index = std::max<int64_t>(topo_max_index, d.dihedrals[k]);
	}
	for (uint64_t k = 0; k < num_restraints; ++k)
		topo_max_index = std::max<int64_t>(topo_max_index, d.restraint_ids[k]);
	for (uint64_t k = 0; k < 2 * num_bonds; ++k)
		if (d.bonds[k] < 0)
			raise(MARSMD_ERR_VALUE, "Patch: negative bond index");
	for (uint64_t k = 0; k < 3 * num_angles; ++k)
		if (d.angles[k] < 0)
			raise(MARSMD_ERR_VALUE, "Patch: negative angle index");
	if (num_bonds) {
		bonds_canon.resize(2 * num_bonds);
		bonds_slot.resize(2 * num_bonds);
		bond_table.resize(num_bonds);
		MMD_HIP(hipMemcpy(bonds_canon.data(), d.bonds, 8 * num_bonds, hipMemcpyHostToDevice));
		std::vector<int32_t> bt(num_bonds, 0);
		if (d.bond_table)
			std::copy(d.bond_table, d.bond_table + num_bonds, bt.begin());
		for (auto v : bt)
			if (v < 0 || v >= d.num_bond_tables)
				raise(MARSMD_ERR_VALUE, "Patch: bond table index out of range");
		MMD_HIP(hipMemcpy(bond_table.data(), bt.data(), 4 * num_bonds, hipMemcpyHostToDevice));
	}
	if (num_angles) {
		std::vector<int32_t> a4(4 * num_angles, 0);
		for (uint64_t k = 0; k < num_angles; ++k)
			for (int q = 0; q < 3; ++q)
				a4[4 * k + uint64_t(q)] = d.angles[3 * k + uint64_t(q)];
		angles_canon.resize(4 * num_angles);
		angles_slot.resize(4 * num_angles);
		angle_table.resize(num_angles);
		MMD_HIP(hipMemcpy(angles_canon.data(), a4.data(), 16 * num_angles, hipMemcpyHostToDevice));
		std::vector<int32_t> at(num_angles, 0);
		if (d.angle_table)
			std::copy(d.angle_table, d.angle_table + num_angles, at.begin());
		for (auto v : at)
			if (v < 0 || v >= d.num_angle_tables)
				raise(MARSMD_ERR_VALUE, "Patch: angle table index out of range");
		MMD_HIP(hipMemcpy(angle_table.data(), at.data(), 4 * num_angles, hipMemcpyHostToDevice));
	}
	if (num_dihedrals) {
		dihedrals_canon.resize(4 * num_dihedrals);
		dihedrals_slot.resize(4 * num_dihedrals);
		dihedral_table.resize(num_dihedrals);
		MMD_HIP(hipMemcpy(dihedrals_canon.data(), d.dihedrals, 16 * num_dihedrals, hipMemcpyHostToDevice));
		std::vector<int32_t> dt(num_dihedrals, 0);
		if (d.dihedral_table)
			std::copy(d.dihedral_table, d.dihedral_table + num_dihedrals, dt.begin());
		for (auto v : dt)
			if (v < 0 || v >= d.num_dihedral_tables)
				raise(MARSMD_ERR_VALUE, "Patch: dihedral table index out of range");
		MMD_HIP(hipMemcpy(dihedral_table.data(), dt.data(), 4 * num_dihedrals, hipMemcpyHostToDevice));
	}
	if (num_restraints) {
		restraint_canon.resize(num_restraints);
		restraint_slot.resize(num_restraints);
		restraint_anchor.resize(num_restraints);
		restraint_k.resize(num_restraints);
		MMD_HIP(hipMemcpy(restraint_canon.data(), d.restraint_ids, 4 * num_restraints, hipMemcpyHostToDevice));
		MMD_HIP(hipMemcpy(restraint_anchor.data(), d.restraint_anchors, 16 * num_restraints, hipMemcpyHostToDevice));
		MMD_HIP(hipMemcpy(restraint_k.data(), d.restraint_k, 4 * num_restraints, hipMemcpyHostToDevice));
	}
	num_excl = d.num_exclusions;
	if (num_excl) {
		std::vector<int32_t> ex(2 * num_excl);
		for (uint64_t k = 0; k < num_excl; ++k) {
			const int32_t a = d.exclusions[2 * k], b = d.exclusions[2 * k + 1];
			if (a < 0 || b < 0 || uint64_t(a) >= capacity || uint64_t(b) >= capacity)
				raise(MARSMD_ERR_VALUE, "Patch: exclusion index out of range");
			ex[2 * k] = std::min(a, b);
			ex[2 * k + 1] = std::max(a, b);
		}
		excl_pairs.resize(2 * num_excl);
		MMD_HIP(hipMemcpy(excl_pairs.data(), ex.data(), 8 * num_excl, hipMemcpyHostToDevice));
	}
	// grids
	num_grids = d.num_grids;
	std::vector<marsmd_grid_view> gv;
	for (int k = 0; k < d.num_grids; ++k) {
		marsmd_grid_view g = d.grids[k];
		const uint64_t cnt = g.dims[0] * g.dims[1] * g.dims[2];
		grid_data.emplace_back(cnt);
		MMD_HIP(hipMemcpy(grid_data.back().data(), g.data, cnt * 4, hipMemcpyHostToDevice));
		g.data = grid_data.back().data();
		gv.push_back(g);
	}
	if (!gv.empty()) {
		grids_dev.resize(gv.size());
		MMD_HIP(hipMemcpy(grids_dev.data(), gv.data(), gv.size() * sizeof(marsmd_grid_view), hipMemcpyHostToDevice));
	}
	pl.reset(new PairlistImpl(capacity, ~uint64_t(0)));
	pl->set_box(box);
}

PatchImpl::~PatchImpl() {
	for (hipEvent_t e : timing_events)
		(void)hipEventDestroy(e);
	delete tables;
}

void PatchImpl::set_kernel_timing(bool enable) {
	timing = enable;
	timing_used = 0;
}

void PatchImpl::kernel_timing(double* total_ms, uint64_t* launches) {
	double sum = 0.0;
	for (size_t k = 0; k + 1 < timing_used; k += 2) {
		MMD_HIP(hipEventSynchronize(timing_events[k + 1]));
		float ms = 0.0f;
		MMD_HIP(hipEventElapsedTime(&ms, timing_events[k], timing_events[k + 1]));
		sum += double(ms);
	}
	*total_ms = sum;
	*launches = timing_used / 2;
}

void PatchImpl::rebuild_exclusions(hipStream_t s) {
	if (num_excl == 0 || n == 0)
		return;
	excl_offsets.resize(n + 1);
	excl_deg.resize(n + 1);
	excl_nbrs.resize(2 * num_excl);
	excl_scan.resize(scan_scratch_bytes(n + 1));
	fill_u32(excl_deg.data(), 0u, n + 1, s);
	k_excl_degree<<<grid_for(num_excl, 256), 256, 0, s>>>(excl_pairs.data(), num_excl, excl_deg.data());
	MMD_LAUNCH_CHECK();
	exclusive_scan_u32(excl_deg.data(), excl_offsets.data(), n, excl_scan.data(), s);
	MMD_HIP(hipMemcpyAsync(excl_deg.data(), excl_offsets.data(), 4 * n, hipMemcpyDeviceToDevice, s));
	k_excl_fill<<<grid_for(num_excl, 256), 256, 0, s>>>(excl_pairs.data(), num_excl, excl_deg.data(),
														 reinterpret_cast<int32_t*>(excl_nbrs.data()));
	MMD_LAUNCH_CHECK();
}

void PatchImpl::copy_from_host(const float* pos4, const float* mom4, const int32_t* type_ids, uint64_t nn) {
	prune_valid = false; // positions or rows change
	if (nn > capacity)
		raise(MARSMD_ERR_RUNTIME, "Cannot copy " + std::to_string(nn) + " particles to patch with capacity " +
									  std::to_string(capacity));
	for (uint64_t i = 0; i < nn; ++i)
		if (type_ids[i] < 0 || type_ids[i] >= types.num_types)
			raise(MARSMD_ERR_VALUE, "Patch: particle type out of range");
	if (topo_max_index >= int64_t(nn))
		raise(MARSMD_ERR_VALUE, "Patch: bonded topology references particle " + std::to_string(topo_max_index) +
									" but only " + std::to_string(nn) + " were copied");
	n = nn;
	for (int b = 0; b < 2; ++b)
		pos[b].resize(std::max<uint64_t>(n, 1));
	mom.resize(std::max<uint64_t>(n, 1));
	mom_tmp.resize(std::max<uint64_t>(n, 1));
	force.resize(std::max<uint64_t>(n, 1));
	ids.resize(std::max<uint64_t>(n, 1));
	ids_tmp.resize(std::max<uint64_t>(n, 1));
	slot_of_canon.resize(std::max<uint64_t>(n, 1));
	cur = 0;
	if (n) {
		DevArray<float4> hp(n), hm(mom4 ? n : 0);
		DevArray<int32_t> ht(n);
		MMD_HIP(hipMemcpy(hp.data(), pos4, 16 * n, hipMemcpyHostToDevice));
		if (mom4)
			MMD_HIP(hipMemcpy(hm.data(), mom4, 16 * n, hipMemcpyHostToDevice));
		MMD_HIP(hipMemcpy(ht.data(), type_ids, 4 * n, hipMemcpyHostToDevice));
		k_import<<<grid_for(n, 256), 256>>>(hp.data(), mom4 ? hm.data() : nullptr, ht.data(), n, pos[0].data(),
											mom.data(), ids.data(), slot_of_canon.data());
		MMD_LAUNCH_CHECK();
		k_zero_force<<<grid_for(n, 256), 256>>>(force.data(), n);
		MMD_LAUNCH_CHECK();
		MMD_HIP(hipDeviceSynchronize());
	}
	built = false;
	force_rebuild = false;
	deferred_kick = false;
	pl->presorted_next = false;
	rebuild_exclusions(nullptr);
	MMD_HIP(hipDeviceSynchronize());
}

void PatchImpl::copy_to_host(float* pos4, float* mom4, float* force4, int32_t* type_ids, int32_t* id_out,
							 uint64_t nn) {
	if (nn > n)
		raise(MARSMD_ERR_RUNTIME, "Cannot copy " + std::to_string(nn) + " particles from patch containing only " +
									  std::to_string(n));
	if (nn == 0)
		return;
	DevArray<float4> op(n), om(n), of(n);
	DevArray<int32_t> ot(n), oi(n);
	MMD_HIP(hipDeviceSynchronize());
	k_export<<<grid_for(n, 256), 256>>>(pos[cur].data(), mom.data(), force.data(), ids.data(), slot_of_canon.data(), n,
										op.data(), om.data(), of.data(), ot.data(), oi.data());
	MMD_LAUNCH_CHECK();
	MMD_HIP(hipDeviceSynchronize());
	if (pos4)
		MMD_HIP(hipMemcpy(pos4, op.data(), 16 * nn, hipMemcpyDeviceToHost));
	if (mom4)
		MMD_HIP(hipMemcpy(mom4, om.data(), 16 * nn, hipMemcpyDeviceToHost));
	if (force4)
		MMD_HIP(hipMemcpy(force4, of.data(), 16 * nn, hipMemcpyDeviceToHost));
	if (type_ids)
		MMD_HIP(hipMemcpy(type_ids, ot.data(), 4 * nn, hipMemcpyDeviceToHost));
	if (id_out)
		MMD_HIP(hipMemcpy(id_out, oi.data(), 4 * nn, hipMemcpyDeviceToHost));
}

void PatchImpl::build(float pairlist_cutoff, hipStream_t s) {
	prune_valid = false; // positions or rows change
	marsmd_exclusion_view ev{};
	if (num_excl) {
		ev.offsets = reinterpret_cast<const int32_t*>(excl_offsets.data());
		ev.excluded = reinterpret_cast<const int32_t*>(excl_nbrs.data());
		ev.num_particles = n;
	}
	pl->excl = ev;
	pl->build(pos[cur].data(), slot_of_canon.data(), true, nullptr, n, pairlist_cutoff, s);
	// The sorted positions become the slot-order state; gather the rest.
	slot_tmp.resize(std::max<uint64_t>(n, 1));
	k_gather_state<<<grid_for(n, 256), 256, 0, s>>>(pl->sorter.sorted_idx.data(), slot_of_canon.data(), n,
													pl->spos.data(), pos[cur].data(), pl->sorter.inverse.data(),
													slot_tmp.data(), mom.data(), mom_tmp.data(), ids.data(),
													ids_tmp.data());
	MMD_LAUNCH_CHECK();
	mom.swap(mom_tmp);
	ids.swap(ids_tmp);
	slot_of_canon.swap(slot_tmp);
	if (num_bonds) {
		k_topo_to_slots<<<grid_for(2 * num_bonds, 256), 256, 0, s>>>(bonds_canon.data(), bonds_slot.data(),
																	  2 * num_bonds, slot_of_canon.data(), n);
		MMD_LAUNCH_CHECK();
	}
	if (num_dihedrals) {
		k_topo_to_slots<<<grid_for(4 * num_dihedrals, 256), 256, 0, s>>>(
			dihedrals_canon.data(), dihedrals_slot.data(), 4 * num_dihedrals, slot_of_canon.data(), n);
		MMD_LAUNCH_CHECK();
	}
	if (num_restraints) {
		k_topo_to_slots<<<grid_for(num_restraints, 256), 256, 0, s>>>(restraint_canon.data(), restraint_slot.data(),
																	   num_restraints, slot_of_canon.data(), n);
		MMD_LAUNCH_CHECK();
	}
	if (num_angles) {
		k_topo_to_slots<<<grid_for(4 * num_angles, 256), 256, 0, s>>>(angles_canon.data(), angles_slot.data(),
																	   4 * num_angles, slot_of_canon.data(), n);
		MMD_LAUNCH_CHECK();
	}
	build_share_index(s);
	built = true;
}

bool PatchImpl::rebuild_policy(float pairlist_cutoff, float cutoff, uint64_t step, uint64_t period, hipStream_t s) {
	// Patch::calculate_nonbonded_forces rebuild decision (Patch.cpp:85-121).
	const bool at_period = (period == 0 || (step - 1) % period == 0);
	bool rebuild = at_period || !built;
	const bool forced = force_rebuild;
	force_rebuild = false;
	rebuild = rebuild || forced;
	const float skin = pairlist_cutoff - cutoff;
	// the displacement record of the pruned rows' life, for the next margin (the build
	// synchronises the stream before prepare_prune reads it)
	prune_life_tracked = false;
	if (rebuild && prune_recording) {
		MMD_HIP(hipMemcpyAsync(prune_host.ptr, prune_disp.data() + 4, 4, hipMemcpyDeviceToHost, s));
		prune_life_tracked = true;
	}
	if (rebuild && !forced && period != 0 && built && skin > 0.0f && n > 0)
		rebuild = pl->max_displacement(pos[cur].data(), nullptr, n, s) > skin * 0.5f;
	if (rebuild)
		build(pairlist_cutoff, s);
	else
		prune_life_tracked = false;
	return rebuild;
}

// (at least two knots: the step kernels read a clamped pair of knots for every pair)
unsigned PatchImpl::lds_bytes() const {
	return tables->nb.lds ? unsigned(std::max(tables->nb.total_knots, 2)) * 4u : 0u;
}

// Workgroups for a persistent launch over the 256-slot tiles: as many as are resident
// at once (occupancy x CUs), a multiple of 8 (one share per XCD), at most the tiles.
unsigned PatchImpl::persistent_grid(const void* kernel, uint64_t nn) {
	const uint64_t tiles = (nn + kBlock - 1) / kBlock;
	const unsigned lds = lds_bytes();
	auto it = occupancy.find(kernel);
	if (it == occupancy.end() || it->second.first != lds) {
		int occ = 0;
		MMD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, kBlock, lds));
		it = occupancy.insert_or_assign(kernel, std::make_pair(lds, std::max(occ, 1))).first;
	}
	if (num_cus == 0) {
		int dev = 0;
		MMD_HIP(hipGetDevice(&dev));
		MMD_HIP(hipDeviceGetAttribute(&num_cus, hipDeviceAttributeMultiprocessorCount, dev));
		num_cus = std::max(num_cus, 1);
	}
	const int per_cu = blocks_per_cu > 0 ? std::min(blocks_per_cu, it->second.second) : it->second.second;
	uint64_t g = uint64_t(per_cu) * uint64_t(num_cus);
	g = std::min<uint64_t>(g, (tiles + 7) / 8 * 8);
	g = std::max<uint64_t>((g + 7) / 8 * 8, 8);
	return unsigned(g);
}

void PatchImpl::calculate_nonbonded(float pairlist_cutoff, float cutoff, uint64_t step, uint64_t period,
									const float ef[3], int interp, bool energy, hipStream_t s) {
	if (n == 0)
		return;
	bonded_prepared = true;
	for (int a = 0; a < 3; ++a)
		efield[a] = ef[a];
	interpolation = interp;
	rebuild_policy(pairlist_cutoff, cutoff, step, period, s);
	const float cut_d = cutoff_threshold(cutoff > 0.0f ? cutoff * cutoff : 0.0f);
	auto kern = (tables->nb.enabled & kAnalyticTerms) ? k_nb_forces<true> : k_nb_forces<false>;
	kern<<<grid_for(n, kBlock), kBlock, lds_bytes(), s>>>(pos[cur].data(), n, pl->list(), tables->nb,
														  tables->nb_knots.data(), make_devbox(box), cut_d,
														  energy ? 1 : 0, force.data());
	MMD_LAUNCH_CHECK();
}

uint64_t PatchImpl::num_shares() const {
	return 2 * num_bonds + 3 * num_angles + 4 * num_dihedrals + num_restraints;
}

// Each slot's bonded shares, in share-index order (after the topology's slot arrays
// change: at every build).
void PatchImpl::build_share_index(hipStream_t s) {
	const uint64_t q = num_shares();
	if (q == 0 || n == 0)
		return;
	if (q >= (uint64_t(1) << 32))
		raise(MARSMD_ERR_CAPACITY, "Patch: more than 2^32 bonded shares");
	share_keys.resize(2 * q);
	share_vals.resize(2 * q);
	share_begin.resize(n);
	share_end.resize(n);
	k_share_keys<<<grid_for(q, 256), 256, 0, s>>>(bonds_slot.data(), num_bonds, angles_slot.data(), num_angles,
												   dihedrals_slot.data(), num_dihedrals, restraint_slot.data(),
												   num_restraints, share_keys.data(), share_vals.data());
	MMD_LAUNCH_CHECK();
	const uint64_t need = radix_sort_scratch_bytes(q);
	share_scratch.resize(std::max<uint64_t>(need, 1));
	// every key bit (shares without a particle carry all ones; once per build)
	radix_sort_pairs(share_keys.data(), share_vals.data(), share_keys.data() + q, share_vals.data() + q, q, 32,
					 share_scratch.data(), need, s);
	fill_u32(share_begin.data(), 0u, n, s);
	fill_u32(share_end.data(), 0u, n, s);
	k_share_ranges<<<grid_for(q, 256), 256, 0, s>>>(share_keys.data() + q, q, n, share_begin.data(),
													 share_end.data());
	MMD_LAUNCH_CHECK();
}

// The bonded terms' forces: every share stored by the launchers, then summed per slot
// in a fixed order into `force` (added to it when add, else written), so the result
// does not depend on the order in which the terms ran.
void PatchImpl::calculate_bonded(bool energy, hipStream_t s, bool add) {
	if (n == 0 || num_bonded_terms() == 0)
		return;
	if (!built)
		raise(MARSMD_ERR_RUNTIME, "Patch: calculate_bonded_forces before the first nonbonded pass");
	const uint64_t q = num_shares();
	shares.resize(q);
	float4* c = shares.data();
	const float* p = reinterpret_cast<const float*>(pos[cur].data());
	if (num_bonds)
		launch_tabulated_bonds(bonds_slot.data(), p, nullptr, tables->bond_dev.data(), bond_table.data(), nullptr, box,
							   energy, num_bonds, s, c);
	c += 2 * num_bonds;
	if (num_angles)
		launch_tabulated_angles(angles_slot.data(), p, nullptr, tables->angle_dev.data(), angle_table.data(), nullptr,
								box, energy, num_angles, s, c);
	c += 3 * num_angles;
	if (num_dihedrals)
		launch_tabulated_dihedrals(dihedrals_slot.data(), p, nullptr, tables->dihedral_dev.data(),
								   dihedral_table.data(), nullptr, box, energy, num_dihedrals, s, c);
	c += 4 * num_dihedrals;
	if (num_restraints)
		launch_harmonic_restraints(restraint_slot.data(), p, nullptr,
								   reinterpret_cast<const float*>(restraint_anchor.data()), restraint_k.data(), box,
								   energy, num_restraints, s, c);
	k_bonded_gather<<<grid_for(n, 256), 256, 0, s>>>(n, share_begin.data(), share_end.data(), share_vals.data() + q,
													  shares.data(), add ? 1 : 0, force.data());
	MMD_LAUNCH_CHECK();
}

void PatchImpl::integrate(float dt, float kT, int integrator, uint64_t step, hipStream_t s) {
	prune_valid = false; // positions or rows change
	// positions move: a build after this one is no longer presorted by a reorder
	pl->presorted_next = false;
	if (n == 0)
		return;
	StepParams S{};
	S.box = make_devbox(box);
	S.dt = dt;
	S.kT = kT;
	S.seed = mix_seed(base_seed, patch_id, step);
	S.ctr = 0;
	S.step = uint32_t(step);
	S.scheme = interpolation;
	S.has_grids = num_grids > 0 ? 1 : 0;
	for (int a = 0; a < 3; ++a)
		S.ef[a] = efield[a];
	const unsigned g = grid_for(n, kBlock);
	if (integrator == MARSMD_INTEGRATOR_BROWNIAN) {
		k_integrate_bd<<<g, kBlock, 0, s>>>(pos[cur].data(), force.data(), n, types, grids_dev.data(), S);
		MMD_LAUNCH_CHECK();
		return;
	}
	if (integrator != MARSMD_INTEGRATOR_LANGEVIN)
		raise(MARSMD_ERR_VALUE, "Unsupported integrator type: " + std::to_string(integrator));
	if (num_grids > 0) {
		k_integrate_pmf<<<g, kBlock, 0, s>>>(pos[cur].data(), force.data(), n, types, grids_dev.data(), S);
		MMD_LAUNCH_CHECK();
	}
	if (deferred_kick) {
		k_integrate_last<<<g, kBlock, 0, s>>>(mom.data(), force.data(), n, dt);
		MMD_LAUNCH_CHECK();
		deferred_kick = false;
	}
	deferred_kick = true;
	k_integrate_baoab<<<g, kBlock, 0, s>>>(pos[cur].data(), mom.data(), force.data(), n, types, S);
	MMD_LAUNCH_CHECK();
}

void PatchImpl::finish_deferred_kick(float dt, hipStream_t s) {
	if (!deferred_kick)
		return;
	deferred_kick = false;
	if (n == 0)
		return;
	k_integrate_last<<<grid_for(n, kBlock), kBlock, 0, s>>>(mom.data(), force.data(), n, dt);
	MMD_LAUNCH_CHECK();
}

void PatchImpl::reorder(hipStream_t s) {
	prune_valid = false; // positions or rows change
	// Patch::reorder_particles (Patch.cpp:456-480): System-mode Morton sort over the
	// periodic box, relabel canonical indices, remap topology, force a rebuild.
	if (n == 0)
		return;
	float bmin[3], bmax[3];
	for (int a = 0; a < 3; ++a) {
		bmin[a] = box.origin[a];
		bmax[a] = box.origin[a] + box.size[a];
		if (!(bmax[a] > bmin[a]))
			return;
	}
	ZSort& z = pl->sorter;
	z.sort(pos[cur].data(), n, bmin, bmax, nullptr, true, s);
	// relabel: canonical c -> rank of c in the System sort = z.inverse[c]
	relabel.resize(n);
	MMD_HIP(hipMemcpyAsync(relabel.data(), z.inverse.data(), 4 * n, hipMemcpyDeviceToDevice, s));
	k_relabel<<<grid_for(n, 256), 256, 0, s>>>(pos[cur].data(), n, relabel.data(), slot_of_canon.data());
	MMD_LAUNCH_CHECK();
	if (num_bonds) {
		k_remap_canon(bonds_canon.data(), 2 * num_bonds, s);
	}
	if (num_angles) {
		k_remap_canon(angles_canon.data(), 4 * num_angles, s);
	}
	if (num_dihedrals)
		k_remap_canon(dihedrals_canon.data(), 4 * num_dihedrals, s);
	if (num_restraints)
		k_remap_canon(restraint_canon.data(), num_restraints, s);
	if (num_excl) {
		k_relabel_pairs<<<grid_for(num_excl, 256), 256, 0, s>>>(excl_pairs.data(), num_excl, relabel.data(), n);
		MMD_LAUNCH_CHECK();
		rebuild_exclusions(s);
	}
	pl->presorted_next = true;
	force_rebuild = true;
}

void PatchImpl::k_remap_canon(int32_t* idx, uint64_t count, hipStream_t s) {
	remap_tmp.resize(count);
	MMD_HIP(hipMemcpyAsync(remap_tmp.data(), idx, 4 * count, hipMemcpyDeviceToDevice, s));
	k_topo_to_slots<<<grid_for(count, 256), 256, 0, s>>>(remap_tmp.data(), idx, count, relabel.data(), n);
	MMD_LAUNCH_CHECK();
}

void PatchImpl::prepare_prune(float cutoff, float pairlist_cutoff, hipStream_t s) {
	if (prune_mode == 0 || n == 0 || !pl->built) {
		prune_valid = false;
		prune_recording = false;
		return;
	}
	if (prune_valid && prune_list_version == pl->list_version)
		return;
	const float skin = pairlist_cutoff - cutoff;
	// margin 0: the rows sorted only, every entry kept, no displacement record
	float margin = 0.0f;
	if (prune_mode >= 2 && cutoff > 0.0f && skin > 0.0f) {
		if (prune_margin_fixed >= 0.0f) {
			margin = prune_margin_fixed;
		} else if (prune_life_tracked) {
			// the previous life's largest displacement, with room to spare
			float d2 = 0.0f;
			std::memcpy(&d2, prune_host.ptr, 4);
			margin = 1.5f * std::sqrt(std::max(d2, 0.0f)) + 0.02f * skin;
		} else {
			margin = prune_margin > 0.0f ? prune_margin : 0.25f * skin;
		}
		margin = std::min(margin, 0.5f * skin); // at skin / 2 the rows keep every entry
	}
	prune_life_tracked = false;
	const uint64_t blocks = (n + 255) / 256;
	const uint32_t cap = (std::max<uint32_t>(pl->max_neighbors, 1u) + 3u) & ~3u;
	if (blocks * 256u * uint64_t(cap) >= (uint64_t(1) << 32))
		raise(MARSMD_ERR_RUNTIME, "Patch: pruned rows beyond 32-bit offsets");
	prune_lane_slot.resize(n);
	prune_count.resize(n);
	prune_tile_base.resize(blocks * 4);
	prune_nbr.resize(blocks * 256u * cap);
	prune_ref.resize(n);
	prune_disp.resize(8);
	// keep entries within cutoff + 2 margin (squared, rounded up)
	const double rk = double(cutoff) + 2.0 * double(margin);
	const float keep2 = float(rk * rk * (1.0 + 1e-6));
	const int keep_all = margin > 0.0f && margin < 0.5f * skin ? 0 : 1;
	prune_keep_all = keep_all != 0;
	k_prune<<<unsigned(blocks), 256, 0, s>>>(pos[cur].data(), uint32_t(n), pl->list(), make_devbox(box), keep2,
											 keep_all, prune_mode == 3 ? 0 : 1, cap, prune_lane_slot.data(), prune_count.data(),
											 prune_tile_base.data(), prune_nbr.data(), prune_ref.data(),
											 prune_disp.data());
	MMD_LAUNCH_CHECK();
	prune_margin = margin;
	prune_recording = margin > 0.0f;
	prune_valid = true;
	prune_list_version = pl->list_version;
	prune_since = 0;
}

void PatchImpl::run(const marsmd_run_params& P, uint64_t first_step, uint64_t num_steps, hipStream_t s,
					bool store_last) {
	if (P.integrator != MARSMD_INTEGRATOR_BROWNIAN && P.integrator != MARSMD_INTEGRATOR_LANGEVIN)
		raise(MARSMD_ERR_VALUE, "Unsupported integrator type: " + std::to_string(P.integrator));
	if (n == 0)
		return;
	const int64_t rp = P.reorder_period > 0 ? P.reorder_period : int64_t(P.rebuild_period);
	const float cut_d = cutoff_threshold(P.cutoff > 0.0f ? P.cutoff * P.cutoff : 0.0f);
	const bool has_bonded = num_bonded_terms() > 0;
	const unsigned g = grid_for(n, kBlock);
	for (uint64_t step = first_step + 1; step <= first_step + num_steps; ++step) {
		// ParticleReorderManager::maybe_reorder (ReorderManager.cpp:8-25)
		if (rp > 0 && step != 0 && (step - 1) % uint64_t(rp) == 0 && bonded_prepared)
			reorder(s);
		const bool energy = P.energy_period > 0 && step % P.energy_period == 0;
		bonded_prepared = true;
		for (int a = 0; a < 3; ++a)
			efield[a] = P.electric_field[a];
		interpolation = P.interpolation;
		rebuild_policy(P.pairlist_cutoff, P.cutoff, step, P.rebuild_period, s);
		if (has_bonded)
			calculate_bonded(energy, s, false); // written, not added: the step adds the rest
		StepParams S{};
		S.box = make_devbox(box);
		S.cut_d = cut_d;
		S.dt = P.timestep;
		S.kT = P.kT;
		S.seed = mix_seed(base_seed, patch_id, step);
		S.ctr = 0;
		S.step = uint32_t(step);
		S.scheme = P.interpolation;
		S.has_grids = num_grids > 0 ? 1 : 0;
		S.deferred = deferred_kick ? 1 : 0;
		for (int a = 0; a < 3; ++a)
			S.ef[a] = P.electric_field[a];
		const int store = (store_last && step == first_step + num_steps) ? 1 : 0;
		float4* pnext = pos[1 - cur].data();
		const bool bd = P.integrator == MARSMD_INTEGRATOR_BROWNIAN;
		const bool an = (tables->nb.enabled & kAnalyticTerms) != 0;
		// with grids, the position-dependent force may come from its own kernel
		const bool pmf_in = split_pmf && num_grids > 0;
		const StepFn kern = pick_step(bd, energy, an, pmf_in);
		const unsigned grid = persistent_grid(reinterpret_cast<const void*>(kern), n);
		prepare_prune(P.cutoff, P.pairlist_cutoff, s);
		PruneView PR{};
		if (prune_valid) {
			PR.list = NbrList{prune_count.data(), prune_tile_base.data(), prune_nbr.data()};
			PR.lane_slot = prune_lane_slot.data();
			PR.ref = prune_ref.data();
			PR.disp = prune_disp.data();
			PR.since = prune_since++;
			if (prune_margin > 0.0f) {
				// the margin less the rounding of FP32 distances and of the periodic wrap
				// (rows that keep every entry hold at any displacement: recorded only)
				const float lmax = std::max({box.size[0], box.size[1], box.size[2]});
				const float lim = prune_margin - (1e-3f + 8.0f * lmax * 0x1p-23f);
				PR.limit2 = prune_keep_all ? 3.0e38f : lim > 0.0f ? lim * lim : 1e-30f;
			}
		}
		if (pmf_in)
			pmf_buf.resize(n);
		if (timing) {
			while (timing_events.size() < timing_used + 2) {
				hipEvent_t e;
				MMD_HIP(hipEventCreate(&e));
				timing_events.push_back(e);
			}
			MMD_HIP(hipEventRecord(timing_events[timing_used], s));
		}
		if (pmf_in) {
			k_pmf_force<<<(g + 7u) / 8u * 8u, kBlock, 0, s>>>(pos[cur].data(), n, types, grids_dev.data(), S,
															   pmf_buf.data());
			MMD_LAUNCH_CHECK();
		}
		kern<<<grid, kBlock, lds_bytes(), s>>>(pos[cur].data(), pnext, mom.data(), force.data(), n, pl->list(),
											   tables->nb, tables->nb_knots.data(), types, grids_dev.data(), S,
											   has_bonded, store, pmf_in ? pmf_buf.data() : nullptr, PR);
		if (timing) {
			MMD_HIP(hipEventRecord(timing_events[timing_used + 1], s));
			timing_used += 2;
		}
		if (!bd)
			deferred_kick = true;
		MMD_LAUNCH_CHECK();
		cur = 1 - cur;
	}
}

void PatchImpl::set_cells(int home_target, int fine_offset) {
	if (home_target < 1 || home_target > 4096)
		raise(MARSMD_ERR_VALUE, "Patch: home_target must be in [1, 4096]");
	if (fine_offset < 1 || fine_offset > 4)
		raise(MARSMD_ERR_VALUE, "Patch: fine_offset must be in [1, 4]");
	pl->home_target = home_target;
	pl->fine_offset = fine_offset;
}

void PatchImpl::autotune_cells(const marsmd_run_params& P, uint64_t first_step, uint64_t num_steps, hipStream_t s,
							   double* walk_out) {
	uint64_t step = first_step;
	if (!pl->built && num_steps > 0) { // the first build
		run(P, step, 1, s);
		++step;
	}
	double best = -1.0;
	if (pl->built && pl->home_level >= 0) {
		const auto measure = [&]() {
			double b = 1e300;
			for (int k = 0; k < 2; ++k)
				b = std::min(b, pl->rewalk_ms(s));
			return b;
		};
		int bt = pl->home_target, bf = pl->fine_offset;
		best = measure();
		for (int t : {8, 12, 16, 24, 48})
			for (int f : {1, 2, 3}) {
				if (t == bt && f == bf)
					continue;
				set_cells(t, f);
				const double ms = measure();
				if (ms < 0.99 * best) {
					best = ms;
					bt = t;
					bf = f;
				}
			}
		set_cells(bt, bf);
		(void)pl->rewalk_ms(s); // the live list as the choice walks it
	}
	*walk_out = best > 0.0 ? best : 0.0;
	if (step < first_step + num_steps)
		run(P, step, first_step + num_steps - step, s);
}

void PatchImpl::set_launch(int bpc, bool split) {
	if (bpc < 0)
		raise(MARSMD_ERR_VALUE, "Patch: blocks_per_cu must be >= 0");
	blocks_per_cu = bpc;
	split_pmf = split;
}

void PatchImpl::autotune(const marsmd_run_params& P, uint64_t first_step, uint64_t num_steps, hipStream_t s,
						 double* best_ms) {
	// candidates: step-kernel workgroups per CU from the occupancy down, with the
	// position-dependent force sampled inside the step kernel or by its own kernel (when
	// there are grids); every candidate computes the same bits
	const bool bd = P.integrator == MARSMD_INTEGRATOR_BROWNIAN;
	const bool an = (tables->nb.enabled & kAnalyticTerms) != 0;
	struct Cand {
		int bpc;
		bool split;
		double ms = 1e300;
	};
	std::vector<Cand> cands;
	for (bool sp : {false, true}) {
		if (sp && num_grids == 0)
			continue;
		// each step kernel's own occupancy (the split one holds fewer registers), then less
		const StepFn probe = pick_step(bd, false, an, sp);
		const int saved = blocks_per_cu;
		blocks_per_cu = 0;
		(void)persistent_grid(reinterpret_cast<const void*>(probe), std::max<uint64_t>(n, 1));
		blocks_per_cu = saved;
		const int occ = occupancy[reinterpret_cast<const void*>(probe)].second;
		std::vector<int> per_cu = {occ};
		for (int v : {occ - 1, occ / 2})
			if (v >= 1 && std::find(per_cu.begin(), per_cu.end(), v) == per_cu.end())
				per_cu.push_back(v);
		for (int v : per_cu)
			cands.push_back(Cand{v, sp});
	}
	// the candidates take turns step by step, one round of steps (one step per candidate)
	// enqueued before its events are read, so the stream stays busy and no launch waits
	// on the host; the order rotates by one every round, so over the rounds each candidate
	// is timed at every point of the rebuild period (the steps after a rebuild are
	// slower: cold caches); only the tuned run's real last step stores its forces. Each
	// candidate keeps the median of its step-kernel times.
	const bool was_timing = timing;
	const uint64_t nc = cands.size();
	const uint64_t rounds = num_steps / nc;
	const uint64_t last = first_step + num_steps;
	std::vector<std::vector<double>> times(nc);
	uint64_t step = first_step;
	for (uint64_t r = 0; r < rounds; ++r) {
		set_kernel_timing(true); // clears the recorded events
		std::vector<size_t> order(nc);
		for (uint64_t k = 0; k < nc; ++k) {
			order[k] = size_t((k + r) % nc);
			set_launch(cands[order[k]].bpc, cands[order[k]].split);
			run(P, step, 1, s, step + 1 == last);
			step += 1;
		}
		for (uint64_t k = 0; k < nc && 2 * k + 1 < timing_used; ++k) {
			MMD_HIP(hipEventSynchronize(timing_events[2 * k + 1]));
			float ms = 0.0f;
			MMD_HIP(hipEventElapsedTime(&ms, timing_events[2 * k], timing_events[2 * k + 1]));
			times[order[k]].push_back(double(ms));
		}
	}
	for (size_t ci = 0; ci < cands.size(); ++ci) {
		std::vector<double>& t = times[ci];
		if (t.empty())
			continue;
		std::sort(t.begin(), t.end());
		cands[ci].ms = t.size() % 2 ? t[t.size() / 2] : 0.5 * (t[t.size() / 2 - 1] + t[t.size() / 2]);
	}
	// the first candidate (the fused kernel at its occupancy) unless another is faster by
	// more than 1% (the medians of neighbouring candidates differ by about that much
	// from run to run)
	const Cand* best = &cands.front();
	for (const Cand& c : cands)
		if (c.ms < 0.99 * best->ms)
			best = &c;
	set_launch(best->bpc, best->split);
	set_kernel_timing(was_timing);
	if (step < first_step + num_steps)
		run(P, step, first_step + num_steps - step, s);
	*best_ms = best->ms < 1e300 ? best->ms : 0.0;
}

} // namespace mmd
