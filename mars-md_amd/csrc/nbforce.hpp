// nbforce.hpp — the nonbonded pair term over the wave-tiled neighbour list, shared by
// the Patch kernels (patch.hip), the replica batch (batch.hip) and force matching
// (fm.hip).
#pragma once

#include "analytic.hpp"
#include "pairlist.hpp"

namespace mmd {

constexpr int kMaxTables = 32;
constexpr int kLdsKnots = 15360; // 60 KB of table knots staged in LDS

struct NbTables {
	int num_tables;
	int num_types;
	uint32_t enabled;
	int lds; // knots staged in LDS
	int total_knots;
	float start[kMaxTables];
	float step_inv[kMaxTables];
	int size[kMaxTables];
	int offset[kMaxTables];
	uint32_t tag[MARSMD_MAX_TYPES * MARSMD_MAX_TYPES];
	PairCoeffs coef; // read by the analytical terms only
};

// What a pair term reads besides positions and knots, staged in LDS once per
// workgroup: parameters indexed per lane (table index, particle type) would otherwise
// be vector loads from the kernel-argument block on every pair.
struct NbLds {
	uint32_t tag[MARSMD_MAX_TYPES * MARSMD_MAX_TYPES]; // [x type << 4 | y type]
	float4 tab[kMaxTables];        // {start, step_inv, size, offset} (size, offset as int bits)
	float4 coef[MARSMD_MAX_TYPES]; // {charge, softcore radius, softcore eps, 0}
	// per type pair: {start, step_inv, last, offset} of its enabled table, last = size - 2
	// (int bits; -2 when the pair has no enabled tabulated term): one read per pair on
	// the tabulated-only path. Indexed from one end of the pair, both orientations:
	// [o << 8 | own type << 4 | other type], o = 1 when the own end is the pair's x
	// (lower canonical index), so no type selects are needed per pair
	float4 pairtab[2 * MARSMD_MAX_TYPES * MARSMD_MAX_TYPES];
};

__device__ __forceinline__ uint32_t fb(float f) { return __float_as_uint(f); }

// One pair term on slot t (PairNonbondedComputer, Pairwise.h:184-305) in canonical
// orientation: r = wrap_diff(pos[higher canon] - pos[lower canon]), the lower takes -F.
// ANALYTIC adds the analytical terms enabled by the pair's tag; without it only the
// tabulated term is evaluated (the kernels are instantiated per case, so the
// tabulated-only path carries none of the analytical code). `knots` points into LDS or
// global memory, as the caller's instantiation says: a pointer that might be either
// would be read with flat loads, whose waits also drain the position gathers in flight.
template<bool ANALYTIC>
__device__ __forceinline__ void nb_pair(const float4& p, uint32_t ct, uint32_t tt, const float4& q,
										const NbTables& T, const float* __restrict__ knots, const NbLds& S,
										const DevBox& box, float cut_d, bool energy, float& fx, float& fy,
										float& fz, float& e) {
	const uint32_t mj = fb(q.w);
	const uint32_t tj = meta_type(mj);
	// ct < cj, from the meta words (canonical index in the high bits; a pad entry is the
	// slot itself: equal words, not lower)
	const bool t_is_x = fb(p.w) < mj;
	const float4 px = t_is_x ? p : q;
	const float4 py = t_is_x ? q : p;
	const uint32_t tx = t_is_x ? tt : tj;
	const uint32_t ty = t_is_x ? tj : tt;
	if constexpr (!ANALYTIC) {
		// Tabulated term only, without branches: every quantity is computed for every
		// pair (table reads clamped into the table) and a pair that the reference would
		// skip adds +-0, which leaves the sums' bits unchanged (they start at +0 and can
		// never become -0). Straight-line code lets the four pairs of a group overlap
		// their LDS reads.
		const float4 tp = S.pairtab[(t_is_x ? 256u : 0u) + ((tt << 4) | tj)];
		const int last = __float_as_int(tp.z); // table size - 2
		const bool tab = last >= 0;
		// py - px is q - p or its negation, exactly (rounding is sign-symmetric)
		const float dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
		const float3 r = wrap_diff_uniform(box, t_is_x ? make_float3(dx, dy, dz) : make_float3(-dx, -dy, -dz));
		// correctly rounded (d picks the table bin); distances below 1e-6 are skipped
		const float d = sqrt_rn_nonneg(len2(r));
		// the unit vector only scales the force: the hardware reciprocal (1 ulp) in place
		// of the correctly rounded division moves the force by ~1e-7 relative
		// (floored at 1e-6, so the zero-distance pads give a finite product with +0)
		const float inv = __builtin_amdgcn_rcpf(fmaxf(d, 1e-6f));
		// the pairs the reference evaluates (d >= 1e-6, within the cutoff); at d == 1e-6
		// exactly its unit vector is 0, so the force needs d > 1e-6
		const bool in_cut = tab && !(d > cut_d);
		// TabulatedPotential::compute (TabulatedPotential.h:19-52), non-periodic; a pair
		// without the term reads knots 0 and 1 (last -2 clamps the bin to 0) and adds 0
		const float* pot = knots + __float_as_int(tp.w);
		float w = (d - tp.x) * tp.y;
		const int home = int(floorf(w));
		w = w - float(home);
		const int hc = max(min(home, last), 0);
		const float U0 = pot[hc];
		const float U1 = pot[hc + 1];
		const float dU = U1 - U0;
		const bool beyond = home > last;
		// (the reference adds these to a +0 accumulator; without that add a -0 can reach
		// the sums only as a product, and adding -0 to a sum leaves its bits unchanged)
		const float fm = beyond ? 0.0f : -dU * tp.y;
		const float en = beyond ? U1 : dU * w + U0;
		// the lower canonical end takes -F: -(fm u) = (-fm) u exactly
		// (force over distance once, then per axis: the product rounds in another order
		// than fm * (r / d), a relative change of about 1e-7)
		const float g = ((in_cut && d > 1e-6f) ? (t_is_x ? -fm : fm) : 0.0f) * inv;
		fx += g * r.x;
		fy += g * r.y;
		fz += g * r.z;
		if (energy)
			e += (in_cut && !(d < 1e-6f)) ? en * 0.5f : 0.0f;
		return;
	}
	const uint32_t tag = S.tag[(tx << 4) | ty];
	const uint32_t terms = tag & T.enabled;
	if (terms == 0u)
		return;
	const float3 r = wrap_diff(box, sub3(py, px));
	const float r2 = len2(r);
	const float d = sqrtf(r2);
	float ux = 0.0f, uy = 0.0f, uz = 0.0f;
	if (d > 1e-6f) {
		const float inv = 1.0f / d;
		ux = inv * r.x;
		uy = inv * r.y;
		uz = inv * r.z;
	}
	if (d < 1e-6f)
		return;
	if (d > cut_d)
		return;
	float fm = 0.0f, en = 0.0f;
	if (ANALYTIC && (terms & kAnalyticTerms)) {
		const float4 cx = S.coef[tx], cy = S.coef[ty];
		analytic_terms(terms, d, r2, cx.x, cy.x, cx.y, cy.y, cx.z, cy.z, T.coef.prm, fm, en);
	}
	if (terms & MARSMD_PAIR_TERM_TABULATED) {
		const float4 tp = S.tab[tag >> MARSMD_PAIR_TABLE_SHIFT];
		const FE v = table_compute(d, knots + __float_as_int(tp.w), __float_as_int(tp.z), tp.x, tp.y, false);
		fm += v.f;
		en += v.e;
	}
	const float gx = fm * ux, gy = fm * uy, gz = fm * uz;
	if (t_is_x) {
		fx += -gx;
		fy += -gy;
		fz += -gz;
	} else {
		fx += gx;
		fy += gy;
		fz += gz;
	}
	if (energy)
		e += en * 0.5f;
}

// Position record of slot j, addressed by a 32-bit byte offset from the array base
// (scalar base + vector offset addressing; every caller holds fewer than 2^28 slots).
__device__ __forceinline__ float4 pos_at(const float4* __restrict__ pos, uint32_t j) {
	return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(pos) + (j << 4));
}

// Sum of the nonbonded pair terms on slot t over its neighbour list, in list order.
// A lane's list is read in groups of 4 indices (one 16-byte load) and pipelined three
// deep: while group g is evaluated, the positions of group g+1 and the indices of
// group g+2 are in flight. The padding entries of a row are t itself (skipped as d = 0).
template<bool ANALYTIC = false>
__device__ __forceinline__ float4 nb_force_list(uint32_t t, const float4& p, const float4* __restrict__ pos,
											   const NbrList& L, const NbTables& T, const float* __restrict__ knots,
											   const NbLds& S, const DevBox& box, float cut_d, bool energy) {
	const uint32_t mt = fb(p.w);
	const uint32_t ct = meta_canon(mt);
	const uint32_t tt = meta_type(mt);
	float fx = 0.0f, fy = 0.0f, fz = 0.0f, e = 0.0f;
	const uint32_t ng = (L.count[t] + 3u) >> 2;
	if (ng == 0)
		return make_float4(fx, fy, fz, e);
	const uint4* __restrict__ row = reinterpret_cast<const uint4*>(L.nbr + L.tile_base[t >> 6]) + (t & 63u);
	// two register sets (a, b) take turns: while one group is evaluated the positions
	// of the next are in flight in the other set, and the indices of the one after;
	// each index register is reloaded in place once its gathers are issued
	uint4 ja = row[0];
	float4 a0 = pos_at(pos, ja.x), a1 = pos_at(pos, ja.y), a2 = pos_at(pos, ja.z), a3 = pos_at(pos, ja.w);
	float4 b0, b1, b2, b3;
	uint4 jb = ng > 1 ? row[64] : ja;
	for (uint32_t g = 0;; g += 2) {
		if (g + 1 < ng) {
			b0 = pos_at(pos, jb.x);
			b1 = pos_at(pos, jb.y);
			b2 = pos_at(pos, jb.z);
			b3 = pos_at(pos, jb.w);
			if (g + 2 < ng)
				ja = row[uint64_t(g + 2) * 64u];
		}
		nb_pair<ANALYTIC>(p, ct, tt, a0, T, knots, S, box, cut_d, energy, fx, fy, fz, e);
		nb_pair<ANALYTIC>(p, ct, tt, a1, T, knots, S, box, cut_d, energy, fx, fy, fz, e);
		nb_pair<ANALYTIC>(p, ct, tt, a2, T, knots, S, box, cut_d, energy, fx, fy, fz, e);
		nb_pair<ANALYTIC>(p, ct, tt, a3, T, knots, S, box, cut_d, energy, fx, fy, fz, e);
		if (g + 1 >= ng)
			break;
		if (g + 2 < ng) {
			a0 = pos_at(pos, ja.x);
			a1 = pos_at(pos, ja.y);
			a2 = pos_at(pos, ja.z);
			a3 = pos_at(pos, ja.w);
			if (g + 3 < ng)
				jb = row[uint64_t(g + 3) * 64u];
		}
		nb_pair<ANALYTIC>(p, ct, tt, b0, T, knots, S, box, cut_d, energy, fx, fy, fz, e);
		nb_pair<ANALYTIC>(p, ct, tt, b1, T, knots, S, box, cut_d, energy, fx, fy, fz, e);
		nb_pair<ANALYTIC>(p, ct, tt, b2, T, knots, S, box, cut_d, energy, fx, fy, fz, e);
		nb_pair<ANALYTIC>(p, ct, tt, b3, T, knots, S, box, cut_d, energy, fx, fy, fz, e);
		if (g + 2 >= ng)
			break;
	}
	return make_float4(fx, fy, fz, e);
}

// Stage the tag matrix, the table parameters and the per-type coefficients in LDS (the
// caller synchronises).
__device__ __forceinline__ void stage_params(const NbTables& T, NbLds& S) {
	const int nt2 = T.num_types * T.num_types;
	for (int i = threadIdx.x; i < nt2; i += blockDim.x) {
		const uint32_t tag = T.tag[i];
		const int x = i / T.num_types, y = i % T.num_types;
		S.tag[(x << 4) | y] = tag;
		const int ti = int(tag >> MARSMD_PAIR_TABLE_SHIFT);
		const float4 v = (tag & T.enabled & MARSMD_PAIR_TERM_TABULATED) && ti < T.num_tables
							 ? make_float4(T.start[ti], T.step_inv[ti], __int_as_float(T.size[ti] - 2),
										   __int_as_float(T.offset[ti]))
							 : make_float4(0.0f, 0.0f, __int_as_float(-2), __int_as_float(0));
		S.pairtab[256 + ((x << 4) | y)] = v; // seen from x
		S.pairtab[(y << 4) | x] = v;         // seen from y
	}
	for (int i = threadIdx.x; i < T.num_tables; i += blockDim.x)
		S.tab[i] = make_float4(T.start[i], T.step_inv[i], __int_as_float(T.size[i]), __int_as_float(T.offset[i]));
	if (T.num_tables == 0 && threadIdx.x == 0) // pairs without the tabulated term read table 0
		S.tab[0] = make_float4(0.0f, 0.0f, __int_as_float(2), __int_as_float(0));
	for (int i = threadIdx.x; i < T.num_types; i += blockDim.x)
		S.coef[i] = make_float4(T.coef.charge[i], T.coef.radius[i], T.coef.eps[i], 0.0f);
}

// Copy the particle-type record into LDS: the integrators and the PMF index it per
// lane (mass, damping, diffusion, charge, grid terms), which from the kernel-argument
// block would be vector loads (the caller synchronises).
__device__ __forceinline__ void stage_types(const marsmd_particle_types& src, marsmd_particle_types& dst) {
	static_assert(sizeof(marsmd_particle_types) % 4 == 0, "4-byte fields");
	const uint32_t* a = reinterpret_cast<const uint32_t*>(&src);
	uint32_t* b = reinterpret_cast<uint32_t*>(&dst);
	for (int i = threadIdx.x; i < int(sizeof(marsmd_particle_types) / 4); i += blockDim.x)
		b[i] = a[i];
}

// Stage table knots (when they fit) and the pair parameters in LDS.
__device__ __forceinline__ void stage_tables(const NbTables& T, const float* __restrict__ knots_g, float* lds_knots,
											 NbLds& S) {
	stage_params(T, S);
	if (T.lds) {
		for (int i = threadIdx.x; i < T.total_knots; i += blockDim.x)
			lds_knots[i] = knots_g[i];
	}
	__syncthreads();
}

// nb_force_list with the knots read from LDS when staged there, else from global
// memory: each branch sees one address space.
template<bool ANALYTIC = false>
__device__ __forceinline__ float4 nb_force(uint32_t t, const float4& p, const float4* __restrict__ pos,
										   const NbrList& L, const NbTables& T, const float* lds_knots,
										   const float* __restrict__ knots_g, const NbLds& S, const DevBox& box,
										   float cut_d, bool energy) {
	if (T.lds)
		return nb_force_list<ANALYTIC>(t, p, pos, L, T, lds_knots, S, box, cut_d, energy);
	return nb_force_list<ANALYTIC>(t, p, pos, L, T, knots_g, S, box, cut_d, energy);
}

} // namespace mmd
