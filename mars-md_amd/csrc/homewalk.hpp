// homewalk.hpp — the home-cell pairlist walk (see k_walk_home) and the reference's
// canonical pair predicate, shared by the pairlist (pairlist.hip) and the replica batch
// (batch.hip).
#pragma once

#include "pairlist.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace mmd {

namespace {

// Canonical pair predicate of ZOrderCellNeighborKernel for sorted lo < hi
// (ZOrderNeighbor.h:204-221): wrap_diff(pos_hi - pos_lo)^2 <= rc^2, not excluded by
// lo's CSR row or a shared rigid body.
__device__ __forceinline__ bool pair_predicate(const DevBox& box, float rc2, const marsmd_exclusion_view& ex,
											   const float4& plo, const float4& phi) {
	const float3 dr = wrap_diff(box, sub3(phi, plo));
	if (!(len2(dr) <= rc2))
		return false;
	const int a = int(meta_canon(__float_as_uint(plo.w)));
	const int b = int(meta_canon(__float_as_uint(phi.w)));
	if (ex.rigid_body_id) {
		const int body = ex.rigid_body_id[a];
		if (body >= 0 && ex.rigid_body_id[b] == body)
			return false;
	}
	if (ex.offsets && a >= 0 && uint64_t(a) < ex.num_particles) {
		const int e0 = ex.offsets[a], e1 = ex.offsets[a + 1];
		for (int e = e0; e < e1; ++e)
			if (ex.excluded[e] == b)
				return false;
	}
	return true;
}

// pair_predicate for a home particle pi (slot ti) and a candidate q (slot j != ti),
// oriented by slot: the difference phi - plo is q - pi or its exact negation (rounding
// is sign-symmetric), so the orientation costs three selects; the metadata of the two
// ends is only looked at when exclusions or rigid bodies are given.
__device__ __forceinline__ bool pair_predicate_home(const DevBox& box, float rc2, const marsmd_exclusion_view& ex,
													const float4& pi, uint32_t ti, const float4& q, uint32_t j) {
	const bool t_lo = j > ti;
	const float dx = q.x - pi.x, dy = q.y - pi.y, dz = q.z - pi.z;
	const float3 dr = wrap_diff(box, t_lo ? make_float3(dx, dy, dz) : make_float3(-dx, -dy, -dz));
	if (!(len2(dr) <= rc2))
		return false;
	if (!ex.rigid_body_id && !ex.offsets)
		return true;
	return pair_predicate(box, rc2, ex, t_lo ? pi : q, t_lo ? q : pi);
}
// ------------------------------------------------------------ home-cell walk
// One wave per home cell: a Morton cell of the home level h, whose particles are
// contiguous in sorted order. The wave enumerates the fine cells (level m = h + s)
// within `reach` fine cells of the home cell, keeps those whose box can lie within rc
// of the home cell's box (both widened by the quantisation margin, so the test is
// conservative across the periodic seam), stores their particle ranges as segments
// in LDS, then streams the candidates 64 at a time (one coalesced load per lane) and
// tests them against the wave's home particles, which are broadcast from LDS.
// Predicate and orientation are k_walk's (the reference's), so the pair set is the
// same; a particle's list is in candidate order. Entries past `cap` are not written
// (the host sees the count and walks again with a larger cap).
constexpr int kHomeWaves = 4;
constexpr int kSegMax = 448;
// LDS per wave for the home particles: 64 positions, then 64 candidate slots
constexpr int kHomeLds = 64 + 16;
// LDS per wave for the queue of candidates kept by the bounding-box test
constexpr int kHomeKeep = 128;

struct HomeParams {
	const float4* spos;
	const uint32_t* home_begin;
	const uint32_t* home_end;
	const uint32_t* fine_begin;
	const uint32_t* fine_end;
	uint32_t num_home;
	int h;         // home level
	int s;         // fine level m = h + s
	int reach[3];  // fine cells enumerated beyond the home cell's span
	int full[3];   // periodic axis enumerated whole at the fine level
	float cw[3];   // fine cell width (quantised)
	float margin[3];
	float rc2;
	DevBox box;
	marsmd_exclusion_view ex;
	uint32_t cap;  // entries per slot the uniform tile rows hold
	// When the reference cell grid misses pairs at the periodic seam, only the pairs its
	// stencil reaches are kept: cells (quantised coordinate >> ref_shift) within
	// ref_r per axis, wrapped on periodic axes; ref_full marks axes it enumerates whole.
	int ref_filter;
	int ref_shift;
	int ref_cells;
	int ref_r[3];
	int ref_full[3];
	float dom_lo[3], dom_hi[3];
	// Home-frame prefilter (pre != 0): each candidate is moved by whole box lengths next
	// to the wave's first home particle, and its plain squared distance d2' to a home
	// particle decides the pair when it is below pre_lo2 (in) or above pre_hi2 (out);
	// between the two, which are rc^2 -+ a bound on |d2' - d2| over pairs near the
	// cutoff, the reference's predicate decides. Set only where the box is wide enough
	// for that move to give the minimum image of every pair within the cutoff.
	int pre;
	float pre_lo2, pre_hi2;
	// The fine cells the walk enumerates, when given: per entry the cell offset from the
	// home cell's first fine cell on each axis (the absolute cell on an axis enumerated
	// whole), in enumeration order, for the cells that pass the gap test; null: every
	// cell of the reach is enumerated and gap-tested per home cell.
	const int4* stencil;
	uint32_t stencil_n;
};

// The fine cells of the walk's enumeration that pass its gap test, as offsets (see
// HomeParams::stencil), in its order (x fastest); empty when there are more than
// max_cells of them. The test is the walk's own: per axis not enumerated whole, the
// number of cells between the cell and the home cell's span times the cell width, less
// twice the margin, floored at 0, squared and summed, against rc^2.
inline std::vector<int4> home_stencil(int h, int s, const int reach[3], const int full[3], const float cw[3],
									  const float margin[3], float rc2, size_t max_cells) {
	const int fine_cells = 1 << (h + s), span = 1 << s;
	int lo[3], nax[3];
	for (int a = 0; a < 3; ++a) {
		lo[a] = full[a] ? 0 : -reach[a];
		nax[a] = full[a] ? fine_cells : span + 2 * reach[a];
	}
	std::vector<int4> out;
	for (int z = 0; z < nax[2]; ++z)
		for (int y = 0; y < nax[1]; ++y)
			for (int x = 0; x < nax[0]; ++x) {
				const int o[3] = {lo[0] + x, lo[1] + y, lo[2] + z};
				float gap2 = 0.0f;
				for (int a = 0; a < 3; ++a) {
					if (full[a])
						continue;
					const int apart = o[a] < 0 ? -o[a] - 1 : (o[a] >= span ? o[a] - span : 0);
					const float g = std::fmax(0.0f, float(apart) * cw[a] - 2.0f * margin[a]);
					gap2 += g * g;
				}
				if (gap2 <= rc2) {
					if (out.size() == max_cells)
						return {};
					out.push_back(make_int4(o[0], o[1], o[2], 0));
				}
			}
	// Morton order of the offsets: 64 consecutive candidates then come from a compact
	// cluster of cells, which a home particle's cutoff sphere often misses as a whole
	auto key = [](const int4& c) {
		uint64_t k = 0;
		const uint32_t v[3] = {uint32_t(c.x + 1024), uint32_t(c.y + 1024), uint32_t(c.z + 1024)};
		for (int b = 10; b >= 0; --b)
			for (int a = 0; a < 3; ++a)
				k = (k << 1) | ((v[a] >> b) & 1u);
		return k;
	};
	std::stable_sort(out.begin(), out.end(), [&](const int4& a, const int4& b) { return key(a) < key(b); });
	return out;
}

// The prefilter's band for a domain [lo, hi) whose home cells (level h) have width
// extent / 2^h: false when a periodic axis is too short for the home frame to give
// minimum images (half the box must exceed a home cell plus the cutoff, with slack).
// The bound on |d2' - d2|: the two differences of an axis differ by a few roundings at
// the coordinates' magnitude C (|x| <= C), i.e. a few ulp(C), over |d| <= rc + 1, plus
// the roundings of the squares and sums at rc^2.
inline bool home_prefilter_band(const marsmd_box& box, const float lo[3], const float hi[3], int h, float rc,
								float* lo2, float* hi2) {
	float C = 0.0f;
	for (int a = 0; a < 3; ++a) {
		const float w = (hi[a] - lo[a]) / float(1 << h);
		C = std::fmax(C, std::fmax(std::fabs(lo[a]), std::fabs(hi[a])));
		if (box.periodic[a] && !(0.5f * box.size[a] > w + rc + 0.01f * box.size[a] + 1.0f))
			return false;
		C = std::fmax(C, box.size[a]);
	}
	const float rc2 = rc * rc;
	const float ulp_c = std::nextafter(C, INFINITY) - C;
	const float ulp_r = std::nextafter(rc2 + 1.0f, INFINITY) - (rc2 + 1.0f);
	const float eps = 64.0f * ulp_c * (rc + 1.0f) + 16.0f * ulp_r;
	if (!(eps < 0.01f * rc2))
		return false;
	*lo2 = rc2 - eps;
	*hi2 = rc2 + eps;
	return true;
}

// Whether the reference grid's stencil (BuildCellNeighborsKernel, ZOrderNeighbor.h:
// 55-139) links the cells of a and b. The relation is symmetric, so it is the pair
// set of ZOrderCellNeighborKernel whichever of the two emits.
__device__ __forceinline__ bool in_reference_stencil(const HomeParams& P, const float4& a, const float4& b) {
	const float pa[3] = {a.x, a.y, a.z}, pb[3] = {b.x, b.y, b.z};
	bool clamped = false;
#pragma unroll
	for (int k = 0; k < 3; ++k) {
		if (P.ref_full[k])
			continue;
		const int ca = int(morton_quantize(pa[k], P.dom_lo[k], P.dom_hi[k], clamped) >> P.ref_shift);
		const int cb = int(morton_quantize(pb[k], P.dom_lo[k], P.dom_hi[k], clamped) >> P.ref_shift);
		int d = ca > cb ? ca - cb : cb - ca;
		if (P.box.periodic[k])
			d = min(d, P.ref_cells - d);
		if (d > P.ref_r[k])
			return false;
	}
	return true;
}

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
	const int lane = threadIdx.x & 63;
#pragma unroll
	for (int o = 1; o < 64; o <<= 1) {
		const uint32_t t = __shfl_up(v, o, 64);
		if (lane >= o)
			v += t;
	}
	return v;
}
__device__ __forceinline__ uint32_t bcast_u32(uint32_t v, uint32_t lane) {
	return uint32_t(__builtin_amdgcn_readlane(int(v), int(lane)));
}
__device__ __forceinline__ float bcast_f32(float v, uint32_t lane) {
	return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), int(lane)));
}
// LDS written by some lanes of this wave and read by others: order the accesses.
__device__ __forceinline__ void wave_lds_sync() {
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The hits of the wave's nh home particles (visited one by one from LDS, the next one
// read ahead) among one block of up to 64 candidates (lane: slot j, position q) by the
// reference's predicate: the mask over the block's lanes of home particle i, in lane i.
// Lane i (< 64, wave-uniform) of the 64-bit per-lane mask {lo, hi} takes the wave-uniform
// mask m: one v_writelane per half, the lane in M0 (a select over lanes would take a
// compare, two moves of the mask into vector registers and two selects).
__device__ __forceinline__ void set_lane_mask(int& lo, int& hi, uint64_t m, uint32_t i) {
	const uint32_t mlo = uint32_t(m), mhi = uint32_t(m >> 32);
	asm("v_writelane_b32 %0, %1, %2" : "+v"(lo) : "s"(mlo), "{m0}"(i));
	asm("v_writelane_b32 %0, %1, %2" : "+v"(hi) : "s"(mhi), "{m0}"(i));
}
__device__ __forceinline__ uint64_t lane_mask(int lo, int hi) {
	return uint64_t(uint32_t(lo)) | (uint64_t(uint32_t(hi)) << 32);
}

__device__ __forceinline__ uint64_t exact_hit_masks(const HomeParams& P, const float4* home_pos, uint32_t nh,
													 uint32_t c0s, bool live, uint32_t j, const float4& q) {
	int mine_lo = 0, mine_hi = 0;
	float4 pn = home_pos[0];
	for (uint32_t i = 0; i < nh; ++i) {
		const float4 pi = pn;
		pn = home_pos[(i + 1u) & 63u];
		const uint32_t ti = c0s + i;
		bool in = live && j != ti && pair_predicate_home(P.box, P.rc2, P.ex, pi, ti, q, j);
		if (in && P.ref_filter)
			in = in_reference_stencil(P, pi, q); // symmetric in its two ends
		set_lane_mask(mine_lo, mine_hi, __ballot(in), i);
	}
	return lane_mask(mine_lo, mine_hi);
}

// The same masks from the home-frame prefilter (the candidate moved into the home frame
// as qx, qy, qz; +inf on lanes without a candidate): a pair with d2' <= pre_lo2 is in,
// one with d2' > pre_hi2 out. A home particle's own slot among the candidates is not
// excluded here (the caller skips it). When any pair of the block falls between the two
// bounds the block is decided again by exact_hit_masks.
__device__ __forceinline__ uint64_t home_frame_hit_masks(const HomeParams& P, const float4* home_pos, uint32_t nh,
														 uint32_t c0s, bool live, uint32_t j, const float4& q,
														 float qx, float qy, float qz) {
	int mine_lo = 0, mine_hi = 0;
	uint64_t band = 0;
	float4 pn = home_pos[0];
	for (uint32_t i = 0; i < nh; ++i) {
		const float4 pi = pn;
		pn = home_pos[(i + 1u) & 63u];
		const float dx = qx - pi.x, dy = qy - pi.y, dz = qz - pi.z;
		const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
		const uint64_t hits = __ballot(d2 <= P.pre_lo2);
		band |= __ballot(d2 <= P.pre_hi2) & ~hits;
		set_lane_mask(mine_lo, mine_hi, hits, i);
	}
	uint64_t mine = lane_mask(mine_lo, mine_hi);
	if (band)
		mine = exact_hit_masks(P, home_pos, nh, c0s, live, j, q);
	return mine;
}

// One block of candidates against the wave's home particles: the hit masks (lane i:
// home particle i's hits over the block's lanes), then every home lane appends its hits
// to its row in lane (candidate) order, reading their slots from `cand` (LDS).
template<bool FAST, bool FILL>
__device__ __forceinline__ void test_candidate_block(const HomeParams& P, const float4* home_pos, uint32_t* cand,
													 uint32_t nh, uint32_t c0, bool live, uint32_t j, const float4& q,
													 float qx, float qy, float qz, uint32_t& cnt,
													 uint32_t* __restrict__ nbr) {
	const uint32_t lane = threadIdx.x & 63u;
	const uint32_t c0s = uint32_t(__builtin_amdgcn_readfirstlane(int(c0))); // wave-uniform
	cand[lane] = j;
	wave_lds_sync();
	const uint64_t mine = FAST ? home_frame_hit_masks(P, home_pos, nh, c0s, live, j, q, qx, qy, qz)
							   : exact_hit_masks(P, home_pos, nh, c0s, live, j, q);
	if (lane < nh) {
		const uint32_t t = c0s + lane;
		uint32_t* __restrict__ row = nbr + (uint64_t(t) >> 6) * 64u * P.cap + (t & 63u) * 4u;
		uint64_t m = mine;
		while (m) {
			const uint32_t b = uint32_t(__builtin_ctzll(m));
			m &= m - 1;
			const uint32_t jb = cand[b];
			if (FAST && jb == t) // never a pair with itself
				continue;
			if (FILL && cnt < P.cap) // entry cnt of slot t (nbr_slot_index)
				row[(cnt >> 2) * 256u + (cnt & 3u)] = jb;
			++cnt;
		}
	}
	wave_lds_sync();
}

// The home-frame coordinates of a candidate at q (see candidate_at).
__device__ __forceinline__ void home_frame(const HomeParams& P, bool live, bool frame, const float4& p0,
										   const float4& q, float& qx, float& qy, float& qz) {
	if (!frame) {
		qx = q.x;
		qy = q.y;
		qz = q.z;
		return;
	}
	qx = qy = qz = INFINITY;
	if (live) {
		qx = P.box.periodic[0] ? q.x - P.box.L[0] * rintf((q.x - p0.x) * P.box.inv_L[0]) : q.x;
		qy = P.box.periodic[1] ? q.y - P.box.L[1] * rintf((q.y - p0.y) * P.box.inv_L[1]) : q.y;
		qz = P.box.periodic[2] ? q.z - P.box.L[2] * rintf((q.z - p0.z) * P.box.inv_L[2]) : q.z;
	}
}

// Candidate qi of segments [0, nseg) (ntot particles): its slot j and position q, and q
// in the home frame (moved by whole box lengths next to the anchor p0, so the image is
// exact; +inf on lanes without a candidate) when `frame`.
__device__ __forceinline__ void candidate_at(const HomeParams& P, const uint32_t* seg_b, const uint32_t* seg_pre,
											 uint32_t nseg, uint32_t qi, bool live, bool frame, const float4& p0,
											 uint32_t& j, float4& q, float& qx, float& qy, float& qz) {
	j = 0xFFFFFFFFu;
	q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	if (live) {
		uint32_t lo = 0, hi = nseg; // seg_pre[lo] <= qi < seg_pre[hi]
		while (hi - lo > 1) {
			const uint32_t mid = (lo + hi) >> 1;
			if (seg_pre[mid] <= qi)
				lo = mid;
			else
				hi = mid;
		}
		j = seg_b[lo] + (qi - seg_pre[lo]);
		q = P.spos[j];
	}
	home_frame(P, live, frame, p0, q, qx, qy, qz);
}

__device__ __forceinline__ float wave_min(float v) {
	for (int o = 32; o > 0; o >>= 1)
		v = fminf(v, __shfl_xor(v, o));
	return v;
}
__device__ __forceinline__ float wave_max(float v) {
	for (int o = 32; o > 0; o >>= 1)
		v = fmaxf(v, __shfl_xor(v, o));
	return v;
}

// Tests the candidates of segments [0, nseg) (ntot particles) against the wave's home
// particles (lane l holds home particle t = c0 + l when valid, with its count) and
// records the hits. The candidates are spread over the lanes, 64 at a time, and the
// home particles are visited one by one from LDS: each (home particle, candidate) pair
// is evaluated by exactly one lane, so a wave whose home cell holds fewer than 64
// particles does not idle lanes on every test. Every row lists its neighbours in
// candidate order.
// With the home-frame prefilter, a candidate whose squared distance to the bounding box
// of the home particles exceeds pre_hi2 is dropped first (the fine cells of the stencil
// reach past that box: at 1M beads most candidates are dropped): that distance is at
// most d2' to every home particle, computed with the same monotone operations, so each
// dropped candidate is one the prefilter would decide "out" for every home particle.
// The kept ones (in candidate order, through a 128-entry LDS queue `keep`) are tested
// 64 at a time, so the pair set and each row's order are unchanged.
template<bool FILL>
__device__ __forceinline__ void home_test_segments(const HomeParams& P, const uint32_t* seg_b, uint32_t* seg_pre,
												   uint32_t nseg, uint32_t ntot, bool valid, uint32_t t,
												   const float4& p, uint32_t& cnt, uint32_t* __restrict__ nbr,
												   float4* home_pos, uint32_t* keep) {
	const uint32_t lane = threadIdx.x & 63u;
	if (lane == 0)
		seg_pre[nseg] = ntot;
	home_pos[lane] = p;
	wave_lds_sync();
	uint32_t* cand = reinterpret_cast<uint32_t*>(home_pos + 64); // the block's candidate slots
	const uint32_t c0 = t - lane;                       // first home slot of this chunk
	const uint32_t nh = uint32_t(__popcll(__ballot(valid))); // valid lanes are a prefix
	const float4 p0 = home_pos[0];                       // the home frame's anchor
	const bool fast = P.pre && !P.ref_filter && !P.ex.rigid_body_id && !P.ex.offsets;
	if (!fast) {
		for (uint32_t q0 = 0; q0 < ntot; q0 += 64) {
			const bool live = q0 + lane < ntot;
			uint32_t j;
			float4 q;
			float qx, qy, qz;
			candidate_at(P, seg_b, seg_pre, nseg, q0 + lane, live, false, p0, j, q, qx, qy, qz);
			test_candidate_block<false, FILL>(P, home_pos, cand, nh, c0, live, j, q, qx, qy, qz, cnt, nbr);
		}
		wave_lds_sync();
		return;
	}
	const float lox = wave_min(valid ? p.x : INFINITY), hix = wave_max(valid ? p.x : -INFINITY);
	const float loy = wave_min(valid ? p.y : INFINITY), hiy = wave_max(valid ? p.y : -INFINITY);
	const float loz = wave_min(valid ? p.z : INFINITY), hiz = wave_max(valid ? p.z : -INFINITY);
	const uint64_t lt = (uint64_t(1) << lane) - 1;
	uint32_t nk = 0; // queued candidates (wave-uniform), fewer than 64 between blocks
	for (uint32_t q0 = 0; q0 < ntot; q0 += 64) {
		const bool live = q0 + lane < ntot;
		uint32_t j;
		float4 q;
		float qx, qy, qz;
		candidate_at(P, seg_b, seg_pre, nseg, q0 + lane, live, true, p0, j, q, qx, qy, qz);
		const float ax = fmaxf(fmaxf(lox - qx, qx - hix), 0.0f);
		const float ay = fmaxf(fmaxf(loy - qy, qy - hiy), 0.0f);
		const float az = fmaxf(fmaxf(loz - qz, qz - hiz), 0.0f);
		const bool kept = live && !(fmaf(ax, ax, fmaf(ay, ay, az * az)) > P.pre_hi2);
		const uint64_t km = __ballot(kept);
		if (kept)
			keep[nk + uint32_t(__popcll(km & lt))] = j;
		nk += uint32_t(__popcll(km));
		if (nk >= 64) {
			wave_lds_sync();
			{
				const uint32_t jb = keep[lane];
				const float4 qb = P.spos[jb];
				float bx, by, bz;
				home_frame(P, true, true, p0, qb, bx, by, bz);
				test_candidate_block<true, FILL>(P, home_pos, cand, nh, c0, true, jb, qb, bx, by, bz, cnt, nbr);
			}
			const uint32_t rest = nk - 64;
			const uint32_t v = lane < rest ? keep[64 + lane] : 0u;
			wave_lds_sync();
			if (lane < rest)
				keep[lane] = v;
			nk = rest;
			wave_lds_sync();
		}
	}
	if (nk > 0) {
		wave_lds_sync();
		const bool live = lane < nk;
		const uint32_t jb = live ? keep[lane] : 0xFFFFFFFFu;
		const float4 qb = live ? P.spos[jb] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		float bx, by, bz;
		home_frame(P, live, true, p0, qb, bx, by, bz);
		test_candidate_block<true, FILL>(P, home_pos, cand, nh, c0, live, jb, qb, bx, by, bz, cnt, nbr);
	}
	wave_lds_sync();
}

template<bool FILL>
__device__ void walk_home_cell(const HomeParams& P, uint32_t home, uint32_t* seg_b, uint32_t* seg_pre, float4* home_pos,
							   uint32_t* keep, uint32_t* __restrict__ counts, uint32_t* __restrict__ nbr) {
	const int lane = threadIdx.x & 63;
	const uint32_t hb = P.home_begin[home], he = P.home_end[home];
	if (hb >= he)
		return;
	const int hc[3] = {int(compact_by3(home >> 2)), int(compact_by3(home >> 1)), int(compact_by3(home))};
	const int fine_cells = 1 << (P.h + P.s);
	const int span = 1 << P.s;
	int lo[3], nax[3];
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		if (P.full[a]) {
			lo[a] = 0;
			nax[a] = fine_cells;
		} else {
			int l = hc[a] * span - P.reach[a];
			int hgh = hc[a] * span + span - 1 + P.reach[a];
			if (!P.box.periodic[a]) { // open axis: only cells that exist (the reach can be huge
				l = max(l, 0);        // when the extent is degenerate)
				hgh = min(hgh, fine_cells - 1);
			}
			lo[a] = l;
			nax[a] = hgh - l + 1;
		}
	}
	const int total = P.stencil ? int(P.stencil_n) : nax[0] * nax[1] * nax[2];
	const uint64_t lt_mask = (uint64_t(1) << lane) - 1;
	// The segments do not depend on the chunk of home particles: when the first chunk's
	// enumeration fits LDS whole, the later chunks test the same segments.
	uint32_t nseg = 0, ntot = 0; // wave-uniform
	bool cached = false;
	for (uint32_t c0 = hb; c0 < he; c0 += 64) {
		const uint32_t t = c0 + uint32_t(lane);
		const bool valid = t < he;
		const float4 p = valid ? P.spos[t] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		uint32_t cnt = 0;
		if (!cached) {
			nseg = 0;
			ntot = 0;
		}
		bool flushed = false;
		for (int cb = cached ? total : 0; cb < total; cb += 64) {
			const int k = cb + lane;
			uint32_t b = 0, e = 0;
			if (P.stencil && k < total) {
				const int4 o = P.stencil[k];
				const int off[3] = {o.x, o.y, o.z};
				bool keep = true;
				uint32_t code = 0;
#pragma unroll
				for (int a = 0; a < 3; ++a) {
					int u = P.full[a] ? off[a] : hc[a] * span + off[a];
					if (P.box.periodic[a])
						u = (u + fine_cells) & (fine_cells - 1);
					else if (u < 0 || u >= fine_cells)
						keep = false;
					code |= split_by3(uint32_t(u) & 1023u) << (2 - a);
				}
				if (keep) {
					b = P.fine_begin[code];
					e = P.fine_end[code];
				}
			} else if (k < total) {
				const int iv[3] = {k % nax[0], (k / nax[0]) % nax[1], k / (nax[0] * nax[1])};
				float gap2 = 0.0f;
				bool keep = true;
				uint32_t code = 0;
#pragma unroll
				for (int a = 0; a < 3; ++a) {
					int u = lo[a] + iv[a];
					if (!P.full[a]) {
						const int h0 = hc[a] * span;
						const int apart = u < h0 ? h0 - u - 1 : (u >= h0 + span ? u - h0 - span : 0);
						const float g = fmaxf(0.0f, float(apart) * P.cw[a] - 2.0f * P.margin[a]);
						gap2 += g * g;
						if (P.box.periodic[a])
							u = (u + fine_cells) & (fine_cells - 1);
						else if (u < 0 || u >= fine_cells)
							keep = false;
					}
					code |= split_by3(uint32_t(u) & 1023u) << (2 - a);
				}
				if (keep && gap2 <= P.rc2) {
					b = P.fine_begin[code];
					e = P.fine_end[code];
				}
			}
			const uint32_t c = e > b ? e - b : 0u;
			const uint64_t mask = __ballot(c > 0);
			const uint32_t nnew = uint32_t(__popcll(mask));
			if (nseg + nnew > uint32_t(kSegMax)) {
				home_test_segments<FILL>(P, seg_b, seg_pre, nseg, ntot, valid, t, p, cnt, nbr, home_pos, keep);
				nseg = 0;
				ntot = 0;
				flushed = true;
			}
			const uint32_t incl = wave_inclusive_scan(c);
			const uint32_t sum = bcast_u32(incl, 63);
			if (c > 0) {
				const uint32_t r = nseg + uint32_t(__popcll(mask & lt_mask));
				seg_b[r] = b;
				seg_pre[r] = ntot + incl - c;
			}
			nseg += nnew;
			ntot += sum;
		}
		home_test_segments<FILL>(P, seg_b, seg_pre, nseg, ntot, valid, t, p, cnt, nbr, home_pos, keep);
		cached = !flushed;
		if (valid) {
			counts[t] = cnt;
			if (FILL && cnt <= P.cap) {
				const uint64_t row = (uint64_t(t) >> 6) * 64u * P.cap;
				for (uint32_t k = cnt; (k & 3u) != 0; ++k)
					nbr[nbr_slot_index(row, t, k)] = t;
			}
		}
	}
}

// Bounded to 80 VGPRs (6 waves per SIMD, a few dwords of spill): the home-particle loop
// waits on an LDS read per home particle, which more waves hide (re-walk at 1M beads
// 0.48 -> 0.46 ms against 90 VGPRs and 5 waves).
template<bool FILL>
__global__ __launch_bounds__(256, 6) void k_walk_home(HomeParams P, uint32_t* __restrict__ counts,
												   uint32_t* __restrict__ nbr) {
	__shared__ uint32_t seg_b[kHomeWaves][kSegMax];
	__shared__ uint32_t seg_pre[kHomeWaves][kSegMax + 1];
	__shared__ float4 home_pos[kHomeWaves][kHomeLds];
	__shared__ uint32_t keep[kHomeWaves][kHomeKeep];
	const int wv = threadIdx.x >> 6;
	const TileRange R = xcd_tiles((P.num_home + kHomeWaves - 1) / kHomeWaves);
	for (uint32_t g = R.tile; g < R.end; g += R.stride) {
		const uint32_t home = g * kHomeWaves + uint32_t(wv);
		if (home < P.num_home)
			walk_home_cell<FILL>(P, home, seg_b[wv], seg_pre[wv], home_pos[wv], keep[wv], counts, nbr);
	}
}

} // namespace

} // namespace mmd
