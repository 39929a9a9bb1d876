// pairlist.hip — Z-order sort and the cell-stencil pairlist on gfx950.
//
// Reference path (src/PatchOperation/PairListKernels/ZOrderPairlist.cpp:102-366,
// ZOrderNeighbor.h:27-225): Morton-encode, stable sort, gather sorted positions,
// cell ranges, then a per-particle stencil walk emitting every pair (i < j in sorted
// order) with wrap_diff(pos_j - pos_i)^2 <= rc^2, minus pairs excluded by the
// emitting particle's CSR row or a shared rigid body, as int2 in canonical indices.
//
// MI355X design:
//   * one lane per sorted particle walks a per-axis stencil whose cells are pruned
//     against the particle's own position (a cell is skipped only when no point of
//     it can lie within rc, with a margin covering Morton quantisation), on a cell
//     grid chosen for this walk; the pair set is grid independent whenever the
//     reference grid covers the cutoff (checked on the host per build, see
//     choose_grids), otherwise the reference grid is walked unpruned;
//   * each lane records its FULL neighbour list (both directions, canonical pair
//     predicate) in CSR form over sorted indices: count pass, exclusive scan, fill
//     pass. No atomics, deterministic order, and the force kernel accumulates each
//     particle's force in registers from its own list;
//   * the int2 half list of Pairlist::get_neighbor_pairs is materialised from the
//     CSR only when asked for.
#include "pairlist.hpp"

#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>

namespace mmd {

namespace {

__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }

__global__ void k_encode(const float4* __restrict__ pos, uint64_t n, float3 bmin, float3 bmax,
						 bool canon_from_meta, uint32_t* __restrict__ keys,
						 uint32_t* __restrict__ vals, uint32_t* __restrict__ outside) {
	const uint64_t s = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (s >= n)
		return;
	const float4 p = pos[s];
	const uint32_t c = canon_from_meta ? meta_canon(fbits(p.w)) : uint32_t(s);
	if (c >= n)
		return;
	bool out = false;
	const uint32_t qx = morton_quantize(p.x, bmin.x, bmax.x, out);
	const uint32_t qy = morton_quantize(p.y, bmin.y, bmax.y, out);
	const uint32_t qz = morton_quantize(p.z, bmin.z, bmax.z, out);
	keys[c] = split_by3(qz) | (split_by3(qy) << 1) | (split_by3(qx) << 2);
	vals[c] = c;
	if (out)
		atomicOr(outside, 1u);
}

__global__ void k_inverse(const uint32_t* __restrict__ sidx, uint64_t n, uint32_t* __restrict__ inv) {
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	const uint32_t c = sidx[t];
	if (c < n)
		inv[c] = uint32_t(t);
}

__global__ void k_finalize(const uint32_t* __restrict__ sidx, uint64_t n, const float4* __restrict__ pos,
						   const uint32_t* __restrict__ slot_of_canon, bool canon_from_meta,
						   const int32_t* __restrict__ types, uint32_t* __restrict__ inv,
						   float4* __restrict__ spos) {
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	const uint32_t c = sidx[t];
	if (c >= n)
		return;
	inv[c] = uint32_t(t);
	const uint32_t s = slot_of_canon ? slot_of_canon[c] : c;
	if (s >= n)
		return;
	const float4 p = pos[s];
	const uint32_t meta = canon_from_meta ? fbits(p.w) : make_meta(types ? uint32_t(types[c]) : 0u, c);
	spos[t] = make_float4(p.x, p.y, p.z, __uint_as_float(meta));
}

template<class T>
__global__ void k_reorder(const T* __restrict__ in, T* __restrict__ out, const uint32_t* __restrict__ sidx,
						  uint64_t n) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	out[i] = in[sidx[i]];
}

__global__ void k_remap(int32_t* __restrict__ idx, uint64_t count, const uint32_t* __restrict__ inv,
						uint64_t map_size) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= count)
		return;
	const int32_t old = idx[i];
	if (old >= 0 && uint64_t(old) < map_size)
		idx[i] = int32_t(inv[old]);
}

// Order-preserving float <-> uint maps for atomic min/max.
__device__ __forceinline__ uint32_t ord(float f) {
	const uint32_t b = __float_as_uint(f);
	return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ inline float unord(uint32_t u) {
	const uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
	float f;
	std::memcpy(&f, &b, 4);
	return f;
}

// BoundingBoxKernel (ZOrderKernels.h:129-200) as a wave/block reduction.
__global__ __launch_bounds__(256) void k_bbox(const float4* __restrict__ pos, uint64_t n,
											  uint32_t* __restrict__ out /* 6 ordered words */) {
	float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
	for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
		const float4 p = pos[i];
		lo[0] = fminf(lo[0], p.x);
		lo[1] = fminf(lo[1], p.y);
		lo[2] = fminf(lo[2], p.z);
		hi[0] = fmaxf(hi[0], p.x);
		hi[1] = fmaxf(hi[1], p.y);
		hi[2] = fmaxf(hi[2], p.z);
	}
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		for (int o = 32; o > 0; o >>= 1) {
			lo[a] = fminf(lo[a], __shfl_xor(lo[a], o, 64));
			hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], o, 64));
		}
	}
	if ((threadIdx.x & 63) == 0) {
#pragma unroll
		for (int a = 0; a < 3; ++a) {
			atomicMin(&out[a], ord(lo[a]));
			atomicMax(&out[3 + a], ord(hi[a]));
		}
	}
}

__global__ void k_init_bbox(uint32_t* out) {
	if (threadIdx.x < 3)
		out[threadIdx.x] = 0xFFFFFFFFu;
	else if (threadIdx.x < 6)
		out[threadIdx.x] = 0u;
}

// BuildCellRangesKernel (ZOrderNeighbor.h:27-43).
__global__ void k_cell_ranges(const uint32_t* __restrict__ codes, uint64_t n, int shift,
							  uint32_t* __restrict__ begin, uint32_t* __restrict__ end) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const uint32_t c = codes[i] >> shift;
	if (i == 0 || (codes[i - 1] >> shift) != c)
		begin[c] = uint32_t(i);
	if (i + 1 == n || (codes[i + 1] >> shift) != c)
		end[c] = uint32_t(i + 1);
}

// BuildCellNeighborsKernel (ZOrderNeighbor.h:55-139), for the reference tables.
__global__ void k_cell_neighbors(uint32_t* __restrict__ out, uint64_t num_cells, int m, int3 r,
								 int slots, int3 per) {
	const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (c >= num_cells)
		return;
	const uint32_t cell = uint32_t(c);
	const uint32_t cx = compact_by3(cell >> 2), cy = compact_by3(cell >> 1), cz = compact_by3(cell);
	const int n = 1 << m;
	const uint32_t mask = uint32_t(n - 1);
	auto range = [&](int p, int rr, int& lo, int& hi) {
		if (p && 2 * rr + 1 >= n) {
			lo = 0;
			hi = n - 1;
		} else {
			lo = -rr;
			hi = rr;
		}
	};
	int xl, xh, yl, yh, zl, zh;
	range(per.x, r.x, xl, xh);
	range(per.y, r.y, yl, yh);
	range(per.z, r.z, zl, zh);
	uint32_t* row = out + c * uint64_t(slots);
	int k = 0;
	for (int dx = xl; dx <= xh; ++dx) {
		int nx = int(cx) + dx;
		if (per.x)
			nx = int(uint32_t(nx + n) & mask);
		else if (nx < 0 || nx >= n)
			continue;
		const uint32_t mx = split_by3(uint32_t(nx)) << 2;
		for (int dy = yl; dy <= yh; ++dy) {
			int ny = int(cy) + dy;
			if (per.y)
				ny = int(uint32_t(ny + n) & mask);
			else if (ny < 0 || ny >= n)
				continue;
			const uint32_t mxy = mx | (split_by3(uint32_t(ny)) << 1);
			for (int dz = zl; dz <= zh; ++dz) {
				int nz = int(cz) + dz;
				if (per.z)
					nz = int(uint32_t(nz + n) & mask);
				else if (nz < 0 || nz >= n)
					continue;
				row[k++] = mxy | split_by3(uint32_t(nz));
			}
		}
	}
	for (; k < slots; ++k)
		row[k] = kInvalidCell;
}

struct WalkParams {
	const float4* spos;
	const uint32_t* codes;
	const uint32_t* cell_begin;
	const uint32_t* cell_end;
	int m;
	int shift;
	int r[3];
	int full[3];     // axis enumerated whole (periodic and 2r+1 >= cells)
	int prune;
	float lo[3];     // domain minimum
	float cw[3];     // cell width along each axis (quantised)
	float margin[3]; // conservative slack of the cell intervals
	float rc2;
	DevBox box;
	marsmd_exclusion_view ex;
};

// Canonical pair predicate of ZOrderCellNeighborKernel for sorted lo < hi
// (ZOrderNeighbor.h:204-221): wrap_diff(pos_hi - pos_lo)^2 <= rc^2, not excluded by
// lo's CSR row or a shared rigid body.
__device__ __forceinline__ bool pair_in_list(const WalkParams& P, const float4& plo, const float4& phi) {
	const float3 dr = wrap_diff(P.box, sub3(phi, plo));
	if (!(len2(dr) <= P.rc2))
		return false;
	const int a = int(meta_canon(__float_as_uint(plo.w)));
	const int b = int(meta_canon(__float_as_uint(phi.w)));
	if (P.ex.rigid_body_id) {
		const int body = P.ex.rigid_body_id[a];
		if (body >= 0 && P.ex.rigid_body_id[b] == body)
			return false;
	}
	if (P.ex.offsets && a >= 0 && uint64_t(a) < P.ex.num_particles) {
		const int e0 = P.ex.offsets[a], e1 = P.ex.offsets[a + 1];
		for (int e = e0; e < e1; ++e)
			if (P.ex.excluded[e] == b)
				return false;
	}
	return true;
}

__device__ __forceinline__ float axis_gap(float prel, int cell, float cw, float margin) {
	const float lo = float(cell) * cw - margin;
	const float hi = float(cell + 1) * cw + margin;
	return fmaxf(0.0f, fmaxf(lo - prel, prel - hi));
}

template<bool FILL>
__global__ __launch_bounds__(256) void k_walk(WalkParams P, uint64_t n, uint32_t* __restrict__ counts,
											  const uint32_t* __restrict__ offsets,
											  uint32_t* __restrict__ nbr) {
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	const float4 p = P.spos[t];
	const uint32_t cell = P.codes[t] >> P.shift;
	const int c[3] = {int(compact_by3(cell >> 2)), int(compact_by3(cell >> 1)), int(compact_by3(cell))};
	const int cells = 1 << P.m;
	const float prel[3] = {p.x - P.lo[0], p.y - P.lo[1], p.z - P.lo[2]};
	uint32_t count = 0;
	uint64_t wpos = FILL ? offsets[t] : 0;
	int lo[3], hi[3];
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		if (P.full[a]) {
			lo[a] = 0;
			hi[a] = cells - 1;
		} else {
			lo[a] = c[a] - P.r[a];
			hi[a] = c[a] + P.r[a];
		}
	}
	for (int ux = lo[0]; ux <= hi[0]; ++ux) {
		int nx = ux;
		if (!P.full[0]) {
			if (P.box.periodic[0])
				nx = (ux + cells) & (cells - 1);
			else if (ux < 0 || ux >= cells)
				continue;
		}
		float gx = 0.0f;
		if (P.prune && !P.full[0]) {
			gx = axis_gap(prel[0], ux, P.cw[0], P.margin[0]);
			gx *= gx;
			if (gx > P.rc2)
				continue;
		}
		const uint32_t mx = split_by3(uint32_t(nx)) << 2;
		for (int uy = lo[1]; uy <= hi[1]; ++uy) {
			int ny = uy;
			if (!P.full[1]) {
				if (P.box.periodic[1])
					ny = (uy + cells) & (cells - 1);
				else if (uy < 0 || uy >= cells)
					continue;
			}
			float gxy = gx;
			if (P.prune && !P.full[1]) {
				const float g = axis_gap(prel[1], uy, P.cw[1], P.margin[1]);
				gxy += g * g;
				if (gxy > P.rc2)
					continue;
			}
			const uint32_t mxy = mx | (split_by3(uint32_t(ny)) << 1);
			for (int uz = lo[2]; uz <= hi[2]; ++uz) {
				int nz = uz;
				if (!P.full[2]) {
					if (P.box.periodic[2])
						nz = (uz + cells) & (cells - 1);
					else if (uz < 0 || uz >= cells)
						continue;
				}
				if (P.prune && !P.full[2]) {
					const float g = axis_gap(prel[2], uz, P.cw[2], P.margin[2]);
					if (gxy + g * g > P.rc2)
						continue;
				}
				const uint32_t nc = mxy | split_by3(uint32_t(nz));
				const uint32_t b = P.cell_begin[nc];
				const uint32_t e = P.cell_end[nc];
				for (uint32_t j = b; j < e; ++j) {
					if (j == uint32_t(t))
						continue;
					const float4 q = P.spos[j];
					const bool in = j > uint32_t(t) ? pair_in_list(P, p, q) : pair_in_list(P, q, p);
					if (in) {
						if (FILL)
							nbr[wpos++] = j;
						++count;
					}
				}
			}
		}
	}
	if (!FILL)
		counts[t] = count;
}

__global__ void k_half_counts(const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ nbr,
							  uint64_t n, uint32_t* __restrict__ half) {
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	uint32_t h = 0;
	for (uint32_t k = offsets[t]; k < offsets[t + 1]; ++k)
		h += nbr[k] > uint32_t(t) ? 1u : 0u;
	half[t] = h;
}

__global__ void k_write_pairs(const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ nbr,
							  const uint32_t* __restrict__ half_off, const float4* __restrict__ spos,
							  uint64_t n, int32_t* __restrict__ pairs) {
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	uint64_t w = half_off[t];
	const int32_t a = int32_t(meta_canon(__float_as_uint(spos[t].w)));
	for (uint32_t k = offsets[t]; k < offsets[t + 1]; ++k) {
		const uint32_t j = nbr[k];
		if (j > uint32_t(t)) {
			const int32_t b = int32_t(meta_canon(__float_as_uint(spos[j].w)));
			pairs[2 * w] = a < b ? a : b;
			pairs[2 * w + 1] = a < b ? b : a;
			++w;
		}
	}
}

__global__ __launch_bounds__(256) void k_max_disp(const float4* __restrict__ cur, const float4* __restrict__ snap,
												  const uint32_t* __restrict__ sorted_of_slot, uint64_t n,
												  float3 plen, uint32_t* __restrict__ out_bits) {
	float m = 0.0f;
	for (uint64_t s = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; s < n; s += uint64_t(gridDim.x) * blockDim.x) {
		const float4 a = cur[s];
		const float4 b = snap[sorted_of_slot ? sorted_of_slot[s] : s];
		float d[3] = {a.x - b.x, a.y - b.y, a.z - b.z};
		const float L[3] = {plen.x, plen.y, plen.z};
#pragma unroll
		for (int k = 0; k < 3; ++k) {
			// DisplacementKernel::min_image (AdaptiveKernels.h:58-66)
			if (L[k] > 0.0f) {
				int guard = 0;
				while (d[k] > 0.5f * L[k] && guard++ < 8)
					d[k] -= L[k];
				while (d[k] < -0.5f * L[k] && guard++ < 16)
					d[k] += L[k];
			}
		}
		const float d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
		m = fmaxf(m, d2);
	}
	for (int o = 32; o > 0; o >>= 1)
		m = fmaxf(m, __shfl_xor(m, o, 64));
	if ((threadIdx.x & 63) == 0)
		atomicMax(out_bits, __float_as_uint(m)); // non-negative floats order as uints
}

__global__ void k_max_count(const uint32_t* __restrict__ offsets, uint64_t n, uint32_t* __restrict__ out) {
	uint32_t m = 0;
	for (uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += uint64_t(gridDim.x) * blockDim.x)
		m = max(m, offsets[t + 1] - offsets[t]);
	for (int o = 32; o > 0; o >>= 1)
		m = max(m, uint32_t(__shfl_xor(int(m), o, 64)));
	if ((threadIdx.x & 63) == 0)
		atomicMax(out, m);
}

} // namespace

// ------------------------------------------------------------------ host side

CellGrid select_cell_grid(float cutoff, uint64_t n, const float extent_f[3], uint64_t max_pairs) {
	CellGrid g;
	const double extent[3] = {double(extent_f[0]), double(extent_f[1]), double(extent_f[2])};
	if (cutoff <= 0.0f || extent[0] <= 0.0 || extent[1] <= 0.0 || extent[2] <= 0.0 || n == 0)
		return g;
	bool have = false;
	double best = 0.0;
	for (int m = 0; m <= 10; ++m) {
		const long long cells = 1LL << m;
		const uint64_t num_cells = uint64_t(1) << (3 * m);
		int r[3];
		long long span[3];
		double volume = 1.0;
		for (int a = 0; a < 3; ++a) {
			const double side = extent[a] / double(cells);
			r[a] = int(std::ceil(double(cutoff) / side));
			span[a] = std::min(2LL * r[a] + 1LL, cells);
			volume *= double(span[a]) * side;
		}
		const long long slots = span[0] * span[1] * span[2];
		if (m > 0 && num_cells > n)
			break;
		const uint64_t bytes = num_cells * (uint64_t(slots) * 4 + 8);
		if (m > 0 && bytes > max_pairs * 8)
			break;
		const double cost = volume * double(slots);
		if (!have || cost < best) {
			have = true;
			best = cost;
			g.m = m;
			g.r[0] = r[0];
			g.r[1] = r[1];
			g.r[2] = r[2];
			g.slots = slots;
		}
	}
	return g;
}

ZSort::ZSort(uint64_t max_n) : max_particles(max_n) { outside_flag.resize(1); }

void ZSort::ensure(uint64_t nn) {
	if (nn > max_particles)
		raise(MARSMD_ERR_VALUE, "Cannot sort " + std::to_string(nn) + " particles, maximum is " +
									std::to_string(max_particles));
	keys_unsorted.resize(nn);
	vals_unsorted.resize(nn);
	codes.resize(nn);
	sorted_idx.resize(nn);
	inverse.resize(nn);
	scratch.resize(radix_sort_scratch_bytes(nn));
}

void ZSort::sort(const float4* pos, uint64_t nn, const float bmin[3], const float bmax[3],
				 const float* domain_dev, bool canon_from_meta, hipStream_t s) {
	(void)domain_dev;
	ensure(nn);
	n = nn;
	if (nn == 0)
		return;
	fill_u32(outside_flag.data(), 0u, 1, s);
	k_encode<<<grid_for(nn, 256), 256, 0, s>>>(pos, nn, make_float3(bmin[0], bmin[1], bmin[2]),
											   make_float3(bmax[0], bmax[1], bmax[2]), canon_from_meta,
											   keys_unsorted.data(), vals_unsorted.data(), outside_flag.data());
	MMD_LAUNCH_CHECK();
	radix_sort_pairs(keys_unsorted.data(), vals_unsorted.data(), codes.data(), sorted_idx.data(), nn, 30,
					 scratch.data(), scratch.count, s);
	k_inverse<<<grid_for(nn, 256), 256, 0, s>>>(sorted_idx.data(), nn, inverse.data());
	MMD_LAUNCH_CHECK();
}

void ZSort::reorder_data(const void* in, void* out, uint64_t nn, uint32_t elem, hipStream_t s) const {
	if (nn > max_particles)
		raise(MARSMD_ERR_VALUE, "Cannot reorder " + std::to_string(nn) + " elements");
	if (nn == 0)
		return;
	const unsigned g = grid_for(nn, 256);
	if (elem == 4)
		k_reorder<uint32_t><<<g, 256, 0, s>>>(static_cast<const uint32_t*>(in), static_cast<uint32_t*>(out), sorted_idx.data(), nn);
	else if (elem == 8)
		k_reorder<uint2><<<g, 256, 0, s>>>(static_cast<const uint2*>(in), static_cast<uint2*>(out), sorted_idx.data(), nn);
	else if (elem == 16)
		k_reorder<uint4><<<g, 256, 0, s>>>(static_cast<const uint4*>(in), static_cast<uint4*>(out), sorted_idx.data(), nn);
	else
		raise(MARSMD_ERR_VALUE, "reorder_data: element size must be 4, 8 or 16");
	MMD_LAUNCH_CHECK();
}

void ZSort::remap_indices(int32_t* idx, uint64_t count, hipStream_t s) const {
	if (count == 0)
		return;
	k_remap<<<grid_for(count, 256), 256, 0, s>>>(idx, count, inverse.data(), n);
	MMD_LAUNCH_CHECK();
}

PairlistImpl::PairlistImpl(uint64_t max_particles_, uint64_t max_pairs_ceiling_)
	: max_particles(max_particles_), max_pairs_ceiling(max_pairs_ceiling_), sorter(max_particles_) {
	ref_max_pairs = std::min<uint64_t>(max_pairs_ceiling, std::max<uint64_t>(max_particles, 1));
	dev_scalars.resize(8);
}

void PairlistImpl::ensure(uint64_t nn) {
	if (nn > max_particles)
		raise(MARSMD_ERR_VALUE, "Cannot build pairlist for " + std::to_string(nn) +
									" particles, maximum is " + std::to_string(max_particles));
	if (nn >= kMaxParticles)
		raise(MARSMD_ERR_VALUE, "pairlist: particle count exceeds 2^24");
	spos.resize(nn);
	counts.resize(nn + 1);
	offsets.resize(nn + 1);
	scan_scratch.resize(scan_scratch_bytes(nn + 1));
}

// The reference grid, whether it covers every pair within rc (the pair set is then
// grid independent), and the grid this build walks.
void PairlistImpl::choose_grids(uint64_t nn) {
	float extent[3];
	for (int a = 0; a < 3; ++a)
		extent[a] = domain_max[a] - domain_min[a];
	const float cut = std::sqrt(cutoff_squared);
	ref_grid = select_cell_grid(cut, nn, extent, ref_max_pairs);
	const double tol_abs = 1e-3;
	auto quant_w = [&](int a) { return double(std::fmax(extent[a], 1e-6f)) / 1023.0; };
	ref_complete = true;
	for (int a = 0; a < 3; ++a) {
		const int cells = 1 << ref_grid.m;
		const int k = 1024 / cells;
		const bool per = box.periodic[a] != 0 && box.size[a] > 0.0f;
		if (per && 2 * ref_grid.r[a] + 1 >= cells)
			continue;
		const double W = quant_w(a);
		const double reach = per ? (double(ref_grid.r[a]) * k - 1.0) * W : double(ref_grid.r[a]) * k * W;
		if (reach < double(cut) * (1.0 + 1e-6) + tol_abs)
			ref_complete = false;
	}
	// Walk grid: largest m with 8^m <= 2n (at most ~2 cells per particle), or the
	// override; radii cover rc with the seam slack of the quantised top cell.
	int mb = 0;
	while (mb < 10 && (uint64_t(1) << (3 * (mb + 1))) <= 2 * std::max<uint64_t>(nn, 1))
		++mb;
	if (build_m_override >= 0)
		mb = std::min(build_m_override, 10);
	build_grid.m = mb;
	const int cells = 1 << mb;
	const int k = 1024 / cells;
	long long slots = 1;
	for (int a = 0; a < 3; ++a) {
		const double W = quant_w(a);
		const bool per = box.periodic[a] != 0 && box.size[a] > 0.0f;
		const double need = (double(cut) * (1.0 + 1e-6) + tol_abs) / W + (per ? 1.0 : 0.0);
		int r = int(std::ceil(need / double(k)));
		if (r < 1)
			r = 1;
		build_grid.r[a] = r;
		slots *= std::min<long long>(2LL * r + 1, cells);
	}
	build_grid.slots = slots;
}

void PairlistImpl::walk(bool fill, bool use_ref, hipStream_t s) {
	const CellGrid& g = use_ref ? ref_grid : build_grid;
	WalkParams P{};
	P.spos = spos.data();
	P.codes = sorter.codes.data();
	P.cell_begin = cell_begin.data();
	P.cell_end = cell_end.data();
	P.m = g.m;
	P.shift = 3 * (10 - g.m);
	const int cells = 1 << g.m;
	const int k = 1024 / cells;
	P.prune = use_ref ? 0 : 1;
	for (int a = 0; a < 3; ++a) {
		const bool per = box.periodic[a] != 0 && box.size[a] > 0.0f;
		P.r[a] = g.r[a];
		P.full[a] = (per && 2 * g.r[a] + 1 >= cells) ? 1 : 0;
		const float extent = std::fmax(domain_max[a] - domain_min[a], 1e-6f);
		P.lo[a] = domain_min[a];
		P.cw[a] = float(double(extent) * k / 1023.0);
		P.margin[a] = float(double(extent) / 1023.0 * 1.001 + 1e-5 * double(extent) + 1e-3);
	}
	P.rc2 = cutoff_squared;
	P.box = make_devbox(box);
	P.ex = excl;
	const unsigned grid = grid_for(n, 256);
	if (fill)
		k_walk<true><<<grid, 256, 0, s>>>(P, n, counts.data(), offsets.data(), nbr.data());
	else
		k_walk<false><<<grid, 256, 0, s>>>(P, n, counts.data(), offsets.data(), nbr.data());
	MMD_LAUNCH_CHECK();
}

void PairlistImpl::build(const float4* pos, const uint32_t* slot_of_canon, bool canon_from_meta,
						 const int32_t* types, uint64_t nn, float pairlist_cutoff, hipStream_t s) {
	const auto t0 = std::chrono::high_resolution_clock::now();
	for (int a = 0; a < 3; ++a)
		if (!(box.size[a] > 0.0f))
			raise(MARSMD_ERR_VALUE, "ZOrderPairlist has no simulation box. Call set_periodic_box() before building.");
	ensure(nn);
	n = nn;
	cutoff = pairlist_cutoff;
	cutoff_squared = pairlist_cutoff * pairlist_cutoff;
	pairs_valid = false;
	built = true;
	if (nn == 0) {
		nbr_entries = 0;
		return;
	}
	// Morton domain (ZOrderPairlist.cpp:131-146, 368-411).
	const bool all_periodic = box.periodic[0] && box.periodic[1] && box.periodic[2];
	float lo[3], hi[3];
	if (!all_periodic) {
		dev_domain.resize(1);
		k_init_bbox<<<1, 64, 0, s>>>(dev_scalars.data());
		MMD_LAUNCH_CHECK();
		k_bbox<<<std::min<unsigned>(grid_for(nn, 256), 1024), 256, 0, s>>>(pos, nn, dev_scalars.data());
		MMD_LAUNCH_CHECK();
		uint32_t words[6];
		MMD_HIP(hipMemcpyAsync(words, dev_scalars.data(), sizeof(words), hipMemcpyDeviceToHost, s));
		MMD_HIP(hipStreamSynchronize(s));
		for (int a = 0; a < 3; ++a) {
			lo[a] = unord(words[a]);
			hi[a] = unord(words[3 + a]);
			const float margin = (hi[a] - lo[a]) * 0.01f;
			lo[a] -= margin;
			hi[a] += margin;
			const float range = hi[a] - lo[a];
			if (range < 1e-4f) {
				const float center = float(double(lo[a]) + double(range) * 0.5);
				lo[a] = float(double(center) - double(1e-4f) * 0.5);
				hi[a] = float(double(center) + double(1e-4f) * 0.5);
			}
		}
	}
	for (int a = 0; a < 3; ++a) {
		if (box.periodic[a]) {
			lo[a] = box.origin[a];
			hi[a] = box.origin[a] + box.size[a];
		}
		domain_min[a] = lo[a];
		domain_max[a] = hi[a];
	}
	sorter.sort(pos, nn, lo, hi, nullptr, canon_from_meta, s);
	k_finalize<<<grid_for(nn, 256), 256, 0, s>>>(sorter.sorted_idx.data(), nn, pos, slot_of_canon, canon_from_meta,
												 types, sorter.inverse.data(), spos.data());
	MMD_LAUNCH_CHECK();
	choose_grids(nn);
	bool use_ref = false;
	for (int attempt = 0; attempt < 2; ++attempt) {
		if (!ref_complete)
			use_ref = true;
		const CellGrid& g = use_ref ? ref_grid : build_grid;
		const uint64_t num_cells = uint64_t(1) << (3 * g.m);
		cell_begin.resize(num_cells);
		cell_end.resize(num_cells);
		fill_u32(cell_begin.data(), 0u, num_cells, s);
		fill_u32(cell_end.data(), 0u, num_cells, s);
		k_cell_ranges<<<grid_for(nn, 256), 256, 0, s>>>(sorter.codes.data(), nn, 3 * (10 - g.m), cell_begin.data(),
														 cell_end.data());
		MMD_LAUNCH_CHECK();
		walk(false, use_ref, s);
		exclusive_scan_u32(counts.data(), offsets.data(), nn, scan_scratch.data(), s);
		uint32_t host[2] = {0, 0};
		MMD_HIP(hipMemcpyAsync(&host[0], offsets.data() + nn, 4, hipMemcpyDeviceToHost, s));
		MMD_HIP(hipMemcpyAsync(&host[1], sorter.outside_flag.data(), 4, hipMemcpyDeviceToHost, s));
		MMD_HIP(hipStreamSynchronize(s));
		if (host[1] != 0 && !use_ref) {
			// A coordinate outside the Morton domain was clamped: cell geometry no
			// longer bounds positions, so walk the reference grid exactly.
			use_ref = true;
			continue;
		}
		nbr_entries = host[0];
		break;
	}
	const uint64_t np = nbr_entries / 2;
	// Pairlist::grow_pair_capacity (Pairlist.h:244-251) bookkeeping.
	if (np > ref_max_pairs) {
		if (np > max_pairs_ceiling)
			raise(MARSMD_ERR_CAPACITY, "Pairlist capacity exceeded: " + std::to_string(np) + " pairs for " +
										   std::to_string(nn) + " particles but the cap holds only " +
										   std::to_string(max_pairs_ceiling));
		ref_max_pairs = std::min<uint64_t>(max_pairs_ceiling, np + np / 4);
	}
	if (nbr_entries > nbr.count)
		nbr.resize(nbr_entries + nbr_entries / 8 + 1024);
	walk(true, use_ref, s);
	// statistics: max neighbours
	fill_u32(dev_scalars.data() + 2, 0u, 1, s);
	k_max_count<<<std::min<unsigned>(grid_for(nn, 256), 512), 256, 0, s>>>(offsets.data(), nn, dev_scalars.data() + 2);
	MMD_LAUNCH_CHECK();
	const auto t1 = std::chrono::high_resolution_clock::now();
	last_build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
}

float PairlistImpl::max_displacement(const float4* pos, const uint32_t* sorted_of_slot, uint64_t nn,
									 hipStream_t s) {
	if (nn == 0)
		return 0.0f;
	fill_u32(dev_scalars.data() + 1, 0u, 1, s);
	const float3 plen = make_float3(box.periodic[0] ? box.size[0] : 0.0f, box.periodic[1] ? box.size[1] : 0.0f,
									box.periodic[2] ? box.size[2] : 0.0f);
	k_max_disp<<<std::min<unsigned>(grid_for(nn, 256), 1024), 256, 0, s>>>(pos, spos.data(), sorted_of_slot, nn, plen,
																			 dev_scalars.data() + 1);
	MMD_LAUNCH_CHECK();
	uint32_t bits = 0;
	MMD_HIP(hipMemcpyAsync(&bits, dev_scalars.data() + 1, 4, hipMemcpyDeviceToHost, s));
	MMD_HIP(hipStreamSynchronize(s));
	float d2;
	std::memcpy(&d2, &bits, 4);
	return std::sqrt(d2);
}

const int32_t* PairlistImpl::neighbor_pairs(hipStream_t s) {
	if (pairs_valid)
		return pairs.data();
	const uint64_t np = num_pairs();
	pairs.resize(2 * np + 2);
	if (n > 0 && np > 0) {
		half_counts.resize(n + 1);
		half_offsets.resize(n + 1);
		k_half_counts<<<grid_for(n, 256), 256, 0, s>>>(offsets.data(), nbr.data(), n, half_counts.data());
		MMD_LAUNCH_CHECK();
		exclusive_scan_u32(half_counts.data(), half_offsets.data(), n, scan_scratch.data(), s);
		k_write_pairs<<<grid_for(n, 256), 256, 0, s>>>(offsets.data(), nbr.data(), half_offsets.data(), spos.data(),
														n, pairs.data());
		MMD_LAUNCH_CHECK();
	}
	pairs_valid = true;
	return pairs.data();
}

void PairlistImpl::reference_cells(hipStream_t s, uint64_t* num_cells_out) {
	const CellGrid& g = ref_grid;
	const uint64_t num_cells = uint64_t(1) << (3 * g.m);
	ref_begin.resize(num_cells);
	ref_end.resize(num_cells);
	ref_neighbors.resize(num_cells * uint64_t(g.slots));
	fill_u32(ref_begin.data(), 0u, num_cells, s);
	fill_u32(ref_end.data(), 0u, num_cells, s);
	if (n > 0) {
		k_cell_ranges<<<grid_for(n, 256), 256, 0, s>>>(sorter.codes.data(), n, 3 * (10 - g.m), ref_begin.data(),
														ref_end.data());
		MMD_LAUNCH_CHECK();
	}
	const int3 per = make_int3(box.periodic[0] && box.size[0] > 0 ? 1 : 0, box.periodic[1] && box.size[1] > 0 ? 1 : 0,
							   box.periodic[2] && box.size[2] > 0 ? 1 : 0);
	k_cell_neighbors<<<grid_for(num_cells, 256), 256, 0, s>>>(ref_neighbors.data(), num_cells, g.m,
															   make_int3(g.r[0], g.r[1], g.r[2]), int(g.slots), per);
	MMD_LAUNCH_CHECK();
	*num_cells_out = num_cells;
}

marsmd_pairlist_stats PairlistImpl::stats() const {
	marsmd_pairlist_stats st{};
	st.num_particles = n;
	st.num_pairs = num_pairs();
	st.average_neighbors_per_particle = n > 0 ? double(float(st.num_pairs) / float(n)) : 0.0;
	uint32_t mx = 0;
	if (n > 0)
		(void)hipMemcpy(&mx, dev_scalars.data() + 2, 4, hipMemcpyDeviceToHost);
	st.max_neighbors_per_particle = mx;
	st.build_time_ms = last_build_ms;
	st.cell_grid_bits = ref_grid.m;
	for (int a = 0; a < 3; ++a) {
		st.cell_radii[a] = ref_grid.r[a];
		st.build_radii[a] = build_grid.r[a];
	}
	st.neighbors_per_cell = int32_t(ref_grid.slots);
	st.build_grid_bits = build_grid.m;
	st.reference_grid_complete = ref_complete ? 1 : 0;
	st.pair_capacity = ref_max_pairs;
	return st;
}

} // namespace mmd
