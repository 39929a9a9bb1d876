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

#include "homewalk.hpp"

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

// The stable sort of nondecreasing keys: codes = keys, sorted and inverse maps the
// identity; a key below its predecessor sets bit 1 of `flag`.
__global__ void k_presorted(const uint32_t* __restrict__ keys, uint64_t n, uint32_t* __restrict__ codes,
							uint32_t* __restrict__ sidx, uint32_t* __restrict__ inv, uint32_t* __restrict__ flag) {
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	const uint32_t k = keys[t];
	codes[t] = k;
	sidx[t] = uint32_t(t);
	inv[t] = uint32_t(t);
	if (t > 0 && keys[t - 1] > k)
		atomicOr(flag, 2u);
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

__device__ __forceinline__ bool pair_in_list(const WalkParams& P, const float4& plo, const float4& phi) {
	return pair_predicate(P.box, P.rc2, P.ex, plo, phi);
}

__device__ __forceinline__ float axis_gap(float prel, int cell, float cw, float margin) {
	const float lo = float(cell) * cw - margin;
	const float hi = float(cell + 1) * cw + margin;
	return fmaxf(0.0f, fmaxf(lo - prel, prel - hi));
}

template<bool FILL>
__global__ __launch_bounds__(256) void k_walk(WalkParams P, uint64_t n, uint32_t* __restrict__ counts,
											  const uint32_t* __restrict__ tile_base,
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
	const uint64_t wbase = FILL ? tile_base[t >> 6] : 0;
	int lo[3], hi[3];
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		if (P.full[a]) {
			lo[a] = 0;
			hi[a] = cells - 1;
		} else {
			lo[a] = c[a] - P.r[a];
			hi[a] = c[a] + P.r[a];
			if (!P.box.periodic[a]) {
				lo[a] = max(lo[a], 0);
				hi[a] = min(hi[a], cells - 1);
			}
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
					const bool t_lo = j > uint32_t(t);
					const bool in = pair_in_list(P, t_lo ? p : q, t_lo ? q : p);
					if (in) {
						if (FILL)
							nbr[nbr_slot_index(wbase, uint32_t(t), count)] = j;
						++count;
					}
				}
			}
		}
	}
	if (FILL) {
		for (uint32_t k = count; (k & 3u) != 0; ++k)
			nbr[nbr_slot_index(wbase, uint32_t(t), k)] = uint32_t(t);
	} else {
		counts[t] = count;
	}
}

// Row length of each wave tile: 64 x the tile's largest count rounded up to 4.
__global__ __launch_bounds__(256) void k_tile_rows(const uint32_t* __restrict__ counts, uint64_t n,
												   uint32_t* __restrict__ rows) {
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	uint32_t m = t < n ? counts[t] : 0u;
	for (int o = 32; o > 0; o >>= 1)
		m = max(m, uint32_t(__shfl_xor(int(m), o, 64)));
	const uint64_t w = t >> 6;
	if ((threadIdx.x & 63) == 0 && w * 64 < n)
		rows[w] = 64u * ((m + 3u) & ~3u);
}

// Largest count and total entries: grid-stride partials reduced in the workgroup, then
// one atomic pair per workgroup (launched with a few hundred workgroups).
// The sorter's flag word is copied to max_count[1], so that one read-back of four words
// returns the three.
__global__ __launch_bounds__(256) void k_count_stats(const uint32_t* __restrict__ counts, uint64_t n,
													 unsigned long long* __restrict__ total,
													 uint32_t* __restrict__ max_count,
													 const uint32_t* __restrict__ flag) {
	if (blockIdx.x == 0 && threadIdx.x == 0)
		max_count[1] = flag[0];
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

// tile_base[w] = w * row for w in [0, tiles] (rows of `row` entries per tile).
__global__ void k_uniform_tile_base(uint32_t* __restrict__ tile_base, uint64_t tiles, uint32_t row) {
	const uint64_t w = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (w <= tiles)
		tile_base[w] = uint32_t(w * row);
}

__global__ void k_half_counts(NbrList L, uint64_t n, uint32_t* __restrict__ half) {
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	const uint64_t base = L.tile_base[t >> 6];
	uint32_t h = 0;
	for (uint32_t k = 0; k < L.count[t]; ++k)
		h += L.nbr[nbr_slot_index(base, uint32_t(t), k)] > uint32_t(t) ? 1u : 0u;
	half[t] = h;
}

__global__ void k_write_pairs(NbrList L, const uint32_t* __restrict__ half_off, const float4* __restrict__ spos,
							  uint64_t n, int32_t* __restrict__ pairs) {
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	uint64_t w = half_off[t];
	const uint64_t base = L.tile_base[t >> 6];
	const int32_t a = int32_t(meta_canon(__float_as_uint(spos[t].w)));
	for (uint32_t k = 0; k < L.count[t]; ++k) {
		const uint32_t j = L.nbr[nbr_slot_index(base, uint32_t(t), k)];
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
				 const float* domain_dev, bool canon_from_meta, hipStream_t s, bool presorted) {
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
	if (presorted) {
		k_presorted<<<grid_for(nn, 256), 256, 0, s>>>(keys_unsorted.data(), nn, codes.data(), sorted_idx.data(),
													  inverse.data(), outside_flag.data());
		MMD_LAUNCH_CHECK();
		return;
	}
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
	const uint64_t tiles = (nn + 63) / 64;
	tile_rows.resize(tiles + 1);
	tile_base.resize(tiles + 1);
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
		int r = int(std::min(std::ceil(need / double(k)), double(cells)));
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
		k_walk<true><<<grid, 256, 0, s>>>(P, n, counts.data(), tile_base.data(), nbr.data());
	else
		k_walk<false><<<grid, 256, 0, s>>>(P, n, counts.data(), tile_base.data(), nbr.data());
	MMD_LAUNCH_CHECK();
}

// Largest count, the sorter's flag word and total entries of `counts` into dev_scalars
// [2], [3] and [4..5].
void PairlistImpl::count_stats(uint64_t nn, hipStream_t s) {
	fill_u32(dev_scalars.data() + 2, 0u, 4, s);
	k_count_stats<<<std::min<unsigned>(grid_for(nn, 256), 240), 256, 0, s>>>(
		counts.data(), nn, reinterpret_cast<unsigned long long*>(dev_scalars.data() + 4), dev_scalars.data() + 2,
		sorter.outside_flag.data());
	MMD_LAUNCH_CHECK();
}

// The home-cell walk on the build grid (k_walk_home), writing uniform tile rows of
// home_cap entries per slot and walking again once with a larger cap when a count
// exceeds it. Returns false when it does not apply (no level below the build grid, or
// rows beyond 32-bit offsets); `outside` reports a clamped coordinate (then the
// caller walks the reference grid).
bool PairlistImpl::walk_home(hipStream_t s, bool& outside) {
	const int m = build_grid.m;
	if (m < 1)
		return false;
	// home level: at least home_target particles per home cell on average (by default
	// about 24..190), below the build grid; the candidate cells are fine_offset levels
	// finer than the home cells
	int h = 0;
	while (h + 1 <= m - 1 && (n >> (3 * (h + 1))) >= uint64_t(std::max(1, home_target)))
		++h;
	const int mf = std::min(10, h + std::max(1, fine_offset));
	const uint64_t tiles = (n + 63) / 64;
	HomeParams P{};
	P.spos = spos.data();
	P.h = h;
	P.s = mf - h;
	P.num_home = uint32_t(1) << (3 * h);
	const int fine_cells = 1 << mf;
	const int span = 1 << (mf - h);
	const int k = 1024 / fine_cells;
	const double cut = std::sqrt(double(cutoff_squared));
	double volume = 1.0;
	for (int a = 0; a < 3; ++a) {
		const bool per = box.periodic[a] != 0 && box.size[a] > 0.0f;
		const float extent = std::fmax(domain_max[a] - domain_min[a], 1e-6f);
		const double W = double(extent) / 1023.0;
		// fine cells covering rc with the seam slack of the quantised top cell
		const double need = (cut * (1.0 + 1e-6) + 1e-3) / W + (per ? 1.0 : 0.0);
		P.reach[a] = std::max(1, int(std::min(std::ceil(need / double(k)), double(fine_cells))));
		P.full[a] = (per && span + 2 * P.reach[a] >= fine_cells) ? 1 : 0;
		volume *= double(extent);
		P.cw[a] = float(double(extent) * k / 1023.0);
		P.margin[a] = float(W * 1.001 + 1e-5 * double(extent) + 1e-3);
	}
	const uint64_t num_fine = uint64_t(1) << (3 * mf);
	cell_begin.resize(num_fine);
	cell_end.resize(num_fine);
	fill_u32(cell_begin.data(), 0u, num_fine, s);
	fill_u32(cell_end.data(), 0u, num_fine, s);
	k_cell_ranges<<<grid_for(n, 256), 256, 0, s>>>(sorter.codes.data(), n, 3 * (10 - mf), cell_begin.data(),
													cell_end.data());
	MMD_LAUNCH_CHECK();
	P.rc2 = cutoff_squared;
	P.box = make_devbox(box);
	P.ex = excl;
	P.pre = home_prefilter_band(box, domain_min, domain_max, h, cutoff, &P.pre_lo2, &P.pre_hi2) ? 1 : 0;
	// the stencil depends on the geometry only: made and uploaded when that changes
	const std::vector<float> key = {float(h),	   float(P.s),		   float(P.reach[0]), float(P.reach[1]),
									float(P.reach[2]), float(P.full[0]),   float(P.full[1]),  float(P.full[2]),
									P.cw[0],		   P.cw[1],			   P.cw[2],			  P.margin[0],
									P.margin[1],	   P.margin[2],		   P.rc2};
	if (key != stencil_key) {
		stencil_key = key;
		stencil_host = home_stencil(h, P.s, P.reach, P.full, P.cw, P.margin, P.rc2, 8192);
		if (!stencil_host.empty()) {
			stencil_dev.resize(stencil_host.size());
			MMD_HIP(hipMemcpy(stencil_dev.data(), stencil_host.data(), sizeof(int4) * stencil_host.size(),
							  hipMemcpyHostToDevice));
		}
	}
	if (!stencil_host.empty()) {
		P.stencil = stencil_dev.data();
		P.stencil_n = uint32_t(stencil_host.size());
	}
	if (!ref_complete) {
		// keep exactly the pairs the reference grid's stencil reaches
		const int cells = 1 << ref_grid.m;
		P.ref_filter = 1;
		P.ref_shift = 10 - ref_grid.m;
		P.ref_cells = cells;
		for (int a = 0; a < 3; ++a) {
			const bool per = box.periodic[a] != 0 && box.size[a] > 0.0f;
			P.ref_r[a] = ref_grid.r[a];
			P.ref_full[a] = (per && 2 * ref_grid.r[a] + 1 >= cells) ? 1 : 0;
			P.dom_lo[a] = domain_min[a];
			P.dom_hi[a] = domain_max[a];
		}
	}
	if (home_cap == 0) {
		// first build: the mean count of a uniform system, with room for its fluctuations
		const double mean = double(n) / volume * 4.0 / 3.0 * 3.141592653589793 * double(cutoff) * cutoff * cutoff;
		home_cap = uint32_t(std::min(1.6 * mean + 16.0, 4096.0) + 3) & ~3u;
	}
	const uint64_t home_cells = uint64_t(P.num_home);
	home_begin.resize(home_cells);
	home_end.resize(home_cells);
	fill_u32(home_begin.data(), 0u, home_cells, s);
	fill_u32(home_end.data(), 0u, home_cells, s);
	k_cell_ranges<<<grid_for(n, 256), 256, 0, s>>>(sorter.codes.data(), n, 3 * (10 - h), home_begin.data(),
													home_end.data());
	MMD_LAUNCH_CHECK();
	P.home_begin = home_begin.data();
	P.home_end = home_end.data();
	P.fine_begin = cell_begin.data();
	P.fine_end = cell_end.data();
	const unsigned groups = unsigned((P.num_home + kHomeWaves - 1) / kHomeWaves);
	const unsigned grid = std::max(8u, (groups + 7u) / 8u * 8u);
	for (int attempt = 0; attempt < 2; ++attempt) {
		const uint64_t storage = tiles * 64 * uint64_t(home_cap);
		if (storage >= (uint64_t(1) << 32))
			return false;
		if (storage > nbr.count)
			nbr.resize(storage + storage / 8 + 1024);
		P.cap = home_cap;
		k_uniform_tile_base<<<grid_for(tiles + 1, 256), 256, 0, s>>>(tile_base.data(), tiles, 64u * home_cap);
		MMD_LAUNCH_CHECK();
		k_walk_home<true><<<grid, 256, 0, s>>>(P, counts.data(), nbr.data());
		MMD_LAUNCH_CHECK();
		count_stats(n, s);
		uint32_t* host = readback.ptr; // max count, flag word, total entries (2 words)
		MMD_HIP(hipMemcpyAsync(host, dev_scalars.data() + 2, 16, hipMemcpyDeviceToHost, s));
		MMD_HIP(hipStreamSynchronize(s));
		if (host[1] & 2u)
			raise(MARSMD_ERR_RUNTIME, "pairlist: keys after a reorder were not in order");
		if (host[1] != 0) {
			outside = true;
			return true;
		}
		max_neighbors = host[0];
		std::memcpy(&nbr_entries, &host[2], 8);
		if (max_neighbors <= home_cap) {
			nbr_storage = storage;
			home_level = h;
			return true;
		}
		home_cap = (max_neighbors + max_neighbors / 8 + 4 + 3) & ~3u;
	}
	return false;
}

void PairlistImpl::build(const float4* pos, const uint32_t* slot_of_canon, bool canon_from_meta,
						 const int32_t* types, uint64_t nn, float pairlist_cutoff, hipStream_t s) {
	const auto t0 = std::chrono::high_resolution_clock::now();
	++list_version;
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
		nbr_storage = 0;
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
	// after a reorder over the periodic box the canonical labels are the ranks of these
	// very keys, so the sort is the identity
	const bool presorted = presorted_next && all_periodic && canon_from_meta;
	presorted_next = false;
	sorter.sort(pos, nn, lo, hi, nullptr, canon_from_meta, s, presorted);
	k_finalize<<<grid_for(nn, 256), 256, 0, s>>>(sorter.sorted_idx.data(), nn, pos, slot_of_canon, canon_from_meta,
												 types, sorter.inverse.data(), spos.data());
	MMD_LAUNCH_CHECK();
	choose_grids(nn);
	bool use_ref = false;
	home_level = -1;
	bool home_done = false;
	{
		bool outside = false;
		home_done = walk_home(s, outside);
		if (outside) {
			// A coordinate outside the Morton domain was clamped: walk the reference grid.
			home_done = false;
			home_level = -1;
			use_ref = true;
		}
	}
	for (int attempt = 0; attempt < 2 && !home_done; ++attempt) {
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
		const uint64_t tiles = (nn + 63) / 64;
		k_tile_rows<<<grid_for(nn, 256), 256, 0, s>>>(counts.data(), nn, tile_rows.data());
		MMD_LAUNCH_CHECK();
		count_stats(nn, s);
		exclusive_scan_u32(tile_rows.data(), tile_base.data(), tiles, scan_scratch.data(), s);
		uint32_t* host = readback.ptr; // storage, then max count, flag word, total entries
		MMD_HIP(hipMemcpyAsync(&host[0], tile_base.data() + tiles, 4, hipMemcpyDeviceToHost, s));
		MMD_HIP(hipMemcpyAsync(&host[2], dev_scalars.data() + 2, 16, hipMemcpyDeviceToHost, s));
		MMD_HIP(hipStreamSynchronize(s));
		host[1] = host[3]; // the flag word, where the checks below read it
		if (host[1] & 2u)
			raise(MARSMD_ERR_RUNTIME, "pairlist: keys after a reorder were not in order");
		if (host[1] != 0 && !use_ref) {
			// A coordinate outside the Morton domain was clamped: cell geometry no
			// longer bounds positions, so walk the reference grid exactly.
			use_ref = true;
			continue;
		}
		uint64_t total = 0;
		std::memcpy(&total, &host[4], 8);
		max_neighbors = host[2];
		// Each row is padded to a multiple of 4 within a tile of 64: the padded
		// storage fits 32-bit offsets when this bound does.
		if (uint64_t(max_neighbors + 3) * 64 * tiles >= (uint64_t(1) << 32))
			raise(MARSMD_ERR_CAPACITY, "Pairlist: neighbour storage exceeds 2^32 entries");
		nbr_entries = total;
		nbr_storage = host[0];
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
	if (!home_done) {
		if (nbr_storage > nbr.count)
			nbr.resize(nbr_storage + nbr_storage / 8 + 1024);
		walk(true, use_ref, s);
	}
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

void ZSort::update_positions_incremental(const float4* pos, uint64_t nn, hipStream_t s) {
	if (nn > max_particles)
		raise(MARSMD_ERR_VALUE, "ZOrderSort: " + std::to_string(nn) + " positions exceed the capacity of " +
									std::to_string(max_particles));
	disp_ref.resize(std::max<uint64_t>(nn, 1));
	if (nn)
		MMD_HIP(hipMemcpyAsync(disp_ref.data(), pos, 16 * nn, hipMemcpyDeviceToDevice, s));
	disp_n = nn;
	disp_valid = true;
}

float ZSort::compute_max_displacement(const float4* pos, uint64_t nn, const float box_len[3], hipStream_t s) {
	if (!disp_valid)
		raise(MARSMD_ERR_RUNTIME, "ZOrderSort: compute_max_displacement before update_positions_incremental");
	if (nn != disp_n)
		raise(MARSMD_ERR_VALUE, "ZOrderSort: " + std::to_string(nn) + " positions against a reference of " +
									std::to_string(disp_n));
	if (nn == 0)
		return 0.0f;
	disp_bits.resize(1);
	fill_u32(disp_bits.data(), 0u, 1, s);
	const float3 plen = make_float3(box_len && box_len[0] > 0.0f ? box_len[0] : 0.0f,
									box_len && box_len[1] > 0.0f ? box_len[1] : 0.0f,
									box_len && box_len[2] > 0.0f ? box_len[2] : 0.0f);
	k_max_disp<<<std::min<unsigned>(grid_for(nn, 256), 1024), 256, 0, s>>>(pos, disp_ref.data(), nullptr, nn, plen,
																			 disp_bits.data());
	MMD_LAUNCH_CHECK();
	uint32_t bits = 0;
	MMD_HIP(hipMemcpyAsync(&bits, disp_bits.data(), 4, hipMemcpyDeviceToHost, s));
	MMD_HIP(hipStreamSynchronize(s));
	float d2;
	std::memcpy(&d2, &bits, 4);
	return std::sqrt(d2);
}

double PairlistImpl::rewalk_ms(hipStream_t s) {
	if (!built || home_level < 0)
		return -1.0;
	hipEvent_t e[2];
	MMD_HIP(hipEventCreate(&e[0]));
	MMD_HIP(hipEventCreate(&e[1]));
	MMD_HIP(hipEventRecord(e[0], s));
	bool outside = false;
	const bool ok = walk_home(s, outside);
	MMD_HIP(hipEventRecord(e[1], s));
	MMD_HIP(hipEventSynchronize(e[1]));
	float ms = 0.0f;
	MMD_HIP(hipEventElapsedTime(&ms, e[0], e[1]));
	(void)hipEventDestroy(e[0]);
	(void)hipEventDestroy(e[1]);
	if (!ok || outside)
		raise(MARSMD_ERR_RUNTIME, "pairlist: the home walk no longer applies to the last build");
	pairs_valid = false;
	++list_version;
	return double(ms);
}

const int32_t* PairlistImpl::neighbor_pairs(hipStream_t s) {
	if (pairs_valid)
		return pairs.data();
	const uint64_t np = num_pairs();
	pairs.resize(2 * np + 2);
	if (n > 0 && np > 0) {
		half_counts.resize(n + 1);
		half_offsets.resize(n + 1);
		k_half_counts<<<grid_for(n, 256), 256, 0, s>>>(list(), n, half_counts.data());
		MMD_LAUNCH_CHECK();
		exclusive_scan_u32(half_counts.data(), half_offsets.data(), n, scan_scratch.data(), s);
		k_write_pairs<<<grid_for(n, 256), 256, 0, s>>>(list(), half_offsets.data(), spos.data(), n, pairs.data());
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
	st.max_neighbors_per_particle = n > 0 ? max_neighbors : 0;
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
