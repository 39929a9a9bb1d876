// fm.hip — force matching of tabulated pair potentials on gfx950: the forces of a fixed
// set of configurations (snapshots) for given table knots, and the adjoint dL/dknot of a
// loss of those forces, accumulated in FP64.
//
// The reference has no gradient path; this is the library's own. The forward force is
// the Patch's pair term (nb_pair: PairNonbondedComputer, Pairwise.h:184-305, with
// TabulatedPotential::compute, TabulatedPotential.h:19-52) over each snapshot's own
// pairlist (ZOrderPairlist semantics), i.e. calculate_nonbonded_forces on that
// configuration. A table force is linear in its knots,
//   fm = -(U[h+1] - U[h]) step_inv      (h = the bin of the pair's distance),
// so with G = dL/dF each pair (x = lower canonical index, y = higher, u = r / d with
// r = wrap(pos_y - pos_x), the forward force's orientation) contributes
//   dL/dfm = (G_y - G_x) . u,   dL/dU[h] += dL/dfm step_inv,   dL/dU[h+1] -= dL/dfm step_inv;
// a distance past the last knot of a (non-periodic) table carries no force and no
// gradient. The expression is symmetric in the two ends (swapping them negates both
// factors), so each pair is visited once, from the end fm_owns picks.
//
// The snapshots are fixed, so everything the adjoint needs of a pair but the gradients
// is fixed too: at creation each snapshot's pairs are reduced, once, to records of the
// pairs that carry a gradient (tabulated, 1e-6 <= d <= cutoff, inside the table) at
// their owning end t, holding the other end's slot o, the knot index of the pair's
// bin and {v, step_inv}, v = u taken from t towards o (-u when t is the higher end:
// the same r, negated exactly), so that dL/dfm = (G_o - G_t) . v. The adjoint then
// streams records instead of the full neighbour rows and positions (half the entries,
// no geometry), with dL/dF first gathered into slot order, so that G_t is read in
// order and G_o from the slots around t (Morton order) instead of from anywhere in the
// snapshot.
#include "fm.hpp"

#include "nbforce.hpp"

#include <algorithm>
#include <cstring>

namespace mmd {

namespace {

constexpr int kFmBlock = 256;
// The adjoint's workgroups: 512 lanes share one workgroup's FP64 knot sums in LDS, so
// the LDS budget admits twice the waves of 256-lane workgroups.
constexpr int kFmAdjBlock = 512;
constexpr int kFmLdsGrad = 7936; // knots whose gradient a workgroup accumulates in LDS (62 KB)

}

// One snapshot's device arrays (its sorted positions, sorted order and pairlist).
struct FmSnapView {
	const float4* spos;
	const uint32_t* sorted_idx;
	const uint32_t* counts;
	const uint32_t* tile_base;
	const uint32_t* nbr;
	// adjoint records: record k of slot t at nbr_slot_index(tile_base[t / 64], t, k)
	const uint32_t* rec_count;
	const uint32_t* rec_y; // slot of the pair's other end
	// {u, code}: code (int bits) = table << kFmKnotBits | knot index of the pair's bin
	// (table offset + bin); the adjoint reads the table's step_inv from LDS
	const float4* rec_u;
};
constexpr int kFmKnotBits = 26;
// step_inv of every table, for the adjoint (a kernel argument staged in LDS)
struct FmSteps {
	float step_inv[kMaxTables];
};

namespace {

// The forces of every snapshot in one persistent launch: work item w is tile w % tiles
// of snapshot w / tiles (snapshot-major, XCD-local chunks), the tables staged once per
// workgroup; written in each snapshot's particle order, or its sorted order.
__global__ __launch_bounds__(kFmBlock) void k_fm_forces(const FmSnapView* __restrict__ snaps, uint64_t n,
														uint32_t tiles, uint32_t items, NbTables T,
														const float* __restrict__ knots_g, DevBox box, float cut_d,
														int sorted, float4* __restrict__ out) {
	extern __shared__ float lds_knots[];
	__shared__ NbLds lds_nb;
	stage_tables(T, knots_g, lds_knots, lds_nb);
	const TileRange R = xcd_tiles(items);
	for (uint32_t w = R.tile; w < R.end; w += R.stride) {
		const uint32_t k = w / tiles;
		const uint64_t t = uint64_t(w - k * tiles) * kFmBlock + threadIdx.x;
		if (t >= n)
			continue;
		const FmSnapView V = snaps[k];
		const NbrList L{V.counts, V.tile_base, V.nbr};
		const float4 f = nb_force(uint32_t(t), V.spos[t], V.spos, L, T, lds_knots, knots_g, lds_nb, box, cut_d, true);
		out[uint64_t(k) * n + (sorted ? t : V.sorted_idx[t])] = f;
	}
}

// The end of pair (a, b) that holds its record: the lower index when a and b have the
// same parity, the higher one otherwise. Exactly one of fm_owns(a, b), fm_owns(b, a)
// holds, and a slot owns about half of its pairs whatever its index (with the lower
// end always owning, slot counts would run from the whole row down to none, and a
// wavefront's loop would take its longest lane's count).
__device__ __forceinline__ bool fm_owns(uint32_t a, uint32_t b) { return ((a ^ b) & 1u) ? a > b : a < b; }

// The adjoint's records of one snapshot (see the top of this file): slot t's row is
// scanned in order and every pair that t owns and that carries a gradient is written as
// the next record of t.
__global__ __launch_bounds__(kFmBlock) void k_fm_records(const float4* __restrict__ spos, uint64_t n,
														 NbrList L, NbTables T, DevBox box, float cutoff2,
														 uint32_t* __restrict__ rec_count, uint32_t* __restrict__ rec_y,
														 float4* __restrict__ rec_u) {
	__shared__ NbLds lds_nb;
	stage_params(T, lds_nb);
	__syncthreads();
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	const float4 p = spos[t];
	const uint32_t mt = __float_as_uint(p.w);
	const uint32_t ct = meta_canon(mt);
	const uint32_t tt = meta_type(mt);
	const uint64_t base = L.tile_base[t >> 6];
	const uint32_t cnt = L.count[t];
	uint32_t m = 0;
	for (uint32_t k = 0; k < cnt; ++k) {
		const float4 q = spos[L.nbr[nbr_slot_index(base, uint32_t(t), k)]];
		const uint32_t mj = __float_as_uint(q.w);
		const uint32_t cj = meta_canon(mj);
		if (!fm_owns(ct, cj))
			continue;
		// the forward's orientation: type pair (lower, higher) and r = pos[higher] -
		// pos[lower] (q - p or its exact negation)
		const bool lower = ct < cj;
		const uint32_t tj = meta_type(mj);
		const uint32_t tag = lds_nb.tag[lower ? (tt << 4) | tj : (tj << 4) | tt];
		if (!(tag & T.enabled & MARSMD_PAIR_TERM_TABULATED))
			continue;
		const float3 dq = sub3(q, p);
		const float3 r = wrap_diff(box, lower ? dq : make_float3(-dq.x, -dq.y, -dq.z));
		const float d = sqrtf(len2(r));
		if (d < 1e-6f)
			continue;
		if (cutoff2 > 0.0f && d * d > cutoff2)
			continue;
		const float inv = lower ? 1.0f / d : -(1.0f / d);
		const float4 tp = lds_nb.tab[tag >> MARSMD_PAIR_TABLE_SHIFT]; // {start, step_inv, size, offset}
		const float w = (d - tp.x) * tp.y;
		int h = int(floorf(w));
		if (h >= __float_as_int(tp.z) - 1)
			continue;
		if (h < 0)
			h = 0;
		const uint64_t at = nbr_slot_index(base, uint32_t(t), m);
		rec_y[at] = L.nbr[nbr_slot_index(base, uint32_t(t), k)];
		const int code = int((tag >> MARSMD_PAIR_TABLE_SHIFT) << kFmKnotBits) | (__float_as_int(tp.w) + h);
		rec_u[at] = make_float4(inv * r.x, inv * r.y, inv * r.z, __int_as_float(code));
		++m;
	}
	rec_count[t] = m;
}

// Rows of W floats per particle (snapshot k = blockIdx.y) from particle order into
// the snapshot's sorted (slot) order: out[k n + t] = in[k n + sorted_idx_k[t]].
template<int W>
__global__ void k_fm_to_sorted(const FmSnapView* __restrict__ snaps, uint64_t n, const float* __restrict__ in,
							   float* __restrict__ out) {
	const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	const uint64_t k = blockIdx.y;
	const uint64_t src = (k * n + snaps[k].sorted_idx[t]) * W, dst = (k * n + t) * W;
	if constexpr (W == 4) {
		*reinterpret_cast<float4*>(out + dst) = *reinterpret_cast<const float4*>(in + src);
	} else {
#pragma unroll
		for (int c = 0; c < W; ++c)
			out[dst + c] = in[src + c];
	}
}

// dL/dknot of every snapshot in one persistent launch (work items as in k_fm_forces),
// from the records and dL/dF in slot order (grad_all): slot t adds, for each record
// (o, k, v) of table j,
// g = ((G_o - G_t) . v) step_inv_j to S[k], in FP64; records are read four at a time.
// Each pair adds g to knot k and -g to knot k + 1, so dL/dU[k] = S[k] - S[k - 1] (a
// table's last knot never starts a record's bin, so its S is 0 and the difference
// never reaches across tables); the reduction forms it. Per-workgroup FP64 sums in LDS
// are written to partial[blockIdx.x][knot], or global FP64 atomics go into partial[0]
// when the tables exceed the LDS budget. Each workgroup visits a fixed sequence of work items, but the
// order of the atomic additions within it is not fixed: the sums are deterministic up
// to FP64 rounding (about 1e-16 relative), not bit for bit.
template<bool LDS>
__global__ __launch_bounds__(kFmAdjBlock) void k_fm_adjoint_rec(const FmSnapView* __restrict__ snaps, uint64_t n,
																uint32_t tiles, uint32_t items, uint32_t total_knots,
																const float4* __restrict__ grad_all, FmSteps steps,
																double* __restrict__ partial) {
	extern __shared__ double lds_grad[];
	__shared__ float lds_step[kMaxTables];
	double* acc = LDS ? lds_grad : partial;
	if (LDS)
		for (uint32_t i = threadIdx.x; i < total_knots; i += blockDim.x)
			lds_grad[i] = 0.0;
	if (threadIdx.x < uint32_t(kMaxTables))
		lds_step[threadIdx.x] = steps.step_inv[threadIdx.x];
	__syncthreads();
	const TileRange R = xcd_tiles(items);
	for (uint32_t w = R.tile; w < R.end; w += R.stride) {
		const uint32_t snap = w / tiles;
		const uint64_t t = uint64_t(w - snap * tiles) * kFmAdjBlock + threadIdx.x;
		if (t >= n)
			continue;
		const FmSnapView V = snaps[snap];
		const float4* __restrict__ grad = grad_all + uint64_t(snap) * n; // slot order
		const uint32_t cnt = V.rec_count[t];
		if (cnt == 0)
			continue;
		const float4 gx = grad[t];
		const uint64_t row = V.tile_base[t >> 6] + uint64_t(t & 63u) * 4u;
		// group g's records (4 at a time) and the gradients at their other ends; the
		// next group's records are in flight while one group is evaluated
		uint4 ys = *reinterpret_cast<const uint4*>(V.rec_y + row);
		float4 uv[4], gy[4];
#pragma unroll
		for (int e = 0; e < 4; ++e)
			uv[e] = V.rec_u[row + uint64_t(e)];
		{
			const uint32_t yv[4] = {ys.x, ys.y, ys.z, ys.w};
#pragma unroll
			for (int e = 0; e < 4; ++e)
				gy[e] = uint32_t(e) < cnt ? grad[yv[e]] : gx; // past the count: unwritten records
		}
		for (uint32_t g0 = 0;; g0 += 4) {
			const bool more = g0 + 4 < cnt;
			const uint64_t nat = row + uint64_t((g0 >> 2) + 1) * 256u;
			uint4 ysn = ys;
			float4 un[4] = {uv[0], uv[1], uv[2], uv[3]};
			if (more) {
				ysn = *reinterpret_cast<const uint4*>(V.rec_y + nat);
#pragma unroll
				for (int e = 0; e < 4; ++e)
					un[e] = V.rec_u[nat + uint64_t(e)];
			}
#pragma unroll
			for (int e = 0; e < 4; ++e) {
				if (g0 + uint32_t(e) >= cnt)
					break;
				const float4 u = uv[e];
				const int code = __float_as_int(u.w);
				const double dfm = (double(gy[e].x) - double(gx.x)) * double(u.x) +
								   (double(gy[e].y) - double(gx.y)) * double(u.y) +
								   (double(gy[e].z) - double(gx.z)) * double(u.z);
				// S[k] += g only: the pair's -g on knot k + 1 is applied once at the end,
				// as dL/dU[k] = S[k] - S[k - 1] (half the atomic additions)
				atomicAdd(&acc[code & ((1 << kFmKnotBits) - 1)], dfm * double(lds_step[code >> kFmKnotBits]));
			}
			if (!more)
				break;
			ys = ysn;
			const uint32_t yv[4] = {ys.x, ys.y, ys.z, ys.w};
#pragma unroll
			for (int e = 0; e < 4; ++e) {
				uv[e] = un[e];
				gy[e] = g0 + 4u + uint32_t(e) < cnt ? grad[yv[e]] : gx;
			}
		}
	}
	if (LDS) {
		__syncthreads();
		double* __restrict__ out = partial + uint64_t(blockIdx.x) * uint64_t(total_knots);
		for (uint32_t i = threadIdx.x; i < total_knots; i += blockDim.x)
			out[i] = lds_grad[i];
	}
}

// seg[y][k] = sum of partial[b][k] over the rows b of segment y (rows in order), then
// out[k] = sum of seg[y][k] over the segments in order.
constexpr uint32_t kFmSegs = 32;
__global__ void k_fm_reduce_rows(const double* __restrict__ partial, uint32_t blocks, uint32_t total,
								 double* __restrict__ seg) {
	const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= total)
		return;
	const uint32_t y = blockIdx.y;
	const uint32_t per = (blocks + kFmSegs - 1) / kFmSegs;
	const uint32_t b0 = y * per, b1 = min(blocks, b0 + per);
	double s = 0.0;
	for (uint32_t b = b0; b < b1; ++b)
		s += partial[uint64_t(b) * total + k];
	seg[uint64_t(y) * total + k] = s;
}
// out[k] = S[k] - S[k - 1], S[k] the sum of seg[y][k] over the segments in order.
__global__ void k_fm_reduce_segs(const double* __restrict__ seg, uint32_t total, double* __restrict__ out) {
	const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= total)
		return;
	double s = 0.0, sp = 0.0;
	for (uint32_t y = 0; y < kFmSegs; ++y) {
		s += seg[uint64_t(y) * total + k];
		if (k > 0)
			sp += seg[uint64_t(y) * total + k - 1];
	}
	out[k] = s - sp;
}

// The mean squared force error and its gradient in one pass over S x n records:
// d = F - target (FP32), G = d * scale (scale = 2 / (3 S n), FP32, the .w lane 0),
// and per-block FP64 sums of d . d (a fixed tree inside the block).
constexpr int kMseBlock = 256;
__global__ __launch_bounds__(kMseBlock) void k_fm_mse(const float4* __restrict__ f, const float* __restrict__ target3,
													  uint64_t count, float scale, float4* __restrict__ grad,
													  double* __restrict__ block_sums) {
	__shared__ double part[kMseBlock];
	double acc = 0.0;
	for (uint64_t i = uint64_t(blockIdx.x) * kMseBlock + threadIdx.x; i < count;
		 i += uint64_t(gridDim.x) * kMseBlock) {
		const float4 v = f[i];
		const float dx = v.x - target3[3 * i], dy = v.y - target3[3 * i + 1], dz = v.z - target3[3 * i + 2];
		grad[i] = make_float4(dx * scale, dy * scale, dz * scale, 0.0f);
		acc += double(dx) * double(dx) + double(dy) * double(dy) + double(dz) * double(dz);
	}
	part[threadIdx.x] = acc;
	__syncthreads();
	for (int o = kMseBlock / 2; o > 0; o >>= 1) {
		if (int(threadIdx.x) < o)
			part[threadIdx.x] += part[threadIdx.x + o];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		block_sums[blockIdx.x] = part[0];
}
// loss = (sum of the block sums) / (3 S n) with one workgroup in a fixed order: thread t
// sums blocks t, t + 256, ..., then a tree over the threads.
__global__ __launch_bounds__(kMseBlock) void k_fm_mse_total(const double* __restrict__ block_sums, uint32_t blocks,
														   double inv_count, double* __restrict__ loss) {
	__shared__ double part[kMseBlock];
	double acc = 0.0;
	for (uint32_t b = threadIdx.x; b < blocks; b += kMseBlock)
		acc += block_sums[b];
	part[threadIdx.x] = acc;
	__syncthreads();
	for (int o = kMseBlock / 2; o > 0; o >>= 1) {
		if (int(threadIdx.x) < o)
			part[threadIdx.x] += part[threadIdx.x + o];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		*loss = part[0] * inv_count;
}

// The loss of every snapshot in sorted order in one persistent launch (work items as
// in k_fm_forces): slot t's force F goes straight into k_fm_mse's terms, d = F - target
// (FP32), grad = d * scale and the FP64 d . d summed per workgroup in a fixed tree, so
// grad has k_fm_mse's bits and F itself is never stored; k_fm_mse_total then forms
// the loss from the block sums (in another grouping than the two-pass path's).
__global__ __launch_bounds__(kFmBlock) void k_fm_loss(const FmSnapView* __restrict__ snaps, uint64_t n,
													  uint32_t tiles, uint32_t items, NbTables T,
													  const float* __restrict__ knots_g, DevBox box, float cut_d,
													  const float* __restrict__ target3, float scale,
													  float4* __restrict__ grad, double* __restrict__ block_sums) {
	static_assert(kFmBlock == kMseBlock, "one tree over the workgroup");
	extern __shared__ float lds_knots[];
	__shared__ NbLds lds_nb;
	__shared__ double part[kFmBlock];
	stage_tables(T, knots_g, lds_knots, lds_nb);
	double acc = 0.0;
	const TileRange R = xcd_tiles(items);
	for (uint32_t w = R.tile; w < R.end; w += R.stride) {
		const uint32_t k = w / tiles;
		const uint64_t t = uint64_t(w - k * tiles) * kFmBlock + threadIdx.x;
		if (t >= n)
			continue;
		const FmSnapView V = snaps[k];
		const NbrList L{V.counts, V.tile_base, V.nbr};
		const float4 f = nb_force(uint32_t(t), V.spos[t], V.spos, L, T, lds_knots, knots_g, lds_nb, box, cut_d, false);
		const uint64_t i = uint64_t(k) * n + t;
		const float dx = f.x - target3[3 * i], dy = f.y - target3[3 * i + 1], dz = f.z - target3[3 * i + 2];
		grad[i] = make_float4(dx * scale, dy * scale, dz * scale, 0.0f);
		acc += double(dx) * double(dx) + double(dy) * double(dy) + double(dz) * double(dz);
	}
	part[threadIdx.x] = acc;
	__syncthreads();
	for (int o = kFmBlock / 2; o > 0; o >>= 1) {
		if (int(threadIdx.x) < o)
			part[threadIdx.x] += part[threadIdx.x + o];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		block_sums[blockIdx.x] = part[0];
}

__global__ void k_fill_f64(double* p, uint64_t n) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = 0.0;
}

} // namespace

struct FMImpl::Snap {
	DevArray<float4> spos;
	DevArray<uint32_t> sorted_idx, counts, tile_base, nbr;
	DevArray<uint32_t> rec_count, rec_y;
	DevArray<float4> rec_u;
};

FMImpl::FMImpl(const marsmd_fm_desc& d) : tables(new NbTables{}) {
	if (d.num_types < 1 || d.num_types > MARSMD_MAX_TYPES)
		raise(MARSMD_ERR_VALUE, "force matching: num_types must be in [1, 16]");
	if (d.num_tables < 1 || d.num_tables > kMaxTables)
		raise(MARSMD_ERR_VALUE, "force matching: num_tables must be in [1, 32]");
	if (d.num_particles == 0 || d.num_snapshots == 0 || !d.positions || !d.type_ids)
		raise(MARSMD_ERR_VALUE, "force matching: snapshots, particles, positions and types are required");
	if (d.num_particles >= kMaxParticles)
		raise(MARSMD_ERR_VALUE, "force matching: particle count must be below 2^24");
	n = d.num_particles;
	S = d.num_snapshots;
	if ((n + kFmBlock - 1) / kFmBlock * S >= (uint64_t(1) << 32))
		raise(MARSMD_ERR_VALUE, "force matching: snapshots x particles / 256 must be below 2^32");
	if (S > 65535)
		raise(MARSMD_ERR_VALUE, "force matching: at most 65535 snapshots");
	box = d.box;
	cutoff2 = d.cutoff > 0.0f ? d.cutoff * d.cutoff : 0.0f;
	NbTables& T = *tables;
	T.num_types = d.num_types;
	T.num_tables = d.num_tables;
	T.enabled = MARSMD_PAIR_TERM_TABULATED;
	for (int i = 0; i < d.num_types * d.num_types; ++i)
		T.tag[i] = d.term_matrix ? d.term_matrix[i] : 0u;
	int total = 0;
	for (int k = 0; k < d.num_tables; ++k) {
		const marsmd_table& t = d.tables[k];
		if (t.size < 2 || t.is_periodic)
			raise(MARSMD_ERR_VALUE, "force matching: tables need >= 2 knots and are not periodic");
		T.start[k] = t.start;
		T.step_inv[k] = t.step_inv;
		T.size[k] = int(t.size);
		T.offset[k] = total;
		total += int(t.size);
	}
	for (int i = 0; i < d.num_types * d.num_types; ++i)
		if ((T.tag[i] & MARSMD_PAIR_TERM_TABULATED) && int(T.tag[i] >> MARSMD_PAIR_TABLE_SHIFT) >= d.num_tables)
			raise(MARSMD_ERR_VALUE, "force matching: a pair tag names a missing table");
	T.total_knots = total;
	T.lds = total <= kLdsKnots ? 1 : 0;
	if (total >= (1 << kFmKnotBits))
		raise(MARSMD_ERR_VALUE, "force matching: at most 2^26 - 1 knots in all");
	PairlistImpl pl(n, ~uint64_t(0));
	pl.set_box(box);
	const float plc = d.pairlist_cutoff > 0.0f ? d.pairlist_cutoff : d.cutoff;
	for (uint64_t s = 0; s < S; ++s) {
		pl.build(reinterpret_cast<const float4*>(d.positions) + s * n, nullptr, false, d.type_ids, n, plc, nullptr);
		Snap* sn = new Snap;
		snaps.push_back(sn);
		const uint64_t tiles = (n + 63) / 64;
		sn->spos.resize(n);
		sn->sorted_idx.resize(n);
		sn->counts.resize(n);
		sn->tile_base.resize(tiles + 1);
		sn->nbr.resize(std::max<uint64_t>(pl.nbr_storage, 1));
		MMD_HIP(hipMemcpyAsync(sn->spos.data(), pl.spos.data(), 16 * n, hipMemcpyDeviceToDevice, nullptr));
		MMD_HIP(hipMemcpyAsync(sn->sorted_idx.data(), pl.sorter.sorted_idx.data(), 4 * n, hipMemcpyDeviceToDevice,
							   nullptr));
		MMD_HIP(hipMemcpyAsync(sn->counts.data(), pl.counts.data(), 4 * n, hipMemcpyDeviceToDevice, nullptr));
		MMD_HIP(hipMemcpyAsync(sn->tile_base.data(), pl.tile_base.data(), 4 * (tiles + 1), hipMemcpyDeviceToDevice,
							   nullptr));
		if (pl.nbr_storage)
			MMD_HIP(hipMemcpyAsync(sn->nbr.data(), pl.nbr.data(), 4 * pl.nbr_storage, hipMemcpyDeviceToDevice, nullptr));
		num_pairs += pl.num_pairs();
		// the adjoint's records, laid out like the rows (at most as many as row entries)
		sn->rec_count.resize(n);
		sn->rec_y.resize(std::max<uint64_t>(pl.nbr_storage, 4));
		sn->rec_u.resize(std::max<uint64_t>(pl.nbr_storage, 4));
		const NbrList L{sn->counts.data(), sn->tile_base.data(), sn->nbr.data()};
		k_fm_records<<<grid_for(n, kFmBlock), kFmBlock, 0, nullptr>>>(sn->spos.data(), n, L, T, make_devbox(box),
																	  cutoff2, sn->rec_count.data(), sn->rec_y.data(),
																	  sn->rec_u.data());
		MMD_LAUNCH_CHECK();
	}
	std::vector<FmSnapView> views(S);
	for (uint64_t k = 0; k < S; ++k)
		views[k] = FmSnapView{snaps[k]->spos.data(),   snaps[k]->sorted_idx.data(), snaps[k]->counts.data(),
							  snaps[k]->tile_base.data(), snaps[k]->nbr.data(),  snaps[k]->rec_count.data(),
							  snaps[k]->rec_y.data(),  snaps[k]->rec_u.data()};
	snap_views.resize(S);
	MMD_HIP(hipMemcpy(snap_views.data(), views.data(), sizeof(FmSnapView) * S, hipMemcpyHostToDevice));
	MMD_HIP(hipDeviceSynchronize());
}

FMImpl::~FMImpl() {
	for (Snap* s : snaps)
		delete s;
	delete tables;
}

// Persistent workgroups for a kernel over the S x tiles work items: as many as are
// resident (occupancy x CUs), a multiple of 8, at most the items rounded up to 8.
static unsigned fm_grid(const void* kernel, unsigned lds, uint64_t items, int block = kFmBlock) {
	int occ = 0, dev = 0, cus = 0;
	MMD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, block, lds));
	MMD_HIP(hipGetDevice(&dev));
	MMD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
	uint64_t g = uint64_t(std::max(occ, 1)) * uint64_t(std::max(cus, 1));
	g = std::min<uint64_t>(g, (items + 7) / 8 * 8);
	return unsigned(std::max<uint64_t>((g + 7) / 8 * 8, 8));
}

void FMImpl::forces(const float* knots, float* out4, hipStream_t s, bool sorted) {
	const unsigned lds = tables->lds ? unsigned(std::max(tables->total_knots, 2)) * 4u : 0u;
	const uint32_t tiles = uint32_t((n + kFmBlock - 1) / kFmBlock);
	const uint64_t items = uint64_t(tiles) * S;
	const unsigned grid = fm_grid(reinterpret_cast<const void*>(k_fm_forces), lds, items);
	k_fm_forces<<<grid, kFmBlock, lds, s>>>(snap_views.data(), n, tiles, uint32_t(items), *tables, knots,
											make_devbox(box), cutoff_threshold(cutoff2), sorted ? 1 : 0,
											reinterpret_cast<float4*>(out4));
	MMD_LAUNCH_CHECK();
}

void FMImpl::mse(const float* force4, const float* target3, float* grad4, double* loss, hipStream_t s) {
	const uint64_t count = S * n;
	const uint32_t blocks = uint32_t(std::min<uint64_t>((count + kMseBlock - 1) / kMseBlock, 2048));
	mse_sums.resize(blocks);
	// the gradient's scale as the FP32 value of 2 / count (FP64 division)
	const float scale = float(2.0 / (3.0 * double(count)));
	k_fm_mse<<<blocks, kMseBlock, 0, s>>>(reinterpret_cast<const float4*>(force4), target3, count, scale,
										  reinterpret_cast<float4*>(grad4), mse_sums.data());
	MMD_LAUNCH_CHECK();
	k_fm_mse_total<<<1, kMseBlock, 0, s>>>(mse_sums.data(), blocks, 1.0 / (3.0 * double(count)), loss);
	MMD_LAUNCH_CHECK();
}

void FMImpl::loss_sorted(const float* knots, const float* target3, float* grad4, double* loss, hipStream_t s) {
	const unsigned lds = tables->lds ? unsigned(std::max(tables->total_knots, 2)) * 4u : 0u;
	const uint32_t tiles = uint32_t((n + kFmBlock - 1) / kFmBlock);
	const uint64_t items = uint64_t(tiles) * S;
	const uint64_t count = S * n;
	const unsigned grid = fm_grid(reinterpret_cast<const void*>(k_fm_loss), lds, items);
	mse_sums.resize(grid);
	const float scale = float(2.0 / (3.0 * double(count))); // as FMImpl::mse
	k_fm_loss<<<grid, kFmBlock, lds, s>>>(snap_views.data(), n, tiles, uint32_t(items), *tables, knots,
										   make_devbox(box), cutoff_threshold(cutoff2), target3, scale,
										   reinterpret_cast<float4*>(grad4), mse_sums.data());
	MMD_LAUNCH_CHECK();
	k_fm_mse_total<<<1, kMseBlock, 0, s>>>(mse_sums.data(), grid, 1.0 / (3.0 * double(count)), loss);
	MMD_LAUNCH_CHECK();
}

void FMImpl::to_sorted(const float* in, int width, float* out, hipStream_t s) {
	if (width != 3 && width != 4)
		raise(MARSMD_ERR_VALUE, "force matching: rows of 3 or 4 floats");
	const dim3 grid(grid_for(n, 256), unsigned(S));
	if (width == 4)
		k_fm_to_sorted<4><<<grid, 256, 0, s>>>(snap_views.data(), n, in, out);
	else
		k_fm_to_sorted<3><<<grid, 256, 0, s>>>(snap_views.data(), n, in, out);
	MMD_LAUNCH_CHECK();
}

void FMImpl::adjoint(const float* knots, const float* grad4, double* out, hipStream_t s, bool sorted) {
	(void)knots; // the gradient of a linear table force does not depend on the knots
	const uint32_t total = uint32_t(tables->total_knots);
	const bool lds = total <= uint32_t(kFmLdsGrad);
	const uint32_t tiles = uint32_t((n + kFmAdjBlock - 1) / kFmAdjBlock);
	const uint64_t items = uint64_t(tiles) * S;
	const float4* g = reinterpret_cast<const float4*>(grad4);
	if (!sorted) {
		grad_slot.resize(S * n);
		to_sorted(grad4, 4, reinterpret_cast<float*>(grad_slot.data()), s);
		g = grad_slot.data();
	}
	FmSteps steps{};
	std::memcpy(steps.step_inv, tables->step_inv, sizeof steps.step_inv);
	uint32_t rows = 1;
	if (lds) {
		const unsigned grid =
			fm_grid(reinterpret_cast<const void*>(k_fm_adjoint_rec<true>), total * 8u, items, kFmAdjBlock);
		rows = grid;
		partial.resize(uint64_t(rows) * total);
		k_fm_adjoint_rec<true><<<grid, kFmAdjBlock, total * 8u, s>>>(snap_views.data(), n, tiles, uint32_t(items),
																	total, g, steps, partial.data());
	} else {
		const unsigned grid =
			fm_grid(reinterpret_cast<const void*>(k_fm_adjoint_rec<false>), 0u, items, kFmAdjBlock);
		partial.resize(total);
		k_fill_f64<<<grid_for(total, 256), 256, 0, s>>>(partial.data(), total);
		MMD_LAUNCH_CHECK();
		k_fm_adjoint_rec<false><<<grid, kFmAdjBlock, 0, s>>>(snap_views.data(), n, tiles, uint32_t(items), total,
															 g, steps, partial.data());
	}
	MMD_LAUNCH_CHECK();
	seg_sums.resize(uint64_t(kFmSegs) * total);
	k_fm_reduce_rows<<<dim3(grid_for(total, 256), kFmSegs), 256, 0, s>>>(partial.data(), rows, total,
																		 seg_sums.data());
	MMD_LAUNCH_CHECK();
	k_fm_reduce_segs<<<grid_for(total, 256), 256, 0, s>>>(seg_sums.data(), total, out);
	MMD_LAUNCH_CHECK();
}

} // namespace mmd
