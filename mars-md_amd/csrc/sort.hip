// sort.hip — stable LSD radix sort of (uint32 key, uint32 value) pairs and an
// exclusive scan, written for gfx950 wave64.
//
// Replaces the CUB DeviceRadixSort::SortPairs call of the reference
// (src/PatchOperation/ZOrderKernels/CUDASort.cu:13-98; stability contract pinned by
// Unit_test/System/zorder_tests.cpp:139-151). One pass per digit of 8 bits, or of 10
// bits when that takes fewer passes (3 instead of 4 for 30-bit Morton keys, 1 instead
// of 2 for keys below 2^10); the last digit is masked to key_bits:
//   1. k_rs_hist     per-tile digit histogram (wave-aggregated LDS counts), and each
//                    digit's total over the tiles (one device-scope addition per tile
//                    and digit present),
//   2. k_rs_offsets  one workgroup per digit: the totals of the lower digits, then the
//                    exclusive scan of its row of tile counts (the digit-major scan
//                    over [digit][tile] in one launch),
//   3. k_rs_scatter  per-tile stable ranks from wave64 ballot peer masks, staged in
//                    LDS in digit order, then written to offset[digit][tile] + rank.
// A tile is 4 waves x 16 rows x 64 lanes = 4096 keys; wave w owns the contiguous
// quarter [w*1024, (w+1)*1024) of its tile, so ranks follow index order and the
// sort is stable. Both kernels issue all 16 row loads of a wave before ranking, so a
// wave waits on memory once per tile rather than once per row; the scatter stages the
// ranked tile in LDS and writes it out in digit order.
#include "internal.hpp"

#include <mutex>

namespace mmd {

namespace {

constexpr int kRsBlock = 256;
constexpr int kRsRows = 16;
constexpr int kRsTile = kRsBlock * kRsRows;
constexpr int kRsMaxBuckets = 1024; // 10-bit digits

__device__ __forceinline__ uint64_t lanemask_lt() {
	const int lane = threadIdx.x & 63;
	return (uint64_t(1) << lane) - 1;
}

// Lanes of this wave holding the same BITS-bit digit (and valid).
// Built on 32-bit halves, sx all ones where the lane's bit b is set: the lanes that agree
// on every bit stay (4 VALU a bit instead of 10 on 64-bit masks).
template<int BITS>
__device__ __forceinline__ uint64_t digit_peers(uint32_t d, bool valid) {
	const uint64_t v = __ballot(valid);
	uint32_t lo = uint32_t(v), hi = uint32_t(v >> 32);
#pragma unroll
	for (int b = 0; b < BITS; ++b) {
		const int32_t sx = int32_t(d << (31 - b)) >> 31;
		const uint64_t m = __ballot(sx != 0);
		lo &= ~(uint32_t(m) ^ uint32_t(sx));
		hi &= ~(uint32_t(m >> 32) ^ uint32_t(sx));
	}
	return valid ? (uint64_t(lo) | (uint64_t(hi) << 32)) : 0;
}

// Digit d of a key is (key >> shift) & mask, mask < 2^BITS. A tile's counts do not
// depend on which wave reads which key, so the histogram spreads a tile over 16 waves
// (4 rows of keys each): one workgroup per tile leaves the GPU mostly idle, and more
// waves per tile hide the loads' latency.
constexpr int kRsHistBlock = 1024;
template<int BITS>
__global__ __launch_bounds__(kRsHistBlock) void k_rs_hist(const uint32_t* __restrict__ keys, uint64_t n,
														  int shift, uint32_t mask,
														  uint32_t* __restrict__ tile_hist,
														  uint32_t num_tiles, uint32_t* __restrict__ totals) {
	constexpr uint32_t B = 1u << BITS;
	constexpr int kWaves = kRsHistBlock / 64;
	constexpr int kRows = kRsTile / kRsHistBlock;
	static_assert(kRows * kRsHistBlock == kRsTile, "whole rows of keys per wave");
	__shared__ uint32_t h[B];
	for (uint32_t i = threadIdx.x; i < B; i += kRsHistBlock)
		h[i] = 0;
	const int lane = threadIdx.x & 63;
	const int w = threadIdx.x >> 6;
	const uint64_t wbase = uint64_t(blockIdx.x) * kRsTile + uint64_t(w) * (kRsTile / kWaves);
	uint32_t mk[kRows];
#pragma unroll
	for (int r = 0; r < kRows; ++r) {
		const uint64_t i = wbase + uint64_t(r) * 64 + lane;
		mk[r] = i < n ? keys[i] : 0u;
	}
	__syncthreads();
#pragma unroll
	for (int r = 0; r < kRows; ++r) {
		const bool valid = wbase + uint64_t(r) * 64 + lane < n;
		const uint32_t d = (mk[r] >> shift) & mask;
		const uint64_t peers = digit_peers<BITS>(d, valid);
		if (valid && (peers & lanemask_lt()) == 0)
			atomicAdd(&h[d], uint32_t(__popcll(peers)));
	}
	__syncthreads();
	for (uint32_t i = threadIdx.x; i < B; i += kRsHistBlock) {
		const uint32_t c = h[i];
		tile_hist[uint64_t(i) * num_tiles + blockIdx.x] = c;
		if (c)
			atomicAdd(&totals[i], c);
	}
}

// Block-wide exclusive scan for blockDim.x = nthreads (multiple of 64, <= 1024).
__device__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* wsum, uint32_t* total) {
	const int lane = threadIdx.x & 63;
	const int w = threadIdx.x >> 6;
	const int nw = blockDim.x >> 6;
	uint32_t incl = v;
#pragma unroll
	for (int o = 1; o < 64; o <<= 1) {
		const uint32_t t = __shfl_up(incl, o, 64);
		if (lane >= o)
			incl += t;
	}
	if (lane == 63)
		wsum[w] = incl;
	__syncthreads();
	if (threadIdx.x == 0) {
		uint32_t acc = 0;
		for (int k = 0; k < nw; ++k) {
			const uint32_t s = wsum[k];
			wsum[k] = acc;
			acc += s;
		}
		wsum[nw] = acc;
	}
	__syncthreads();
	const uint32_t out = wsum[w] + incl - v;
	if (total)
		*total = wsum[nw];
	__syncthreads();
	return out;
}

// Digit d = blockIdx.x: row d of tile counts becomes its offsets, (sum of the totals of
// the digits below d) + (exclusive scan of the row), in place; the next pass's total of
// digit d is cleared (that buffer is not read here).
__global__ __launch_bounds__(kRsBlock) void k_rs_offsets(uint32_t* __restrict__ tile_hist, uint32_t num_tiles,
														 const uint32_t* __restrict__ totals,
														 uint32_t* __restrict__ totals_next) {
	__shared__ uint32_t wsum[kRsBlock / 64 + 1];
	const uint32_t d = blockIdx.x;
	uint32_t below = 0;
	for (uint32_t i = threadIdx.x; i < d; i += kRsBlock)
		below += totals[i];
	uint32_t base;
	(void)block_exclusive_scan(below, wsum, &base);
	if (threadIdx.x == 0)
		totals_next[d] = 0u;
	uint32_t* __restrict__ row = tile_hist + uint64_t(d) * num_tiles;
	for (uint32_t c0 = 0; c0 < num_tiles; c0 += kRsBlock) {
		const uint32_t i = c0 + threadIdx.x;
		const uint32_t v = i < num_tiles ? row[i] : 0u;
		uint32_t chunk;
		const uint32_t ex = block_exclusive_scan(v, wsum, &chunk);
		if (i < num_tiles)
			row[i] = base + ex;
		base += chunk;
	}
}

// In-place exclusive scan of `count` values with one block of 1024 threads.
__global__ __launch_bounds__(1024) void k_scan_one_block(uint32_t* __restrict__ data, uint64_t count,
														 uint32_t* __restrict__ total_out) {
	__shared__ uint32_t wsum[17];
	const uint64_t per = (count + blockDim.x - 1) / blockDim.x;
	const uint64_t b = uint64_t(threadIdx.x) * per;
	const uint64_t e = b + per < count ? b + per : count;
	uint32_t s = 0;
	for (uint64_t i = b; i < e; ++i)
		s += data[i];
	uint32_t total;
	uint32_t run = block_exclusive_scan(s, wsum, &total);
	for (uint64_t i = b; i < e; ++i) {
		const uint32_t v = data[i];
		data[i] = run;
		run += v;
	}
	if (threadIdx.x == 0 && total_out)
		*total_out = total;
}

// Two-phase scatter: ranks keys within the tile by digit (stable), stages the tile in
// LDS in digit order, then writes it out with consecutive lanes on consecutive output
// addresses (one run per digit present in the tile) instead of one scattered address
// per lane.
template<int BITS>
__global__ __launch_bounds__(kRsBlock) void k_rs_scatter(const uint32_t* __restrict__ keys,
														 const uint32_t* __restrict__ vals,
														 uint32_t* __restrict__ keys_out,
														 uint32_t* __restrict__ vals_out, uint64_t n,
														 int shift, uint32_t mask,
														 const uint32_t* __restrict__ tile_offsets,
														 uint32_t num_tiles) {
	constexpr uint32_t B = 1u << BITS;
	constexpr int kPer = int(B) / kRsBlock; // digits per thread in the tile-local scan
	static_assert(kPer >= 1 && kPer * kRsBlock == int(B), "digit count must be a multiple of the block");
	__shared__ uint32_t adj[B];		// output offset of the digit's run minus its start in the tile
	__shared__ uint32_t wcnt[4][B]; // per-wave digit counts, then per-wave starts in the tile
	__shared__ uint32_t stage_k[kRsTile];
	__shared__ uint32_t stage_v[kRsTile];
	__shared__ uint32_t wsum[kRsBlock / 64 + 1];
	const int lane = threadIdx.x & 63;
	const int w = threadIdx.x >> 6;
	uint32_t mk[kRsRows], mv[kRsRows], mr[kRsRows];
	const uint64_t tile0 = uint64_t(blockIdx.x) * kRsTile;
	const uint64_t wbase = tile0 + uint64_t(w) * (kRsTile / 4);
#pragma unroll
	for (int r = 0; r < kRsRows; ++r) {
		const uint64_t i = wbase + uint64_t(r) * 64 + lane;
		mk[r] = i < n ? keys[i] : 0u;
		mv[r] = i < n ? vals[i] : 0u;
	}
	for (uint32_t i = threadIdx.x; i < B; i += kRsBlock)
		adj[i] = tile_offsets[uint64_t(i) * num_tiles + blockIdx.x];
	for (uint32_t i = uint32_t(lane); i < B; i += 64)
		wcnt[w][i] = 0;
	__syncthreads();
#pragma unroll
	for (int r = 0; r < kRsRows; ++r) {
		const bool valid = wbase + uint64_t(r) * 64 + lane < n;
		const uint32_t d = (mk[r] >> shift) & mask;
		const uint64_t peers = digit_peers<BITS>(d, valid);
		const uint64_t below = peers & lanemask_lt();
		const uint32_t rank = wcnt[w][d] + uint32_t(__popcll(below));
		__builtin_amdgcn_wave_barrier();
		if (valid && below == 0)
			wcnt[w][d] += uint32_t(__popcll(peers));
		__builtin_amdgcn_wave_barrier();
		mr[r] = valid ? rank : 0xFFFFFFFFu;
	}
	__syncthreads();
	// Tile-local start of each digit (exclusive scan over digits), then each wave's start
	// within the digit; thread t owns digits [t*kPer, (t+1)*kPer).
	uint32_t total = 0;
#pragma unroll
	for (int k = 0; k < kPer; ++k) {
		const uint32_t d = uint32_t(threadIdx.x) * kPer + k;
		total += wcnt[0][d] + wcnt[1][d] + wcnt[2][d] + wcnt[3][d];
	}
	uint32_t run = block_exclusive_scan(total, wsum, nullptr);
#pragma unroll
	for (int k = 0; k < kPer; ++k) {
		const uint32_t d = uint32_t(threadIdx.x) * kPer + k;
		adj[d] -= run;
#pragma unroll
		for (int q = 0; q < 4; ++q) {
			const uint32_t c = wcnt[q][d];
			wcnt[q][d] = run;
			run += c;
		}
	}
	__syncthreads();
#pragma unroll
	for (int r = 0; r < kRsRows; ++r) {
		if (mr[r] == 0xFFFFFFFFu)
			continue;
		const uint32_t d = (mk[r] >> shift) & mask;
		const uint32_t pos = wcnt[w][d] + mr[r];
		stage_k[pos] = mk[r];
		stage_v[pos] = mv[r];
	}
	__syncthreads();
	const uint32_t tile_n = n - tile0 < uint64_t(kRsTile) ? uint32_t(n - tile0) : uint32_t(kRsTile);
#pragma unroll
	for (int r = 0; r < kRsRows; ++r) {
		const uint32_t j = uint32_t(r) * kRsBlock + threadIdx.x;
		if (j < tile_n) {
			const uint32_t k = stage_k[j];
			const uint32_t dst = adj[(k >> shift) & mask] + j;
			keys_out[dst] = k;
			vals_out[dst] = stage_v[j];
		}
	}
}

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const uint32_t* __restrict__ in, uint64_t n,
															uint32_t* __restrict__ block_sums) {
	__shared__ uint32_t wsum[5];
	const uint64_t base = uint64_t(blockIdx.x) * kScanTile + uint64_t(threadIdx.x) * kScanItems;
	uint32_t s = 0;
#pragma unroll
	for (int k = 0; k < kScanItems; ++k)
		if (base + k < n)
			s += in[base + k];
	uint32_t total;
	(void)block_exclusive_scan(s, wsum, &total);
	if (threadIdx.x == 0)
		block_sums[blockIdx.x] = total;
}

// `in` may alias `out` (in-place scan): each thread reads its items before writing them.
__global__ __launch_bounds__(kScanBlock) void k_scan_apply(const uint32_t* in, uint64_t n,
														   const uint32_t* __restrict__ block_offsets,
														   uint32_t* out) {
	__shared__ uint32_t wsum[5];
	const uint64_t base = uint64_t(blockIdx.x) * kScanTile + uint64_t(threadIdx.x) * kScanItems;
	uint32_t v[kScanItems];
	uint32_t s = 0;
#pragma unroll
	for (int k = 0; k < kScanItems; ++k) {
		v[k] = base + k < n ? in[base + k] : 0u;
		s += v[k];
	}
	uint32_t run = block_exclusive_scan(s, wsum, nullptr) + block_offsets[blockIdx.x];
#pragma unroll
	for (int k = 0; k < kScanItems; ++k) {
		if (base + k < n)
			out[base + k] = run;
		run += v[k];
	}
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = v;
}
__global__ void k_fill_f32(float* p, float v, uint64_t n) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = v;
}

inline uint64_t align256(uint64_t b) { return (b + 255) & ~uint64_t(255); }

} // namespace

uint64_t radix_sort_scratch_bytes(uint64_t n) {
	const uint64_t tiles = (n + kRsTile - 1) / kRsTile;
	return align256(n * 4) * 2 + align256(tiles * kRsMaxBuckets * 4 + 4) + align256(2 * kRsMaxBuckets * 4);
}

void radix_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
					  uint32_t* vals_out, uint64_t n, int key_bits, void* scratch,
					  uint64_t scratch_bytes, hipStream_t s) {
	if (n == 0)
		return;
	if (n > 0xFFFFFFFFull)
		raise(MARSMD_ERR_VALUE, "radix sort: more than 2^32 keys");
	if (scratch_bytes < radix_sort_scratch_bytes(n))
		raise(MARSMD_ERR_VALUE, "radix sort: scratch too small");
	if (key_bits < 1 || key_bits > 32)
		raise(MARSMD_ERR_VALUE, "radix sort: key_bits must be in [1, 32]");
	// the first pass may write the outputs while it reads the inputs
	const auto overlap = [n](const uint32_t* a, const uint32_t* b) { return a < b + n && b < a + n; };
	if (overlap(keys_in, keys_out) || overlap(keys_in, vals_out) || overlap(vals_in, keys_out) ||
		overlap(vals_in, vals_out) || overlap(keys_out, vals_out))
		raise(MARSMD_ERR_VALUE, "radix sort: output buffers must not overlap the inputs or each other");
	const uint32_t tiles = uint32_t((n + kRsTile - 1) / kRsTile);
	char* sp = static_cast<char*>(scratch);
	uint32_t* tmp_k = reinterpret_cast<uint32_t*>(sp);
	uint32_t* tmp_v = reinterpret_cast<uint32_t*>(sp + align256(n * 4));
	uint32_t* hist = reinterpret_cast<uint32_t*>(sp + 2 * align256(n * 4));
	// digit totals of this pass and the next (k_rs_offsets clears the next pass's)
	uint32_t* totals = reinterpret_cast<uint32_t*>(sp + 2 * align256(n * 4) +
												   align256(uint64_t(tiles) * kRsMaxBuckets * 4 + 4));
	const int bits = (key_bits + 9) / 10 < (key_bits + 7) / 8 ? 10 : 8;
	const int passes = (key_bits + bits - 1) / bits;
	const uint64_t buckets = uint64_t(1) << bits;
	const uint32_t* src_k = keys_in;
	const uint32_t* src_v = vals_in;
	fill_u32(totals, 0u, buckets, s);
	for (int p = 0; p < passes; ++p) {
		uint32_t* tot = totals + (p & 1) * kRsMaxBuckets;
		uint32_t* tot_next = totals + ((p + 1) & 1) * kRsMaxBuckets;
		const bool to_out = ((passes - 1 - p) % 2) == 0;
		uint32_t* dk = to_out ? keys_out : tmp_k;
		uint32_t* dv = to_out ? vals_out : tmp_v;
		const int shift = bits * p;
		const int width = key_bits - shift < bits ? key_bits - shift : bits;
		const uint32_t mask = (1u << width) - 1u;
		if (bits == 10)
			k_rs_hist<10><<<tiles, kRsHistBlock, 0, s>>>(src_k, n, shift, mask, hist, tiles, tot);
		else
			k_rs_hist<8><<<tiles, kRsHistBlock, 0, s>>>(src_k, n, shift, mask, hist, tiles, tot);
		MMD_LAUNCH_CHECK();
		k_rs_offsets<<<unsigned(buckets), kRsBlock, 0, s>>>(hist, tiles, tot, tot_next);
		MMD_LAUNCH_CHECK();
		if (bits == 10)
			k_rs_scatter<10><<<tiles, kRsBlock, 0, s>>>(src_k, src_v, dk, dv, n, shift, mask, hist, tiles);
		else
			k_rs_scatter<8><<<tiles, kRsBlock, 0, s>>>(src_k, src_v, dk, dv, n, shift, mask, hist, tiles);
		MMD_LAUNCH_CHECK();
		src_k = dk;
		src_v = dv;
	}
}

uint64_t scan_scratch_bytes(uint64_t n) {
	const uint64_t blocks = (n + kScanTile - 1) / kScanTile;
	return align256((blocks + 1) * 4);
}

// ------------------------------------------------- sort without allocation
// A workspace for radix_sort_pairs_static, fixed at build time: the same layout as
// radix_sort_scratch_bytes(kStaticSortMaxKeys) (two key/value staging arrays, the
// [digit][tile] histogram, two rows of digit totals). Every GPU that loads the library
// holds one copy until the process ends. Nothing in it has to start at zero: each pass
// writes every histogram entry before reading it, and the digit totals are cleared by a
// kernel at the start of each sort (and by each pass for the next one).
namespace {
constexpr uint64_t cx_align256(uint64_t b) { return (b + 255) & ~uint64_t(255); }
constexpr uint64_t cx_sort_scratch_bytes(uint64_t n) {
	return cx_align256(n * 4) * 2 + cx_align256((n + kRsTile - 1) / kRsTile * kRsMaxBuckets * 4 + 4) +
		   cx_align256(2 * kRsMaxBuckets * 4);
}
constexpr uint64_t kStaticSortScratchBytes = cx_sort_scratch_bytes(kStaticSortMaxKeys);

static __device__ __align__(256) unsigned char scratch[kStaticSortScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void* scratch_addr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per device
// (each GPU holds the array at its own address; its host-side name is not one).
void* static_sort_scratch() {
	int dev = 0;
	MMD_HIP(hipGetDevice(&dev));
	if (dev < 0 || dev >= kMaxDevices)
		raise(MARSMD_ERR_VALUE, "radix sort (static workspace): device ordinal " + std::to_string(dev) +
									" beyond the " + std::to_string(kMaxDevices) + " supported");
	std::lock_guard<std::mutex> lock(scratch_mu);
	if (!scratch_addr[dev]) {
		void* p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p)
			raise(MARSMD_ERR_DEVICE, std::string("radix sort (static workspace): hipGetSymbolAddress failed: ") +
									  hipGetErrorString(e));
		scratch_addr[dev] = p;
	}
	return scratch_addr[dev];
}
} // namespace

uint64_t static_sort_scratch_bytes() { return kStaticSortScratchBytes; }

void radix_sort_pairs_static(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
							 uint32_t* vals_out, uint64_t n, int key_bits, hipStream_t s) {
	if (n > kStaticSortMaxKeys)
		raise(MARSMD_ERR_VALUE, "radix sort (static workspace): " + std::to_string(n) + " keys exceed the maximum of " +
									std::to_string(kStaticSortMaxKeys));
	static_assert(cx_sort_scratch_bytes(kStaticSortMaxKeys) >= 2 * kStaticSortMaxKeys * 4, "workspace layout");
	void* ws = static_sort_scratch();
	if (radix_sort_scratch_bytes(n) > kStaticSortScratchBytes)
		raise(MARSMD_ERR_RUNTIME, "radix sort (static workspace): layout mismatch");
	radix_sort_pairs(keys_in, vals_in, keys_out, vals_out, n, key_bits, ws, kStaticSortScratchBytes, s);
}

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, void* scratch, hipStream_t s) {
	uint32_t* sums = static_cast<uint32_t*>(scratch);
	if (n == 0) {
		k_fill_u32<<<1, 64, 0, s>>>(out, 0u, 1);
		MMD_LAUNCH_CHECK();
		return;
	}
	const uint64_t blocks = (n + kScanTile - 1) / kScanTile;
	k_scan_reduce<<<unsigned(blocks), kScanBlock, 0, s>>>(in, n, sums);
	MMD_LAUNCH_CHECK();
	k_scan_one_block<<<1, 1024, 0, s>>>(sums, blocks, out + n);
	MMD_LAUNCH_CHECK();
	k_scan_apply<<<unsigned(blocks), kScanBlock, 0, s>>>(in, n, sums, out);
	MMD_LAUNCH_CHECK();
}

void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
	if (n == 0)
		return;
	k_fill_u32<<<grid_for(n, 256), 256, 0, s>>>(p, v, n);
	MMD_LAUNCH_CHECK();
}
void fill_f32(float* p, float v, uint64_t n, hipStream_t s) {
	if (n == 0)
		return;
	k_fill_f32<<<grid_for(n, 256), 256, 0, s>>>(p, v, n);
	MMD_LAUNCH_CHECK();
}

} // namespace mmd
