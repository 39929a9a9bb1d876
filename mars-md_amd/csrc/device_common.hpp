// device_common.hpp — gfx950 device helpers shared by every kernel of the library.
//
// All floating-point geometry here is written operation for operation like the
// reference functors it cites, and the library is compiled with -ffp-contract=off,
// so keys, pair predicates and per-pair terms are bit-identical to a CPU FP32
// evaluation of the same expressions. Citations are reference file:line.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/marsmd.h"

namespace mmd {

constexpr int kWave = 64;
constexpr uint32_t kInvalidCell = 0xFFFFFFFFu;
// Meta word of a position record: canon << 8 | type. The canonical index sits in the
// high bits, so two particles' meta words order as their canonical indices (one
// compare decides a pair's orientation), and the type is one AND away. The word is
// only ever moved bitwise (as a float it may be any pattern, NaN included).
constexpr uint32_t kCanonBits = 24;
constexpr uint32_t kTypeBits = 8;
constexpr uint32_t kCanonMask = (1u << kCanonBits) - 1u;
constexpr uint64_t kMaxParticles = (1ull << kCanonBits);

__host__ __device__ inline uint32_t make_meta(uint32_t type, uint32_t canon) {
	return ((canon & kCanonMask) << kTypeBits) | (type & ((1u << kTypeBits) - 1u));
}
__host__ __device__ inline uint32_t meta_canon(uint32_t m) { return m >> kTypeBits; }
__host__ __device__ inline uint32_t meta_type(uint32_t m) { return m & ((1u << kTypeBits) - 1u); }

// The pair predicate d * d > cutoff2 (an FP32 product, d >= 0) as d > t: fl(d * d) never
// decreases as d grows, so {d : fl(d * d) > cutoff2} = {d > t} with t the largest float
// whose square rounds to at most cutoff2 (+inf when cutoff2 <= 0: no cutoff). Host code.
inline float cutoff_threshold(float cutoff2) {
	if (!(cutoff2 > 0.0f) || !(cutoff2 < 3.0e38f))
		return __builtin_inff();
	volatile float t = __builtin_sqrtf(cutoff2);
	for (;;) {
		const volatile float sq = t * t;
		if (!(sq > cutoff2))
			break;
		t = __builtin_nextafterf(t, 0.0f);
	}
	for (;;) {
		const volatile float up = __builtin_nextafterf(t, __builtin_inff());
		const volatile float sq = up * up;
		if (!(sq <= cutoff2))
			break;
		t = up;
	}
	return t;
}

// Orthorhombic box as kernel argument; inv_size is 1/L computed once on the host
// (the same correctly rounded value PeriodicBox::wrap_diff_scalar forms each call).
struct DevBox {
	float L[3];
	float inv_L[3];
	float origin[3];
	int periodic[3];
	// wrap_diff without a branch per axis: on an open axis a length of 3e38 with an
	// inverse of 0 (the image is +-0 and the difference is returned unchanged)
	float wrap_L[3], wrap_inv[3], wrap_half[3];
};

inline DevBox make_devbox(const marsmd_box& b) {
	DevBox d;
	for (int a = 0; a < 3; ++a) {
		d.L[a] = b.size[a];
		d.inv_L[a] = b.size[a] > 0.0f ? 1.0f / b.size[a] : 0.0f;
		d.origin[a] = b.origin[a];
		d.periodic[a] = b.periodic[a] != 0 && b.size[a] > 0.0f;
		d.wrap_L[a] = d.periodic[a] ? d.L[a] : 3e38f;
		d.wrap_inv[a] = d.periodic[a] ? d.inv_L[a] : 0.0f;
		d.wrap_half[a] = 0.5f * d.wrap_L[a];
	}
	return d;
}

// PeriodicBox::wrap_diff_scalar (src/System/PeriodicBox.h:453-461).
__device__ __forceinline__ float wrap_diff_scalar(float x, float l, float inv_l) {
	// float(int(floor(y))) == floor(y) for the |y| < 2^31 of any in-box difference
	const float image = floorf(x * inv_l);
	x -= image * l;
	if (x >= 0.5f * l)
		x -= l;
	return x;
}
// PeriodicBox::wrap_scalar (PeriodicBox.h:445-451).
__device__ __forceinline__ float wrap_scalar(float x, float o, float l, float inv_l) {
	const float rel = x - o;
	const int image = int(floorf(rel * inv_l));
	return x - float(image) * l;
}
// wrap_diff(b - a) of the orthorhombic fast path (PeriodicBox.h:134-148).
__device__ __forceinline__ float3 wrap_diff(const DevBox& box, float3 d) {
	if (box.periodic[0])
		d.x = wrap_diff_scalar(d.x, box.L[0], box.inv_L[0]);
	if (box.periodic[1])
		d.y = wrap_diff_scalar(d.y, box.L[1], box.inv_L[1]);
	if (box.periodic[2])
		d.z = wrap_diff_scalar(d.z, box.L[2], box.inv_L[2]);
	return d;
}
// Correctly rounded sqrt(x) for x = 0 and x >= 2^-96: the hardware square root (within
// one ulp) corrected by the signs of the residuals x - s'.s of its two neighbours s', as
// the compiler's lowering of sqrtf does, without that lowering's rescaling of inputs
// below 2^-96 and its zero / infinity select (0 comes out as 0). Equal to sqrtf on every
// float from 2^-96 to 2^24 and on 0, bit for bit (marsmd_selftest_pair_sqrt compares the
// two over any range of bit patterns). Below 2^-96 the result may differ from sqrtf;
// callers use it where such distances (< 4e-15) are skipped whatever their value.
__device__ __forceinline__ float sqrt_rn_nonneg(float x) {
	const float s = __builtin_amdgcn_sqrtf(x);
	const float sm = __uint_as_float(__float_as_uint(s) - 1u);
	const float sp = __uint_as_float(__float_as_uint(s) + 1u);
	float r = s;
	if (fmaf(-sm, s, x) <= 0.0f)
		r = sm;
	if (fmaf(-sp, s, x) > 0.0f)
		r = sp;
	return r;
}

// wrap_diff with the same operations on every axis (see DevBox::wrap_L): on a periodic
// axis they are wrap_diff_scalar's, with x - image * l as one fused multiply-add, which
// is the same value whenever image * l is exact (every |image| <= 2, so every difference
// of two positions within a box length of each other); on an open axis
// x - (+-0) * 3e38 = x, and x never reaches 1.5e38.
__device__ __forceinline__ float wrap_axis(float x, float l, float inv_l, float half_l) {
	const float image = floorf(x * inv_l);
	x = fmaf(-image, l, x);
	return x >= half_l ? x - l : x;
}
__device__ __forceinline__ float3 wrap_diff_uniform(const DevBox& box, float3 d) {
	d.x = wrap_axis(d.x, box.wrap_L[0], box.wrap_inv[0], box.wrap_half[0]);
	d.y = wrap_axis(d.y, box.wrap_L[1], box.wrap_inv[1], box.wrap_half[1]);
	d.z = wrap_axis(d.z, box.wrap_L[2], box.wrap_inv[2], box.wrap_half[2]);
	return d;
}
__device__ __forceinline__ float3 sub3(const float4& a, const float4& b) {
	return make_float3(a.x - b.x, a.y - b.y, a.z - b.z);
}
__device__ __forceinline__ float len2(const float3& v) { return v.x * v.x + v.y * v.y + v.z * v.z; }

// Persistent schedule over tiles (of slots, or of home cells). Workgroups are
// dispatched round-robin over the 8 XCDs, so workgroup b runs on XCD b % 8; the tiles
// are cut into 8 contiguous (Morton-ordered, hence spatially compact) chunks and chunk
// x goes to the workgroups of XCD x, keeping each XCD's L2 on one region of the
// system. gridDim.x is a multiple of 8.
struct TileRange {
	uint32_t tile, end, stride;
};
__device__ __forceinline__ TileRange xcd_tiles(uint32_t tiles) {
	const uint32_t x = blockIdx.x & 7u;
	const uint32_t r = blockIdx.x >> 3;
	const uint32_t cs = uint32_t(uint64_t(tiles) * x / 8u);
	const uint32_t ce = uint32_t(uint64_t(tiles) * (x + 1u) / 8u);
	return TileRange{cs + r, ce, gridDim.x >> 3};
}

// ------------------------------------------------------------------ Morton
// MortonCode (src/PatchOperation/ZOrderKernels/MortonCode.h:95-180).
__host__ __device__ __forceinline__ uint32_t split_by3(uint32_t a) {
	uint32_t x = a & 1023u;
	x = (x | x << 16) & 0x30000ffu;
	x = (x | x << 8) & 0x300f00fu;
	x = (x | x << 4) & 0x30c30c3u;
	x = (x | x << 2) & 0x9249249u;
	return x;
}
__host__ __device__ __forceinline__ uint32_t compact_by3(uint32_t x) {
	x &= 0x9249249u;
	x = (x ^ (x >> 2)) & 0x30c30c3u;
	x = (x ^ (x >> 4)) & 0x300f00fu;
	x = (x ^ (x >> 8)) & 0x30000ffu;
	x = (x ^ (x >> 16)) & 1023u;
	return x;
}
__device__ __forceinline__ uint32_t morton_quantize(float p, float lo, float hi, bool& outside) {
	const float range = fmaxf(hi - lo, 1e-6f);
	const float normalized = (p - lo) / range;
	const float q = normalized * 1023.0f;
	outside = outside || !(q >= 0.0f && q <= 1023.0f);
	return uint32_t(fminf(fmaxf(q, 0.0f), 1023.0f));
}

// ------------------------------------------------------------------ tables
// TabulatedPotential::compute (src/Interactions/TabulatedPotential.h:19-52) for a
// non-periodic or periodic table whose knots sit at `pot` (global or LDS).
struct FE {
	float f, e;
};
__device__ __forceinline__ FE table_compute(float dx, const float* __restrict__ pot, int size,
											float start, float step_inv, bool periodic) {
	float w = (dx - start) * step_inv;
	int home = int(floorf(w));
	w = w - float(home);
	if (periodic) {
		home %= size;
		if (home < 0)
			home += size;
	} else {
		if (home >= size - 1)
			return FE{0.0f, pot[size - 1]};
		if (home < 0)
			home = 0;
	}
	int home1 = periodic ? (home + 1) % size : home + 1;
	if (!periodic && home1 >= size)
		home1 = size - 1;
	const float U0 = pot[home];
	const float dU = pot[home1] - U0;
	return FE{-dU * step_inv, dU * w + U0};
}

// FP32 arccosine of the angle term, from the arcsine Maclaurin series
// asin(t) = t + t z sum_k c_k z^k (z = t^2, c_k = (2k+2)! / (4^(k+1) ((k+1)!)^2 (2k+3))):
// acos(x) = pi/2 - asin(x) for |x| <= 1/2, acos(|x|) = 2 asin(sqrt((1-|x|)/2)) and
// acos(-|x|) = pi - acos(|x|) otherwise. The same expression is the CPU checker's, so
// the angle (and its table bin) is the same bits on both sides.
__host__ __device__ __forceinline__ float acos_series(float x) {
	const float ax = fabsf(x);
	float t, z;
	if (ax <= 0.5f) {
		t = x;
		z = x * x;
	} else {
		z = (1.0f - ax) * 0.5f;
		t = sqrtf(z);
	}
	float p = 7.312525874e-03f;
	p = p * z + 8.390335810e-03f;
	p = p * z + 9.761609529e-03f;
	p = p * z + 1.155180090e-02f;
	p = p * z + 1.396484375e-02f;
	p = p * z + 1.735276442e-02f;
	p = p * z + 2.237215909e-02f;
	p = p * z + 3.038194444e-02f;
	p = p * z + 4.464285714e-02f;
	p = p * z + 7.500000000e-02f;
	p = p * z + 1.666666667e-01f;
	const float as = t + t * z * p;
	if (ax <= 0.5f)
		return 1.57079637f - as;
	const float r = 2.0f * as;
	return x < 0.0f ? 3.14159274f - r : r;
}

// FP32 two-argument arctangent, shared with the CPU checker for the same reason as
// acos_series (the dihedral angle and its table bin are the same bits on both sides).
// With t = min(|x|,|y|) / max(|x|,|y|) in [0, 1]: for t > tan(pi/12) use
// atan(t) = pi/6 + atan((sqrt(3) t - 1) / (t + sqrt(3))), which maps t into
// |t'| <= tan(pi/12); then the Maclaurin series atan(t') = t' - t'^3/3 + ... + t'^13/13
// (next term below 3e-10). Octant and quadrant follow from |y| > |x| and the signs
// (signed zeros as IEEE atan2: atan2(+-0, -0) = +-pi). NaN propagates.
__host__ __device__ __forceinline__ float atan2_series(float y, float x) {
	const float ax = fabsf(x), ay = fabsf(y);
	const float mx = ax > ay ? ax : ay;
	const float mn = ax > ay ? ay : ax;
	float t = mx == 0.0f ? 0.0f : mn / mx;
	float base = 0.0f;
	if (t > 0.267949194f) {
		t = (t * 1.73205081f - 1.0f) / (t + 1.73205081f);
		base = 0.523598776f;
	}
	const float z = t * t;
	float p = 7.69230769e-02f;
	p = p * z - 9.09090909e-02f;
	p = p * z + 1.11111111e-01f;
	p = p * z - 1.42857143e-01f;
	p = p * z + 2.00000000e-01f;
	p = p * z - 3.33333333e-01f;
	float r = base + (t + t * z * p);
	if (ay > ax)
		r = 1.57079637f - r;
	const bool xneg = x < 0.0f || (x == 0.0f && 1.0f / x < 0.0f);
	if (xneg)
		r = 3.14159274f - r;
	const bool yneg = y < 0.0f || (y == 0.0f && 1.0f / y < 0.0f);
	return yneg ? -r : r;
}

// ------------------------------------------------------------------ Philox
// openrand::Philox (src/PatchOperation/Random/philox.h:60-139) and u01
// (base_state.h:299-305).
__device__ __forceinline__ uint4 philox4x32_10(uint32_t k0, uint32_t k1, uint4 c) {
#pragma unroll
	for (int r = 0; r < 10; ++r) {
		if (r > 0) {
			k0 += 0x9E3779B9u;
			k1 += 0xBB67AE85u;
		}
		const uint32_t lo0 = 0xD2511F53u * c.x;
		const uint32_t hi0 = __umulhi(0xD2511F53u, c.x);
		const uint32_t lo1 = 0xCD9E8D57u * c.z;
		const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z);
		c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
	}
	return c;
}
// The first draw of Philox(seed, ctr, global_seed, ctr1): counter {ctr, ctr1, global_seed, 0}.
__device__ __forceinline__ uint4 philox_first(uint64_t seed, uint32_t ctr, uint32_t global_seed,
											  uint32_t ctr1) {
	return philox4x32_10(uint32_t(seed >> 32), uint32_t(seed & 0xFFFFFFFFu),
						 make_uint4(ctr, ctr1, global_seed, 0u));
}
__device__ __forceinline__ float u01(uint32_t x) {
	return float(x) * 2.3283064365386963e-10f + 1.1641532182693481e-10f;
}
// Box-Muller of BDIntegrate / BAOABIntegrate (BD.h:70-78; BAOAB.h:97-104) on the
// hardware's transcendentals. The reference builds with -use_fast_math, so its logf,
// sinf, cosf and sqrtf are approximate intrinsics as well; here v_log_f32 (log2: ln u =
// log2(u) ln 2), v_sin_f32 / v_cos_f32 (the angle in revolutions, which is the uniform
// draw itself: the reference's 2 pi u) and v_sqrt_f32, each within about 1e-6 relative
// of the precise functions, at a tenth of their instructions.
__device__ __forceinline__ float3 box_muller(uint4 r) {
	const float ux = u01(r.x), uy = u01(r.y), uz = u01(r.z), uw = u01(r.w);
	constexpr float kMinus2Ln2 = -1.38629436f; // -2 ln 2
	const float r1 = __builtin_amdgcn_sqrtf(kMinus2Ln2 * __builtin_amdgcn_logf(ux));
	const float r2 = __builtin_amdgcn_sqrtf(kMinus2Ln2 * __builtin_amdgcn_logf(uz));
	return make_float3(r1 * __builtin_amdgcn_cosf(uy), r1 * __builtin_amdgcn_sinf(uy), r2 * __builtin_amdgcn_cosf(uw));
}
__host__ __device__ inline uint64_t mix_seed(uint64_t seed, uint64_t patch_id, uint64_t step) {
	return (seed * 0x9E3779B97F4A7C15ULL) ^ (patch_id << 32) ^ (step * 0xBF58476D1CE4E5B9ULL);
}

// ------------------------------------------------------------------ grids
// Sampling of BaseGridView (src/Types/BaseGridDevice.h:218-475).
__device__ __forceinline__ bool map_grid_index(int& j, int n, int bc) {
	if (j >= 0 && j < n)
		return true;
	if (bc == MARSMD_BC_PERIODIC) {
		j = ((j % n) + n) % n;
		return true;
	}
	if (bc == MARSMD_BC_NEUMANN) {
		j = j < 0 ? 0 : n - 1;
		return true;
	}
	return false;
}
__device__ __forceinline__ void mat_transform(const float* m, const float v[3], float out[3]) {
#pragma unroll
	for (int k = 0; k < 3; ++k)
		out[k] = v[0] * m[0 + k] + v[1] * m[3 + k] + v[2] * m[6 + k];
}
// basis_inv.transpose().transform(g): out[k] = g0*m[3k] + g1*m[3k+1] + g2*m[3k+2].
__device__ __forceinline__ void mat_transpose_transform(const float* m, const float v[3], float out[3]) {
#pragma unroll
	for (int k = 0; k < 3; ++k)
		out[k] = v[0] * m[3 * k + 0] + v[1] * m[3 * k + 1] + v[2] * m[3 * k + 2];
}

struct GridSample {
	float value;
	float grad[3];
};

// Offset of a tap plane/row along one axis (mapped index x stride), or -1 when the
// tap lies outside a Dirichlet grid (its value is then 0).
__device__ __forceinline__ int64_t tap_offset(int j, int n, int bc, int64_t stride) {
	return map_grid_index(j, n, bc) ? int64_t(j) * stride : int64_t(-1);
}

// 4-byte aligned runs of grid values along z: one 8- or 16-byte load each.
struct __attribute__((aligned(4))) GridRun2 {
	float v[2];
};
struct __attribute__((aligned(4))) GridRun4 {
	float v[4];
};

// Grid values live in global memory; the view's pointer comes from memory, so it is
// cast to the global address space (global loads instead of flat ones, whose waits
// would also cover LDS traffic).
// (The host pass of the compile only parses these device functions.)
#if defined(__HIP_DEVICE_COMPILE__)
#define MMD_GLOBAL_AS __attribute__((address_space(1)))
#define MMD_CONST_AS __attribute__((address_space(4)))
#else
#define MMD_GLOBAL_AS
#define MMD_CONST_AS
#endif
typedef MMD_GLOBAL_AS const float gfloat;
__device__ __forceinline__ gfloat* grid_data(const marsmd_grid_view& g) { return (gfloat*)g.data; }

// K consecutive z taps starting at jz0 of the row at `row`: a single load when the run
// lies inside the grid, else tap by tap through the boundary mapping.
template<int K>
__device__ __forceinline__ void load_run(gfloat* __restrict__ row, int jz0, int nz, int bc, float out[K]) {
	if (jz0 >= 0 && jz0 + K <= nz) {
		if constexpr (K == 4) {
			const GridRun4 r = *(MMD_GLOBAL_AS const GridRun4*)(row + jz0);
#pragma unroll
			for (int q = 0; q < 4; ++q)
				out[q] = r.v[q];
		} else {
			const GridRun2 r = *(MMD_GLOBAL_AS const GridRun2*)(row + jz0);
#pragma unroll
			for (int q = 0; q < 2; ++q)
				out[q] = r.v[q];
		}
	} else {
#pragma unroll
		for (int q = 0; q < K; ++q) {
			int jz = jz0 + q;
			out[q] = map_grid_index(jz, nz, bc) ? row[jz] : 0.0f;
		}
	}
}

// sample_grid_linear (BaseGridDevice.h:256-365).
__device__ inline GridSample sample_linear(const marsmd_grid_view& g, float px, float py, float pz, int bc) {
	const float v[3] = {px - g.origin[0], py - g.origin[1], pz - g.origin[2]};
	float gp[3];
	mat_transform(g.basis_inv, v, gp);
	const int i0 = int(floorf(gp[0])), j0 = int(floorf(gp[1])), k0 = int(floorf(gp[2]));
	const float fx = gp[0] - float(i0), fy = gp[1] - float(j0), fz = gp[2] - float(k0);
	const int nx = int(g.dims[0]), ny = int(g.dims[1]), nz = int(g.dims[2]);
	const int64_t sx = int64_t(ny) * int64_t(nz);
	float r[2][2][2]; // [x][y][z]
#pragma unroll
	for (int ix = 0; ix < 2; ++ix) {
		const int64_t ox = tap_offset(i0 + ix, nx, bc, sx);
#pragma unroll
		for (int iy = 0; iy < 2; ++iy) {
			const int64_t oy = tap_offset(j0 + iy, ny, bc, int64_t(nz));
			if (ox < 0 || oy < 0) {
				r[ix][iy][0] = 0.0f;
				r[ix][iy][1] = 0.0f;
			} else {
				load_run<2>(grid_data(g) + ox + oy, k0, nz, bc, r[ix][iy]);
			}
		}
	}
	const float v000 = r[0][0][0], v001 = r[0][0][1];
	const float v010 = r[0][1][0], v011 = r[0][1][1];
	const float v100 = r[1][0][0], v101 = r[1][0][1];
	const float v110 = r[1][1][0], v111 = r[1][1][1];
	const float v00 = v000 * (1.0f - fx) + v100 * fx;
	const float v01 = v001 * (1.0f - fx) + v101 * fx;
	const float v10 = v010 * (1.0f - fx) + v110 * fx;
	const float v11 = v011 * (1.0f - fx) + v111 * fx;
	const float v0 = v00 * (1.0f - fy) + v10 * fy;
	const float v1 = v01 * (1.0f - fy) + v11 * fy;
	GridSample s;
	s.value = v0 * (1.0f - fz) + v1 * fz;
	const float gg[3] = {
		(1.0f - fy) * (1.0f - fz) * (v100 - v000) + fy * (1.0f - fz) * (v110 - v010) +
			(1.0f - fy) * fz * (v101 - v001) + fy * fz * (v111 - v011),
		(1.0f - fx) * (1.0f - fz) * (v010 - v000) + fx * (1.0f - fz) * (v110 - v100) +
			(1.0f - fx) * fz * (v011 - v001) + fx * fz * (v111 - v101),
		(1.0f - fx) * (1.0f - fy) * (v001 - v000) + fx * (1.0f - fy) * (v101 - v100) +
			(1.0f - fx) * fy * (v011 - v010) + fx * fy * (v111 - v110)};
	mat_transpose_transform(g.basis_inv, gg, s.grad);
	return s;
}

// Catmull-Rom interpolation along one axis as weights of the four taps: the cubic
// ((a3 t + a2) t + a1) t + v1 with a3 = (-v0 + 3 v1 - 3 v2 + v3) / 2,
// a2 = (2 v0 - 5 v1 + 4 v2 - v3) / 2, a1 = (v2 - v0) / 2 (BaseGridDevice.h:405-475) is
// sum_i w_i(t) v_i with the weights below, and its t-derivative sum_i w_i'(t) v_i. The
// weights are formed once per axis and each pass is then four multiply-adds per output,
// where the reference's form re-derives the coefficients for every tap line; the sums
// agree with it to rounding.
struct CrWeights {
	float w[4], d[4];
};
__device__ __forceinline__ CrWeights cr_weights(float t) {
	const float t2 = t * t, t3 = t2 * t;
	CrWeights c;
	c.w[0] = 0.5f * (2.0f * t2 - t3 - t);
	c.w[1] = 0.5f * (3.0f * t3 - 5.0f * t2 + 2.0f);
	c.w[2] = 0.5f * (4.0f * t2 - 3.0f * t3 + t);
	c.w[3] = 0.5f * (t3 - t2);
	c.d[0] = 0.5f * (4.0f * t - 3.0f * t2 - 1.0f);
	c.d[1] = 0.5f * (9.0f * t2 - 10.0f * t);
	c.d[2] = 0.5f * (8.0f * t - 9.0f * t2 + 1.0f);
	c.d[3] = 0.5f * (3.0f * t2 - 2.0f * t);
	return c;
}
__device__ __forceinline__ float dot4(const float w[4], float v0, float v1, float v2, float v3) {
	return fmaf(w[3], v3, fmaf(w[2], v2, fmaf(w[1], v1, w[0] * v0)));
}

// The stencil of sample_grid_cubic (BaseGridDevice.h:405-475): home index h =
// floor(grid coordinate) per axis, the fractions w, and the taps h - 1 .. h + 2.
struct CubicStencil {
	int hx, hy, hz;
	float wx, wy, wz;
};
__device__ __forceinline__ CubicStencil cubic_stencil(const marsmd_grid_view& g, float px, float py, float pz) {
	const float v[3] = {px - g.origin[0], py - g.origin[1], pz - g.origin[2]};
	float gp[3];
	mat_transform(g.basis_inv, v, gp);
	CubicStencil c;
	c.hx = int(floorf(gp[0]));
	c.hy = int(floorf(gp[1]));
	c.hz = int(floorf(gp[2]));
	c.wx = gp[0] - float(c.hx);
	c.wy = gp[1] - float(c.hy);
	c.wz = gp[2] - float(c.hz);
	return c;
}

// The 64 taps rr[y][x][z] of a stencil inside a grid of fewer than 2^30 values: 16 runs
// of 4 z taps at 32-bit byte offsets from the grid's (wave-uniform) base, the first run's
// offset per lane and the others a uniform stride from it, so each load is a scalar base
// plus a vector offset and the addresses take a handful of 32-bit adds.
__device__ __forceinline__ void cubic_taps_interior(const marsmd_grid_view& g, const CubicStencil& c,
													float rr[4][4][4]) {
	const uint32_t ny = uint32_t(g.dims[1]), nz = uint32_t(g.dims[2]);
	const uint32_t sy = 4u * nz, sx = sy * ny; // bytes per y row, per x plane
	const uint32_t first = uint32_t(c.hx - 1) * sx + uint32_t(c.hy - 1) * sy + 4u * uint32_t(c.hz - 1);
	MMD_GLOBAL_AS const char* __restrict__ base = (MMD_GLOBAL_AS const char*)g.data;
#pragma unroll
	for (int ix = 0; ix < 4; ++ix) {
		const uint32_t ox = first + uint32_t(ix) * sx;
#pragma unroll
		for (int iy = 0; iy < 4; ++iy) {
			const GridRun4 r = *(MMD_GLOBAL_AS const GridRun4*)(base + (ox + uint32_t(iy) * sy));
#pragma unroll
			for (int q = 0; q < 4; ++q)
				rr[iy][ix][q] = r.v[q];
		}
	}
}

// The same taps of any stencil: each tap axis is mapped once (boundary condition) and
// the four z taps of an (x, y) line are one 16-byte load when they do not cross a grid
// face.
__device__ __forceinline__ void cubic_taps_mapped(const marsmd_grid_view& g, const CubicStencil& c, int bc,
												  float rr[4][4][4]) {
	const int nx = int(g.dims[0]), ny = int(g.dims[1]), nz = int(g.dims[2]);
	const int64_t sx = int64_t(ny) * int64_t(nz);
	int64_t ox[4];
#pragma unroll
	for (int i = 0; i < 4; ++i)
		ox[i] = tap_offset(c.hx - 1 + i, nx, bc, sx);
#pragma unroll
	for (int iy = 0; iy < 4; ++iy) {
		const int64_t oy = tap_offset(c.hy - 1 + iy, ny, bc, int64_t(nz));
#pragma unroll
		for (int ix = 0; ix < 4; ++ix) {
			if (ox[ix] < 0 || oy < 0) {
#pragma unroll
				for (int q = 0; q < 4; ++q)
					rr[iy][ix][q] = 0.0f;
			} else {
				load_run<4>(grid_data(g) + ox[ix] + oy, c.hz - 1, nz, bc, rr[iy][ix]);
			}
		}
	}
}

// sample_grid_cubic (BaseGridDevice.h:405-475): the x pass runs over the 4 x 4 (y, z)
// tap lines, then y, then z, as the reference does. Stencils inside a grid of fewer than
// 2^30 values read their taps at 32-bit offsets, the others through the boundary
// mapping: the same values, so the same bits either way.
__device__ inline GridSample sample_cubic(const marsmd_grid_view& g, float px, float py, float pz, int bc) {
	const CubicStencil c = cubic_stencil(g, px, py, pz);
	const uint64_t nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
	float rr[4][4][4]; // [y][x][z] taps
	if (nx * ny * nz < (1ull << 30) && c.hx >= 1 && int64_t(c.hx) + 2 < int64_t(nx) && c.hy >= 1 &&
		int64_t(c.hy) + 2 < int64_t(ny) && c.hz >= 1 && int64_t(c.hz) + 2 < int64_t(nz))
		cubic_taps_interior(g, c, rr);
	else
		cubic_taps_mapped(g, c, bc, rr);
	const CrWeights Wx = cr_weights(c.wx), Wy = cr_weights(c.wy), Wz = cr_weights(c.wz);
	float d1[4][4], v1[4][4];
#pragma unroll
	for (int iy = 0; iy < 4; ++iy) {
#pragma unroll
		for (int iz = 0; iz < 4; ++iz) {
			d1[iy][iz] = dot4(Wx.d, rr[iy][0][iz], rr[iy][1][iz], rr[iy][2][iz], rr[iy][3][iz]);
			v1[iy][iz] = dot4(Wx.w, rr[iy][0][iz], rr[iy][1][iz], rr[iy][2][iz], rr[iy][3][iz]);
		}
	}
	float d2x[4], d2y[4], v2[4];
#pragma unroll
	for (int iz = 0; iz < 4; ++iz) {
		d2x[iz] = dot4(Wy.w, d1[0][iz], d1[1][iz], d1[2][iz], d1[3][iz]);
		d2y[iz] = dot4(Wy.d, v1[0][iz], v1[1][iz], v1[2][iz], v1[3][iz]);
		v2[iz] = dot4(Wy.w, v1[0][iz], v1[1][iz], v1[2][iz], v1[3][iz]);
	}
	const float gg[3] = {dot4(Wz.w, d2x[0], d2x[1], d2x[2], d2x[3]), dot4(Wz.w, d2y[0], d2y[1], d2y[2], d2y[3]),
						 dot4(Wz.d, v2[0], v2[1], v2[2], v2[3])};
	GridSample s;
	s.value = dot4(Wz.w, v2[0], v2[1], v2[2], v2[3]);
	mat_transpose_transform(g.basis_inv, gg, s.grad);
	return s;
}

// Runs body(grid view) for the active lanes whose grid index is gid, one distinct
// index at a time: the index is wave-uniform inside, so the view is read with scalar
// loads instead of occupying vector registers in every lane. Lanes with gid < 0 skip.
template<class Body>
__device__ __forceinline__ void for_each_grid(int gid, const marsmd_grid_view* __restrict__ grids, Body&& body) {
	uint64_t pending = __ballot(gid >= 0);
	while (pending) {
		const int lane = __ffsll((unsigned long long)pending) - 1;
		const int ug = __builtin_amdgcn_readlane(gid, lane);
		if (gid == ug) {
			// the view through a wave-uniform pointer in the constant address space, so it
			// is read with scalar loads into scalar registers (a vector load of the same
			// record per lane puts a memory round trip and 32 VGPRs before the sampler)
			const uint64_t a = reinterpret_cast<uint64_t>(grids + ug);
			const uint64_t ua = uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(a))))) |
								(uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(a >> 32))))) << 32);
			const marsmd_grid_view g = *reinterpret_cast<MMD_CONST_AS const marsmd_grid_view*>(ua);
			body(g);
		}
		pending &= ~__ballot(gid == ug);
	}
}

// compute_position_dependent_force (src/Interactions/Nonbonded/Pmf.h:68-124) for one
// interpolation scheme. The returned .w is the PMF energy (the reference's .t). Each
// lane adds its type's terms in term order, as the reference does.
template<bool CUBIC>
__device__ inline float4 position_dependent_force_t(float px, float py, float pz, int type,
													const marsmd_particle_types& types,
													const marsmd_grid_view* __restrict__ grids, const float ef[3]) {
	const float charge = types.charge[type];
	float4 f = make_float4(charge * ef[0], charge * ef[1], charge * ef[2], 0.0f);
	if (grids == nullptr)
		return f;
	int max_terms = 0;
	for (int y = 0; y < types.num_types; ++y)
		max_terms = max(max_terms, types.pmf_grid_count[y]);
	const int b = types.pmf_grid_offset[type];
	const int cnt = types.pmf_grid_count[type];
#pragma unroll 1
	for (int k = 0; k < max_terms; ++k) {
		marsmd_grid_term term{-1, 0.0f, 0.0f, -1};
		if (k < cnt)
			term = types.pmf_grid_terms[b + k];
		for_each_grid(term.grid_id, grids, [&](const marsmd_grid_view& g) {
			if (g.data == nullptr)
				return;
			const int bc = term.boundary_condition >= 0 ? term.boundary_condition : g.boundary_condition;
			const GridSample s = CUBIC ? sample_cubic(g, px, py, pz, bc) : sample_linear(g, px, py, pz, bc);
			const float ns = -term.scale;
			f.x += s.grad[0] * ns;
			f.y += s.grad[1] * ns;
			f.z += s.grad[2] * ns;
			f.w += term.scale * s.value;
		});
	}
#pragma unroll 1
	for (int a = 0; a < 3; ++a) {
		for_each_grid(types.force_grid_id[type][a], grids, [&](const marsmd_grid_view& g) {
			if (g.data == nullptr)
				return;
			const float v = CUBIC ? sample_cubic(g, px, py, pz, g.boundary_condition).value
								  : sample_linear(g, px, py, pz, g.boundary_condition).value;
			const float add = types.force_grid_scale[type][a] * v;
			if (a == 0)
				f.x += add;
			else if (a == 1)
				f.y += add;
			else
				f.z += add;
		});
	}
	return f;
}

__device__ inline float4 position_dependent_force(float px, float py, float pz, int type,
												  const marsmd_particle_types& types,
												  const marsmd_grid_view* __restrict__ grids, const float ef[3],
												  int scheme) {
	return scheme == MARSMD_INTERP_LINEAR
			   ? position_dependent_force_t<false>(px, py, pz, type, types, grids, ef)
			   : position_dependent_force_t<true>(px, py, pz, type, types, grids, ef);
}

__device__ __forceinline__ float atomic_add_f32(float* p, float v) {
	return unsafeAtomicAdd(p, v);
}

} // namespace mmd
