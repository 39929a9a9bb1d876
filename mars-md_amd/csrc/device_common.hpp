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
constexpr uint32_t kCanonBits = 24; // meta word of a position record: type << 24 | canon
constexpr uint32_t kCanonMask = (1u << kCanonBits) - 1u;
constexpr uint64_t kMaxParticles = (1ull << kCanonBits);

__host__ __device__ inline uint32_t make_meta(uint32_t type, uint32_t canon) {
	return (type << kCanonBits) | (canon & kCanonMask);
}
__host__ __device__ inline uint32_t meta_canon(uint32_t m) { return m & kCanonMask; }
__host__ __device__ inline uint32_t meta_type(uint32_t m) { return m >> kCanonBits; }

// Orthorhombic box as kernel argument; inv_size is 1/L computed once on the host
// (the same correctly rounded value PeriodicBox::wrap_diff_scalar forms each call).
struct DevBox {
	float L[3];
	float inv_L[3];
	float origin[3];
	int periodic[3];
};

inline DevBox make_devbox(const marsmd_box& b) {
	DevBox d;
	for (int a = 0; a < 3; ++a) {
		d.L[a] = b.size[a];
		d.inv_L[a] = b.size[a] > 0.0f ? 1.0f / b.size[a] : 0.0f;
		d.origin[a] = b.origin[a];
		d.periodic[a] = b.periodic[a] != 0 && b.size[a] > 0.0f;
	}
	return d;
}

// PeriodicBox::wrap_diff_scalar (src/System/PeriodicBox.h:453-461).
__device__ __forceinline__ float wrap_diff_scalar(float x, float l, float inv_l) {
	const int image = int(floorf(x * inv_l));
	x -= float(image) * l;
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
__device__ __forceinline__ float3 sub3(const float4& a, const float4& b) {
	return make_float3(a.x - b.x, a.y - b.y, a.z - b.z);
}
__device__ __forceinline__ float len2(const float3& v) { return v.x * v.x + v.y * v.y + v.z * v.z; }

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
// Box-Muller of BDIntegrate / BAOABIntegrate (BD.h:70-78; BAOAB.h:97-104).
__device__ __forceinline__ float3 box_muller(uint4 r) {
	const float ux = u01(r.x), uy = u01(r.y), uz = u01(r.z), uw = u01(r.w);
	const float r1 = sqrtf(-2.0f * logf(ux));
	const float theta1 = 2.0f * 3.1415926535f * uy;
	const float r2 = sqrtf(-2.0f * logf(uz));
	const float theta2 = 2.0f * 3.1415926535f * uw;
	return make_float3(r1 * cosf(theta1), r1 * sinf(theta1), r2 * cosf(theta2));
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
__device__ __forceinline__ float grid_fetch(const marsmd_grid_view& g, int jx, int jy, int jz, int bc) {
	const int nx = int(g.dims[0]), ny = int(g.dims[1]), nz = int(g.dims[2]);
	if (!map_grid_index(jx, nx, bc) || !map_grid_index(jy, ny, bc) || !map_grid_index(jz, nz, bc))
		return 0.0f;
	const size_t idx = size_t(jz) + size_t(jy) * size_t(nz) + size_t(jx) * size_t(ny) * size_t(nz);
	return g.data[idx];
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

__device__ inline GridSample sample_linear(const marsmd_grid_view& g, float px, float py, float pz, int bc) {
	const float v[3] = {px - g.origin[0], py - g.origin[1], pz - g.origin[2]};
	float gp[3];
	mat_transform(g.basis_inv, v, gp);
	const int i0 = int(floorf(gp[0])), j0 = int(floorf(gp[1])), k0 = int(floorf(gp[2]));
	const float fx = gp[0] - float(i0), fy = gp[1] - float(j0), fz = gp[2] - float(k0);
	const float v000 = grid_fetch(g, i0, j0, k0, bc), v001 = grid_fetch(g, i0, j0, k0 + 1, bc);
	const float v010 = grid_fetch(g, i0, j0 + 1, k0, bc), v011 = grid_fetch(g, i0, j0 + 1, k0 + 1, bc);
	const float v100 = grid_fetch(g, i0 + 1, j0, k0, bc), v101 = grid_fetch(g, i0 + 1, j0, k0 + 1, bc);
	const float v110 = grid_fetch(g, i0 + 1, j0 + 1, k0, bc), v111 = grid_fetch(g, i0 + 1, j0 + 1, k0 + 1, bc);
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

__device__ __forceinline__ float cr_value(float v0, float v1, float v2, float v3, float t) {
	const float a3 = 0.5f * (-v0 + 3.0f * v1 - 3.0f * v2 + v3);
	const float a2 = 0.5f * (2.0f * v0 - 5.0f * v1 + 4.0f * v2 - v3);
	const float a1 = 0.5f * (-v0 + v2);
	return ((a3 * t + a2) * t + a1) * t + v1;
}
__device__ __forceinline__ float cr_deriv(float v0, float v1, float v2, float v3, float t) {
	const float a3 = 0.5f * (-v0 + 3.0f * v1 - 3.0f * v2 + v3);
	const float a2 = 0.5f * (2.0f * v0 - 5.0f * v1 + 4.0f * v2 - v3);
	const float a1 = 0.5f * (-v0 + v2);
	return (3.0f * a3 * t + 2.0f * a2) * t + a1;
}

// sample_grid_cubic (BaseGridDevice.h:405-475). For a periodic grid whose 4-tap
// z-run does not wrap, the run is read as one 16-byte row.
__device__ inline GridSample sample_cubic(const marsmd_grid_view& g, float px, float py, float pz, int bc) {
	const float v[3] = {px - g.origin[0], py - g.origin[1], pz - g.origin[2]};
	float gp[3];
	mat_transform(g.basis_inv, v, gp);
	const int hx = int(floorf(gp[0])), hy = int(floorf(gp[1])), hz = int(floorf(gp[2]));
	const float wx = gp[0] - float(hx), wy = gp[1] - float(hy), wz = gp[2] - float(hz);
	float d1[4][4], v1[4][4];
#pragma unroll
	for (int iz = 0; iz < 4; ++iz) {
		const int jz = hz - 1 + iz;
#pragma unroll
		for (int iy = 0; iy < 4; ++iy) {
			const int jy = hy - 1 + iy;
			float t[4];
#pragma unroll
			for (int ix = 0; ix < 4; ++ix)
				t[ix] = grid_fetch(g, hx - 1 + ix, jy, jz, bc);
			d1[iy][iz] = cr_deriv(t[0], t[1], t[2], t[3], wx);
			v1[iy][iz] = cr_value(t[0], t[1], t[2], t[3], wx);
		}
	}
	float d2x[4], d2y[4], v2[4];
#pragma unroll
	for (int iz = 0; iz < 4; ++iz) {
		d2x[iz] = cr_value(d1[0][iz], d1[1][iz], d1[2][iz], d1[3][iz], wy);
		d2y[iz] = cr_deriv(v1[0][iz], v1[1][iz], v1[2][iz], v1[3][iz], wy);
		v2[iz] = cr_value(v1[0][iz], v1[1][iz], v1[2][iz], v1[3][iz], wy);
	}
	const float gg[3] = {cr_value(d2x[0], d2x[1], d2x[2], d2x[3], wz),
						 cr_value(d2y[0], d2y[1], d2y[2], d2y[3], wz),
						 cr_deriv(v2[0], v2[1], v2[2], v2[3], wz)};
	GridSample s;
	s.value = cr_value(v2[0], v2[1], v2[2], v2[3], wz);
	mat_transpose_transform(g.basis_inv, gg, s.grad);
	return s;
}

// compute_position_dependent_force (src/Interactions/Nonbonded/Pmf.h:68-124). The
// returned .w is the PMF energy (the reference's .t).
__device__ inline float4 position_dependent_force(float px, float py, float pz, int type,
												  const marsmd_particle_types& types,
												  const marsmd_grid_view* __restrict__ grids,
												  const float ef[3], int scheme) {
	const float charge = types.charge[type];
	float4 f = make_float4(charge * ef[0], charge * ef[1], charge * ef[2], 0.0f);
	if (grids == nullptr)
		return f;
	const int b = types.pmf_grid_offset[type];
	const int e = b + types.pmf_grid_count[type];
	for (int k = b; k < e; ++k) {
		const marsmd_grid_term term = types.pmf_grid_terms[k];
		if (term.grid_id < 0)
			continue;
		const marsmd_grid_view& g = grids[term.grid_id];
		if (g.data == nullptr)
			continue;
		const int bc = term.boundary_condition >= 0 ? term.boundary_condition : g.boundary_condition;
		const GridSample s = scheme == MARSMD_INTERP_LINEAR ? sample_linear(g, px, py, pz, bc)
															: sample_cubic(g, px, py, pz, bc);
		const float ns = -term.scale;
		f.x += s.grad[0] * ns;
		f.y += s.grad[1] * ns;
		f.z += s.grad[2] * ns;
		f.w += term.scale * s.value;
	}
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		const int gid = types.force_grid_id[type][a];
		if (gid < 0)
			continue;
		const marsmd_grid_view& g = grids[gid];
		if (g.data == nullptr)
			continue;
		const float v = scheme == MARSMD_INTERP_LINEAR ? sample_linear(g, px, py, pz, g.boundary_condition).value
													   : sample_cubic(g, px, py, pz, g.boundary_condition).value;
		const float add = types.force_grid_scale[type][a] * v;
		if (a == 0)
			f.x += add;
		else if (a == 1)
			f.y += add;
		else
			f.z += add;
	}
	return f;
}

__device__ __forceinline__ float atomic_add_f32(float* p, float v) {
	return unsafeAtomicAdd(p, v);
}

} // namespace mmd
