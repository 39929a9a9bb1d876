// step.hpp — the per-particle integrator updates (BD with the PMF fused, BAOAB) and
// their parameters, shared by the Patch (patch.hip) and the replica batch (batch.hip).
#pragma once

#include "device_common.hpp"

namespace mmd {

struct StepParams {
	DevBox box;
	float cutoff2;
	float dt;
	float kT;
	uint64_t seed;
	uint32_t ctr;
	uint32_t step;
	int scheme;
	int has_grids;
	int deferred;
	float ef[3];
};

__device__ __forceinline__ float4 wrap_pos(const DevBox& box, float4 p) {
	if (box.periodic[0])
		p.x = wrap_scalar(p.x, box.origin[0], box.L[0], box.inv_L[0]);
	if (box.periodic[1])
		p.y = wrap_scalar(p.y, box.origin[1], box.L[1], box.inv_L[1]);
	if (box.periodic[2])
		p.z = wrap_scalar(p.z, box.origin[2], box.L[2], box.inv_L[2]);
	return p;
}

// BDIntegrate (BD.h:44-92) given the particle's position-dependent force `pmf`
// (compute_position_dependent_force, Pmf.h:68-124: grid terms and the electric field);
// f is ForceEnergy (xyz force, w energy).
__device__ __forceinline__ float4 bd_update(float4 p, float4& f, const marsmd_particle_types& types, const float4 pmf,
										   const StepParams& S) {
	const uint32_t m = fb(p.w);
	const int ty = int(meta_type(m));
	const float Fx = f.x + pmf.x, Fy = f.y + pmf.y, Fz = f.z + pmf.z;
	f.w += pmf.w;
	const float* D = types.diffusion[ty];
	const float3 rn = box_muller(philox_first(S.seed, S.ctr + meta_canon(m), S.step, 0x5324120u));
	p.x += (D[0] / S.kT) * Fx * S.dt + sqrtf(2.0f * D[0] * S.dt) * rn.x;
	p.y += (D[1] / S.kT) * Fy * S.dt + sqrtf(2.0f * D[1] * S.dt) * rn.y;
	p.z += (D[2] / S.kT) * Fz * S.dt + sqrtf(2.0f * D[2] * S.dt) * rn.z;
	return wrap_pos(S.box, p);
}

// Deferred closing kick (BAOAB_LastUpdate, BAOAB.h:151-172) when pending, then
// B-A-O-A (BAOABIntegrate, BAOAB.h:46-120). f already includes the PMF.
__device__ __forceinline__ float4 baoab_update(float4 p, float4& mo, const float4& f,
											   const marsmd_particle_types& types, const StepParams& S) {
	const uint32_t m = fb(p.w);
	const int ty = int(meta_type(m));
	const float FCF = 4.18679994e4f;
	const float SQ = 2.046167337f;
	const float hdt = 0.5f * S.dt;
	if (S.deferred) {
		mo.x += hdt * f.x * FCF;
		mo.y += hdt * f.y * FCF;
		mo.z += hdt * f.z * FCF;
	}
	const float mass = types.mass[ty];
	const float* g = types.trans_damping[ty];
	mo.x += hdt * f.x * FCF;
	mo.y += hdt * f.y * FCF;
	mo.z += hdt * f.z * FCF;
	const float inv_mass = 1.0f / mass;
	p.x += hdt * mo.x * inv_mass * 10000.0f;
	p.y += hdt * mo.y * inv_mass * 10000.0f;
	p.z += hdt * mo.z * inv_mass * 10000.0f;
	const float c0 = expf(-g[0] * S.dt), c1 = expf(-g[1] * S.dt), c2 = expf(-g[2] * S.dt);
	const float n0 = sqrtf(S.kT * mass * (1.0f - c0 * c0)) * SQ;
	const float n1 = sqrtf(S.kT * mass * (1.0f - c1 * c1)) * SQ;
	const float n2 = sqrtf(S.kT * mass * (1.0f - c2 * c2)) * SQ;
	const float3 rn = box_muller(philox_first(S.seed, S.ctr + meta_canon(m), S.step, 0x1356914u));
	mo.x = c0 * mo.x + n0 * rn.x;
	mo.y = c1 * mo.y + n1 * rn.y;
	mo.z = c2 * mo.z + n2 * rn.z;
	p.x += hdt * mo.x * inv_mass * 10000.0f;
	p.y += hdt * mo.y * inv_mass * 10000.0f;
	p.z += hdt * mo.z * inv_mass * 10000.0f;
	return wrap_pos(S.box, p);
}

} // namespace mmd
