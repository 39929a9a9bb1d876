// step.hpp — the per-particle integrator updates (BD with the PMF fused, BAOAB) and
// their parameters, shared by the Patch (patch.hip) and the replica batch (batch.hip).
#pragma once

#include "device_common.hpp"

namespace mmd {

struct StepParams {
	DevBox box;
	float cut_d; // pairs farther than this are beyond the cutoff (cutoff_threshold)
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

// The per-type constants of one integrator step, the same expressions whether formed
// per particle or once per type (a kernel may keep them in LDS):
//   BD:    a = D / kT, b = sqrt(2 D dt)                       (BD.h:70-78)
//   BAOAB: a = exp(-gamma dt), b = sqrt(kT m (1 - a^2)) SQ,    (BAOAB.h:85-104)
//          inv_mass = 1 / m
struct TypeStep {
	float a[3], b[3], inv_mass, pad;
};
__device__ __forceinline__ TypeStep bd_type_step(const marsmd_particle_types& types, int ty, const StepParams& S) {
	TypeStep t;
	const float* D = types.diffusion[ty];
#pragma unroll
	for (int k = 0; k < 3; ++k) {
		t.a[k] = D[k] / S.kT;
		t.b[k] = sqrtf(2.0f * D[k] * S.dt);
	}
	t.inv_mass = 0.0f;
	t.pad = 0.0f;
	return t;
}
constexpr float kBaoabSQ = 2.046167337f;
__device__ __forceinline__ TypeStep baoab_type_step(const marsmd_particle_types& types, int ty, const StepParams& S) {
	TypeStep t;
	const float mass = types.mass[ty];
	const float* g = types.trans_damping[ty];
#pragma unroll
	for (int k = 0; k < 3; ++k) {
		t.a[k] = expf(-g[k] * S.dt);
		t.b[k] = sqrtf(S.kT * mass * (1.0f - t.a[k] * t.a[k])) * kBaoabSQ;
	}
	t.inv_mass = 1.0f / mass;
	t.pad = 0.0f;
	return t;
}

// BDIntegrate (BD.h:44-92) given the particle's position-dependent force `pmf`
// (compute_position_dependent_force, Pmf.h:68-124: grid terms and the electric field);
// f is ForceEnergy (xyz force, w energy); ts = bd_type_step of the particle's type.
__device__ __forceinline__ float4 bd_update(float4 p, float4& f, const TypeStep& ts, const float4 pmf,
										   const StepParams& S) {
	const uint32_t m = fb(p.w);
	const float Fx = f.x + pmf.x, Fy = f.y + pmf.y, Fz = f.z + pmf.z;
	f.w += pmf.w;
	const float3 rn = box_muller(philox_first(S.seed, S.ctr + meta_canon(m), S.step, 0x5324120u));
	p.x += ts.a[0] * Fx * S.dt + ts.b[0] * rn.x;
	p.y += ts.a[1] * Fy * S.dt + ts.b[1] * rn.y;
	p.z += ts.a[2] * Fz * S.dt + ts.b[2] * rn.z;
	return wrap_pos(S.box, p);
}
__device__ __forceinline__ float4 bd_update(float4 p, float4& f, const marsmd_particle_types& types, const float4 pmf,
										   const StepParams& S) {
	return bd_update(p, f, bd_type_step(types, int(meta_type(fb(p.w))), S), pmf, S);
}

// Deferred closing kick (BAOAB_LastUpdate, BAOAB.h:151-172) when pending, then
// B-A-O-A (BAOABIntegrate, BAOAB.h:46-120). f already includes the PMF; ts =
// baoab_type_step of the particle's type.
__device__ __forceinline__ float4 baoab_update(float4 p, float4& mo, const float4& f, const TypeStep& ts,
											   const StepParams& S) {
	const uint32_t m = fb(p.w);
	const float FCF = 4.18679994e4f;
	const float hdt = 0.5f * S.dt;
	if (S.deferred) {
		mo.x += hdt * f.x * FCF;
		mo.y += hdt * f.y * FCF;
		mo.z += hdt * f.z * FCF;
	}
	mo.x += hdt * f.x * FCF;
	mo.y += hdt * f.y * FCF;
	mo.z += hdt * f.z * FCF;
	const float inv_mass = ts.inv_mass;
	p.x += hdt * mo.x * inv_mass * 10000.0f;
	p.y += hdt * mo.y * inv_mass * 10000.0f;
	p.z += hdt * mo.z * inv_mass * 10000.0f;
	const float3 rn = box_muller(philox_first(S.seed, S.ctr + meta_canon(m), S.step, 0x1356914u));
	mo.x = ts.a[0] * mo.x + ts.b[0] * rn.x;
	mo.y = ts.a[1] * mo.y + ts.b[1] * rn.y;
	mo.z = ts.a[2] * mo.z + ts.b[2] * rn.z;
	p.x += hdt * mo.x * inv_mass * 10000.0f;
	p.y += hdt * mo.y * inv_mass * 10000.0f;
	p.z += hdt * mo.z * inv_mass * 10000.0f;
	return wrap_pos(S.box, p);
}
__device__ __forceinline__ float4 baoab_update(float4 p, float4& mo, const float4& f,
											   const marsmd_particle_types& types, const StepParams& S) {
	return baoab_update(p, mo, f, baoab_type_step(types, int(meta_type(fb(p.w))), S), S);
}

} // namespace mmd
