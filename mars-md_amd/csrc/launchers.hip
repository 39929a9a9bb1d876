// launchers.hip — the reference's free launch functions over caller-owned arrays in
// the reference layout (float4 Vector3 records, canonical order):
// launch_resolve_pair_tables, launch_pairwise_nonbonded, launch_tabulated_bonds,
// launch_tabulated_angles, launch_BD, launch_BAOAB, launch_BAOAB_LastUpdate,
// launch_PMF. These keep the reference's one-thread-per-item shape and its atomic
// scatter, because callers own the pair/bond arrays; the Patch path (patch.hip)
// replaces them with the fused, atomic-free per-particle kernel.
#include "launchers.hpp"

#include "analytic.hpp"

namespace mmd {

namespace {

// ResolvePairTableKernel (Pairwise.h:104-120).
__global__ void k_resolve(const int2* __restrict__ pairs, const int32_t* __restrict__ types,
						  const uint32_t* __restrict__ matrix, uint64_t nt, uint32_t* __restrict__ tag,
						  uint64_t np) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= np)
		return;
	const int2 p = pairs[i];
	tag[i] = matrix[uint64_t(types[p.x]) * nt + uint64_t(types[p.y])];
}

__device__ __forceinline__ void atomic_add4(float4* p, float x, float y, float z) {
	float* f = reinterpret_cast<float*>(p);
	atomic_add_f32(f + 0, x);
	atomic_add_f32(f + 1, y);
	atomic_add_f32(f + 2, z);
}

// One particle's share of a bonded term: added to fe[atom] with FP32 atomics (the
// reference's sum, in no fixed order), or, with contrib, stored at its contribution
// slot for a gather that sums each particle's shares in a fixed order.
__device__ __forceinline__ void bonded_emit(float4* fe, float4* contrib, uint64_t slot, int atom, float x, float y,
											float z, bool energy, float e) {
	if (contrib) {
		contrib[slot] = make_float4(x, y, z, energy ? e : 0.0f);
		return;
	}
	atomic_add4(&fe[atom], x, y, z);
	if (energy)
		atomic_add_f32(reinterpret_cast<float*>(&fe[atom]) + 3, e);
}
// The shares of a term that adds nothing (skipped form, zero length, no particle).
__device__ __forceinline__ void bonded_zero(float4* contrib, uint64_t first, int roles) {
	if (contrib)
		for (int r = 0; r < roles; ++r)
			contrib[first + uint64_t(r)] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// PairNonbondedComputer::operator() (Pairwise.h:214-304) with ANALYTIC / TABULATED
// selecting its branches; AnalyticalPairKernel::operator() (AnalyticalPairKernels.h:
// 44-104) is the ANALYTIC-only instance. The pair's own (x, y) orientation is kept.
template<bool ANALYTIC, bool TABULATED>
__global__ void k_pairwise(const int2* __restrict__ pairs, const float4* __restrict__ pos, float4* fe,
						   const int32_t* __restrict__ type_ids, const uint32_t* __restrict__ tags,
						   const marsmd_table* __restrict__ tables, PairCoeffs C, DevBox box, int get_energy,
						   uint64_t np, float cutoff2, uint32_t enabled) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= np)
		return;
	const uint32_t tag = tags ? tags[i] : enabled;
	const uint32_t terms = tag & enabled & (TABULATED ? MARSMD_PAIR_TERM_MASK : kAnalyticTerms);
	if (terms == 0u)
		return;
	const int2 ij = pairs[i];
	const float3 r = wrap_diff(box, sub3(pos[ij.y], pos[ij.x]));
	const float r2 = len2(r);
	const float d = sqrtf(r2);
	float3 u = make_float3(0.0f, 0.0f, 0.0f);
	if (d > 1e-6f) {
		const float inv = 1.0f / d;
		u = make_float3(inv * r.x, inv * r.y, inv * r.z);
	}
	if (d < 1e-6f)
		return;
	if (cutoff2 > 0.0f && d * d > cutoff2)
		return;
	float fm = 0.0f, en = 0.0f;
	if (ANALYTIC && (terms & kAnalyticTerms)) {
		const int ti = type_ids[ij.x], tj = type_ids[ij.y];
		analytic_terms(terms, d, r2, C.charge[ti], C.charge[tj], C.radius[ti], C.radius[tj], C.eps[ti], C.eps[tj],
					   C.prm, fm, en);
	}
	if (TABULATED && (terms & MARSMD_PAIR_TERM_TABULATED)) {
		const marsmd_table t = tables[tag >> MARSMD_PAIR_TABLE_SHIFT];
		const FE v = table_compute(d, t.pot, int(t.size), t.start, t.step_inv, t.is_periodic != 0);
		fm += v.f;
		en += v.e;
	}
	const float fx = fm * u.x, fy = fm * u.y, fz = fm * u.z;
	atomic_add4(&fe[ij.x], -fx, -fy, -fz);
	atomic_add4(&fe[ij.y], fx, fy, fz);
	if (get_energy) {
		const float h = en * 0.5f;
		atomic_add_f32(reinterpret_cast<float*>(&fe[ij.x]) + 3, h);
		atomic_add_f32(reinterpret_cast<float*>(&fe[ij.y]) + 3, h);
	}
}

// TabulatedBondComputer (BondComputer.h:163-191).
__global__ void k_bonds(const int2* __restrict__ bonds, const float4* __restrict__ pos, float4* fe,
						const marsmd_table* __restrict__ tables, const int32_t* __restrict__ tidx,
						const int32_t* __restrict__ forms, DevBox box, int get_energy, uint64_t nb,
						float4* __restrict__ contrib) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= nb)
		return;
	if (forms && forms[i] != MARSMD_FORM_TABULATED) {
		bonded_zero(contrib, 2 * i, 2);
		return;
	}
	const int2 ij = bonds[i];
	const float3 r = wrap_diff(box, sub3(pos[ij.y], pos[ij.x]));
	const float d = sqrtf(len2(r));
	float3 u = make_float3(0.0f, 0.0f, 0.0f);
	if (d > 1e-6f) {
		const float inv = 1.0f / d;
		u = make_float3(inv * r.x, inv * r.y, inv * r.z);
	}
	if (d < 1e-6f) {
		bonded_zero(contrib, 2 * i, 2);
		return;
	}
	const marsmd_table t = tables[tidx ? tidx[i] : 0];
	const FE v = table_compute(d, t.pot, int(t.size), t.start, t.step_inv, t.is_periodic != 0);
	const float fx = v.f * u.x, fy = v.f * u.y, fz = v.f * u.z;
	const float h = v.e * 0.5f;
	bonded_emit(fe, contrib, 2 * i, ij.x, -fx, -fy, -fz, get_energy != 0, h);
	bonded_emit(fe, contrib, 2 * i + 1, ij.y, fx, fy, fz, get_energy != 0, h);
}

// AngleGeometry + TabulatedAngleComputer (BondGeometry.h:13-45; BondComputer.h:231-268).
__device__ __forceinline__ void angle_forces(const DevBox& box, const float4& pa, const float4& pb,
											 const float4& pc, const marsmd_table& t, float3& f1, float3& f3,
											 float& third) {
	const float3 ab = wrap_diff(box, sub3(pb, pa));
	const float3 bc = wrap_diff(box, sub3(pc, pb));
	const float3 ac = wrap_diff(box, sub3(pc, pa));
	const float ab2 = len2(ab), bc2 = len2(bc), ac2 = len2(ac);
	float c = (ab2 + bc2 - ac2) * 0.5f / (sqrtf(ab2) * sqrtf(bc2));
	if (c < -1.0f)
		c = -1.0f;
	if (c > 1.0f)
		c = 1.0f;
	const float angle = acos_series(c);
	const float sn = sqrtf(1.0f - c * c);
	const FE v = table_compute(angle, t.pot, int(t.size), t.start, t.step_inv, t.is_periodic != 0);
	const float s = sn > 1e-3f ? sn : 1e-3f;
	const float dUdtheta = -v.f / s;
	const float inv_ab = 1.0f / sqrtf(len2(ab));
	const float inv_bc = 1.0f / sqrtf(len2(bc));
	const float s1 = dUdtheta * inv_ab;
	const float t1 = c * inv_ab;
	f1 = make_float3(s1 * (t1 * ab.x + inv_bc * bc.x), s1 * (t1 * ab.y + inv_bc * bc.y),
					 s1 * (t1 * ab.z + inv_bc * bc.z));
	const float s3 = -(dUdtheta * inv_bc);
	const float t3 = c * inv_bc;
	f3 = make_float3(s3 * (t3 * bc.x + inv_ab * ab.x), s3 * (t3 * bc.y + inv_ab * ab.y),
					 s3 * (t3 * bc.z + inv_ab * ab.z));
	third = v.e * float(1.0 / 3.0);
}

__global__ void k_angles(const int4* __restrict__ angles, const float4* __restrict__ pos, float4* fe,
						 const marsmd_table* __restrict__ tables, const int32_t* __restrict__ tidx,
						 const int32_t* __restrict__ forms, DevBox box, int get_energy, uint64_t na,
						 float4* __restrict__ contrib) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= na)
		return;
	if (forms && forms[i] != MARSMD_FORM_TABULATED) {
		bonded_zero(contrib, 3 * i, 3);
		return;
	}
	const int4 a = angles[i];
	float3 f1, f3;
	float third;
	angle_forces(box, pos[a.x], pos[a.y], pos[a.z], tables[tidx ? tidx[i] : 0], f1, f3, third);
	const bool e = get_energy != 0;
	bonded_emit(fe, contrib, 3 * i, a.x, f1.x, f1.y, f1.z, e, third);
	bonded_emit(fe, contrib, 3 * i + 1, a.y, -(f1.x + f3.x), -(f1.y + f3.y), -(f1.z + f3.z), e, third);
	bonded_emit(fe, contrib, 3 * i + 2, a.z, f3.x, f3.y, f3.z, e, third);
}

__device__ __forceinline__ float3 cross3(const float3& v, const float3& w) {
	return make_float3(v.y * w.z - v.z * w.y, v.z * w.x - v.x * w.z, v.x * w.y - v.y * w.x);
}
__device__ __forceinline__ float dot3(const float3& v, const float3& w) { return v.x * w.x + v.y * w.y + v.z * w.z; }
// Vector3::rLength2 (src/Types/Vector3.h:233-236).
__device__ __forceinline__ float rlen2(const float3& v) {
	const float l2 = len2(v);
	return l2 != 0.0f ? 1.0f / l2 : 0.0f;
}

// DihedralGeometry (src/Interactions/Bonded/BondGeometry.h:56-87) and
// TabulatedDihedralComputer (BondComputer.h:296-343): forces on the four particles
// (f1, f2 - f1, f3 - f2, -f3, each scaled by the clamped table force) and U/4.
__device__ __forceinline__ void dihedral_forces(const DevBox& box, const float4& p0, const float4& p1,
												const float4& p2, const float4& p3, const marsmd_table& t,
												float3 out[4], float& quarter) {
	const float3 ab = wrap_diff(box, sub3(p1, p0));
	const float3 bc = wrap_diff(box, sub3(p2, p1));
	const float3 cd = wrap_diff(box, sub3(p3, p2));
	const float3 c_abc = cross3(ab, bc);
	const float3 c_bcd = cross3(bc, cd);
	const float3 c_x = cross3(bc, c_abc);
	const float cos_phi = dot3(c_abc, c_bcd) / (sqrtf(len2(c_abc)) * sqrtf(len2(c_bcd)));
	const float sin_phi = dot3(c_x, c_bcd) / (sqrtf(len2(c_x)) * sqrtf(len2(c_bcd)));
	const float phi = atan2_series(sin_phi, cos_phi);
	const float bcl = sqrtf(len2(bc));
	const float r_abc = rlen2(c_abc), r_bcd = rlen2(c_bcd), r_bc = rlen2(bc);
	const float s1 = -bcl * r_abc;
	const float s3 = -bcl * r_bcd;
	const float3 f1 = make_float3(s1 * c_abc.x, s1 * c_abc.y, s1 * c_abc.z);
	const float3 f3 = make_float3(s3 * c_bcd.x, s3 * c_bcd.y, s3 * c_bcd.z);
	const float a2 = -(dot3(ab, bc) * r_bc);
	const float b2 = dot3(bc, cd) * r_bc;
	const float3 f2 = make_float3(a2 * f1.x - b2 * f3.x, a2 * f1.y - b2 * f3.y, a2 * f1.z - b2 * f3.z);
	const bool degenerate = (len2(ab) * len2(bc) * r_abc > 100.0f) || (len2(bc) * len2(cd) * r_bcd > 100.0f);
	const FE v = table_compute(phi, t.pot, int(t.size), t.start, t.step_inv, t.is_periodic != 0);
	float mag = degenerate ? 0.0f : v.f;
	if (mag > 1000.0f)
		mag = 1000.0f;
	else if (mag < -1000.0f)
		mag = -1000.0f;
	const float3 g1 = make_float3(f1.x * mag, f1.y * mag, f1.z * mag);
	const float3 g2 = make_float3(f2.x * mag, f2.y * mag, f2.z * mag);
	const float3 g3 = make_float3(f3.x * mag, f3.y * mag, f3.z * mag);
	out[0] = g1;
	out[1] = make_float3(g2.x - g1.x, g2.y - g1.y, g2.z - g1.z);
	out[2] = make_float3(g3.x - g2.x, g3.y - g2.y, g3.z - g2.z);
	out[3] = make_float3(-g3.x, -g3.y, -g3.z);
	quarter = v.e * 0.25f;
}

__global__ void k_dihedrals(const int4* __restrict__ dih, const float4* __restrict__ pos, float4* fe,
							const marsmd_table* __restrict__ tables, const int32_t* __restrict__ tidx,
							const int32_t* __restrict__ forms, DevBox box, int get_energy, uint64_t nd,
							float4* __restrict__ contrib) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= nd)
		return;
	if (forms && forms[i] != MARSMD_FORM_TABULATED) {
		bonded_zero(contrib, 4 * i, 4);
		return;
	}
	const int4 q = dih[i];
	float3 f[4];
	float quarter;
	dihedral_forces(box, pos[q.x], pos[q.y], pos[q.z], pos[q.w], tables[tidx ? tidx[i] : 0], f, quarter);
	const int idx[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
	for (int k = 0; k < 4; ++k)
		bonded_emit(fe, contrib, 4 * i + uint64_t(k), idx[k], f[k].x, f[k].y, f[k].z, get_energy != 0, quarter);
}

// HarmonicRestraintComputer (BondComputer.h:109-129) with CalcDistance
// (Interactions.h:42-53) and the harmonic form AnalyticalForceComputer<0>
// (Bonded/Analytical.h:18-30) at r0 = 0: F = unit(wrap_diff(x - x0)) * (-k d), full U.
__global__ void k_restraints(const int32_t* __restrict__ ids, const float4* __restrict__ pos, float4* fe,
							 const float4* __restrict__ anchors, const float* __restrict__ kspring, DevBox box,
							 int get_energy, uint64_t nr, float4* __restrict__ contrib) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= nr)
		return;
	const int id = ids[i];
	if (id < 0) {
		bonded_zero(contrib, i, 1);
		return;
	}
	const float3 r = wrap_diff(box, sub3(pos[id], anchors[i]));
	const float d = sqrtf(len2(r));
	float3 u = make_float3(0.0f, 0.0f, 0.0f);
	if (d > 1e-6f) {
		const float inv = 1.0f / d;
		u = make_float3(r.x * inv, r.y * inv, r.z * inv);
	}
	const float k = kspring[i];
	const float dr = d - 0.0f;
	const float force = -k * dr;
	bonded_emit(fe, contrib, i, id, u.x * force, u.y * force, u.z * force, get_energy != 0, 0.5f * k * dr * dr);
}

// BDIntegrate::operator() (BD.h:44-92).
__global__ void k_bd(marsmd_particle_view v, marsmd_particle_types types, DevBox box, float dt, uint64_t step,
					 float kT, uint64_t n, uint64_t seed, uint32_t ctr, const marsmd_grid_view* grids,
					 float3 ef, int scheme) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	if (v.flags && (v.flags[i] & MARSMD_FLAG_RB_ATTACHED))
		return;
	float4* pos = reinterpret_cast<float4*>(v.pos);
	float4* fe = reinterpret_cast<float4*>(v.force_energy);
	float4 p = pos[i];
	float4 f = fe[i];
	const int ty = v.type_id[i];
	const float efa[3] = {ef.x, ef.y, ef.z};
	const float4 pmf = position_dependent_force(p.x, p.y, p.z, ty, types, grids, efa, scheme);
	f.x += pmf.x;
	f.y += pmf.y;
	f.z += pmf.z;
	fe[i].w += pmf.w;
	const float* D = types.diffusion[ty];
	const float3 rn = box_muller(philox_first(seed, ctr + uint32_t(i), uint32_t(step), 0x5324120u));
	p.x += (D[0] / kT) * f.x * dt + sqrtf(2.0f * D[0] * dt) * rn.x;
	p.y += (D[1] / kT) * f.y * dt + sqrtf(2.0f * D[1] * dt) * rn.y;
	p.z += (D[2] / kT) * f.z * dt + sqrtf(2.0f * D[2] * dt) * rn.z;
	if (box.periodic[0])
		p.x = wrap_scalar(p.x, box.origin[0], box.L[0], box.inv_L[0]);
	if (box.periodic[1])
		p.y = wrap_scalar(p.y, box.origin[1], box.L[1], box.inv_L[1]);
	if (box.periodic[2])
		p.z = wrap_scalar(p.z, box.origin[2], box.L[2], box.inv_L[2]);
	pos[i] = p;
}

// BAOABIntegrate::operator() (BAOAB.h:46-120).
__global__ void k_baoab(marsmd_particle_view v, marsmd_particle_types types, DevBox box, float dt, uint64_t step,
						float kT, uint64_t n, uint64_t seed, uint32_t ctr) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	if (v.flags && (v.flags[i] & MARSMD_FLAG_RB_ATTACHED))
		return;
	float4* pos = reinterpret_cast<float4*>(v.pos);
	float4* mom = reinterpret_cast<float4*>(v.mom);
	const float4* fe = reinterpret_cast<const float4*>(v.force_energy);
	float4 p = pos[i], m = mom[i];
	const float4 f = fe[i];
	const int ty = v.type_id[i];
	const float mass = types.mass[ty];
	const float* g = types.trans_damping[ty];
	const float FCF = 4.18679994e4f;
	const float SQ = 2.046167337f;
	const float hdt = 0.5f * dt;
	m.x += hdt * f.x * FCF;
	m.y += hdt * f.y * FCF;
	m.z += hdt * f.z * FCF;
	const float inv_mass = 1.0f / mass;
	p.x += hdt * m.x * inv_mass * 10000.0f;
	p.y += hdt * m.y * inv_mass * 10000.0f;
	p.z += hdt * m.z * inv_mass * 10000.0f;
	const float c0 = expf(-g[0] * dt), c1 = expf(-g[1] * dt), c2 = expf(-g[2] * dt);
	const float n0 = sqrtf(kT * mass * (1.0f - c0 * c0)) * SQ;
	const float n1 = sqrtf(kT * mass * (1.0f - c1 * c1)) * SQ;
	const float n2 = sqrtf(kT * mass * (1.0f - c2 * c2)) * SQ;
	const float3 rn = box_muller(philox_first(seed, ctr + uint32_t(i), uint32_t(step), 0x1356914u));
	m.x = c0 * m.x + n0 * rn.x;
	m.y = c1 * m.y + n1 * rn.y;
	m.z = c2 * m.z + n2 * rn.z;
	p.x += hdt * m.x * inv_mass * 10000.0f;
	p.y += hdt * m.y * inv_mass * 10000.0f;
	p.z += hdt * m.z * inv_mass * 10000.0f;
	if (box.periodic[0])
		p.x = wrap_scalar(p.x, box.origin[0], box.L[0], box.inv_L[0]);
	if (box.periodic[1])
		p.y = wrap_scalar(p.y, box.origin[1], box.L[1], box.inv_L[1]);
	if (box.periodic[2])
		p.z = wrap_scalar(p.z, box.origin[2], box.L[2], box.inv_L[2]);
	pos[i] = p;
	mom[i] = m;
}

// BAOAB_LastUpdate::operator() (BAOAB.h:151-172).
__global__ void k_baoab_last(marsmd_particle_view v, float dt, uint64_t n) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	if (v.flags && (v.flags[i] & MARSMD_FLAG_RB_ATTACHED))
		return;
	float4* mom = reinterpret_cast<float4*>(v.mom);
	const float4 f = reinterpret_cast<const float4*>(v.force_energy)[i];
	float4 m = mom[i];
	const float hdt = 0.5f * dt;
	const float FCF = 4.18679994e4f;
	m.x += hdt * f.x * FCF;
	m.y += hdt * f.y * FCF;
	m.z += hdt * f.z * FCF;
	mom[i] = m;
}

// ComputePMFKernel (Pmf.h:126-152).
__global__ void k_pmf(marsmd_particle_view v, marsmd_particle_types types, uint64_t n,
					  const marsmd_grid_view* grids, float3 ef, int scheme) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const float4 p = reinterpret_cast<const float4*>(v.pos)[i];
	const float efa[3] = {ef.x, ef.y, ef.z};
	const float4 f = position_dependent_force(p.x, p.y, p.z, v.type_id[i], types, grids, efa, scheme);
	float4* fe = reinterpret_cast<float4*>(v.force_energy);
	float4 q = fe[i];
	q.x += f.x;
	q.y += f.y;
	q.z += f.z;
	q.w += f.w;
	fe[i] = q;
}

__global__ void k_philox(uint64_t seed, uint32_t base_ctr, uint32_t gs, uint32_t c1, uint32_t draws, uint64_t n,
						 uint4* out) {
	const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	for (uint32_t d = 0; d < draws; ++d)
		out[i * draws + d] = philox4x32_10(uint32_t(seed >> 32), uint32_t(seed & 0xFFFFFFFFu),
										   make_uint4(base_ctr + uint32_t(i), c1, gs, d));
}

} // namespace

void launch_resolve_pair_tables(const int32_t* pairs, const int32_t* types, const uint32_t* matrix, uint64_t nt,
								uint32_t* tag, uint64_t np, hipStream_t s) {
	if (np == 0)
		return;
	k_resolve<<<grid_for(np, 256), 256, 0, s>>>(reinterpret_cast<const int2*>(pairs), types, matrix, nt, tag, np);
	MMD_LAUNCH_CHECK();
}

void launch_pairwise_nonbonded(const int32_t* pairs, const marsmd_particle_view& v, const uint32_t* tags,
							   const marsmd_table* tables, const marsmd_particle_types* types, const marsmd_box& box,
							   bool energy, uint64_t np, float cutoff2, uint32_t enabled,
							   const marsmd_pair_params* params, hipStream_t s) {
	if (np == 0 || (enabled & MARSMD_PAIR_TERM_MASK) == 0u)
		return;
	const bool an = (enabled & kAnalyticTerms) != 0;
	if (an && (!types || !v.type_id))
		raise(MARSMD_ERR_VALUE, "launch_pairwise_nonbonded: the analytical terms need the particle types and type ids");
	if ((enabled & MARSMD_PAIR_TERM_TABULATED) && !tables)
		raise(MARSMD_ERR_VALUE, "launch_pairwise_nonbonded: the tabulated term needs the tables");
	const PairCoeffs C = types ? make_pair_coeffs(*types, params) : PairCoeffs{};
	auto kern = an ? k_pairwise<true, true> : k_pairwise<false, true>;
	kern<<<grid_for(np, 256), 256, 0, s>>>(reinterpret_cast<const int2*>(pairs), reinterpret_cast<const float4*>(v.pos),
										   reinterpret_cast<float4*>(v.force_energy), v.type_id, tags, tables, C,
										   make_devbox(box), energy ? 1 : 0, np, cutoff2, enabled);
	MMD_LAUNCH_CHECK();
}

void launch_analytical_pair_nonbonded(const int32_t* pairs, const float* pos, float* fe, const int32_t* type_ids,
									  const uint32_t* tags, const marsmd_particle_types& types, const marsmd_box& box,
									  uint32_t enabled, const marsmd_pair_params* params, bool energy, uint64_t np,
									  float cutoff2, hipStream_t s) {
	if (np == 0 || (enabled & kAnalyticTerms) == 0u)
		return;
	k_pairwise<true, false><<<grid_for(np, 256), 256, 0, s>>>(
		reinterpret_cast<const int2*>(pairs), reinterpret_cast<const float4*>(pos), reinterpret_cast<float4*>(fe),
		type_ids, tags, nullptr, make_pair_coeffs(types, params), make_devbox(box), energy ? 1 : 0, np, cutoff2,
		enabled);
	MMD_LAUNCH_CHECK();
}

void launch_tabulated_bonds(const int32_t* bonds, const float* pos, float* fe, const marsmd_table* tables,
							const int32_t* tidx, const int32_t* forms, const marsmd_box& box, bool energy,
							uint64_t nb, hipStream_t s, float4* contrib) {
	if (nb == 0)
		return;
	k_bonds<<<grid_for(nb, 256), 256, 0, s>>>(reinterpret_cast<const int2*>(bonds),
											   reinterpret_cast<const float4*>(pos), reinterpret_cast<float4*>(fe),
											   tables, tidx, forms, make_devbox(box), energy ? 1 : 0, nb, contrib);
	MMD_LAUNCH_CHECK();
}

void launch_tabulated_angles(const int32_t* angles, const float* pos, float* fe, const marsmd_table* tables,
							 const int32_t* tidx, const int32_t* forms, const marsmd_box& box, bool energy,
							 uint64_t na, hipStream_t s, float4* contrib) {
	if (na == 0)
		return;
	k_angles<<<grid_for(na, 256), 256, 0, s>>>(reinterpret_cast<const int4*>(angles),
												reinterpret_cast<const float4*>(pos), reinterpret_cast<float4*>(fe),
												tables, tidx, forms, make_devbox(box), energy ? 1 : 0, na, contrib);
	MMD_LAUNCH_CHECK();
}

void launch_tabulated_dihedrals(const int32_t* dihedrals, const float* pos, float* fe, const marsmd_table* tables,
								const int32_t* tidx, const int32_t* forms, const marsmd_box& box, bool energy,
								uint64_t nd, hipStream_t s, float4* contrib) {
	if (nd == 0)
		return;
	k_dihedrals<<<grid_for(nd, 256), 256, 0, s>>>(reinterpret_cast<const int4*>(dihedrals),
												   reinterpret_cast<const float4*>(pos),
												   reinterpret_cast<float4*>(fe), tables, tidx, forms,
												   make_devbox(box), energy ? 1 : 0, nd, contrib);
	MMD_LAUNCH_CHECK();
}

void launch_harmonic_restraints(const int32_t* ids, const float* pos, float* fe, const float* anchors,
								const float* kspring, const marsmd_box& box, bool energy, uint64_t nr,
								hipStream_t s, float4* contrib) {
	if (nr == 0)
		return;
	k_restraints<<<grid_for(nr, 256), 256, 0, s>>>(ids, reinterpret_cast<const float4*>(pos),
													reinterpret_cast<float4*>(fe),
													reinterpret_cast<const float4*>(anchors), kspring,
													make_devbox(box), energy ? 1 : 0, nr, contrib);
	MMD_LAUNCH_CHECK();
}

void launch_bd(const marsmd_particle_view& v, const marsmd_particle_types& t, float dt, uint64_t step, float kT,
			   uint64_t n, const marsmd_box& box, uint64_t seed, uint32_t ctr, const marsmd_grid_view* grids,
			   const float ef[3], int scheme, hipStream_t s) {
	if (n == 0)
		return;
	k_bd<<<grid_for(n, 256), 256, 0, s>>>(v, t, make_devbox(box), dt, step, kT, n, seed, ctr, grids,
										   make_float3(ef[0], ef[1], ef[2]), scheme);
	MMD_LAUNCH_CHECK();
}

void launch_baoab(const marsmd_particle_view& v, const marsmd_particle_types& t, const marsmd_box& box, float dt,
				  uint64_t step, float kT, uint64_t n, uint64_t seed, uint32_t ctr, hipStream_t s) {
	if (n == 0)
		return;
	k_baoab<<<grid_for(n, 256), 256, 0, s>>>(v, t, make_devbox(box), dt, step, kT, n, seed, ctr);
	MMD_LAUNCH_CHECK();
}

void launch_baoab_last_update(const marsmd_particle_view& v, float dt, uint64_t n, hipStream_t s) {
	if (n == 0)
		return;
	k_baoab_last<<<grid_for(n, 256), 256, 0, s>>>(v, dt, n);
	MMD_LAUNCH_CHECK();
}

void launch_pmf(const marsmd_particle_view& v, const marsmd_particle_types& t, uint64_t n,
				const marsmd_grid_view* grids, const float ef[3], int scheme, hipStream_t s) {
	if (n == 0)
		return;
	k_pmf<<<grid_for(n, 256), 256, 0, s>>>(v, t, n, grids, make_float3(ef[0], ef[1], ef[2]), scheme);
	MMD_LAUNCH_CHECK();
}

namespace {
// sqrt_rn_nonneg against sqrtf on the bit patterns lo + i, i < count.
__global__ void k_selftest_sqrt(uint32_t lo, uint32_t count, unsigned long long* __restrict__ bad) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= count)
		return;
	const float x = __uint_as_float(lo + i);
	if (__float_as_uint(sqrt_rn_nonneg(x)) != __float_as_uint(sqrtf(x)))
		atomicAdd(bad, 1ull);
}
} // namespace

uint64_t selftest_pair_sqrt(uint32_t lo_bits, uint32_t hi_bits, hipStream_t s) {
	if (hi_bits <= lo_bits)
		return 0;
	DevArray<unsigned long long> bad(1);
	fill_u32(reinterpret_cast<uint32_t*>(bad.data()), 0u, 2, s);
	const uint32_t count = hi_bits - lo_bits;
	k_selftest_sqrt<<<(count + 255u) / 256u, 256, 0, s>>>(lo_bits, count, bad.data());
	MMD_LAUNCH_CHECK();
	unsigned long long host = 0;
	MMD_HIP(hipMemcpyAsync(&host, bad.data(), sizeof(host), hipMemcpyDeviceToHost, s));
	MMD_HIP(hipStreamSynchronize(s));
	return uint64_t(host);
}

void philox_draw(uint64_t seed, uint32_t base_ctr, uint32_t gs, uint32_t c1, uint32_t draws, uint64_t n,
				 uint32_t* out, hipStream_t s) {
	if (n == 0)
		return;
	k_philox<<<grid_for(n, 256), 256, 0, s>>>(seed, base_ctr, gs, c1, draws, n, reinterpret_cast<uint4*>(out));
	MMD_LAUNCH_CHECK();
}

} // namespace mmd
