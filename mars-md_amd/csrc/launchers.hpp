// launchers.hpp — host entry points of launchers.hip.
#pragma once

#include "internal.hpp"

namespace mmd {

void launch_resolve_pair_tables(const int32_t* pairs, const int32_t* types, const uint32_t* matrix, uint64_t nt,
								uint32_t* tag, uint64_t np, hipStream_t s);
void launch_pairwise_nonbonded(const int32_t* pairs, const marsmd_particle_view& v, const uint32_t* tags,
							   const marsmd_table* tables, const marsmd_particle_types* types, const marsmd_box& box,
							   bool energy, uint64_t np, float cutoff2, uint32_t enabled,
							   const marsmd_pair_params* params, hipStream_t s);
void launch_analytical_pair_nonbonded(const int32_t* pairs, const float* pos, float* fe, const int32_t* type_ids,
									  const uint32_t* tags, const marsmd_particle_types& types, const marsmd_box& box,
									  uint32_t enabled, const marsmd_pair_params* params, bool energy, uint64_t np,
									  float cutoff2, hipStream_t s);
// The bonded launchers add each term's forces (and energy shares) to fe with FP32
// atomics; with contrib they store them instead, share r of term i of a kind with R
// particles per term at contrib[R i + r] (fe unused), for a fixed-order gather.
void launch_tabulated_bonds(const int32_t* bonds, const float* pos, float* fe, const marsmd_table* tables,
							const int32_t* tidx, const int32_t* forms, const marsmd_box& box, bool energy,
							uint64_t nb, hipStream_t s, float4* contrib = nullptr);
void launch_tabulated_angles(const int32_t* angles, const float* pos, float* fe, const marsmd_table* tables,
							 const int32_t* tidx, const int32_t* forms, const marsmd_box& box, bool energy,
							 uint64_t na, hipStream_t s, float4* contrib = nullptr);
void launch_tabulated_dihedrals(const int32_t* dihedrals, const float* pos, float* fe, const marsmd_table* tables,
								const int32_t* tidx, const int32_t* forms, const marsmd_box& box, bool energy,
								uint64_t nd, hipStream_t s, float4* contrib = nullptr);
void launch_harmonic_restraints(const int32_t* ids, const float* pos, float* fe, const float* anchors,
								const float* kspring, const marsmd_box& box, bool energy, uint64_t nr,
								hipStream_t s, float4* contrib = nullptr);
void launch_bd(const marsmd_particle_view& v, const marsmd_particle_types& t, float dt, uint64_t step, float kT,
			   uint64_t n, const marsmd_box& box, uint64_t seed, uint32_t ctr, const marsmd_grid_view* grids,
			   const float ef[3], int scheme, hipStream_t s);
void launch_baoab(const marsmd_particle_view& v, const marsmd_particle_types& t, const marsmd_box& box, float dt,
				  uint64_t step, float kT, uint64_t n, uint64_t seed, uint32_t ctr, hipStream_t s);
void launch_baoab_last_update(const marsmd_particle_view& v, float dt, uint64_t n, hipStream_t s);
void launch_pmf(const marsmd_particle_view& v, const marsmd_particle_types& t, uint64_t n,
				const marsmd_grid_view* grids, const float ef[3], int scheme, hipStream_t s);
uint64_t selftest_pair_sqrt(uint32_t lo_bits, uint32_t hi_bits, hipStream_t s);
void philox_draw(uint64_t seed, uint32_t base_ctr, uint32_t gs, uint32_t c1, uint32_t draws, uint64_t n,
				 uint32_t* out, hipStream_t s);

} // namespace mmd
