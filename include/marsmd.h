/*
 * marsmd.h — C ABI of the MI355X (gfx950) implementation of MARS-MD's per-step
 * particle hot path: Z-order keys + stable radix sort, the cell-stencil pairlist,
 * tabulated nonbonded / bond / angle forces, grid-potential (PMF) sampling and the
 * Brownian (BD) and Langevin (BAOAB) integrators with the Philox counter RNG.
 *
 * Conventions (all entry points):
 *   - Every function returns marsmd_status (0 = success). On failure
 *     marsmd_last_error() returns a thread-local message; the C++ mirror rethrows it
 *     as MARS::Exception, as the reference's MARS_Exception does.
 *   - Array pointers are DEVICE pointers unless the parameter is documented "host".
 *   - Vector3 arrays are 16-byte {x, y, z, t} records (the reference's alignas(16)
 *     Vector3_t<float>, src/Types/Vector3.h:41-56); ForceEnergy.t carries energy.
 *   - int2 pair arrays are {x, y} with x < y in the caller's (canonical) indices.
 *   - Launches are ordered on the hipStream_t passed (NULL = legacy default stream)
 *     and never synchronise implicitly, except where a function returns a host value.
 *   - The caller owns every buffer it passes; scratch belongs to opaque handles.
 *   - Structures cross the boundary by pointer, never by value (some FFIs, CPython
 *     3.10 ctypes among them, mis-classify by-value structs that contain arrays).
 *
 * Each declaration cites the reference interface it replaces (paths relative to the
 * reference tree).
 */
#ifndef MARSMD_H_
#define MARSMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int marsmd_status;
typedef void* marsmd_stream; /* hipStream_t */

#define MARSMD_OK 0
#define MARSMD_ERR_VALUE 1   /* ExceptionType::ValueError   (src/MARSException.h:103-130) */
#define MARSMD_ERR_RUNTIME 2 /* ExceptionType::RuntimeError                            */
#define MARSMD_ERR_DEVICE 3  /* a HIP runtime call failed (CUDA_CHECK, KernelHelper.cuh:170) */
#define MARSMD_ERR_CAPACITY 4 /* pairlist capacity ceiling (ZOrderPairlist.cpp:292-335) */

#define MARSMD_MAX_TYPES 16
#define MARSMD_MAX_PMF_TERMS 32

/* Pair-term bits of a pair tag (src/Interactions/Nonbonded/Pairwise.h:19-45). */
#define MARSMD_PAIR_TERM_NONE 0u
#define MARSMD_PAIR_TERM_COULOMB (1u << 0)
#define MARSMD_PAIR_TERM_DEBYE_HUCKEL (1u << 1)
#define MARSMD_PAIR_TERM_ONCK (1u << 2)
#define MARSMD_PAIR_TERM_GAUSSIAN (1u << 3)
#define MARSMD_PAIR_TERM_SOFTCORE (1u << 4)
#define MARSMD_PAIR_TERM_TABULATED (1u << 5)
#define MARSMD_PAIR_TERM_MASK 0xFFu
#define MARSMD_PAIR_TABLE_SHIFT 8

/* Integrators (IntegratorType) and grid enums (src/Types/GridTerm.h:26-36). */
#define MARSMD_INTEGRATOR_BROWNIAN 0
#define MARSMD_INTEGRATOR_LANGEVIN 1
#define MARSMD_INTERP_LINEAR 1
#define MARSMD_INTERP_CUBIC 3
#define MARSMD_BC_DIRICHLET 0
#define MARSMD_BC_NEUMANN 1
#define MARSMD_BC_PERIODIC 2

/* Particle flag FLAG_RB_ATTACHED (src/Objects/DeviceParticle.h): not integrated. */
#define MARSMD_FLAG_RB_ATTACHED (1u << 2)

/* InteractionForm::Tabulated as stored in the bonded `forms` arrays. */
#define MARSMD_FORM_TABULATED 1

/* ------------------------------------------------------------------ POD views */

/* Orthorhombic PeriodicBox (src/System/PeriodicBox.h:16-209, 445-461). A zero
 * periodic flag marks an open axis. Triclinic boxes are not part of this path. */
typedef struct marsmd_box {
    float size[3];
    float origin[3];
    int32_t periodic[3];
} marsmd_box;

/* TabulatedPotential (src/Interactions/TabulatedPotential.h:8-13); pot is a device
 * pointer to `size` knots. */
typedef struct marsmd_table {
    const float* pot;
    float step_inv;
    uint32_t size;
    float start;
    int32_t is_periodic;
} marsmd_table;

/* ExclusionView (src/Interactions/DeviceExclusions.h:22-75): symmetric CSR over
 * particles [0, num_particles), plus an optional per-particle rigid-body id. */
typedef struct marsmd_exclusion_view {
    const int32_t* offsets;
    const int32_t* excluded;
    uint64_t num_particles;
    const int32_t* rigid_body_id;
} marsmd_exclusion_view;

/* GridTerm (src/Types/GridTerm.h:55-69). */
typedef struct marsmd_grid_term {
    int32_t grid_id;
    float scale;
    float scale_slope;
    int32_t boundary_condition;
} marsmd_grid_term;

/* BaseGridView<float> (src/Types/BaseGridDevice.h:24-30, 105+). basis and
 * basis_inv are column-major 3x3 (cols[0] = first basis vector, as Matrix3_t).
 * Layout: value(ix,iy,iz) = data[iz + iy*nz + ix*ny*nz]. */
typedef struct marsmd_grid_view {
    const float* data;
    float origin[3];
    float basis[9];
    float basis_inv[9];
    uint64_t dims[3];
    int32_t grid_id;
    int32_t boundary_condition;
} marsmd_grid_view;

/* ParticleTypeView (src/Objects/DeviceParticle.h:49-65), held BY VALUE (host
 * struct, passed to kernels as kernel arguments). */
typedef struct marsmd_particle_types {
    int32_t num_types;
    float mass[MARSMD_MAX_TYPES];
    float charge[MARSMD_MAX_TYPES];
    float radius[MARSMD_MAX_TYPES]; /* softcore pair term                        */
    float eps[MARSMD_MAX_TYPES];    /* softcore pair term                        */
    float diffusion[MARSMD_MAX_TYPES][3];
    float trans_damping[MARSMD_MAX_TYPES][3];
    int32_t pmf_grid_offset[MARSMD_MAX_TYPES];
    int32_t pmf_grid_count[MARSMD_MAX_TYPES];
    marsmd_grid_term pmf_grid_terms[MARSMD_MAX_PMF_TERMS];
    int32_t force_grid_id[MARSMD_MAX_TYPES][3];
    float force_grid_scale[MARSMD_MAX_TYPES][3];
} marsmd_particle_types;

/* Constants of the analytical pair terms: SolventParams (NonBondedInteraction.h:121-127)
 * for Debye-Hueckel and ONC, and GaussianPotential's members (Pairwise.h:53-57), which
 * launch_pairwise_nonbonded leaves at their defaults. A NULL pointer wherever one is
 * taken means these defaults: debye_length 10, dielectric 80, onck_kappa 0.1,
 * onck_sz 80, onck_z 6.86, gaussian_amp 1, gaussian_sigma 1, gaussian_cutoff_squared 0
 * (no cutoff of its own). */
typedef struct marsmd_pair_params {
    float debye_length;
    float dielectric;
    float onck_kappa;
    float onck_sz;
    float onck_z;
    float gaussian_amp;
    float gaussian_sigma;
    float gaussian_cutoff_squared;
} marsmd_pair_params;

/* ParticleView (src/Objects/DeviceParticle.h:18-28). Vector3 fields are float4
 * records; flags and attached_rigid_body_id may be NULL. */
typedef struct marsmd_particle_view {
    int32_t* id;
    int32_t* type_id;
    float* pos;
    float* mom;
    float* force_energy;
    uint32_t* flags;
    int32_t* attached_rigid_body_id;
} marsmd_particle_view;

/* Pairlist::Statistics (src/PatchOperation/Pairlist.h:177-199) plus the cell grid
 * of ZOrderPairlist::select_cell_grid (ZOrderPairlist.cpp:24-81). */
typedef struct marsmd_pairlist_stats {
    uint64_t num_particles;
    uint64_t num_pairs;
    double average_neighbors_per_particle;
    uint32_t max_neighbors_per_particle;
    double build_time_ms;
    int32_t cell_grid_bits;     /* the reference's choice (select_cell_grid)        */
    int32_t cell_radii[3];
    int32_t neighbors_per_cell;
    int32_t build_grid_bits;    /* the grid this build actually walked              */
    int32_t build_radii[3];
    int32_t reference_grid_complete; /* 1: the reference's stencil covers every pair */
    uint64_t pair_capacity;
} marsmd_pairlist_stats;

/* ------------------------------------------------------------------ errors */

const char* marsmd_last_error(void);
/* Library identification: returns a static string "marsmd-hip <version> gfx950". */
const char* marsmd_version(void);

/* ------------------------------------------------------------------ Z-order sort
 * Replaces class ZOrderSort (src/PatchOperation/ZOrderKernels/ZOrderSort.h:43-313):
 * MortonEncodeKernel (ZOrderKernels.h:27-43) + stable radix sort of (key, index)
 * (ZOrderSort.h:255-294, CUB SortPairs in CUDASort.cu:13-98) + InverseIndexKernel
 * (ZOrderKernels.h:68-83). */
typedef struct marsmd_zorder marsmd_zorder;

marsmd_status marsmd_zorder_create(uint64_t max_particles, marsmd_zorder** out);
marsmd_status marsmd_zorder_destroy(marsmd_zorder* z);
/* ZOrderSort::sort_particles(positions, n, box_min, box_max); box_min/max host. */
marsmd_status marsmd_zorder_sort_particles(marsmd_zorder* z, const float* pos4, uint64_t n,
                                           const float box_min[3], const float box_max[3],
                                           marsmd_stream stream);
/* get_morton_codes / get_sorted_indices / get_inverse_indices (device pointers,
 * valid until the next sort). */
marsmd_status marsmd_zorder_get_buffers(marsmd_zorder* z, const uint32_t** morton_codes,
                                        const uint32_t** sorted_indices,
                                        const uint32_t** inverse_indices);
/* ZOrderSort::reorder_data<T>: out[i] = in[sorted_indices[i]]; elem_size 4, 8 or 16. */
marsmd_status marsmd_zorder_reorder_data(marsmd_zorder* z, const void* in, void* out,
                                         uint64_t n, uint32_t elem_size, marsmd_stream stream);
/* ZOrderSort::remap_indices (RemapIndicesKernel ZOrderKernels.h:87-101). */
marsmd_status marsmd_zorder_remap_indices(marsmd_zorder* z, int32_t* indices, uint64_t n,
                                          marsmd_stream stream);
/* ZOrderSort::update_positions_incremental (ZOrderSort.h:172-184): keep a copy of n
 * device positions (float4 records) as the reference of compute_max_displacement. */
marsmd_status marsmd_zorder_update_positions_incremental(marsmd_zorder* z, const float* pos4, uint64_t n,
                                                         marsmd_stream stream);
/* ZOrderSort::compute_max_displacement (ZOrderSort.h:160-170): the largest distance a
 * particle moved since the last update_positions_incremental (same n), minimum image on
 * axes whose box_len is > 0 (box_len may be NULL: all axes open). Reads the result back
 * (synchronises the stream); MARSMD_ERR_RUNTIME when no reference was kept. */
marsmd_status marsmd_zorder_compute_max_displacement(marsmd_zorder* z, const float* pos4, uint64_t n,
                                                     const float box_len[3], float* max_displacement,
                                                     marsmd_stream stream);
/* Stand-alone stable key/index sort used by the above (device keys in, sorted out):
 * orders by bits [0, key_bits) of the keys; the four buffers must not overlap
 * (MARSMD_ERR_VALUE otherwise). */
marsmd_status marsmd_radix_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in,
                                      uint32_t* keys_out, uint32_t* vals_out, uint64_t n,
                                      int32_t key_bits, void* scratch, uint64_t scratch_bytes,
                                      marsmd_stream stream);
/* Scratch bytes marsmd_radix_sort_pairs needs for n keys. */
uint64_t marsmd_radix_sort_scratch_bytes(uint64_t n);
/* No counterpart in the reference. For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): marsmd_radix_sort_pairs with the library's own fixed workspace, a device
 * array of marsmd_radix_sort_static_scratch_bytes() bytes that every GPU loading the
 * library holds until the process ends. Same results and buffer rules as
 * marsmd_radix_sort_pairs; at most MARSMD_STATIC_SORT_MAX_KEYS keys (MARSMD_ERR_VALUE
 * beyond). It allocates nothing; the first call on a device looks the workspace's
 * address up. Every call on one GPU shares that workspace: calls that overlap in time
 * on one GPU (other streams, other threads) race. */
#define MARSMD_STATIC_SORT_MAX_KEYS (1u << 21)
marsmd_status marsmd_radix_sort_pairs_static(const uint32_t* keys_in, const uint32_t* vals_in,
                                             uint32_t* keys_out, uint32_t* vals_out, uint64_t n,
                                             int32_t key_bits, marsmd_stream stream);
uint64_t marsmd_radix_sort_static_scratch_bytes(void);

/* ------------------------------------------------------------------ pairlist
 * Replaces ZOrderPairlist (src/PatchOperation/PairListKernels/ZOrderPairlist.h:27-190,
 * ZOrderPairlist.cpp:102-411; ZOrderNeighbor.h:27-225) behind Pairlist
 * (src/PatchOperation/Pairlist.h:56-252) and create_pairlist (Pairlist.cpp:7-31). */
typedef struct marsmd_pairlist marsmd_pairlist;

marsmd_status marsmd_pairlist_create(uint64_t max_particles, uint64_t max_pairs,
                                     marsmd_pairlist** out);
marsmd_status marsmd_pairlist_destroy(marsmd_pairlist* pl);
/* Pairlist::set_periodic_box (host struct). */
marsmd_status marsmd_pairlist_set_periodic_box(marsmd_pairlist* pl, const marsmd_box* box);
/* Pairlist::set_exclusions(ExclusionView) — pointers borrowed until the next build. */
marsmd_status marsmd_pairlist_set_exclusions(marsmd_pairlist* pl, const marsmd_exclusion_view* view);
/* Pairlist::set_max_pairs_ceiling (Pairlist.h:24-35): the regrow cap in pairs. */
marsmd_status marsmd_pairlist_set_max_pairs_ceiling(marsmd_pairlist* pl, uint64_t ceiling);
/* ZOrderPairlist::build_pairlist(positions, n, pairlist_cutoff). type_ids may be NULL
 * (types read as 0). Synchronises once to size the list (as the reference's
 * pair_count read-back at ZOrderPairlist.cpp:294). */
marsmd_status marsmd_pairlist_build(marsmd_pairlist* pl, const float* pos4,
                                    const int32_t* type_ids, uint64_t n, float pairlist_cutoff,
                                    marsmd_stream stream);
/* ZOrderPairlist::needs_update (ZOrderPairlist.cpp:184-190): max displacement since
 * the last build > skin/2. Host result; synchronises. */
marsmd_status marsmd_pairlist_needs_update(marsmd_pairlist* pl, const float* pos4, uint64_t n,
                                           float skin, int32_t* needs, float* max_disp,
                                           marsmd_stream stream);
/* get_num_pairs (host). */
marsmd_status marsmd_pairlist_get_num_pairs(marsmd_pairlist* pl, uint64_t* num_pairs);
/* get_neighbor_pairs(): int2 pairs in the caller's indices, x < y. The list is
 * materialised on first request after a build. *pairs is a device pointer. */
marsmd_status marsmd_pairlist_get_neighbor_pairs(marsmd_pairlist* pl, const int32_t** pairs,
                                                 marsmd_stream stream);
/* get_statistics(). */
marsmd_status marsmd_pairlist_get_statistics(marsmd_pairlist* pl, marsmd_pairlist_stats* out);
/* The sorter of the last build (get_morton_codes etc. of ZOrderPairlist's sorter_). */
marsmd_status marsmd_pairlist_get_sorter(marsmd_pairlist* pl, marsmd_zorder** out);
/* Cell tables of the reference grid for the last build (BuildCellRangesKernel,
 * BuildCellNeighborsKernel): device pointers + sizes. Built on request. */
marsmd_status marsmd_pairlist_get_reference_cells(marsmd_pairlist* pl, const uint32_t** begin,
                                                  const uint32_t** end,
                                                  const uint32_t** neighbors,
                                                  uint64_t* num_cells, int32_t* slots,
                                                  marsmd_stream stream);

/* ------------------------------------------------------------------ launchers */

/* launch_resolve_pair_tables (src/Interactions/NonBondedInteraction.h:205-261;
 * ResolvePairTableKernel Pairwise.h:104-120). */
marsmd_status marsmd_launch_resolve_pair_tables(const int32_t* pairs, const int32_t* type_ids,
                                                const uint32_t* term_matrix,
                                                uint64_t num_types, uint32_t* pair_tag,
                                                uint64_t num_pairs, marsmd_stream stream);
/* launch_pairwise_nonbonded (NonBondedInteraction.h:262-296; PairNonbondedComputer
 * Pairwise.h:184-305): every enabled term (Coulomb, Debye-Hueckel, ONC, Gaussian,
 * softcore, tabulated) of each pair in one pass. tables: device array of marsmd_table
 * (may be NULL without the tabulated term). pair_tag may be NULL (every pair takes
 * enabled_terms). types: host record (charges, softcore radius and eps); params: the
 * solvent constants (NULL: defaults). */
marsmd_status marsmd_launch_pairwise_nonbonded(const int32_t* pairs, const marsmd_particle_view* view,
                                               const uint32_t* pair_tag,
                                               const marsmd_table* tables, const marsmd_particle_types* types,
                                               const marsmd_box* box, int32_t get_energy, uint64_t num_pairs,
                                               float cutoff_squared, uint32_t enabled_terms,
                                               const marsmd_pair_params* params, marsmd_stream stream);
/* launch_analytical_pair_nonbonded (AnalyticalPairKernels.h:21-118): the analytical
 * terms only, all of their constants taken from params (NULL: defaults). pos4 and
 * force_energy4 are float4 records; type_ids int32. */
marsmd_status marsmd_launch_analytical_pair_nonbonded(const int32_t* pairs, const float* pos4, float* force_energy4,
                                                      const int32_t* type_ids, const uint32_t* pair_tag,
                                                      const marsmd_particle_types* types, const marsmd_box* box,
                                                      uint32_t enabled_terms, const marsmd_pair_params* params,
                                                      int32_t get_energy, uint64_t num_pairs, float cutoff_squared,
                                                      marsmd_stream stream);
/* launch_tabulated_bonds (src/Interactions/DeviceBondedInteraction.h:495-518;
 * TabulatedBondComputer BondComputer.h:136-192). bonds: int2 array. */
marsmd_status marsmd_launch_tabulated_bonds(const int32_t* bonds, const float* pos4,
                                            float* force_energy, const marsmd_table* tables,
                                            const int32_t* table_indices, const int32_t* forms,
                                            const marsmd_box* box, int32_t get_energy,
                                            uint64_t num_bonds, marsmd_stream stream);
/* launch_tabulated_angles (DeviceBondedInteraction.h:520-546; TabulatedAngleComputer
 * BondComputer.h:204-269). angles: 16-byte int records {i, j, k, pad}. */
marsmd_status marsmd_launch_tabulated_angles(const int32_t* angles, const float* pos4,
                                             float* force_energy, const marsmd_table* tables,
                                             const int32_t* table_indices, const int32_t* forms,
                                             const marsmd_box* box, int32_t get_energy,
                                             uint64_t num_angles, marsmd_stream stream);
/* launch_tabulated_dihedrals (DeviceBondedInteraction.h:548-573; TabulatedDihedralComputer
 * BondComputer.h:281-346 with DihedralGeometry BondGeometry.h:56-87). dihedrals: 16-byte
 * int records {i, j, k, l}; tables periodic, abscissa in radians. */
marsmd_status marsmd_launch_tabulated_dihedrals(const int32_t* dihedrals, const float* pos4,
                                                float* force_energy, const marsmd_table* tables,
                                                const int32_t* table_indices, const int32_t* forms,
                                                const marsmd_box* box, int32_t get_energy,
                                                uint64_t num_dihedrals, marsmd_stream stream);
/* launch_harmonic_restraints (DeviceBondedInteraction.h:595-620; HarmonicRestraintComputer
 * BondComputer.h:84-131): particle_ids (< 0 skipped), anchors as float4 records,
 * one spring constant per restraint. */
marsmd_status marsmd_launch_harmonic_restraints(const int32_t* particle_ids, const float* pos4,
                                                float* force_energy, const float* anchors4,
                                                const float* spring_constants, const marsmd_box* box,
                                                int32_t get_energy, uint64_t num_restraints,
                                                marsmd_stream stream);
/* launch_BD (src/PatchOperation/Integrator.h:25-54; BDIntegrate BD.h:44-92). grids:
 * device array of marsmd_grid_view (NULL = none). */
marsmd_status marsmd_launch_bd(const marsmd_particle_view* view, const marsmd_particle_types* types,
                               float timestep, uint64_t current_step, float kT,
                               uint64_t num_particles, const marsmd_box* box, uint64_t base_seed,
                               uint32_t base_ctr, const marsmd_grid_view* grids,
                               const float electric_field[3], int32_t interpolation,
                               marsmd_stream stream);
/* launch_BAOAB (Integrator.h:56-85; BAOABIntegrate BAOAB.h:46-120). */
marsmd_status marsmd_launch_baoab(const marsmd_particle_view* view, const marsmd_particle_types* types,
                                  const marsmd_box* box, float timestep, uint64_t current_step, float kT,
                                  uint64_t num_particles, uint64_t base_seed, uint32_t base_ctr,
                                  marsmd_stream stream);
/* launch_BAOAB_LastUpdate (Integrator.h:97-123; BAOAB.h:151-172). */
marsmd_status marsmd_launch_baoab_last_update(const marsmd_particle_view* view,
                                              const marsmd_particle_types* types,
                                              float timestep, uint64_t num_particles,
                                              marsmd_stream stream);
/* launch_PMF (Integrator.h:127-144; ComputePMFKernel Pmf.h:126-152). */
marsmd_status marsmd_launch_pmf(const marsmd_particle_view* view, const marsmd_particle_types* types,
                                uint64_t num_particles, const marsmd_grid_view* grids,
                                const float electric_field[3], int32_t interpolation,
                                marsmd_stream stream);
/* Raw Philox4x32-10 draws (openrand::Philox, src/PatchOperation/Random/philox.h:60-139)
 * for counters ctr0 = base_ctr + i, i in [0, n): out is n uint4 records. Used to pin
 * the device RNG bit-for-bit. */
marsmd_status marsmd_philox_draw(uint64_t seed, uint32_t base_ctr, uint32_t global_seed,
                                 uint32_t ctr1, uint32_t draws, uint64_t n, uint32_t* out,
                                 marsmd_stream stream);
/* No counterpart in the reference: a self-test of the square root the step kernels use
 * for pair distances (the hardware square root corrected to correct rounding). Compares
 * it with the correctly rounded sqrtf on every float whose bit pattern lies in
 * [lo_bits, hi_bits) and writes the number of differences to *mismatches (host). */
marsmd_status marsmd_selftest_pair_sqrt(uint32_t lo_bits, uint32_t hi_bits, uint64_t* mismatches,
                                        marsmd_stream stream);

/* ------------------------------------------------------------------ Patch
 * Replaces class Patch (src/PatchOperation/Patch.h:48-645, Patch.cpp:43-480) for one
 * patch on one GPU. The patch owns its particle state in HBM in its own (Morton)
 * layout; copy_particles_{from,to}_host translate to the caller's canonical order. */
typedef struct marsmd_patch marsmd_patch;

typedef struct marsmd_patch_desc {
    uint64_t capacity;           /* Patch(patch_t, capacity, ...)                     */
    uint64_t patch_id;           /* patch_t, enters the seed mix (Patch.cpp:252-254)   */
    marsmd_box box;
    marsmd_particle_types types;
    /* nonbonded: tag matrix [num_types^2] (host) and tables (host knots) */
    const uint32_t* term_matrix;
    uint32_t enabled_terms;
    int32_t num_nb_tables;
    const marsmd_table* nb_tables; /* host records; .pot host pointers to knots */
    /* bonded topology, host arrays in canonical indices (BondedInteractions) */
    uint64_t num_bonds;
    const int32_t* bonds;          /* 2 per bond */
    const int32_t* bond_table;     /* table index per bond */
    uint64_t num_angles;
    const int32_t* angles;         /* 3 per angle */
    const int32_t* angle_table;
    int32_t num_bond_tables;
    const marsmd_table* bond_tables;  /* host */
    int32_t num_angle_tables;
    const marsmd_table* angle_tables; /* host, is_periodic set, radians */
    uint64_t num_exclusions;
    const int32_t* exclusions;     /* 2 per exclusion (host) */
    /* grids (host data, copied to HBM) */
    int32_t num_grids;
    const marsmd_grid_view* grids; /* host records; .data host pointers */
    /* dihedrals and harmonic restraints (BondedInteractions), canonical indices */
    uint64_t num_dihedrals;
    const int32_t* dihedrals;      /* 4 per dihedral */
    const int32_t* dihedral_table;
    int32_t num_dihedral_tables;
    const marsmd_table* dihedral_tables; /* host, is_periodic set, radians */
    uint64_t num_restraints;
    const int32_t* restraint_ids;     /* particle per restraint */
    const float* restraint_anchors;   /* 4 floats per restraint (x, y, z, unused) */
    const float* restraint_k;         /* spring constant per restraint */
    /* constants of the analytical pair terms (NULL: defaults) */
    const marsmd_pair_params* pair_params;
} marsmd_patch_desc;

marsmd_status marsmd_patch_create(const marsmd_patch_desc* desc, marsmd_patch** out);
marsmd_status marsmd_patch_destroy(marsmd_patch* p);
/* Patch::set_base_seed. */
marsmd_status marsmd_patch_set_base_seed(marsmd_patch* p, uint64_t seed);
/* Patch::copy_particles_from_host (host arrays, canonical order; mom may be NULL). */
marsmd_status marsmd_patch_copy_particles_from_host(marsmd_patch* p, const float* pos4,
                                                    const float* mom4, const int32_t* type_ids,
                                                    uint64_t n);
/* Patch::copy_particles_to_host (host arrays, canonical order; any may be NULL). The
 * canonical order changes at each reorder_particles, as the reference's SoA does;
 * ids returns each slot's global id (HostParticleData::global_id). */
marsmd_status marsmd_patch_copy_particles_to_host(marsmd_patch* p, float* pos4, float* mom4,
                                                  float* force_energy4, int32_t* type_ids,
                                                  int32_t* ids, uint64_t n);
/* Patch::calculate_nonbonded_forces (Patch.cpp:43-154): clear forces, rebuild policy,
 * pairwise force. interpolation/electric_field are stashed for integrate_motion. */
marsmd_status marsmd_patch_calculate_nonbonded_forces(marsmd_patch* p, float pairlist_cutoff,
                                                      float interaction_cutoff, uint64_t step,
                                                      uint64_t rebuild_period,
                                                      const float electric_field[3],
                                                      int32_t interpolation,
                                                      int32_t compute_energy,
                                                      marsmd_stream stream);
/* Patch::calculate_bonded_forces (Patch.cpp:156-236). */
marsmd_status marsmd_patch_calculate_bonded_forces(marsmd_patch* p, int32_t compute_energy,
                                                   marsmd_stream stream);
/* Patch::integrate_motion (Patch.cpp:238-352). */
marsmd_status marsmd_patch_integrate_motion(marsmd_patch* p, float dt, float kT,
                                            int32_t integrator, uint64_t step,
                                            marsmd_stream stream);
/* Patch::finish_deferred_kick (Patch.cpp:354-373). */
marsmd_status marsmd_patch_finish_deferred_kick(marsmd_patch* p, float dt, marsmd_stream stream);
/* Patch::has_deferred_kick (Patch.h:300-304): 1 while a BAOAB closing half-kick is
 * outstanding (after a Langevin step, until finish_deferred_kick or the next step). */
marsmd_status marsmd_patch_has_deferred_kick(marsmd_patch* p, int32_t* pending);
/* Patch::reorder_particles (Patch.cpp:456-480) via ParticleReorderManager. */
marsmd_status marsmd_patch_reorder_particles(marsmd_patch* p, marsmd_stream stream);
/* Patch::get_pairlist(). */
marsmd_status marsmd_patch_get_pairlist(marsmd_patch* p, marsmd_pairlist** out);

/* The per-step loop of SimManager::run (SimManager.cpp:271-296, 337-504) for steps
 * first_step+1 .. first_step+num_steps, without output: reorder cadence
 * (ReorderManager.cpp:8-25), rebuild policy (Patch.cpp:85-121), forces and
 * integration, with force + PMF + integrator fused into one kernel per step. */
typedef struct marsmd_run_params {
    float pairlist_cutoff;
    float cutoff;
    uint64_t rebuild_period;   /* decompPeriod */
    int64_t reorder_period;    /* reorderPeriod (<= 0: falls back to rebuild_period) */
    float timestep;
    float kT;
    int32_t integrator;
    int32_t interpolation;
    float electric_field[3];
    uint64_t energy_period;    /* outputEnergyPeriod (0 = never) */
} marsmd_run_params;

marsmd_status marsmd_patch_run(marsmd_patch* p, const marsmd_run_params* params,
                               uint64_t first_step, uint64_t num_steps, marsmd_stream stream);

/* Launch timing of the fused step kernel inside marsmd_patch_run: while enabled, each
 * launch is bracketed by HIP events on the run's stream. Enabling (or disabling)
 * clears the record; the getter waits for the recorded launches and returns their
 * summed duration in milliseconds and their number. A measurement hook with no
 * counterpart in the reference. */
marsmd_status marsmd_patch_set_kernel_timing(marsmd_patch* p, int32_t enable);
marsmd_status marsmd_patch_get_kernel_timing(marsmd_patch* p, double* total_ms, uint64_t* launches);

/* Launch configuration of marsmd_patch_run's step (no counterpart in the reference).
 * Results do not depend on it: each particle's step is one lane's work whatever the
 * grid, and a precomputed position-dependent force is the same function of the same
 * inputs. With split_pmf, the timing events above bracket both kernels of a step. */
typedef struct marsmd_patch_launch {
    int32_t blocks_per_cu; /* persistent step-kernel workgroups per CU; 0: as many as fit */
    int32_t split_pmf;     /* with grids: the position-dependent force (grid sampling and
                              field) by a kernel of its own, read by the step kernel */
} marsmd_patch_launch;

marsmd_status marsmd_patch_set_launch(marsmd_patch* p, const marsmd_patch_launch* launch);
marsmd_status marsmd_patch_get_launch(marsmd_patch* p, marsmd_patch_launch* launch);
/* Steps first_step+1 .. first_step+num_steps like marsmd_patch_run while timing every
 * candidate launch configuration on the step kernels (two passes over the candidates),
 * then keeps the fastest: *best is set to it and *best_ms_per_step to its time. */
marsmd_status marsmd_patch_autotune(marsmd_patch* p, const marsmd_run_params* params, uint64_t first_step,
                                    uint64_t num_steps, marsmd_patch_launch* best, double* best_ms_per_step,
                                    marsmd_stream stream);
/* Cell configuration of the Patch's pairlist walk: home cells (one wave's tile of home
 * particles) of the finest Morton level that holds home_target particles on average
 * (1 .. 4096, default 24), candidate cells fine_offset levels below (1 .. 4, default
 * 2). The pair set does not depend on it; the order of neighbours in a row, and so the
 * FP32 summation order of the forces, does. Applies from the next build. No
 * counterpart in the reference (its cell grid is select_cell_grid's,
 * ZOrderPairlist.cpp:24-81). */
marsmd_status marsmd_patch_set_cells(marsmd_patch* p, int32_t home_target, int32_t fine_offset);
marsmd_status marsmd_patch_get_cells(marsmd_patch* p, int32_t* home_target, int32_t* fine_offset);
/* Runs first_step+1 .. first_step+num_steps as marsmd_patch_run, after choosing the cell
 * configuration (home_target 8/12/16/24/48 x fine_offset 1/2/3), each timed re-walking
 * the last build's positions; the live list is then walked with the choice.
 * *walk_ms: the choice's walk time (0 when the home walk does not apply). */
marsmd_status marsmd_patch_autotune_cells(marsmd_patch* p, const marsmd_run_params* params,
                                          uint64_t first_step, uint64_t num_steps, double* walk_ms,
                                          marsmd_stream stream);
/* Pruned rows of the fused step (marsmd_patch_run), for systems whose particles move
 * little within a list's life: after each build the step walks a copy of the
 * neighbour rows, each keeping (in list order) the entries within cutoff + 2 margin,
 * and falls back to the full rows, on the device, once a particle has moved a margin
 * from its position at the prune (less the rounding of FP32 distances), so the forces
 * are the full rows' bits in every mode.
 * mode 0 (default): off; 1: rows placed within each block of 256 slots by ascending
 * length (every entry kept); 2: placed so and pruned; 3: pruned at their own slots.
 * The margin is 1.5x the previous list life's largest displacement (+ skin / 50, at
 * most skin / 2; a quarter of the skin for the first), or margin when margin >= 0.
 * Applies from the next prune (the next build or position change). No counterpart in
 * the reference. */
marsmd_status marsmd_patch_set_pruning(marsmd_patch* p, int32_t mode, float margin);
/* The live configuration: mode, the margin of the pruned rows (0: none, rows sorted or
 * off), whether they keep every entry, and the largest displacement from the positions
 * at the prune recorded by the fused steps since (0 when none is recorded; a
 * synchronous read). */
marsmd_status marsmd_patch_get_pruning(marsmd_patch* p, int32_t* mode, float* margin, int32_t* keep_all,
                                       float* displacement);

/* ------------------------------------------------------------------ replica batch
 * Many independent replicas of a periodic nonbonded system, each one the reference's
 * Patch (Patch.h:48-645) driven by SimManager::run's per-step sequence
 * (SimManager.cpp:271-296, 337-504), stepped together: one launch per step over every
 * replica and one batched rebuild for all. Replica r's id enters its seed mix as the
 * patch id (Patch.cpp:252-254). The replicas share particle types and pair tables and
 * carry no bonded topology, exclusions or grids; each box is periodic on all axes. No
 * counterpart in the reference, which steps patches one at a time. */
typedef struct marsmd_batch marsmd_batch;

typedef struct marsmd_batch_desc {
    uint64_t num_replicas;
    const uint64_t* num_particles; /* host, per replica                                   */
    const marsmd_box* boxes;       /* host, per replica                                   */
    const uint64_t* patch_ids;     /* host, per replica (seed-mix word); NULL: 0, 1, ...  */
    marsmd_particle_types types;   /* shared by the replicas (no grid terms)             */
    const uint32_t* term_matrix;   /* host, num_types^2 pair tags                        */
    uint32_t enabled_terms;
    int32_t num_nb_tables;
    const marsmd_table* nb_tables; /* host records; .pot host pointers to knots          */
    const marsmd_pair_params* pair_params; /* analytical pair terms (NULL: defaults)     */
} marsmd_batch_desc;

marsmd_status marsmd_batch_create(const marsmd_batch_desc* desc, marsmd_batch** out);
marsmd_status marsmd_batch_destroy(marsmd_batch* b);
marsmd_status marsmd_batch_set_base_seed(marsmd_batch* b, uint64_t seed);
/* Every replica's particles, replica after replica, each in its own canonical order
 * (host arrays; mom may be NULL). Positions must lie inside their replica's box. */
marsmd_status marsmd_batch_copy_particles_from_host(marsmd_batch* b, const float* pos4, const float* mom4,
                                                    const int32_t* type_ids, uint64_t total);
marsmd_status marsmd_batch_copy_particles_to_host(marsmd_batch* b, float* pos4, float* mom4, float* force_energy4,
                                                  int32_t* type_ids, int32_t* ids, uint64_t total);
/* Steps first_step+1 .. first_step+num_steps of every replica (marsmd_patch_run). */
marsmd_status marsmd_batch_run(marsmd_batch* b, const marsmd_run_params* params, uint64_t first_step,
                               uint64_t num_steps, marsmd_stream stream);
/* Patch::finish_deferred_kick (Patch.cpp:354-373) of every replica. */
marsmd_status marsmd_batch_finish_deferred_kick(marsmd_batch* b, float dt, marsmd_stream stream);
/* The current pair set of one replica: int2 {x < y} in its canonical indices (device
 * pointer, valid until the next call); *num_pairs on the host. */
marsmd_status marsmd_batch_get_neighbor_pairs(marsmd_batch* b, uint64_t replica, const int32_t** pairs,
                                              uint64_t* num_pairs, marsmd_stream stream);
/* Launch timing of the batched step kernel (as marsmd_patch_set/get_kernel_timing):
 * event pairs around each launch or replayed segment; *steps counts the steps they
 * span. */
marsmd_status marsmd_batch_set_kernel_timing(marsmd_batch* b, int32_t enable);
marsmd_status marsmd_batch_get_kernel_timing(marsmd_batch* b, double* total_ms, uint64_t* steps);

/* Launch configuration of the batched step (no counterpart in the reference). Results
 * do not depend on it: each particle's step is one lane's work whatever the grid, and
 * a replayed segment launches the same kernels with the same arguments. */
typedef struct marsmd_batch_launch {
    int32_t blocks_per_cu; /* persistent step-kernel workgroups per CU; 0: as many as fit */
    int32_t use_graph;     /* replay the steps between host decisions (rebuild checks,
                              reorders, energy steps, a run's last step) as hipGraphs */
} marsmd_batch_launch;

marsmd_status marsmd_batch_set_launch(marsmd_batch* b, const marsmd_batch_launch* launch);
marsmd_status marsmd_batch_get_launch(marsmd_batch* b, marsmd_batch_launch* launch);
/* Steps first_step+1 .. first_step+num_steps like marsmd_batch_run while timing every
 * candidate launch configuration on runs of steps (two passes over the candidates),
 * then keeps the fastest: *best is set to it and *best_ms_per_step to its time. */
marsmd_status marsmd_batch_autotune(marsmd_batch* b, const marsmd_run_params* params, uint64_t first_step,
                                    uint64_t num_steps, marsmd_batch_launch* best, double* best_ms_per_step,
                                    marsmd_stream stream);
/* Cell configuration of the batch's pairlist walk per replica size bucket (bucket b:
 * 2^b <= n < 2^(b+1) particles): the home cells (one wave's tile of home particles)
 * are the finest Morton cells that still hold home_target[b] particles on average,
 * the candidate cells lie fine_offset[b] Morton levels below them. The pair set does
 * not depend on it; the order of neighbours in a row, and so the FP32 summation order
 * of the forces, does. Defaults: 24 and 2. No counterpart in the reference (its cell
 * grid is ZOrderPairlist::select_cell_grid's, ZOrderPairlist.cpp:24-81). */
#define MARSMD_BATCH_BUCKETS 32
typedef struct marsmd_batch_cells {
    int32_t home_target[MARSMD_BATCH_BUCKETS]; /* 1 .. 4096 */
    int32_t fine_offset[MARSMD_BATCH_BUCKETS]; /* 1 .. 4    */
} marsmd_batch_cells;
marsmd_status marsmd_batch_set_cells(marsmd_batch* b, const marsmd_batch_cells* cells);
marsmd_status marsmd_batch_get_cells(marsmd_batch* b, marsmd_batch_cells* cells);
/* Runs first_step+1 .. first_step+num_steps as marsmd_batch_run, after choosing the
 * cell configuration bucket by bucket (candidates home_target 8/16/24/48/96 x
 * fine_offset 1/2/3, each timed on the cell tables and walk of the last build's
 * positions into scratch buffers, the other buckets held); the choice applies from
 * the next build on. *walk_ms: the chosen configuration's walk time. */
marsmd_status marsmd_batch_autotune_cells(marsmd_batch* b, const marsmd_run_params* params,
                                          uint64_t first_step, uint64_t num_steps, double* walk_ms,
                                          marsmd_stream stream);

/* ------------------------------------------------------------------ force matching
 * No counterpart in the reference, which has no gradient path: the forces of a fixed
 * set of configurations (snapshots) under tabulated pair potentials whose knots are
 * trainable, and the adjoint dL/dknot of a loss of those forces in FP64. The force is
 * the Patch's pair term (PairNonbondedComputer, Pairwise.h:184-305) on each snapshot's
 * own pairlist, built once at creation and kept in HBM. */
typedef struct marsmd_fm marsmd_fm;

typedef struct marsmd_fm_desc {
    marsmd_box box;
    uint64_t num_particles;        /* per snapshot                                   */
    uint64_t num_snapshots;
    const float* positions;        /* device, num_snapshots x num_particles float4    */
    const int32_t* type_ids;       /* device, num_particles (the same every snapshot) */
    int32_t num_types;
    const uint32_t* term_matrix;   /* host, num_types^2 pair tags (make_pair_tag)     */
    int32_t num_tables;
    const marsmd_table* tables;    /* host: start, step_inv, size (non-periodic); .pot unused */
    float cutoff;                  /* force cutoff (d^2 > cutoff^2 skipped)           */
    float pairlist_cutoff;         /* pairs listed at d <= pairlist_cutoff (0: cutoff) */
} marsmd_fm_desc;

marsmd_status marsmd_fm_create(const marsmd_fm_desc* desc, marsmd_fm** out);
marsmd_status marsmd_fm_destroy(marsmd_fm* fm);
/* Listed pairs summed over the snapshots. */
marsmd_status marsmd_fm_num_pairs(marsmd_fm* fm, uint64_t* num_pairs);
/* Forces for the given knots (device FP32, all tables concatenated in table order):
 * force_energy4 is num_snapshots x num_particles float4 {F, energy}, each snapshot in
 * its own particle order. */
marsmd_status marsmd_fm_forces(marsmd_fm* fm, const float* knots, float* force_energy4, marsmd_stream stream);
/* dL/dknot (device FP64, one per knot) for dL/dF = grad_force4 (device, the layout of
 * force_energy4; .w ignored). Overwrites grad_knots. */
marsmd_status marsmd_fm_adjoint(marsmd_fm* fm, const float* knots, const float* grad_force4, double* grad_knots,
                                marsmd_stream stream);
/* The mean squared force error over every snapshot, particle and component, and its
 * gradient, in one pass: *loss (device FP64) = sum |F - target|^2 / (3 S n) with the
 * differences in FP32 and the squares summed in FP64 in a fixed order; grad_force4
 * (S x n float4) = (F - target) * (FP32 value of 2 / (3 S n)), .w = 0. force_energy4 as
 * marsmd_fm_forces writes it; target3: S x n x 3 floats in the same order. No
 * counterpart in the reference (it has no gradient path). */
marsmd_status marsmd_fm_mse(marsmd_fm* fm, const float* force_energy4, const float* target3, float* grad_force4,
                            double* loss, marsmd_stream stream);
/* The same forces and adjoint with each snapshot in its sorted order (the Morton order
 * of its pairlist build) instead of its particle order, and the permutation of S x n
 * rows of width 3 or 4 floats from particle order into sorted order (in != out). A
 * loss evaluated in sorted order (target permuted once) reads and writes every array
 * in order and lets the adjoint read dL/dF around each particle's slot. No counterpart
 * in the reference. */
marsmd_status marsmd_fm_forces_sorted(marsmd_fm* fm, const float* knots, float* force_energy4,
                                      marsmd_stream stream);
marsmd_status marsmd_fm_adjoint_sorted(marsmd_fm* fm, const float* knots, const float* grad_force4,
                                       double* grad_knots, marsmd_stream stream);
marsmd_status marsmd_fm_to_sorted(marsmd_fm* fm, const float* in, int32_t width, float* out, marsmd_stream stream);
/* marsmd_fm_forces_sorted followed by marsmd_fm_mse in one pass, the forces never
 * stored: target3 and grad_force4 in each snapshot's sorted order; grad_force4 has
 * marsmd_fm_mse's bits, the loss the same FP64 terms summed in another fixed grouping.
 * No counterpart in the reference (it has no gradient path). */
marsmd_status marsmd_fm_loss_sorted(marsmd_fm* fm, const float* knots, const float* target3, float* grad_force4,
                                    double* loss, marsmd_stream stream);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* MARSMD_H_ */
