// oracle_capi.cpp — extern "C" surface of the parity oracle (see oracle.hpp).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg. The product library never links it.
//
// What pins each part (beyond restating the reference):
//   philox / u01       golden vectors from the reference's own philox.h compiled in
//                      oracle/_ref (tests/golden/philox_ref.json) + the published
//                      Random123 Philox4x32-10 known-answer vectors.
//   morton             closed forms: corner codes, decode(encode(q)) == q, axis bit order.
//   sort               std::stable_sort equivalence (zorder_tests.cpp:139-151); the
//                      oracle IS std::stable_sort, checked for stability on duplicates.
//   pairlist           O(N^2) brute force of the same predicate, and the known-answer
//                      sets of Zorder_pairlist2.cpp:387-403, zorder_pairlist_tests.cpp:224-342.
//   tables / forces    closed forms (NonbondedEnergy.cpp:149-265), FP64 twin, central
//                      finite differences of the energy, momentum conservation.
//   bonds / angles     chain fixture bond-19.190-3.800.dat (BDKernels.cpp:222-320),
//                      finite differences, momentum/torque sums (TabulatedBonded.cpp).
//   grids              linear-ramp exactness, trilinear/Catmull-Rom reproduce polynomials
//                      they must, finite differences of the sampled value.
//   integrators        single-step closed forms with given noise, MSD = 6Dt,
//                      BAOAB equipartition <p^2> = kT m SQRT_CAL_TO_JOULE^2.
#include "oracle.hpp"

#include "../include/marsmd.h"

#include <cstdio>
#include <omp.h>
#include <map>
#include <memory>
#include <string>

using namespace orc;

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
	g_err = msg;
	return 1;
}

const V4f* as_v4(const float* p) { return reinterpret_cast<const V4f*>(p); }
V4f* as_v4(float* p) { return reinterpret_cast<V4f*>(p); }

Box make_box(const marsmd_box& b) {
	Box out;
	for (int a = 0; a < 3; ++a) {
		out.L[a] = b.size[a];
		out.origin[a] = b.origin[a];
		out.periodic[a] = b.periodic[a] != 0;
	}
	return out;
}

template<class T>
Table<T> make_table(const marsmd_table& t) {
	Table<T> out;
	out.pot.assign(t.pot, t.pot + t.size);
	out.step_inv = T(t.step_inv);
	out.start = T(t.start);
	out.periodic = t.is_periodic != 0;
	return out;
}

template<class T>
PairCoef<T> make_coef(const marsmd_particle_types& t, const marsmd_pair_params* p) {
	PairCoef<T> c;
	if (p) {
		c.debye_length = T(p->debye_length);
		c.dielectric = T(p->dielectric);
		c.onck_kappa = T(p->onck_kappa);
		c.onck_sz = T(p->onck_sz);
		c.onck_z = T(p->onck_z);
		c.gaussian_amp = T(p->gaussian_amp);
		c.gaussian_sigma = T(p->gaussian_sigma);
		c.gaussian_cutoff_squared = T(p->gaussian_cutoff_squared);
	}
	for (int k = 0; k < MARSMD_MAX_TYPES; ++k) {
		c.charge.push_back(T(t.charge[k]));
		c.radius.push_back(T(t.radius[k]));
		c.eps.push_back(T(t.eps[k]));
	}
	return c;
}

template<class T>
Grid<T> make_grid(const marsmd_grid_view& g) {
	Grid<T> out;
	out.data = g.data;
	for (int k = 0; k < 3; ++k) {
		out.origin[k] = T(g.origin[k]);
		out.dims[k] = size_t(g.dims[k]);
	}
	for (int k = 0; k < 9; ++k) {
		out.basis[k] = T(g.basis[k]);
		out.basis_inv[k] = T(g.basis_inv[k]);
	}
	out.bc = g.boundary_condition;
	return out;
}

// ------------------------------------------------------------- pairlist pieces

// ZOrderPairlist::compute_particle_extent (ZOrderPairlist.cpp:368-411) and the
// periodic-axis override of build_pairlist (:131-146).
void pairlist_domain(const V4f* pos, size_t n, const Box& box, float bmin[3], float bmax[3]) {
	float lo[3], hi[3];
	for (int a = 0; a < 3; ++a) {
		lo[a] = std::numeric_limits<float>::max();
		hi[a] = std::numeric_limits<float>::lowest();
	}
	for (size_t i = 0; i < n; ++i) {
		const float c[3] = {pos[i].x, pos[i].y, pos[i].z};
		for (int a = 0; a < 3; ++a) {
			lo[a] = std::min(lo[a], c[a]);
			hi[a] = std::max(hi[a], c[a]);
		}
	}
	for (int a = 0; a < 3; ++a) {
		const float margin = (hi[a] - lo[a]) * 0.01f;
		lo[a] -= margin;
		hi[a] += margin;
	}
	const float kMinExtent = 1e-4f;
	for (int a = 0; a < 3; ++a) {
		const float range = hi[a] - lo[a];
		if (range < kMinExtent) {
			const float center = float(double(lo[a]) + double(range) * 0.5);
			lo[a] = float(double(center) - double(kMinExtent) * 0.5);
			hi[a] = float(double(center) + double(kMinExtent) * 0.5);
		}
	}
	for (int a = 0; a < 3; ++a) {
		if (box.periodic[a]) {
			lo[a] = box.origin[a];
			hi[a] = box.origin[a] + box.L[a];
		}
		bmin[a] = lo[a];
		bmax[a] = hi[a];
	}
}

struct CellGrid {
	int m = 0;
	int r[3] = {1, 1, 1};
	long long slots = 1;
};

// ZOrderPairlist::select_cell_grid (ZOrderPairlist.cpp:24-81).
CellGrid select_cell_grid(float cutoff, size_t n, const float extent_f[3], uint64_t max_pairs) {
	CellGrid g;
	const double extent[3] = {double(extent_f[0]), double(extent_f[1]), double(extent_f[2])};
	if (cutoff <= 0.0f || extent[0] <= 0.0 || extent[1] <= 0.0 || extent[2] <= 0.0 || n == 0)
		return g;
	bool have = false;
	double best_cost = 0.0;
	for (int m = 0; m <= 10; ++m) {
		const long long cells_per_dim = 1LL << m;
		const size_t num_cells = size_t(1) << (3 * m);
		int r[3];
		long long span[3];
		double volume = 1.0;
		for (int a = 0; a < 3; ++a) {
			const double side = extent[a] / double(cells_per_dim);
			r[a] = int(std::ceil(double(cutoff) / side));
			span[a] = std::min(2LL * r[a] + 1LL, cells_per_dim);
			volume *= double(span[a]) * side;
		}
		const long long slots = span[0] * span[1] * span[2];
		if (m > 0 && num_cells > n)
			break;
		const size_t bytes = num_cells * (size_t(slots) * 4 + 2 * 4);
		if (m > 0 && bytes > size_t(max_pairs) * 8)
			break;
		const double cost = volume * double(slots);
		if (!have || cost < best_cost) {
			have = true;
			best_cost = cost;
			g.m = m;
			g.r[0] = r[0];
			g.r[1] = r[1];
			g.r[2] = r[2];
			g.slots = slots;
		}
	}
	return g;
}

// BuildCellNeighborsKernel (ZOrderNeighbor.h:55-139).
void cell_neighbors(const CellGrid& g, const bool per[3], std::vector<uint32_t>& out) {
	const size_t num_cells = size_t(1) << (3 * g.m);
	out.assign(num_cells * size_t(g.slots), 0xFFFFFFFFu);
	const int n = 1 << g.m;
	const uint32_t mask = uint32_t(n - 1);
	auto axis_range = [&](bool p, int r, int& lo, int& hi) {
		if (p && 2 * r + 1 >= n) {
			lo = 0;
			hi = n - 1;
		} else {
			lo = -r;
			hi = r;
		}
	};
	int xl, xh, yl, yh, zl, zh;
	axis_range(per[0], g.r[0], xl, xh);
	axis_range(per[1], g.r[1], yl, yh);
	axis_range(per[2], g.r[2], zl, zh);
	for (size_t c = 0; c < num_cells; ++c) {
		const uint32_t cell = uint32_t(c);
		const uint32_t cx = compact_by3(cell >> 2), cy = compact_by3(cell >> 1), cz = compact_by3(cell);
		uint32_t* row = out.data() + c * size_t(g.slots);
		int k = 0;
		for (int dx = xl; dx <= xh; ++dx) {
			int nx = int(cx) + dx;
			if (per[0])
				nx = int(uint32_t(nx + n) & mask);
			else if (nx < 0 || nx >= n)
				continue;
			const uint32_t mx = split_by3(uint32_t(nx)) << 2;
			for (int dy = yl; dy <= yh; ++dy) {
				int ny = int(cy) + dy;
				if (per[1])
					ny = int(uint32_t(ny + n) & mask);
				else if (ny < 0 || ny >= n)
					continue;
				const uint32_t mxy = mx | (split_by3(uint32_t(ny)) << 1);
				for (int dz = zl; dz <= zh; ++dz) {
					int nz = int(cz) + dz;
					if (per[2])
						nz = int(uint32_t(nz + n) & mask);
					else if (nz < 0 || nz >= n)
						continue;
					row[k++] = mxy | split_by3(uint32_t(nz));
				}
			}
		}
	}
}

struct ExclCSR {
	std::vector<int32_t> offsets, excluded;
	uint64_t num_particles = 0;
	const int32_t* body = nullptr;
	int row_begin(int p) const {
		return (p >= 0 && uint64_t(p) < num_particles) ? offsets[size_t(p)] : 0;
	}
	int row_end(int p) const {
		return (p >= 0 && uint64_t(p) < num_particles) ? offsets[size_t(p) + 1] : 0;
	}
};

// DeviceBondedInteractions::copy_from_host exclusion CSR (DeviceBondedInteraction.h:111-167).
ExclCSR build_exclusion_csr(std::vector<std::pair<int, int>> pairs) {
	ExclCSR csr;
	if (pairs.empty())
		return csr;
	int max_idx = 0;
	for (auto& p : pairs) {
		if (p.first > p.second)
			std::swap(p.first, p.second);
		max_idx = std::max(max_idx, p.second);
	}
	std::sort(pairs.begin(), pairs.end());
	csr.num_particles = uint64_t(max_idx) + 1;
	csr.offsets.assign(csr.num_particles + 1, 0);
	for (auto& p : pairs) {
		csr.offsets[size_t(p.first) + 1]++;
		csr.offsets[size_t(p.second) + 1]++;
	}
	for (size_t i = 0; i < csr.num_particles; ++i)
		csr.offsets[i + 1] += csr.offsets[i];
	csr.excluded.assign(size_t(csr.offsets[csr.num_particles]), 0);
	std::vector<int32_t> cursor(csr.offsets.begin(), csr.offsets.end() - 1);
	for (auto& p : pairs) {
		csr.excluded[size_t(cursor[size_t(p.first)]++)] = p.second;
		csr.excluded[size_t(cursor[size_t(p.second)]++)] = p.first;
	}
	return csr;
}

struct PairlistResult {
	std::vector<uint32_t> codes, sorted_idx, inverse;
	std::vector<std::pair<int, int>> pairs; // (min, max) canonical, sorted
	CellGrid grid;
	float bmin[3], bmax[3];
};

// Exclusion test of ZOrderCellNeighborKernel: the emitting (lower sorted index)
// particle's CSR row and rigid body only (ZOrderNeighbor.h:188-213).
inline bool excluded_by(const marsmd_exclusion_view& ex, int a, int b) {
	const int body = ex.rigid_body_id ? ex.rigid_body_id[a] : -1;
	if (body >= 0 && ex.rigid_body_id[b] == body)
		return true;
	if (ex.offsets && a >= 0 && uint64_t(a) < ex.num_particles) {
		for (int e = ex.offsets[a]; e < ex.offsets[a + 1]; ++e)
			if (ex.excluded[e] == b)
				return true;
	}
	return false;
}

// ZOrderPairlist::build_pairlist (ZOrderPairlist.cpp:102-178) with the stencil walk
// of ZOrderCellNeighborKernel (ZOrderNeighbor.h:174-224), single-threaded. The
// emission order of the reference is nondeterministic, so the result is the set.
void build_pairlist(const V4f* pos, size_t n, const Box& box, float pairlist_cutoff,
					const marsmd_exclusion_view& ex, uint64_t max_pairs, PairlistResult& out,
					bool brute_force) {
	pairlist_domain(pos, n, box, out.bmin, out.bmax);
	out.codes.resize(n);
	out.sorted_idx.resize(n);
	out.inverse.resize(n);
	std::vector<uint32_t> keys(n);
	for (size_t i = 0; i < n; ++i)
		keys[i] = morton_encode(pos[i], out.bmin, out.bmax);
	std::iota(out.sorted_idx.begin(), out.sorted_idx.end(), 0u);
	std::stable_sort(out.sorted_idx.begin(), out.sorted_idx.end(),
					 [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
	for (size_t i = 0; i < n; ++i) {
		out.codes[i] = keys[out.sorted_idx[i]];
		out.inverse[out.sorted_idx[i]] = uint32_t(i);
	}
	std::vector<V4f> sp(n);
	for (size_t i = 0; i < n; ++i)
		sp[i] = pos[out.sorted_idx[i]];
	const float cutoff_squared = pairlist_cutoff * pairlist_cutoff;
	out.pairs.clear();
	auto test = [&](size_t i, size_t j) {
		const V4f dr = wrap_diff<float>(box, sub<float>(sp[j], sp[i]));
		const float d2 = length2<float>(dr);
		if (d2 <= cutoff_squared) {
			const int a = int(out.sorted_idx[i]);
			const int b = int(out.sorted_idx[j]);
			if (excluded_by(ex, a, b))
				return;
			out.pairs.emplace_back(std::min(a, b), std::max(a, b));
		}
	};
	if (brute_force) {
		for (size_t i = 0; i < n; ++i)
			for (size_t j = i + 1; j < n; ++j)
				test(i, j);
	} else {
		const float extent[3] = {out.bmax[0] - out.bmin[0], out.bmax[1] - out.bmin[1],
								 out.bmax[2] - out.bmin[2]};
		const float cutoff = std::sqrt(cutoff_squared);
		out.grid = select_cell_grid(cutoff, n, extent, max_pairs);
		const CellGrid& g = out.grid;
		const size_t num_cells = size_t(1) << (3 * g.m);
		std::vector<uint32_t> begin(num_cells, 0), end(num_cells, 0), nbr;
		const int shift = 3 * (10 - g.m);
		for (size_t i = 0; i < n; ++i) { // BuildCellRangesKernel (ZOrderNeighbor.h:27-43)
			const uint32_t c = out.codes[i] >> shift;
			if (i == 0 || (out.codes[i - 1] >> shift) != c)
				begin[c] = uint32_t(i);
			if (i + 1 == n || (out.codes[i + 1] >> shift) != c)
				end[c] = uint32_t(i + 1);
		}
		const bool per[3] = {box.periodic[0] && box.L[0] > 0, box.periodic[1] && box.L[1] > 0,
							 box.periodic[2] && box.L[2] > 0};
		cell_neighbors(g, per, nbr);
		for (size_t i = 0; i < n; ++i) {
			const uint32_t cell = out.codes[i] >> shift;
			const uint32_t* row = nbr.data() + size_t(cell) * size_t(g.slots);
			for (long long k = 0; k < g.slots; ++k) {
				const uint32_t nc = row[k];
				if (nc == 0xFFFFFFFFu)
					continue;
				const uint32_t lo = std::max<uint32_t>(begin[nc], uint32_t(i) + 1u);
				for (uint32_t j = lo; j < end[nc]; ++j)
					test(i, j);
			}
		}
	}
	std::sort(out.pairs.begin(), out.pairs.end());
}

// ------------------------------------------------------------- simulation state

struct Sim {
	uint64_t capacity = 0, patch_id = 0, base_seed = 0;
	Box box;
	marsmd_particle_types types{};
	std::vector<uint32_t> term_matrix;
	uint32_t enabled_terms = MARSMD_PAIR_TERM_TABULATED;
	std::vector<Table<float>> nb_tables, bond_tables, angle_tables;
	std::vector<Table<double>> nb_tables64;
	PairCoef<float> coef;
	std::vector<std::pair<int, int>> bonds;
	std::vector<int32_t> bond_table;
	std::vector<std::array<int, 3>> angles;
	std::vector<int32_t> angle_table;
	std::vector<Table<float>> dihedral_tables;
	std::vector<std::array<int, 4>> dihedrals;
	std::vector<int32_t> dihedral_table;
	std::vector<int32_t> restraint_ids;
	std::vector<V4f> restraint_anchors;
	std::vector<float> restraint_k;
	std::vector<std::pair<int, int>> excl_pairs;
	std::vector<std::vector<float>> grid_data;
	std::vector<Grid<float>> grids;
	// particle state (canonical slots)
	size_t n = 0;
	std::vector<V4f> pos, mom, force, pos_at_build;
	std::vector<V4<double>> force64;
	std::vector<int32_t> type, id;
	// patch state
	bool built = false, force_rebuild = false, deferred_kick = false, bonded_prepared = false;
	uint64_t max_pairs = 0, max_pairs_ceiling = ~uint64_t(0);
	PairlistResult pl;
	int interp = MARSMD_INTERP_LINEAR;
	float efield[3] = {0, 0, 0};
	// > 1: the OpenMP timing mode of the CPU baseline (per-particle loops split over
	// threads, pair forces into per-thread buffers summed in thread order); the FP64
	// force twin is not kept in that mode, and the FP32 pair sums round differently
	int threads = 1;
	std::vector<std::vector<V4f>> thread_force;
	bool has_grids() const { return !grids.empty(); }
};

// compute_position_dependent_force (src/Interactions/Nonbonded/Pmf.h:68-124).
V4f position_dependent_force(const Sim& s, const V4f& p, int type) {
	const marsmd_particle_types& t = s.types;
	const float charge = t.charge[type];
	V4f f{charge * s.efield[0], charge * s.efield[1], charge * s.efield[2], 0.0f};
	if (!s.has_grids())
		return f;
	const int b = t.pmf_grid_offset[type];
	const int e = b + t.pmf_grid_count[type];
	for (int k = b; k < e; ++k) {
		const marsmd_grid_term term = t.pmf_grid_terms[k];
		if (term.grid_id < 0)
			continue;
		const Grid<float>& g = s.grids[size_t(term.grid_id)];
		const int bc = term.boundary_condition >= 0 ? term.boundary_condition : g.bc;
		const GridSample<float> smp =
			s.interp == MARSMD_INTERP_LINEAR ? sample_linear<float>(g, p, bc) : sample_cubic<float>(g, p, bc);
		const float ns = -term.scale;
		f.x += smp.grad[0] * ns;
		f.y += smp.grad[1] * ns;
		f.z += smp.grad[2] * ns;
		f.t += term.scale * smp.value;
	}
	for (int a = 0; a < 3; ++a) {
		const int gid = t.force_grid_id[type][a];
		if (gid < 0)
			continue;
		const Grid<float>& g = s.grids[size_t(gid)];
		const float v = s.interp == MARSMD_INTERP_LINEAR ? sample_linear<float>(g, p, g.bc).value
														  : sample_cubic<float>(g, p, g.bc).value;
		(&f.x)[a] += t.force_grid_scale[type][a] * v;
	}
	return f;
}

// The pair terms of sim_pair_forces with one FP32 force buffer per thread, each thread
// taking a contiguous share of the pairs; the buffers are summed in thread order, so
// the result does not depend on scheduling.
void sim_pair_forces_threads(Sim& s, float cutoff, bool compute_energy) {
	const float cutoff2 = cutoff > 0.0f ? cutoff * cutoff : 0.0f;
	const int T = s.types.num_types;
	const int nt = s.threads;
	s.thread_force.resize(size_t(nt));
	const size_t P = s.pl.pairs.size();
#pragma omp parallel num_threads(nt)
	{
		const int t = omp_get_thread_num();
		std::vector<V4f>& f = s.thread_force[size_t(t)];
		f.assign(s.n, V4f{});
		const size_t b = P * size_t(t) / size_t(nt), e = P * size_t(t + 1) / size_t(nt);
		for (size_t k = b; k < e; ++k) {
			const int x = s.pl.pairs[k].first, y = s.pl.pairs[k].second;
			const uint32_t tag = s.term_matrix[size_t(s.type[x] * T + s.type[y])];
			const uint32_t terms = tag & s.enabled_terms;
			if (!terms)
				continue;
			const PairTerm<float> pt = pair_term<float>(
				s.box, s.pos[x], s.pos[y],
				(terms & MARSMD_PAIR_TERM_TABULATED) ? &s.nb_tables[tag >> MARSMD_PAIR_TABLE_SHIFT] : nullptr, cutoff2,
				terms, s.type[x], s.type[y], &s.coef);
			if (!pt.valid)
				continue;
			f[size_t(x)].x += -pt.force.x;
			f[size_t(x)].y += -pt.force.y;
			f[size_t(x)].z += -pt.force.z;
			f[size_t(y)].x += pt.force.x;
			f[size_t(y)].y += pt.force.y;
			f[size_t(y)].z += pt.force.z;
			if (compute_energy) {
				f[size_t(x)].t += pt.half_energy;
				f[size_t(y)].t += pt.half_energy;
			}
		}
	}
#pragma omp parallel for num_threads(nt)
	for (size_t i = 0; i < s.n; ++i)
		for (int t = 0; t < nt; ++t) {
			const V4f& f = s.thread_force[size_t(t)][i];
			s.force[i].x += f.x;
			s.force[i].y += f.y;
			s.force[i].z += f.z;
			s.force[i].t += f.t;
		}
}

void sim_pair_forces(Sim& s, float cutoff, bool compute_energy) {
	if (s.threads > 1) {
		sim_pair_forces_threads(s, cutoff, compute_energy);
		return;
	}
	const float cutoff2 = cutoff > 0.0f ? cutoff * cutoff : 0.0f;
	const int T = s.types.num_types;
	for (const auto& pr : s.pl.pairs) {
		const int x = pr.first, y = pr.second;
		const uint32_t tag = s.term_matrix[size_t(s.type[x] * T + s.type[y])];
		const uint32_t terms = tag & s.enabled_terms;
		if (!terms)
			continue;
		const PairTerm<float> pt = pair_term<float>(
			s.box, s.pos[x], s.pos[y],
			(terms & MARSMD_PAIR_TERM_TABULATED) ? &s.nb_tables[tag >> MARSMD_PAIR_TABLE_SHIFT] : nullptr, cutoff2,
			terms, s.type[x], s.type[y], &s.coef);
		if (!pt.valid)
			continue;
		s.force[x].x += -pt.force.x;
		s.force[x].y += -pt.force.y;
		s.force[x].z += -pt.force.z;
		s.force[y].x += pt.force.x;
		s.force[y].y += pt.force.y;
		s.force[y].z += pt.force.z;
		s.force64[x].x -= double(pt.force.x);
		s.force64[x].y -= double(pt.force.y);
		s.force64[x].z -= double(pt.force.z);
		s.force64[y].x += double(pt.force.x);
		s.force64[y].y += double(pt.force.y);
		s.force64[y].z += double(pt.force.z);
		if (compute_energy) {
			s.force[x].t += pt.half_energy;
			s.force[y].t += pt.half_energy;
			s.force64[x].t += double(pt.half_energy);
			s.force64[y].t += double(pt.half_energy);
		}
	}
}

void sim_dihedral_restraint_forces(Sim& s, bool compute_energy);

void sim_bonded_forces(Sim& s, bool compute_energy) {
	for (size_t b = 0; b < s.bonds.size(); ++b) {
		const int x = s.bonds[b].first, y = s.bonds[b].second;
		const PairTerm<float> pt =
			bond_term<float>(s.box, s.pos[x], s.pos[y], s.bond_tables[size_t(s.bond_table[b])]);
		if (!pt.valid)
			continue;
		s.force[x].x += -pt.force.x;
		s.force[x].y += -pt.force.y;
		s.force[x].z += -pt.force.z;
		s.force[y].x += pt.force.x;
		s.force[y].y += pt.force.y;
		s.force[y].z += pt.force.z;
		s.force64[x].x -= pt.force.x;
		s.force64[x].y -= pt.force.y;
		s.force64[x].z -= pt.force.z;
		s.force64[y].x += pt.force.x;
		s.force64[y].y += pt.force.y;
		s.force64[y].z += pt.force.z;
		if (compute_energy) {
			s.force[x].t += pt.half_energy;
			s.force[y].t += pt.half_energy;
			s.force64[x].t += pt.half_energy;
			s.force64[y].t += pt.half_energy;
		}
	}
	for (size_t a = 0; a < s.angles.size(); ++a) {
		const int i = s.angles[a][0], j = s.angles[a][1], k = s.angles[a][2];
		const AngleTerm<float> at = angle_term<float>(s.box, s.pos[i], s.pos[j], s.pos[k],
													   s.angle_tables[size_t(s.angle_table[a])]);
		const V4f* fs[3] = {&at.f1, &at.f2, &at.f3};
		const int ids[3] = {i, j, k};
		for (int q = 0; q < 3; ++q) {
			s.force[ids[q]].x += fs[q]->x;
			s.force[ids[q]].y += fs[q]->y;
			s.force[ids[q]].z += fs[q]->z;
			s.force64[ids[q]].x += fs[q]->x;
			s.force64[ids[q]].y += fs[q]->y;
			s.force64[ids[q]].z += fs[q]->z;
			if (compute_energy) {
				s.force[ids[q]].t += at.third_energy;
				s.force64[ids[q]].t += at.third_energy;
			}
		}
	}
	sim_dihedral_restraint_forces(s, compute_energy);
}

// Dihedrals and harmonic restraints of Patch::calculate_bonded_forces (Patch.cpp:208-236).
void sim_dihedral_restraint_forces(Sim& s, bool compute_energy) {
	for (size_t d = 0; d < s.dihedrals.size(); ++d) {
		const auto& q = s.dihedrals[d];
		const DihedralTerm<float> dt = dihedral_term<float>(s.box, s.pos[size_t(q[0])], s.pos[size_t(q[1])],
															 s.pos[size_t(q[2])], s.pos[size_t(q[3])],
															 s.dihedral_tables[size_t(s.dihedral_table[d])]);
		for (int k = 0; k < 4; ++k) {
			const size_t id = size_t(q[size_t(k)]);
			s.force[id].x += dt.f[k].x;
			s.force[id].y += dt.f[k].y;
			s.force[id].z += dt.f[k].z;
			s.force64[id].x += dt.f[k].x;
			s.force64[id].y += dt.f[k].y;
			s.force64[id].z += dt.f[k].z;
			if (compute_energy) {
				s.force[id].t += dt.quarter_energy;
				s.force64[id].t += dt.quarter_energy;
			}
		}
	}
	for (size_t r = 0; r < s.restraint_ids.size(); ++r) {
		const int id = s.restraint_ids[r];
		if (id < 0)
			continue;
		const RestraintTerm<float> rt = restraint_term<float>(s.box, s.pos[size_t(id)], s.restraint_anchors[r],
															   s.restraint_k[r]);
		s.force[size_t(id)].x += rt.f.x;
		s.force[size_t(id)].y += rt.f.y;
		s.force[size_t(id)].z += rt.f.z;
		s.force64[size_t(id)].x += rt.f.x;
		s.force64[size_t(id)].y += rt.f.y;
		s.force64[size_t(id)].z += rt.f.z;
		if (compute_energy) {
			s.force[size_t(id)].t += rt.energy;
			s.force64[size_t(id)].t += rt.energy;
		}
	}
}

// DisplacementKernel::min_image (AdaptiveKernels.h:58-66).
float min_image(float d, float L) {
	if (L <= 0.0f)
		return d;
	while (d > 0.5f * L)
		d -= L;
	while (d < -0.5f * L)
		d += L;
	return d;
}
float max_displacement(const Sim& s) { // ZOrderSort::compute_max_displacement
	float m = 0.0f;
	for (size_t i = 0; i < s.n; ++i) {
		const V4f raw = sub<float>(s.pos[i], s.pos_at_build[i]);
		const float L[3] = {s.box.periodic[0] ? s.box.L[0] : 0.0f,
							s.box.periodic[1] ? s.box.L[1] : 0.0f,
							s.box.periodic[2] ? s.box.L[2] : 0.0f};
		const V4f dr{min_image(raw.x, L[0]), min_image(raw.y, L[1]), min_image(raw.z, L[2]), 0.0f};
		const float d2 = length2<float>(dr);
		if (d2 > m)
			m = d2;
	}
	return std::sqrt(m);
}

marsmd_exclusion_view view_of(const ExclCSR& c) {
	marsmd_exclusion_view v{};
	if (c.num_particles) {
		v.offsets = c.offsets.data();
		v.excluded = c.excluded.data();
		v.num_particles = c.num_particles;
	}
	return v;
}

// Patch::reorder_particles (Patch.cpp:456-480): System-mode sort over the box,
// permute the SoA, remap bonded indices, rebuild the exclusion CSR, force a rebuild.
void sim_reorder(Sim& s) {
	if (s.n == 0)
		return;
	float bmin[3], bmax[3];
	for (int a = 0; a < 3; ++a) {
		bmin[a] = s.box.origin[a];
		bmax[a] = s.box.origin[a] + s.box.L[a];
		if (!(bmax[a] > bmin[a]))
			return;
	}
	std::vector<uint32_t> keys(s.n), order(s.n), inv(s.n);
	for (size_t i = 0; i < s.n; ++i)
		keys[i] = morton_encode(s.pos[i], bmin, bmax);
	std::iota(order.begin(), order.end(), 0u);
	std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
	for (size_t i = 0; i < s.n; ++i)
		inv[order[i]] = uint32_t(i);
	auto permute = [&](auto& v) {
		auto old = v;
		for (size_t i = 0; i < s.n; ++i)
			v[i] = old[order[i]];
	};
	permute(s.pos);
	permute(s.mom);
	permute(s.type);
	permute(s.id);
	for (auto& b : s.bonds) {
		b.first = int(inv[size_t(b.first)]);
		b.second = int(inv[size_t(b.second)]);
	}
	for (auto& a : s.angles)
		for (int q = 0; q < 3; ++q)
			a[q] = int(inv[size_t(a[q])]);
	for (auto& d : s.dihedrals)
		for (int q = 0; q < 4; ++q)
			d[size_t(q)] = int(inv[size_t(d[size_t(q)])]);
	for (auto& r : s.restraint_ids)
		if (r >= 0)
			r = int(inv[size_t(r)]);
	for (auto& e : s.excl_pairs) {
		e.first = int(inv[size_t(e.first)]);
		e.second = int(inv[size_t(e.second)]);
	}
	s.force_rebuild = true;
}

void sim_calculate_nonbonded(Sim& s, float pairlist_cutoff, float cutoff, uint64_t step,
							 uint64_t rebuild_period, const float ef[3], int interp,
							 bool compute_energy) {
	std::fill(s.force.begin(), s.force.end(), V4f{});
	std::fill(s.force64.begin(), s.force64.end(), V4<double>{});
	if (s.n == 0)
		return;
	s.bonded_prepared = true;
	s.interp = interp;
	for (int a = 0; a < 3; ++a)
		s.efield[a] = ef[a];
	const bool at_period = (rebuild_period == 0 || (step - 1) % rebuild_period == 0);
	bool rebuild = at_period || !s.built;
	const bool forced = s.force_rebuild;
	s.force_rebuild = false;
	rebuild = rebuild || forced;
	const float skin = pairlist_cutoff - cutoff;
	if (rebuild && !forced && rebuild_period != 0 && s.built && skin > 0.0f)
		rebuild = max_displacement(s) > skin * 0.5f;
	if (rebuild) {
		const ExclCSR csr = build_exclusion_csr(s.excl_pairs);
		build_pairlist(s.pos.data(), s.n, s.box, pairlist_cutoff, view_of(csr), s.max_pairs, s.pl, false);
		// Pairlist::grow_pair_capacity (Pairlist.h:244-251)
		const uint64_t needed = s.pl.pairs.size();
		if (needed > s.max_pairs) {
			if (needed > s.max_pairs_ceiling)
				throw std::runtime_error("Pairlist capacity exceeded");
			s.max_pairs = std::min<uint64_t>(s.max_pairs_ceiling, needed + needed / 4);
		}
		s.built = true;
		s.pos_at_build = s.pos;
	}
	sim_pair_forces(s, cutoff, compute_energy);
}

void sim_integrate(Sim& s, float dt, float kT, int integrator, uint64_t step) {
	const uint64_t base_seed = mix_seed(s.base_seed, s.patch_id, step);
	const uint32_t base_ctr = 0;
	const marsmd_particle_types& t = s.types;
	const int nt = s.threads;
	if (integrator == MARSMD_INTEGRATOR_BROWNIAN) {
#pragma omp parallel for num_threads(nt) if (nt > 1)
		for (size_t i = 0; i < s.n; ++i) { // BDIntegrate::operator() (BD.h:44-92)
			V4f p = s.pos[i];
			V4f f = s.force[i];
			const int ty = s.type[i];
			const V4f pmf = position_dependent_force(s, p, ty);
			f.x += pmf.x;
			f.y += pmf.y;
			f.z += pmf.z;
			s.force[i].t += pmf.t;
			s.force64[i].t += pmf.t;
			const float* D = t.diffusion[ty];
			Philox rng(base_seed, base_ctr + uint32_t(i), uint32_t(step), 0x5324120u);
			uint32_t raw[4];
			rng.generate(raw);
			const float u[4] = {u01(raw[0]), u01(raw[1]), u01(raw[2]), u01(raw[3])};
			float rn[3];
			box_muller(u, rn);
			p.x += (D[0] / kT) * f.x * dt + std::sqrt(2.0f * D[0] * dt) * rn[0];
			p.y += (D[1] / kT) * f.y * dt + std::sqrt(2.0f * D[1] * dt) * rn[1];
			p.z += (D[2] / kT) * f.z * dt + std::sqrt(2.0f * D[2] * dt) * rn[2];
			p = wrap<float>(s.box, p);
			s.pos[i].x = p.x;
			s.pos[i].y = p.y;
			s.pos[i].z = p.z;
		}
		return;
	}
	// Langevin (Patch.cpp:290-342): PMF, deferred closing kick, then B-A-O-A.
	if (s.has_grids()) {
#pragma omp parallel for num_threads(nt) if (nt > 1)
		for (size_t i = 0; i < s.n; ++i) { // ComputePMFKernel (Pmf.h:131-151)
			const V4f f = position_dependent_force(s, s.pos[i], s.type[i]);
			s.force[i].x += f.x;
			s.force[i].y += f.y;
			s.force[i].z += f.z;
			s.force[i].t += f.t;
			s.force64[i].x += f.x;
			s.force64[i].y += f.y;
			s.force64[i].z += f.z;
			s.force64[i].t += f.t;
		}
	}
	const float FCF = 4.18679994e4f;
	const float SQ = 2.046167337f;
	const float half_dt = 0.5f * dt;
	if (s.deferred_kick) { // BAOAB_LastUpdate (BAOAB.h:151-172)
#pragma omp parallel for num_threads(nt) if (nt > 1)
		for (size_t i = 0; i < s.n; ++i) {
			s.mom[i].x += half_dt * s.force[i].x * FCF;
			s.mom[i].y += half_dt * s.force[i].y * FCF;
			s.mom[i].z += half_dt * s.force[i].z * FCF;
		}
		s.deferred_kick = false;
	}
	s.deferred_kick = true;
#pragma omp parallel for num_threads(nt) if (nt > 1)
	for (size_t i = 0; i < s.n; ++i) { // BAOABIntegrate::operator() (BAOAB.h:46-120)
		V4f p = s.pos[i], m = s.mom[i];
		const V4f f = s.force[i];
		const int ty = s.type[i];
		const float mass = t.mass[ty];
		const float* g = t.trans_damping[ty];
		m.x += half_dt * f.x * FCF;
		m.y += half_dt * f.y * FCF;
		m.z += half_dt * f.z * FCF;
		const float inv_mass = 1.0f / mass;
		p.x += half_dt * m.x * inv_mass * 10000.0f;
		p.y += half_dt * m.y * inv_mass * 10000.0f;
		p.z += half_dt * m.z * inv_mass * 10000.0f;
		const float c[3] = {std::exp(-g[0] * dt), std::exp(-g[1] * dt), std::exp(-g[2] * dt)};
		float ns[3];
		for (int a = 0; a < 3; ++a)
			ns[a] = std::sqrt(kT * mass * (1.0f - c[a] * c[a])) * SQ;
		Philox rng(base_seed, base_ctr + uint32_t(i), uint32_t(step), 0x1356914u);
		uint32_t raw[4];
		rng.generate(raw);
		const float u[4] = {u01(raw[0]), u01(raw[1]), u01(raw[2]), u01(raw[3])};
		float rn[3];
		box_muller(u, rn);
		m.x = c[0] * m.x + ns[0] * rn[0];
		m.y = c[1] * m.y + ns[1] * rn[1];
		m.z = c[2] * m.z + ns[2] * rn[2];
		p.x += half_dt * m.x * inv_mass * 10000.0f;
		p.y += half_dt * m.y * inv_mass * 10000.0f;
		p.z += half_dt * m.z * inv_mass * 10000.0f;
		p = wrap<float>(s.box, p);
		s.pos[i].x = p.x;
		s.pos[i].y = p.y;
		s.pos[i].z = p.z;
		s.mom[i].x = m.x;
		s.mom[i].y = m.y;
		s.mom[i].z = m.z;
	}
}

void sim_finish_deferred_kick(Sim& s, float dt) {
	if (!s.deferred_kick)
		return;
	s.deferred_kick = false;
	const float FCF = 4.18679994e4f;
	const float half_dt = 0.5f * dt;
	for (size_t i = 0; i < s.n; ++i) {
		s.mom[i].x += half_dt * s.force[i].x * FCF;
		s.mom[i].y += half_dt * s.force[i].y * FCF;
		s.mom[i].z += half_dt * s.force[i].z * FCF;
	}
}

} // namespace

// ===================================================================== C API
extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

// ---- RNG
void orc_philox_draw(uint64_t seed, uint32_t ctr0, uint32_t global_seed, uint32_t ctr1,
					 uint32_t draws, uint32_t* out) {
	Philox p(seed, ctr0, global_seed, ctr1);
	for (uint32_t d = 0; d < draws; ++d)
		p.generate(out + 4 * d);
}
float orc_u01(uint32_t x) { return u01(x); }
void orc_box_muller(const float u[4], float out[3]) { box_muller(u, out); }
uint64_t orc_mix_seed(uint64_t seed, uint64_t patch_id, uint64_t step) {
	return mix_seed(seed, patch_id, step);
}

// ---- Morton / sort
void orc_morton_encode(const float* pos4, uint64_t n, const float bmin[3], const float bmax[3],
					   uint32_t* codes) {
	const V4f* p = as_v4(pos4);
	for (uint64_t i = 0; i < n; ++i)
		codes[i] = morton_encode(p[i], bmin, bmax);
}
uint32_t orc_morton_from_coords(uint32_t x, uint32_t y, uint32_t z) { return morton_from_coords(x, y, z); }
void orc_morton_decode(uint32_t code, uint32_t out[3]) {
	out[0] = compact_by3(code >> 2);
	out[1] = compact_by3(code >> 1);
	out[2] = compact_by3(code);
}
// Stable sort of (key, index) — the semantics of ZOrderSort::sort_morton_codes.
void orc_stable_sort_pairs(const uint32_t* keys, uint64_t n, uint32_t* sorted_keys,
						   uint32_t* sorted_vals) {
	std::vector<uint32_t> idx(n);
	std::iota(idx.begin(), idx.end(), 0u);
	std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
	for (uint64_t i = 0; i < n; ++i) {
		sorted_keys[i] = keys[idx[i]];
		sorted_vals[i] = idx[i];
	}
}

// ---- pairlist
int orc_pairlist_domain(const float* pos4, uint64_t n, const marsmd_box* box, float bmin[3],
						float bmax[3]) {
	pairlist_domain(as_v4(pos4), n, make_box(*box), bmin, bmax);
	return 0;
}
void orc_select_cell_grid(float cutoff, uint64_t n, const float extent[3], uint64_t max_pairs,
						  int32_t* m, int32_t r[3], int32_t* slots) {
	const CellGrid g = select_cell_grid(cutoff, n, extent, max_pairs);
	*m = g.m;
	r[0] = g.r[0];
	r[1] = g.r[1];
	r[2] = g.r[2];
	*slots = int32_t(g.slots);
}
// Cell tables for grid (m, r) and periodicity: begin/end need 8^m entries,
// neighbors 8^m * slots. codes are the sorted Morton codes.
void orc_cell_tables(const uint32_t* sorted_codes, uint64_t n, int32_t m, const int32_t r[3],
					 int32_t slots, const int32_t periodic[3], uint32_t* begin, uint32_t* end,
					 uint32_t* neighbors) {
	CellGrid g;
	g.m = m;
	g.r[0] = r[0];
	g.r[1] = r[1];
	g.r[2] = r[2];
	g.slots = slots;
	const size_t num_cells = size_t(1) << (3 * m);
	std::fill(begin, begin + num_cells, 0u);
	std::fill(end, end + num_cells, 0u);
	const int shift = 3 * (10 - m);
	for (uint64_t i = 0; i < n; ++i) {
		const uint32_t c = sorted_codes[i] >> shift;
		if (i == 0 || (sorted_codes[i - 1] >> shift) != c)
			begin[c] = uint32_t(i);
		if (i + 1 == n || (sorted_codes[i + 1] >> shift) != c)
			end[c] = uint32_t(i + 1);
	}
	const bool per[3] = {periodic[0] != 0, periodic[1] != 0, periodic[2] != 0};
	std::vector<uint32_t> nbr;
	cell_neighbors(g, per, nbr);
	std::copy(nbr.begin(), nbr.end(), neighbors);
}

// Builds the pair set. Returns the number of pairs (or -1 on error); writes at most
// `cap` (x, y) pairs, sorted, canonical indices. sorted_idx/inverse/codes/info may be NULL.
int64_t orc_build_pairlist(const float* pos4, uint64_t n, const marsmd_box* box,
						   float pairlist_cutoff, const marsmd_exclusion_view* exp, uint64_t max_pairs,
						   int32_t brute_force, int32_t* pairs_out, uint64_t cap,
						   uint32_t* codes, uint32_t* sorted_idx, uint32_t* inverse,
						   int32_t* grid_info /* m, rx, ry, rz, slots */) {
	try {
		PairlistResult r;
		const marsmd_exclusion_view ex = exp ? *exp : marsmd_exclusion_view{};
		build_pairlist(as_v4(pos4), n, make_box(*box), pairlist_cutoff, ex, max_pairs, r,
					   brute_force != 0);
		const uint64_t np = r.pairs.size();
		for (uint64_t k = 0; k < np && k < cap; ++k) {
			pairs_out[2 * k] = r.pairs[k].first;
			pairs_out[2 * k + 1] = r.pairs[k].second;
		}
		if (codes)
			std::copy(r.codes.begin(), r.codes.end(), codes);
		if (sorted_idx)
			std::copy(r.sorted_idx.begin(), r.sorted_idx.end(), sorted_idx);
		if (inverse)
			std::copy(r.inverse.begin(), r.inverse.end(), inverse);
		if (grid_info) {
			grid_info[0] = r.grid.m;
			grid_info[1] = r.grid.r[0];
			grid_info[2] = r.grid.r[1];
			grid_info[3] = r.grid.r[2];
			grid_info[4] = int32_t(r.grid.slots);
		}
		return int64_t(np);
	} catch (const std::exception& e) {
		fail(e.what());
		return -1;
	}
}

// Exclusion CSR from an exclusion pair list (DeviceBondedInteraction.h:111-167).
// Returns num_particles of the CSR; offsets needs num_particles+1, excluded 2*n_excl.
int64_t orc_exclusion_csr(const int32_t* pairs, uint64_t n_excl, int32_t* offsets,
						  int32_t* excluded) {
	std::vector<std::pair<int, int>> v(n_excl);
	for (uint64_t k = 0; k < n_excl; ++k)
		v[k] = {pairs[2 * k], pairs[2 * k + 1]};
	const ExclCSR c = build_exclusion_csr(v);
	std::copy(c.offsets.begin(), c.offsets.end(), offsets);
	std::copy(c.excluded.begin(), c.excluded.end(), excluded);
	return int64_t(c.num_particles);
}

// ---- forces on explicit arrays
// mode 0: FP32 terms, FP32 sums in pair order; 1: FP32 terms, FP64 sums;
// 2: FP64 twin (all arithmetic in double). out is n x 4 doubles (force xyz, energy).
int orc_pair_forces(const int32_t* pairs, uint64_t num_pairs, const float* pos4,
					const int32_t* types, const uint32_t* term_matrix, int32_t num_types,
					const marsmd_table* tables, int32_t num_tables, const marsmd_box* box,
					float cutoff_squared, uint32_t enabled_terms, int32_t mode, uint64_t n,
					double* out, const marsmd_particle_types* ptypes, const marsmd_pair_params* params) {
	const Box b = make_box(*box);
	const marsmd_particle_types no_types{};
	const PairCoef<float> c32 = make_coef<float>(ptypes ? *ptypes : no_types, params);
	const PairCoef<double> c64 = make_coef<double>(ptypes ? *ptypes : no_types, params);
	std::vector<Table<float>> t32;
	std::vector<Table<double>> t64;
	for (int k = 0; k < num_tables; ++k) {
		t32.push_back(make_table<float>(tables[k]));
		t64.push_back(make_table<double>(tables[k]));
	}
	std::vector<V4f> f32(n);
	std::fill(out, out + 4 * n, 0.0);
	const V4f* p = as_v4(pos4);
	for (uint64_t k = 0; k < num_pairs; ++k) {
		const int x = pairs[2 * k], y = pairs[2 * k + 1];
		const uint32_t tag = term_matrix ? term_matrix[size_t(types[x] * num_types + types[y])] : enabled_terms;
		const uint32_t terms = tag & enabled_terms;
		if (!terms)
			continue;
		const int ti = int(tag >> MARSMD_PAIR_TABLE_SHIFT);
		const bool tab = (terms & MARSMD_PAIR_TERM_TABULATED) != 0;
		if (mode == 2) {
			const V4<double> px{p[x].x, p[x].y, p[x].z, 0}, py{p[y].x, p[y].y, p[y].z, 0};
			const PairTerm<double> pt = pair_term<double>(b, px, py, tab ? &t64[size_t(ti)] : nullptr,
														  double(cutoff_squared), terms, types[x], types[y], &c64);
			if (!pt.valid)
				continue;
			out[4 * x + 0] -= pt.force.x;
			out[4 * x + 1] -= pt.force.y;
			out[4 * x + 2] -= pt.force.z;
			out[4 * x + 3] += pt.half_energy;
			out[4 * y + 0] += pt.force.x;
			out[4 * y + 1] += pt.force.y;
			out[4 * y + 2] += pt.force.z;
			out[4 * y + 3] += pt.half_energy;
			continue;
		}
		const PairTerm<float> pt = pair_term<float>(b, p[x], p[y], tab ? &t32[size_t(ti)] : nullptr, cutoff_squared,
													 terms, types[x], types[y], &c32);
		if (!pt.valid)
			continue;
		if (mode == 1) {
			out[4 * x + 0] -= pt.force.x;
			out[4 * x + 1] -= pt.force.y;
			out[4 * x + 2] -= pt.force.z;
			out[4 * x + 3] += pt.half_energy;
			out[4 * y + 0] += pt.force.x;
			out[4 * y + 1] += pt.force.y;
			out[4 * y + 2] += pt.force.z;
			out[4 * y + 3] += pt.half_energy;
		} else {
			f32[size_t(x)].x += -pt.force.x;
			f32[size_t(x)].y += -pt.force.y;
			f32[size_t(x)].z += -pt.force.z;
			f32[size_t(x)].t += pt.half_energy;
			f32[size_t(y)].x += pt.force.x;
			f32[size_t(y)].y += pt.force.y;
			f32[size_t(y)].z += pt.force.z;
			f32[size_t(y)].t += pt.half_energy;
		}
	}
	if (mode == 0)
		for (uint64_t i = 0; i < n; ++i) {
			out[4 * i + 0] = f32[i].x;
			out[4 * i + 1] = f32[i].y;
			out[4 * i + 2] = f32[i].z;
			out[4 * i + 3] = f32[i].t;
		}
	return 0;
}

// Single-table lookup (TabulatedPotential::compute), FP32 and FP64.
void orc_table_compute(const marsmd_table* t, float x, float out[2]) {
	const FE<float> fe = table_compute<float>(x, make_table<float>(*t));
	out[0] = fe.f;
	out[1] = fe.e;
}
void orc_table_compute64(const marsmd_table* t, double x, double out[2]) {
	const FE<double> fe = table_compute<double>(x, make_table<double>(*t));
	out[0] = fe.f;
	out[1] = fe.e;
}

// Bonds (int2) and angles (3 ints per angle) on explicit arrays; mode as above (0/1/2).
int orc_bonded_forces(const int32_t* bonds, const int32_t* bond_table, uint64_t num_bonds,
					  const int32_t* angles, const int32_t* angle_table, uint64_t num_angles,
					  const float* pos4, const marsmd_table* btabs, int32_t nbt,
					  const marsmd_table* atabs, int32_t nat, const marsmd_box* box, int32_t mode,
					  uint64_t n, double* out) {
	const Box b = make_box(*box);
	const V4f* p = as_v4(pos4);
	std::fill(out, out + 4 * n, 0.0);
	std::vector<V4f> f32(n);
	auto add = [&](int i, double fx, double fy, double fz, double e, float sfx, float sfy, float sfz, float se) {
		if (mode == 0) {
			f32[size_t(i)].x += sfx;
			f32[size_t(i)].y += sfy;
			f32[size_t(i)].z += sfz;
			f32[size_t(i)].t += se;
		} else {
			out[4 * i + 0] += fx;
			out[4 * i + 1] += fy;
			out[4 * i + 2] += fz;
			out[4 * i + 3] += e;
		}
	};
	for (uint64_t k = 0; k < num_bonds; ++k) {
		const int x = bonds[2 * k], y = bonds[2 * k + 1];
		if (mode == 2) {
			const V4<double> px{p[x].x, p[x].y, p[x].z, 0}, py{p[y].x, p[y].y, p[y].z, 0};
			const PairTerm<double> pt = bond_term<double>(b, px, py, make_table<double>(btabs[bond_table[k]]));
			if (!pt.valid)
				continue;
			add(x, -pt.force.x, -pt.force.y, -pt.force.z, pt.half_energy, 0, 0, 0, 0);
			add(y, pt.force.x, pt.force.y, pt.force.z, pt.half_energy, 0, 0, 0, 0);
			continue;
		}
		const PairTerm<float> pt = bond_term<float>(b, p[x], p[y], make_table<float>(btabs[bond_table[k]]));
		if (!pt.valid)
			continue;
		add(x, -pt.force.x, -pt.force.y, -pt.force.z, pt.half_energy, -pt.force.x, -pt.force.y, -pt.force.z, pt.half_energy);
		add(y, pt.force.x, pt.force.y, pt.force.z, pt.half_energy, pt.force.x, pt.force.y, pt.force.z, pt.half_energy);
	}
	for (uint64_t k = 0; k < num_angles; ++k) {
		const int i = angles[3 * k], j = angles[3 * k + 1], l = angles[3 * k + 2];
		if (mode == 2) {
			const V4<double> pa{p[i].x, p[i].y, p[i].z, 0}, pb{p[j].x, p[j].y, p[j].z, 0},
				pc{p[l].x, p[l].y, p[l].z, 0};
			const AngleTerm<double> at = angle_term<double>(b, pa, pb, pc, make_table<double>(atabs[angle_table[k]]));
			add(i, at.f1.x, at.f1.y, at.f1.z, at.third_energy, 0, 0, 0, 0);
			add(j, at.f2.x, at.f2.y, at.f2.z, at.third_energy, 0, 0, 0, 0);
			add(l, at.f3.x, at.f3.y, at.f3.z, at.third_energy, 0, 0, 0, 0);
			continue;
		}
		const AngleTerm<float> at = angle_term<float>(b, p[i], p[j], p[l], make_table<float>(atabs[angle_table[k]]));
		add(i, at.f1.x, at.f1.y, at.f1.z, at.third_energy, at.f1.x, at.f1.y, at.f1.z, at.third_energy);
		add(j, at.f2.x, at.f2.y, at.f2.z, at.third_energy, at.f2.x, at.f2.y, at.f2.z, at.third_energy);
		add(l, at.f3.x, at.f3.y, at.f3.z, at.third_energy, at.f3.x, at.f3.y, at.f3.z, at.third_energy);
	}
	(void)nbt;
	(void)nat;
	if (mode == 0)
		for (uint64_t q = 0; q < n; ++q) {
			out[4 * q + 0] = f32[q].x;
			out[4 * q + 1] = f32[q].y;
			out[4 * q + 2] = f32[q].z;
			out[4 * q + 3] = f32[q].t;
		}
	return 0;
}

// Grid sample at points; scheme 1 linear / 3 cubic; fp64 != 0 uses the FP64 twin.
// out: n x 4 (value, grad x, grad y, grad z).
void orc_grid_sample(const marsmd_grid_view* g, const float* pos4, uint64_t n, int32_t scheme,
					 int32_t bc, int32_t fp64, double* out) {
	const V4f* p = as_v4(pos4);
	if (fp64) {
		const Grid<double> gd = make_grid<double>(*g);
		for (uint64_t i = 0; i < n; ++i) {
			const V4<double> q{p[i].x, p[i].y, p[i].z, 0};
			const GridSample<double> s = scheme == 1 ? sample_linear<double>(gd, q, bc) : sample_cubic<double>(gd, q, bc);
			out[4 * i] = s.value;
			for (int a = 0; a < 3; ++a)
				out[4 * i + 1 + a] = s.grad[a];
		}
		return;
	}
	const Grid<float> gf = make_grid<float>(*g);
	for (uint64_t i = 0; i < n; ++i) {
		const GridSample<float> s = scheme == 1 ? sample_linear<float>(gf, p[i], bc) : sample_cubic<float>(gf, p[i], bc);
		out[4 * i] = s.value;
		for (int a = 0; a < 3; ++a)
			out[4 * i + 1 + a] = s.grad[a];
	}
}

// ---- simulation (Patch + SimManager loop)
void* orc_sim_create(const marsmd_patch_desc* d) {
	try {
		auto s = std::make_unique<Sim>();
		s->capacity = d->capacity;
		s->patch_id = d->patch_id;
		s->box = make_box(d->box);
		s->types = d->types;
		const int T = d->types.num_types;
		if (d->term_matrix)
			s->term_matrix.assign(d->term_matrix, d->term_matrix + size_t(T) * size_t(T));
		else
			s->term_matrix.assign(size_t(T) * size_t(T), 0u);
		s->enabled_terms = d->enabled_terms;
		s->coef = make_coef<float>(d->types, d->pair_params);
		for (int k = 0; k < d->num_nb_tables; ++k) {
			s->nb_tables.push_back(make_table<float>(d->nb_tables[k]));
			s->nb_tables64.push_back(make_table<double>(d->nb_tables[k]));
		}
		for (int k = 0; k < d->num_bond_tables; ++k)
			s->bond_tables.push_back(make_table<float>(d->bond_tables[k]));
		for (int k = 0; k < d->num_angle_tables; ++k)
			s->angle_tables.push_back(make_table<float>(d->angle_tables[k]));
		for (uint64_t k = 0; k < d->num_bonds; ++k) {
			s->bonds.emplace_back(d->bonds[2 * k], d->bonds[2 * k + 1]);
			s->bond_table.push_back(d->bond_table ? d->bond_table[k] : 0);
		}
		for (uint64_t k = 0; k < d->num_angles; ++k) {
			s->angles.push_back({d->angles[3 * k], d->angles[3 * k + 1], d->angles[3 * k + 2]});
			s->angle_table.push_back(d->angle_table ? d->angle_table[k] : 0);
		}
		for (uint64_t k = 0; k < d->num_exclusions; ++k)
			s->excl_pairs.emplace_back(d->exclusions[2 * k], d->exclusions[2 * k + 1]);
		for (int k = 0; k < d->num_dihedral_tables; ++k)
			s->dihedral_tables.push_back(make_table<float>(d->dihedral_tables[k]));
		for (uint64_t k = 0; k < d->num_dihedrals; ++k) {
			s->dihedrals.push_back({d->dihedrals[4 * k], d->dihedrals[4 * k + 1], d->dihedrals[4 * k + 2],
									d->dihedrals[4 * k + 3]});
			s->dihedral_table.push_back(d->dihedral_table ? d->dihedral_table[k] : 0);
		}
		for (uint64_t k = 0; k < d->num_restraints; ++k) {
			s->restraint_ids.push_back(d->restraint_ids[k]);
			s->restraint_anchors.push_back(V4f{d->restraint_anchors[4 * k], d->restraint_anchors[4 * k + 1],
											   d->restraint_anchors[4 * k + 2], 0.0f});
			s->restraint_k.push_back(d->restraint_k[k]);
		}
		for (int k = 0; k < d->num_grids; ++k) {
			const marsmd_grid_view& g = d->grids[k];
			const size_t cnt = size_t(g.dims[0] * g.dims[1] * g.dims[2]);
			s->grid_data.emplace_back(g.data, g.data + cnt);
		}
		for (int k = 0; k < d->num_grids; ++k) {
			marsmd_grid_view g = d->grids[k];
			g.data = s->grid_data[size_t(k)].data();
			s->grids.push_back(make_grid<float>(g));
		}
		s->max_pairs = std::max<uint64_t>(d->capacity, 1);
		return s.release();
	} catch (const std::exception& e) {
		fail(e.what());
		return nullptr;
	}
}
void orc_sim_destroy(void* h) { delete static_cast<Sim*>(h); }
void orc_sim_set_base_seed(void* h, uint64_t seed) { static_cast<Sim*>(h)->base_seed = seed; }
int orc_sim_set_particles(void* h, const float* pos4, const float* mom4, const int32_t* type, uint64_t n) {
	Sim& s = *static_cast<Sim*>(h);
	s.n = n;
	s.pos.assign(as_v4(pos4), as_v4(pos4) + n);
	if (mom4)
		s.mom.assign(as_v4(mom4), as_v4(mom4) + n);
	else
		s.mom.assign(n, V4f{});
	s.type.assign(type, type + n);
	s.id.resize(n);
	std::iota(s.id.begin(), s.id.end(), 0);
	s.force.assign(n, V4f{});
	s.force64.assign(n, V4<double>{});
	return 0;
}
// Overwrite the particle state (canonical order) and the deferred-kick flag, so a
// single operation can be replayed from another implementation's state.
int orc_sim_set_state(void* h, const float* pos4, const float* mom4, const float* force4, const int32_t* type,
					  const int32_t* ids, uint64_t n, int32_t deferred_kick) {
	Sim& s = *static_cast<Sim*>(h);
	if (n != s.n)
		return fail("set_state: particle count differs");
	for (uint64_t i = 0; i < n; ++i) {
		if (pos4)
			s.pos[i] = as_v4(pos4)[i];
		if (mom4)
			s.mom[i] = as_v4(mom4)[i];
		if (force4) {
			s.force[i] = as_v4(force4)[i];
			s.force64[i] = V4<double>{s.force[i].x, s.force[i].y, s.force[i].z, s.force[i].t};
		}
		if (type)
			s.type[i] = type[i];
		if (ids)
			s.id[i] = ids[i];
	}
	if (deferred_kick >= 0)
		s.deferred_kick = deferred_kick != 0;
	return 0;
}
// force64 (n x 4 doubles): FP32 terms accumulated in double (the FP64 recomputation).
int orc_sim_get_particles(void* h, float* pos4, float* mom4, float* force4, double* force64,
						  int32_t* type, int32_t* ids, uint64_t n) {
	Sim& s = *static_cast<Sim*>(h);
	if (n > s.n)
		return fail("n too large");
	for (uint64_t i = 0; i < n; ++i) {
		if (pos4)
			as_v4(pos4)[i] = s.pos[i];
		if (mom4)
			as_v4(mom4)[i] = s.mom[i];
		if (force4)
			as_v4(force4)[i] = s.force[i];
		if (force64) {
			force64[4 * i] = s.force64[i].x;
			force64[4 * i + 1] = s.force64[i].y;
			force64[4 * i + 2] = s.force64[i].z;
			force64[4 * i + 3] = s.force64[i].t;
		}
		if (type)
			type[i] = s.type[i];
		if (ids)
			ids[i] = s.id[i];
	}
	return 0;
}
int orc_sim_calculate_nonbonded_forces(void* h, float pl_cut, float cut, uint64_t step,
									   uint64_t period, const float ef[3], int32_t interp,
									   int32_t energy) {
	try {
		sim_calculate_nonbonded(*static_cast<Sim*>(h), pl_cut, cut, step, period, ef, interp, energy != 0);
		return 0;
	} catch (const std::exception& e) {
		return fail(e.what());
	}
}
int orc_sim_calculate_bonded_forces(void* h, int32_t energy) {
	sim_bonded_forces(*static_cast<Sim*>(h), energy != 0);
	return 0;
}
int orc_sim_integrate_motion(void* h, float dt, float kT, int32_t integrator, uint64_t step) {
	sim_integrate(*static_cast<Sim*>(h), dt, kT, integrator, step);
	return 0;
}
int orc_sim_finish_deferred_kick(void* h, float dt) {
	sim_finish_deferred_kick(*static_cast<Sim*>(h), dt);
	return 0;
}
int orc_sim_reorder_particles(void* h) {
	sim_reorder(*static_cast<Sim*>(h));
	return 0;
}
// One SimManager step sequence per step (SimManager.cpp:271-296, 337-504).
// Threads for the timing mode of the CPU baseline (1: the serial checker).
int orc_sim_set_threads(void* h, int32_t threads) {
	if (threads < 1)
		return fail("threads must be >= 1");
	static_cast<Sim*>(h)->threads = threads;
	return 0;
}

int orc_sim_run(void* h, const marsmd_run_params* prm, uint64_t first_step, uint64_t num_steps) {
	Sim& s = *static_cast<Sim*>(h);
	try {
		const int64_t rp = prm->reorder_period > 0 ? prm->reorder_period : int64_t(prm->rebuild_period);
		for (uint64_t step = first_step + 1; step <= first_step + num_steps; ++step) {
			if (rp > 0 && step != 0 && (step - 1) % uint64_t(rp) == 0 && s.bonded_prepared)
				sim_reorder(s);
			const bool energy = prm->energy_period > 0 && step % prm->energy_period == 0;
			sim_calculate_nonbonded(s, prm->pairlist_cutoff, prm->cutoff, step, prm->rebuild_period,
									prm->electric_field, prm->interpolation, energy);
			sim_bonded_forces(s, energy);
			sim_integrate(s, prm->timestep, prm->kT, prm->integrator, step);
		}
		return 0;
	} catch (const std::exception& e) {
		return fail(e.what());
	}
}
// Pairs of the last build (sorted, canonical). Returns the count.
int64_t orc_sim_get_pairs(void* h, int32_t* out, uint64_t cap) {
	Sim& s = *static_cast<Sim*>(h);
	const uint64_t np = s.pl.pairs.size();
	for (uint64_t k = 0; k < np && k < cap; ++k) {
		out[2 * k] = s.pl.pairs[k].first;
		out[2 * k + 1] = s.pl.pairs[k].second;
	}
	return int64_t(np);
}
int orc_sim_get_topology(void* h, int32_t* bonds, int32_t* angles, int32_t* excl) {
	Sim& s = *static_cast<Sim*>(h);
	for (size_t k = 0; k < s.bonds.size(); ++k) {
		bonds[2 * k] = s.bonds[k].first;
		bonds[2 * k + 1] = s.bonds[k].second;
	}
	for (size_t k = 0; k < s.angles.size(); ++k)
		for (int q = 0; q < 3; ++q)
			angles[3 * k + size_t(q)] = s.angles[k][size_t(q)];
	for (size_t k = 0; k < s.excl_pairs.size(); ++k) {
		excl[2 * k] = s.excl_pairs[k].first;
		excl[2 * k + 1] = s.excl_pairs[k].second;
	}
	return 0;
}

} // extern "C"

extern "C" {
// The shared FP32 arccosine of the angle term.
float orc_acos(float x) { return acos_series(x); }
// The shared FP32 arctangent of the dihedral term.
float orc_atan2(float y, float x) { return atan2_series(y, x); }

// Dihedrals (4 ints each) and harmonic restraints on explicit arrays, as
// launch_tabulated_dihedrals / launch_harmonic_restraints. mode 1: FP32 terms summed
// in FP64; mode 2: FP64 terms (libm atan2). out: n x 4 doubles (force, energy) and,
// when phi_out is not null, each dihedral's angle.
int orc_dihedral_restraint_forces(const int32_t* dihedrals, const int32_t* dihedral_table, uint64_t num_dihedrals,
								  const marsmd_table* dtabs, const int32_t* restraint_ids, const float* anchors4,
								  const float* ks, uint64_t num_restraints, const float* pos4, const marsmd_box* box,
								  int32_t mode, uint64_t n, double* out, double* phi_out) {
	const Box b = make_box(*box);
	const V4f* p = as_v4(pos4);
	std::fill(out, out + 4 * n, 0.0);
	auto add = [&](int i, double fx, double fy, double fz, double e) {
		out[4 * i + 0] += fx;
		out[4 * i + 1] += fy;
		out[4 * i + 2] += fz;
		out[4 * i + 3] += e;
	};
	auto v64 = [&](int i) { return V4<double>{p[i].x, p[i].y, p[i].z, 0.0}; };
	for (uint64_t k = 0; k < num_dihedrals; ++k) {
		const int q[4] = {dihedrals[4 * k], dihedrals[4 * k + 1], dihedrals[4 * k + 2], dihedrals[4 * k + 3]};
		const marsmd_table& t = dtabs[dihedral_table ? dihedral_table[k] : 0];
		if (mode == 2) {
			const DihedralTerm<double> dt =
				dihedral_term<double>(b, v64(q[0]), v64(q[1]), v64(q[2]), v64(q[3]), make_table<double>(t));
			for (int a = 0; a < 4; ++a)
				add(q[a], dt.f[a].x, dt.f[a].y, dt.f[a].z, dt.quarter_energy);
			if (phi_out)
				phi_out[k] = dt.phi;
			continue;
		}
		const DihedralTerm<float> dt = dihedral_term<float>(b, p[q[0]], p[q[1]], p[q[2]], p[q[3]], make_table<float>(t));
		for (int a = 0; a < 4; ++a)
			add(q[a], dt.f[a].x, dt.f[a].y, dt.f[a].z, dt.quarter_energy);
		if (phi_out)
			phi_out[k] = dt.phi;
	}
	for (uint64_t k = 0; k < num_restraints; ++k) {
		const int id = restraint_ids[k];
		if (id < 0)
			continue;
		const V4f anc{anchors4[4 * k], anchors4[4 * k + 1], anchors4[4 * k + 2], 0.0f};
		if (mode == 2) {
			const RestraintTerm<double> rt = restraint_term<double>(
				b, v64(id), V4<double>{anc.x, anc.y, anc.z, 0.0}, double(ks[k]));
			add(id, rt.f.x, rt.f.y, rt.f.z, rt.energy);
			continue;
		}
		const RestraintTerm<float> rt = restraint_term<float>(b, p[id], anc, ks[k]);
		add(id, rt.f.x, rt.f.y, rt.f.z, rt.energy);
	}
	return 0;
}
// Stash the field and interpolation order the integrator applies (the state
// calculate_nonbonded_forces records, Patch.cpp:71-73).
int orc_sim_set_field(void* h, const float ef[3], int32_t interp) {
	Sim& s = *static_cast<Sim*>(h);
	for (int a = 0; a < 3; ++a)
		s.efield[a] = ef[a];
	s.interp = interp;
	return 0;
}
} // extern "C"

// ---- force matching of tabulated pair tables (the gradient path of this library; the
// reference has none). Per snapshot: the oracle's pair set at `cutoff`; each pair's
// term as PairNonbondedComputer (Pairwise.h:184-305) forms it -- FP32 wrap_diff,
// distance, unit vector and table bin (TabulatedPotential.h:19-52, non-periodic) --
// but with FP64 knots and FP64 sums, so the forces are linear in the knots exactly:
//   fm = -(U[h+1] - U[h]) step_inv.
// Loss L = mean over snapshots, particles and components of (F - F_target)^2 and
// dL/dU by the chain rule through that linear map (each pair once, x = lower index):
//   dL/dfm = (G_y - G_x) . u, G = dL/dF;  dL/dU[h] += dL/dfm step_inv,
//   dL/dU[h+1] -= dL/dfm step_inv.
// knots64: all tables' knots, concatenated in table order. Any output may be NULL.
extern "C" int orc_fm_loss_grad(uint64_t S, uint64_t n, const float* pos4, const int32_t* types, int32_t num_types,
								const uint32_t* term_matrix, uint32_t enabled, int32_t num_tables,
								const marsmd_table* tables, const double* knots64, const marsmd_box* box,
								float cutoff, const double* target, double* forces_out, double* loss_out,
								double* grad_out) {
	try {
		const Box b = make_box(*box);
		std::vector<size_t> off(size_t(num_tables) + 1, 0);
		for (int k = 0; k < num_tables; ++k) {
			if (tables[k].is_periodic)
				return -1;
			off[size_t(k) + 1] = off[size_t(k)] + tables[k].size;
		}
		if (grad_out)
			std::fill(grad_out, grad_out + off[size_t(num_tables)], 0.0);
		const float cut2 = cutoff * cutoff;
		const double norm = 1.0 / (double(S) * double(n) * 3.0);
		double loss = 0.0;
		std::vector<double> F(3 * n), G(3 * n);
		struct Hit {
			int x, y;
			size_t k0;
			double sinv;
			double ux, uy, uz;
		};
		std::vector<Hit> hits;
		for (uint64_t s = 0; s < S; ++s) {
			const V4f* p = as_v4(pos4 + 4 * n * s);
			PairlistResult r;
			build_pairlist(p, n, b, cutoff, marsmd_exclusion_view{}, ~uint64_t(0), r, false);
			std::fill(F.begin(), F.end(), 0.0);
			hits.clear();
			for (const auto& pr : r.pairs) {
				const int x = pr.first, y = pr.second;
				const uint32_t tag = term_matrix[size_t(types[x]) * size_t(num_types) + size_t(types[y])];
				if (!(tag & enabled & MARSMD_PAIR_TERM_TABULATED))
					continue;
				const V4f dr = wrap_diff<float>(b, sub<float>(p[y], p[x]));
				const float d = std::sqrt(length2<float>(dr));
				if (d < 1e-6f)
					continue;
				if (cut2 > 0.0f && d * d > cut2)
					continue;
				const float inv = 1.0f / d;
				const float ux = inv * dr.x, uy = inv * dr.y, uz = inv * dr.z;
				const int ti = int(tag >> MARSMD_PAIR_TABLE_SHIFT);
				const marsmd_table& t = tables[ti];
				const float w = (d - t.start) * t.step_inv;
				int h = int(std::floor(w));
				if (h >= int(t.size) - 1)
					continue; // past the last knot: no force
				if (h < 0)
					h = 0;
				const size_t k0 = off[size_t(ti)] + size_t(h);
				const double fm = -(knots64[k0 + 1] - knots64[k0]) * double(t.step_inv);
				F[3 * size_t(x) + 0] += -fm * ux;
				F[3 * size_t(x) + 1] += -fm * uy;
				F[3 * size_t(x) + 2] += -fm * uz;
				F[3 * size_t(y) + 0] += fm * ux;
				F[3 * size_t(y) + 1] += fm * uy;
				F[3 * size_t(y) + 2] += fm * uz;
				hits.push_back(Hit{x, y, k0, double(t.step_inv), ux, uy, uz});
			}
			if (forces_out)
				std::copy(F.begin(), F.end(), forces_out + 3 * n * s);
			if (!target)
				continue;
			const double* T = target + 3 * n * s;
			for (size_t i = 0; i < 3 * n; ++i) {
				const double diff = F[i] - T[i];
				loss += diff * diff * norm;
				G[i] = 2.0 * diff * norm;
			}
			if (grad_out)
				for (const Hit& hit : hits) {
					const size_t X = 3 * size_t(hit.x), Y = 3 * size_t(hit.y);
					const double dfm =
						(G[Y] - G[X]) * hit.ux + (G[Y + 1] - G[X + 1]) * hit.uy + (G[Y + 2] - G[X + 2]) * hit.uz;
					grad_out[hit.k0] += dfm * hit.sinv;
					grad_out[hit.k0 + 1] -= dfm * hit.sinv;
				}
		}
		if (loss_out)
			*loss_out = loss;
		return 0;
	} catch (...) {
		return -1;
	}
}
