// oracle.hpp — CPU restatement of MARS-MD's per-step particle hot path.
//
// TEST INFRASTRUCTURE ONLY. This is the parity oracle: it is called by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and only as the
// checker (or the timed CPU baseline). Nothing in the product path links or calls it.
//
// Single-threaded, FP32 op for op in the reference's operation order (built with
// -ffp-contract=off, no fast-math) and templated on the real type so an FP64 twin
// exists for every floating-point function. Citations are reference file:line.
//
// Pinning (what checks each part beyond restating it) is listed in the header of
// oracle_capi.cpp and in the tests that exercise it.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <vector>

namespace orc {

template<class T>
struct V4 {
	T x{0}, y{0}, z{0}, t{0};
};
using V4f = V4<float>;

// ---------------------------------------------------------------------- Philox
// openrand::Philox, src/PatchOperation/Random/philox.h:34-139. Key = {seed_hi,
// seed_lo}; counter = {ctr, ctr1, global_seed, internal}; 10 rounds, key bumped by
// the Weyl constants before every round but the first.
struct Philox {
	uint32_t key0, key1, c0, c1, c2, internal = 0;
	Philox(uint64_t seed, uint32_t ctr, uint32_t global_seed, uint32_t ctr1)
		: key0(uint32_t(seed >> 32)), key1(uint32_t(seed & 0xFFFFFFFFu)), c0(ctr), c1(ctr1),
		  c2(global_seed) {}
	static void round(const uint32_t k[2], uint32_t c[4]) {
		const uint64_t p0 = uint64_t(0xD2511F53u) * uint64_t(c[0]);
		const uint64_t p1 = uint64_t(0xCD9E8D57u) * uint64_t(c[2]);
		const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
		const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
		const uint32_t n0 = hi1 ^ c[1] ^ k[0];
		const uint32_t n2 = hi0 ^ c[3] ^ k[1];
		c[0] = n0;
		c[1] = lo1;
		c[2] = n2;
		c[3] = lo0;
	}
	void generate(uint32_t out[4]) {
		uint32_t k[2] = {key0, key1};
		out[0] = c0;
		out[1] = c1;
		out[2] = c2;
		out[3] = internal;
		for (int r = 0; r < 10; ++r) {
			if (r > 0) {
				k[0] += 0x9E3779B9u;
				k[1] += 0xBB67AE85u;
			}
			round(k, out);
		}
		++internal;
	}
};

// u01<float, uint32_t>: x * 2^-32 + 2^-33 (src/PatchOperation/Random/base_state.h:299-305).
inline float u01(uint32_t x) {
	const float factor = 1.0f / (float(~uint32_t(0)) + 1.0f);
	const float half = 0.5f * factor;
	return float(x) * factor + half;
}

// Box-Muller of BDIntegrate / BAOABIntegrate (BD.h:70-78; BAOAB.h:97-104).
inline void box_muller(const float u[4], float out[3]) {
	const float r1 = std::sqrt(-2.0f * std::log(u[0]));
	const float theta1 = 2.0f * 3.1415926535f * u[1];
	const float r2 = std::sqrt(-2.0f * std::log(u[2]));
	const float theta2 = 2.0f * 3.1415926535f * u[3];
	out[0] = r1 * std::cos(theta1);
	out[1] = r1 * std::sin(theta1);
	out[2] = r2 * std::cos(theta2);
}

// Seed mix of Patch::integrate_motion (Patch.cpp:252-254).
inline uint64_t mix_seed(uint64_t seed, uint64_t patch_id, uint64_t step) {
	return (seed * 0x9E3779B97F4A7C15ULL) ^ (patch_id << 32) ^ (step * 0xBF58476D1CE4E5B9ULL);
}

// ---------------------------------------------------------------------- box
// Orthorhombic PeriodicBox (src/System/PeriodicBox.h:134-209, 445-461).
struct Box {
	float L[3]{0, 0, 0};
	float origin[3]{0, 0, 0};
	bool periodic[3]{false, false, false};
};

template<class T>
inline T wrap_diff_scalar(T x, T l) { // PeriodicBox.h:453-461
	if (l <= T(0))
		return x;
	const int image = int(std::floor(x * (T(1) / l)));
	x -= T(image) * l;
	if (x >= T(0.5) * l)
		x -= l;
	return x;
}
template<class T>
inline T wrap_scalar(T x, T o, T l) { // PeriodicBox.h:445-451
	if (l <= T(0))
		return x;
	const T rel = x - o;
	const int image = int(std::floor(rel * (T(1) / l)));
	return x - T(image) * l;
}
template<class T>
inline V4<T> wrap_diff(const Box& b, V4<T> d) { // PeriodicBox.h:134-148
	if (b.periodic[0])
		d.x = wrap_diff_scalar<T>(d.x, T(b.L[0]));
	if (b.periodic[1])
		d.y = wrap_diff_scalar<T>(d.y, T(b.L[1]));
	if (b.periodic[2])
		d.z = wrap_diff_scalar<T>(d.z, T(b.L[2]));
	return d;
}
template<class T>
inline V4<T> wrap(const Box& b, V4<T> r) { // PeriodicBox.h:172-182
	if (b.periodic[0])
		r.x = wrap_scalar<T>(r.x, T(b.origin[0]), T(b.L[0]));
	if (b.periodic[1])
		r.y = wrap_scalar<T>(r.y, T(b.origin[1]), T(b.L[1]));
	if (b.periodic[2])
		r.z = wrap_scalar<T>(r.z, T(b.origin[2]), T(b.L[2]));
	return r;
}
template<class T>
inline V4<T> sub(const V4<T>& a, const V4<T>& b) { // Vector3_t::operator- (.t dropped)
	return V4<T>{a.x - b.x, a.y - b.y, a.z - b.z, T(0)};
}
template<class T>
inline T length2(const V4<T>& v) { // Vector3.h length2: (x*x + y*y) + z*z
	return v.x * v.x + v.y * v.y + v.z * v.z;
}

// ---------------------------------------------------------------------- Morton
// MortonCode (src/PatchOperation/ZOrderKernels/MortonCode.h:91-180).
inline uint32_t split_by3(uint32_t a) {
	uint32_t x = a & 1023u;
	x = (x | x << 16) & 0x30000ffu;
	x = (x | x << 8) & 0x300f00fu;
	x = (x | x << 4) & 0x30c30c3u;
	x = (x | x << 2) & 0x9249249u;
	return x;
}
inline uint32_t compact_by3(uint32_t x) {
	x &= 0x9249249u;
	x = (x ^ (x >> 2)) & 0x30c30c3u;
	x = (x ^ (x >> 4)) & 0x300f00fu;
	x = (x ^ (x >> 8)) & 0x30000ffu;
	x = (x ^ (x >> 16)) & 1023u;
	return x;
}
inline uint32_t morton_from_coords(uint32_t x, uint32_t y, uint32_t z) {
	return split_by3(z) | (split_by3(y) << 1) | (split_by3(x) << 2);
}
inline uint32_t morton_quantize(float p, float lo, float hi) {
	const float range = std::fmax(hi - lo, 1e-6f);
	const float normalized = (p - lo) / range;
	return uint32_t(std::fmin(std::fmax(normalized * 1023.0f, 0.0f), 1023.0f));
}
// MortonCode::encode(pos, box_min, box_max), MortonCode.h:95-131.
inline uint32_t morton_encode(const V4f& p, const float bmin[3], const float bmax[3]) {
	return morton_from_coords(morton_quantize(p.x, bmin[0], bmax[0]),
							  morton_quantize(p.y, bmin[1], bmax[1]),
							  morton_quantize(p.z, bmin[2], bmax[2]));
}

// ---------------------------------------------------------------------- tables
// TabulatedPotential (src/Interactions/TabulatedPotential.h:8-52).
template<class T>
struct Table {
	std::vector<T> pot;
	T step_inv{0};
	T start{0};
	bool periodic{false};
};

template<class T>
struct FE {
	T f{0}, e{0};
};

template<class T>
inline FE<T> table_compute(T dx, const Table<T>& tab) { // TabulatedPotential.h:19-52
	const int size = int(tab.pot.size());
	T w = (dx - tab.start) * tab.step_inv;
	int home = int(std::floor(w));
	w = w - T(home);
	if (tab.periodic) {
		home %= size;
		if (home < 0)
			home += size;
	} else {
		if (home >= size - 1)
			return FE<T>{T(0), tab.pot[size - 1]};
		if (home < 0)
			home = 0;
	}
	int home1 = tab.periodic ? (home + 1) % size : home + 1;
	if (!tab.periodic && home1 >= size)
		home1 = size - 1;
	const T U0 = tab.pot[home];
	const T dU = tab.pot[home1] - U0;
	return FE<T>{-dU * tab.step_inv, dU * w + U0};
}

// ---------------------------------------------------------------------- pair force
// ------------------------------------------------------- analytical pair terms
// Constants of the analytical terms: SolventParams (NonBondedInteraction.h:121-127) and
// GaussianPotential (Pairwise.h:53-57), per-type charge, softcore radius and eps.
template<class T>
struct PairCoef {
	T debye_length{10}, dielectric{80}, onck_kappa{T(0.1)}, onck_sz{80}, onck_z{T(6.86)};
	T gaussian_amp{1}, gaussian_sigma{1}, gaussian_cutoff_squared{0};
	std::vector<T> charge, radius, eps;
};

constexpr uint32_t kTermCoulomb = 1u, kTermDH = 2u, kTermOnck = 4u, kTermGauss = 8u, kTermSoftcore = 16u,
				   kTermTabulated = 32u;

inline float sinh_t(float x) { return std::sinh(x); }
inline double sinh_t(double x) { return std::sinh(x); }
inline float cosh_t(float x) { return std::cosh(x); }
inline double cosh_t(double x) { return std::cosh(x); }
inline float exp_t(float x) { return std::exp(x); }
inline double exp_t(double x) { return std::exp(x); }

// sinh(v) - v and v cosh(v) - sinh(v): for |v| < 1 the Taylor series to v^9, as
// math::sinh_minus_x / x_cosh_minus_sinh (Types/Math.h:198-215); the FP64 twin uses
// the same form.
template<class T>
inline T sinh_minus_v(T v) {
	if (v >= T(1) || v <= T(-1))
		return sinh_t(v) - v;
	const T w = v * v;
	return v * w * (T(1.0 / 6) + w * (T(1.0 / 120) + w * (T(1.0 / 5040) + w * T(1.0 / 362880))));
}
template<class T>
inline T v_cosh_minus_sinh(T v) {
	if (v >= T(1) || v <= T(-1))
		return v * cosh_t(v) - sinh_t(v);
	const T w = v * v;
	return v * w * (T(1.0 / 3) + w * (T(1.0 / 30) + w * (T(1.0 / 840) + w * T(1.0 / 45360))));
}

// The analytical branches of PairNonbondedComputer::operator() (Pairwise.h:226-277) in
// order, with the functors of Columb.h:11-101 (Coulomb, Debye-Hueckel, ONC) and
// Pairwise.h:47-99 (Gaussian, softcore). d = |r|, d2 = r.r; adds -dU/dr to fm and U to en.
template<class T>
inline void analytic_terms(uint32_t terms, T d, T d2, int ti, int tj, const PairCoef<T>& c, T& fm, T& en) {
	const T C = T(332.0636f); // constants::COULOMB (Constants.h:24)
	T qi = 0, qj = 0;
	if (terms & (kTermCoulomb | kTermDH | kTermOnck)) {
		qi = c.charge[size_t(ti)];
		qj = c.charge[size_t(tj)];
	}
	if (terms & kTermCoulomb) {
		const T inv_d2 = T(1) / (d * d);
		const T inv_d = T(1) / d;
		fm += qi * qj * inv_d2 * C;
		en += qi * qj * inv_d * C;
	}
	if (terms & kTermDH) {
		const T inv_r = T(1) / d;
		const T a = qi * qj * C / c.dielectric;
		const T screen = exp_t(-d / c.debye_length);
		const T u = a * screen * inv_r;
		fm += u * (inv_r + T(1) / c.debye_length);
		en += u;
	}
	if (terms & kTermOnck) {
		const T r = d < T(1e-2f) ? T(1e-2f) : d;
		const T a = qi * qj * C;
		const T screen = exp_t(-c.onck_kappa * r);
		const T v = r / (T(2) * c.onck_z);
		T eps, deps;
		if (v < T(1e-5f)) {
			eps = c.onck_sz * v * v / T(3);
			deps = c.onck_sz * v / (T(3) * c.onck_z);
		} else {
			const T sh = sinh_t(v);
			eps = c.onck_sz * sinh_minus_v(v) * (sh + v) / (sh * sh);
			deps = c.onck_sz * v * v_cosh_minus_sinh(v) / (c.onck_z * sh * sh * sh);
		}
		en += a * screen / (eps * r);
		fm += a * screen * (c.onck_kappa / (eps * r) + (deps * r + eps) / (eps * eps * r * r));
	}
	if (terms & kTermGauss) {
		const T dd = d * d;
		if (!(c.gaussian_cutoff_squared > T(0) && dd > c.gaussian_cutoff_squared)) {
			const T inv_s2 = T(1) / (c.gaussian_sigma * c.gaussian_sigma);
			const T u = c.gaussian_amp * exp_t(-dd * inv_s2);
			fm += T(2) * d * u * inv_s2;
			en += u;
		}
	}
	if (terms & kTermSoftcore) {
		const T rad = T(0.5) * (c.radius[size_t(ti)] + c.radius[size_t(tj)]);
		const T rad2 = rad * rad;
		const T rad6 = rad2 * rad2 * rad2;
		const T pe = std::sqrt(c.eps[size_t(ti)] * c.eps[size_t(tj)]);
		const T d6 = d2 * d2 * d2;
		T f = T(-12) * pe * (rad6 * rad6 / (d6 * d6 * d2) - rad6 / (d6 * d2));
		T u = 0;
		if (d6 < rad6) {
			const T d12 = d6 * d6, rad12 = rad6 * rad6;
			u = pe * ((rad12 / d12) - (T(2) * rad6 / d6)) + pe;
			f = T(-12) * pe * (rad12 / (d12 * d2) - rad6 / (d6 * d2));
		}
		fm += -f * d;
		en += u;
	}
}

// CalcDistance::compute (src/Interactions/Interactions.h:37-53) followed by the term
// branches of PairNonbondedComputer (Pairwise.h:214-304): the analytical terms in
// `terms` (with types ti, tj of x and y) and the tabulated one when `tab` is given.
template<class T>
struct PairTerm {
	bool valid{false};
	V4<T> force; // force on y; -force on x
	T half_energy{0};
};

template<class T>
inline PairTerm<T> pair_term(const Box& box, const V4<T>& px, const V4<T>& py, const Table<T>* tab,
							 T cutoff_squared, uint32_t terms = kTermTabulated, int ti = 0, int tj = 0,
							 const PairCoef<T>* coef = nullptr) {
	PairTerm<T> out;
	const V4<T> r = wrap_diff<T>(box, sub<T>(py, px));
	const T d2 = length2<T>(r);
	const T d = std::sqrt(d2);
	V4<T> unit{};
	if (d > T(1e-6)) {
		const T inv = T(1) / d;
		unit = V4<T>{inv * r.x, inv * r.y, inv * r.z, T(0)};
	}
	if (d < T(1e-6))
		return out;
	if (cutoff_squared > T(0) && d * d > cutoff_squared)
		return out;
	// force_magnitude and energy start at 0 and take each enabled term in order
	T fm = 0, en = 0;
	if (coef && (terms & (kTermCoulomb | kTermDH | kTermOnck | kTermGauss | kTermSoftcore)))
		analytic_terms<T>(terms, d, d2, ti, tj, *coef, fm, en);
	if (tab && (terms & kTermTabulated)) {
		const FE<T> fe = table_compute<T>(d, *tab);
		fm += fe.f;
		en += fe.e;
	}
	out.valid = true;
	out.force = V4<T>{fm * unit.x, fm * unit.y, fm * unit.z, T(0)};
	out.half_energy = en * T(0.5);
	return out;
}

// TabulatedBondComputer (src/Interactions/Bonded/BondComputer.h:163-191).
template<class T>
inline PairTerm<T> bond_term(const Box& box, const V4<T>& px, const V4<T>& py,
							 const Table<T>& tab) {
	PairTerm<T> out;
	const V4<T> r = wrap_diff<T>(box, sub<T>(py, px));
	const T d = std::sqrt(length2<T>(r));
	V4<T> unit{};
	if (d > T(1e-6)) {
		const T inv = T(1) / d;
		unit = V4<T>{inv * r.x, inv * r.y, inv * r.z, T(0)};
	}
	if (d < T(1e-6))
		return out;
	const FE<T> fe = table_compute<T>(d, tab);
	out.valid = true;
	out.force = V4<T>{fe.f * unit.x, fe.f * unit.y, fe.f * unit.z, T(0)};
	out.half_energy = fe.e * T(0.5);
	return out;
}

// FP32 arccosine shared with the device kernels (device_common.hpp acos_series), so
// the angle term is the same bits on both sides: the arcsine Maclaurin series
// asin(t) = t + t z sum_k c_k z^k, z = t^2, c_k = (2k+2)! / (4^(k+1) ((k+1)!)^2 (2k+3)),
// with acos(x) = pi/2 - asin(x) for |x| <= 1/2 and acos(|x|) = 2 asin(sqrt((1-|x|)/2)),
// acos(-|x|) = pi - acos(|x|) otherwise (z <= 1/4: the series converges to < 1e-9).
inline float acos_series(float x) {
	const float ax = std::fabs(x);
	float t, z;
	if (ax <= 0.5f) {
		t = x;
		z = x * x;
	} else {
		z = (1.0f - ax) * 0.5f;
		t = std::sqrt(z);
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
inline float acos_t(float x) { return acos_series(x); }
inline double acos_t(double x) { return std::acos(x); }

// AngleGeometry (src/Interactions/Bonded/BondGeometry.h:13-45) and
// TabulatedAngleComputer (BondComputer.h:231-268). The arccosine is acos_t.
template<class T>
struct AngleTerm {
	V4<T> f1, f2, f3;
	T third_energy{0};
};
template<class T>
inline AngleTerm<T> angle_term(const Box& box, const V4<T>& pa, const V4<T>& pb,
							   const V4<T>& pc, const Table<T>& tab) {
	const V4<T> ab = wrap_diff<T>(box, sub<T>(pb, pa));
	const V4<T> bc = wrap_diff<T>(box, sub<T>(pc, pb));
	const V4<T> ac = wrap_diff<T>(box, sub<T>(pc, pa));
	const T ab2 = length2<T>(ab), bc2 = length2<T>(bc), ac2 = length2<T>(ac);
	T cosv = (ab2 + bc2 - ac2) * T(0.5) / (std::sqrt(ab2) * std::sqrt(bc2));
	if (cosv < T(-1))
		cosv = T(-1);
	if (cosv > T(1))
		cosv = T(1);
	const T angle = acos_t(cosv);
	const T sinv = std::sqrt(T(1) - cosv * cosv);
	const FE<T> fe = table_compute<T>(angle, tab);
	const T sin_floor = T(1e-3);
	const T s = sinv > sin_floor ? sinv : sin_floor;
	const T dUdtheta = -fe.f / s;
	const T inv_ab = T(1) / std::sqrt(length2<T>(ab));
	const T inv_bc = T(1) / std::sqrt(length2<T>(bc));
	const T s1 = dUdtheta * inv_ab;
	const T t1 = cosv * inv_ab;
	const V4<T> v1{t1 * ab.x + inv_bc * bc.x, t1 * ab.y + inv_bc * bc.y,
				   t1 * ab.z + inv_bc * bc.z, T(0)};
	const T s3 = -(dUdtheta * inv_bc);
	const T t3 = cosv * inv_bc;
	const V4<T> v3{t3 * bc.x + inv_ab * ab.x, t3 * bc.y + inv_ab * ab.y,
				   t3 * bc.z + inv_ab * ab.z, T(0)};
	AngleTerm<T> out;
	out.f1 = V4<T>{s1 * v1.x, s1 * v1.y, s1 * v1.z, T(0)};
	out.f3 = V4<T>{s3 * v3.x, s3 * v3.y, s3 * v3.z, T(0)};
	out.f2 = V4<T>{-(out.f1.x + out.f3.x), -(out.f1.y + out.f3.y), -(out.f1.z + out.f3.z), T(0)};
	out.third_energy = fe.e * T(1.0 / 3.0);
	return out;
}

// FP32 two-argument arctangent shared with the device kernels (device_common.hpp
// atan2_series): t = min(|x|,|y|)/max(|x|,|y|); for t > tan(pi/12),
// atan(t) = pi/6 + atan((sqrt(3) t - 1)/(t + sqrt(3))); then the Maclaurin series of
// atan to the t^13 term; octant/quadrant from |y| > |x| and the signs (signed zeros as
// IEEE atan2).
inline float atan2_series(float y, float x) {
	const float ax = std::fabs(x), ay = std::fabs(y);
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
inline float atan2_t(float y, float x) { return atan2_series(y, x); }
inline double atan2_t(double y, double x) { return std::atan2(y, x); }

template<class T>
inline V4<T> cross3(const V4<T>& v, const V4<T>& w) {
	return V4<T>{v.y * w.z - v.z * w.y, v.z * w.x - v.x * w.z, v.x * w.y - v.y * w.x, T(0)};
}
template<class T>
inline T dot3(const V4<T>& v, const V4<T>& w) {
	return v.x * w.x + v.y * w.y + v.z * w.z;
}
// Vector3::rLength2 (src/Types/Vector3.h:233-236).
template<class T>
inline T rlen2(const V4<T>& v) {
	const T l2 = length2<T>(v);
	return l2 != T(0) ? T(1) / l2 : T(0);
}

// DihedralGeometry (src/Interactions/Bonded/BondGeometry.h:56-87) and
// TabulatedDihedralComputer (BondComputer.h:296-343): f[0..3] on i, j, k, l and U/4.
template<class T>
struct DihedralTerm {
	V4<T> f[4];
	T quarter_energy{0};
	T phi{0};
};
template<class T>
inline DihedralTerm<T> dihedral_term(const Box& box, const V4<T>& p0, const V4<T>& p1, const V4<T>& p2,
									 const V4<T>& p3, const Table<T>& tab) {
	const V4<T> ab = wrap_diff<T>(box, sub<T>(p1, p0));
	const V4<T> bc = wrap_diff<T>(box, sub<T>(p2, p1));
	const V4<T> cd = wrap_diff<T>(box, sub<T>(p3, p2));
	const V4<T> c_abc = cross3<T>(ab, bc);
	const V4<T> c_bcd = cross3<T>(bc, cd);
	const V4<T> c_x = cross3<T>(bc, c_abc);
	const T cos_phi = dot3<T>(c_abc, c_bcd) / (std::sqrt(length2<T>(c_abc)) * std::sqrt(length2<T>(c_bcd)));
	const T sin_phi = dot3<T>(c_x, c_bcd) / (std::sqrt(length2<T>(c_x)) * std::sqrt(length2<T>(c_bcd)));
	const T phi = atan2_t(sin_phi, cos_phi);
	const T bcl = std::sqrt(length2<T>(bc));
	const T r_abc = rlen2<T>(c_abc), r_bcd = rlen2<T>(c_bcd), r_bc = rlen2<T>(bc);
	const T s1 = -bcl * r_abc;
	const T s3 = -bcl * r_bcd;
	const V4<T> f1{s1 * c_abc.x, s1 * c_abc.y, s1 * c_abc.z, T(0)};
	const V4<T> f3{s3 * c_bcd.x, s3 * c_bcd.y, s3 * c_bcd.z, T(0)};
	const T a2 = -(dot3<T>(ab, bc) * r_bc);
	const T b2 = dot3<T>(bc, cd) * r_bc;
	const V4<T> f2{a2 * f1.x - b2 * f3.x, a2 * f1.y - b2 * f3.y, a2 * f1.z - b2 * f3.z, T(0)};
	const bool degenerate =
		(length2<T>(ab) * length2<T>(bc) * r_abc > T(100)) || (length2<T>(bc) * length2<T>(cd) * r_bcd > T(100));
	const FE<T> fe = table_compute<T>(phi, tab);
	T mag = degenerate ? T(0) : fe.f;
	if (mag > T(1000))
		mag = T(1000);
	else if (mag < T(-1000))
		mag = T(-1000);
	const V4<T> g1{f1.x * mag, f1.y * mag, f1.z * mag, T(0)};
	const V4<T> g2{f2.x * mag, f2.y * mag, f2.z * mag, T(0)};
	const V4<T> g3{f3.x * mag, f3.y * mag, f3.z * mag, T(0)};
	DihedralTerm<T> out;
	out.f[0] = g1;
	out.f[1] = V4<T>{g2.x - g1.x, g2.y - g1.y, g2.z - g1.z, T(0)};
	out.f[2] = V4<T>{g3.x - g2.x, g3.y - g2.y, g3.z - g2.z, T(0)};
	out.f[3] = V4<T>{-g3.x, -g3.y, -g3.z, T(0)};
	out.quarter_energy = fe.e * T(0.25);
	out.phi = phi;
	return out;
}

// HarmonicRestraintComputer (BondComputer.h:109-129): CalcDistance from the anchor
// (Interactions.h:42-53) and AnalyticalForceComputer<0> (Bonded/Analytical.h:18-30)
// at r0 = 0. Full U on the particle.
template<class T>
struct RestraintTerm {
	V4<T> f;
	T energy{0};
};
template<class T>
inline RestraintTerm<T> restraint_term(const Box& box, const V4<T>& p, const V4<T>& anchor, T k) {
	const V4<T> r = wrap_diff<T>(box, sub<T>(p, anchor));
	const T d = std::sqrt(length2<T>(r));
	V4<T> u{};
	if (d > T(1e-6)) {
		const T inv = T(1) / d;
		u = V4<T>{r.x * inv, r.y * inv, r.z * inv, T(0)};
	}
	const T dr = d - T(0);
	const T force = -k * dr;
	RestraintTerm<T> out;
	out.f = V4<T>{u.x * force, u.y * force, u.z * force, T(0)};
	out.energy = T(0.5) * k * dr * dr;
	return out;
}

// ---------------------------------------------------------------------- grids
// BaseGridView sampling (src/Types/BaseGridDevice.h:218-475).
template<class T>
struct Grid {
	const float* data{nullptr};
	T origin[3]{0, 0, 0};
	T basis[9]{};	  // column-major
	T basis_inv[9]{}; // column-major
	size_t dims[3]{0, 0, 0};
	int bc{0};
};

inline bool map_grid_index(int& j, size_t n, int bc) { // BaseGridDevice.h:218-232
	const int ni = int(n);
	if (j >= 0 && j < ni)
		return true;
	if (bc == 2) {
		j = ((j % ni) + ni) % ni;
		return true;
	}
	if (bc == 1) {
		j = j < 0 ? 0 : ni - 1;
		return true;
	}
	return false;
}
template<class T>
inline T fetch(const Grid<T>& g, int jx, int jy, int jz, int bc) { // :236-251
	if (!map_grid_index(jx, g.dims[0], bc) || !map_grid_index(jy, g.dims[1], bc) ||
		!map_grid_index(jz, g.dims[2], bc))
		return T(0);
	const size_t idx = size_t(jz) + size_t(jy) * g.dims[2] + size_t(jx) * g.dims[1] * g.dims[2];
	return T(g.data[idx]);
}
// Matrix3_t::transform: cols[0]*v.x + cols[1]*v.y + cols[2]*v.z (Matrix3.h:74-76).
template<class T>
inline void mat_transform(const T m[9], const T v[3], T out[3]) {
	for (int k = 0; k < 3; ++k)
		out[k] = v[0] * m[0 + k] + v[1] * m[3 + k] + v[2] * m[6 + k];
}
template<class T>
inline void mat_transpose(const T m[9], T out[9]) {
	for (int c = 0; c < 3; ++c)
		for (int r = 0; r < 3; ++r)
			out[c * 3 + r] = m[r * 3 + c];
}

template<class T>
struct GridSample {
	T value{0};
	T grad[3]{0, 0, 0};
};

template<class T>
inline GridSample<T> sample_linear(const Grid<T>& g, const V4<T>& p, int bc) { // :256-365
	const T v[3] = {p.x - g.origin[0], p.y - g.origin[1], p.z - g.origin[2]};
	T gp[3];
	mat_transform<T>(g.basis_inv, v, gp);
	const int i0 = int(std::floor(gp[0])), j0 = int(std::floor(gp[1])), k0 = int(std::floor(gp[2]));
	const T fx = gp[0] - T(i0), fy = gp[1] - T(j0), fz = gp[2] - T(k0);
	const T v000 = fetch(g, i0, j0, k0, bc), v001 = fetch(g, i0, j0, k0 + 1, bc);
	const T v010 = fetch(g, i0, j0 + 1, k0, bc), v011 = fetch(g, i0, j0 + 1, k0 + 1, bc);
	const T v100 = fetch(g, i0 + 1, j0, k0, bc), v101 = fetch(g, i0 + 1, j0, k0 + 1, bc);
	const T v110 = fetch(g, i0 + 1, j0 + 1, k0, bc), v111 = fetch(g, i0 + 1, j0 + 1, k0 + 1, bc);
	const T one = T(1);
	const T v00 = v000 * (one - fx) + v100 * fx;
	const T v01 = v001 * (one - fx) + v101 * fx;
	const T v10 = v010 * (one - fx) + v110 * fx;
	const T v11 = v011 * (one - fx) + v111 * fx;
	const T v0 = v00 * (one - fy) + v10 * fy;
	const T v1 = v01 * (one - fy) + v11 * fy;
	GridSample<T> out;
	out.value = v0 * (one - fz) + v1 * fz;
	const T gx = (one - fy) * (one - fz) * (v100 - v000) + fy * (one - fz) * (v110 - v010) +
				 (one - fy) * fz * (v101 - v001) + fy * fz * (v111 - v011);
	const T gy = (one - fx) * (one - fz) * (v010 - v000) + fx * (one - fz) * (v110 - v100) +
				 (one - fx) * fz * (v011 - v001) + fx * fz * (v111 - v101);
	const T gz = (one - fx) * (one - fy) * (v001 - v000) + fx * (one - fy) * (v101 - v100) +
				 (one - fx) * fy * (v011 - v010) + fx * fy * (v111 - v110);
	T bt[9];
	mat_transpose<T>(g.basis_inv, bt);
	const T gg[3] = {gx, gy, gz};
	mat_transform<T>(bt, gg, out.grad);
	return out;
}

template<class T>
inline T cr_value(T v0, T v1, T v2, T v3, T t) { // BaseGridDevice.h:376-382
	const T a3 = T(0.5) * (-v0 + T(3) * v1 - T(3) * v2 + v3);
	const T a2 = T(0.5) * (T(2) * v0 - T(5) * v1 + T(4) * v2 - v3);
	const T a1 = T(0.5) * (-v0 + v2);
	return ((a3 * t + a2) * t + a1) * t + v1;
}
template<class T>
inline T cr_deriv(T v0, T v1, T v2, T v3, T t) { // :387-392
	const T a3 = T(0.5) * (-v0 + T(3) * v1 - T(3) * v2 + v3);
	const T a2 = T(0.5) * (T(2) * v0 - T(5) * v1 + T(4) * v2 - v3);
	const T a1 = T(0.5) * (-v0 + v2);
	return (T(3) * a3 * t + T(2) * a2) * t + a1;
}
template<class T>
inline GridSample<T> sample_cubic(const Grid<T>& g, const V4<T>& p, int bc) { // :405-475
	const T v[3] = {p.x - g.origin[0], p.y - g.origin[1], p.z - g.origin[2]};
	T gp[3];
	mat_transform<T>(g.basis_inv, v, gp);
	const int hx = int(std::floor(gp[0])), hy = int(std::floor(gp[1])), hz = int(std::floor(gp[2]));
	const T wx = gp[0] - T(hx), wy = gp[1] - T(hy), wz = gp[2] - T(hz);
	T d1[4][4], v1[4][4];
	for (int iz = 0; iz < 4; ++iz) {
		const int jz = hz - 1 + iz;
		for (int iy = 0; iy < 4; ++iy) {
			const int jy = hy - 1 + iy;
			T t[4];
			for (int ix = 0; ix < 4; ++ix)
				t[ix] = fetch(g, hx - 1 + ix, jy, jz, bc);
			d1[iy][iz] = cr_deriv<T>(t[0], t[1], t[2], t[3], wx);
			v1[iy][iz] = cr_value<T>(t[0], t[1], t[2], t[3], wx);
		}
	}
	T d2x[4], d2y[4], v2[4];
	for (int iz = 0; iz < 4; ++iz) {
		d2x[iz] = cr_value<T>(d1[0][iz], d1[1][iz], d1[2][iz], d1[3][iz], wy);
		d2y[iz] = cr_deriv<T>(v1[0][iz], v1[1][iz], v1[2][iz], v1[3][iz], wy);
		v2[iz] = cr_value<T>(v1[0][iz], v1[1][iz], v1[2][iz], v1[3][iz], wy);
	}
	const T gg[3] = {cr_value<T>(d2x[0], d2x[1], d2x[2], d2x[3], wz),
					 cr_value<T>(d2y[0], d2y[1], d2y[2], d2y[3], wz),
					 cr_deriv<T>(v2[0], v2[1], v2[2], v2[3], wz)};
	GridSample<T> out;
	out.value = cr_value<T>(v2[0], v2[1], v2[2], v2[3], wz);
	T bt[9];
	mat_transpose<T>(g.basis_inv, bt);
	mat_transform<T>(bt, gg, out.grad);
	return out;
}

} // namespace orc
