// analytic.hpp — the analytical nonbonded pair terms (Coulomb, Debye-Hueckel, ONC
// electrostatics, Gaussian, softcore) of PairNonbondedComputer / AnalyticalPairKernel
// (src/Interactions/Nonbonded/Pairwise.h:226-296, AnalyticalPairKernels.h:44-104; the
// functors of Columb.h:11-101 and Pairwise.h:47-99), in FP32 and in the reference's
// order of accumulation.
#pragma once

#include "internal.hpp"

namespace mmd {

constexpr uint32_t kAnalyticTerms = MARSMD_PAIR_TERM_COULOMB | MARSMD_PAIR_TERM_DEBYE_HUCKEL |
									MARSMD_PAIR_TERM_ONCK | MARSMD_PAIR_TERM_GAUSSIAN | MARSMD_PAIR_TERM_SOFTCORE;
constexpr float kCoulomb = 332.0636f; // constants::COULOMB (src/Constants.h:24)

// Per-type charges and softcore parameters with the solvent constants: what the
// analytical terms read besides the pair geometry.
struct PairCoeffs {
	float charge[MARSMD_MAX_TYPES];
	float radius[MARSMD_MAX_TYPES];
	float eps[MARSMD_MAX_TYPES];
	marsmd_pair_params prm;
};

inline marsmd_pair_params default_pair_params() {
	marsmd_pair_params p;
	p.debye_length = 10.0f;
	p.dielectric = 80.0f;
	p.onck_kappa = 0.1f;
	p.onck_sz = 80.0f;
	p.onck_z = 6.86f;
	p.gaussian_amp = 1.0f;
	p.gaussian_sigma = 1.0f;
	p.gaussian_cutoff_squared = 0.0f;
	return p;
}

inline PairCoeffs make_pair_coeffs(const marsmd_particle_types& t, const marsmd_pair_params* p) {
	PairCoeffs c{};
	for (int k = 0; k < MARSMD_MAX_TYPES; ++k) {
		c.charge[k] = t.charge[k];
		c.radius[k] = t.radius[k];
		c.eps[k] = t.eps[k];
	}
	c.prm = p ? *p : default_pair_params();
	return c;
}

// sinh(v) - v and v cosh(v) - sinh(v) without the cancellation near 0: the Taylor
// series sum_k v^(2k+1) / (2k+1)! (k >= 1), resp. sum_k 2k v^(2k+1) / (2k+1)!, to
// v^9 for |v| < 1 (the same truncation as math::sinh_minus_x, Types/Math.h:198-215).
__device__ __forceinline__ float sinh_minus_v(float v) {
	if (v >= 1.0f || v <= -1.0f)
		return sinhf(v) - v;
	const float w = v * v;
	return v * w * (1.0f / 6.0f + w * (1.0f / 120.0f + w * (1.0f / 5040.0f + w * (1.0f / 362880.0f))));
}
__device__ __forceinline__ float v_cosh_minus_sinh(float v) {
	if (v >= 1.0f || v <= -1.0f)
		return v * coshf(v) - sinhf(v);
	const float w = v * v;
	return v * w * (1.0f / 3.0f + w * (1.0f / 30.0f + w * (1.0f / 840.0f + w * (1.0f / 45360.0f))));
}

// The analytical terms enabled in `terms` for a pair at distance d (d2 = the squared
// length of the wrapped difference) whose particles carry charges ci, cj and softcore
// radii and eps (ri, rj, ei, ej): force magnitude (-dU/dr) and energy added to fm and
// en in the reference's order.
__device__ __forceinline__ void analytic_terms(uint32_t terms, float d, float d2, float ci, float cj, float ri,
											   float rj, float ei, float ej, const marsmd_pair_params& prm, float& fm,
											   float& en) {
	constexpr uint32_t charged = MARSMD_PAIR_TERM_COULOMB | MARSMD_PAIR_TERM_DEBYE_HUCKEL | MARSMD_PAIR_TERM_ONCK;
	float qi = 0.0f, qj = 0.0f;
	if (terms & charged) {
		qi = ci;
		qj = cj;
	}
	if (terms & MARSMD_PAIR_TERM_COULOMB) { // ColumbPotential::compute
		const float inv_d2 = 1.0f / (d * d);
		const float inv_d = 1.0f / d;
		fm += qi * qj * inv_d2 * kCoulomb;
		en += qi * qj * inv_d * kCoulomb;
	}
	if (terms & MARSMD_PAIR_TERM_DEBYE_HUCKEL) { // U = A exp(-r/l) / r, A = C qi qj / eps
		const float inv_r = 1.0f / d;
		const float a = qi * qj * kCoulomb / prm.dielectric;
		const float screen = expf(-d / prm.debye_length);
		const float u = a * screen * inv_r;
		fm += u * (inv_r + 1.0f / prm.debye_length);
		en += u;
	}
	if (terms & MARSMD_PAIR_TERM_ONCK) { // U = C qi qj exp(-k r) / (eps(r) r)
		const float r = d < 1e-2f ? 1e-2f : d;
		const float a = qi * qj * kCoulomb;
		const float screen = expf(-prm.onck_kappa * r);
		const float sz = prm.onck_sz, z = prm.onck_z;
		const float v = r / (2.0f * z);
		float eps, deps;
		if (v < 1e-5f) {
			eps = sz * v * v / 3.0f;
			deps = sz * v / (3.0f * z);
		} else {
			const float s = sinhf(v);
			eps = sz * sinh_minus_v(v) * (s + v) / (s * s);
			deps = sz * v * v_cosh_minus_sinh(v) / (z * s * s * s);
		}
		en += a * screen / (eps * r);
		fm += a * screen * (prm.onck_kappa / (eps * r) + (deps * r + eps) / (eps * eps * r * r));
	}
	if (terms & MARSMD_PAIR_TERM_GAUSSIAN) { // U = amp exp(-r^2 / sigma^2)
		const float dd = d * d;
		if (!(prm.gaussian_cutoff_squared > 0.0f && dd > prm.gaussian_cutoff_squared)) {
			const float inv_s2 = 1.0f / (prm.gaussian_sigma * prm.gaussian_sigma);
			const float u = prm.gaussian_amp * expf(-dd * inv_s2);
			fm += 2.0f * d * u * inv_s2;
			en += u;
		}
	}
	if (terms & MARSMD_PAIR_TERM_SOFTCORE) { // softcoreForce returns dU/dr / r
		const float rad = 0.5f * (ri + rj);
		const float rad2 = rad * rad;
		const float rad6 = rad2 * rad2 * rad2;
		const float pe = sqrtf(ei * ej);
		const float d6 = d2 * d2 * d2;
		float f = -12.0f * pe * (rad6 * rad6 / (d6 * d6 * d2) - rad6 / (d6 * d2));
		float u = 0.0f;
		if (d6 < rad6) {
			const float d12 = d6 * d6, rad12 = rad6 * rad6;
			u = pe * ((rad12 / d12) - (2.0f * rad6 / d6)) + pe;
			f = -12.0f * pe * (rad12 / (d12 * d2) - rad6 / (d6 * d2));
		}
		fm += -f * d;
		en += u;
	}
}

} // namespace mmd
