// fm.hpp — host class behind marsmd_fm (see fm.hip).
#pragma once

#include <vector>

#include "pairlist.hpp"

namespace mmd {

struct NbTables;
struct FmSnapView;

struct FMImpl {
	struct Snap;
	uint64_t n = 0, S = 0, num_pairs = 0;
	marsmd_box box{};
	float cutoff2 = 0.0f;
	NbTables* tables = nullptr;
	std::vector<Snap*> snaps;
	DevArray<double> partial, seg_sums, mse_sums;
	DevArray<float4> grad_slot; // the adjoint's dL/dF in each snapshot's slot order
	DevArray<FmSnapView> snap_views; // per snapshot: its device arrays

	explicit FMImpl(const marsmd_fm_desc& d);
	~FMImpl();
	FMImpl(const FMImpl&) = delete;
	FMImpl& operator=(const FMImpl&) = delete;

	// Forces of every snapshot for the given knots (device FP32, tables concatenated):
	// out4 is S x n float4 records {F, energy} in each snapshot's particle order, or in
	// its sorted order (the order of its pairlist build) when sorted.
	void forces(const float* knots, float* out4, hipStream_t s, bool sorted = false);
	// S x n rows of width (3 or 4) floats from particle order into sorted order.
	void to_sorted(const float* in, int width, float* out, hipStream_t s);
	// Mean squared error of the forces (force4: S x n float4 as forces() writes them)
	// against target3 (S x n x 3 FP32): *loss (device FP64) = sum |F - target|^2 / (3 S n),
	// and grad4 (S x n float4) = (F - target) 2 / (3 S n), the loss's dL/dF.
	void mse(const float* force4, const float* target3, float* grad4, double* loss, hipStream_t s);
	// The same loss and dL/dF of the forces for the given knots, each snapshot in its
	// sorted order (target3 too) and without storing the forces: grad4 has mse()'s bits;
	// the loss sums the same FP64 terms in another fixed grouping.
	void loss_sorted(const float* knots, const float* target3, float* grad4, double* loss, hipStream_t s);
	// dL/dknot (device FP64) for dL/dF = grad4 (S x n float4, .w ignored; in sorted
	// order when sorted, else particle order).
	void adjoint(const float* knots, const float* grad4, double* out, hipStream_t s, bool sorted = false);
};

} // namespace mmd
