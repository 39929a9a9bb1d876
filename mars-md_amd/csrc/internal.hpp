// internal.hpp — host-side internals shared by the library's translation units:
// error plumbing, device buffers, and the launch wrappers each .hip file exports.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/marsmd.h"
#include "device_common.hpp"

namespace mmd {

// MARS::Exception counterpart: carries a marsmd_status code.
struct Error : std::runtime_error {
	int code;
	Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void raise(int code, const std::string& msg) { throw Error(code, msg); }

#define MMD_HIP(call)                                                                         \
	do {                                                                                      \
		hipError_t e_ = (call);                                                               \
		if (e_ != hipSuccess)                                                                 \
			::mmd::raise(MARSMD_ERR_DEVICE, std::string(#call " failed: ") + hipGetErrorString(e_) + \
												" at " + __FILE__ + ":" + std::to_string(__LINE__)); \
	} while (0)

#define MMD_LAUNCH_CHECK() MMD_HIP(hipGetLastError())

inline unsigned grid_for(uint64_t n, unsigned block) {
	const uint64_t g = (n + block - 1) / block;
	if (g > 0x7FFFFFFFull)
		raise(MARSMD_ERR_VALUE, "launch too large");
	return unsigned(g == 0 ? 1 : g);
}

// RAII device allocation (DeviceBuffer<T> counterpart, src/Backend/Buffer.h:123+).
template<class T>
struct DevArray {
	T* ptr = nullptr;
	uint64_t count = 0;
	DevArray() = default;
	explicit DevArray(uint64_t n) { resize(n); }
	DevArray(const DevArray&) = delete;
	DevArray& operator=(const DevArray&) = delete;
	DevArray(DevArray&& o) noexcept : ptr(o.ptr), count(o.count) {
		o.ptr = nullptr;
		o.count = 0;
	}
	DevArray& operator=(DevArray&& o) noexcept {
		if (this != &o) {
			release();
			ptr = o.ptr;
			count = o.count;
			o.ptr = nullptr;
			o.count = 0;
		}
		return *this;
	}
	~DevArray() { release(); }
	void release() {
		if (ptr)
			(void)hipFree(ptr);
		ptr = nullptr;
		count = 0;
	}
	// Grow-only: keeps the allocation when it is already large enough.
	void resize(uint64_t n) {
		if (n <= count && ptr)
			return;
		release();
		if (n == 0)
			return;
		MMD_HIP(hipMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T)));
		count = n;
	}
	T* data() const { return ptr; }
	void swap(DevArray& o) {
		std::swap(ptr, o.ptr);
		std::swap(count, o.count);
	}
};

// Pinned host words for device read-backs: a copy into them is asynchronous, so a
// read-back costs one stream synchronisation (a copy into pageable memory waits itself).
struct PinnedWords {
	uint32_t* ptr = nullptr;
	explicit PinnedWords(size_t n) { MMD_HIP(hipHostMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(uint32_t))); }
	PinnedWords(const PinnedWords&) = delete;
	PinnedWords& operator=(const PinnedWords&) = delete;
	~PinnedWords() {
		if (ptr)
			(void)hipHostFree(ptr);
	}
};

// ----------------------------------------------------------- sort.hip
uint64_t radix_sort_scratch_bytes(uint64_t n);
// Stable LSD radix sort of (key, value) over bits [0, key_bits). keys_out/vals_out
// receive the result; the inputs are left untouched.
void radix_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
					  uint32_t* vals_out, uint64_t n, int key_bits, void* scratch,
					  uint64_t scratch_bytes, hipStream_t s);
// The same sort with the library's own fixed device workspace instead of a caller's:
// no allocation, at most kStaticSortMaxKeys keys (MARSMD_ERR_VALUE beyond). All calls
// on one GPU share that workspace, so calls that overlap in time on one GPU race.
constexpr uint64_t kStaticSortMaxKeys = uint64_t(1) << 21;
void radix_sort_pairs_static(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
							 uint32_t* vals_out, uint64_t n, int key_bits, hipStream_t s);
uint64_t static_sort_scratch_bytes();
uint64_t scan_scratch_bytes(uint64_t n);
// Exclusive scan of n counts into out[0..n]; out[n] holds the total.
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, void* scratch,
						hipStream_t s);
void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
void fill_f32(float* p, float v, uint64_t n, hipStream_t s);

} // namespace mmd
