// Driver for the reference's own Philox header (compiled from the reference tree by
// oracle/Makefile into oracle/_ref/). Prints draws of openrand::Philox for fixed
// (seed, ctr, global_seed, ctr1) tuples so the oracle's restatement can be pinned
// against them. TEST INFRASTRUCTURE ONLY.
#include "philox.h"

#include <cinttypes>
#include <cstdio>

int main() {
	struct Case {
		uint64_t seed;
		uint32_t ctr, global_seed, ctr1;
	};
	const Case cases[] = {
		{0ull, 0u, 0u, 0u},
		{0xFFFFFFFFFFFFFFFFull, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu},
		{0xA4093822299F31D0ull, 0x243F6A88u, 0x13198A2Eu, 0x85A308D3u},
		{42ull, 7u, 1u, 0x5324120u},
		{12345ull, 0u, 0u, 0x12345u},
		{0x9E3779B97F4A7C15ull, 123456u, 1000u, 0x1356914u},
		{214ull, 999999u, 20u, 0x5324120u},
	};
	std::printf("[\n");
	bool first = true;
	for (const Case& c : cases) {
		openrand::Philox rng(c.seed, c.ctr, c.global_seed, c.ctr1);
		for (int d = 0; d < 3; ++d) {
			const openrand::uint4 u = rng.draw_int4();
			std::printf("%s  {\"seed\": %" PRIu64 ", \"ctr\": %u, \"global_seed\": %u, \"ctr1\": %u, "
						"\"draw\": %d, \"out\": [%u, %u, %u, %u]}",
						first ? "" : ",\n", c.seed, c.ctr, c.global_seed, c.ctr1, d, u.x, u.y, u.z, u.w);
			first = false;
		}
		openrand::Philox frng(c.seed, c.ctr, c.global_seed, c.ctr1);
		const openrand::float4 f = frng.draw_float4();
		std::printf(",\n  {\"seed\": %" PRIu64 ", \"ctr\": %u, \"global_seed\": %u, \"ctr1\": %u, "
					"\"float4\": [%.9g, %.9g, %.9g, %.9g]}",
					c.seed, c.ctr, c.global_seed, c.ctr1, (double)f.x, (double)f.y, (double)f.z, (double)f.w);
	}
	std::printf("\n]\n");
	return 0;
}
