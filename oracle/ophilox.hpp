// ORACLE — test infrastructure only (see raycaster.cpp's header).
//
// Counter-based random numbers for the oracle's path tracer: Philox4x32-10 (Salmon, Moraes,
// Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), restated from the paper's
// definition. The reference draws from per-thread std::mt19937 seeded in thread-creation order
// (src/engine/generation/Random.cpp:72-85), which no second implementation can reproduce, so
// the oracle and the GPU share this keyed stream instead and can be compared sample by sample.
// Pinned by the Random123 known-answer vectors in tests/test_oracle_rng.py.
#pragma once
#include <cstdint>

namespace orc {

struct u4 {
	uint32_t x, y, z, w;
};

inline void mulhilo(uint32_t a, uint32_t b, uint32_t & hi, uint32_t & lo) {
	const uint64_t p = uint64_t(a) * uint64_t(b);
	hi = uint32_t(p >> 32);
	lo = uint32_t(p);
}

inline u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1) {
	const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
	const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
	for(int r = 0; r < 10; ++r) {
		uint32_t hi0, lo0, hi1, lo1;
		mulhilo(M0, c.x, hi0, lo0);
		mulhilo(M1, c.z, hi1, lo1);
		c = {hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
		k0 += W0;
		k1 += W1;
	}
	return c;
}

// 24 high bits -> [0, 1): exactly representable, never 1.
inline float u01(uint32_t x) { return float(x >> 8) * (1.0f / 16777216.0f); }

// Stream of one path sample: key = (seed, pixel), counter = (sample, bounce, 0, 0).
// Each call to draw4 yields 4 uniforms for one bounce (or for the camera jitter at bounce ~0u).
struct PathRng {
	uint32_t seed, pixel, sample;
	inline void draw4(uint32_t slot, float out[4]) const {
		const u4 r = philox4x32_10({sample, slot, 0x52445531u, 0u}, seed, pixel);
		out[0] = u01(r.x);
		out[1] = u01(r.y);
		out[2] = u01(r.z);
		out[3] = u01(r.w);
	}
};

} // namespace orc
