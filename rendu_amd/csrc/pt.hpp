// Host-visible interface of the wavefront path tracer kernels (pt.hip).
#pragma once
#include "pt_device.hpp"

namespace rdu {

// Per-render constants: image size, random stream key, stratification grid and camera frame
// (Camera::pixelShifts, src/engine/input/Camera.cpp:42-49).
struct PtFrame {
	uint32_t width = 0, height = 0;
	uint32_t seed = 0;
	int32_t cellX = 1;
	float cellW = 1.0f, cellH = 1.0f;
	f3 eye{0, 0, 0}, corner{0, 0, 0}, dx{0, 0, 0}, dy{0, 0, 0};
};

// Device counters (uint32 each).
enum : int { kCountActive = 0, kCountNext = 1, kCountShadow = 2, kCountShadowNext = 3, kCountSlots = 4 };

// Path state, structure of arrays indexed by path id p = sample_in_pass * pixels + pixel.
struct PtPaths {
	float4 * rays = nullptr;        // 2 float4 per path: rdu_ray
	rdu_hit * hits = nullptr;
	float4 * atten = nullptr;       // throughput
	float4 * color = nullptr;       // accumulated sample colour
	uint32_t * queue[2] = {nullptr, nullptr};
	uint32_t * counts = nullptr;    // kCountSlots
	unsigned long long * totals = nullptr; // [extension rays, shadow rays]
	float4 * shadowRays = nullptr;  // 2 float4 per path
	float4 * shadowContrib = nullptr;
	uint32_t * shadowQueue[2] = {nullptr, nullptr};
	uint8_t * occluded = nullptr;
	rdu_hit * shadowHits = nullptr;
};

struct PtOptions {
	uint32_t depth = 5;
	int32_t estimator = RDU_ESTIMATOR_REFERENCE;
	int32_t rrStart = 3;
	int32_t maxShadowLayers = 16;
};

struct PtScratch {
	int * gstack = nullptr; // traversal stack spill, globalStackInts() for the pass size
};

int ptRenderPass(const DevScene & S, const DevBVH & bvh, const PtFrame & f, const PtPaths & P, uint32_t sampleBase,
	uint32_t samplesInPass, const PtOptions & opt, float * dImage, uint32_t components, PtScratch & scratch,
	hipStream_t stream);
int ptNormalizeImage(float * dImage, uint32_t pixels, uint32_t components, uint32_t spp, hipStream_t stream);

} // namespace rdu
