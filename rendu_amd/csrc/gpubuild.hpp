// GPU BVH build (gpubuild.hip): a linear BVH over all triangles, collapsed to the 4-wide near-first
// layout of bvh.hpp. An alternative to the host build of Rendu's tree, not a reproduction of it.
#pragma once
#include "bvh.hpp"
#include "capi_common.hpp"

#include <hip/hip_runtime.h>

namespace rdu {

struct GpuBvh {
	DevBuffer tris;  // 3 float4 per triangle (layout of DeviceLayout::tris), Morton order
	DevBuffer wide;  // 8 float4 per 4-wide node, root at slot 0
	uint32_t triangles = 0, innerCount = 0, wideCount = 0, wideDepth = 0;
	int32_t topCode = 0; // 0 (wide root) or the leaf code of a one-triangle scene
	float topMin[3] = {0, 0, 0}, topMax[3] = {0, 0, 0};
	float sceneMin[3] = {0, 0, 0}, sceneMax[3] = {0, 0, 0};
};

// Builds from the scene's transformed vertices and triangle list; returns RDU_OK or an RDU_ERR_* status.
int gpuBuildBvh(const HostScene & hs, GpuBvh & out, hipStream_t stream);

} // namespace rdu
