// Device-side ray/BVH traversal: shared primitives (usable from other kernels of the library)
// and the launchers of the batch kernels.
#pragma once
#include "rdu_vec.hpp"
#include <cstddef>
#include <cstdint>

namespace rdu {

// Device copy of a flattened BVH (layout documented in bvh.hpp).
struct DevBVH {
	const float4 * nodes = nullptr;
	const float4 * tris = nullptr;
	const float4 * roots = nullptr;
	const float4 * wide = nullptr;
	// Instanced scenes: 4 float4 per instance, rows 0-2 of the inverse model matrix (object from
	// world) and {root code of its mesh's tree, object index, 0, 0 as int bits}; the top level's
	// leaves carry kInstCode0 + instance. Null for flattened scenes.
	const float4 * inst = nullptr;
	uint32_t nRoots = 0;
	uint32_t maxDepth = 0;
	uint32_t wideDepth = 0;
	float sceneMin[3] = {0, 0, 0}, sceneMax[3] = {0, 0, 0};
	// Single root of the 4-wide tree: a top level over the mesh roots when there are several.
	float wideTopMin[3] = {0, 0, 0}, wideTopMax[3] = {0, 0, 0};
	int32_t wideTopCode = 0;
};

struct DevRay {
	f3 pos, dir, inv;
	float mini, maxi;
};

// Ray::Ray (Intersection.cpp:3-5): dir = normalize(d), invdir = 1 / normalize(d).
RDU_HD DevRay makeRay(f3 o, f3 d, float mini, float maxi) {
	DevRay r;
	r.pos = o;
	r.dir = normalize(d);
	r.inv = rcp3(normalize(d));
	r.mini = mini;
	r.maxi = maxi;
	return r;
}

// Intersection::box (Intersection.cpp:22-32): slab test with select-based min/max.
RDU_HD bool slab(const DevRay & r, f3 bmin, f3 bmax, float mini, float maxi, float & entry) {
	const f3 a = (bmin - r.pos) * r.inv;
	const f3 b = (bmax - r.pos) * r.inv;
	const f3 mn = gmin(a, b);
	const f3 mx = gmax(a, b);
	const float closest = gmax(mn.x, gmax(mn.y, mn.z));
	const float furthest = gmin(mx.x, gmin(mx.y, mx.z));
	entry = gmax(closest, mini);
	return entry <= gmin(furthest, maxi);
}

// Moller-Trumbore (Raycaster.cpp:232-262) on a precomputed {v0, e1 = v1 - v0, e2 = v2 - v0}.
RDU_HD bool mollerTrumbore(const DevRay & r, f3 v0, f3 e1, f3 e2, float mini, float maxi, float & t, float & u,
	float & v) {
	const f3 p = cross(r.dir, e2);
	const float det = dot(e1, p);
	if(__builtin_fabsf(det) < 1.19209290e-07f) // FLT_EPSILON
		return false;
	const float invDet = 1.0f / det;
	const f3 q = r.pos - v0;
	u = invDet * dot(q, p);
	if(u < 0.0f || u > 1.0f)
		return false;
	const f3 rr = cross(q, e1);
	v = invDet * dot(r.dir, rr);
	if(v < 0.0f || (u + v) > 1.0f)
		return false;
	t = invDet * dot(e2, rr);
	return t > mini && t < maxi;
}

// How a batch is traversed.
struct TraceOptions {
	bool referenceOrder = false; // Rendu's exact depth-first order over the binary tree
	bool wide = true;            // near-first order over the 4-wide collapse (else binary)
	const uint32_t * perm = nullptr; // optional ray permutation: lane i handles ray perm[i]
	const uint32_t * countDev = nullptr; // optional device count: only the first *countDev entries
	bool spatialOrder = false; // perm is a spatial (Morton) order: interleave wavefronts finely
	uint32_t refillMin = 0;    // idle lanes before a wavefront refills (0 = default for the order)
	uint32_t chunk = 0;        // rays per interleaved run of a persistent wavefront (0 = default for the order)
	uint32_t blocksPerCu = 0;  // resident workgroups per CU of the persistent grid (0 = occupancy limit)
	bool nearFar = false;      // fetch node planes ordered near/far by the ray's direction signs (fewer
	                           // instructions per node, lower occupancy: for incoherent rays); any hit
	                           // uses it when unorderedAny is off
	bool unorderedAny = false; // any hit: visit the children that pass in node order, no near-first sort
	                           // (for coherent rays)
	// Batches of several scenes over one node arena (persistent 4-wide kernels only): ray k starts
	// at the root of scene rayJob[k], jobRoots[2 j] = {top box min, root code bits}, [2 j + 1] =
	// {top box max, 0}.
	const uint32_t * rayJob = nullptr;
	const float4 * jobRoots = nullptr;
};

// Launch parameters of the persistent traversal picked by the autotuner (0 = default).
struct TraceTuning {
	uint32_t refillMin = 0, chunk = 0, blocksPerCu = 0;
	bool operator==(const TraceTuning & o) const {
		return refillMin == o.refillMin && chunk == o.chunk && blocksPerCu == o.blocksPerCu;
	}
};
inline void applyTuning(TraceOptions & o, const TraceTuning & t) {
	o.refillMin = t.refillMin;
	o.chunk = t.chunk;
	o.blocksPerCu = t.blocksPerCu;
}
// Workgroups per CU the persistent closest-hit (any = false) or any-hit kernel can keep resident.
int persistentOccupancy(bool any);

// Launchers. rays: device rdu_ray array (float4 pairs); outputs are device pointers, indexed by
// the original ray index. Return hipError_t as int.
int launchClosest(const DevBVH & bvh, const float4 * rays, size_t n, void * hits, const TraceOptions & opt,
	int * gstack, uint64_t * stats, hipStream_t stream);
int launchAny(const DevBVH & bvh, const float4 * rays, size_t n, uint8_t * occluded, const TraceOptions & opt,
	int * gstack, uint64_t * stats, hipStream_t stream);

// Number of stack entries the kernels keep in LDS per ray; deeper stacks spill to `gstack`,
// which then needs globalStackInts() ints for the launch.
constexpr int kLdsStack = 32;
int ldsStackDepth();
// Worst-case stack entries of a traversal: binary tree depth + 2, or 3 per 4-wide level + 2.
int stackNeed(const DevBVH & bvh, const TraceOptions & opt);
size_t globalStackInts(const DevBVH & bvh, const TraceOptions & opt, size_t n);

// Ray reordering for incoherent batches: keys = Morton code of the origin cell in the scene box
// (kOrderBits per axis), sorted with the ray index. perm receives the ray indices in key order.
constexpr int kOrderBits = 7;
size_t sortTempBytes(size_t n);
int launchRayOrder(const DevBVH & bvh, const float4 * rays, size_t n, uint32_t * keys, uint32_t * keysAlt,
	uint32_t * perm, uint32_t * permAlt, void * temp, size_t tempBytes, hipStream_t stream);

} // namespace rdu
