// Host-side scene store and Rendu's BVH build, plus the flattening into the device layout.
//
// The tree is Rendu's own (src/engine/raycaster/Raycaster.cpp:54-142): one root per mesh, mean
// centroid split on the largest axis with an nth_element median fallback, leaves of 1-2
// triangles, nodes appended in the order of a LIFO work stack. The device layout keeps every
// box and every leaf of that tree; only the storage changes (see DeviceLayout).
#pragma once
#include "rdu_vec.hpp"
#include <cstddef>
#include <cstdint>
#include <vector>

namespace rdu {

struct HBox {
	f3 mn, mx;
};

// One triangle as the build moves it (Raycaster.hpp:117-124 TriangleInfos, plus its centroid,
// which Rendu recomputes as 0.5*(min+max) at every comparison: storing it changes no value).
struct HTri {
	HBox box;
	f3 c;
	uint32_t v0, v1, v2;
	uint32_t localId;
	uint32_t meshId;
};

// Raycaster.hpp:127-132 Node.
struct HNode {
	HBox box;
	uint64_t left = 0, right = 0;
	bool leaf = true;
};

// Device layout, all float4-aligned arrays:
//  nodes: 4 float4 per internal node (depth-first preorder, left subtree first):
//         {lmin.xyz, lmax.x} {lmax.yz, rmin.xy} {rmin.z, rmax.xyz} {lcode, rcode, 0, 0 as int bits}
//         child code >= 0: internal node index; < 0: leaf ~(first << 1 | (count - 1)).
//  tris:  3 float4 per triangle in Rendu's permuted order:
//         {v0.xyz, e1.x} {e1.yz, e2.xy} {e2.z, meshId, localId, 0 as int bits}, e1 = v1 - v0 and
//         e2 = v2 - v0 computed in float exactly as the Moller-Trumbore test does.
//  roots: 2 float4 per mesh {min.xyz, code} {max.xyz, wide code}.
//  wide:  the same tree with every internal node's children expanded (largest surface area
//         first) until it holds up to 4 of Rendu's nodes: 8 float4 per 4-wide node,
//         {lo.x[4]} {hi.x[4]} {lo.y[4]} {hi.y[4]} {lo.z[4]} {hi.z[4]} {code[4]} {pad}; every box is
//         a box of Rendu's tree, so a child passes the slab test only if all the boxes between it
//         and its 4-wide parent do (the test is monotone in the box under IEEE rounding); an
//         unused slot has code kWideEmpty (rdu_vec.hpp). Used by the near-first kernels.

struct DeviceLayout {
	std::vector<float> nodes, tris, roots, wide;
	// flatten(out, true): each mesh root collapsed on its own, with no level above the meshes; the
	// 4-wide code of mesh m's root (an instanced scene's per-mesh tree).
	std::vector<int32_t> meshWideCodes;
	uint32_t internalCount = 0;
	uint32_t maxDepth = 0;
	uint32_t wideCount = 0;
	uint32_t wideDepth = 0;
	HBox wideTopBox{{0, 0, 0}, {0, 0, 0}}; // the single root of the 4-wide tree (top level over meshes)
	int32_t wideTopCode = 0;
	f3 sceneMin{0, 0, 0}, sceneMax{0, 0, 0};
};

class HostScene {
public:
	std::vector<f3> vertices;
	std::vector<HTri> tris;
	std::vector<HNode> nodes;
	uint32_t meshCount = 0;
	bool built = false;

	// Raycaster.cpp:14-52. model: glm column-major 4x4 or nullptr.
	void addMesh(const float * xyz, size_t nv, const uint32_t * idx, size_t ni, const float * model);
	// Raycaster.cpp:54-142.
	void build();
	// perMeshRoots: no top level over the meshes; every mesh root gets a 4-wide root of its own
	// (out.meshWideCodes), for trees shared by instances.
	void flatten(DeviceLayout & out, bool perMeshRoots = false) const;
	// Sizes the stores for meshes still to come (vertices, triangles, meshes in total), so that
	// adding many meshes does not regrow them mesh by mesh. Changes no content.
	void reserve(size_t nVertices, size_t nTriangles, size_t nMeshes);
};

// Top level of an instanced scene: a 4-wide tree over object boxes (median split of the box
// centres on the largest axis, as the top level over mesh roots), appended to `wide` (8 float4 per
// node, wideCount counting them); object i's leaf code is leafCodes[i] (an instance code
// kInstCode0 + i, or the 4-wide code of a tree already in world space). Returns the root's code
// (object 0's leaf code when there is one object) and sets its box and the tree's depth.
int32_t buildInstanceTop(const std::vector<HBox> & boxes, const std::vector<int32_t> & leafCodes, std::vector<float> & wide,
	uint32_t & wideCount, uint32_t & depth, HBox & topBox);

} // namespace rdu
