// ORACLE — test infrastructure only (see raycaster.cpp's header).
//
// The oracle's restatement of Rendu's raycaster (src/engine/raycaster/Raycaster.cpp,
// Intersection.cpp, src/engine/resources/Bounds.cpp) and of System::forParallel
// (src/engine/system/System.hpp:72-104), shared by raycaster.cpp and pathtracer.cpp.
#pragma once
#include "omath.hpp"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cstring>
#include <limits>
#include <stack>
#include <thread>
#include <vector>

namespace orc {

// Bounds.hpp:98-99, Bounds.cpp:12-20, 33-35, 61-63.
struct Box {
	v3 minis = v3(std::numeric_limits<float>::max());
	v3 maxis = v3(std::numeric_limits<float>::lowest());
	Box() = default;
	Box(v3 a, v3 b, v3 c) {
		minis = gmin(gmin(a, b), c);
		maxis = gmax(gmax(a, b), c);
	}
	void merge(const Box & o) {
		minis = gmin(minis, o.minis);
		maxis = gmax(maxis, o.maxis);
	}
	v3 size() const { return maxis - minis; }
	v3 centroid() const { return 0.5f * (minis + maxis); }
};

// Raycaster.hpp:117-124.
struct TriangleInfos {
	Box box;
	unsigned long v0 = 0, v1 = 0, v2 = 0;
	unsigned long localId = 0;
	unsigned int meshId = 0;
};

// Raycaster.hpp:127-132.
struct Node {
	Box box;
	size_t left = 0;
	size_t right = 0;
	bool leaf = true;
};

// Intersection.hpp:10-20, Intersection.cpp:3-5.
struct Ray {
	v3 pos, dir, invdir;
	Ray(v3 o, v3 d) : pos(o), dir(normalize(d)), invdir(1.0f / normalize(d)) {}
};

// Raycaster.hpp:16-42, Raycaster.cpp:6-12.
struct Hit {
	bool hit = false;
	float dist = std::numeric_limits<float>::max();
	float u = 0.f, v = 0.f, w = 0.f;
	unsigned long localId = 0, meshId = 0;
};

// Intersection.cpp:22-32. std::max(a,b) = (a < b) ? b : a; std::min(a,b) = (b < a) ? b : a.
inline bool boxTest(const Ray & ray, const Box & box, float mini, float maxi) {
	const v3 minRatio = (box.minis - ray.pos) * ray.invdir;
	const v3 maxRatio = (box.maxis - ray.pos) * ray.invdir;
	const v3 minFinal = gmin(minRatio, maxRatio);
	const v3 maxFinal = gmax(minRatio, maxRatio);
	const float closest = gmax(minFinal[0], gmax(minFinal[1], minFinal[2]));
	const float furthest = gmin(maxFinal[0], gmin(maxFinal[1], maxFinal[2]));
	return gmax(closest, mini) <= gmin(furthest, maxi);
}

struct Stats {
	uint64_t nodes = 0; // internal-node expansions (two child box tests each)
	uint64_t tris = 0;	// Moller-Trumbore tests
};

class Raycaster {
public:
	std::vector<TriangleInfos> triangles;
	std::vector<v3> vertices;
	std::vector<Node> hierarchy;
	unsigned int meshCount = 0;
	bool built = false;

	// Raycaster.cpp:14-52. model is glm column-major (m[c][r] = model[4*c + r]); identity skips.
	void addMesh(const float * xyz, size_t nv, const uint32_t * idx, size_t ni, const float * model) {
		const unsigned long indexOffset = static_cast<unsigned long>(vertices.size());
		const size_t start = vertices.size();
		for(size_t i = 0; i < nv; ++i)
			vertices.emplace_back(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
		const size_t end = vertices.size();
		bool identity = true;
		if(model) {
			for(int c = 0; c < 4; ++c)
				for(int r = 0; r < 4; ++r)
					identity = identity && (model[4 * c + r] == (c == r ? 1.0f : 0.0f));
		}
		if(model && !identity) {
			for(size_t vid = start; vid < end; ++vid) {
				const v3 p = vertices[vid];
				// glm mat4 * vec4 (type_mat4x4.inl:558-569): (m0*x + m1*y) + (m2*z + m3*1)
				float out[3];
				for(int r = 0; r < 3; ++r) {
					const float a0 = model[0 * 4 + r] * p.x + model[1 * 4 + r] * p.y;
					const float a1 = model[2 * 4 + r] * p.z + model[3 * 4 + r] * 1.0f;
					out[r] = a0 + a1;
				}
				vertices[vid] = v3(out[0], out[1], out[2]);
			}
		}
		const size_t startTriangle = triangles.size();
		const size_t count = ni / 3;
		for(size_t tid = 0; tid < count; ++tid) {
			const size_t localId = 3 * tid;
			TriangleInfos t;
			t.v0 = indexOffset + idx[localId + 0];
			t.v1 = indexOffset + idx[localId + 1];
			t.v2 = indexOffset + idx[localId + 2];
			t.localId = static_cast<unsigned long>(localId);
			t.meshId = meshCount;
			t.box = Box(vertices[t.v0], vertices[t.v1], vertices[t.v2]);
			triangles.push_back(t);
		}
		hierarchy.emplace_back();
		hierarchy.back().left = startTriangle;
		hierarchy.back().right = count;
		++meshCount;
	}

	// Raycaster.cpp:54-142.
	void updateHierarchy() {
		struct SetInfos {
			size_t id, begin, count;
		};
		std::stack<SetInfos> remaining;
		for(size_t nid = 0; nid < meshCount; ++nid)
			remaining.push({nid, hierarchy[nid].left, hierarchy[nid].right});
		while(!remaining.empty()) {
			const SetInfos cur = remaining.top();
			remaining.pop();
			const size_t begin = cur.begin, count = cur.count;
			Box global(triangles[begin].box);
			for(size_t tid = 1; tid < count; ++tid)
				global.merge(triangles[begin + tid].box);
			Node & node = hierarchy[cur.id];
			node.box = global;
			if(count < 3) {
				node.leaf = true;
				node.left = begin;
				node.right = count;
				continue;
			}
			node.leaf = false;
			size_t splitCount = 0;
			const v3 s = global.size();
			const int axis = (s.x >= s.y && s.x >= s.z) ? 0 : (s.y >= s.z ? 1 : 2);
			if(count >= 5) {
				float abscisse = 0.0f;
				for(size_t tid = 0; tid < count; ++tid)
					abscisse += triangles[begin + tid].box.centroid()[axis];
				abscisse /= float(count);
				const auto split = std::partition(triangles.begin() + begin, triangles.begin() + begin + count,
					[abscisse, axis](const TriangleInfos & t0) { return t0.box.centroid()[axis] < abscisse; });
				splitCount = std::distance(triangles.begin() + begin, split);
			}
			if(splitCount == 0 || splitCount == count || count < 5) {
				splitCount = count / 2;
				std::nth_element(triangles.begin() + begin, triangles.begin() + begin + splitCount, triangles.begin() + begin + count,
					[axis](const TriangleInfos & t0, const TriangleInfos & t1) { return t0.box.centroid()[axis] < t1.box.centroid()[axis]; });
			}
			hierarchy.emplace_back();
			const size_t leftPos = hierarchy.size() - 1;
			hierarchy[cur.id].left = leftPos;
			remaining.push({leftPos, begin, splitCount});
			hierarchy.emplace_back();
			const size_t rightPos = hierarchy.size() - 1;
			hierarchy[cur.id].right = rightPos;
			remaining.push({rightPos, begin + splitCount, count - splitCount});
		}
		built = true;
	}

	// Raycaster.cpp:232-262.
	Hit triTest(const Ray & ray, const TriangleInfos & tri, float mini, float maxi) const {
		const v3 & v0 = vertices[tri.v0];
		const v3 v01 = vertices[tri.v1] - v0;
		const v3 v02 = vertices[tri.v2] - v0;
		const v3 p = cross(ray.dir, v02);
		const float det = dot(v01, p);
		if(std::abs(det) < std::numeric_limits<float>::epsilon())
			return {};
		const float invDet = 1.0f / det;
		const v3 q = ray.pos - v0;
		const float u = invDet * dot(q, p);
		if(u < 0.0f || u > 1.0f)
			return {};
		const v3 r = cross(q, v01);
		const float v = invDet * dot(ray.dir, r);
		if(v < 0.0f || (u + v) > 1.0f)
			return {};
		const float t = invDet * dot(v02, r);
		if(t > mini && t < maxi) {
			Hit h;
			h.hit = true;
			h.dist = t;
			h.u = u;
			h.v = v;
			h.w = 1.0f - u - v;
			h.localId = tri.localId;
			h.meshId = tri.meshId;
			return h;
		}
		return {};
	}

	// Raycaster.cpp:144-185 (anyHit=false) and 187-224 (anyHit=true).
	Hit traverse(v3 o, v3 d, float mini, float maxi, bool anyHit, Stats * st) const {
		const Ray ray(o, d);
		std::stack<size_t> nodes;
		for(size_t nid = 0; nid < meshCount; ++nid) {
			if(!boxTest(ray, hierarchy[nid].box, mini, maxi))
				continue;
			nodes.push(nid);
		}
		Hit best;
		while(!nodes.empty()) {
			const Node & node = hierarchy[nodes.top()];
			nodes.pop();
			if(node.leaf) {
				for(size_t tid = 0; tid < node.right; ++tid) {
					if(st)
						st->tris++;
					const Hit h = triTest(ray, triangles[node.left + tid], mini, maxi);
					if(anyHit && h.hit)
						return h;
					if(h.hit && h.dist < best.dist) {
						best = h;
						maxi = best.dist;
					}
				}
				continue;
			}
			if(st)
				st->nodes++;
			if(boxTest(ray, hierarchy[node.left].box, mini, maxi))
				nodes.push(node.left);
			if(boxTest(ray, hierarchy[node.right].box, mini, maxi))
				nodes.push(node.right);
		}
		return best;
	}

	// Every triangle through the same Moller-Trumbore test, first strict minimum in addMesh order
	// (mesh, then triangle). Used only to pin traversal; not a reference function.
	Hit brute(v3 o, v3 d, float mini, float maxi) const {
		const Ray ray(o, d);
		std::vector<const TriangleInfos *> order(triangles.size());
		for(size_t i = 0; i < triangles.size(); ++i)
			order[i] = &triangles[i];
		std::sort(order.begin(), order.end(), [](const TriangleInfos * a, const TriangleInfos * b) {
			return a->meshId != b->meshId ? a->meshId < b->meshId : a->localId < b->localId;
		});
		Hit best;
		for(const TriangleInfos * t : order) {
			const Hit h = triTest(ray, *t, mini, maxi);
			if(h.hit && h.dist < best.dist) {
				best = h;
				maxi = best.dist;
			}
		}
		return best;
	}

	int maxDepth() const {
		int best = 0;
		std::vector<std::pair<size_t, int>> st;
		for(size_t r = 0; r < meshCount; ++r)
			st.push_back({r, 1});
		while(!st.empty()) {
			auto cur = st.back();
			st.pop_back();
			best = std::max(best, cur.second);
			const Node & n = hierarchy[cur.first];
			if(!n.leaf) {
				st.push_back({n.left, cur.second + 1});
				st.push_back({n.right, cur.second + 1});
			}
		}
		return best;
	}
};

// System.hpp:72-104: max(hw-1, 1) threads over contiguous spans, the last one taking the tail.
template<typename F>
void forParallel(size_t low, size_t high, int threadsWanted, F func) {
	if(high < low)
		std::swap(low, high);
	if(high == low)
		return;
	size_t avail = threadsWanted > 0 ? size_t(threadsWanted)
									 : size_t(std::max(int(std::thread::hardware_concurrency()) - 1, 1));
	const size_t count = std::min(avail, high - low);
	const size_t span = std::max(size_t(1), (high - low) / count);
	std::vector<std::thread> threads;
	threads.reserve(count);
	for(size_t tid = 0; tid < count; ++tid) {
		const size_t a = low + tid * span;
		const size_t b = tid == (count - 1) ? high : low + (tid + 1) * span;
		threads.emplace_back([&func, a, b]() {
			for(size_t i = a; i < b; ++i)
				func(i);
		});
	}
	for(auto & t : threads)
		t.join();
}

} // namespace orc
