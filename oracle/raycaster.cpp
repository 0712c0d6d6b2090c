// ORACLE — test infrastructure only. Nothing in the product (rendu_amd/) may include, link
// or call this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
// and only as the checker or the timed CPU baseline.
//
// CPU restatement of Rendu's BVH raycaster (src/engine/raycaster/Raycaster.cpp,
// Intersection.cpp, src/engine/resources/Bounds.cpp). Same data flow, same operation order,
// same libstdc++ std::partition / std::nth_element calls (GCC 11.4 libstdc++, the one the
// reference would be built against here) so that the BVH node array and every hit are the
// reference's own. Exposed through a small extern "C" API for ctypes.
//
// Pinning (tests/test_oracle_raycaster.py): closed-form Moller-Trumbore KATs, IEEE edge cases
// of the slab test, brute force over every triangle on seeded scenes, and the structural
// invariants of updateHierarchy (partition rule, leaf sizes, exact merged boxes, LIFO node order).
#include "omath.hpp"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cstring>
#include <limits>
#include <stack>
#include <thread>
#include <vector>

using namespace orc;

namespace {

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

} // namespace

extern "C" {

void * orc_rc_create() { return new Raycaster(); }
void orc_rc_destroy(void * rc) { delete static_cast<Raycaster *>(rc); }

int orc_rc_add_mesh(void * h, const float * xyz, size_t nv, const uint32_t * idx, size_t ni, const float * model16) {
	Raycaster * rc = static_cast<Raycaster *>(h);
	if(rc->built)
		return -1;
	rc->addMesh(xyz, nv, idx, ni, model16);
	return 0;
}

int orc_rc_update(void * h) {
	Raycaster * rc = static_cast<Raycaster *>(h);
	for(size_t m = 0; m < rc->meshCount; ++m)
		if(rc->hierarchy[m].right == 0)
			return -2; // a 0-triangle mesh reads _triangles[begin] out of range in the reference
	rc->updateHierarchy();
	return 0;
}

size_t orc_rc_node_count(void * h) { return static_cast<Raycaster *>(h)->hierarchy.size(); }
size_t orc_rc_tri_count(void * h) { return static_cast<Raycaster *>(h)->triangles.size(); }
int orc_rc_max_depth(void * h) { return static_cast<Raycaster *>(h)->maxDepth(); }

// boxes: N*6 floats (min xyz, max xyz); left/right: N; leaf: N.
void orc_rc_nodes(void * h, float * boxes, uint64_t * left, uint64_t * right, uint8_t * leaf) {
	const Raycaster * rc = static_cast<Raycaster *>(h);
	for(size_t i = 0; i < rc->hierarchy.size(); ++i) {
		const Node & n = rc->hierarchy[i];
		const float b[6] = {n.box.minis.x, n.box.minis.y, n.box.minis.z, n.box.maxis.x, n.box.maxis.y, n.box.maxis.z};
		std::memcpy(boxes + 6 * i, b, sizeof(b));
		left[i] = n.left;
		right[i] = n.right;
		leaf[i] = n.leaf ? 1 : 0;
	}
}

// Permuted triangle list after the build: global vertex ids, localId, meshId, box.
void orc_rc_triangles(void * h, uint64_t * vids, uint64_t * localId, uint32_t * meshId, float * boxes) {
	const Raycaster * rc = static_cast<Raycaster *>(h);
	for(size_t i = 0; i < rc->triangles.size(); ++i) {
		const TriangleInfos & t = rc->triangles[i];
		vids[3 * i] = t.v0;
		vids[3 * i + 1] = t.v1;
		vids[3 * i + 2] = t.v2;
		localId[i] = t.localId;
		meshId[i] = t.meshId;
		const float b[6] = {t.box.minis.x, t.box.minis.y, t.box.minis.z, t.box.maxis.x, t.box.maxis.y, t.box.maxis.z};
		std::memcpy(boxes + 6 * i, b, sizeof(b));
	}
}

// rays: n x {ox, oy, oz, tmin, dx, dy, dz, tmax}. mode: 0 closest, 1 any, 2 brute-force closest.
// Outputs are per-ray; w is 1-u-v as in Hit(). stats (nullable): [node expansions, tri tests].
void orc_rc_intersect(void * h, const float * rays, size_t n, int mode, int threads, float * t, float * u, float * v,
	float * w, uint32_t * hit, uint32_t * mesh, uint32_t * local, uint64_t * stats) {
	const Raycaster * rc = static_cast<Raycaster *>(h);
	std::atomic<uint64_t> nodes(0), tris(0);
	// Chunk the ray range so the thread policy spans contiguous blocks of rays like rows.
	const size_t chunk = 256;
	const size_t nchunks = (n + chunk - 1) / chunk;
	forParallel(0, nchunks, threads, [&](size_t c) {
		Stats st;
		const size_t a = c * chunk, b = std::min(n, a + chunk);
		for(size_t i = a; i < b; ++i) {
			const float * r = rays + 8 * i;
			const v3 o(r[0], r[1], r[2]), d(r[4], r[5], r[6]);
			Hit hh;
			if(mode == 2)
				hh = rc->brute(o, d, r[3], r[7]);
			else
				hh = rc->traverse(o, d, r[3], r[7], mode == 1, stats ? &st : nullptr);
			t[i] = hh.dist;
			u[i] = hh.u;
			v[i] = hh.v;
			w[i] = hh.w;
			hit[i] = hh.hit ? 1u : 0u;
			mesh[i] = uint32_t(hh.meshId);
			local[i] = uint32_t(hh.localId);
		}
		if(stats) {
			nodes += st.nodes;
			tris += st.tris;
		}
	});
	if(stats) {
		stats[0] = nodes.load();
		stats[1] = tris.load();
	}
}

// Single Moller-Trumbore test against explicit vertices (Raycaster.cpp:232-262), for KATs.
int orc_mt(const float * o, const float * d, const float * a, const float * b, const float * c, float mini, float maxi,
	float * out /* t,u,v */) {
	Raycaster rc;
	rc.vertices = {v3(a[0], a[1], a[2]), v3(b[0], b[1], b[2]), v3(c[0], c[1], c[2])};
	TriangleInfos tri;
	tri.v0 = 0;
	tri.v1 = 1;
	tri.v2 = 2;
	const Hit hh = rc.triTest(Ray(v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2])), tri, mini, maxi);
	out[0] = hh.dist;
	out[1] = hh.u;
	out[2] = hh.v;
	return hh.hit ? 1 : 0;
}

// Intersection::box (Intersection.cpp:22-32) on a raw box, plus the Ray construction (dir, invdir).
int orc_box(const float * o, const float * d, const float * bmin, const float * bmax, float mini, float maxi) {
	Box b;
	b.minis = v3(bmin[0], bmin[1], bmin[2]);
	b.maxis = v3(bmax[0], bmax[1], bmax[2]);
	return boxTest(Ray(v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2])), b, mini, maxi) ? 1 : 0;
}

void orc_ray_setup(const float * d, float * dir, float * invdir) {
	const Ray r(v3(0.f), v3(d[0], d[1], d[2]));
	dir[0] = r.dir.x;
	dir[1] = r.dir.y;
	dir[2] = r.dir.z;
	invdir[0] = r.invdir.x;
	invdir[1] = r.invdir.y;
	invdir[2] = r.invdir.z;
}

} // extern "C"
