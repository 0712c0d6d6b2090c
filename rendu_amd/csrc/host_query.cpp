// Per-ray queries answered on the host: Raycaster::intersects / intersectsAny for one ray
// (src/engine/raycaster/Raycaster.cpp:144-224), the latency-bound calls Rendu's PathTracer makes
// from its forParallel workers (PathTracer.cpp:186, 204). A single ray's traversal is a chain of
// dependent reads; a GPU round trip for one ray costs tens of microseconds, more than the whole
// traversal takes on a host core. These queries therefore walk the same flattened binary tree that
// was uploaded for the GPU (DeviceLayout::nodes / tris / roots, kept on the host after the upload),
// in Rendu's visiting order, with the same float operations as the GPU's reference-order kernel
// (makeRay, the select-based slab test, Moller-Trumbore on v0 / e1 / e2): every field of the
// result is bit-identical to that kernel's and to Rendu's. Batches and the path tracer always run
// on the GPU; these calls need the uploaded tree too, so without a device they fail like the rest.
//
// No lock and no allocation per call (the traversal stack lives on the caller's stack frame, or in
// a per-thread buffer for trees deeper than it), so any number of threads may query at once.
#include "raycaster_impl.hpp"

#include <cfloat>
#include <cstring>
#include <vector>

namespace rdu {

namespace {

inline int asInt(float f) {
	int i;
	std::memcpy(&i, &f, sizeof(i));
	return i;
}

inline uint32_t asU32(float f) {
	uint32_t i;
	std::memcpy(&i, &f, sizeof(i));
	return i;
}

struct HostBest {
	float maxi;
	int best = -1;
	float t = FLT_MAX, u = 0.0f, v = 0.0f;
};

// Leaf code c < 0: ~(first << 1 | (count - 1)). Returns true when an any-hit query is answered.
template<bool ANY>
inline bool leaf(const float * tris, const DevRay & r, int c, float mini, HostBest & b) {
	const int code = ~c;
	const int first = code >> 1;
	const int cnt = (code & 1) + 1;
	for(int k = 0; k < cnt; ++k) {
		const float * t = tris + 12 * size_t(first + k);
		float tt, uu, vv;
		if(mollerTrumbore(r, mk(t[0], t[1], t[2]), mk(t[3], t[4], t[5]), mk(t[6], t[7], t[8]), mini, b.maxi, tt, uu, vv)) {
			if(ANY)
				return true;
			// hit.dist < bestHit.dist holds whenever the test passed: t < maxi = best.
			b.best = first + k;
			b.t = tt;
			b.u = uu;
			b.v = vv;
			b.maxi = tt;
		}
	}
	return false;
}

// Rendu's order: every root whose box passes with the query's maxi is pushed, the last popped
// first (roots from M-1 down to 0); below a root, left is pushed before right and right is popped
// first; boxes below the roots are tested with the current maxi.
template<bool ANY>
bool walk(const DeviceLayout & L, uint32_t nRoots, const DevRay & r, HostBest & b, int * stack, size_t cap) {
	const float mini = r.mini, maxi0 = r.maxi;
	const float * nodes = L.nodes.data();
	const float * tris = L.tris.data();
	for(int ri = int(nRoots) - 1; ri >= 0; --ri) {
		const float * rb = L.roots.data() + 8 * size_t(ri);
		float e;
		if(!slab(r, mk(rb[0], rb[1], rb[2]), mk(rb[4], rb[5], rb[6]), mini, maxi0, e))
			continue;
		int c = asInt(rb[3]);
		size_t sp = 0;
		while(true) {
			if(c >= 0) {
				const float * n = nodes + 16 * size_t(c);
				float eL, eR;
				const bool hL = slab(r, mk(n[0], n[1], n[2]), mk(n[3], n[4], n[5]), mini, b.maxi, eL);
				const bool hR = slab(r, mk(n[6], n[7], n[8]), mk(n[9], n[10], n[11]), mini, b.maxi, eR);
				if(hL) {
					if(sp == cap)
						return false; // cannot happen: cap covers the tree's depth
					stack[sp++] = asInt(n[12]);
				}
				if(hR) {
					c = asInt(n[13]);
					continue;
				}
			} else if(leaf<ANY>(tris, r, c, mini, b)) {
				return true;
			}
			if(sp == 0)
				break;
			c = stack[--sp];
		}
	}
	return false;
}

template<bool ANY>
bool query(const rdu_raycaster * rc, const DevRay & r, HostBest & b) {
	b.maxi = r.maxi;
	// Depth below a root bounds the stack: at most one pending left child per level.
	const size_t need = size_t(rc->layout.maxDepth) + 2;
	constexpr size_t kFrame = 96;
	if(need <= kFrame) {
		int stack[kFrame];
		return walk<ANY>(rc->layout, rc->scene.meshCount, r, b, stack, kFrame);
	}
	thread_local std::vector<int> deep;
	if(deep.size() < need)
		deep.resize(need);
	return walk<ANY>(rc->layout, rc->scene.meshCount, r, b, deep.data(), deep.size());
}

} // namespace

void hostIntersect(const rdu_raycaster * rc, const rdu_ray & ray, rdu_hit * out) {
	const DevRay r = makeRay(mk(ray.ox, ray.oy, ray.oz), mk(ray.dx, ray.dy, ray.dz), ray.tmin, ray.tmax);
	HostBest b;
	query<false>(rc, r, b);
	rdu_hit h{FLT_MAX, 0.0f, 0.0f, 0u, 0u, 0u};
	if(b.best >= 0) {
		const float * t = rc->layout.tris.data() + 12 * size_t(b.best);
		h.t = b.t;
		h.u = b.u;
		h.v = b.v;
		h.hit = 1u;
		h.mesh_id = asU32(t[9]);
		h.local_id = asU32(t[10]);
	}
	*out = h;
}

bool hostIntersectAny(const rdu_raycaster * rc, const rdu_ray & ray) {
	const DevRay r = makeRay(mk(ray.ox, ray.oy, ray.oz), mk(ray.dx, ray.dy, ray.dz), ray.tmin, ray.tmax);
	HostBest b;
	return query<true>(rc, r, b);
}

} // namespace rdu
