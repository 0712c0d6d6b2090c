// Tangent frames of triangle meshes, as Rendu's Mesh::computeTangentsAndBitangents(force) builds
// them (Mesh.cpp:356-508) when its scenes are loaded for the path tracer
// (ResourcesManager.cpp:319-326, pathtracer/main.cpp:90).
//
// Two steps:
// 1. A tangent per triangle corner by MikkTSpace (Mikkelsen, "Simulation of wrinkled surfaces
//    revisited", 2008; the algorithm of the mikktspace library that Rendu calls through
//    genTangSpaceDefault, i.e. an angular threshold of 180 degrees, Mesh.cpp:386-394): corners
//    with identical position, normal and texture coordinate are welded; triangles with a
//    repeated position are degenerate; the others get the first-order texture-space derivatives
//    (Os, Ot) and an orientation from the sign of their texture area; triangles are linked across
//    shared welded edges of opposite winding; around every welded vertex, the triangles reachable
//    through those links with the same orientation form a group (depth-first from the first
//    triangle in index order, the next triangle across the corner's outgoing edge before the one
//    across its incoming edge; triangles without usable derivatives take the orientation of the
//    first group that reaches them); a group's tangent is the angle-weighted sum of its
//    triangles' Os projected into the tangent plane of the vertex normal. Degenerate triangles
//    copy the tangent of the first healthy corner with the same welded vertex.
// 2. Rendu's vertex split (Mesh.cpp:395-505): each vertex keeps the tangent of its first corner;
//    a later corner with the same tangent (and sign) as the corner just before it reuses that
//    corner's vertex, another one gets a new vertex appended after the existing ones; bitangents
//    are -sign * cross(n, t), both normalised, with a fixed tangent where t is parallel to n.
//
// Single-precision arithmetic in the operation order of the library (no contraction: the library
// is built with -ffp-contract=off), so the frames are bit-identical to the reference's on the
// same input.
#include "capi_common.hpp"
#include "../../include/rdu.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

struct V3 {
	float x, y, z;
};
inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 mul(float s, V3 v) { return V3{s * v.x, s * v.y, s * v.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline bool nonZero(float f) { return std::fabs(f) > FLT_MIN; }
inline bool anyNonZero(V3 v) { return nonZero(v.x) || nonZero(v.y) || nonZero(v.z); }
inline V3 unit(V3 v) { return mul(1.0f / std::sqrt(dot(v, v)), v); }
inline V3 unitIfNonZero(V3 v) { return anyNonZero(v) ? unit(v) : v; }
// v minus its component along the unit normal n, normalised unless it vanishes.
inline V3 inPlane(V3 v, V3 n) { return unitIfNonZero(v - mul(dot(n, v), n)); }

constexpr uint32_t kNone = 0xffffffffu;

// Per healthy triangle: texture-space derivatives and flags.
struct Face {
	V3 os{0.0f, 0.0f, 0.0f}, ot{0.0f, 0.0f, 0.0f};
	bool orient = false; // texture area positive (orientation preserving)
	bool any = true;     // no usable derivatives: joins any group that reaches it
	uint32_t nbr[3] = {kNone, kNone, kNone}; // across edge k: corner k -> corner k+1
	uint32_t group[3] = {kNone, kNone, kNone};
};

struct Group {
	uint32_t rep; // welded vertex the group turns around
	bool orient;
	std::vector<uint32_t> faces; // in the order they joined
};

struct Frame {
	V3 t{1.0f, 0.0f, 0.0f};
	bool orient = false;
};

class Mikk {
public:
	Mikk(const float * pos, const float * nrm, const float * uv, const uint32_t * idx, size_t nTris)
		: pos_(pos), nrm_(nrm), uv_(uv), idx_(idx), nTris_(nTris) {}

	// Corner frames (3 per triangle, input order).
	std::vector<Frame> run() {
		std::vector<Frame> out(nTris_ * 3);
		weld();
		splitDegenerate();
		deriveFaces();
		linkFaces();
		buildGroups();
		frames(out);
		degenerateFrames(out);
		return out;
	}

private:
	V3 position(uint32_t corner) const {
		const float * p = pos_ + size_t(idx_[corner]) * 3;
		return V3{p[0], p[1], p[2]};
	}
	V3 normal(uint32_t corner) const {
		const float * p = nrm_ + size_t(idx_[corner]) * 3;
		return V3{p[0], p[1], p[2]};
	}
	V3 texcoord(uint32_t corner) const {
		if(!uv_)
			return V3{0.5f, 0.5f, 1.0f};
		const float * p = uv_ + size_t(idx_[corner]) * 2;
		return V3{p[0], p[1], 1.0f};
	}

	// Welded id of every corner: the smallest corner with the same position, normal and texture
	// coordinate under float equality (-0 equals +0; a NaN equals nothing).
	void weld() {
		const size_t n = nTris_ * 3;
		weld_.resize(n);
		struct Key {
			uint32_t k[8];
			uint32_t corner;
		};
		std::vector<Key> keys;
		keys.reserve(n);
		for(uint32_t c = 0; c < n; ++c) {
			const V3 p = position(c), q = normal(c), t = texcoord(c);
			const float v[8] = {p.x, p.y, p.z, q.x, q.y, q.z, t.x, t.y};
			Key key;
			bool nan = false;
			for(int i = 0; i < 8; ++i) {
				nan |= v[i] != v[i];
				const float z = v[i] == 0.0f ? 0.0f : v[i];
				std::memcpy(&key.k[i], &z, 4);
			}
			key.corner = c;
			if(nan)
				weld_[c] = c;
			else
				keys.push_back(key);
		}
		std::sort(keys.begin(), keys.end(), [](const Key & a, const Key & b) {
			const int d = std::memcmp(a.k, b.k, sizeof(a.k));
			return d != 0 ? d < 0 : a.corner < b.corner;
		});
		for(size_t i = 0; i < keys.size();) {
			size_t j = i;
			while(j < keys.size() && std::memcmp(keys[j].k, keys[i].k, sizeof(keys[i].k)) == 0)
				++j;
			for(size_t m = i; m < j; ++m)
				weld_[keys[m].corner] = keys[i].corner;
			i = j;
		}
	}

	// Healthy triangles (no two corners at one position) in input order, then the degenerate ones.
	void splitDegenerate() {
		for(uint32_t t = 0; t < nTris_; ++t) {
			const V3 a = position(weld_[3 * t]), b = position(weld_[3 * t + 1]), c = position(weld_[3 * t + 2]);
			auto same = [](V3 u, V3 v) { return u.x == v.x && u.y == v.y && u.z == v.z; };
			(same(a, b) || same(a, c) || same(b, c) ? degenerate_ : healthy_).push_back(t);
		}
	}

	uint32_t w(uint32_t f, int k) const { return weld_[3 * healthy_[f] + k]; }

	void deriveFaces() {
		faces_.assign(healthy_.size(), Face());
		for(size_t f = 0; f < healthy_.size(); ++f) {
			const uint32_t c0 = w(uint32_t(f), 0), c1 = w(uint32_t(f), 1), c2 = w(uint32_t(f), 2);
			const V3 p1 = position(c0), p2 = position(c1), p3 = position(c2);
			const V3 t1 = texcoord(c0), t2 = texcoord(c1), t3 = texcoord(c2);
			const float s21 = t2.x - t1.x, u21 = t2.y - t1.y, s31 = t3.x - t1.x, u31 = t3.y - t1.y;
			const V3 d1 = p2 - p1, d2 = p3 - p1;
			const float area2 = s21 * u31 - u21 * s31;
			const V3 os = mul(u31, d1) - mul(u21, d2);
			const V3 ot = mul(-s31, d1) + mul(s21, d2);
			Face & F = faces_[f];
			F.orient = area2 > 0.0f;
			if(!nonZero(area2))
				continue;
			const float absArea = std::fabs(area2);
			const float lenS = std::sqrt(dot(os, os)), lenT = std::sqrt(dot(ot, ot));
			const float sign = F.orient ? 1.0f : -1.0f;
			if(nonZero(lenS))
				F.os = mul(sign / lenS, os);
			if(nonZero(lenT))
				F.ot = mul(sign / lenT, ot);
			if(nonZero(lenS / absArea) && nonZero(lenT / absArea))
				F.any = false;
		}
	}

	// Links triangles across welded edges of opposite winding. The edges of one welded vertex pair
	// are visited in triangle order; each takes the first later unlinked edge of the pair wound the
	// other way.
	void linkFaces() {
		struct Edge {
			uint32_t lo, hi, face;
		};
		std::vector<Edge> edges;
		edges.reserve(healthy_.size() * 3);
		for(uint32_t f = 0; f < healthy_.size(); ++f)
			for(int k = 0; k < 3; ++k) {
				const uint32_t a = w(f, k), b = w(f, (k + 1) % 3);
				edges.push_back(Edge{std::min(a, b), std::max(a, b), f});
			}
		std::sort(edges.begin(), edges.end(), [](const Edge & x, const Edge & y) {
			return x.lo != y.lo ? x.lo < y.lo : (x.hi != y.hi ? x.hi < y.hi : x.face < y.face);
		});
		// Edge number of (lo, hi) in face f, and its endpoints in winding order.
		auto locate = [this](uint32_t f, uint32_t lo, uint32_t hi, uint32_t & from, uint32_t & to) {
			const uint32_t a = w(f, 0), b = w(f, 1), c = w(f, 2);
			if(a == lo || a == hi) {
				if(b == lo || b == hi) {
					from = a, to = b;
					return 0;
				}
				from = c, to = a;
				return 2;
			}
			from = b, to = c;
			return 1;
		};
		for(size_t i = 0; i < edges.size(); ++i) {
			const Edge & e = edges[i];
			uint32_t fromA, toA;
			const int ka = locate(e.face, e.lo, e.hi, fromA, toA);
			if(faces_[e.face].nbr[ka] != kNone)
				continue;
			for(size_t j = i + 1; j < edges.size() && edges[j].lo == e.lo && edges[j].hi == e.hi; ++j) {
				uint32_t fromB, toB;
				const int kb = locate(edges[j].face, e.lo, e.hi, fromB, toB);
				if(fromA == toB && toA == fromB && faces_[edges[j].face].nbr[kb] == kNone) {
					faces_[e.face].nbr[ka] = edges[j].face;
					faces_[edges[j].face].nbr[kb] = e.face;
					break;
				}
			}
		}
	}

	// Offers face f to group g (which turns around a welded vertex of f); on joining, f's
	// neighbours across the two edges at that vertex are offered next (outgoing edge first).
	void offer(uint32_t f, uint32_t g, std::vector<uint32_t> & stack) {
		stack.push_back(f);
		while(!stack.empty()) {
			const uint32_t t = stack.back();
			stack.pop_back();
			Face & F = faces_[t];
			const uint32_t rep = groups_[g].rep;
			const int k = w(t, 0) == rep ? 0 : (w(t, 1) == rep ? 1 : 2);
			if(F.group[k] != kNone)
				continue;
			if(F.any && F.group[0] == kNone && F.group[1] == kNone && F.group[2] == kNone)
				F.orient = groups_[g].orient;
			if(F.orient != groups_[g].orient)
				continue;
			groups_[g].faces.push_back(t);
			F.group[k] = g;
			if(F.nbr[(k + 2) % 3] != kNone)
				stack.push_back(F.nbr[(k + 2) % 3]);
			if(F.nbr[k] != kNone)
				stack.push_back(F.nbr[k]);
		}
	}

	void buildGroups() {
		std::vector<uint32_t> stack;
		for(uint32_t f = 0; f < faces_.size(); ++f)
			for(int k = 0; k < 3; ++k) {
				Face & F = faces_[f];
				if(F.any || F.group[k] != kNone)
					continue;
				const uint32_t g = uint32_t(groups_.size());
				groups_.push_back(Group{w(f, k), F.orient, {f}});
				F.group[k] = g;
				const uint32_t left = F.nbr[k], right = F.nbr[(k + 2) % 3];
				if(left != kNone)
					offer(left, g, stack);
				if(right != kNone)
					offer(right, g, stack);
			}
	}

	// Angle-weighted tangent of the faces `members` (ascending) around welded vertex rep.
	V3 groupTangent(const std::vector<uint32_t> & members, uint32_t rep) const {
		V3 sum{0.0f, 0.0f, 0.0f};
		for(uint32_t f : members) {
			const Face & F = faces_[f];
			if(F.any)
				continue;
			const int k = w(f, 0) == rep ? 0 : (w(f, 1) == rep ? 1 : 2);
			const V3 n = normal(rep);
			const V3 os = inPlane(F.os, n);
			const V3 prev = position(w(f, (k + 2) % 3)), at = position(w(f, k)), next = position(w(f, (k + 1) % 3));
			const V3 e1 = inPlane(prev - at, n), e2 = inPlane(next - at, n);
			float c = dot(e1, e2);
			c = c > 1.0f ? 1.0f : (c < -1.0f ? -1.0f : c);
			const float angle = float(std::acos(double(c)));
			sum = sum + mul(angle, os);
		}
		return unitIfNonZero(sum);
	}

	void frames(std::vector<Frame> & out) {
		const float thresholdAngle = (180.0f * float(M_PI)) / 180.0f;
		const float cosThreshold = float(std::cos(double(thresholdAngle)));
		std::vector<uint32_t> members;
		std::vector<std::vector<uint32_t>> subsets;
		std::vector<V3> subsetTangent;
		for(uint32_t g = 0; g < groups_.size(); ++g) {
			const Group & G = groups_[g];
			subsets.clear();
			subsetTangent.clear();
			const V3 n = normal(G.rep);
			for(uint32_t f : G.faces) {
				const Face & F = faces_[f];
				const int k = F.group[0] == g ? 0 : (F.group[1] == g ? 1 : 2);
				const V3 os = inPlane(F.os, n), ot = inPlane(F.ot, n);
				members.clear();
				for(uint32_t t : G.faces) {
					const Face & T = faces_[t];
					const V3 os2 = inPlane(T.os, n), ot2 = inPlane(T.ot, n);
					if(F.any || T.any || f == t || (dot(os, os2) > cosThreshold && dot(ot, ot2) > cosThreshold))
						members.push_back(t);
				}
				std::sort(members.begin(), members.end());
				size_t s = 0;
				while(s < subsets.size() && subsets[s] != members)
					++s;
				if(s == subsets.size()) {
					subsets.push_back(members);
					subsetTangent.push_back(groupTangent(members, G.rep));
				}
				Frame & fr = out[3 * size_t(healthy_[f]) + k];
				fr.t = subsetTangent[s];
				fr.orient = G.orient;
			}
		}
	}

	void degenerateFrames(std::vector<Frame> & out) const {
		if(degenerate_.empty())
			return;
		// First healthy corner (triangle order, then corner order) of every welded vertex.
		std::vector<std::pair<uint32_t, uint32_t>> first;
		first.reserve(healthy_.size() * 3);
		for(uint32_t f = 0; f < healthy_.size(); ++f)
			for(int k = 0; k < 3; ++k)
				first.emplace_back(w(f, k), 3 * healthy_[f] + k);
		std::stable_sort(first.begin(), first.end(),
			[](const std::pair<uint32_t, uint32_t> & a, const std::pair<uint32_t, uint32_t> & b) { return a.first < b.first; });
		for(uint32_t t : degenerate_)
			for(int k = 0; k < 3; ++k) {
				const uint32_t wc = weld_[3 * t + k];
				auto it = std::lower_bound(first.begin(), first.end(), std::make_pair(wc, 0u),
					[](const std::pair<uint32_t, uint32_t> & a, const std::pair<uint32_t, uint32_t> & b) { return a.first < b.first; });
				if(it != first.end() && it->first == wc)
					out[3 * size_t(t) + k] = out[it->second];
			}
	}

	const float * pos_, * nrm_, * uv_;
	const uint32_t * idx_;
	size_t nTris_;
	std::vector<uint32_t> weld_, healthy_, degenerate_;
	std::vector<Face> faces_;
	std::vector<Group> groups_;
};

// glm's single-precision forms (Mesh.cpp uses glm::cross, glm::length, glm::normalize).
inline V3 cross(V3 a, V3 b) { return V3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
inline float glmDot(V3 a, V3 b) {
	const V3 m{a.x * b.x, a.y * b.y, a.z * b.z};
	return (m.x + m.y) + m.z;
}
inline V3 normalise(V3 v) {
	const float inv = 1.0f / std::sqrt(glmDot(v, v));
	return V3{v.x * inv, v.y * inv, v.z * inv};
}

// Mesh.cpp:456-470: tangent of a corner frame at vertex normal n, and the bitangent.
void storeFrame(V3 t, float sign, V3 n, float * tangent, float * bitangent) {
	const V3 c = cross(n, t);
	if(std::sqrt(glmDot(c, c)) < 0.01f)
		t = std::fabs(n.z) > 0.01f ? V3{1.0f, 0.0f, 0.0f} : V3{0.0f, 0.0f, 1.0f};
	const V3 b0 = cross(n, t);
	const float s = -sign;
	const V3 b = normalise(V3{s * b0.x, s * b0.y, s * b0.z});
	const V3 tn = normalise(t);
	tangent[0] = tn.x, tangent[1] = tn.y, tangent[2] = tn.z;
	bitangent[0] = b.x, bitangent[1] = b.y, bitangent[2] = b.z;
}

int checkMesh(const float * positions, const float * normals, size_t nv, const uint32_t * indices, size_t ni) {
	if(!positions || !normals || (ni && !indices))
		return rdu::fail(RDU_ERR_INVALID, "null mesh array");
	if(ni % 3 != 0)
		return rdu::fail(RDU_ERR_INVALID, "index count is not a multiple of 3");
	if(ni > (size_t(1) << 31))
		return rdu::fail(RDU_ERR_UNSUPPORTED, "more than 2^31 indices");
	for(size_t i = 0; i < ni; ++i)
		if(indices[i] >= nv)
			return rdu::fail(RDU_ERR_INVALID, "index out of range");
	return RDU_OK;
}

} // namespace

extern "C" {

int rdu_mesh_corner_tangents(const float * positions, const float * normals, const float * texcoords, size_t n_vertices,
	const uint32_t * indices, size_t n_indices, float * corner_tangents) {
	rdu::clearError();
	int st = checkMesh(positions, normals, n_vertices, indices, n_indices);
	if(st)
		return st;
	if(n_indices && !corner_tangents)
		return rdu::fail(RDU_ERR_INVALID, "null output");
	try {
		const std::vector<Frame> fr = Mikk(positions, normals, texcoords, indices, n_indices / 3).run();
		for(size_t c = 0; c < fr.size(); ++c) {
			corner_tangents[4 * c] = fr[c].t.x;
			corner_tangents[4 * c + 1] = fr[c].t.y;
			corner_tangents[4 * c + 2] = fr[c].t.z;
			corner_tangents[4 * c + 3] = fr[c].orient ? 1.0f : -1.0f;
		}
	} catch(const std::bad_alloc &) {
		return rdu::fail(RDU_ERR_ALLOC, "host allocation failed");
	}
	return RDU_OK;
}

int rdu_mesh_tangent_frames(const float * positions, const float * normals, const float * texcoords, size_t n_vertices,
	uint32_t * indices, size_t n_indices, uint32_t * source_vertex, float * tangents, float * bitangents, size_t capacity,
	size_t * out_vertices) {
	rdu::clearError();
	int st = checkMesh(positions, normals, n_vertices, indices, n_indices);
	if(st)
		return st;
	if(!source_vertex || !tangents || !bitangents || !out_vertices)
		return rdu::fail(RDU_ERR_INVALID, "null output");
	if(capacity < n_vertices + n_indices)
		return rdu::fail(RDU_ERR_INVALID, "capacity below n_vertices + n_indices");
	for(size_t v = 0; v < n_vertices; ++v)
		source_vertex[v] = uint32_t(v);
	if(n_indices == 0) {
		// No triangle: the library reports a failure and Mesh.cpp:388-392 fills fixed axes.
		for(size_t v = 0; v < n_vertices; ++v) {
			tangents[3 * v] = 1.0f, tangents[3 * v + 1] = 0.0f, tangents[3 * v + 2] = 0.0f;
			bitangents[3 * v] = 0.0f, bitangents[3 * v + 1] = 1.0f, bitangents[3 * v + 2] = 0.0f;
		}
		*out_vertices = n_vertices;
		return RDU_OK;
	}
	try {
		const std::vector<Frame> fr = Mikk(positions, normals, texcoords, indices, n_indices / 3).run();
		auto sameFrame = [&fr](uint32_t a, uint32_t b) {
			return fr[a].t.x == fr[b].t.x && fr[a].t.y == fr[b].t.y && fr[a].t.z == fr[b].t.z && fr[a].orient == fr[b].orient;
		};
		// Corners of every vertex in index order (counting sort).
		std::vector<uint32_t> start(n_vertices + 1, 0), corners(n_indices);
		for(size_t i = 0; i < n_indices; ++i)
			++start[indices[i] + 1];
		for(size_t v = 0; v < n_vertices; ++v)
			start[v + 1] += start[v];
		{
			std::vector<uint32_t> fill(start.begin(), start.end() - 1);
			for(size_t i = 0; i < n_indices; ++i)
				corners[fill[indices[i]]++] = uint32_t(i);
		}
		auto frameOf = [&](uint32_t corner, size_t vertex) {
			const float * n = normals + 3 * size_t(source_vertex[vertex]);
			storeFrame(fr[corner].t, fr[corner].orient ? 1.0f : -1.0f, V3{n[0], n[1], n[2]}, tangents + 3 * vertex,
				bitangents + 3 * vertex);
		};
		size_t added = 0;
		for(size_t v = 0; v < n_vertices; ++v) {
			const uint32_t b = start[v], e = start[v + 1];
			if(b == e) {
				// A vertex no triangle uses has no frame in Rendu (Mesh.cpp:417 reads past an empty
				// list); it gets the frame of the library's default corner, tangent +x, sign -1.
				const float * n = normals + 3 * v;
				storeFrame(V3{1.0f, 0.0f, 0.0f}, -1.0f, V3{n[0], n[1], n[2]}, tangents + 3 * v, bitangents + 3 * v);
				continue;
			}
			frameOf(corners[b], v);
			for(uint32_t r = b + 1; r < e; ++r) {
				if(sameFrame(corners[r], corners[r - 1])) {
					indices[corners[r]] = indices[corners[r - 1]];
					continue;
				}
				const size_t nv = n_vertices + added++;
				source_vertex[nv] = uint32_t(v);
				indices[corners[r]] = uint32_t(nv);
				frameOf(corners[r], nv);
			}
		}
		*out_vertices = n_vertices + added;
	} catch(const std::bad_alloc &) {
		return rdu::fail(RDU_ERR_ALLOC, "host allocation failed");
	}
	return RDU_OK;
}

} // extern "C"
