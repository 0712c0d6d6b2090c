// Rendu's BVH build on the host and its flattening into the device layout (see bvh.hpp).
#include "bvh.hpp"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <limits>
#include <thread>

namespace rdu {

namespace {

HBox triBox(f3 a, f3 b, f3 c) {
	// Bounds.cpp:12-15
	return {gmin(gmin(a, b), c), gmax(gmax(a, b), c)};
}

void mergeInto(HBox & dst, const HBox & o) {
	// Bounds.cpp:17-20
	dst.mn = gmin(dst.mn, o.mn);
	dst.mx = gmax(dst.mx, o.mx);
}

int32_t leafCode(uint64_t first, uint64_t count) { return ~int32_t((first << 1) | (count - 1)); }

uint32_t asBits(float f) {
	uint32_t u;
	std::memcpy(&u, &f, 4);
	return u;
}
float fromBits(uint32_t u) {
	float f;
	std::memcpy(&f, &u, 4);
	return f;
}

} // namespace

void HostScene::addMesh(const float * xyz, size_t nv, const uint32_t * idx, size_t ni, const float * model) {
	const uint32_t indexOffset = uint32_t(vertices.size());
	const size_t start = vertices.size();
	// Geometric growth: reserving exactly per mesh would copy everything at every call.
	if(vertices.capacity() < start + nv)
		vertices.reserve(std::max(start + nv, 2 * vertices.capacity()));
	for(size_t i = 0; i < nv; ++i)
		vertices.push_back(mk(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
	bool identity = true;
	if(model)
		for(int c = 0; c < 4; ++c)
			for(int r = 0; r < 4; ++r)
				identity = identity && (model[4 * c + r] == (c == r ? 1.0f : 0.0f));
	if(model && !identity) {
		// vec3(model * vec4(v, 1)), glm mat4*vec4 order (type_mat4x4.inl:558-569):
		// (m[0]*x + m[1]*y) + (m[2]*z + m[3]*1) per component.
		for(size_t vid = start; vid < vertices.size(); ++vid) {
			const f3 p = vertices[vid];
			float o[3];
			for(int r = 0; r < 3; ++r) {
				const float a0 = model[r] * p.x + model[4 + r] * p.y;
				const float a1 = model[8 + r] * p.z + model[12 + r] * 1.0f;
				o[r] = a0 + a1;
			}
			vertices[vid] = mk(o[0], o[1], o[2]);
		}
	}
	const size_t startTri = tris.size();
	const size_t count = ni / 3;
	if(tris.capacity() < startTri + count)
		tris.reserve(std::max(startTri + count, 2 * tris.capacity()));
	for(size_t t = 0; t < count; ++t) {
		HTri tr;
		tr.v0 = indexOffset + idx[3 * t + 0];
		tr.v1 = indexOffset + idx[3 * t + 1];
		tr.v2 = indexOffset + idx[3 * t + 2];
		tr.localId = uint32_t(3 * t);
		tr.meshId = meshCount;
		tr.box = triBox(vertices[tr.v0], vertices[tr.v1], vertices[tr.v2]);
		tr.c = 0.5f * (tr.box.mn + tr.box.mx); // Bounds.cpp:61-63
		tris.push_back(tr);
	}
	HNode root;
	root.left = startTri;
	root.right = count;
	nodes.push_back(root);
	++meshCount;
}

void HostScene::reserve(size_t nVertices, size_t nTriangles, size_t nMeshes) {
	vertices.reserve(vertices.size() + nVertices);
	tris.reserve(tris.size() + nTriangles);
	nodes.reserve(nodes.size() + nMeshes);
}

namespace {

// One iteration of Raycaster::updateHierarchy's loop (Raycaster.cpp:54-142) on a node covering
// tris[begin, begin + count): its box; a leaf below 3 triangles; else the split of the range (mean
// of the centroids on the largest axis and std::partition, or the count / 2 nth_element fallback).
// Returns whether the node is internal, with `split` triangles going left.
bool splitNode(std::vector<HTri> & tris, HNode & node, uint64_t begin, uint64_t count, uint64_t & split) {
	HBox global = tris[begin].box;
	for(uint64_t t = 1; t < count; ++t)
		mergeInto(global, tris[begin + t].box);
	node.box = global;
	if(count < 3) {
		node.leaf = true;
		node.left = begin;
		node.right = count;
		return false;
	}
	node.leaf = false;
	split = 0;
	const f3 s = global.mx - global.mn;
	const int axis = (s.x >= s.y && s.x >= s.z) ? 0 : (s.y >= s.z ? 1 : 2);
	HTri * first = tris.data() + begin;
	HTri * last = first + count;
	if(count >= 5) {
		// Sequential float sum in triangle order, then one division (Raycaster.cpp:102-106).
		float mean = 0.0f;
		for(uint64_t t = 0; t < count; ++t)
			mean += comp(first[t].c, axis);
		mean /= float(count);
		HTri * mid = std::partition(first, last, [mean, axis](const HTri & a) { return comp(a.c, axis) < mean; });
		split = uint64_t(mid - first);
	}
	if(split == 0 || split == count || count < 5) {
		split = count / 2;
		std::nth_element(first, first + split, last,
			[axis](const HTri & a, const HTri & b) { return comp(a.c, axis) < comp(b.c, axis); });
	}
	return true;
}

// Subtrees at least this large are built on their own thread (while spawning depth is left).
constexpr uint64_t kParallelMin = 32768;

// Appends `sub` (a subtree list: its root at sub[0], descendants after, child links local to sub)
// to `out`: sub[0] goes to out[at], the descendants to the end of out, links remapped.
void stitch(std::vector<HNode> & out, uint64_t at, const std::vector<HNode> & sub) {
	const uint64_t base = out.size(); // where sub[1] lands
	auto remap = [base](HNode n) {
		if(!n.leaf) {
			n.left = base + (n.left - 1);
			n.right = base + (n.right - 1);
		}
		return n;
	};
	out[at] = remap(sub[0]);
	for(size_t i = 1; i < sub.size(); ++i)
		out.push_back(remap(sub[i]));
}

// Builds the subtree whose root record is out[0] (covering tris[begin, begin + count)), appending
// its descendants to `out` in exactly the order Rendu's single LIFO loop allocates them: a node's
// two children together, then everything below the right child (pushed last, so popped first),
// then everything below the left child. Large subtrees split into threads for `spawn` more
// levels; their lists are stitched in that same order, so the result does not depend on the
// number of threads.
void buildSubtree(std::vector<HTri> & tris, std::vector<HNode> & out, uint64_t begin, uint64_t count, int spawn) {
	struct Work {
		uint64_t id, begin, count;
	};
	if(spawn > 0 && count >= kParallelMin) {
		uint64_t split = 0;
		if(!splitNode(tris, out[0], begin, count, split))
			return;
		out.resize(3); // left child at 1, right child at 2
		out[0].left = 1;
		out[0].right = 2;
		std::vector<HNode> right(1), left(1);
		right.reserve(2 * (count - split));
		left.reserve(2 * split);
		std::thread worker([&] { buildSubtree(tris, right, begin + split, count - split, spawn - 1); });
		buildSubtree(tris, left, begin, split, spawn - 1);
		worker.join();
		stitch(out, 2, right);
		stitch(out, 1, left);
		return;
	}
	std::vector<Work> stack;
	stack.push_back({0, begin, count});
	while(!stack.empty()) {
		const Work cur = stack.back();
		stack.pop_back();
		uint64_t split = 0;
		if(!splitNode(tris, out[cur.id], cur.begin, cur.count, split))
			continue;
		out.emplace_back();
		const uint64_t leftPos = out.size() - 1;
		out[cur.id].left = leftPos;
		stack.push_back({leftPos, cur.begin, split});
		out.emplace_back();
		const uint64_t rightPos = out.size() - 1;
		out[cur.id].right = rightPos;
		stack.push_back({rightPos, cur.begin + split, cur.count - split});
	}
}

} // namespace

void HostScene::build() {
	// Raycaster.cpp:54-142: one LIFO loop seeded with the mesh roots 0..M-1, so the last mesh's
	// subtree is built (and its nodes allocated) first. The meshes' triangle ranges are disjoint:
	// each subtree is built on its own (on a pool of threads), then the lists are stitched in the
	// loop's order (root M-1 first), which reproduces its node array exactly.
	const uint64_t m = meshCount;
	std::vector<std::vector<HNode>> subs(m);
	// Threads only pay off for big scenes (starting one costs tens of microseconds).
	const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
	const unsigned workers = tris.size() < kParallelMin ? 1u : unsigned(std::min<uint64_t>(hw, m));
	const int spawn = 4; // up to 16 threads inside a mesh, only for subtrees of kParallelMin or more
	std::atomic<uint64_t> next{0};
	auto run = [&] {
		for(uint64_t i = next.fetch_add(1); i < m; i = next.fetch_add(1)) {
			const uint64_t r = m - 1 - i; // the loop's order: biggest index first
			subs[r].assign(1, nodes[r]);
			subs[r].reserve(2 * nodes[r].right);
			buildSubtree(tris, subs[r], nodes[r].left, nodes[r].right, spawn);
		}
	};
	std::vector<std::thread> pool;
	for(unsigned t = 1; t < workers; ++t)
		pool.emplace_back(run);
	run();
	for(std::thread & t : pool)
		t.join();
	nodes.reserve(nodes.size() + 2 * tris.size());
	for(uint64_t i = 0; i < m; ++i) {
		const uint64_t r = m - 1 - i;
		stitch(nodes, r, subs[r]);
		std::vector<HNode>().swap(subs[r]);
	}
	built = true;
}

namespace {
// body(begin, end) over [0, n): on up to 16 threads (contiguous chunks) for large n, else inline.
// Each index is written by exactly one chunk, so the result does not depend on the threads.
template<typename F>
void parallelFor(size_t n, const F & body) {
	const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
	const size_t workers = n < kParallelMin ? 1 : std::min<size_t>(hw, 16);
	if(workers <= 1) {
		body(size_t(0), n);
		return;
	}
	const size_t step = (n + workers - 1) / workers;
	std::vector<std::thread> pool;
	for(size_t w = 1; w < workers; ++w) {
		const size_t b = w * step, e = std::min(n, b + step);
		if(b < e)
			pool.emplace_back([&body, b, e] { body(b, e); });
	}
	body(size_t(0), std::min(n, step));
	for(std::thread & t : pool)
		t.join();
}
} // namespace

// Collapses the binary tree `ext` below node `top` into 4-wide nodes appended to `wide` (wideCount
// counting them): each 4-wide node takes its binary node's two children and keeps opening its
// internal child of largest surface area until it holds 4 children, numbered depth-first in
// preorder; a leaf child gets leafCodeOf(child). Returns the code of `top` (its leaf code when it
// is a leaf) and sets `depth` to the number of 4-wide levels.
template<typename LeafFn>
int32_t collapseWide(const std::vector<HNode> & ext, uint64_t top, const LeafFn & leafCodeOf, std::vector<float> & wide,
	uint32_t & wideCount, uint32_t & depth) {
	depth = 0;
	if(ext[top].leaf)
		return leafCodeOf(ext[top]);
	auto area = [&](uint64_t id) {
		const f3 s = ext[id].box.mx - ext[id].box.mn;
		return s.x * s.y + s.y * s.z + s.z * s.x;
	};
	struct WItem {
		uint64_t id; // internal node heading this 4-wide node
		uint32_t slot; // wide index assigned to it
		uint32_t depth;
	};
	std::vector<WItem> wst;
	const int32_t topCode = int32_t(wideCount++);
	wide.resize(size_t(wideCount) * 32, 0.0f);
	wst.push_back({top, uint32_t(topCode), 1});
	while(!wst.empty()) {
		const WItem it = wst.back();
		wst.pop_back();
		depth = std::max(depth, it.depth);
		uint64_t kids[4] = {ext[it.id].left, ext[it.id].right, 0, 0};
		size_t nk = 2;
		while(nk < 4) {
			int best = -1;
			float bestArea = -1.0f;
			for(size_t k = 0; k < nk; ++k)
				if(!ext[kids[k]].leaf && area(kids[k]) > bestArea) {
					best = int(k);
					bestArea = area(kids[k]);
				}
			if(best < 0)
				break;
			const uint64_t e = kids[size_t(best)];
			for(size_t k = nk; k > size_t(best) + 1; --k) // open a slot after `best`
				kids[k] = kids[k - 1];
			kids[size_t(best)] = ext[e].left;
			kids[size_t(best) + 1] = ext[e].right;
			++nk;
		}
		int32_t codes[4] = {kWideEmpty, kWideEmpty, kWideEmpty, kWideEmpty};
		float * o = wide.data() + size_t(it.slot) * 32;
		WItem children[4];
		size_t nc = 0;
		for(size_t k = 0; k < nk; ++k) {
			const HNode & c = ext[kids[k]];
			o[0 + k] = c.box.mn.x;
			o[4 + k] = c.box.mx.x;
			o[8 + k] = c.box.mn.y;
			o[12 + k] = c.box.mx.y;
			o[16 + k] = c.box.mn.z;
			o[20 + k] = c.box.mx.z;
			if(c.leaf) {
				codes[k] = leafCodeOf(c);
			} else {
				codes[k] = int32_t(wideCount++);
				children[nc++] = {kids[k], uint32_t(codes[k]), it.depth + 1};
			}
		}
		wide.resize(size_t(wideCount) * 32, 0.0f);
		o = wide.data() + size_t(it.slot) * 32;
		for(int k = 0; k < 4; ++k)
			o[24 + k] = fromBits(uint32_t(codes[k]));
		for(size_t k = nc; k > 0; --k)
			wst.push_back(children[k - 1]);
	}
	return topCode;
}

int32_t buildInstanceTop(const std::vector<HBox> & boxes, const std::vector<int32_t> & leafCodes, std::vector<float> & wide,
	uint32_t & wideCount, uint32_t & depth, HBox & topBox) {
	const uint64_t n = boxes.size();
	std::vector<HNode> ext(n);
	for(uint64_t i = 0; i < n; ++i) {
		ext[i].box = boxes[i];
		ext[i].leaf = true;
		ext[i].left = i;
		ext[i].right = 1;
	}
	uint64_t top = 0;
	if(n > 1) {
		struct TWork {
			uint64_t id, begin, count;
		};
		std::vector<uint64_t> ids(n);
		for(uint64_t i = 0; i < n; ++i)
			ids[i] = i;
		ext.reserve(2 * n);
		ext.emplace_back();
		top = ext.size() - 1;
		std::vector<TWork> tw{{top, 0, n}};
		while(!tw.empty()) {
			const TWork w = tw.back();
			tw.pop_back();
			HBox box = boxes[ids[w.begin]];
			for(uint64_t k = 1; k < w.count; ++k)
				mergeInto(box, boxes[ids[w.begin + k]]);
			ext[w.id].box = box;
			ext[w.id].leaf = false;
			const f3 sz = box.mx - box.mn;
			const int axis = (sz.x >= sz.y && sz.x >= sz.z) ? 0 : (sz.y >= sz.z ? 1 : 2);
			const uint64_t half = w.count / 2;
			auto cen = [&](uint64_t id) { return comp(boxes[id].mn, axis) + comp(boxes[id].mx, axis); };
			std::nth_element(ids.begin() + w.begin, ids.begin() + w.begin + half, ids.begin() + w.begin + w.count,
				[&](uint64_t x, uint64_t y) { return cen(x) < cen(y); });
			uint64_t kid[2];
			const uint64_t beg[2] = {w.begin, w.begin + half}, cnt[2] = {half, w.count - half};
			for(int k = 0; k < 2; ++k) {
				if(cnt[k] == 1) {
					kid[k] = ids[beg[k]];
				} else {
					ext.emplace_back();
					kid[k] = ext.size() - 1;
					tw.push_back({kid[k], beg[k], cnt[k]});
				}
			}
			ext[w.id].left = kid[0];
			ext[w.id].right = kid[1];
		}
	}
	topBox = ext[top].box;
	const auto objectLeaf = [&](const HNode & l) { return leafCodes[l.left]; };
	return collapseWide(ext, top, objectLeaf, wide, wideCount, depth);
}

void HostScene::flatten(DeviceLayout & out, bool perMeshRoots) const {
	// Depth-first preorder numbering of internal nodes (left subtree first) so that a parent and
	// its near subtree share cache lines as traversal descends.
	std::vector<int32_t> gpuIndex(nodes.size(), -1);
	std::vector<uint64_t> order;
	order.reserve(nodes.size() / 2 + 1);
	uint32_t maxDepth = 0;
	std::vector<std::pair<uint64_t, uint32_t>> st;
	for(uint64_t r = 0; r < meshCount; ++r) {
		st.push_back({r, 1});
		while(!st.empty()) {
			const auto cur = st.back();
			st.pop_back();
			maxDepth = std::max(maxDepth, cur.second);
			const HNode & n = nodes[cur.first];
			if(n.leaf)
				continue;
			gpuIndex[cur.first] = int32_t(order.size());
			order.push_back(cur.first);
			st.push_back({n.right, cur.second + 1});
			st.push_back({n.left, cur.second + 1});
		}
	}
	auto code = [&](uint64_t id) -> int32_t {
		const HNode & n = nodes[id];
		return n.leaf ? leafCode(n.left, n.right) : gpuIndex[id];
	};
	out.internalCount = uint32_t(order.size());
	out.maxDepth = maxDepth;
	out.nodes.assign(order.size() * 16, 0.0f);
	parallelFor(order.size(), [&](size_t begin, size_t end) {
		for(size_t i = begin; i < end; ++i) {
			const HNode & n = nodes[order[i]];
			const HBox & L = nodes[n.left].box;
			const HBox & R = nodes[n.right].box;
			float * o = out.nodes.data() + 16 * i;
			const float v[12] = {L.mn.x, L.mn.y, L.mn.z, L.mx.x, L.mx.y, L.mx.z, R.mn.x, R.mn.y, R.mn.z, R.mx.x, R.mx.y, R.mx.z};
			std::memcpy(o, v, sizeof(v));
			o[12] = fromBits(uint32_t(code(n.left)));
			o[13] = fromBits(uint32_t(code(n.right)));
		}
	});
	out.tris.assign(tris.size() * 12, 0.0f);
	parallelFor(tris.size(), [&](size_t begin, size_t end) {
		for(size_t i = begin; i < end; ++i) {
			const HTri & t = tris[i];
			const f3 v0 = vertices[t.v0];
			const f3 e1 = vertices[t.v1] - v0; // Raycaster.cpp:235
			const f3 e2 = vertices[t.v2] - v0; // Raycaster.cpp:236
			float * o = out.tris.data() + 12 * i;
			const float v[9] = {v0.x, v0.y, v0.z, e1.x, e1.y, e1.z, e2.x, e2.y, e2.z};
			std::memcpy(o, v, sizeof(v));
			o[9] = fromBits(t.meshId);
			o[10] = fromBits(t.localId);
		}
	});
	// Top level over the mesh roots (near-first queries only): Rendu tests every root box in turn
	// (Raycaster.cpp:149-155); here the roots are grouped under a binary tree of exact box merges,
	// split at the median root centroid on the largest axis. A group's box contains its members'
	// boxes, so it culls nothing that a member's own test would keep.
	// Near-first layout only: a 2-triangle leaf becomes a node over two 1-triangle leaves with the
	// triangles' own boxes (Bounds.cpp:12-15), so that the 4-wide node culls each triangle. Sized
	// once for those leaves and the top level, so the copy never reallocates.
	size_t pairs = 0;
	for(const HNode & n : nodes)
		pairs += (n.leaf && n.right == 2) ? 1 : 0;
	std::vector<HNode> ext;
	ext.reserve(nodes.size() + 2 * pairs + 2 * size_t(meshCount) + 1);
	ext.assign(nodes.begin(), nodes.end());
	const size_t baseCount = ext.size();
	for(size_t i = 0; i < baseCount; ++i) {
		if(!ext[i].leaf || ext[i].right != 2)
			continue;
		const uint64_t first = ext[i].left;
		HNode a, b;
		a.leaf = b.leaf = true;
		a.left = first;
		a.right = 1;
		a.box = tris[first].box;
		b.left = first + 1;
		b.right = 1;
		b.box = tris[first + 1].box;
		ext.push_back(a);
		ext.push_back(b);
		ext[i].leaf = false;
		ext[i].left = ext.size() - 2;
		ext[i].right = ext.size() - 1;
	}
	uint64_t top = 0;
	if(meshCount > 1 && !perMeshRoots) {
		struct TWork {
			uint64_t id, begin, count;
		};
		std::vector<uint64_t> ids(meshCount);
		for(uint32_t r = 0; r < meshCount; ++r)
			ids[r] = r;
		std::vector<TWork> tw;
		ext.emplace_back();
		top = ext.size() - 1;
		tw.push_back({top, 0, meshCount});
		while(!tw.empty()) {
			const TWork w = tw.back();
			tw.pop_back();
			HBox box = nodes[ids[w.begin]].box;
			for(uint64_t k = 1; k < w.count; ++k)
				mergeInto(box, nodes[ids[w.begin + k]].box);
			ext[w.id].box = box;
			ext[w.id].leaf = false;
			const f3 sz = box.mx - box.mn;
			const int axis = (sz.x >= sz.y && sz.x >= sz.z) ? 0 : (sz.y >= sz.z ? 1 : 2);
			const uint64_t half = w.count / 2;
			auto cen = [&](uint64_t id) { return comp(nodes[id].box.mn, axis) + comp(nodes[id].box.mx, axis); };
			std::nth_element(ids.begin() + w.begin, ids.begin() + w.begin + half, ids.begin() + w.begin + w.count,
				[&](uint64_t x, uint64_t y) { return cen(x) < cen(y); });
			uint64_t kid[2];
			const uint64_t beg[2] = {w.begin, w.begin + half}, cnt[2] = {half, w.count - half};
			for(int k = 0; k < 2; ++k) {
				if(cnt[k] == 1) {
					kid[k] = ids[beg[k]];
				} else {
					ext.emplace_back();
					kid[k] = ext.size() - 1;
					tw.push_back({kid[k], beg[k], cnt[k]});
				}
			}
			ext[w.id].left = kid[0];
			ext[w.id].right = kid[1];
		}
	}
	out.wideTopBox = ext[top].box;

	// 4-wide collapse, depth-first preorder numbering: from the single top node, or from each mesh
	// root on its own.
	out.wide.clear();
	out.wide.reserve((ext.size() / 2 + 1) * 32);
	out.wideCount = 0;
	out.wideDepth = 0;
	const auto triLeaf = [](const HNode & n) { return leafCode(n.left, n.right); };
	if(perMeshRoots) {
		out.meshWideCodes.assign(meshCount, 0);
		for(uint32_t r = 0; r < meshCount; ++r) {
			uint32_t d = 0;
			out.meshWideCodes[r] = collapseWide(ext, r, triLeaf, out.wide, out.wideCount, d);
			out.wideDepth = std::max(out.wideDepth, d);
		}
	} else {
		out.wideTopCode = collapseWide(ext, top, triLeaf, out.wide, out.wideCount, out.wideDepth);
	}

	out.roots.assign(size_t(meshCount) * 8, 0.0f);
	for(uint32_t r = 0; r < meshCount; ++r) {
		const HNode & n = nodes[r];
		float * o = out.roots.data() + 8 * r;
		o[0] = n.box.mn.x;
		o[1] = n.box.mn.y;
		o[2] = n.box.mn.z;
		o[3] = fromBits(uint32_t(code(r)));
		o[4] = n.box.mx.x;
		o[5] = n.box.mx.y;
		o[6] = n.box.mx.z;
		if(r == 0) {
			out.sceneMin = n.box.mn;
			out.sceneMax = n.box.mx;
		} else {
			out.sceneMin = gmin(out.sceneMin, n.box.mn);
			out.sceneMax = gmax(out.sceneMax, n.box.mx);
		}
	}
	(void)asBits;
}

} // namespace rdu
