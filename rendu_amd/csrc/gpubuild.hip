// GPU BVH build for gfx950: an alternative to Rendu's host build (Raycaster::updateHierarchy,
// src/engine/raycaster/Raycaster.cpp:54-142) that produces another tree over the same triangles.
//
// Linear BVH (Karras, "Maximizing Parallelism in the Construction of BVHs, Octrees, and k-d Trees",
// HPG 2012): 30-bit Morton codes of the triangle centroids (centroid = 0.5 (min + max) of the
// triangle box, Bounds.cpp:61-63) in the centroid bounds, a radix sort of (code, triangle), one
// binary radix tree over all triangles of all meshes (duplicate codes split by index), one
// triangle per leaf, boxes bottom-up (the second child to finish computes its parent), then the
// same 4-wide collapse as the host layout (largest surface area expanded first, up to 4 children),
// level by level with an atomic slot counter. The output is the near-first layout of bvh.hpp
// (triangle records and 4-wide nodes); Rendu's binary tree and visiting order are not produced.
//
// What it keeps: every box is an exact float merge of triangle boxes (Bounds.cpp:12-20), so a
// closest-hit query returns the same t, u, v and triangle as the host-built tree except where
// several triangles tie at the same t; any-hit results are identical.
#include "gpubuild.hpp"
#include "../../include/rdu.h"

#include <rocprim/device/device_radix_sort.hpp>

#include <vector>

namespace rdu {

namespace {

constexpr int kGbBlock = 256;

inline unsigned gbBlocks(size_t n) { return unsigned((n + kGbBlock - 1) / kGbBlock); }

// Order-preserving map of a float onto a signed int (for atomic min / max of floats).
__device__ __forceinline__ int orderedInt(float f) {
	const int i = __float_as_int(f);
	return i >= 0 ? i : i ^ 0x7fffffff;
}
__host__ __device__ inline float fromOrderedInt(int i) {
	const int b = i >= 0 ? i : i ^ 0x7fffffff;
	float f;
	__builtin_memcpy(&f, &b, 4);
	return f;
}

// Bounds: [0..2] scene min, [3..5] scene max, [6..8] centroid min, [9..11] centroid max (ordered ints).
__global__ void gbInitBounds(int * b) {
	const int t = threadIdx.x;
	if(t < 12)
		b[t] = (t % 6) < 3 ? 0x7fffffff : int(0x80000000);
}

// Per triangle: the record of the device layout ({v0, e1, e2, meshId, localId}, e1 = v1 - v0 and
// e2 = v2 - v0 as Moller-Trumbore computes them), its box and centroid; scene and centroid bounds.
__global__ __launch_bounds__(kGbBlock) void gbPrep(const float * __restrict__ vtx, const uint32_t * __restrict__ tri,
	uint32_t n, float4 * __restrict__ rec, float4 * __restrict__ box, float4 * __restrict__ cen, int * bounds) {
	const uint32_t i = blockIdx.x * kGbBlock + threadIdx.x;
	f3 mn = {__int_as_float(0x7f800000), __int_as_float(0x7f800000), __int_as_float(0x7f800000)};
	f3 mx = -mn, cmn = mn, cmx = mx;
	if(i < n) {
		const uint32_t * t = tri + 5 * size_t(i);
		const f3 a = {vtx[3 * size_t(t[0])], vtx[3 * size_t(t[0]) + 1], vtx[3 * size_t(t[0]) + 2]};
		const f3 b = {vtx[3 * size_t(t[1])], vtx[3 * size_t(t[1]) + 1], vtx[3 * size_t(t[1]) + 2]};
		const f3 c = {vtx[3 * size_t(t[2])], vtx[3 * size_t(t[2]) + 1], vtx[3 * size_t(t[2]) + 2]};
		const f3 e1 = b - a, e2 = c - a;
		rec[3 * size_t(i)] = make_float4(a.x, a.y, a.z, e1.x);
		rec[3 * size_t(i) + 1] = make_float4(e1.y, e1.z, e2.x, e2.y);
		rec[3 * size_t(i) + 2] = make_float4(e2.z, __uint_as_float(t[3]), __uint_as_float(t[4]), 0.0f);
		mn = gmin(gmin(a, b), c); // Bounds.cpp:12-15
		mx = gmax(gmax(a, b), c);
		const f3 ce = 0.5f * (mn + mx); // Bounds.cpp:61-63
		box[2 * size_t(i)] = make_float4(mn.x, mn.y, mn.z, 0.0f);
		box[2 * size_t(i) + 1] = make_float4(mx.x, mx.y, mx.z, 0.0f);
		cen[i] = make_float4(ce.x, ce.y, ce.z, 0.0f);
		cmn = ce;
		cmx = ce;
	}
	// Wavefront reductions, then one atomic per wavefront and bound.
	float v[12] = {mn.x, mn.y, mn.z, mx.x, mx.y, mx.z, cmn.x, cmn.y, cmn.z, cmx.x, cmx.y, cmx.z};
	for(int k = 0; k < 12; ++k) {
		const bool isMin = (k % 6) < 3;
		for(int off = 32; off > 0; off >>= 1) {
			const float o = __shfl_xor(v[k], off, 64);
			v[k] = isMin ? fminf(v[k], o) : fmaxf(v[k], o);
		}
	}
	if(__lane_id() == 0)
		for(int k = 0; k < 12; ++k) {
			if(!(v[k] == v[k]) || v[k] == __int_as_float(0x7f800000) || v[k] == -__int_as_float(0x7f800000))
				continue;
			if((k % 6) < 3)
				atomicMin(bounds + k, orderedInt(v[k]));
			else
				atomicMax(bounds + k, orderedInt(v[k]));
		}
}

__device__ __forceinline__ uint32_t spread10(uint32_t x) {
	x &= 0x3ffu;
	x = (x | (x << 16)) & 0x030000ffu;
	x = (x | (x << 8)) & 0x0300f00fu;
	x = (x | (x << 4)) & 0x030c30c3u;
	x = (x | (x << 2)) & 0x09249249u;
	return x;
}

__global__ __launch_bounds__(kGbBlock) void gbMorton(const float4 * __restrict__ cen, uint32_t n, const int * bounds,
	uint32_t * __restrict__ keys, uint32_t * __restrict__ vals) {
	const uint32_t i = blockIdx.x * kGbBlock + threadIdx.x;
	if(i >= n)
		return;
	const float4 c = cen[i];
	const float cv[3] = {c.x, c.y, c.z};
	uint32_t q[3];
	for(int a = 0; a < 3; ++a) {
		const float lo = fromOrderedInt(bounds[6 + a]), hi = fromOrderedInt(bounds[9 + a]);
		const float ext = hi - lo;
		float t = ext > 0.0f ? (cv[a] - lo) / ext : 0.0f;
		t = t > 0.0f ? (t < 1.0f ? t : 1.0f) : 0.0f;
		const float s = t * 1023.0f;
		q[a] = uint32_t(s);
	}
	keys[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
	vals[i] = i;
}

// Records and leaf boxes in sorted order.
__global__ __launch_bounds__(kGbBlock) void gbGather(const uint32_t * __restrict__ order, uint32_t n,
	const float4 * __restrict__ recIn, const float4 * __restrict__ boxIn, float4 * __restrict__ recOut,
	float4 * __restrict__ leafBox) {
	const uint32_t j = blockIdx.x * kGbBlock + threadIdx.x;
	if(j >= n)
		return;
	const uint32_t s = order[j];
	for(int k = 0; k < 3; ++k)
		recOut[3 * size_t(j) + k] = recIn[3 * size_t(s) + k];
	leafBox[2 * size_t(j)] = boxIn[2 * size_t(s)];
	leafBox[2 * size_t(j) + 1] = boxIn[2 * size_t(s) + 1];
}

// Common-prefix length of keys i and j (index bits break ties), -1 outside [0, n).
__device__ __forceinline__ int delta(const uint32_t * keys, uint32_t n, int i, int j) {
	if(j < 0 || j >= int(n))
		return -1;
	const uint32_t a = keys[i], b = keys[j];
	if(a == b)
		return 32 + __clz(uint32_t(i) ^ uint32_t(j));
	return __clz(a ^ b);
}

// Internal node i of the radix tree: its key range, split and children. Child codes: >= 0
// internal node, < 0 leaf -(index + 1).
__global__ __launch_bounds__(kGbBlock) void gbKarras(const uint32_t * __restrict__ keys, uint32_t n,
	int2 * __restrict__ child, uint32_t * __restrict__ parentInner, uint32_t * __restrict__ parentLeaf) {
	const int i = int(blockIdx.x * kGbBlock + threadIdx.x);
	if(i >= int(n) - 1)
		return;
	const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
	const int dmin = delta(keys, n, i, i - d);
	int lmax = 2;
	while(delta(keys, n, i, i + lmax * d) > dmin)
		lmax <<= 1;
	int l = 0;
	for(int t = lmax >> 1; t >= 1; t >>= 1)
		if(delta(keys, n, i, i + (l + t) * d) > dmin)
			l += t;
	const int j = i + l * d;
	const int dnode = delta(keys, n, i, j);
	int s = 0, t = l;
	do {
		t = (t + 1) >> 1;
		if(delta(keys, n, i, i + (s + t) * d) > dnode)
			s += t;
	} while(t > 1);
	const int split = i + s * d + (d < 0 ? d : 0);
	const int lo = i < j ? i : j, hi = i < j ? j : i;
	int2 c;
	if(lo == split) {
		c.x = -(split + 1);
		parentLeaf[split] = uint32_t(i);
	} else {
		c.x = split;
		parentInner[split] = uint32_t(i);
	}
	if(hi == split + 1) {
		c.y = -(split + 2);
		parentLeaf[split + 1] = uint32_t(i);
	} else {
		c.y = split + 1;
		parentInner[split + 1] = uint32_t(i);
	}
	child[i] = c;
}

// Agent-scope loads: boxes written by other workgroups, read without a stale L1 copy.
__device__ __forceinline__ float4 loadCoherent(const float4 * p) {
	const float * f = reinterpret_cast<const float *>(p);
	return make_float4(__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
		__hip_atomic_load(f + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
		__hip_atomic_load(f + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
		__hip_atomic_load(f + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

__device__ __forceinline__ void childBox(int code, const float4 * leafBox, const float4 * innerBox, float4 & mn, float4 & mx) {
	if(code < 0) {
		mn = leafBox[2 * size_t(-code - 1)];
		mx = leafBox[2 * size_t(-code - 1) + 1];
	} else {
		mn = loadCoherent(innerBox + 2 * size_t(code));
		mx = loadCoherent(innerBox + 2 * size_t(code) + 1);
	}
}

// Bottom-up boxes: every leaf walks towards the root; at each node the first arrival stops and the
// second computes the node's box from both children and continues. Exactly one thread finishes
// each internal node; no thread waits on another.
__global__ __launch_bounds__(kGbBlock) void gbBounds(uint32_t n, const int2 * __restrict__ child,
	const uint32_t * __restrict__ parentInner, const uint32_t * __restrict__ parentLeaf, const float4 * leafBox,
	float4 * innerBox, uint32_t * flags) {
	const uint32_t j = blockIdx.x * kGbBlock + threadIdx.x;
	if(j >= n)
		return;
	uint32_t p = parentLeaf[j];
	while(true) {
		__threadfence();
		if(atomicAdd(flags + p, 1u) == 0u)
			return;
		__threadfence();
		const int2 c = child[p];
		float4 amn, amx, bmn, bmx;
		childBox(c.x, leafBox, innerBox, amn, amx);
		childBox(c.y, leafBox, innerBox, bmn, bmx);
		// Bounds.cpp:17-20 merge.
		const f3 mn = gmin(mk(amn.x, amn.y, amn.z), mk(bmn.x, bmn.y, bmn.z));
		const f3 mx = gmax(mk(amx.x, amx.y, amx.z), mk(bmx.x, bmx.y, bmx.z));
		innerBox[2 * size_t(p)] = make_float4(mn.x, mn.y, mn.z, 0.0f);
		innerBox[2 * size_t(p) + 1] = make_float4(mx.x, mx.y, mx.z, 0.0f);
		if(p == 0)
			return;
		p = parentInner[p];
	}
}

__device__ __forceinline__ float halfArea(float4 mn, float4 mx) {
	const float sx = mx.x - mn.x, sy = mx.y - mn.y, sz = mx.z - mn.z;
	return sx * sy + sy * sz + sz * sx;
}

// One level of the 4-wide collapse: each queued (binary node, wide slot) takes its two children and
// expands the internal child of largest surface area until it holds 4 (or only leaves remain),
// writes the wide node and queues its internal children with fresh slots.
__device__ __forceinline__ void collapseOne(const uint2 it, uint2 * __restrict__ qOut, uint32_t * __restrict__ countOut,
	uint32_t * __restrict__ wideCount, const int2 * __restrict__ child, const float4 * __restrict__ leafBox,
	const float4 * __restrict__ innerBox, float4 * __restrict__ wide) {
	int kids[4];
	const int2 c0 = child[it.x];
	kids[0] = c0.x;
	kids[1] = c0.y;
	int nk = 2;
	// Leaves stay single triangles: measured faster than merging sibling leaves into 2-triangle
	// leaves (the 4-wide node then culls each triangle by its own box).
	while(nk < 4) {
		int best = -1;
		float bestArea = -1.0f;
		for(int k = 0; k < nk; ++k)
			if(kids[k] >= 0) {
				const float a = halfArea(innerBox[2 * size_t(kids[k])], innerBox[2 * size_t(kids[k]) + 1]);
				if(a > bestArea) {
					best = k;
					bestArea = a;
				}
			}
		if(best < 0)
			break;
		const int2 cc = child[kids[best]];
		for(int k = nk; k > best + 1; --k)
			kids[k] = kids[k - 1];
		kids[best] = cc.x;
		kids[best + 1] = cc.y;
		++nk;
	}
	float lo[3][4], hi[3][4];
	int codes[4];
	for(int k = 0; k < 4; ++k) {
		float4 mn = make_float4(0.0f, 0.0f, 0.0f, 0.0f), mx = mn;
		int code = kWideEmpty;
		if(k < nk) {
			const int c = kids[k];
			if(c < 0) {
				const uint32_t leaf = uint32_t(-c - 1);
				mn = leafBox[2 * size_t(leaf)];
				mx = leafBox[2 * size_t(leaf) + 1];
				code = ~int32_t(leaf << 1); // one triangle: ~(first << 1 | 0)
			} else {
				mn = innerBox[2 * size_t(c)];
				mx = innerBox[2 * size_t(c) + 1];
				const uint32_t slot = atomicAdd(wideCount, 1u);
				const uint32_t q = atomicAdd(countOut, 1u);
				qOut[q] = make_uint2(uint32_t(c), slot);
				code = int32_t(slot);
			}
		}
		lo[0][k] = mn.x;
		lo[1][k] = mn.y;
		lo[2][k] = mn.z;
		hi[0][k] = mx.x;
		hi[1][k] = mx.y;
		hi[2][k] = mx.z;
		codes[k] = code;
	}
	float4 * o = wide + 8 * size_t(it.y);
	for(int a = 0; a < 3; ++a) {
		o[2 * a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
		o[2 * a + 1] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
	}
	o[6] = make_float4(__int_as_float(codes[0]), __int_as_float(codes[1]), __int_as_float(codes[2]),
		__int_as_float(codes[3]));
	o[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Grid-stride over the level's queue, whose length is read on the device: levels can be launched
// back to back without the host knowing their sizes.
__global__ __launch_bounds__(kGbBlock) void gbCollapse(const uint2 * __restrict__ qIn, const uint32_t * __restrict__ countIn,
	uint2 * __restrict__ qOut, uint32_t * __restrict__ countOut, uint32_t * __restrict__ wideCount,
	const int2 * __restrict__ child, const float4 * __restrict__ leafBox, const float4 * __restrict__ innerBox,
	float4 * __restrict__ wide) {
	const uint32_t count = *countIn;
	for(uint32_t i = blockIdx.x * kGbBlock + threadIdx.x; i < count; i += gridDim.x * kGbBlock)
		collapseOne(qIn[i], qOut, countOut, wideCount, child, leafBox, innerBox, wide);
}

__global__ void gbZero(uint32_t * p, size_t n) {
	const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if(i < n)
		p[i] = 0u;
}

__global__ void gbSeedQueue(uint2 * q, uint32_t * count, uint32_t * wideCount) {
	if(threadIdx.x == 0) {
		q[0] = make_uint2(0u, 0u);
		*count = 1u;
		*wideCount = 1u;
	}
}

// One stream-ordered block (from the library's pool) for every temporary of a build: allocation
// and free order with the build's stream, where a plain hipFree would wait for the whole device,
// so builds overlap other streams' renders. Sub-allocations are 256-B aligned.
class Arena {
public:
	explicit Arena(hipStream_t s) : s_(s) {}
	Arena(const Arena &) = delete;
	Arena & operator=(const Arena &) = delete;
	~Arena() {
		if(base_)
			(void)hipFreeAsync(base_, s_);
	}
	size_t add(size_t bytes) {
		const size_t o = size_;
		size_ = (size_ + bytes + 255) & ~size_t(255);
		return o;
	}
	int allocate() {
		return poolAlloc(&base_, size_ ? size_ : 256, s_);
	}
	template<typename T>
	T * at(size_t off) const {
		return reinterpret_cast<T *>(static_cast<char *>(base_) + off);
	}

private:
	hipStream_t s_;
	void * base_ = nullptr;
	size_t size_ = 0;
};

#define GB_CHECK(x)                                 \
	do {                                            \
		const hipError_t e_ = (x);                  \
		if(e_ != hipSuccess)                        \
			return hipCheck(e_, "GPU BVH build");   \
	} while(0)

} // namespace

int gpuBuildBvh(const HostScene & hs, GpuBvh & out, hipStream_t s) {
	const size_t n = hs.tris.size();
	if(n == 0)
		return fail(RDU_ERR_STATE, "no triangles to build over");
	if(n >= (size_t(1) << 30))
		return fail(RDU_ERR_UNSUPPORTED, "GPU build above 2^30 triangles");
	// Inputs: transformed vertices and per-triangle {v0, v1, v2, meshId, localId}.
	std::vector<uint32_t> tri(5 * n);
	for(size_t i = 0; i < n; ++i) {
		const HTri & t = hs.tris[i];
		tri[5 * i] = t.v0;
		tri[5 * i + 1] = t.v1;
		tri[5 * i + 2] = t.v2;
		tri[5 * i + 3] = t.meshId;
		tri[5 * i + 4] = t.localId;
	}
	// Every temporary lives in one stream-ordered block (see Arena), laid out before it is allocated.
	const size_t ni = n - 1;
	size_t tb = 0;
	GB_CHECK(rocprim::radix_sort_pairs(nullptr, tb, static_cast<uint32_t *>(nullptr), static_cast<uint32_t *>(nullptr),
		static_cast<uint32_t *>(nullptr), static_cast<uint32_t *>(nullptr), n, 0, 30, s));
	Arena ar(s);
	const size_t oVtx = ar.add(hs.vertices.size() * sizeof(f3)), oTri = ar.add(tri.size() * 4), oRec = ar.add(n * 48),
				 oBox = ar.add(n * 32), oCen = ar.add(n * 16), oKeys = ar.add(n * 4), oKeysAlt = ar.add(n * 4),
				 oVals = ar.add(n * 4), oValsAlt = ar.add(n * 4), oTemp = ar.add(tb), oLeaf = ar.add(n * 32),
				 oBounds = ar.add(12 * 4), oCounts = ar.add(4 * 4), oInner = ar.add(ni * 32), oChild = ar.add(ni * 8),
				 oParI = ar.add(ni * 4), oParL = ar.add(n * 4), oFlags = ar.add(ni * 4), oQA = ar.add(ni * 8),
				 oQB = ar.add(ni * 8);
	int st;
	if((st = ar.allocate()) || (st = out.tris.reserve(n * 48)) || (ni && (st = out.wide.reserve(ni * 128))))
		return st;
	GB_CHECK(hipMemcpyAsync(ar.at<void>(oVtx), hs.vertices.data(), hs.vertices.size() * sizeof(f3), hipMemcpyHostToDevice, s));
	GB_CHECK(hipMemcpyAsync(ar.at<void>(oTri), tri.data(), tri.size() * 4, hipMemcpyHostToDevice, s));
	int * dB = ar.at<int>(oBounds);
	hipLaunchKernelGGL(gbInitBounds, dim3(1), dim3(64), 0, s, dB);
	hipLaunchKernelGGL(gbPrep, dim3(gbBlocks(n)), dim3(kGbBlock), 0, s, ar.at<const float>(oVtx),
		ar.at<const uint32_t>(oTri), uint32_t(n), ar.at<float4>(oRec), ar.at<float4>(oBox), ar.at<float4>(oCen), dB);
	hipLaunchKernelGGL(gbMorton, dim3(gbBlocks(n)), dim3(kGbBlock), 0, s, ar.at<const float4>(oCen), uint32_t(n), dB,
		ar.at<uint32_t>(oKeysAlt), ar.at<uint32_t>(oValsAlt));
	GB_CHECK(hipGetLastError());
	GB_CHECK(rocprim::radix_sort_pairs(ar.at<void>(oTemp), tb, ar.at<uint32_t>(oKeysAlt), ar.at<uint32_t>(oKeys),
		ar.at<uint32_t>(oValsAlt), ar.at<uint32_t>(oVals), n, 0, 30, s));
	hipLaunchKernelGGL(gbGather, dim3(gbBlocks(n)), dim3(kGbBlock), 0, s, ar.at<const uint32_t>(oVals), uint32_t(n),
		ar.at<const float4>(oRec), ar.at<const float4>(oBox), static_cast<float4 *>(out.tris.ptr), ar.at<float4>(oLeaf));
	GB_CHECK(hipGetLastError());
	int bh[12];
	GB_CHECK(hipMemcpyAsync(bh, dB, sizeof(bh), hipMemcpyDeviceToHost, s));
	GB_CHECK(hipStreamSynchronize(s));
	for(int a = 0; a < 3; ++a) {
		out.sceneMin[a] = fromOrderedInt(bh[a]);
		out.sceneMax[a] = fromOrderedInt(bh[3 + a]);
	}
	out.triangles = uint32_t(n);
	if(n == 1) {
		// A single leaf: no internal node, no 4-wide node.
		out.wideCount = 0;
		out.wideDepth = 0;
		out.topCode = ~int32_t(0);
		out.innerCount = 0;
		return 0;
	}
	const int2 * childC = ar.at<const int2>(oChild);
	const float4 * leafC = ar.at<const float4>(oLeaf);
	const float4 * innerC = ar.at<const float4>(oInner);
	hipLaunchKernelGGL(gbKarras, dim3(gbBlocks(ni)), dim3(kGbBlock), 0, s, ar.at<const uint32_t>(oKeys), uint32_t(n),
		ar.at<int2>(oChild), ar.at<uint32_t>(oParI), ar.at<uint32_t>(oParL));
	hipLaunchKernelGGL(gbZero, dim3(gbBlocks(ni)), dim3(kGbBlock), 0, s, ar.at<uint32_t>(oFlags), ni);
	hipLaunchKernelGGL(gbBounds, dim3(gbBlocks(n)), dim3(kGbBlock), 0, s, uint32_t(n), childC, ar.at<const uint32_t>(oParI),
		ar.at<const uint32_t>(oParL), leafC, ar.at<float4>(oInner), ar.at<uint32_t>(oFlags));
	GB_CHECK(hipGetLastError());
	// Collapse, one level per launch; the host reads each level's count (a few bytes) to stop.
	uint32_t * dc = ar.at<uint32_t>(oCounts); // [0], [1]: queue counts; [2]: wide nodes
	hipLaunchKernelGGL(gbSeedQueue, dim3(1), dim3(64), 0, s, ar.at<uint2>(oQA), dc, dc + 2);
	uint2 * q[2] = {ar.at<uint2>(oQA), ar.at<uint2>(oQB)};
	uint32_t levels = 0, pending = 1;
	int cur = 0;
	// Levels go in batches of 16 launches with a fixed grid (grid-stride inside); the host reads the
	// last level's count once per batch. Levels past the last non-empty one find an empty queue.
	const unsigned grid = gbBlocks(ni) < 2048u ? gbBlocks(ni) : 2048u;
	uint32_t sizes[16];
	while(pending > 0) {
		if(levels > 4096)
			return fail(RDU_ERR_UNSUPPORTED, "GPU build: tree deeper than 4096 levels");
		for(int k = 0; k < 16; ++k) {
			hipLaunchKernelGGL(gbZero, dim3(1), dim3(64), 0, s, dc + (1 - cur), size_t(1));
			hipLaunchKernelGGL(gbCollapse, dim3(grid), dim3(kGbBlock), 0, s, q[cur], dc + cur, q[1 - cur], dc + (1 - cur),
				dc + 2, childC, leafC, innerC, static_cast<float4 *>(out.wide.ptr));
			GB_CHECK(hipMemcpyAsync(sizes + k, dc + (1 - cur), 4, hipMemcpyDeviceToHost, s));
			cur = 1 - cur;
		}
		GB_CHECK(hipGetLastError());
		GB_CHECK(hipStreamSynchronize(s));
		// Depth = levels with a non-empty input queue.
		for(int k = 0; k < 16 && pending > 0; ++k) {
			++levels;
			pending = sizes[k];
		}
	}
	uint32_t wideCount = 0;
	float4 root[2];
	GB_CHECK(hipMemcpyAsync(&wideCount, dc + 2, 4, hipMemcpyDeviceToHost, s));
	GB_CHECK(hipMemcpyAsync(root, innerC, 32, hipMemcpyDeviceToHost, s));
	GB_CHECK(hipStreamSynchronize(s));
	out.wideCount = wideCount;
	out.wideDepth = levels;
	out.topCode = 0;
	out.innerCount = uint32_t(ni);
	out.topMin[0] = root[0].x;
	out.topMin[1] = root[0].y;
	out.topMin[2] = root[0].z;
	out.topMax[0] = root[1].x;
	out.topMax[1] = root[1].y;
	out.topMax[2] = root[1].z;
	return 0;
}

} // namespace rdu
