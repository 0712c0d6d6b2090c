// Batch ray/BVH traversal kernels for gfx950: closest hit (Raycaster::intersects,
// src/engine/raycaster/Raycaster.cpp:144-185) and any hit (Raycaster::intersectsAny,
// Raycaster.cpp:187-224) over the flattened tree of bvh.hpp.
//
// One ray per lane, 64-lane wavefronts. Each lane keeps its traversal stack in LDS, laid out
// [depth][lane] so that a wavefront's pushes and pops of one level hit 64 consecutive banks;
// trees deeper than the LDS budget spill the excess levels to a global array laid out
// [level][ray] (coalesced). Internal nodes hold both children's boxes, so one node fetch
// (4 x 16 B) feeds the two slab tests of an expansion; triangles are 3 x 16 B records.
//
// Two visiting orders:
//  - reference order: Rendu's depth-first order exactly (roots pushed 0..M-1 and tested with the
//    query's maxi, children tested left then right with the current maxi when their parent is
//    expanded, right popped first). Same boxes, same tests, same order: hits are the reference's
//    bit for bit, including which of several equal-t triangles is returned.
//  - near-first order: the nearer child (by slab entry) is visited first and the farther one is
//    pushed. Box tests are conservative, so the closest hit is the same triangle with the same
//    t, u, v, except where two triangles tie at equal t (shared edges), where the reference keeps
//    whichever its order met first. Any-hit results are order independent.
#include "trace.hpp"
#include "../../include/rdu.h"

#include <cfloat>
#include <cstdlib>

namespace rdu {

namespace {

constexpr int kBlock = 128;


__device__ __forceinline__ f3 xyz(float4 v) { return mk(v.x, v.y, v.z); }

template<bool SPILL>
__device__ __forceinline__ void push(int (*ls)[kBlock], int * gs, size_t gstride, size_t gid, int lds, int & sp, int v) {
	if(!SPILL || sp < lds)
		ls[sp][threadIdx.x] = v;
	else
		gs[size_t(sp - lds) * gstride + gid] = v;
	++sp;
}

template<bool SPILL>
__device__ __forceinline__ int pop(int (*ls)[kBlock], const int * gs, size_t gstride, size_t gid, int lds, int & sp) {
	--sp;
	if(!SPILL || sp < lds)
		return ls[sp][threadIdx.x];
	return gs[size_t(sp - lds) * gstride + gid];
}

__device__ __forceinline__ uint64_t waveSum(uint64_t v) {
	for(int off = 32; off > 0; off >>= 1)
		v += __shfl_xor(v, off, 64);
	return v;
}

template<bool ANY, bool REF, bool SPILL, bool STATS>
__global__ __launch_bounds__(kBlock) void traceKernel(DevBVH bvh, const float4 * __restrict__ rays, size_t n,
	rdu_hit * __restrict__ hits, uint8_t * __restrict__ occ, int * __restrict__ gstack, int lds, uint64_t * stats) {
	__shared__ int ls[kLdsStack][kBlock];
	const size_t gid = size_t(blockIdx.x) * kBlock + threadIdx.x;
	const size_t gstride = size_t(gridDim.x) * kBlock;
	uint64_t nodeVisits = 0, triTests = 0;
	if(gid < n) {
		const float4 ra = rays[2 * gid];
		const float4 rb = rays[2 * gid + 1];
		const DevRay r = makeRay(xyz(ra), xyz(rb), ra.w, rb.w);
		const float mini = r.mini;
		const float maxi0 = r.maxi;
		float maxi = maxi0;
		int best = -1;
		float bt = FLT_MAX, bu = 0.0f, bv = 0.0f;
		bool found = false;

		// Rendu pushes every root whose box passes (tested with the query's maxi), then pops the
		// last one first: visiting roots from M-1 down to 0 is the same order.
		for(int ri = int(bvh.nRoots) - 1; ri >= 0 && !found; --ri) {
			const float4 rmin = bvh.roots[2 * ri];
			const float4 rmax = bvh.roots[2 * ri + 1];
			float e;
			if(!slab(r, xyz(rmin), xyz(rmax), mini, REF ? maxi0 : maxi, e))
				continue;
			int sp = 0;
			int c = __float_as_int(rmin.w);
			while(true) {
				if(c >= 0) {
					if(STATS)
						++nodeVisits;
					const float4 n0 = bvh.nodes[4 * c + 0];
					const float4 n1 = bvh.nodes[4 * c + 1];
					const float4 n2 = bvh.nodes[4 * c + 2];
					const float4 n3 = bvh.nodes[4 * c + 3];
					float eL, eR;
					const bool hL = slab(r, mk(n0.x, n0.y, n0.z), mk(n0.w, n1.x, n1.y), mini, maxi, eL);
					const bool hR = slab(r, mk(n1.z, n1.w, n2.x), mk(n2.y, n2.z, n2.w), mini, maxi, eR);
					const int lc = __float_as_int(n3.x);
					const int rc = __float_as_int(n3.y);
					if(REF) {
						// Push left then right; the right child is popped first.
						if(hL)
							push<SPILL>(ls, gstack, gstride, gid, lds, sp, lc);
						if(hR) {
							c = rc;
							continue;
						}
					} else {
						if(hL && hR) {
							const bool leftNear = eL <= eR;
							push<SPILL>(ls, gstack, gstride, gid, lds, sp, leftNear ? rc : lc);
							c = leftNear ? lc : rc;
							continue;
						}
						if(hL) {
							c = lc;
							continue;
						}
						if(hR) {
							c = rc;
							continue;
						}
					}
				} else {
					const int code = ~c;
					const int first = code >> 1;
					const int cnt = (code & 1) + 1;
					for(int k = 0; k < cnt; ++k) {
						if(STATS)
							++triTests;
						const int ti = first + k;
						const float4 t0 = bvh.tris[3 * ti + 0];
						const float4 t1 = bvh.tris[3 * ti + 1];
						const float4 t2 = bvh.tris[3 * ti + 2];
						float tt, uu, vv;
						if(mollerTrumbore(r, xyz(t0), mk(t0.w, t1.x, t1.y), mk(t1.z, t1.w, t2.x), mini, maxi, tt, uu, vv)) {
							if(ANY) {
								found = true;
								break;
							}
							// hit.dist < bestHit.dist holds whenever the test passed: t < maxi = best.
							best = ti;
							bt = tt;
							bu = uu;
							bv = vv;
							maxi = tt;
						}
					}
					if(ANY && found)
						break;
				}
				if(sp == 0)
					break;
				c = pop<SPILL>(ls, gstack, gstride, gid, lds, sp);
			}
		}
		if(ANY) {
			occ[gid] = found ? 1 : 0;
		} else {
			rdu_hit h;
			if(best >= 0) {
				const float4 t2 = bvh.tris[3 * best + 2];
				h.t = bt;
				h.u = bu;
				h.v = bv;
				h.hit = 1u;
				h.mesh_id = __float_as_uint(t2.y);
				h.local_id = __float_as_uint(t2.z);
			} else {
				h.t = FLT_MAX;
				h.u = 0.0f;
				h.v = 0.0f;
				h.hit = 0u;
				h.mesh_id = 0u;
				h.local_id = 0u;
			}
			hits[gid] = h;
		}
	}
	if(STATS) {
		const uint64_t a = waveSum(nodeVisits);
		const uint64_t b = waveSum(triTests);
		if((threadIdx.x & 63) == 0) {
			atomicAdd(reinterpret_cast<unsigned long long *>(stats), (unsigned long long)a);
			atomicAdd(reinterpret_cast<unsigned long long *>(stats + 1), (unsigned long long)b);
		}
	}
}

template<bool ANY>
int launchImpl(const DevBVH & bvh, const float4 * rays, size_t n, rdu_hit * hits, uint8_t * occ, bool ref, int * gstack,
	uint64_t * stats, hipStream_t stream) {
	if(n == 0)
		return 0;
	const int lds = ldsStackDepth();
	const bool spill = int(bvh.maxDepth) + 2 > lds;
	const dim3 grid(unsigned((n + kBlock - 1) / kBlock));
	const dim3 block(kBlock);
#define RDU_LAUNCH(R, S, ST) \
	hipLaunchKernelGGL((traceKernel<ANY, R, S, ST>), grid, block, 0, stream, bvh, rays, n, hits, occ, gstack, lds, stats)
	if(stats) {
		if(ref) {
			if(spill) RDU_LAUNCH(true, true, true); else RDU_LAUNCH(true, false, true);
		} else {
			if(spill) RDU_LAUNCH(false, true, true); else RDU_LAUNCH(false, false, true);
		}
	} else {
		if(ref) {
			if(spill) RDU_LAUNCH(true, true, false); else RDU_LAUNCH(true, false, false);
		} else {
			if(spill) RDU_LAUNCH(false, true, false); else RDU_LAUNCH(false, false, false);
		}
	}
#undef RDU_LAUNCH
	return int(hipGetLastError());
}

} // namespace

int ldsStackDepth() {
	// RDU_DEBUG_LDS_STACK=k (1 <= k <= 32) keeps only k stack levels in LDS so that tests can
	// exercise the global-memory spill path on ordinary trees.
	const char * v = std::getenv("RDU_DEBUG_LDS_STACK");
	const int k = v ? std::atoi(v) : kLdsStack;
	return (k >= 1 && k <= kLdsStack) ? k : kLdsStack;
}

size_t globalStackInts(const DevBVH & bvh, size_t n) {
	const int need = int(bvh.maxDepth) + 2 - ldsStackDepth();
	if(need <= 0)
		return 0;
	const size_t slots = ((n + kBlock - 1) / kBlock) * kBlock;
	return size_t(need) * slots;
}

int launchClosest(const DevBVH & bvh, const float4 * rays, size_t n, void * hits, bool referenceOrder, int * gstack,
	uint64_t * stats, hipStream_t stream) {
	return launchImpl<false>(bvh, rays, n, static_cast<rdu_hit *>(hits), nullptr, referenceOrder, gstack, stats, stream);
}

int launchAny(const DevBVH & bvh, const float4 * rays, size_t n, uint8_t * occluded, bool referenceOrder, int * gstack,
	uint64_t * stats, hipStream_t stream) {
	return launchImpl<true>(bvh, rays, n, nullptr, occluded, referenceOrder, gstack, stats, stream);
}

} // namespace rdu
