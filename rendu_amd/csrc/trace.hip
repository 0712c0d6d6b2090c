// Batch ray/BVH traversal kernels for gfx950: closest hit (Raycaster::intersects,
// src/engine/raycaster/Raycaster.cpp:144-185) and any hit (Raycaster::intersectsAny,
// Raycaster.cpp:187-224) over the flattened trees of bvh.hpp.
//
// One ray per lane, 64-lane wavefronts. Each lane keeps its traversal stack in LDS, laid out
// [depth][lane] so that a wavefront's pushes and pops of one level hit 64 consecutive banks;
// stacks deeper than the LDS budget spill the excess levels to a global array laid out
// [level][lane slot] (coalesced).
//
// Visiting orders:
//  - reference order (binary tree): Rendu's depth-first order exactly (roots pushed 0..M-1 and
//    tested with the query's maxi, children tested left then right with the current maxi when
//    their parent is expanded, right popped first). Same boxes, same tests, same order: hits are
//    the reference's bit for bit, including which of several equal-t triangles is returned.
//  - near-first order (binary or 4-wide tree): the nearest child by slab entry is visited first,
//    the others are pushed farthest first. Every box is one of Rendu's and the tests are the
//    same, so the closest hit is the same triangle with the same t, u, v, except where
//    triangles tie within rounding (shared edges), where the reference keeps whichever its order
//    met first. Any-hit results are order independent.
//
// Slab tests use Rendu's select-based min/max, whose NaN behaviour differs from IEEE minNum.
// When no lane of a wavefront has an infinite reciprocal direction or a non-finite origin, no
// slab product can be NaN and the selects equal v_min/v_max (values are only compared, so the
// sign of a zero result is irrelevant): such wavefronts take the min3/max3 path, the others the
// select path. The choice is wave-uniform (ballot), so it costs no divergence.
#include "trace.hpp"
#include "../../include/rdu.h"

#include <cfloat>
#include <cstdlib>

#include <rocprim/device/device_radix_sort.hpp>

namespace rdu {

namespace {

constexpr int kBlock = 256;       // persistent kernels (4 wavefronts per workgroup)
constexpr int kSimpleBlock = 128; // one-ray-per-lane kernel
static_assert(kBlock % kSimpleBlock == 0, "spill-stack slots are sized for the persistent grid");
constexpr int kWideLds = 16;

__device__ __forceinline__ f3 xyz(float4 v) { return mk(v.x, v.y, v.z); }

template<bool EXACT>
__device__ __forceinline__ float vmin(float a, float b) {
	if constexpr(EXACT)
		return gmin(a, b);
	else
		return __builtin_fminf(a, b);
}
template<bool EXACT>
__device__ __forceinline__ float vmax(float a, float b) {
	if constexpr(EXACT)
		return gmax(a, b);
	else
		return __builtin_fmaxf(a, b);
}

// Intersection::box (Intersection.cpp:22-32) on one box given by its six bounds.
template<bool EXACT>
__device__ __forceinline__ bool slab6(const DevRay & r, float lx, float ly, float lz, float hx, float hy, float hz,
	float mini, float maxi, float & entry) {
	const float ax = (lx - r.pos.x) * r.inv.x, bx = (hx - r.pos.x) * r.inv.x;
	const float ay = (ly - r.pos.y) * r.inv.y, by = (hy - r.pos.y) * r.inv.y;
	const float az = (lz - r.pos.z) * r.inv.z, bz = (hz - r.pos.z) * r.inv.z;
	const float mnx = vmin<EXACT>(ax, bx), mny = vmin<EXACT>(ay, by), mnz = vmin<EXACT>(az, bz);
	const float mxx = vmax<EXACT>(ax, bx), mxy = vmax<EXACT>(ay, by), mxz = vmax<EXACT>(az, bz);
	const float closest = vmax<EXACT>(mnx, vmax<EXACT>(mny, mnz));
	const float furthest = vmin<EXACT>(mxx, vmin<EXACT>(mxy, mxz));
	entry = vmax<EXACT>(closest, mini);
	return entry <= vmin<EXACT>(furthest, maxi);
}

// Fast-path slab test on planes already ordered per axis by the sign of the reciprocal
// direction (n = near plane, f = far plane): for finite operands (a - p) * inv is monotonic in a,
// so the near plane's product is min(a, b) and the far plane's is max(a, b) of the reference.
__device__ __forceinline__ bool slabNearFar(const DevRay & r, float nx, float ny, float nz, float fx, float fy, float fz,
	float mini, float maxi, float & entry) {
	const float ax = (nx - r.pos.x) * r.inv.x, bx = (fx - r.pos.x) * r.inv.x;
	const float ay = (ny - r.pos.y) * r.inv.y, by = (fy - r.pos.y) * r.inv.y;
	const float az = (nz - r.pos.z) * r.inv.z, bz = (fz - r.pos.z) * r.inv.z;
	entry = __builtin_fmaxf(__builtin_fmaxf(ax, __builtin_fmaxf(ay, az)), mini);
	return entry <= __builtin_fminf(__builtin_fminf(bx, __builtin_fminf(by, bz)), maxi);
}

// 16-byte load from a global-memory address held as an integer (keeps it a global, not a flat, load).
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ldGlobal(uintptr_t a) {
	const f4v v = *reinterpret_cast<const __attribute__((address_space(1))) f4v *>(a);
	return make_float4(v.x, v.y, v.z, v.w);
}

template<int B>
struct LaneStack {
	int (*ls)[B];
	int * gs;
	size_t gstride, slot;
	int lds;
	int sp = 0;
	template<bool SPILL>
	__device__ __forceinline__ void push(int v) {
		if(!SPILL || sp < lds)
			ls[sp][threadIdx.x] = v;
		else
			gs[size_t(sp - lds) * gstride + slot] = v;
		++sp;
	}
	template<bool SPILL>
	__device__ __forceinline__ int pop() {
		--sp;
		if(!SPILL || sp < lds)
			return ls[sp][threadIdx.x];
		return gs[size_t(sp - lds) * gstride + slot];
	}
};

__device__ __forceinline__ uint64_t waveSum(uint64_t v) {
	for(int off = 32; off > 0; off >>= 1)
		v += __shfl_xor(v, off, 64);
	return v;
}

struct TraceState {
	float maxi;
	int best = -1;
	float bt = FLT_MAX, bu = 0.0f, bv = 0.0f;
	bool found = false;
	uint64_t nodeVisits = 0, triTests = 0;
};

// Leaf: Moller-Trumbore on its 1-2 triangles in order. Returns true when an any-hit query is done.
template<bool ANY, bool STATS>
__device__ __forceinline__ bool testLeaf(const DevBVH & bvh, const DevRay & r, int c, float mini, TraceState & s) {
	const int code = ~c;
	const int first = code >> 1;
	const int cnt = (code & 1) + 1;
	for(int k = 0; k < cnt; ++k) {
		if(STATS)
			++s.triTests;
		const int ti = first + k;
		const float4 t0 = bvh.tris[3 * ti + 0];
		const float4 t1 = bvh.tris[3 * ti + 1];
		const float4 t2 = bvh.tris[3 * ti + 2];
		float tt, uu, vv;
		if(mollerTrumbore(r, xyz(t0), mk(t0.w, t1.x, t1.y), mk(t1.z, t1.w, t2.x), mini, s.maxi, tt, uu, vv)) {
			if(ANY) {
				s.found = true;
				return true;
			}
			// hit.dist < bestHit.dist holds whenever the test passed: t < maxi = best.
			s.best = ti;
			s.bt = tt;
			s.bu = uu;
			s.bv = vv;
			s.maxi = tt;
		}
	}
	return false;
}

// Binary tree, reference or near-first order, from one root code.
template<bool ANY, bool REF, bool SPILL, bool EXACT, bool STATS>
__device__ __forceinline__ void traverseBinary(const DevBVH & bvh, const DevRay & r, int c, float mini,
	LaneStack<kSimpleBlock> & st, TraceState & s) {
	st.sp = 0;
	while(true) {
		if(c >= 0) {
			if(STATS)
				++s.nodeVisits;
			const float4 n0 = bvh.nodes[4 * c + 0];
			const float4 n1 = bvh.nodes[4 * c + 1];
			const float4 n2 = bvh.nodes[4 * c + 2];
			const float4 n3 = bvh.nodes[4 * c + 3];
			float eL, eR;
			const bool hL = slab6<EXACT>(r, n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, mini, s.maxi, eL);
			const bool hR = slab6<EXACT>(r, n1.z, n1.w, n2.x, n2.y, n2.z, n2.w, mini, s.maxi, eR);
			const int lc = __float_as_int(n3.x);
			const int rc = __float_as_int(n3.y);
			if(REF) {
				// Push left then right; the right child is popped first.
				if(hL)
					st.push<SPILL>(lc);
				if(hR) {
					c = rc;
					continue;
				}
			} else {
				if(hL && hR) {
					const bool leftNear = eL <= eR;
					st.push<SPILL>(leftNear ? rc : lc);
					c = leftNear ? lc : rc;
					continue;
				}
				if(hL || hR) {
					c = hL ? lc : rc;
					continue;
				}
			}
		} else if(testLeaf<ANY, STATS>(bvh, r, c, mini, s)) {
			return;
		}
		if(st.sp == 0)
			return;
		c = st.pop<SPILL>();
	}
}

// fp16 value not above e: round towards zero (v_cvt_pkrtz; beyond the fp16 range it gives the
// largest finite half), which is towards -inf for e >= 0; negative entries (a negative tmin) are
// stored as -inf, which never culls.
__device__ __forceinline__ _Float16 roundDownHalf(float e) {
	return e < 0.0f ? _Float16(-__builtin_inff()) : __builtin_amdgcn_cvt_pkrtz(e, e)[0];
}

__device__ __forceinline__ void cswap(float & a, int & ca, float & b, int & cb) {
	const bool sw = b < a;
	const float ta = sw ? b : a, tb = sw ? a : b;
	const int tca = sw ? cb : ca, tcb = sw ? ca : cb;
	a = ta;
	b = tb;
	ca = tca;
	cb = tcb;
}

// 4-wide tree, near-first order, from one root code.
template<bool ANY, bool SPILL, bool EXACT, bool STATS>
__device__ __forceinline__ void traverseWide(const DevBVH & bvh, const DevRay & r, int c, float mini, LaneStack<kSimpleBlock> & st,
	TraceState & s) {
	st.sp = 0;
	while(true) {
		if(c >= 0) {
			if(STATS)
				++s.nodeVisits;
			const float4 * nd = bvh.wide + 8 * size_t(c);
			const float4 lx = nd[0], hx = nd[1], ly = nd[2], hy = nd[3], lz = nd[4], hz = nd[5];
			const float4 cf = nd[6];
			int c0 = __float_as_int(cf.x), c1 = __float_as_int(cf.y), c2 = __float_as_int(cf.z),
				c3 = __float_as_int(cf.w);
			float e0, e1, e2 = 0.0f, e3 = 0.0f;
			const bool h0 = slab6<EXACT>(r, lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, mini, s.maxi, e0);
			const bool h1 = slab6<EXACT>(r, lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, mini, s.maxi, e1);
			const bool h2 = (c2 != kWideEmpty) && slab6<EXACT>(r, lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, mini, s.maxi, e2);
			const bool h3 = (c3 != kWideEmpty) && slab6<EXACT>(r, lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, mini, s.maxi, e3);
			const int nh = int(h0) + int(h1) + int(h2) + int(h3);
			if(nh == 0) {
				if(st.sp == 0)
					return;
				c = st.pop<SPILL>();
				continue;
			}
			// Hit entries are finite (clamped below +inf), misses are +inf: after the sorting
			// network the nh hit children come first, nearest first.
			e0 = h0 ? __builtin_fminf(e0, FLT_MAX) : __builtin_inff();
			e1 = h1 ? __builtin_fminf(e1, FLT_MAX) : __builtin_inff();
			e2 = h2 ? __builtin_fminf(e2, FLT_MAX) : __builtin_inff();
			e3 = h3 ? __builtin_fminf(e3, FLT_MAX) : __builtin_inff();
			cswap(e0, c0, e1, c1);
			cswap(e2, c2, e3, c3);
			cswap(e0, c0, e2, c2);
			cswap(e1, c1, e3, c3);
			cswap(e1, c1, e2, c2);
			if(nh > 3)
				st.push<SPILL>(c3);
			if(nh > 2)
				st.push<SPILL>(c2);
			if(nh > 1)
				st.push<SPILL>(c1);
			c = c0;
			continue;
		}
		if(testLeaf<ANY, STATS>(bvh, r, c, mini, s))
			return;
		if(st.sp == 0)
			return;
		c = st.pop<SPILL>();
	}
}

template<bool ANY, bool REF, bool WIDE, bool SPILL, bool EXACT, bool STATS>
__device__ __forceinline__ void traceRay(const DevBVH & bvh, const DevRay & r, LaneStack<kSimpleBlock> & st, TraceState & s) {
	const float mini = r.mini;
	const float maxi0 = r.maxi;
	// Rendu pushes every root whose box passes (tested with the query's maxi), then pops the last
	// one first: visiting roots from M-1 down to 0 is the same order.
	if(WIDE) {
		// One root: the top of the 4-wide tree.
		float e;
		if(slab6<EXACT>(r, bvh.wideTopMin[0], bvh.wideTopMin[1], bvh.wideTopMin[2], bvh.wideTopMax[0], bvh.wideTopMax[1],
			   bvh.wideTopMax[2], mini, s.maxi, e))
			traverseWide<ANY, SPILL, EXACT, STATS>(bvh, r, bvh.wideTopCode, mini, st, s);
		return;
	}
	for(int ri = int(bvh.nRoots) - 1; ri >= 0; --ri) {
		const float4 rmin = bvh.roots[2 * ri];
		const float4 rmax = bvh.roots[2 * ri + 1];
		float e;
		if(!slab6<EXACT>(r, rmin.x, rmin.y, rmin.z, rmax.x, rmax.y, rmax.z, mini, REF ? maxi0 : s.maxi, e))
			continue;
		traverseBinary<ANY, REF, SPILL, EXACT, STATS>(bvh, r, __float_as_int(rmin.w), mini, st, s);
		if(ANY && s.found)
			return;
	}
}

__device__ __forceinline__ bool finite3(f3 v) {
	return __builtin_isfinite(v.x) && __builtin_isfinite(v.y) && __builtin_isfinite(v.z);
}

template<bool ANY, bool REF, bool WIDE, bool SPILL, bool STATS>
__global__ __launch_bounds__(kSimpleBlock) void traceKernel(DevBVH bvh, const float4 * __restrict__ rays, size_t n,
	const uint32_t * __restrict__ perm, const uint32_t * __restrict__ countDev, rdu_hit * __restrict__ hits,
	uint8_t * __restrict__ occ, int * __restrict__ gstack, int lds, uint64_t * stats) {
	__shared__ int ls[WIDE ? kWideLds : kLdsStack][kSimpleBlock];
	const size_t gid = size_t(blockIdx.x) * kSimpleBlock + threadIdx.x;
	LaneStack<kSimpleBlock> st{ls, gstack, size_t(gridDim.x) * kSimpleBlock, gid, lds};
	TraceState s;
	if(countDev)
		n = size_t(*countDev) < n ? size_t(*countDev) : n;
	const bool active = gid < n;
	size_t ri = gid;
	DevRay r;
	if(active) {
		ri = perm ? size_t(perm[gid]) : gid;
		const float4 ra = rays[2 * ri];
		const float4 rb = rays[2 * ri + 1];
		r = makeRay(xyz(ra), xyz(rb), ra.w, rb.w);
	} else {
		r = makeRay(mk(0, 0, 0), mk(0, 0, 1), 0.0f, 0.0f);
	}
	s.maxi = r.maxi;
	const bool clean = finite3(r.inv) && finite3(r.pos);
	if(active) {
		if(__all(clean))
			traceRay<ANY, REF, WIDE, SPILL, false, STATS>(bvh, r, st, s);
		else
			traceRay<ANY, REF, WIDE, SPILL, true, STATS>(bvh, r, st, s);
		if(ANY) {
			occ[ri] = s.found ? 1 : 0;
		} else {
			rdu_hit h;
			if(s.best >= 0) {
				const float4 t2 = bvh.tris[3 * s.best + 2];
				h.t = s.bt;
				h.u = s.bu;
				h.v = s.bv;
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
			hits[ri] = h;
		}
	}
	if(STATS) {
		const uint64_t a = waveSum(s.nodeVisits);
		const uint64_t b = waveSum(s.triTests);
		if((threadIdx.x & 63) == 0) {
			atomicAdd(reinterpret_cast<unsigned long long *>(stats), (unsigned long long)a);
			atomicAdd(reinterpret_cast<unsigned long long *>(stats + 1), (unsigned long long)b);
		}
	}
}

// ALT selects the kernel's variant: for closest hit the near/far plane fetch, for any hit the
// unsorted visiting order; NFA selects the near/far plane fetch for (sorted) any hit.
// BATCH: rays of several scenes whose trees share one node arena; ray k starts at the root of scene
// rayJob[k], whose top box and code are jobRoots[2 rayJob[k]] (min, code bits) and [+1] (max).
// INST: an instanced scene (bvh.inst): the top level's leaves are objects; a shared tree is
// traversed with the ray moved into the object's space, a world-space tree with the ray as it is.
// Waves per SIMD the instanced variants are held to, which a few more registers would otherwise
// cost a wave: any hit 7 (64 registers; 5 with the near/far fetch, 88), closest hit 5 (94
// registers with the near/far fetch; without it, 30 KB of LDS per workgroup with the world rays
// allows no more).
template<bool ANY, bool NFA, bool INST>
constexpr int persistentMinWaves() {
	return !INST ? 1 : ANY ? (NFA ? 5 : 7) : 5;
}
template<bool ANY, bool SPILL, bool STATS, bool ALT, bool NFA = false, bool BATCH = false, bool INST = false>
__global__ __launch_bounds__(kBlock, (persistentMinWaves<ANY, NFA, INST>())) void tracePersistent(DevBVH bvh, const float4 * __restrict__ rays, uint32_t n,
	const uint32_t * __restrict__ perm, const uint32_t * __restrict__ countDev, rdu_hit * __restrict__ hits,
	uint8_t * __restrict__ occ, int * __restrict__ gstack, int lds, uint32_t xcds, uint32_t refillMin, uint32_t chunk,
	const uint32_t * __restrict__ rayJob, const float4 * __restrict__ jobRoots, uint64_t * stats);

// MI355X (gfx950) has 8 XCDs; workgroups are dispatched to them round-robin.
constexpr uint32_t kXcds = 8;

struct PersistentGrid {
	unsigned blocks = 0;
};

// Idle lanes needed before a wavefront refills (RDU_REFILL_MIN overrides, 1..64).
// Refill threshold of a launch: RDU_REFILL_MIN (1..64) when set, else the batch's tuned value,
// else 32 for spatially sorted (incoherent) batches and 16 for input order.
uint32_t refillMin(const TraceOptions & opt) {
	const char * v = std::getenv("RDU_REFILL_MIN");
	if(v) {
		const int k = std::atoi(v);
		if(k >= 1 && k <= 64)
			return uint32_t(k);
	}
	if(opt.refillMin >= 1 && opt.refillMin <= 64)
		return opt.refillMin;
	return opt.spatialOrder ? 32u : 16u;
}

template<bool ANY, bool SPILL, bool STATS, bool ALT, bool NFA, bool BATCH, bool INST = false>
PersistentGrid persistentGrid(size_t n, uint32_t blocksPerCuCap);

template<bool ANY>
int launchImpl(const DevBVH & bvh, const float4 * rays, size_t n, rdu_hit * hits, uint8_t * occ, const TraceOptions & opt,
	int * gstack, uint64_t * stats, hipStream_t stream) {
	if(n == 0)
		return 0;
	const bool wide = opt.wide && !opt.referenceOrder && bvh.wide != nullptr;
	const int lds = wide ? (ldsStackDepth() < kWideLds ? ldsStackDepth() : kWideLds) : ldsStackDepth();
	TraceOptions o = opt;
	o.wide = wide;
	const bool spill = stackNeed(bvh, o) > lds;
	const dim3 grid(unsigned((n + kSimpleBlock - 1) / kSimpleBlock));
	const dim3 block(kSimpleBlock);
	const uint32_t * perm = opt.perm;
#define RDU_LAUNCH(R, W, S, ST)                                                                               \
	hipLaunchKernelGGL((traceKernel<ANY, R, W, S, ST>), grid, block, 0, stream, bvh, rays, n, perm, opt.countDev, hits, \
		occ, gstack, lds, stats)
#define RDU_BY_SPILL(R, W, ST)     \
	if(spill)                      \
		RDU_LAUNCH(R, W, true, ST); \
	else                           \
		RDU_LAUNCH(R, W, false, ST)
	const char * kern = std::getenv("RDU_TRACE_KERNEL");
	const bool persistent = wide && !(kern && kern[0] == 's'); // RDU_TRACE_KERNEL=simple: one ray per lane
	if(bvh.inst && !persistent)
		return int(hipErrorInvalidValue); // instanced trees: persistent 4-wide traversal only
	if(persistent) {
		if(n > 0xffffffffull)
			return int(hipErrorInvalidValue);
#define RDU_PERSIST_X(S, ST, ALT, NFA, B, I)                                                                      \
	{                                                                                                            \
		const PersistentGrid g = persistentGrid<ANY, S, ST, ALT, NFA, B, I>(n, opt.blocksPerCu);                      \
		const uint32_t waves = (g.blocks / kXcds) * (kBlock / 64);                                               \
		const uint64_t slice = (uint64_t(n) + kXcds - 1) / kXcds;                                                \
		/* sorted batches: interleaved 64-ray chunks; input order: one contiguous run per wavefront; */        \
		/* a tuned chunk overrides either */                                                                   \
		uint64_t chunk = opt.spatialOrder ? 64 : (slice + waves - 1) / (waves ? waves : 1);                     \
		chunk = chunk < 64 ? 64 : chunk;                                                                        \
		if(opt.chunk >= 16 && opt.chunk <= (1u << 20))                                                          \
			chunk = opt.chunk;                                                                                  \
		hipLaunchKernelGGL((tracePersistent<ANY, S, ST, ALT, NFA, B, I>), dim3(g.blocks), dim3(kBlock), 0, stream, bvh,  \
			rays, uint32_t(n), perm, opt.countDev, hits, occ, gstack, lds, kXcds, refillMin(opt), uint32_t(chunk),      \
			opt.rayJob, opt.jobRoots, stats);                                                                   \
	}
#define RDU_PERSIST_B(S, ST, ALT, NFA, B) RDU_PERSIST_X(S, ST, ALT, NFA, B, false)
#define RDU_PERSIST_I(S, ALT, NFA, B) RDU_PERSIST_X(S, false, ALT, NFA, B, true)
#define RDU_PERSIST(S, ST, ALT, NFA) RDU_PERSIST_B(S, ST, ALT, NFA, false)
		// Kernel variant: the near/far plane fetch (fewer instructions per node, more registers:
		// 5 waves per SIMD instead of 6 for closest hit and of 7 for any hit; pays on incoherent rays), or
		// for any hit on coherent rays the unsorted visiting order.
		const bool alt = ANY ? opt.unorderedAny : opt.nearFar;
		const bool nfa = ANY && !opt.unorderedAny && opt.nearFar;
		if(bvh.inst && opt.jobRoots) {
			// Batch passes with instanced jobs in the arena: as the batch variants below.
			if(stats || (ANY && !nfa))
				return int(hipErrorInvalidValue);
			if(ANY) {
				if(spill) RDU_PERSIST_I(true, false, true, true) else RDU_PERSIST_I(false, false, true, true)
			} else if(alt) {
				if(spill) RDU_PERSIST_I(true, true, false, true) else RDU_PERSIST_I(false, true, false, true)
			} else {
				if(spill) RDU_PERSIST_I(true, false, false, true) else RDU_PERSIST_I(false, false, false, true)
			}
		} else if(bvh.inst) {
			// Instanced scenes (the path tracer's passes): closest hit with either fetch, any hit
			// with the near/far fetch or in node order; no statistics.
			if(stats)
				return int(hipErrorInvalidValue);
			if(ANY && nfa) {
				if(spill) RDU_PERSIST_I(true, false, true, false) else RDU_PERSIST_I(false, false, true, false)
			} else if(alt) {
				if(spill) RDU_PERSIST_I(true, true, false, false) else RDU_PERSIST_I(false, true, false, false)
			} else {
				if(spill) RDU_PERSIST_I(true, false, false, false) else RDU_PERSIST_I(false, false, false, false)
			}
		} else if(opt.jobRoots) {
			// Batch passes of the path tracer: closest hit with either fetch, any hit with the
			// near/far fetch; no statistics.
			if(stats || (ANY && !nfa))
				return int(hipErrorInvalidValue);
			if(ANY) {
				if(spill) RDU_PERSIST_B(true, false, false, true, true) else RDU_PERSIST_B(false, false, false, true, true)
			} else if(alt) {
				if(spill) RDU_PERSIST_B(true, false, true, false, true) else RDU_PERSIST_B(false, false, true, false, true)
			} else {
				if(spill) RDU_PERSIST_B(true, false, false, false, true) else RDU_PERSIST_B(false, false, false, false, true)
			}
		} else if(alt) {
			if(stats) {
				if(spill) RDU_PERSIST(true, true, true, false) else RDU_PERSIST(false, true, true, false)
			} else {
				if(spill) RDU_PERSIST(true, false, true, false) else RDU_PERSIST(false, false, true, false)
			}
		} else if(nfa) {
			if(stats) {
				if(spill) RDU_PERSIST(true, true, false, true) else RDU_PERSIST(false, true, false, true)
			} else {
				if(spill) RDU_PERSIST(true, false, false, true) else RDU_PERSIST(false, false, false, true)
			}
		} else {
			if(stats) {
				if(spill) RDU_PERSIST(true, true, false, false) else RDU_PERSIST(false, true, false, false)
			} else {
				if(spill) RDU_PERSIST(true, false, false, false) else RDU_PERSIST(false, false, false, false)
			}
		}
#undef RDU_PERSIST
#undef RDU_PERSIST_I
#undef RDU_PERSIST_B
#undef RDU_PERSIST_X
		return int(hipGetLastError());
	}
	if(stats) {
		if(opt.referenceOrder) {
			RDU_BY_SPILL(true, false, true);
		} else if(wide) {
			RDU_BY_SPILL(false, true, true);
		} else {
			RDU_BY_SPILL(false, false, true);
		}
	} else {
		if(opt.referenceOrder) {
			RDU_BY_SPILL(true, false, false);
		} else if(wide) {
			RDU_BY_SPILL(false, true, false);
		} else {
			RDU_BY_SPILL(false, false, false);
		}
	}
#undef RDU_BY_SPILL
#undef RDU_LAUNCH
	return int(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
// Persistent near-first traversal of the 4-wide tree with lane refill.
//
// The one-ray-per-lane kernel above makes every wavefront wait for its slowest ray, and each
// loop trip can pay two dependent memory round trips (a node fetch, then a triangle fetch by
// other lanes). Here every lane is a small state machine that advances by exactly one step per
// trip — one wide node, one triangle, or one root box — so a trip issues at most one fetch per
// lane and waits once. A lane whose ray is finished keeps its result until the wavefront
// refills: once at least `refillMin` lanes are idle (or all are), the idle lanes write their
// results and take the next rays of the wavefront's sequence together, so ray set-up and result
// stores run for many lanes at a time instead of one lane per trip.
//
// Ray assignment follows the XCD layout: workgroups are dispatched round-robin over the `xcds`
// dies, so die k runs blocks k, k + xcds, ...; die k owns the k-th contiguous slice of the
// (optionally Morton-sorted) batch, and its W wavefronts take that slice in interleaved batches
// of 64 (wavefront v takes batches v, v + W, v + 2W, ...). At any moment a die's wavefronts
// work on one window of about 64 W neighbouring rays, which keeps the part of the tree they
// touch small enough for that die's own L2. Every wavefront exits when its sequence is done.

// Closest-hit queries also keep each pushed child's slab entry in LDS (fp16 rounded towards -inf,
// so never above the true entry): a popped child whose entry already exceeds the current maxi
// is dropped without being fetched, exactly as its own box test would drop it (the test fails
// once entry > maxi). Levels spilled to global memory carry no entry and are always fetched.
template<bool ANY, bool SPILL, bool STATS, bool ALT, bool NFA, bool BATCH, bool INST>
__global__ __launch_bounds__(kBlock, (persistentMinWaves<ANY, NFA, INST>())) void tracePersistent(DevBVH bvh, const float4 * __restrict__ rays, uint32_t n,
	const uint32_t * __restrict__ perm, const uint32_t * __restrict__ countDev, rdu_hit * __restrict__ hits,
	uint8_t * __restrict__ occ, int * __restrict__ gstack, int lds, uint32_t xcds, uint32_t refillMin, uint32_t chunk,
	const uint32_t * __restrict__ rayJob, const float4 * __restrict__ jobRoots, uint64_t * stats) {
	constexpr bool NF = (ALT && !ANY) || (ANY && NFA && !ALT); // near/far plane fetch
	constexpr bool UNSORTED = ALT && ANY; // children in node order
	__shared__ int ls[kWideLds][kBlock];
	__shared__ _Float16 le[ANY ? 1 : kWideLds][kBlock];
	// Instanced scenes: the lane's world ray (position, direction) while it is in an object's space.
	__shared__ float wr[INST ? 6 : 1][INST ? kBlock : 1];
	if(countDev)
		n = *countDev < n ? *countDev : n;
	const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
	const uint32_t die = blockIdx.x % xcds;
	const uint64_t W = uint64_t(gridDim.x / xcds) * (kBlock / 64);
	const uint64_t v = uint64_t(blockIdx.x / xcds) * (kBlock / 64) + (threadIdx.x >> 6);
	const uint64_t sliceBegin = uint64_t(n) * die / xcds;
	const uint64_t sliceEnd = uint64_t(n) * (die + 1) / xcds;
	uint64_t cursor = 0; // position in this wavefront's sequence (wave-uniform)
	LaneStack<kBlock> st{ls, gstack, size_t(gridDim.x) * kBlock, gtid, lds};
	uint64_t nodeVisits = 0, triTests = 0;

	bool has = false, pend = false;
	uint32_t ri = 0;
	uint32_t ox = 0, oy = 32, oz = 64;
	DevRay r = makeRay(mk(0, 0, 0), mk(0, 0, 1), 0.0f, 0.0f);
	bool clean = true;
	int c = 0, kk = 0;
	float maxi = 0.0f, bt = FLT_MAX, bu = 0.0f, bv = 0.0f;
	int best = -1;
	// Instanced scenes: the object whose shared tree the lane is in (-1 in world space, where the
	// triangles carry their object's id), and that of the best hit.
	int inst = -1, bestInst = -1;
	auto writeResult = [&]() {
		if(ANY) {
			occ[ri] = best >= 0 ? 1 : 0;
		} else {
			rdu_hit h;
			if(best >= 0) {
				const float4 t2 = bvh.tris[3 * best + 2];
				h.t = bt;
				h.u = bu;
				h.v = bv;
				h.hit = 1u;
				h.mesh_id = INST && bestInst >= 0 ? uint32_t(bestInst) : __float_as_uint(t2.y);
				h.local_id = __float_as_uint(t2.z);
			} else {
				h.t = FLT_MAX;
				h.u = 0.0f;
				h.v = 0.0f;
				h.hit = 0u;
				h.mesh_id = 0u;
				h.local_id = 0u;
			}
			hits[ri] = h;
		}
	};
	while(true) {
		// Refill idle lanes from this wavefront's sequence (wave-uniform cursor, no atomics).
		const uint64_t idle = __ballot(!has);
		const uint32_t nIdle = uint32_t(__popcll(idle));
		if(nIdle >= refillMin || nIdle == 64u) {
			if(!has) {
				if(pend) {
					writeResult();
					pend = false;
				}
				const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(idle >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(idle), 0u));
				const uint64_t j = cursor + rank;
				const uint64_t kpos = sliceBegin + ((j / chunk) * W + v) * chunk + (j % chunk);
				if(kpos < sliceEnd) {
					const uint32_t k = uint32_t(kpos);
					ri = perm ? perm[k] : k;
					const float4 ra = rays[2 * size_t(ri)];
					const float4 rb = rays[2 * size_t(ri) + 1];
					r = makeRay(xyz(ra), xyz(rb), ra.w, rb.w);
					clean = finite3(r.inv) && finite3(r.pos);
					// Byte offsets of each axis' near plane record within a node (x: 0/16, y: 32/48, z: 64/80).
					if constexpr(NF) {
						ox = r.inv.x < 0.0f ? 16u : 0u;
						oy = r.inv.y < 0.0f ? 48u : 32u;
						oz = r.inv.z < 0.0f ? 80u : 64u;
					}
					maxi = r.maxi;
					best = -1;
					inst = bestInst = -1;
					bt = FLT_MAX;
					bu = bv = 0.0f;
					st.sp = 0;
					kk = 0;
					// The top box of the 4-wide tree (select semantics: correct for any ray).
					float e;
					bool top;
					if constexpr(BATCH) {
						const uint32_t jb = rayJob[ri];
						const float4 lo = jobRoots[2 * size_t(jb)], hi = jobRoots[2 * size_t(jb) + 1];
						top = slab6<true>(r, lo.x, lo.y, lo.z, hi.x, hi.y, hi.z, r.mini, maxi, e);
						c = __float_as_int(lo.w);
					} else {
						top = slab6<true>(r, bvh.wideTopMin[0], bvh.wideTopMin[1], bvh.wideTopMin[2], bvh.wideTopMax[0],
							bvh.wideTopMax[1], bvh.wideTopMax[2], r.mini, maxi, e);
						c = bvh.wideTopCode;
					}
					if(top) {
						has = true;
					} else {
						pend = true; // a miss, written at the next refill
					}
				}
			}
			cursor += nIdle;
		}
		if(!__any(has)) {
			// Leave only when this wavefront's sequence is exhausted too (positions only grow).
			const uint64_t next = sliceBegin + ((cursor / chunk) * W + v) * chunk + (cursor % chunk);
			if(next >= sliceEnd)
				break;
			continue;
		}

		// Fetch: one record per lane.
		float4 a0 = make_float4(0, 0, 0, 0), a1 = a0, a2 = a0, a3 = a0, a4 = a0, a5 = a0, a6 = a0;
		if(has) {
			if(c >= 0) {
				const float4 * nd = bvh.wide + 8 * size_t(c);
				// Near planes in a0, a2, a4 and far planes in a1, a3, a5 (nodes are 128-byte aligned,
				// so the offsets combine with the address by OR and XOR).
				if constexpr(NF) {
					const uintptr_t b = reinterpret_cast<uintptr_t>(nd);
					a0 = ldGlobal(b | ox);
					a1 = ldGlobal(b | (ox ^ 16u));
					a2 = ldGlobal(b | oy);
					a3 = ldGlobal(b | (oy ^ 16u));
					a4 = ldGlobal(b | oz);
					a5 = ldGlobal(b | (oz ^ 16u));
				} else {
					a0 = nd[0];
					a1 = nd[1];
					a2 = nd[2];
					a3 = nd[3];
					a4 = nd[4];
					a5 = nd[5];
				}
				a6 = nd[6];
			} else if(INST && c < kInstLimit) {
				// An instance record (its inverse model rows, then {tree root code, object, world flag}).
				const float4 * ir = bvh.inst + 4 * size_t(c - kInstCode0);
				a0 = ir[0];
				a1 = ir[1];
				a2 = ir[2];
				a3 = ir[3];
			} else {
				const int ti = ((~c) >> 1) + kk;
				a0 = bvh.tris[3 * ti + 0];
				a1 = bvh.tris[3 * ti + 1];
				a2 = bvh.tris[3 * ti + 2];
			}
		}
		if(!has)
			continue;
		// Process.
		const bool exact = !__all(clean);
		bool next = false; // pop, or go to the next root, or finish
		if(INST && c < kInstLimit) {
			// Into an object's tree, with a marker pushed under its nodes (its entry never culls it)
			// whose pop brings the world ray back from LDS and leaves the object: then the top
			// level's entries resume. For an instance the ray moves to object space, p' = M^-1 p
			// and d' = M^-1 d (not renormalised, so t keeps its world value); an object with an
			// identity model takes the ray as it is.
			const f3 p = r.pos, d = r.dir;
			wr[0][threadIdx.x] = p.x;
			wr[1][threadIdx.x] = p.y;
			wr[2][threadIdx.x] = p.z;
			wr[3][threadIdx.x] = d.x;
			wr[4][threadIdx.x] = d.y;
			wr[5][threadIdx.x] = d.z;
			if(!__float_as_int(a3.z)) {
				r.pos = mk(a0.x * p.x + a0.y * p.y + a0.z * p.z + a0.w, a1.x * p.x + a1.y * p.y + a1.z * p.z + a1.w,
					a2.x * p.x + a2.y * p.y + a2.z * p.z + a2.w);
				r.dir = mk(a0.x * d.x + a0.y * d.y + a0.z * d.z, a1.x * d.x + a1.y * d.y + a1.z * d.z,
					a2.x * d.x + a2.y * d.y + a2.z * d.z);
				r.inv = rcp3(r.dir);
			}
			if(!ANY && (!SPILL || st.sp < st.lds))
				le[st.sp][threadIdx.x] = _Float16(-__builtin_inff());
			st.push<SPILL>(kInstReturn);
			inst = __float_as_int(a3.y);
			c = __float_as_int(a3.x);
			kk = 0;
			clean = finite3(r.inv) && finite3(r.pos);
			if constexpr(NF) {
				ox = r.inv.x < 0.0f ? 16u : 0u;
				oy = r.inv.y < 0.0f ? 48u : 32u;
				oz = r.inv.z < 0.0f ? 80u : 64u;
			}
		} else if(c >= 0) {
			if(STATS)
				++nodeVisits;
			int c0 = __float_as_int(a6.x), c1 = __float_as_int(a6.y), c2 = __float_as_int(a6.z), c3 = __float_as_int(a6.w);
			float e0, e1, e2 = 0.0f, e3 = 0.0f;
			bool h0, h1, h2, h3;
			if(exact) {
				// Back to (min, max) plane order for the reference's select semantics.
				if(NF && ox) {
					const float4 t = a0;
					a0 = a1;
					a1 = t;
				}
				if(NF && oy != 32u) {
					const float4 t = a2;
					a2 = a3;
					a3 = t;
				}
				if(NF && oz != 64u) {
					const float4 t = a4;
					a4 = a5;
					a5 = t;
				}
				h0 = slab6<true>(r, a0.x, a2.x, a4.x, a1.x, a3.x, a5.x, r.mini, maxi, e0);
				h1 = slab6<true>(r, a0.y, a2.y, a4.y, a1.y, a3.y, a5.y, r.mini, maxi, e1);
				h2 = (c2 != kWideEmpty) && slab6<true>(r, a0.z, a2.z, a4.z, a1.z, a3.z, a5.z, r.mini, maxi, e2);
				h3 = (c3 != kWideEmpty) && slab6<true>(r, a0.w, a2.w, a4.w, a1.w, a3.w, a5.w, r.mini, maxi, e3);
			} else {
				if constexpr(NF) {
					h0 = slabNearFar(r, a0.x, a2.x, a4.x, a1.x, a3.x, a5.x, r.mini, maxi, e0);
					h1 = slabNearFar(r, a0.y, a2.y, a4.y, a1.y, a3.y, a5.y, r.mini, maxi, e1);
					h2 = (c2 != kWideEmpty) && slabNearFar(r, a0.z, a2.z, a4.z, a1.z, a3.z, a5.z, r.mini, maxi, e2);
					h3 = (c3 != kWideEmpty) && slabNearFar(r, a0.w, a2.w, a4.w, a1.w, a3.w, a5.w, r.mini, maxi, e3);
				} else {
					h0 = slab6<false>(r, a0.x, a2.x, a4.x, a1.x, a3.x, a5.x, r.mini, maxi, e0);
					h1 = slab6<false>(r, a0.y, a2.y, a4.y, a1.y, a3.y, a5.y, r.mini, maxi, e1);
					h2 = (c2 != kWideEmpty) && slab6<false>(r, a0.z, a2.z, a4.z, a1.z, a3.z, a5.z, r.mini, maxi, e2);
					h3 = (c3 != kWideEmpty) && slab6<false>(r, a0.w, a2.w, a4.w, a1.w, a3.w, a5.w, r.mini, maxi, e3);
				}
			}
			const int nh = int(h0) + int(h1) + int(h2) + int(h3);
			if(nh == 0) {
				next = true;
			} else if constexpr(UNSORTED) {
				// Any hit: the answer does not depend on the visiting order, so no sorting: go on
				// with the first child that passed and push the others (cheaper per node; near-first
				// finds a hit sooner on incoherent rays, this pays on coherent ones).
				const int first = h0 ? 0 : (h1 ? 1 : (h2 ? 2 : 3));
				if(h1 && first != 1)
					st.push<SPILL>(c1);
				if(h2 && first != 2)
					st.push<SPILL>(c2);
				if(h3 && first != 3)
					st.push<SPILL>(c3);
				c = first == 0 ? c0 : (first == 1 ? c1 : (first == 2 ? c2 : c3));
				kk = 0;
			} else {
				// A child that passed its test has a finite entry (mini <= entry <= maxi); the others
				// sort last.
				e0 = h0 ? e0 : __builtin_inff();
				e1 = h1 ? e1 : __builtin_inff();
				e2 = h2 ? e2 : __builtin_inff();
				e3 = h3 ? e3 : __builtin_inff();
				cswap(e0, c0, e1, c1);
				cswap(e2, c2, e3, c3);
				cswap(e0, c0, e2, c2);
				cswap(e1, c1, e3, c3);
				cswap(e1, c1, e2, c2);
				auto pushE = [&](int code, float e) {
					if(!ANY && (!SPILL || st.sp < st.lds))
						le[st.sp][threadIdx.x] = roundDownHalf(e);
					st.push<SPILL>(code);
				};
				if(nh > 3)
					pushE(c3, e3);
				if(nh > 2)
					pushE(c2, e2);
				if(nh > 1)
					pushE(c1, e1);
				c = c0;
				kk = 0;
			}
		} else {
			if(STATS)
				++triTests;
			const int code = ~c;
			const int ti = (code >> 1) + kk;
			float tt, uu, vv;
			if(mollerTrumbore(r, xyz(a0), mk(a0.w, a1.x, a1.y), mk(a1.z, a1.w, a2.x), r.mini, maxi, tt, uu, vv)) {
				best = ti;
				bt = tt;
				bu = uu;
				bv = vv;
				maxi = tt;
				if(INST)
					bestInst = inst;
			}
			if(ANY && best >= 0) {
				st.sp = 0;
				next = true;
			} else if(kk < (code & 1)) {
				++kk;
			} else {
				next = true;
			}
		}
		if(next && !ANY) {
			// Drop stacked children whose entry is already beyond the closest hit.
			while(st.sp > 0 && st.sp <= st.lds && float(le[st.sp - 1][threadIdx.x]) > maxi)
				--st.sp;
		}
		if(next) {
			bool more = st.sp > 0;
			if(more) {
				c = st.pop<SPILL>();
				kk = 0;
			}
			if(INST && more && c == kInstReturn) {
				// Out of an object's space: the world ray again (its direction already normalised,
				// so the reciprocals are the ones the query started with), then on with the top
				// level's entries.
				r.pos = mk(wr[0][threadIdx.x], wr[1][threadIdx.x], wr[2][threadIdx.x]);
				r.dir = mk(wr[3][threadIdx.x], wr[4][threadIdx.x], wr[5][threadIdx.x]);
				r.inv = rcp3(r.dir);
				inst = -1;
				clean = finite3(r.inv) && finite3(r.pos);
				if constexpr(NF) {
					ox = r.inv.x < 0.0f ? 16u : 0u;
					oy = r.inv.y < 0.0f ? 48u : 32u;
					oz = r.inv.z < 0.0f ? 80u : 64u;
				}
				if(!ANY)
					while(st.sp > 0 && st.sp <= st.lds && float(le[st.sp - 1][threadIdx.x]) > maxi)
						--st.sp;
				more = st.sp > 0;
				if(more)
					c = st.pop<SPILL>();
			}
			if(!more) {
				// Done: the result is written at the ray's own index at the next refill.
				has = false;
				pend = true;
			}
		}
	}
	if(pend)
		writeResult();
	if(STATS) {
		const uint64_t a = waveSum(nodeVisits);
		const uint64_t b = waveSum(triTests);
		if((threadIdx.x & 63) == 0) {
			atomicAdd(reinterpret_cast<unsigned long long *>(stats), (unsigned long long)a);
			atomicAdd(reinterpret_cast<unsigned long long *>(stats + 1), (unsigned long long)b);
		}
	}
}

template<bool ANY, bool SPILL, bool STATS, bool ALT, bool NFA, bool BATCH, bool INST>
PersistentGrid persistentGrid(size_t n, uint32_t blocksPerCuCap) {
	struct Occupancy {
		int blocksPerCu = 1, cus = 1;
	};
	// Queried once per kernel variant (thread-safe initialisation of a function-local static).
	static const Occupancy occ = [] {
		Occupancy o;
		int dev = 0;
		(void)hipGetDevice(&dev);
		(void)hipDeviceGetAttribute(&o.cus, hipDeviceAttributeMultiprocessorCount, dev);
		(void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&o.blocksPerCu, tracePersistent<ANY, SPILL, STATS, ALT, NFA, BATCH, INST>,
			kBlock, 0);
		o.blocksPerCu = o.blocksPerCu < 1 ? 1 : o.blocksPerCu;
		o.cus = o.cus < 1 ? 1 : o.cus;
		return o;
	}();
	const int blocksPerCu = (blocksPerCuCap > 0 && int(blocksPerCuCap) < occ.blocksPerCu) ? int(blocksPerCuCap) : occ.blocksPerCu;
	const int cus = occ.cus;
	PersistentGrid g;
	// One resident grid (optionally fewer workgroups per CU than fit), trimmed for small batches
	// (about one 64-ray batch per wavefront), always a whole number of blocks per die.
	size_t blocks = size_t(blocksPerCu) * size_t(cus);
	const size_t needed = (n + kBlock - 1) / kBlock;
	blocks = needed < blocks ? needed : blocks;
	blocks = ((blocks + kXcds - 1) / kXcds) * kXcds;
	g.blocks = unsigned(blocks);
	return g;
}

// Morton code of the origin's cell in the scene box, kOrderBits per axis.
__device__ __forceinline__ uint32_t spread3(uint32_t x) {
	x &= 0x3ff;
	x = (x | (x << 16)) & 0x030000ff;
	x = (x | (x << 8)) & 0x0300f00f;
	x = (x | (x << 4)) & 0x030c30c3;
	x = (x | (x << 2)) & 0x09249249;
	return x;
}

__global__ __launch_bounds__(256) void rayKeysKernel(const float4 * __restrict__ rays, size_t n, float3 lo, float3 scale,
	uint32_t * __restrict__ keys, uint32_t * __restrict__ idx) {
	const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
	if(i >= n)
		return;
	const float4 o = rays[2 * i];
	const float cells = float(1u << kOrderBits);
	auto q = [cells](float v) {
		// NaN and out-of-box origins clamp to the border cells.
		const float c = __builtin_fminf(__builtin_fmaxf(v, 0.0f), cells - 1.0f);
		return uint32_t(c);
	};
	const uint32_t x = q((o.x - lo.x) * scale.x), y = q((o.y - lo.y) * scale.y), z = q((o.z - lo.z) * scale.z);
	keys[i] = (spread3(x) << 2) | (spread3(y) << 1) | spread3(z);
	idx[i] = uint32_t(i);
}

} // namespace

int ldsStackDepth() {
	// RDU_DEBUG_LDS_STACK=k (1 <= k <= 32) keeps only k stack levels in LDS so that tests can
	// exercise the global-memory spill path on ordinary trees.
	const char * v = std::getenv("RDU_DEBUG_LDS_STACK");
	const int k = v ? std::atoi(v) : kLdsStack;
	return (k >= 1 && k <= kLdsStack) ? k : kLdsStack;
}

int persistentOccupancy(bool any) {
	int b = 0;
	const hipError_t e = any ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, tracePersistent<true, false, false, false>, kBlock, 0)
							 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, tracePersistent<false, false, false, false>, kBlock, 0);
	return (e != hipSuccess || b < 1) ? 1 : b;
}

int stackNeed(const DevBVH & bvh, const TraceOptions & opt) {
	const bool wide = opt.wide && !opt.referenceOrder && bvh.wide != nullptr;
	return wide ? 3 * int(bvh.wideDepth) + 2 : int(bvh.maxDepth) + 2;
}

size_t globalStackInts(const DevBVH & bvh, const TraceOptions & opt, size_t n) {
	const bool wide = opt.wide && !opt.referenceOrder && bvh.wide != nullptr;
	const int lds = wide ? (ldsStackDepth() < kWideLds ? ldsStackDepth() : kWideLds) : ldsStackDepth();
	const int need = stackNeed(bvh, opt) - lds;
	if(need <= 0)
		return 0;
	// One slot per lane of the widest grid a launch of n rays uses: the persistent grid rounds
	// its block count up to a whole number per XCD, so it can exceed ceil(n / kBlock); the
	// one-ray-per-lane grid, ceil(n / kSimpleBlock) blocks of kSimpleBlock, never exceeds that.
	const size_t blocks = (n + kBlock - 1) / kBlock;
	const size_t slots = ((blocks + kXcds - 1) / kXcds) * kXcds * kBlock;
	return size_t(need) * slots;
}

int launchClosest(const DevBVH & bvh, const float4 * rays, size_t n, void * hits, const TraceOptions & opt, int * gstack,
	uint64_t * stats, hipStream_t stream) {
	return launchImpl<false>(bvh, rays, n, static_cast<rdu_hit *>(hits), nullptr, opt, gstack, stats, stream);
}

int launchAny(const DevBVH & bvh, const float4 * rays, size_t n, uint8_t * occluded, const TraceOptions & opt,
	int * gstack, uint64_t * stats, hipStream_t stream) {
	return launchImpl<true>(bvh, rays, n, nullptr, occluded, opt, gstack, stats, stream);
}

size_t sortTempBytes(size_t n) {
	size_t bytes = 0;
	(void)rocprim::radix_sort_pairs(nullptr, bytes, static_cast<uint32_t *>(nullptr), static_cast<uint32_t *>(nullptr),
		static_cast<uint32_t *>(nullptr), static_cast<uint32_t *>(nullptr), n, 0, 3 * kOrderBits);
	return bytes;
}

int launchRayOrder(const DevBVH & bvh, const float4 * rays, size_t n, uint32_t * keys, uint32_t * keysAlt,
	uint32_t * perm, uint32_t * permAlt, void * temp, size_t tempBytes, hipStream_t stream) {
	if(n == 0)
		return 0;
	const float cells = float(1u << kOrderBits);
	float3 lo, scale;
	float ext[3];
	for(int a = 0; a < 3; ++a) {
		ext[a] = bvh.sceneMax[a] - bvh.sceneMin[a];
		ext[a] = ext[a] > 0.0f ? ext[a] : 1.0f;
	}
	lo = make_float3(bvh.sceneMin[0], bvh.sceneMin[1], bvh.sceneMin[2]);
	scale = make_float3(cells / ext[0], cells / ext[1], cells / ext[2]);
	hipLaunchKernelGGL(rayKeysKernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream, rays, n, lo, scale, keysAlt,
		permAlt);
	hipError_t e = hipGetLastError();
	if(e != hipSuccess)
		return int(e);
	size_t bytes = tempBytes;
	e = rocprim::radix_sort_pairs(temp, bytes, keysAlt, keys, permAlt, perm, n, 0, 3 * kOrderBits, stream);
	return int(e);
}

} // namespace rdu
