// C ABI of the raycaster (include/rdu.h): the drop-in boundary for Rendu's Raycaster
// (src/engine/raycaster/Raycaster.hpp:10-157). Host-side bookkeeping only; all queries run on
// the GPU through trace.hip. There is no CPU query path: without a HIP device the queries fail
// with RDU_ERR_HIP.
#include "raycaster_impl.hpp"

#include <cfloat>
#include <cstdlib>
#include <new>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace rdu;

namespace {

int checkBuilt(const rdu_raycaster * rc) {
	if(!rc)
		return fail(RDU_ERR_INVALID, "null raycaster");
	if(!rc->scene.built && !rc->gpuBuilt)
		return fail(RDU_ERR_STATE, "query before updateHierarchy");
	if(!rc->uploaded)
		return fail(RDU_ERR_HIP, "the hierarchy is not resident on a GPU (no HIP device or upload failed)");
	return RDU_OK;
}

int ensureGlobalStack(const rdu_raycaster * rc, const TraceOptions & opt, size_t n, hipStream_t s) {
	const size_t ints = globalStackInts(rc->dev, opt, n);
	if(ints == 0)
		return RDU_OK;
	return rc->gstack.reserveOn(ints * sizeof(int), s);
}

// Traversal choice for a batch. RDU_TRACE_LAYOUT=binary selects the binary near-first kernels
// (comparison runs); the 4-wide tree is the default for near-first queries.
TraceOptions traceOptions(uint32_t flags) {
	TraceOptions o;
	o.referenceOrder = (flags & RDU_REFERENCE_ORDER) != 0;
	const char * lay = std::getenv("RDU_TRACE_LAYOUT");
	o.wide = !(lay && std::strcmp(lay, "binary") == 0);
	return o;
}

// Per-ray queries follow Rendu's visiting order on the host-built tree, near-first on a GPU-built one.
uint32_t perRayFlags(const rdu_raycaster * rc) { return (rc && rc->gpuBuilt) ? 0u : RDU_REFERENCE_ORDER; }

// Incoherent batches below this size are traced in input order: the sort does not pay off.
constexpr size_t kReorderMin = size_t(1) << 16;

TuneKey tuneKey(const rdu_raycaster * rc, size_t n, uint32_t flags, bool any) {
	const uint32_t cls = ((flags & RDU_INCOHERENT) ? kTuneIncoherent : 0u) | (any ? kTuneAny : 0u) |
						 (rc->gpuBuilt ? kTuneGpuTree : 0u);
	return makeTuneKey(rc->scene.tris.size(), rc->dev.wideDepth, rc->scene.meshCount, n, cls);
}

// Launch parameters of a batch: this raycaster's own tuning for the class, else the process-wide
// cache's entry for the scene statistics, else the defaults. source: 0 default, 1 raycaster, 2 cache.
TraceTuning tuningFor(const rdu_raycaster * rc, size_t n, uint32_t flags, bool any, uint32_t * source) {
	const int c = (flags & RDU_INCOHERENT) ? 1 : 0, a = any ? 1 : 0;
	TraceTuning t;
	uint32_t src = 0;
	if(rc->tunedSet[c][a]) {
		t = rc->tuned[c][a];
		src = 1;
	} else if(tuneCacheLookup(tuneKey(rc, n, flags, any), t)) {
		src = 2;
	}
	if(source)
		*source = src;
	return t;
}

template<bool ANY>
int query(const rdu_raycaster * rc, const rdu_ray * rays, size_t n, void * out, uint32_t flags, void * stream,
	uint64_t * statsOut) {
	int st = checkBuilt(rc);
	if(st)
		return st;
	if(n == 0)
		return RDU_OK;
	if(!rays || !out)
		return fail(RDU_ERR_INVALID, "null rays or output");
	if(n > (size_t(1) << 31))
		return fail(RDU_ERR_UNSUPPORTED, "batch larger than 2^31 rays");
	TraceOptions opt = traceOptions(flags);
	if(rc->gpuBuilt && (opt.referenceOrder || !opt.wide))
		return fail(RDU_ERR_STATE, "a GPU-built tree has no reference visiting order or binary layout");
	applyTuning(opt, tuningFor(rc, n, flags, ANY, nullptr));
	opt.nearFar = (flags & RDU_INCOHERENT) != 0;
	opt.unorderedAny = (flags & RDU_INCOHERENT) == 0;
	const size_t outBytes = ANY ? n : n * sizeof(rdu_hit);
	std::lock_guard<std::mutex> guard(rc->lock);
	if((st = setDevice(rc->device)))
		return st;
	hipStream_t s = static_cast<hipStream_t>(stream);
	// Order this batch after the previous batch's use of the scratch buffers (on whatever stream
	// it ran), and record the end of this batch's use on `s` before returning.
	if(!rc->scratchDone &&
		(st = hipCheck(hipEventCreateWithFlags(&rc->scratchDone, hipEventDisableTiming), "hipEventCreate")))
		return st;
	if((st = hipCheck(hipStreamWaitEvent(s, rc->scratchDone, 0), "hipStreamWaitEvent")))
		return st;
	struct MarkDone {
		const rdu_raycaster * rc;
		hipStream_t s;
		~MarkDone() { (void)hipEventRecord(rc->scratchDone, s); }
	} markDone{rc, s};
	if((st = ensureGlobalStack(rc, opt, n, s)))
		return st;
	uint64_t * dstats = nullptr;
	if(statsOut) {
		if((st = rc->stats.reserveOn(2 * sizeof(uint64_t), s)))
			return st;
		dstats = static_cast<uint64_t *>(rc->stats.ptr);
		if((st = hipCheck(hipMemsetAsync(dstats, 0, 2 * sizeof(uint64_t), s), "hipMemsetAsync")))
			return st;
	}
	const float4 * dRays;
	void * dOut;
	const bool device = (flags & RDU_DEVICE_POINTERS) != 0;
	if(device) {
		dRays = reinterpret_cast<const float4 *>(rays);
		dOut = out;
	} else {
		if((st = rc->stageIn.reserveOn(n * sizeof(rdu_ray), s)) || (st = rc->stageOut.reserveOn(outBytes, s)))
			return st;
		if((st = hipCheck(hipMemcpyAsync(rc->stageIn.ptr, rays, n * sizeof(rdu_ray), hipMemcpyHostToDevice, s),
				"hipMemcpyAsync H2D")))
			return st;
		dRays = static_cast<const float4 *>(rc->stageIn.ptr);
		dOut = rc->stageOut.ptr;
	}
	if((flags & RDU_INCOHERENT) && !opt.referenceOrder && n >= kReorderMin) {
		const size_t tb = sortTempBytes(n);
		if((st = rc->keys.reserveOn(n * 4, s)) || (st = rc->keysAlt.reserveOn(n * 4, s)) || (st = rc->perm.reserveOn(n * 4, s)) ||
			(st = rc->permAlt.reserveOn(n * 4, s)) || (st = rc->sortTemp.reserveOn(tb, s)))
			return st;
		if((st = hipCheck(hipError_t(launchRayOrder(rc->dev, dRays, n, static_cast<uint32_t *>(rc->keys.ptr),
								  static_cast<uint32_t *>(rc->keysAlt.ptr), static_cast<uint32_t *>(rc->perm.ptr),
								  static_cast<uint32_t *>(rc->permAlt.ptr), rc->sortTemp.ptr, tb, s)),
				"ray ordering")))
			return st;
		opt.perm = static_cast<const uint32_t *>(rc->perm.ptr);
		opt.spatialOrder = true;
	}
	hipEvent_t tStart = nullptr, tStop = nullptr;
	if(rc->timing && rc->timedLaunches < rdu_raycaster::kMaxTimed) {
		if(rc->timingEvents.size() < 2 * (rc->timedLaunches + 1)) {
			hipEvent_t a, b;
			if((st = hipCheck(hipEventCreate(&a), "hipEventCreate")))
				return st;
			if((st = hipCheck(hipEventCreate(&b), "hipEventCreate"))) {
				(void)hipEventDestroy(a);
				return st;
			}
			rc->timingEvents.push_back(a);
			rc->timingEvents.push_back(b);
		}
		tStart = rc->timingEvents[2 * rc->timedLaunches];
		tStop = rc->timingEvents[2 * rc->timedLaunches + 1];
		if((st = hipCheck(hipEventRecord(tStart, s), "hipEventRecord")))
			return st;
	}
	const int e = ANY ? launchAny(rc->dev, dRays, n, static_cast<uint8_t *>(dOut), opt, static_cast<int *>(rc->gstack.ptr),
							  dstats, s)
					  : launchClosest(rc->dev, dRays, n, dOut, opt, static_cast<int *>(rc->gstack.ptr), dstats, s);
	if((st = hipCheck(hipError_t(e), "trace kernel launch")))
		return st;
	if(tStop) {
		if((st = hipCheck(hipEventRecord(tStop, s), "hipEventRecord")))
			return st;
		rc->timedLaunches += 1;
	}
	if(!device) {
		if((st = hipCheck(hipMemcpyAsync(out, dOut, outBytes, hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H")))
			return st;
	}
	if(statsOut) {
		if((st = hipCheck(hipMemcpyAsync(statsOut, dstats, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, s), "stats copy")))
			return st;
	}
	if(!device || statsOut) {
		if((st = hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize")))
			return st;
	}
	return RDU_OK;
}

// Per-thread resources of the GPU per-ray entry points, one set per device: a stream, a pinned
// host-mapped record (the ray, then the result) that the kernel reads and writes in place, and a
// spill stack. Created on a thread's first GPU per-ray call on a device and released when the
// thread exits, so that per-ray calls from many host threads need no lock and no copies and run
// concurrently on the GPU.
struct PerRayCtx {
	int device = 0;
	hipStream_t stream = nullptr;
	unsigned char * host = nullptr; // rdu_ray at 0, rdu_hit at 32, occlusion byte at 56
	unsigned char * dev = nullptr;  // device address of `host`
	DevBuffer gstack;               // in `stream`'s order
	~PerRayCtx() {
		if(!stream)
			return;
		int current = -1;
		(void)hipGetDevice(&current);
		(void)hipSetDevice(device);
		(void)hipStreamSynchronize(stream);
		gstack.releaseOn(stream);
		(void)hipStreamSynchronize(stream);
		(void)hipStreamDestroy(stream);
		if(host)
			(void)hipHostFree(host);
		if(current >= 0)
			(void)hipSetDevice(current);
	}
};
constexpr int kMaxPerRayDevices = 16;
struct PerRayThread {
	PerRayCtx * ctx[kMaxPerRayDevices] = {};
	~PerRayThread() {
		for(PerRayCtx * c : ctx)
			delete c;
	}
};
thread_local PerRayThread tlsPerRay;

int perRayCtx(int device, PerRayCtx ** out) {
	if(device < 0 || device >= kMaxPerRayDevices)
		return fail(RDU_ERR_UNSUPPORTED, "per-ray queries support devices 0..15");
	PerRayCtx *& slot = tlsPerRay.ctx[device];
	if(!slot) {
		PerRayCtx * c = new(std::nothrow) PerRayCtx();
		if(!c)
			return fail(RDU_ERR_ALLOC, "host allocation failed");
		c->device = device;
		int st;
		if((st = hipCheck(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate")) ||
			(st = hipCheck(hipHostMalloc(reinterpret_cast<void **>(&c->host), 64, hipHostMallocMapped), "hipHostMalloc")) ||
			(st = hipCheck(hipHostGetDevicePointer(reinterpret_cast<void **>(&c->dev), c->host, 0), "hipHostGetDevicePointer"))) {
			delete c;
			return st;
		}
		slot = c;
	}
	*out = slot;
	return RDU_OK;
}

// One ray, synchronous, with no raycaster lock: on a host-built tree the host walk of the same
// flattened tree (host_query.cpp), else (GPU-built tree, or per-ray device mode) a one-ray launch on
// the calling thread's own stream and pinned record.
template<bool ANY>
int queryOne(const rdu_raycaster * rc, const rdu_ray & ray, rdu_hit * hit, uint8_t * occ) {
	int st = checkBuilt(rc);
	if(st)
		return st;
	if(!rc->gpuBuilt && !rc->perRayDevice) {
		if(ANY)
			*occ = hostIntersectAny(rc, ray) ? 1 : 0;
		else
			hostIntersect(rc, ray, hit);
		return RDU_OK;
	}
	TraceOptions opt = traceOptions(perRayFlags(rc));
	if((st = setDevice(rc->device)))
		return st;
	PerRayCtx * c = nullptr;
	if((st = perRayCtx(rc->device, &c)))
		return st;
	const size_t ints = globalStackInts(rc->dev, opt, 1);
	if(ints && (st = c->gstack.reserveOn(ints * sizeof(int), c->stream)))
		return st;
	std::memcpy(c->host, &ray, sizeof(rdu_ray));
	const float4 * dRay = reinterpret_cast<const float4 *>(c->dev);
	int * gs = static_cast<int *>(c->gstack.ptr);
	const int e = ANY ? launchAny(rc->dev, dRay, 1, c->dev + 56, opt, gs, nullptr, c->stream)
					  : launchClosest(rc->dev, dRay, 1, c->dev + 32, opt, gs, nullptr, c->stream);
	if((st = hipCheck(hipError_t(e), "trace kernel launch")) ||
		(st = hipCheck(hipStreamSynchronize(c->stream), "hipStreamSynchronize")))
		return st;
	if(ANY)
		*occ = c->host[56];
	else
		std::memcpy(hit, c->host + 32, sizeof(rdu_hit));
	return RDU_OK;
}

// Scratch of the allocation-free batch entry (rdu_raycaster_intersect_noalloc): a file-scope
// device array, one copy per device, held by every device the library runs on until the process
// ends. Host code never passes `scratch` itself to a kernel (on the host the name stands for a
// host-side placeholder, not a device address): each device's address comes from
// hipGetSymbolAddress, looked up once per device with that device current, and is what the
// kernels receive. Every call on a device shares its array, so calls that overlap on one device
// race; nothing in it needs to start at zero (every kernel writes its part before reading it).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kNoAllocDevices = 64;
std::mutex noAllocMutex;
unsigned char * noAllocScratch[kNoAllocDevices] = {};

// The calling thread's current device must be `device`.
int noAllocScratchFor(int device, unsigned char ** out) {
	if(device < 0 || device >= kNoAllocDevices)
		return fail(RDU_ERR_UNSUPPORTED, "the allocation-free entry supports devices 0..63");
	std::lock_guard<std::mutex> g(noAllocMutex);
	if(!noAllocScratch[device]) {
		void * p = nullptr;
		const int st = hipCheck(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
		if(st)
			return st;
		noAllocScratch[device] = static_cast<unsigned char *>(p);
	}
	*out = noAllocScratch[device];
	return RDU_OK;
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

} // namespace

extern "C" {

int rdu_raycaster_intersect_noalloc(const rdu_raycaster * rc, const rdu_ray * rays, size_t n, rdu_hit * hits,
	uint32_t flags, void * stream) {
	clearError();
	int st = checkBuilt(rc);
	if(st)
		return st;
	if(!(flags & RDU_DEVICE_POINTERS))
		return fail(RDU_ERR_INVALID, "the allocation-free entry takes device pointers only (RDU_DEVICE_POINTERS)");
	if(n == 0)
		return RDU_OK;
	if(!rays || !hits)
		return fail(RDU_ERR_INVALID, "null rays or output");
	if(n > (size_t(1) << 31))
		return fail(RDU_ERR_UNSUPPORTED, "batch larger than 2^31 rays");
	TraceOptions opt = traceOptions(flags);
	if(rc->gpuBuilt && (opt.referenceOrder || !opt.wide))
		return fail(RDU_ERR_STATE, "a GPU-built tree has no reference visiting order or binary layout");
	applyTuning(opt, tuningFor(rc, n, flags, false, nullptr));
	opt.nearFar = (flags & RDU_INCOHERENT) != 0;
	opt.unorderedAny = (flags & RDU_INCOHERENT) == 0;
	// The same work as rdu_raycaster_intersect, with every scratch buffer carved out of `scratch`:
	// the spill stack, then (sorted incoherent batches) keys, keys', order, order' and the sort's
	// temporary storage, each 256-byte aligned.
	const bool order = (flags & RDU_INCOHERENT) && !opt.referenceOrder && n >= kReorderMin;
	const size_t stackBytes = align256(globalStackInts(rc->dev, opt, n) * sizeof(int));
	const size_t keyBytes = align256(n * 4);
	const size_t tempBytes = order ? sortTempBytes(n) : 0;
	const size_t need = stackBytes + (order ? 4 * keyBytes + align256(tempBytes) : 0);
	if(need > kNoAllocScratchBytes)
		return fail(RDU_ERR_UNSUPPORTED, "batch needs " + std::to_string(need) + " bytes of scratch; the allocation-free entry holds " +
											  std::to_string(kNoAllocScratchBytes) + " (use rdu_raycaster_intersect, or smaller batches)");
	if((st = setDevice(rc->device)))
		return st;
	unsigned char * base = nullptr;
	if((st = noAllocScratchFor(rc->device, &base)))
		return st;
	hipStream_t s = static_cast<hipStream_t>(stream);
	const float4 * dRays = reinterpret_cast<const float4 *>(rays);
	int * gstack = stackBytes ? reinterpret_cast<int *>(base) : nullptr;
	if(order) {
		unsigned char * k = base + stackBytes;
		uint32_t * keys = reinterpret_cast<uint32_t *>(k);
		uint32_t * keysAlt = reinterpret_cast<uint32_t *>(k + keyBytes);
		uint32_t * perm = reinterpret_cast<uint32_t *>(k + 2 * keyBytes);
		uint32_t * permAlt = reinterpret_cast<uint32_t *>(k + 3 * keyBytes);
		void * temp = k + 4 * keyBytes;
		if((st = hipCheck(hipError_t(launchRayOrder(rc->dev, dRays, n, keys, keysAlt, perm, permAlt, temp, tempBytes, s)),
				"ray ordering")))
			return st;
		opt.perm = perm;
		opt.spatialOrder = true;
	}
	return hipCheck(hipError_t(launchClosest(rc->dev, dRays, n, hits, opt, gstack, nullptr, s)), "trace kernel launch");
}

int rdu_raycaster_create(rdu_raycaster ** out) {
	clearError();
	if(!out)
		return fail(RDU_ERR_INVALID, "null output handle");
	*out = new(std::nothrow) rdu_raycaster();
	if(!*out)
		return fail(RDU_ERR_ALLOC, "host allocation failed");
	return RDU_OK;
}

void rdu_raycaster_destroy(rdu_raycaster * rc) {
	if(!rc)
		return;
	if(rc->device >= 0)
		setDevice(rc->device);
	delete rc;
}

int rdu_raycaster_add_mesh(rdu_raycaster * rc, const float * xyz, size_t nv, const uint32_t * idx, size_t ni,
	const float * model) {
	clearError();
	if(!rc)
		return fail(RDU_ERR_INVALID, "null raycaster");
	if(rc->scene.built)
		return fail(RDU_ERR_STATE, "addMesh after updateHierarchy: mesh roots must precede all child nodes");
	if(ni % 3 != 0)
		return fail(RDU_ERR_INVALID, "index count is not a multiple of 3");
	if((nv && !xyz) || (ni && !idx))
		return fail(RDU_ERR_INVALID, "null vertex or index pointer");
	for(size_t i = 0; i < ni; ++i)
		if(idx[i] >= nv)
			return fail(RDU_ERR_INVALID, "index out of range of the mesh vertices");
	if(rc->scene.vertices.size() + nv >= (size_t(1) << 32) || rc->scene.tris.size() + ni / 3 >= (size_t(1) << 30))
		return fail(RDU_ERR_UNSUPPORTED, "more than 2^30 triangles or 2^32 vertices");
	rc->scene.addMesh(xyz, nv, idx, ni, model);
	return RDU_OK;
}

int rdu_raycaster_update_hierarchy(rdu_raycaster * rc) {
	clearError();
	if(!rc)
		return fail(RDU_ERR_INVALID, "null raycaster");
	if(rc->scene.built)
		return fail(RDU_ERR_STATE, "updateHierarchy called twice");
	for(size_t m = 0; m < rc->scene.meshCount; ++m)
		if(rc->scene.nodes[m].right == 0)
			return fail(RDU_ERR_INVALID, "mesh with no triangle (Rendu reads past its triangle list here)");
	rc->scene.build();
	rc->scene.flatten(rc->layout);
	int st = currentDevice(&rc->device);
	if(st)
		return st;
	if((st = rc->nodes.upload(rc->layout.nodes.data(), rc->layout.nodes.size() * 4)) ||
		(st = rc->tris.upload(rc->layout.tris.data(), rc->layout.tris.size() * 4)) ||
		(st = rc->roots.upload(rc->layout.roots.data(), rc->layout.roots.size() * 4)) ||
		(st = rc->wide.upload(rc->layout.wide.data(), rc->layout.wide.size() * 4)))
		return st;
	rc->dev.nodes = static_cast<const float4 *>(rc->nodes.ptr);
	rc->dev.tris = static_cast<const float4 *>(rc->tris.ptr);
	rc->dev.roots = static_cast<const float4 *>(rc->roots.ptr);
	rc->dev.wide = static_cast<const float4 *>(rc->wide.ptr);
	rc->dev.nRoots = rc->scene.meshCount;
	rc->dev.maxDepth = rc->layout.maxDepth;
	rc->dev.wideDepth = rc->layout.wideDepth;
	const f3 lo = rc->layout.sceneMin, hi = rc->layout.sceneMax;
	const float l3[3] = {lo.x, lo.y, lo.z}, h3[3] = {hi.x, hi.y, hi.z};
	std::memcpy(rc->dev.sceneMin, l3, sizeof(l3));
	std::memcpy(rc->dev.sceneMax, h3, sizeof(h3));
	const HBox & tb = rc->layout.wideTopBox;
	const float tl[3] = {tb.mn.x, tb.mn.y, tb.mn.z}, th[3] = {tb.mx.x, tb.mx.y, tb.mx.z};
	std::memcpy(rc->dev.wideTopMin, tl, sizeof(tl));
	std::memcpy(rc->dev.wideTopMax, th, sizeof(th));
	rc->dev.wideTopCode = rc->layout.wideTopCode;
	rc->uploaded = true;
	return RDU_OK;
}

int rdu_raycaster_update_hierarchy_gpu(rdu_raycaster * rc, void * stream) {
	clearError();
	if(!rc)
		return fail(RDU_ERR_INVALID, "null raycaster");
	if(rc->scene.built || rc->gpuBuilt)
		return fail(RDU_ERR_STATE, "updateHierarchy called twice");
	if(rc->scene.tris.empty())
		return fail(RDU_ERR_INVALID, "no triangles");
	for(size_t m = 0; m < rc->scene.meshCount; ++m)
		if(rc->scene.nodes[m].right == 0)
			return fail(RDU_ERR_INVALID, "mesh with no triangle (Rendu reads past its triangle list here)");
	int st = currentDevice(&rc->device);
	if(st)
		return st;
	hipStream_t s = static_cast<hipStream_t>(stream);
	if((st = gpuBuildBvh(rc->scene, rc->gpu, s)))
		return st;
	if(!rc->gpu.wide.ptr && (st = rc->gpu.wide.reserve(128))) // one-triangle scene: a placeholder node
		return st;
	rc->dev = DevBVH{};
	rc->dev.tris = static_cast<const float4 *>(rc->gpu.tris.ptr);
	rc->dev.wide = static_cast<const float4 *>(rc->gpu.wide.ptr);
	rc->dev.wideDepth = rc->gpu.wideDepth;
	std::memcpy(rc->dev.sceneMin, rc->gpu.sceneMin, sizeof(rc->gpu.sceneMin));
	std::memcpy(rc->dev.sceneMax, rc->gpu.sceneMax, sizeof(rc->gpu.sceneMax));
	std::memcpy(rc->dev.wideTopMin, rc->gpu.topMin, sizeof(rc->gpu.topMin));
	std::memcpy(rc->dev.wideTopMax, rc->gpu.topMax, sizeof(rc->gpu.topMax));
	rc->dev.wideTopCode = rc->gpu.topCode;
	if(rc->gpu.triangles == 1) {
		std::memcpy(rc->dev.wideTopMin, rc->gpu.sceneMin, sizeof(rc->gpu.sceneMin));
		std::memcpy(rc->dev.wideTopMax, rc->gpu.sceneMax, sizeof(rc->gpu.sceneMax));
	}
	rc->gpuBuilt = true;
	rc->uploaded = true;
	return RDU_OK;
}

int rdu_raycaster_intersect(const rdu_raycaster * rc, const rdu_ray * rays, size_t n, rdu_hit * hits, uint32_t flags,
	void * stream) {
	clearError();
	return query<false>(rc, rays, n, hits, flags, stream, nullptr);
}

int rdu_raycaster_intersect_any(const rdu_raycaster * rc, const rdu_ray * rays, size_t n, uint8_t * occluded,
	uint32_t flags, void * stream) {
	clearError();
	return query<true>(rc, rays, n, occluded, flags, stream, nullptr);
}

int rdu_raycaster_intersect1(const rdu_raycaster * rc, const float o[3], const float d[3], float mini, float maxi,
	rdu_hit * out) {
	clearError();
	if(!o || !d || !out)
		return fail(RDU_ERR_INVALID, "null argument");
	const rdu_ray r = {o[0], o[1], o[2], mini, d[0], d[1], d[2], maxi};
	return queryOne<false>(rc, r, out, nullptr);
}

int rdu_raycaster_intersect_any1(const rdu_raycaster * rc, const float o[3], const float d[3], float mini, float maxi,
	int * occluded) {
	clearError();
	if(!o || !d || !occluded)
		return fail(RDU_ERR_INVALID, "null argument");
	const rdu_ray r = {o[0], o[1], o[2], mini, d[0], d[1], d[2], maxi};
	uint8_t occ = 0;
	const int st = queryOne<true>(rc, r, nullptr, &occ);
	*occluded = occ;
	return st;
}

int rdu_raycaster_set_per_ray_device(rdu_raycaster * rc, int on_device) {
	clearError();
	if(!rc)
		return fail(RDU_ERR_INVALID, "null raycaster");
	rc->perRayDevice = on_device != 0;
	return RDU_OK;
}

int rdu_raycaster_visible(const rdu_raycaster * rc, const float p0[3], const float p1[3], int * visible) {
	clearError();
	if(!p0 || !p1 || !visible)
		return fail(RDU_ERR_INVALID, "null argument");
	// Raycaster.cpp:226-230
	const f3 dir = mk(p1[0], p1[1], p1[2]) - mk(p0[0], p0[1], p0[2]);
	const float maxi = length(dir);
	const float d[3] = {dir.x, dir.y, dir.z};
	int occ = 0;
	const int st = rdu_raycaster_intersect_any1(rc, p0, d, 0.0001f, maxi, &occ);
	*visible = occ ? 0 : 1;
	return st;
}

int rdu_raycaster_info(const rdu_raycaster * rc, rdu_bvh_info * out) {
	clearError();
	if(!rc || !out)
		return fail(RDU_ERR_INVALID, "null argument");
	std::memset(out, 0, sizeof(*out));
	out->meshes = rc->scene.meshCount;
	out->triangles = rc->scene.tris.size();
	out->vertices = rc->scene.vertices.size();
	if(rc->gpuBuilt) {
		// One binary radix tree over all triangles (one per leaf); depth is not tracked.
		const uint64_t n = rc->gpu.triangles;
		out->nodes = 2 * n - 1;
		out->internal_nodes = rc->gpu.innerCount;
		out->wide_nodes = rc->gpu.wideCount;
		out->wide_depth = rc->gpu.wideDepth;
		out->wide_bytes = uint64_t(rc->gpu.wideCount) * 128 + n * 48;
		out->device_bytes = out->wide_bytes;
		return RDU_OK;
	}
	out->nodes = rc->scene.nodes.size();
	out->internal_nodes = rc->layout.internalCount;
	out->max_depth = rc->layout.maxDepth;
	const uint64_t tr = (rc->layout.tris.size() + rc->layout.roots.size()) * 4;
	out->binary_bytes = rc->layout.nodes.size() * 4 + tr;
	out->wide_bytes = rc->layout.wide.size() * 4 + tr;
	out->device_bytes = out->binary_bytes + rc->layout.wide.size() * 4;
	out->wide_nodes = rc->layout.wideCount;
	out->wide_depth = rc->layout.wideDepth;
	return RDU_OK;
}

int rdu_raycaster_export_nodes(const rdu_raycaster * rc, float * boxes, uint64_t * left, uint64_t * right, uint8_t * leaf) {
	clearError();
	if(!rc || !boxes || !left || !right || !leaf)
		return fail(RDU_ERR_INVALID, "null argument");
	if(!rc->scene.built)
		return fail(RDU_ERR_STATE, "export before updateHierarchy (or on a GPU-built tree: no Rendu node array)");
	for(size_t i = 0; i < rc->scene.nodes.size(); ++i) {
		const HNode & nd = rc->scene.nodes[i];
		const float b[6] = {nd.box.mn.x, nd.box.mn.y, nd.box.mn.z, nd.box.mx.x, nd.box.mx.y, nd.box.mx.z};
		std::memcpy(boxes + 6 * i, b, sizeof(b));
		left[i] = nd.left;
		right[i] = nd.right;
		leaf[i] = nd.leaf ? 1 : 0;
	}
	return RDU_OK;
}

int rdu_raycaster_export_triangles(const rdu_raycaster * rc, uint32_t * mesh_id, uint32_t * local_id) {
	clearError();
	if(!rc || !mesh_id || !local_id)
		return fail(RDU_ERR_INVALID, "null argument");
	if(!rc->scene.built)
		return fail(RDU_ERR_STATE, "export before updateHierarchy (or on a GPU-built tree: no Rendu triangle order)");
	for(size_t i = 0; i < rc->scene.tris.size(); ++i) {
		mesh_id[i] = rc->scene.tris[i].meshId;
		local_id[i] = rc->scene.tris[i].localId;
	}
	return RDU_OK;
}

int rdu_raycaster_autotune(rdu_raycaster * rc, const rdu_ray * rays, size_t n, int any_hit, uint32_t flags, void * stream,
	rdu_tune_report * report) {
	clearError();
	int st = checkBuilt(rc);
	if(st)
		return st;
	if(!rays || n == 0)
		return fail(RDU_ERR_INVALID, "autotune needs a non-empty batch");
	if(flags & RDU_REFERENCE_ORDER)
		return fail(RDU_ERR_INVALID, "the reference visiting order has nothing to tune");
	const bool device = (flags & RDU_DEVICE_POINTERS) != 0;
	hipStream_t s = static_cast<hipStream_t>(stream);
	// Device copies of the batch so that every candidate is timed on the GPU work alone.
	DevBuffer inBuf, outBuf;
	const rdu_ray * dRays = rays;
	if(!device) {
		if((st = inBuf.upload(rays, n * sizeof(rdu_ray))))
			return st;
		dRays = static_cast<const rdu_ray *>(inBuf.ptr);
	}
	if((st = outBuf.reserve(any_hit ? n : n * sizeof(rdu_hit))))
		return st;
	const uint32_t qflags = (flags & ~RDU_DEVICE_POINTERS) | RDU_DEVICE_POINTERS;
	const int c = (flags & RDU_INCOHERENT) ? 1 : 0, a = any_hit ? 1 : 0;
	hipEvent_t e0, e1;
	if((st = hipCheck(hipEventCreate(&e0), "hipEventCreate")))
		return st;
	if((st = hipCheck(hipEventCreate(&e1), "hipEventCreate"))) {
		(void)hipEventDestroy(e0);
		return st;
	}
	// One candidate: a warm-up query, then the mean of two timed ones.
	auto timeOne = [&](const TraceTuning & t, float & ms) -> int {
		rc->tuned[c][a] = t;
		rc->tunedSet[c][a] = true;
		ms = 0.0f;
		int e = RDU_OK;
		for(int rep = 0; rep < 3 && !e; ++rep) {
			if(rep > 0 && (e = hipCheck(hipEventRecord(e0, s), "hipEventRecord")))
				break;
			e = any_hit ? query<true>(rc, dRays, n, outBuf.ptr, qflags, s, nullptr)
						: query<false>(rc, dRays, n, outBuf.ptr, qflags, s, nullptr);
			if(e || rep == 0)
				continue;
			float t1 = 0.0f;
			if(!(e = hipCheck(hipEventRecord(e1, s), "hipEventRecord")) &&
				!(e = hipCheck(hipEventSynchronize(e1), "hipEventSynchronize")) &&
				!(e = hipCheck(hipEventElapsedTime(&t1, e0, e1), "hipEventElapsedTime")))
				ms += 0.5f * t1;
		}
		return e;
	};
	TuneResult res;
	st = tuneSearch(timeOne, persistentOccupancy(any_hit != 0), res);
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	rc->tuned[c][a] = st ? TraceTuning{} : res.best;
	rc->tunedSet[c][a] = st == RDU_OK;
	if(st)
		return st;
	tuneCacheStore(tuneKey(rc, n, qflags, any_hit != 0), res.best);
	if(report) {
		report->refill_min = res.best.refillMin;
		report->best_ms = res.bestMs;
		report->default_ms = res.defaultMs;
		report->candidates = res.timed;
		report->chunk = res.best.chunk;
		report->blocks_per_cu = res.best.blocksPerCu;
		report->source = 1;
	}
	return RDU_OK;
}

int rdu_raycaster_tuning(const rdu_raycaster * rc, size_t n, int any_hit, uint32_t flags, rdu_tune_report * out) {
	clearError();
	if(!rc || !out)
		return fail(RDU_ERR_INVALID, "null raycaster or report");
	if(!rc->scene.built && !rc->gpuBuilt)
		return fail(RDU_ERR_STATE, "tuning before updateHierarchy");
	uint32_t src = 0;
	const TraceTuning t = tuningFor(rc, n, flags, any_hit != 0, &src);
	*out = rdu_tune_report{};
	out->refill_min = t.refillMin;
	out->chunk = t.chunk;
	out->blocks_per_cu = t.blocksPerCu;
	out->source = src;
	return RDU_OK;
}

int rdu_raycaster_set_tuning(rdu_raycaster * rc, int any_hit, uint32_t flags, const rdu_tune_report * params) {
	clearError();
	if(!rc || !params)
		return fail(RDU_ERR_INVALID, "null raycaster or parameters");
	TraceTuning t;
	t.refillMin = params->refill_min;
	t.chunk = params->chunk;
	t.blocksPerCu = params->blocks_per_cu;
	if(!tuningValid(t))
		return fail(RDU_ERR_INVALID, "launch parameters out of range");
	const int c = (flags & RDU_INCOHERENT) ? 1 : 0, a = any_hit ? 1 : 0;
	rc->tuned[c][a] = t;
	rc->tunedSet[c][a] = true;
	return RDU_OK;
}

int rdu_tune_cache_save(const char * path) {
	clearError();
	return tuneCacheSave(path);
}

int rdu_tune_cache_load(const char * path) {
	clearError();
	return tuneCacheLoad(path);
}

int rdu_tune_cache_clear(void) {
	clearError();
	tuneCacheClear();
	return RDU_OK;
}

int rdu_tune_cache_size(size_t * entries) {
	clearError();
	if(!entries)
		return fail(RDU_ERR_INVALID, "null output");
	*entries = tuneCacheSize();
	return RDU_OK;
}

int rdu_raycaster_set_kernel_timing(rdu_raycaster * rc, int enable) {
	clearError();
	if(!rc)
		return fail(RDU_ERR_INVALID, "null raycaster");
	std::lock_guard<std::mutex> guard(rc->lock);
	rc->timing = enable != 0;
	rc->timedLaunches = 0;
	return RDU_OK;
}

int rdu_raycaster_kernel_time(rdu_raycaster * rc, double * total_ms, uint64_t * launches) {
	clearError();
	if(!rc || !total_ms || !launches)
		return fail(RDU_ERR_INVALID, "null argument");
	std::lock_guard<std::mutex> guard(rc->lock);
	int st = setDevice(rc->device);
	if(st)
		return st;
	double sum = 0.0;
	for(size_t i = 0; i < rc->timedLaunches; ++i) {
		float ms = 0.0f;
		if((st = hipCheck(hipEventSynchronize(rc->timingEvents[2 * i + 1]), "hipEventSynchronize")) ||
			(st = hipCheck(hipEventElapsedTime(&ms, rc->timingEvents[2 * i], rc->timingEvents[2 * i + 1]),
				 "hipEventElapsedTime")))
			return st;
		sum += ms;
	}
	*total_ms = sum;
	*launches = rc->timedLaunches;
	rc->timedLaunches = 0;
	return RDU_OK;
}

int rdu_raycaster_trace_stats(const rdu_raycaster * rc, const rdu_ray * rays, size_t n, int any_hit, uint32_t flags,
	uint64_t * node_visits, uint64_t * tri_tests) {
	clearError();
	if(!node_visits || !tri_tests)
		return fail(RDU_ERR_INVALID, "null argument");
	if(flags & RDU_DEVICE_POINTERS)
		return fail(RDU_ERR_INVALID, "trace_stats takes host rays");
	uint64_t s[2] = {0, 0};
	std::vector<uint8_t> out(any_hit ? n : n * sizeof(rdu_hit));
	const int st = any_hit ? query<true>(rc, rays, n, out.data(), flags, nullptr, s)
						   : query<false>(rc, rays, n, out.data(), flags, nullptr, s);
	*node_visits = s[0];
	*tri_tests = s[1];
	return st;
}

} // extern "C"

int rdu::buildInstanceBase(rdu_raycaster * rc, const std::vector<uint32_t> & objectOf) {
	if(rc->scene.built || rc->gpuBuilt)
		return fail(RDU_ERR_STATE, "updateHierarchy called twice");
	for(size_t m = 0; m < rc->scene.meshCount; ++m)
		if(rc->scene.nodes[m].right == 0)
			return fail(RDU_ERR_INVALID, "mesh with no triangle (Rendu reads past its triangle list here)");
	if(rc->scene.tris.size() >= (size_t(1) << 29))
		return fail(RDU_ERR_UNSUPPORTED, "instanced scene with 2^29 distinct triangles or more");
	rc->scene.build();
	rc->scene.flatten(rc->layout, true);
	// Triangle record {e2.z, mesh id, local id, 0}: the mesh id is float 9 of the 12.
	std::vector<float> & T = rc->layout.tris;
	for(size_t t = 0; t < T.size() / 12; ++t) {
		uint32_t m;
		std::memcpy(&m, &T[12 * t + 9], 4);
		if(m < objectOf.size() && objectOf[m] != UINT32_MAX)
			std::memcpy(&T[12 * t + 9], &objectOf[m], 4);
	}
	int st = currentDevice(&rc->device);
	if(st)
		return st;
	if((st = rc->tris.upload(rc->layout.tris.data(), rc->layout.tris.size() * 4)) ||
		(st = rc->wide.upload(rc->layout.wide.data(), rc->layout.wide.size() * 4)))
		return st;
	rc->dev = DevBVH{};
	rc->dev.tris = static_cast<const float4 *>(rc->tris.ptr);
	rc->dev.wide = static_cast<const float4 *>(rc->wide.ptr);
	rc->dev.nRoots = rc->scene.meshCount;
	rc->dev.wideDepth = rc->layout.wideDepth;
	rc->uploaded = true;
	return RDU_OK;
}
