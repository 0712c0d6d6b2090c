// The raycaster handle behind include/rdu.h, shared by the C ABI translation units.
#pragma once
#include "bvh.hpp"
#include "capi_common.hpp"
#include "gpubuild.hpp"
#include "trace.hpp"
#include "tuning.hpp"
#include "../../include/rdu.h"

#include <mutex>
#include <vector>

struct rdu_raycaster;

namespace rdu {
// capi_raycaster.cpp: Rendu's host build of every added mesh with each mesh's root collapsed into
// a 4-wide tree of its own and no level above (the trees an instanced scene's objects use); uploads
// triangles and nodes. The triangles of mesh m carry objectOf[m] as their mesh id where that is
// not UINT32_MAX (a mesh in world space, used by that one object). The handle then answers no
// queries of its own.
int buildInstanceBase(rdu_raycaster * rc, const std::vector<uint32_t> & objectOf);
// host_query.cpp: one ray in Rendu's visiting order over the host copy of the flattened tree.
void hostIntersect(const rdu_raycaster * rc, const rdu_ray & ray, rdu_hit * out);
bool hostIntersectAny(const rdu_raycaster * rc, const rdu_ray & ray);
} // namespace rdu

struct rdu_raycaster {
	rdu::HostScene scene;
	rdu::DeviceLayout layout;
	rdu::DevBVH dev;
	rdu::DevBuffer nodes, tris, roots, wide;
	bool uploaded = false;
	int device = -1;
	// GPU-built tree (rdu_raycaster_update_hierarchy_gpu): near-first queries only.
	rdu::GpuBvh gpu;
	bool gpuBuilt = false;
	// Per-ray calls (intersect1, intersect_any1, visible) on a host-built tree walk the host copy
	// of the flattened tree (host_query.cpp) unless this is set (rdu_raycaster_set_per_ray_device):
	// then each call is a one-ray GPU launch. GPU-built trees always take the GPU.
	bool perRayDevice = false;
	// Staging for host-pointer queries and global stack spill; guarded by `lock`. Batches use
	// them on their own streams: each batch's stream first waits for `scratchDone`, recorded on
	// the previous batch's stream after its last use of them, so batches on different streams
	// never overlap on this memory, and the buffers grow and shrink in that stream order.
	mutable std::mutex lock;
	mutable rdu::DevBuffer stageIn, stageOut, gstack, stats;
	mutable rdu::DevBuffer keys, keysAlt, perm, permAlt, sortTemp; // incoherent-batch ordering
	mutable hipEvent_t scratchDone = nullptr;
	// Autotuned launch parameters per [incoherent][any-hit] class; classes this raycaster has not
	// tuned take the process-wide tuning cache's entry for the scene's statistics, if any.
	rdu::TraceTuning tuned[2][2];
	bool tunedSet[2][2] = {{false, false}, {false, false}};
	// Optional timing of the traversal kernel alone: event pairs recorded on the query's stream
	// around each traversal launch, read back and summed by rdu_raycaster_kernel_time.
	static constexpr size_t kMaxTimed = 4096;
	bool timing = false;
	mutable std::vector<hipEvent_t> timingEvents; // 2 per launch
	mutable size_t timedLaunches = 0;
	~rdu_raycaster() {
		// Batches still in flight use the scratch buffers and the tree until `scratchDone`;
		// renders and per-ray queries have finished (their callers wait for them). After that
		// every block is idle and is freed without waiting for the rest of the device.
		if(scratchDone) {
			(void)hipEventSynchronize(scratchDone);
			(void)hipEventDestroy(scratchDone);
		}
		for(rdu::DevBuffer * b : {&stageIn, &stageOut, &gstack, &stats, &keys, &keysAlt, &perm, &permAlt, &sortTemp,
				 &nodes, &tris, &roots, &wide, &gpu.tris, &gpu.wide})
			b->releaseOn(nullptr);
		for(hipEvent_t e : timingEvents)
			(void)hipEventDestroy(e);
	}
};
