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
	// Staging for host-pointer queries and global stack spill; guarded by `lock`.
	mutable std::mutex lock;
	mutable rdu::DevBuffer stageIn, stageOut, gstack, stats;
	mutable rdu::DevBuffer keys, keysAlt, perm, permAlt, sortTemp; // incoherent-batch ordering
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
		for(hipEvent_t e : timingEvents)
			(void)hipEventDestroy(e);
	}
};
