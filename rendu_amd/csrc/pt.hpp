// Host-visible interface of the wavefront path tracer kernels (pt.hip).
#pragma once
#include "pt_device.hpp"
#include "rcache.hpp"

namespace rdu {

// Per-render constants: image size, stratification grid and camera frame
// (Camera::pixelShifts, src/engine/input/Camera.cpp:42-49). The random stream key is a device
// counter (kCountSeed).
struct PtFrame {
	uint32_t width = 0, height = 0;
	int32_t cellX = 1;
	float cellW = 1.0f, cellH = 1.0f;
	f3 eye{0, 0, 0}, corner{0, 0, 0}, dx{0, 0, 0}, dy{0, 0, 0};
};

// Device counters (uint32 each). kCountBase is the first sample index of the running pass and
// kCountBaseNext that of the next one: each pass reads its base from the device and advances it,
// so a captured pass replays as the next pass with no host-side arguments changing. kCountSeed is
// the render's random-stream key, read by the kernels from here for the same reason.
enum : int {
	kCountActive = 0,
	kCountNext = 1,
	kCountShadow = 2,
	kCountShadowNext = 3,
	kCountBase = 4,
	kCountBaseNext = 5,
	kCountSeed = 6,
	kCountOverflow = 7, // batch passes: queue entries left for the second shading launch of a bounce
	kCountSlots = 8
};

// Path state, structure of arrays indexed by path id p = sample_in_pass * pixels + pixel.
struct PtPaths {
	float4 * rays = nullptr;        // 2 float4 per path: rdu_ray
	rdu_hit * hits = nullptr;
	float4 * atten = nullptr;       // throughput
	float4 * color = nullptr;       // accumulated sample colour
	uint32_t * queue[2] = {nullptr, nullptr};
	uint32_t * counts = nullptr;    // kCountSlots
	unsigned long long * totals = nullptr; // [extension rays, shadow rays]
	float4 * shadowRays = nullptr;  // 2 float4 per path
	float4 * shadowContrib = nullptr;
	uint32_t * shadowQueue[2] = {nullptr, nullptr};
	uint8_t * occluded = nullptr;
	rdu_hit * shadowHits = nullptr;
	float2 * mis = nullptr;         // extension: BSDF density of the last scattering direction (0: no
	                                // MIS weight) and the distance travelled since, through cut-outs
};

struct PtOptions {
	uint32_t depth = 5;
	int32_t estimator = RDU_ESTIMATOR_REFERENCE;
	int32_t rrStart = 3;
	int32_t maxShadowLayers = 16;
};

struct PtScratch {
	int * gstack = nullptr; // traversal stack spill, globalStackInts() for the pass size
	uint32_t * walkLeft = nullptr; // pinned host word: shadow rays still walking (scenes with cut-outs)
};

// Radiance-cache mode of a pass (an extension; plain tracing when absent). Paths of the pass's
// training pixels ((x + trainX) % 4 == 0 and (y + trainY) % 4 == 0: one pixel in 16) are traced in
// full and record every surface vertex from bounce `terminateBounce` on; at the end of the pass each
// record becomes a (query, target radiance) pair, target = (final colour - colour on arrival) /
// throughput on arrival. Every other path ends at its first surface vertex from that bounce on where
// the cache is trusted, adding throughput x cache(position, outgoing direction).
// Records with index % 8 == 7 are held out of training; their predictions against their targets
// (luminance sums per cell of a fine 32^3 and a coarse 8^3 grid, scaled by kCacheDecay every pass)
// decide trust: a cell is trusted with at least `minCount` records and |sum(prediction) -
// sum(target)| <= maxBias * max(sums) + 1e-3 count; a vertex is trusted when its fine cell is, or
// when its fine cell has too few records and its coarse cell is.
constexpr int kCacheCellRes = kRcCellRes;     // fine cells per axis over the cache's box
constexpr int kCacheCoarseRes = kRcCoarseRes; // coarse cells per axis
constexpr int kCacheFineCells = kRcCellRes * kRcCellRes * kRcCellRes;
constexpr float kCacheDecay = 0.9f;
enum : int { kCacheQueries = 0, kCacheRecords = 1, kCacheTrain = 2, kCacheHold = 3, kCacheSlots = 4 };
struct PtCache {
	int32_t terminateBounce = 2;
	uint32_t trainX = 0, trainY = 0;
	float maxBias = 0.2f;
	float minCount = 8.0f;
	uint32_t alwaysTrust = 0;
	uint32_t records = 1; // 0 (frozen cache, no trust test): no training pixels, no records
	RcGrid grid{};
	const RcParams * params = nullptr;
	uint32_t * counts = nullptr;       // kCacheSlots
	float4 * query = nullptr;          // terminated paths: 2 float4 (position, outgoing direction) per slot
	float4 * queryAtten = nullptr;     // throughput per slot
	uint32_t * queryPath = nullptr;    // path per slot
	float4 * pred = nullptr;           // cache output per slot (also reused for held-out predictions)
	float4 * rec = nullptr;            // records: position, outgoing direction, colour on arrival, throughput
	uint32_t * recPath = nullptr;
	uint32_t recCapacity = 0;
	float4 * trainQ = nullptr, * trainT = nullptr; // compact training pairs (2 float4 + 1 float4)
	float4 * holdQ = nullptr, * holdT = nullptr;   // held-out pairs
	float4 * cells = nullptr;          // kRcCells x {sum prediction, sum target, count, unused}: fine, then coarse
};

// One pass of samplesInPass samples per pixel, starting at the sample index the device holds in
// kCountBaseNext, with the random key in kCountSeed (both set with ptSetSampleBase before the
// first pass). Traversal launches use `tune`.
// Enqueues only kernels (no copies, fills or host synchronisation), so it can be captured in a graph,
// except in scenes with alpha cut-outs: their shadow walk reads the number of rays still walking
// into scratch.walkLeft between rounds of steps (a host synchronisation), so it runs eagerly.
int ptRenderPass(const DevScene & S, const DevBVH & bvh, const PtFrame & f, const PtPaths & P, uint32_t samplesInPass,
	const PtOptions & opt, float * dImage, uint32_t components, PtScratch & scratch, const PtCache * cache,
	const TraceTuning & tune, hipStream_t stream,
	hipEvent_t accumulateAfter = nullptr);
int ptSetSampleBase(uint32_t * counts, uint32_t base, uint32_t seed, hipStream_t stream);

// Batch passes: the paths of several jobs (scene, camera, image, samples, depth) in one set of
// launches. A pass holds chunks: some consecutive samples of one job, whose paths are the id range
// [pathBegin, pathBegin + samples x pixels) laid out like a single render's pass (sample-major).
// Each job's tree lives in one node arena shared by the pass (codes rebased), so the traversal
// starts every ray at its job's root. Paths record their chunk and job; the shading kernels read a
// wavefront's job descriptor with scalar loads (once per distinct job among its lanes).
struct PtJob {
	DevScene S;
	PtFrame f;
	PtOptions opt;
	float * image = nullptr; // device accumulation image of the job (pixels x components)
	uint32_t components = 3;
};
struct PtChunk {
	uint32_t job = 0;
	uint32_t pathBegin = 0;  // first path id of the chunk in the pass
	uint32_t samples = 0;    // samples of the chunk
	uint32_t sampleBase = 0; // index of its first sample in the job's render
	uint32_t seed = 0;       // the job's random-stream key
	uint32_t slotBegin = 0;  // first accumulation slot of the chunk (running sum of pixels over chunks)
};
struct PtBatch {
	const PtJob * jobs = nullptr;     // device table
	const PtChunk * chunks = nullptr; // device table of the pass
	uint32_t nChunks = 0;
	uint32_t nSlots = 0;              // pixels summed over the chunks
	uint32_t * pathChunk = nullptr;   // per path
	uint32_t * pathJob = nullptr;     // per path (the traversal's root index)
	uint32_t * overflow = nullptr;    // per path: queue entries whose wavefront served another chunk
	const float4 * roots = nullptr;   // per job: {top min, root code bits}, {top max, 0}
};
// One batch pass of nPaths paths over the arena tree; maxDepth is the largest job depth;
// shadows: some job samples lights; masked: some job has cut-outs (then every job's shadow rays
// take the cut-out walk, which gives the any-hit answer on scenes without cut-outs).
// accumulateAfter (or null): an event the pass's accumulation into the job images waits for.
int ptRenderBatchPass(const PtBatch & B, const DevBVH & arena, const PtPaths & P, uint32_t nPaths, uint32_t maxDepth,
	bool shadows, bool masked, int maxShadowLayers, PtScratch & scratch, const TraceTuning & tune, hipStream_t stream,
	hipEvent_t accumulateAfter);
// Copies a tree's 4-wide nodes into an arena at node offset nodeOffset with its triangles at
// triangle offset triOffset, rebasing the child codes.
int ptRebaseWide(float4 * dst, const float4 * src, uint32_t nodes, int32_t nodeOffset, int32_t triOffset,
	hipStream_t stream);
int ptNormalizeImage(float * dImage, uint32_t pixels, uint32_t components, uint32_t spp, hipStream_t stream);

} // namespace rdu
