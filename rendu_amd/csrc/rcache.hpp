// Learned radiance cache: a multiresolution hash-grid encoding of the position (16 levels x 2^19
// entries x 2 features) and a degree-3 spherical-harmonics encoding of the outgoing direction (16
// features), followed by a 5 x 64 ReLU MLP (48 -> 64 -> 64 -> 64 -> 64 -> 64 -> 3), evaluated and
// trained with bf16 MFMA (v_mfma_f32_16x16x16_bf16) accumulating in fp32. Tables are kept in fp32
// (master, Adam) with an fp16 copy for the forward gathers; MLP weights in fp32 (master, Adam) with a
// bf16 copy for the MFMA operands.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace rdu {

constexpr int kRcLevels = 16;
constexpr int kRcFeatures = 2;
constexpr int kRcLog2Table = 19;
constexpr uint32_t kRcTable = 1u << kRcLog2Table;
constexpr int kRcEnc = kRcLevels * kRcFeatures; // 32 hash-grid features
constexpr int kRcDir = 16;                      // spherical harmonics, degrees 0..3
constexpr int kRcIn = kRcEnc + kRcDir;          // 48
constexpr int kRcWidth = 64;
constexpr int kRcHidden = 5;  // hidden layers of width 64
constexpr int kRcOut = 3;
constexpr int kRcOutPad = 16; // output rows padded to one MFMA tile
// Parameter counts: W0 64x48, W1..W4 64x64, W5 16x64 (rows >= 3 are zero and never updated).
constexpr int kRcW0 = kRcWidth * kRcIn;
constexpr int kRcWh = kRcWidth * kRcWidth;
constexpr int kRcW5 = kRcOutPad * kRcWidth;
constexpr int kRcMlpParams = kRcW0 + 4 * kRcWh + kRcW5; // 20480 (19648 used + 832 padding)
constexpr size_t kRcTableParams = size_t(kRcLevels) * kRcTable * kRcFeatures;
// Trust cells of the path tracer's cache mode: a fine 32^3 grid and a coarse 8^3 grid over the box.
constexpr int kRcCellRes = 32;
constexpr int kRcCoarseRes = 8;
constexpr int kRcCells = kRcCellRes * kRcCellRes * kRcCellRes + kRcCoarseRes * kRcCoarseRes * kRcCoarseRes;

struct RcGrid {
	float lo[3], inv[3]; // position normalisation to [0, 1]^3
	int32_t res[kRcLevels];
	int32_t dense[kRcLevels]; // 1: (res+1)^3 <= table size, direct indexing
};

struct RcParams {
	float * w32 = nullptr;      // MLP master weights, kRcMlpParams
	uint16_t * w16 = nullptr;   // bf16 copy
	float * gw = nullptr;       // MLP gradient accumulators
	float * mw = nullptr, * vw = nullptr;
	float * t32 = nullptr;      // tables master, kRcTableParams
	uint16_t * t16 = nullptr;   // fp16 copy
	float * gt = nullptr;       // table gradient accumulators (rcGradients only)
	float * mt = nullptr, * vt = nullptr;
};

// Per-sample activations and output gradients of a training batch (bf16, feature-major
// [feature][sample]), used by the weight-gradient GEMMs.
// Two feature gradients of one table entry.
struct RcGradPair {
	float a, b;
};

struct RcTrainBuffers {
	uint16_t * x = nullptr; // X0 (48 rows) then X1..X5 (64 rows each): 368 rows
	uint16_t * g = nullptr; // G0..G4 (64 rows each, pre-activation gradients) then GY (16 rows): 336 rows
	float * loss = nullptr; // per-sample loss (for monitoring)
	uint32_t capacity = 0;
	// Table-gradient records, one per (sample, level, corner): key = level * kRcTable + entry and
	// the two weighted feature gradients. They are sorted by key and summed per key (in a fixed
	// order, so the table gradients are run-to-run deterministic) instead of being added with atomics.
	uint32_t * recKey = nullptr, * recKeyAlt = nullptr, * uniqKey = nullptr, * uniqCount = nullptr;
	RcGradPair * recVal = nullptr, * recValAlt = nullptr, * uniqVal = nullptr;
	void * temp = nullptr;
	size_t tempBytes = 0;
};
constexpr uint32_t kRcRecordsPerSample = kRcLevels * 8;
// Scratch bytes the sort and the reduction need for `records` records.
size_t rcReduceTempBytes(uint32_t records);

struct RcAdam {
	// lr 3e-3: one-sample path-tracing targets are sparse and noisy; 1e-2 overshoots on them.
	float lr = 3e-3f, beta1 = 0.9f, beta2 = 0.99f, eps = 1e-15f;
	float l2 = 1e-6f; // weight decay on MLP weights only
	uint32_t step = 0;
};

// Queries: 2 float4 per query, {position xyz, unused} then {outgoing direction xyz, unused} (the
// rdu_ray layout); out: n float4 (rgb radiance). A device count, when given, bounds n.
int rcInfer(const RcGrid & grid, const RcParams & p, const float4 * queries, const uint32_t * countDev, uint32_t n,
	float4 * out, hipStream_t stream);
// One optimisation step on n (query, target radiance) pairs: dense Adam over the MLP weights (with
// the l2 term), lazy Adam over the table entries the batch touched (untouched entries keep their
// weights and moments). p.gt is neither read nor written. Returns hipError_t.
int rcTrainStep(const RcGrid & grid, RcParams & p, RcTrainBuffers & tb, RcAdam & adam, const float4 * queries,
	const float4 * targets, uint32_t n, hipStream_t stream);
// Gradients only (no update), for checks: fills p.gw and p.gt for the batch.
int rcGradients(const RcGrid & grid, RcParams & p, RcTrainBuffers & tb, const float4 * queries, const float4 * targets,
	uint32_t n, hipStream_t stream);
int rcSyncCopies(RcParams & p, hipStream_t stream); // refresh bf16/fp16 copies from the masters

} // namespace rdu

// Internal access to a cache handle for the path tracer's cache mode (capi_rcache.cpp).
struct rdu_rcache;
namespace rdu {
const RcGrid * rcacheGrid(const rdu_rcache * c);
RcParams * rcacheParams(rdu_rcache * c);
RcTrainBuffers * rcacheTrain(rdu_rcache * c, uint32_t n); // grown to n pairs; nullptr on failure
RcAdam * rcacheAdam(rdu_rcache * c);
float4 * rcacheCells(rdu_rcache * c); // trust statistics of the coarse cells (persist with the cache)
int rcacheDevice(const rdu_rcache * c);
} // namespace rdu
