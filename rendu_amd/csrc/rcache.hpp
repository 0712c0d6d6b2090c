// Learned radiance cache: a multiresolution hash-grid encoding (16 levels x 2^19 entries x 2
// features) followed by a 5 x 64 ReLU MLP (32 -> 64 -> 64 -> 64 -> 64 -> 64 -> 3), evaluated and
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
constexpr int kRcIn = kRcLevels * kRcFeatures; // 32
constexpr int kRcWidth = 64;
constexpr int kRcHidden = 5;  // hidden layers of width 64
constexpr int kRcOut = 3;
constexpr int kRcOutPad = 16; // output rows padded to one MFMA tile
// Parameter counts: W0 64x32, W1..W4 64x64, W5 16x64 (rows >= 3 are zero and never updated).
constexpr int kRcW0 = kRcWidth * kRcIn;
constexpr int kRcWh = kRcWidth * kRcWidth;
constexpr int kRcW5 = kRcOutPad * kRcWidth;
constexpr int kRcMlpParams = kRcW0 + 4 * kRcWh + kRcW5; // 19456 (18944 used + 512 padding)
constexpr size_t kRcTableParams = size_t(kRcLevels) * kRcTable * kRcFeatures;

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
	float * gt = nullptr;       // table gradient accumulators
	float * mt = nullptr, * vt = nullptr;
};

// Per-sample activations and output gradients of a training batch (bf16, feature-major
// [feature][sample]), used by the weight-gradient GEMMs.
struct RcTrainBuffers {
	uint16_t * x = nullptr; // X0 (32 rows) then X1..X5 (64 rows each): 352 rows
	uint16_t * g = nullptr; // G0..G4 (64 rows each, pre-activation gradients) then GY (16 rows): 336 rows
	float * loss = nullptr; // per-sample loss (for monitoring)
	uint32_t capacity = 0;
};

struct RcAdam {
	float lr = 1e-2f, beta1 = 0.9f, beta2 = 0.99f, eps = 1e-15f;
	float l2 = 1e-6f; // weight decay on MLP weights only
	uint32_t step = 0;
};

// positions: n float4 (xyz used); out: n float4 (rgb radiance).
int rcInfer(const RcGrid & grid, const RcParams & p, const float4 * positions, const uint32_t * countDev, uint32_t n,
	float4 * out, hipStream_t stream);
// One optimisation step on n (position, target) pairs. Returns hipError_t.
int rcTrainStep(const RcGrid & grid, RcParams & p, RcTrainBuffers & tb, RcAdam & adam, const float4 * positions,
	const float4 * targets, uint32_t n, hipStream_t stream);
// Gradients only (no update), for checks: fills p.gw and p.gt for the batch.
int rcGradients(const RcGrid & grid, RcParams & p, RcTrainBuffers & tb, const float4 * positions, const float4 * targets,
	uint32_t n, hipStream_t stream);
int rcSyncCopies(RcParams & p, hipStream_t stream); // refresh bf16/fp16 copies from the masters

} // namespace rdu
