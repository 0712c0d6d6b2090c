// Radiance cache kernels for gfx950: hash-grid and spherical-harmonics encodings, fused 5 x 64 MLP
// forward / backward on bf16 MFMA with fp32 accumulation, split-K weight-gradient GEMMs and Adam.
//
// MFMA mapping (one wavefront). gfx950's v_mfma_f32_16x16x32_bf16: D[16x16] += A[16x32] B[32x16], with
//   A[m][k] in lane m + 16 * (k / 8), element k % 8;   B[k][n] in lane n + 16 * (k / 8), element k % 8;
//   D[m][n] in lane n + 16 * (m / 4), element m % 4.
// A wavefront evaluates 16 samples at a time with layers computed transposed, Y^T = W X^T:
// A = weight tile (m = output feature, k = input feature), B = activations (k = feature,
// n = sample). The D layout of output tile t (lane = sample + 16 g, element e: feature 16 t + 4 g + e)
// is the layout of 16-feature chunk t of the next layer's input. One 32-deep product takes chunks t
// and t + 1 side by side (elements 0-3 from chunk t, 4-7 from chunk t + 1, in both A and B): the
// product sums over k in any order A and B agree on, so activations stay in registers between
// layers with no shuffle. A remaining 16-deep chunk (the 48-wide input, the 16-row output layer's
// transpose) uses v_mfma_f32_16x16x16_bf16 (same D layout; k = 4 g + element). The backward pass
// uses the same structure with transposed weight tiles.
#include "rcache.hpp"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce_by_key.hpp>

namespace rdu {

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4v4 __attribute__((ext_vector_type(4)));

constexpr int kRcBlock = 256; // 4 wavefronts
constexpr int kOffW[6] = {0, kRcW0, kRcW0 + kRcWh, kRcW0 + 2 * kRcWh, kRcW0 + 3 * kRcWh, kRcW0 + 4 * kRcWh};
constexpr int kInDim[6] = {kRcIn, kRcWidth, kRcWidth, kRcWidth, kRcWidth, kRcWidth};
constexpr int kOutDim[6] = {kRcWidth, kRcWidth, kRcWidth, kRcWidth, kRcWidth, kRcOutPad};
// Row offsets of the per-sample buffers (feature-major).
constexpr int kXRow[6] = {0, 48, 112, 176, 240, 304}; // X0..X5
constexpr int kGRow[6] = {0, 64, 128, 192, 256, 320}; // G1..G5, GY

__device__ __forceinline__ uint16_t f2bf(float x) {
	uint32_t u = __float_as_uint(x);
	if((u & 0x7fffffffu) > 0x7f800000u)
		return 0x7fc0;
	u += 0x7fffu + ((u >> 16) & 1u); // round to nearest even
	return uint16_t(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }
__device__ __forceinline__ float h2f(uint16_t h) { return float(__builtin_bit_cast(_Float16, h)); }
__device__ __forceinline__ uint16_t f2h(float x) { return __builtin_bit_cast(uint16_t, _Float16(x)); }

__device__ __forceinline__ s4 pack4(const float v[4]) {
	s4 r;
	r.x = short(f2bf(v[0]));
	r.y = short(f2bf(v[1]));
	r.z = short(f2bf(v[2]));
	r.w = short(f2bf(v[3]));
	return r;
}
__device__ __forceinline__ float lane4(const s4 & v, int e) {
	const short s = e == 0 ? v.x : (e == 1 ? v.y : (e == 2 ? v.z : v.w));
	return bf2f(uint16_t(s));
}

// Stores one B-layout tile (features 16 t + 4 g + e of sample idx) into a feature-major bf16 buffer.
__device__ __forceinline__ void storeTile(uint16_t * buf, int row0, int t, int g, uint32_t cap, uint32_t idx, s4 v) {
	const size_t base = size_t(row0 + 16 * t + 4 * g) * cap + idx;
	buf[base] = uint16_t(v.x);
	buf[base + cap] = uint16_t(v.y);
	buf[base + 2 * size_t(cap)] = uint16_t(v.z);
	buf[base + 3 * size_t(cap)] = uint16_t(v.w);
}
// Bits e of a B-layout tile whose bf16 value is non-zero (post-ReLU activations are >= 0).
__device__ __forceinline__ uint32_t nonzero4(s4 v) {
	return (v.x != 0 ? 1u : 0u) | (v.y != 0 ? 2u : 0u) | (v.z != 0 ? 4u : 0u) | (v.w != 0 ? 8u : 0u);
}

// 16-deep product (one chunk of A and B).
__device__ __forceinline__ f4v4 mfma(s4 a, s4 b, f4v4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
// 32-deep product over two chunks (a0, b0) and (a1, b1).
__device__ __forceinline__ f4v4 mfma2(s4 a0, s4 a1, s4 b0, s4 b1, f4v4 c) {
	const s8 a = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
	const s8 b = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
	return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c, 0, 0, 0);
}

// A operand: W[mt*16 + m][kt*16 + 4g .. +3] (row-major [out][in], in = ld).
__device__ __forceinline__ s4 loadA(const uint16_t * W, int ld, int mt, int kt, int lane) {
	const int m = mt * 16 + (lane & 15), k = kt * 16 + 4 * (lane >> 4);
	const uint16_t * p = W + m * ld + k;
	s4 r;
	r.x = short(p[0]);
	r.y = short(p[1]);
	r.z = short(p[2]);
	r.w = short(p[3]);
	return r;
}
// A operand of W^T: A[m][k] = W[k][m], m = input feature (tile mt), k = output feature (tile kt).
__device__ __forceinline__ s4 loadAT(const uint16_t * W, int ld, int mt, int kt, int lane) {
	const int m = mt * 16 + (lane & 15), k = kt * 16 + 4 * (lane >> 4);
	s4 r;
	r.x = short(W[(k + 0) * ld + m]);
	r.y = short(W[(k + 1) * ld + m]);
	r.z = short(W[(k + 2) * ld + m]);
	r.w = short(W[(k + 3) * ld + m]);
	return r;
}

// Hash-grid features of one level at normalised position p (trilinear over 8 corners).
__device__ __forceinline__ void cornerSetup(const RcGrid & grid, int level, const float p[3], int c, uint32_t & idx,
	float & w) {
	const int res = grid.res[level];
	int cell[3];
	float fr[3];
	for(int a = 0; a < 3; ++a) {
		const float x = p[a] * float(res);
		int ci = int(floorf(x));
		ci = ci < 0 ? 0 : (ci > res - 1 ? res - 1 : ci);
		cell[a] = ci;
		fr[a] = x - float(ci);
	}
	const int dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
	const uint32_t cx = uint32_t(cell[0] + dx), cy = uint32_t(cell[1] + dy), cz = uint32_t(cell[2] + dz);
	const float wx = dx ? fr[0] : 1.0f - fr[0];
	const float wy = dy ? fr[1] : 1.0f - fr[1];
	const float wz = dz ? fr[2] : 1.0f - fr[2];
	w = (wx * wy) * wz;
	if(grid.dense[level]) {
		const uint32_t r1 = uint32_t(res + 1);
		idx = cx + cy * r1 + cz * r1 * r1;
	} else {
		idx = (cx * 1u ^ cy * 2654435761u ^ cz * 805459861u) & (kRcTable - 1u);
	}
}

__device__ __forceinline__ void encodeLevel(const RcGrid & grid, const uint16_t * t16, int level, const float p[3],
	float & f0, float & f1) {
	f0 = 0.0f;
	f1 = 0.0f;
	const uint16_t * tl = t16 + size_t(level) * kRcTable * kRcFeatures;
	for(int c = 0; c < 8; ++c) {
		uint32_t idx;
		float w;
		cornerSetup(grid, level, p, c, idx, w);
		const uint32_t pair = reinterpret_cast<const uint32_t *>(tl)[idx]; // 2 fp16 features
		f0 += w * h2f(uint16_t(pair & 0xffffu));
		f1 += w * h2f(uint16_t(pair >> 16));
	}
}

// Table-gradient records of one sample and level: record (sample, level, corner) holds the
// entry's key and the corner-weighted gradients of its two features.
__device__ __forceinline__ void recordLevel(const RcGrid & grid, const RcTrainBuffers & tb, uint32_t sample, int level,
	const float p[3], float g0, float g1) {
	const size_t base = (size_t(sample) * kRcLevels + size_t(level)) * 8;
	for(int c = 0; c < 8; ++c) {
		uint32_t idx;
		float w;
		cornerSetup(grid, level, p, c, idx, w);
		tb.recKey[base + c] = uint32_t(level) * kRcTable + idx;
		tb.recVal[base + c] = RcGradPair{w * g0, w * g1};
	}
}

struct GradPairSum {
	__device__ __forceinline__ RcGradPair operator()(const RcGradPair & x, const RcGradPair & y) const {
		return RcGradPair{x.a + y.a, x.b + y.b};
	}
};

// gt[key] += per-key sums (keys are unique, so no two threads touch one entry).
__global__ void rcApplyTableGrad(const uint32_t * __restrict__ keys, const RcGradPair * __restrict__ sums,
	const uint32_t * __restrict__ count, float * __restrict__ gt) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= *count)
		return;
	const size_t k = keys[i];
	gt[2 * k] += sums[i].a;
	gt[2 * k + 1] += sums[i].b;
}

// Real spherical harmonics of degrees 0..3 of a unit direction, coefficients 4 * chunk .. + 3
// (order: l = 0; l = 1, m = -1, 0, 1; l = 2, m = -2..2; l = 3, m = -3..3).
__device__ __forceinline__ void shChunk(int chunk, float x, float y, float z, float e[4]) {
	const float xx = x * x, yy = y * y, zz = z * z;
	if(chunk == 0) {
		e[0] = 0.28209479177387814f;
		e[1] = -0.4886025119029199f * y;
		e[2] = 0.4886025119029199f * z;
		e[3] = -0.4886025119029199f * x;
	} else if(chunk == 1) {
		e[0] = 1.0925484305920792f * (x * y);
		e[1] = -1.0925484305920792f * (y * z);
		e[2] = 0.31539156525252005f * (3.0f * zz - 1.0f);
		e[3] = -1.0925484305920792f * (x * z);
	} else if(chunk == 2) {
		e[0] = 0.5462742152960396f * (xx - yy);
		e[1] = -0.5900435899266435f * (y * (3.0f * xx - yy));
		e[2] = 2.890611442640554f * ((x * y) * z);
		e[3] = -0.4570457994644658f * (y * (5.0f * zz - 1.0f));
	} else {
		e[0] = 0.3731763325901154f * (z * (5.0f * zz - 3.0f));
		e[1] = -0.4570457994644658f * (x * (5.0f * zz - 1.0f));
		e[2] = 1.445305721320277f * (z * (xx - yy));
		e[3] = -0.5900435899266435f * (x * (xx - 3.0f * yy));
	}
}

// Unit direction of a query (a zero or non-finite direction encodes as +z).
__device__ __forceinline__ void queryDir(float4 q, float d[3]) {
	const float l2 = (q.x * q.x + q.y * q.y) + q.z * q.z;
	if(!(l2 > 0.0f) || !(l2 < 3.0e38f)) {
		d[0] = 0.0f;
		d[1] = 0.0f;
		d[2] = 1.0f;
		return;
	}
	const float inv = 1.0f / sqrtf(l2);
	d[0] = q.x * inv;
	d[1] = q.y * inv;
	d[2] = q.z * inv;
}

__device__ __forceinline__ void normalisePos(const RcGrid & grid, float4 q, float p[3]) {
	const float v[3] = {q.x, q.y, q.z};
	for(int a = 0; a < 3; ++a) {
		const float t = (v[a] - grid.lo[a]) * grid.inv[a];
		p[a] = t > 0.0f ? (t < 1.0f ? t : 1.0f) : 0.0f; // NaN -> 0
	}
}

__device__ __forceinline__ void loadWeights(uint16_t * W, const uint16_t * w16) {
	for(int i = threadIdx.x; i < kRcMlpParams / 4; i += kRcBlock)
		reinterpret_cast<uint2 *>(W)[i] = reinterpret_cast<const uint2 *>(w16)[i];
}

// Fused per-sample pass. TRAIN: forward, relative-L2 loss gradient, backward to the encoding
// (scattered into the table gradients), activations and output gradients stored for the
// weight-gradient GEMMs. Otherwise: forward only, radiance written to `out`.
template<bool TRAIN>
__global__ __launch_bounds__(kRcBlock) void rcFused(RcGrid grid, const uint16_t * __restrict__ w16,
	const uint16_t * __restrict__ t16, const float4 * __restrict__ queries, const float4 * __restrict__ targets,
	const uint32_t * __restrict__ countDev, uint32_t n, float4 * __restrict__ out, RcTrainBuffers tb,
	float * __restrict__ gt, float lossScale) {
	__shared__ uint16_t W[kRcMlpParams];
	loadWeights(W, w16);
	__syncthreads();
	if(countDev)
		n = *countDev < n ? *countDev : n;
	const int lane = threadIdx.x & 63;
	const int q = lane & 15, g = lane >> 4;
	const uint32_t waveGlobal = blockIdx.x * (kRcBlock / 64) + (threadIdx.x >> 6);
	const uint32_t waves = gridDim.x * (kRcBlock / 64);
	const uint32_t nPad = (n + 15u) & ~15u;
	for(uint32_t base = waveGlobal * 16u; base < nPad; base += waves * 16u) {
		const uint32_t idx = base + uint32_t(q);
		const bool valid = idx < n;
		float p[3] = {0.0f, 0.0f, 0.0f}, d[3] = {0.0f, 0.0f, 1.0f};
		if(valid) {
			normalisePos(grid, queries[2 * size_t(idx)], p);
			queryDir(queries[2 * size_t(idx) + 1], d);
		}
		// Encoding: this lane computes levels 2g, 2g+1 (chunk 0), 8+2g, 9+2g (chunk 1) and the
		// spherical harmonics 4g..4g+3 (chunk 2).
		// Training stores each activation tile (for the weight gradients) as soon as it exists and
		// keeps only a bit per element of whether it is non-zero (the ReLU derivative), so that
		// activations do not stay live in registers through the backward pass.
		const uint32_t cap = tb.capacity;
		uint32_t active[6] = {0, 0, 0, 0, 0, 0}; // bit 4 mt + e: X[l][mt] element e > 0
		s4 X[6][4];
		{
			float e[4];
			encodeLevel(grid, t16, 2 * g, p, e[0], e[1]);
			encodeLevel(grid, t16, 2 * g + 1, p, e[2], e[3]);
			X[0][0] = pack4(e);
			encodeLevel(grid, t16, 8 + 2 * g, p, e[0], e[1]);
			encodeLevel(grid, t16, 9 + 2 * g, p, e[2], e[3]);
			X[0][1] = pack4(e);
			shChunk(g, d[0], d[1], d[2], e);
			X[0][2] = pack4(e);
			if(TRAIN)
				for(int t = 0; t < 3; ++t)
					storeTile(tb.x, kXRow[0], t, g, cap, idx, X[0][t]);
		}
		// Hidden layers 0..4 (ReLU).
#pragma unroll
		for(int l = 0; l < 5; ++l) {
			const uint16_t * Wl = W + kOffW[l];
			const int kts = kInDim[l] / 16;
#pragma unroll
			for(int mt = 0; mt < 4; ++mt) {
				f4v4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
				for(int kt = 0; kt + 1 < kts; kt += 2)
					acc = mfma2(loadA(Wl, kInDim[l], mt, kt, lane), loadA(Wl, kInDim[l], mt, kt + 1, lane), X[l][kt],
						X[l][kt + 1], acc);
				if(kts & 1)
					acc = mfma(loadA(Wl, kInDim[l], mt, kts - 1, lane), X[l][kts - 1], acc);
				const float r[4] = {acc.x > 0.0f ? acc.x : 0.0f, acc.y > 0.0f ? acc.y : 0.0f, acc.z > 0.0f ? acc.z : 0.0f,
					acc.w > 0.0f ? acc.w : 0.0f};
				X[l + 1][mt] = pack4(r);
				if(TRAIN) {
					storeTile(tb.x, kXRow[l + 1], mt, g, cap, idx, X[l + 1][mt]);
					active[l + 1] |= nonzero4(X[l + 1][mt]) << (4 * mt);
				}
			}
		}
		// Output layer (rows 0..2 used).
		f4v4 Y = {0.0f, 0.0f, 0.0f, 0.0f};
		for(int kt = 0; kt < 4; kt += 2)
			Y = mfma2(loadA(W + kOffW[5], kRcWidth, 0, kt, lane), loadA(W + kOffW[5], kRcWidth, 0, kt + 1, lane), X[5][kt],
				X[5][kt + 1], Y);
		if(!TRAIN) {
			if(valid && g == 0)
				out[idx] = make_float4(Y.x, Y.y, Y.z, 0.0f);
			continue;
		}
		// Relative L2 loss gradient on the three outputs: d/dy (y - t)^2 / (sg(y)^2 + 0.01).
		float gy[4] = {0.0f, 0.0f, 0.0f, 0.0f};
		if(valid && g == 0) {
			const float4 t = targets[idx];
			const float y3[3] = {Y.x, Y.y, Y.z}, t3[3] = {t.x, t.y, t.z};
			float loss = 0.0f;
			for(int j = 0; j < 3; ++j) {
				const float den = y3[j] * y3[j] + 0.01f;
				const float d = y3[j] - t3[j];
				gy[j] = 2.0f * d / den * lossScale;
				loss += d * d / den;
			}
			tb.loss[idx] = loss / 3.0f;
		}
		s4 G[6][4]; // G[l][t]: gradient w.r.t. the pre-activation of layer l's output, tile t
		G[5][0] = pack4(gy);
		storeTile(tb.g, kGRow[5], 0, g, cap, idx, valid ? G[5][0] : s4{0, 0, 0, 0});
		// Back through the output layer: dX5 = W5^T GY, then through each hidden layer.
#pragma unroll
		for(int l = 5; l >= 1; --l) {
			const uint16_t * Wl = W + kOffW[l];
			const int kts = kOutDim[l] / 16; // K of the transposed product
#pragma unroll
			for(int mt = 0; mt < 4; ++mt) {
				f4v4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
				for(int kt = 0; kt + 1 < kts; kt += 2)
					acc = mfma2(loadAT(Wl, kInDim[l], mt, kt, lane), loadAT(Wl, kInDim[l], mt, kt + 1, lane), G[l][kt],
						G[l][kt + 1], acc);
				if(kts & 1)
					acc = mfma(loadAT(Wl, kInDim[l], mt, kts - 1, lane), G[l][kts - 1], acc);
				// ReLU derivative: X[l][mt] > 0 (post-activation of layer l - 1).
				float r[4];
				const float a4[4] = {acc.x, acc.y, acc.z, acc.w};
				for(int e = 0; e < 4; ++e)
					r[e] = ((active[l] >> (4 * mt + e)) & 1u) ? a4[e] : 0.0f;
				G[l - 1][mt] = pack4(r);
				storeTile(tb.g, kGRow[l - 1], mt, g, cap, idx, valid ? G[l - 1][mt] : s4{0, 0, 0, 0});
			}
		}
		// dX0 = W0^T G0 for the 32 hash-grid features, scattered into the tables (the spherical
		// harmonics have no parameters).
		for(int mt = 0; mt < 2; ++mt) {
			f4v4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
			for(int kt = 0; kt < 4; kt += 2)
				acc = mfma2(loadAT(W, kRcIn, mt, kt, lane), loadAT(W, kRcIn, mt, kt + 1, lane), G[0][kt], G[0][kt + 1], acc);
			if(valid) {
				const int l0 = 8 * mt + 2 * g;
				recordLevel(grid, tb, idx, l0, p, acc.x, acc.y);
				recordLevel(grid, tb, idx, l0 + 1, p, acc.z, acc.w);
			}
		}
		// X0..X5 and G0..G4, GY (bf16, feature-major) are stored above for the weight gradients
		// dW_l = G_l X_l^T.
	}
}

// dW_l[out][in] += sum over samples of G_l[out][s] X_l[in][s]: one 16x16 tile per wavefront over a
// range of samples (split K), accumulated with atomics.
constexpr uint32_t kGradTiles = 80;
__global__ __launch_bounds__(kRcBlock) void rcWeightGrad(RcTrainBuffers tb, uint32_t nPad, uint32_t chunk, float * gw) {
	// Tile list: layer 0: 4x3, layers 1..4: 4x4 each, layer 5: 1x4 -> 12 + 64 + 4 = 80 tiles.
	const int lane = threadIdx.x & 63;
	const uint32_t wave = blockIdx.x * (kRcBlock / 64) + (threadIdx.x >> 6);
	const uint32_t tile = wave % kGradTiles, part = wave / kGradTiles;
	int layer, mt, nt;
	if(tile < 12) {
		layer = 0;
		mt = int(tile) / 3;
		nt = int(tile) % 3;
	} else if(tile < 76) {
		layer = 1 + int(tile - 12) / 16;
		mt = int((tile - 12) % 16) / 4;
		nt = int((tile - 12) % 16) % 4;
	} else {
		layer = 5;
		mt = 0;
		nt = int(tile - 76);
	}
	const uint32_t s0 = part * chunk;
	if(s0 >= nPad)
		return;
	const uint32_t s1 = s0 + chunk < nPad ? s0 + chunk : nPad;
	const uint32_t cap = tb.capacity;
	const uint16_t * Gr = tb.g + size_t(kGRow[layer] + mt * 16 + (lane & 15)) * cap;
	const uint16_t * Xr = tb.x + size_t(kXRow[layer] + nt * 16 + (lane & 15)) * cap;
	// k = sample: lane group g takes samples s + 4 g .. + 3 of each 16-sample chunk.
	const auto chunkAt = [&](const uint16_t * row, uint32_t s) {
		const uint2 u = *reinterpret_cast<const uint2 *>(row + s + 4u * uint32_t(lane >> 4));
		const s4 r = {short(u.x & 0xffffu), short(u.x >> 16), short(u.y & 0xffffu), short(u.y >> 16)};
		return r;
	};
	f4v4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
	uint32_t s = s0;
	for(; s + 32u <= s1; s += 32u)
		acc = mfma2(chunkAt(Gr, s), chunkAt(Gr, s + 16u), chunkAt(Xr, s), chunkAt(Xr, s + 16u), acc);
	if(s < s1)
		acc = mfma(chunkAt(Gr, s), chunkAt(Xr, s), acc);
	// D[m][n]: lane n + 16 (m / 4), element m % 4 -> gw[layer][mt*16 + m][nt*16 + n].
	const int ld = kInDim[layer];
	float * base = gw + kOffW[layer];
	const int n = nt * 16 + (lane & 15);
	const float a4[4] = {acc.x, acc.y, acc.z, acc.w};
	for(int e = 0; e < 4; ++e) {
		const int m = mt * 16 + 4 * (lane >> 4) + e;
		if(layer == 5 && m >= kRcOut)
			continue;
		atomicAdd(base + m * ld + n, a4[e]);
	}
}

// Adam on one parameter array; refreshes the low-precision copy and clears the gradient.
template<bool BF16>
__global__ void rcAdam(float * w, uint16_t * w16, float * gsum, float * m, float * v, size_t n, float lr, float b1, float b2,
	float eps, float l2, float c1, float c2) {
	const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if(i >= n)
		return;
	float gr = gsum[i] + l2 * w[i];
	const float mi = b1 * m[i] + (1.0f - b1) * gr;
	const float vi = b2 * v[i] + (1.0f - b2) * gr * gr;
	m[i] = mi;
	v[i] = vi;
	const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
	const float wi = w[i] - upd;
	w[i] = wi;
	w16[i] = BF16 ? f2bf(wi) : f2h(wi);
	gsum[i] = 0.0f;
}

// Adam on the hash-table entries one step touched (the keys of the reduction, each with its summed
// gradient pair), read straight from the reduction: entries no sample touched keep their moments
// and weights. This is the lazy Adam of instant neural graphics primitives (Mueller et al. 2022,
// section 4: table entries with a zero gradient are skipped), and it makes a step's cost follow the
// batch instead of the 2^19 x 16 levels x 2 table parameters.
__global__ void rcAdamTouched(const uint32_t * __restrict__ keys, const RcGradPair * __restrict__ sums,
	const uint32_t * __restrict__ count, float * w, uint16_t * w16, float * m, float * v, float lr, float b1, float b2,
	float eps, float c1, float c2) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= *count)
		return;
	const size_t k = keys[i];
	const RcGradPair g = sums[i];
	const float gr[2] = {g.a, g.b};
	for(int j = 0; j < 2; ++j) {
		const size_t e = 2 * k + j;
		const float mi = b1 * m[e] + (1.0f - b1) * gr[j];
		const float vi = b2 * v[e] + (1.0f - b2) * gr[j] * gr[j];
		m[e] = mi;
		v[e] = vi;
		const float wi = w[e] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
		w[e] = wi;
		w16[e] = f2h(wi);
	}
}

template<bool BF16>
__global__ void rcCopy(const float * w, uint16_t * w16, size_t n) {
	const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if(i < n)
		w16[i] = BF16 ? f2bf(w[i]) : f2h(w[i]);
}

// Sums the table-gradient records of n samples per entry: radix sort by key (stable), then a
// deterministic reduction by key into (tb.uniqKey, tb.uniqVal, *tb.uniqCount); with intoGt, one add
// per touched entry into gt.
int reduceTableGradients(RcParams & p, RcTrainBuffers & tb, uint32_t n, bool intoGt, hipStream_t stream) {
	const uint32_t recs = n * kRcRecordsPerSample;
	rocprim::double_buffer<uint32_t> keys(tb.recKey, tb.recKeyAlt);
	rocprim::double_buffer<RcGradPair> vals(tb.recVal, tb.recValAlt);
	size_t bytes = tb.tempBytes;
	hipError_t e = rocprim::radix_sort_pairs(tb.temp, bytes, keys, vals, recs, 0, kRcLog2Table + 4, stream);
	if(e != hipSuccess)
		return int(e);
	bytes = tb.tempBytes;
	e = rocprim::deterministic_reduce_by_key(tb.temp, bytes, keys.current(), vals.current(), size_t(recs), tb.uniqKey,
		tb.uniqVal, tb.uniqCount, GradPairSum(), rocprim::equal_to<uint32_t>(), stream);
	if(e != hipSuccess)
		return int(e);
	if(!intoGt)
		return 0;
	hipLaunchKernelGGL(rcApplyTableGrad, dim3((recs + 255) / 256), dim3(256), 0, stream, tb.uniqKey, tb.uniqVal,
		tb.uniqCount, p.gt);
	return int(hipGetLastError());
}

unsigned fusedBlocks(uint32_t n) {
	const uint32_t groups = (n + 15u) / 16u;
	uint32_t blocks = (groups + 3u) / 4u;
	return blocks < 1 ? 1u : (blocks > 4096u ? 4096u : blocks);
}

int weightGradients(RcParams & p, RcTrainBuffers & tb, uint32_t n, hipStream_t stream) {
	const uint32_t nPad = (n + 15u) & ~15u;
	const uint32_t chunk = 1024;
	const uint32_t parts = (nPad + chunk - 1) / chunk;
	const uint32_t waves = kGradTiles * parts;
	hipLaunchKernelGGL(rcWeightGrad, dim3((waves + 3) / 4), dim3(kRcBlock), 0, stream, tb, nPad, chunk, p.gw);
	return int(hipGetLastError());
}

} // namespace

size_t rcReduceTempBytes(uint32_t records) {
	size_t a = 0, b = 0;
	rocprim::double_buffer<uint32_t> keys(nullptr, nullptr);
	rocprim::double_buffer<RcGradPair> vals(nullptr, nullptr);
	(void)rocprim::radix_sort_pairs(nullptr, a, keys, vals, records, 0, kRcLog2Table + 4);
	(void)rocprim::deterministic_reduce_by_key(nullptr, b, static_cast<const uint32_t *>(nullptr),
		static_cast<const RcGradPair *>(nullptr), size_t(records), static_cast<uint32_t *>(nullptr),
		static_cast<RcGradPair *>(nullptr), static_cast<uint32_t *>(nullptr), GradPairSum(), rocprim::equal_to<uint32_t>());
	return a > b ? a : b;
}

int rcInfer(const RcGrid & grid, const RcParams & p, const float4 * queries, const uint32_t * countDev, uint32_t n,
	float4 * out, hipStream_t stream) {
	if(n == 0)
		return 0;
	RcTrainBuffers none;
	hipLaunchKernelGGL((rcFused<false>), dim3(fusedBlocks(n)), dim3(kRcBlock), 0, stream, grid, p.w16, p.t16, queries,
		nullptr, countDev, n, out, none, nullptr, 0.0f);
	return int(hipGetLastError());
}

namespace {

int gradientsImpl(const RcGrid & grid, RcParams & p, RcTrainBuffers & tb, const float4 * queries, const float4 * targets,
	uint32_t n, bool intoGt, hipStream_t stream) {
	if(n == 0 || n > tb.capacity)
		return int(hipErrorInvalidValue);
	const float lossScale = 1.0f / (3.0f * float(n));
	hipLaunchKernelGGL((rcFused<true>), dim3(fusedBlocks(n)), dim3(kRcBlock), 0, stream, grid, p.w16, p.t16, queries,
		targets, nullptr, n, nullptr, tb, p.gt, lossScale);
	int e = int(hipGetLastError());
	if(e)
		return e;
	if((e = reduceTableGradients(p, tb, n, intoGt, stream)))
		return e;
	return weightGradients(p, tb, n, stream);
}

} // namespace

int rcGradients(const RcGrid & grid, RcParams & p, RcTrainBuffers & tb, const float4 * queries, const float4 * targets,
	uint32_t n, hipStream_t stream) {
	return gradientsImpl(grid, p, tb, queries, targets, n, true, stream);
}

int rcTrainStep(const RcGrid & grid, RcParams & p, RcTrainBuffers & tb, RcAdam & adam, const float4 * queries,
	const float4 * targets, uint32_t n, hipStream_t stream) {
	int e = gradientsImpl(grid, p, tb, queries, targets, n, false, stream);
	if(e)
		return e;
	adam.step += 1;
	const float c1 = 1.0f - powf(adam.beta1, float(adam.step));
	const float c2 = 1.0f - powf(adam.beta2, float(adam.step));
	hipLaunchKernelGGL((rcAdam<true>), dim3((kRcMlpParams + 255) / 256), dim3(256), 0, stream, p.w32, p.w16, p.gw, p.mw,
		p.vw, size_t(kRcMlpParams), adam.lr, adam.beta1, adam.beta2, adam.eps, adam.l2, c1, c2);
	const uint32_t recs = n * kRcRecordsPerSample;
	hipLaunchKernelGGL(rcAdamTouched, dim3((recs + 255) / 256), dim3(256), 0, stream, tb.uniqKey, tb.uniqVal, tb.uniqCount,
		p.t32, p.t16, p.mt, p.vt, adam.lr, adam.beta1, adam.beta2, adam.eps, c1, c2);
	return int(hipGetLastError());
}

int rcSyncCopies(RcParams & p, hipStream_t stream) {
	hipLaunchKernelGGL((rcCopy<true>), dim3((kRcMlpParams + 255) / 256), dim3(256), 0, stream, p.w32, p.w16,
		size_t(kRcMlpParams));
	hipLaunchKernelGGL((rcCopy<false>), dim3(unsigned((kRcTableParams + 255) / 256)), dim3(256), 0, stream, p.t32, p.t16,
		kRcTableParams);
	return int(hipGetLastError());
}

} // namespace rdu
