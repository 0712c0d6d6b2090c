// Float3 arithmetic shared by the host BVH build and the gfx950 kernels.
//
// Every formula keeps the operation order of the vendored glm 0.9.9.7 that Rendu's raycaster is
// written against, so that device results are bit-identical to the reference's CPU results:
//   dot       (a*b).x + (a*b).y, then + (a*b).z        src/libs/glm/detail/func_geometric.inl:48-55
//   cross     x.y*y.z - y.y*x.z, ...                    func_geometric.inl:68-80
//   normalize v * (1 / sqrt(dot(v, v)))                 func_geometric.inl:82-90, func_exponential.inl:134-139
//   min / max (y < x) ? y : x / (x < y) ? y : x         func_common.inl:17-30
// The library is compiled with -ffp-contract=off and IEEE division and square root
// (hipcc's default -fhip-fp32-correctly-rounded-divide-sqrt), and without denormal flushing.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define RDU_HD __host__ __device__ __forceinline__

namespace rdu {

// Child code of an unused slot of a 4-wide node.
constexpr int32_t kWideEmpty = 0x7fffffff;
// Instanced scenes: below kInstLimit, codes name instances (kInstCode0 + i, i < kInstMax) or the
// marker that returns a ray from an instance's tree to the top level; triangle leaf codes
// (~(first << 1 | count - 1), first < 2^29) stay at kInstLimit or above.
constexpr int32_t kInstCode0 = int32_t(0x80000000u);
constexpr int32_t kInstMax = 0x3ffffffe;
constexpr int32_t kInstReturn = int32_t(0x80000000u) + 0x3fffffff;
constexpr int32_t kInstLimit = -(1 << 30);

struct f3 {
	float x, y, z;
};

RDU_HD f3 mk(float x, float y, float z) { return {x, y, z}; }
RDU_HD f3 splat(float a) { return {a, a, a}; }
RDU_HD f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
RDU_HD f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
RDU_HD f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
RDU_HD f3 operator/(f3 a, f3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
RDU_HD f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
RDU_HD f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
RDU_HD f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
RDU_HD f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
RDU_HD f3 rcp3(f3 a) { return {1.0f / a.x, 1.0f / a.y, 1.0f / a.z}; }

// Comparison selects: NaN handling follows the first operand, as in glm/std.
RDU_HD float gmin(float x, float y) { return (y < x) ? y : x; }
RDU_HD float gmax(float x, float y) { return (x < y) ? y : x; }
RDU_HD f3 gmin(f3 a, f3 b) { return {gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
RDU_HD f3 gmax(f3 a, f3 b) { return {gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }

RDU_HD float dot(f3 a, f3 b) {
	const f3 t = a * b;
	return (t.x + t.y) + t.z;
}
RDU_HD f3 cross(f3 x, f3 y) {
	return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
RDU_HD float comp(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// Correctly rounded square root (IEEE): sqrtss on the host; on gfx950 the correctly rounded
// lowering selected by -fhip-fp32-correctly-rounded-divide-sqrt.
RDU_HD float sqrt_rn(float x) { return __builtin_sqrtf(x); }

RDU_HD f3 normalize(f3 v) {
	const float inv = 1.0f / sqrt_rn(dot(v, v));
	return v * inv;
}
RDU_HD float length(f3 v) { return sqrt_rn(dot(v, v)); }

} // namespace rdu
