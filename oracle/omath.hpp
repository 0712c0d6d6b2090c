// ORACLE — test infrastructure only. Nothing in the product (rendu_amd/) may include, link
// or call this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
// and only as the checker or the timed CPU baseline.
//
// Scalar vector helpers restating the vendored glm 0.9.9.7 formulas the reference path uses,
// operation for operation, so that the CPU oracle reproduces the reference's float results:
//   dot      : tmp = a*b; (tmp.x + tmp.y) + tmp.z        src/libs/glm/detail/func_geometric.inl:48-55
//   cross    : (x.y*y.z - y.y*x.z, x.z*y.x - y.z*x.x, x.x*y.y - y.x*x.y)            :68-80
//   normalize: v * (1 / sqrt(dot(v, v)))                   :82-90, func_exponential.inl:134-139
//   length   : sqrt(dot(v, v))                             :8-14
//   min / max: (y < x) ? y : x  /  (x < y) ? y : x         func_common.inl:17-30 (std:: the same)
// Built with -O2 -ffp-contract=off and no -march, like the reference (premake5.lua:49-51, 66).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstddef>

namespace orc {

struct v3 {
	float x, y, z;
	v3() : x(0.f), y(0.f), z(0.f) {}
	v3(float a) : x(a), y(a), z(a) {}
	v3(float a, float b, float c) : x(a), y(b), z(c) {}
	float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
	float & operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
struct v2 {
	float x, y;
};
struct v4 {
	float x, y, z, w;
};

inline v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline v3 operator*(v3 a, v3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline v3 operator/(v3 a, v3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline v3 operator*(float s, v3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline v3 operator*(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline v3 operator/(v3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline v3 operator/(float s, v3 a) { return {s / a.x, s / a.y, s / a.z}; }
inline v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }

inline float gmin(float x, float y) { return (y < x) ? y : x; }
inline float gmax(float x, float y) { return (x < y) ? y : x; }
inline v3 gmin(v3 a, v3 b) { return {gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
inline v3 gmax(v3 a, v3 b) { return {gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }

inline float dot(v3 a, v3 b) {
	const v3 t = a * b;
	return t.x + t.y + t.z; // left to right: (x + y) + z
}
inline v3 cross(v3 x, v3 y) {
	return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
inline v3 normalize(v3 v) {
	const float inv = 1.0f / std::sqrt(dot(v, v));
	return v * inv;
}
inline float length(v3 v) { return std::sqrt(dot(v, v)); }

// glm::mat3 / mat4 are column-major: m[c][r].
struct m3 {
	v3 c[3];
};
inline v3 mul(const m3 & m, v3 v) {
	// glm mat3 * vec3 (type_mat3x3.inl operator*): per row, (m0*v.x + m1*v.y) + m2*v.z
	return {m.c[0].x * v.x + m.c[1].x * v.y + m.c[2].x * v.z,
			m.c[0].y * v.x + m.c[1].y * v.y + m.c[2].y * v.z,
			m.c[0].z * v.x + m.c[1].z * v.y + m.c[2].z * v.z};
}
inline m3 transpose(const m3 & m) {
	m3 r;
	for(int i = 0; i < 3; ++i)
		for(int j = 0; j < 3; ++j)
			r.c[i][j] = m.c[j][i];
	return r;
}

} // namespace orc
