// ORACLE — test infrastructure only. Nothing in the product (rendu_amd/) may include, link
// or call this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
// and only as the checker or the timed CPU baseline.
//
// CPU restatement of Rendu's path tracer: PathTracer::render, buildLocalFrame, checkVisibility,
// evalBackground, getSampleGrid/getSamplePosition (src/apps/pathtracer/PathTracer.cpp:11-299),
// MaterialGGX (MaterialGGX.cpp:5-101), MaterialSky (MaterialSky.cpp:6-98), Image::rgbl/rgbal
// (src/engine/resources/Image.cpp:64-106), Texture::sampleCubemap (Texture.cpp:141-190), the light
// samplers (src/engine/scene/lights/{Point,Directional,Spot}Light.cpp) and Camera::pixelShifts
// (src/engine/input/Camera.cpp:42-49), over the same scene description as the product's C ABI
// (include/rdu.h, used here for its struct layouts only).
//
// Random numbers: Rendu draws from per-thread std::mt19937 in thread-creation order
// (Random.cpp:72-85), which cannot be replayed; this oracle draws the same quantities in the same
// places from the Philox stream of ophilox.hpp keyed by (seed, pixel, sample, bounce), so the GPU
// renderer can be compared with it sample by sample.
//
// Pinning (tests/test_oracle_pathtracer.py): GGX closed forms and pdf normalisation, concentric
// disk and cosine-hemisphere sampling distributions, cubemap face table, bilinear closed forms,
// camera closed form, the white-furnace scene and convergence of the estimator.
#include "oraycaster.hpp"
#include "ophilox.hpp"
#include "../include/rdu.h"

#include <cmath>
#include <memory>

using namespace orc;

namespace {

constexpr float kPi = 3.14159265358979323846264338327950288f;     // glm::pi<float>()
constexpr float kOneOverPi = 0.318309886183790671537767526745028724f; // glm::one_over_pi<float>()
constexpr float kTwoPi = 6.28318530717958647692528676655900576f;     // glm::two_pi<float>()
constexpr float kHalfPi = 1.57079632679489661923132169163975144f;
constexpr float kQuarterPi = 0.785398163397448309615660845819875721f;

inline v4 operator+(v4 a, v4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline v4 operator*(float s, v4 a) { return {s * a.x, s * a.y, s * a.z, s * a.w}; }
inline float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; } // glm compute_mix
inline v3 mix3(v3 x, v3 y, float a) { return {mixf(x.x, y.x, a), mixf(x.y, y.y, a), mixf(x.z, y.z, a)}; }
inline float clampf(float v, float lo, float hi) { return gmin(gmax(v, lo), hi); } // glm::clamp
inline int modPos(int x, int w) { return ((x % w) + w) % w; }                    // Image.hpp:189-191

struct Texture {
	int w = 0, h = 0;
	std::vector<float> px; // RGBA
	v4 at(int x, int y) const {
		const float * p = &px[size_t(y * w + x) * 4];
		return {p[0], p[1], p[2], p[3]};
	}
	// Image::rgbal (Image.cpp:86-106); rgbl is the same on the first three channels.
	v4 rgbal(float x, float y) const {
		const float xi = x * float(w);
		const float yi = y * float(h);
		const float xb = std::floor(xi);
		const float yb = std::floor(yi);
		const float dx = xi - xb;
		const float dy = yi - yb;
		const int x0 = modPos(int(xb), w);
		const int y0 = modPos(int(yb), h);
		const int x1 = modPos(int(xb) + 1, w);
		const int y1 = modPos(int(yb) + 1, h);
		const v4 p00 = at(x0, y0), p01 = at(x0, y1), p10 = at(x1, y0), p11 = at(x1, y1);
		return (1.0f - dx) * ((1.0f - dy) * p00 + dy * p01) + dx * ((1.0f - dy) * p10 + dy * p11);
	}
	v3 rgbl(float x, float y) const {
		const v4 c = rgbal(x, y);
		return {c.x, c.y, c.z};
	}
};

struct MeshData {
	std::vector<v3> nrm, tan, bit;
	std::vector<v2> uv;
	std::vector<uint32_t> idx;
	bool hasUV = false;
};

struct Object {
	uint32_t mesh = 0, mat = 0;
	bool useUV = false;
	float model[16];
	m3 invtp;
};

// Extension (extended estimator only): light sampling of the lat-long environment and of emissive
// triangles, combined with BSDF sampling by the power heuristic. Tables are built in double
// precision in a fixed order, as the product builds them, so both pick the same samples.
struct AreaLights {
	int envW = 0, envH = 0;
	std::vector<float> envPmf;    // per texel: probability of the texel
	std::vector<float> envRowCdf; // envH + 1
	std::vector<float> envColCdf; // envH rows of envW + 1
	struct Tri {
		v3 v0, e1, e2, n; // n = cross(e1, e2)
		float area;
		uint32_t object, localId;
	};
	std::vector<Tri> tris;
	std::vector<float> triPmf, triCdf;  // triCdf: tris + 1
	std::vector<int32_t> objectBase;    // first emissive triangle of each object, or -1
	float pEnv = 0.0f;                  // environment vs triangles, within area-light sampling
	float pArea = 0.0f;                 // area lights vs analytic lights
};

struct Scene {
	std::vector<Texture> textures;
	std::vector<MeshData> meshes;
	std::vector<rdu_mesh_desc> meshDescs; // positions kept for the raycaster
	std::vector<std::vector<float>> meshPos;
	std::vector<rdu_material_desc> materials;
	std::vector<Object> objects;
	std::vector<rdu_light_desc> lights;
	rdu_background_desc bg{};
	Raycaster rc;
	AreaLights area;
	bool committed = false;
};

// glm inverse(mat3) (func_matrix.inl:269-291) of transpose(mat3(model)).
m3 inverseTransposeUpper(const float * m4) {
	m3 m; // m = transpose(mat3(model)): m.c[c][r] = model[r][c] = m4[4*r + c]
	for(int c = 0; c < 3; ++c)
		for(int r = 0; r < 3; ++r)
			m.c[c][r] = m4[4 * r + c];
	const float ood = 1.0f / (+m.c[0][0] * (m.c[1][1] * m.c[2][2] - m.c[2][1] * m.c[1][2]) -
								 m.c[1][0] * (m.c[0][1] * m.c[2][2] - m.c[2][1] * m.c[0][2]) +
								 m.c[2][0] * (m.c[0][1] * m.c[1][2] - m.c[1][1] * m.c[0][2]));
	m3 inv;
	inv.c[0][0] = +(m.c[1][1] * m.c[2][2] - m.c[2][1] * m.c[1][2]) * ood;
	inv.c[1][0] = -(m.c[1][0] * m.c[2][2] - m.c[2][0] * m.c[1][2]) * ood;
	inv.c[2][0] = +(m.c[1][0] * m.c[2][1] - m.c[2][0] * m.c[1][1]) * ood;
	inv.c[0][1] = -(m.c[0][1] * m.c[2][2] - m.c[2][1] * m.c[0][2]) * ood;
	inv.c[1][1] = +(m.c[0][0] * m.c[2][2] - m.c[2][0] * m.c[0][2]) * ood;
	inv.c[2][1] = -(m.c[0][0] * m.c[2][1] - m.c[2][0] * m.c[0][1]) * ood;
	inv.c[0][2] = +(m.c[0][1] * m.c[1][2] - m.c[1][1] * m.c[0][2]) * ood;
	inv.c[1][2] = -(m.c[0][0] * m.c[1][2] - m.c[1][0] * m.c[0][2]) * ood;
	inv.c[2][2] = +(m.c[0][0] * m.c[1][1] - m.c[1][0] * m.c[0][1]) * ood;
	return inv;
}

// ---------------------------------------------------------------------------------------------
// MaterialGGX (MaterialGGX.cpp:5-101)
// ---------------------------------------------------------------------------------------------
v3 ggxF(v3 F0, float VdotH) {
	return F0 + std::pow(1.0f - VdotH, 5.0f) * (v3(1.0f) - F0);
}
float ggxD(float NdotH, float alpha) {
	const float halfDenum = NdotH * NdotH * (alpha * alpha - 1.0f) + 1.0f;
	const float halfTerm = alpha / gmax(0.0001f, halfDenum);
	return halfTerm * halfTerm * kOneOverPi;
}
float ggxV(float NdotL, float NdotV, float alpha) {
	const float alpha2 = alpha * alpha;
	const float visL = NdotV * std::sqrt((-NdotL * alpha2 + NdotL) * NdotL + alpha2);
	const float visV = NdotL * std::sqrt((-NdotV * alpha2 + NdotV) * NdotV + alpha2);
	return 0.5f / gmax(0.0001f, visV + visL);
}
float alphaFromRoughness(float roughness) {
	const float rc = gmax(0.045f, roughness);
	return gmax(0.0001f, rc * rc);
}
v3 ggxBrdf(v3 wo, v3 base, float alpha, float metallic, v3 wi, float * pdf) {
	const v3 h = normalize(wi + wo);
	const float NdotH = gmax(h.z, 0.0f);
	const float VdotH = gmax(dot(wi, h), 0.0f);
	const float NdotL = gmax(wo.z, 0.0f);
	const float NdotV = gmax(wi.z, 0.0f);
	const float Dh = ggxD(NdotH, alpha);
	if(pdf) {
		const float hPdf = Dh * NdotH;
		*pdf = hPdf / (4.0f * gmax(0.0001f, VdotH));
	}
	const v3 F0 = mix3(v3(0.04f), base, metallic);
	const v3 specular = Dh * ggxV(NdotL, NdotV, alpha) * ggxF(F0, VdotH);
	const v3 diffuse = (1.0f - metallic) * kOneOverPi * base * (v3(1.0f) - F0);
	const v3 multiAdj = v3(1.0f) + (2.0f * alpha * alpha * NdotV) * F0;
	return (diffuse + specular * multiAdj) * NdotV;
}
// Random::sampleDisk / sampleCosineHemisphere (Random.cpp:41-70) on two given uniforms.
v2 sampleDisk(float u0, float u1) {
	const float x = 2.0f * u0 - 1.0f;
	const float y = 2.0f * u1 - 1.0f;
	if(x == 0.0f && y == 0.0f)
		return {0.0f, 0.0f};
	float angle, radius;
	if(std::abs(x) > std::abs(y)) {
		radius = x;
		angle = kQuarterPi * y / x;
	} else {
		radius = y;
		angle = kHalfPi - kQuarterPi * x / y;
	}
	return {radius * std::cos(angle), radius * std::sin(angle)};
}
v3 sampleCosineHemisphere(float u0, float u1) {
	const v2 xy = sampleDisk(u0, u1);
	const float z = std::sqrt(gmax(0.0f, 1.0f - xy.x * xy.x - xy.y * xy.y));
	return {xy.x, xy.y, z};
}
// u: [lobe choice, dir0, dir1]
v3 ggxSampleAndEval(v3 wo, v3 base, float roughness, float metallic, const float * u, v3 & wi, float * pdfOut = nullptr) {
	const float probaSpecular = mixf(1.0f / (dot(base, v3(1.0f)) / 3.0f + 1.0f), 1.0f, metallic);
	const float alpha = alphaFromRoughness(roughness);
	if(u[0] < probaSpecular) {
		const float a2 = alpha * alpha;
		const float x = u[1];
		const float phiH = u[2] * kTwoPi;
		const float cosThetaHSqr = gmin((1.0f - x) / ((a2 - 1.0f) * x + 1.0f), 1.0f);
		const float cosThetaH = std::sqrt(cosThetaHSqr);
		const float sinThetaH = std::sqrt(1.0f - cosThetaHSqr);
		const v3 lh(sinThetaH * std::cos(phiH), sinThetaH * std::sin(phiH), cosThetaH);
		wi = 2.0f * dot(wo, lh) * lh - wo;
	} else {
		wi = sampleCosineHemisphere(u[1], u[2]);
		if(wo.z < 0.0f)
			wi.z *= -1.0f;
	}
	if(wi.z < 0.0f)
		return v3(0.0f);
	float pdfSpec = 0.0f;
	const v3 brdf = ggxBrdf(wo, base, alpha, metallic, wi, &pdfSpec);
	const float pdf = mixf(kOneOverPi * gmax(wi.z, 0.0f), pdfSpec, probaSpecular);
	if(pdfOut)
		*pdfOut = pdf;
	if(pdf == 0.0f)
		return v3(0.0f);
	return brdf / pdf;
}
// Extension: the mixture density of sampleAndEval for a given direction (0 below the horizon).
float ggxPdf(v3 wo, v3 base, float roughness, float metallic, v3 wi) {
	if(wi.z < 0.0f)
		return 0.0f;
	const float probaSpecular = mixf(1.0f / (dot(base, v3(1.0f)) / 3.0f + 1.0f), 1.0f, metallic);
	float pdfSpec = 0.0f;
	ggxBrdf(wo, base, alphaFromRoughness(roughness), metallic, wi, &pdfSpec);
	return mixf(kOneOverPi * gmax(wi.z, 0.0f), pdfSpec, probaSpecular);
}
v3 ggxEval(v3 wo, v3 base, float roughness, float metallic, v3 wi) {
	if(wi.z < 0.0f)
		return v3(0.0f);
	return ggxBrdf(wo, base, alphaFromRoughness(roughness), metallic, wi, nullptr);
}

// ---------------------------------------------------------------------------------------------
// Backgrounds
// ---------------------------------------------------------------------------------------------
// Texture::sampleCubemap (Texture.cpp:141-190); faces px, nx, py, ny, pz, nz.
v3 sampleCubemap(const Scene & s, const int32_t * faces, v3 dir) {
	const v3 a(std::abs(dir.x), std::abs(dir.y), std::abs(dir.z));
	int side = 0;
	float x = 0.0f, y = 0.0f, denom = 1.0f;
	if(a.x >= a.y && a.x >= a.z) {
		denom = a.x;
		y = dir.y;
		if(dir.x >= 0.0f) {
			side = 0;
			x = -dir.z;
		} else {
			side = 1;
			x = dir.z;
		}
	} else if(a.y >= a.x && a.y >= a.z) {
		denom = a.y;
		x = dir.x;
		if(dir.y >= 0.0f) {
			side = 2;
			y = -dir.z;
		} else {
			side = 3;
			y = dir.z;
		}
	} else if(a.z >= a.x && a.z >= a.y) {
		denom = a.z;
		y = dir.y;
		if(dir.z >= 0.0f) {
			side = 4;
			x = dir.x;
		} else {
			side = 5;
			x = -dir.x;
		}
	}
	x = 0.5f * (x / denom) + 0.5f;
	y = 0.5f * (-y / denom) + 0.5f;
	const Texture & t = s.textures[size_t(faces[side])];
	const float eps = 1.0f / float(std::min(t.w, t.h));
	x = clampf(x, 0.0f + eps, 1.0f - eps);
	y = clampf(y, 0.0f + eps, 1.0f - eps);
	return t.rgbl(x, y);
}

// Intersection::sphere (Intersection.cpp:7-20).
bool sphereRoots(v3 o, v3 d, float radius, v2 & roots) {
	const float a = dot(d, d);
	const float b = dot(o, d);
	const float c = dot(o, o) - radius * radius;
	const float delta = b * b - a * c;
	if(delta < 0.0f)
		return false;
	const float ds = std::sqrt(delta);
	roots = {(-b - ds) / a, (-b + ds) / a};
	return true;
}

// MaterialSky::rayleighPhase / miePhase (MaterialSky.cpp:89-98), gMie = 0.758 (Sky.hpp:53-65).
float skyRayleighPhase(float cosAngle) {
	const float k = 1.0f / (4.0f * kPi);
	return k * 3.0f / 4.0f * (1.0f + cosAngle * cosAngle);
}
float skyMiePhase(float cosAngle) {
	const float gMie = 0.758f;
	const float k = 1.0f / (4.0f * kPi);
	const float g2 = gMie * gMie;
	return k * 3.0f * (1.0f - g2) / (2.0f * (2.0f + g2)) * (1.0f + cosAngle * cosAngle) /
		   std::pow(1.0f + g2 - 2.0f * gMie * cosAngle, 3.0f / 2.0f);
}

// MaterialSky::eval (MaterialSky.cpp:6-98) with Sky::AtmosphereParameters (Sky.hpp:53-65).
v3 skyEval(const Texture & table, v3 rayOrigin, v3 rayDir, v3 sunDir) {
	const v3 sunColor(1.474f, 1.8504f, 1.91198f);
	const v3 kRayleigh(5.5e-6f, 13.0e-6f, 22.4e-6f);
	const float groundRadius = 6371e3f, topRadius = 6471e3f, sunIntensity = 20.0f, kMie = 21e-6f;
	const float heightRayleigh = 8000.0f, heightMie = 1200.0f, sunRadius = 0.04675f;
	const float sunRadiusCos = 0.998f;
	const unsigned samplesCount = 16;
	const v3 planetPos = rayOrigin + v3(0.0f, groundRadius, 0.0f) + v3(0.0f, 1.0f, 0.0f);
	v2 interTop{0, 0}, interGround{0, 0};
	if(!sphereRoots(planetPos, rayDir, topRadius, interTop))
		return v3(0.0f);
	const bool didHitGround = sphereRoots(planetPos, rayDir, groundRadius, interGround);
	const float distanceToInter = gmin(interTop.y, didHitGround ? interGround.x : 0.0f);
	const float stepSize = (distanceToInter - interTop.x) / float(samplesCount);
	const float cosViewSun = dot(rayDir, sunDir);
	float rayleighDist = 0.0f, mieDist = 0.0f;
	v3 rayleighScatt(0.0f), mieScatt(0.0f), transmittance(0.0f);
	for(unsigned i = 0; i < samplesCount; ++i) {
		const v3 currPos = planetPos + ((float(i) + 0.5f) * stepSize) * rayDir;
		float currHeight = length(currPos) - groundRadius;
		if(i == (samplesCount - 1) && currHeight < 0.0f)
			currHeight = 0.0f;
		const float rayleighStep = std::exp(-currHeight / heightRayleigh) * stepSize;
		const float mieStep = std::exp(-currHeight / heightMie) * stepSize;
		rayleighDist += rayleighStep;
		mieDist += mieStep;
		const v3 ex = -(mieDist * v3(kMie) + rayleighDist * kRayleigh);
		const v3 directAttenuation(std::exp(ex.x), std::exp(ex.y), std::exp(ex.z));
		const float relativeHeight = (length(currPos) - groundRadius) / (topRadius - groundRadius);
		const float relativeCosAngle = -0.5f * sunDir.y + 0.5f;
		float ux = (511.0f / 512.0f) * relativeHeight + 0.5f / 512.0f;
		float uy = (511.0f / 512.0f) * relativeCosAngle + 0.5f / 512.0f;
		ux = clampf(ux, 0.5f / 512.0f, 511.0f / 512.0f);
		uy = clampf(uy, 0.5f / 512.0f, 511.0f / 512.0f);
		const v3 secondary = table.rgbl(ux, uy);
		const v3 attenuation = directAttenuation * secondary;
		rayleighScatt = rayleighScatt + rayleighStep * attenuation;
		mieScatt = mieScatt + mieStep * attenuation;
		transmittance = transmittance + directAttenuation;
	}
	const v3 rayleighPart = skyRayleighPhase(cosViewSun) * kRayleigh * rayleighScatt;
	const v3 miePart = kMie * skyMiePhase(cosViewSun) * mieScatt;
	v3 sunRadiance(0.0f);
	const bool didHitGroundForward = didHitGround && interGround.y > 0.0f;
	if(!didHitGroundForward && dot(rayDir, sunDir) > sunRadiusCos)
		sunRadiance = sunColor / (kPi * sunRadius * sunRadius);
	return sunIntensity * (rayleighPart + miePart) + transmittance * sunRadiance;
}

// Extension: latitude-longitude environment map, u along the azimuth from -z towards +x,
// v from +y (top row) to -y.
v3 sampleLatLong(const Texture & t, v3 d) {
	const float u = 0.5f + std::atan2(d.x, -d.z) * (0.5f * kOneOverPi);
	const float v = std::acos(clampf(d.y, -1.0f, 1.0f)) * kOneOverPi;
	return t.rgbl(u, v);
}

// ---------------------------------------------------------------------------------------------
// Extension: area-light sampling (extended estimator)
// ---------------------------------------------------------------------------------------------
inline double lumd(double r, double g, double b) { return 0.2126 * r + 0.7152 * g + 0.0722 * b; }
inline float lumf(v3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
constexpr float kTwoPiSq = 19.7392088021787171f; // 2 pi^2

// Largest k in [0, n) with cdf[k] <= u (cdf: n + 1 entries from 0 to 1, u in [0, 1)).
int cdfSearch(const float * cdf, int n, float u) {
	int lo = 0, hi = n;
	while(hi - lo > 1) {
		const int mid = (lo + hi) >> 1;
		if(cdf[mid] <= u)
			lo = mid;
		else
			hi = mid;
	}
	return lo;
}

// Tables: lat-long texel weights = the largest luminance of the texel's bilinear footprint times
// sin(theta) at the texel centre; emissive triangles weighted by area times the mean luminance of
// their emission texture.
void buildAreaLights(Scene & s) {
	AreaLights & A = s.area;
	A = AreaLights{};
	double envPower = 0.0;
	if(s.bg.mode == RDU_BACKGROUND_LATLONG) {
		const Texture & t = s.textures[size_t(s.bg.texture)];
		const int W = t.w, H = t.h;
		A.envW = W;
		A.envH = H;
		std::vector<double> f(size_t(W) * size_t(H)), rows(static_cast<size_t>(H));
		double total = 0.0;
		for(int j = 0; j < H; ++j) {
			const double sinT = std::sin(3.14159265358979323846 * (double(j) + 0.5) / double(H));
			double row = 0.0;
			for(int i = 0; i < W; ++i) {
				double m = 0.0;
				for(int dj = 0; dj < 2; ++dj)
					for(int di = 0; di < 2; ++di) {
						const v4 c = t.at((i + di) % W, (j + dj) % H);
						m = std::max(m, lumd(c.x, c.y, c.z));
					}
				f[size_t(j) * W + i] = m * sinT;
				row += m * sinT;
			}
			rows[size_t(j)] = row;
			total += row;
		}
		if(total > 0.0) {
			A.envPmf.resize(size_t(W) * H);
			A.envRowCdf.resize(size_t(H) + 1);
			A.envColCdf.resize(size_t(H) * (W + 1));
			double cum = 0.0;
			A.envRowCdf[0] = 0.0f;
			for(int j = 0; j < H; ++j) {
				cum += rows[size_t(j)];
				A.envRowCdf[size_t(j) + 1] = j + 1 == H ? 1.0f : float(cum / total);
				float * cc = &A.envColCdf[size_t(j) * (W + 1)];
				cc[0] = 0.0f;
				double c2 = 0.0;
				for(int i = 0; i < W; ++i) {
					c2 += f[size_t(j) * W + i];
					cc[i + 1] = i + 1 == W ? 1.0f : (rows[size_t(j)] > 0.0 ? float(c2 / rows[size_t(j)]) : float(double(i + 1) / W));
					A.envPmf[size_t(j) * W + i] = float(f[size_t(j) * W + i] / total);
				}
			}
			envPower = total;
		} else {
			A.envW = A.envH = 0;
		}
	}
	// Emissive triangles, objects in order, triangles in mesh order; world vertices as the
	// raycaster transformed them (objects were added in order, so object o's vertices follow
	// those of the objects before it).
	A.objectBase.assign(s.objects.size(), -1);
	std::vector<double> w;
	size_t vbase = 0;
	for(size_t o = 0; o < s.objects.size(); ++o) {
		const Object & obj = s.objects[o];
		const MeshData & m = s.meshes[obj.mesh];
		const size_t nv = s.meshPos[obj.mesh].size() / 3;
		const rdu_material_desc & mat = s.materials[obj.mat];
		if(mat.type == RDU_MATERIAL_EMISSIVE) {
			const Texture & t = s.textures[size_t(mat.textures[0])];
			double mean = 0.0;
			for(int y = 0; y < t.h; ++y)
				for(int x = 0; x < t.w; ++x) {
					const v4 c = t.at(x, y);
					mean += lumd(c.x, c.y, c.z);
				}
			mean /= double(t.w) * double(t.h);
			A.objectBase[o] = int32_t(A.tris.size());
			for(size_t k = 0; k + 2 < m.idx.size(); k += 3) {
				AreaLights::Tri tr;
				tr.v0 = s.rc.vertices[vbase + m.idx[k]];
				const v3 v1 = s.rc.vertices[vbase + m.idx[k + 1]];
				const v3 v2 = s.rc.vertices[vbase + m.idx[k + 2]];
				tr.e1 = v1 - tr.v0;
				tr.e2 = v2 - tr.v0;
				tr.n = cross(tr.e1, tr.e2);
				const double nx = tr.n.x, ny = tr.n.y, nz = tr.n.z;
				const double area = 0.5 * std::sqrt(nx * nx + ny * ny + nz * nz);
				tr.area = float(area);
				tr.object = uint32_t(o);
				tr.localId = uint32_t(k);
				A.tris.push_back(tr);
				w.push_back(area * mean);
			}
		}
		vbase += nv;
	}
	double triPower = 0.0;
	for(double x : w)
		triPower += x;
	if(triPower > 0.0) {
		A.triPmf.resize(w.size());
		A.triCdf.resize(w.size() + 1);
		double cum = 0.0;
		A.triCdf[0] = 0.0f;
		for(size_t k = 0; k < w.size(); ++k) {
			cum += w[k];
			A.triPmf[k] = float(w[k] / triPower);
			A.triCdf[k + 1] = k + 1 == w.size() ? 1.0f : float(cum / triPower);
		}
	} else {
		A.tris.clear();
		A.objectBase.assign(s.objects.size(), -1);
	}
	const bool env = envPower > 0.0, tri = !A.tris.empty();
	A.pEnv = env ? (tri ? 0.5f : 1.0f) : 0.0f;
	A.pArea = (env || tri) ? (s.lights.empty() ? 1.0f : 0.5f) : 0.0f;
}

// Density (solid angle) of the environment strategy for direction d.
float envPdf(const Scene & s, v3 d) {
	const AreaLights & A = s.area;
	if(A.envW == 0)
		return 0.0f;
	const float u = 0.5f + std::atan2(d.x, -d.z) * (0.5f * kOneOverPi);
	const float v = std::acos(clampf(d.y, -1.0f, 1.0f)) * kOneOverPi;
	int i = int(u * float(A.envW)), j = int(v * float(A.envH));
	i = i < 0 ? 0 : (i >= A.envW ? A.envW - 1 : i);
	j = j < 0 ? 0 : (j >= A.envH ? A.envH - 1 : j);
	const float sinT = std::sqrt(gmax(0.0f, 1.0f - d.y * d.y));
	if(!(sinT > 0.0f))
		return 0.0f;
	return A.envPmf[size_t(j) * A.envW + i] * float(A.envW * A.envH) / (kTwoPiSq * sinT);
}

// One direction towards the environment; pdf = the density it was drawn with.
v3 envSample(const Scene & s, float u1, float u2, float & pdf) {
	const AreaLights & A = s.area;
	const int j = cdfSearch(A.envRowCdf.data(), A.envH, u1);
	const float a = A.envRowCdf[size_t(j)], b = A.envRowCdf[size_t(j) + 1];
	const float dv = (u1 - a) / (b - a);
	const float * cc = &A.envColCdf[size_t(j) * (A.envW + 1)];
	const int i = cdfSearch(cc, A.envW, u2);
	const float du = (u2 - cc[i]) / (cc[i + 1] - cc[i]);
	const float v = (float(j) + dv) / float(A.envH), u = (float(i) + du) / float(A.envW);
	const float theta = v * kPi, phi = (u - 0.5f) * kTwoPi;
	const float sinT = std::sin(theta), cosT = std::cos(theta);
	pdf = sinT > 0.0f ? A.envPmf[size_t(j) * A.envW + i] * float(A.envW * A.envH) / (kTwoPiSq * sinT) : 0.0f;
	return {sinT * std::sin(phi), cosT, -sinT * std::cos(phi)};
}

// Density (solid angle, from distance dist along a ray of direction dir) of the triangle strategy
// for a point on emissive triangle t.
float triPdf(const AreaLights & A, int t, float dist, v3 dir) {
	const AreaLights::Tri & T = A.tris[size_t(t)];
	const float cosL = std::abs(dot(T.n, dir)) / length(T.n);
	return A.triPmf[size_t(t)] / T.area * (dist * dist) / cosL;
}

// Power heuristic weight of a strategy with density p against one with density q.
inline float powerWeight(float p, float q) {
	const float r = q / p;
	return 1.0f / (1.0f + r * r);
}

// PathTracer::evalBackground (PathTracer.cpp:23-55).
v3 evalBackground(const Scene & s, v3 rayDir, v3 rayPos, v2 ndc, bool direct) {
	const rdu_background_desc & b = s.bg;
	if(direct) {
		if(b.mode == RDU_BACKGROUND_IMAGE)
			return s.textures[size_t(b.texture)].rgbl(ndc.x, ndc.y);
		if(b.mode == RDU_BACKGROUND_SKYBOX)
			return sampleCubemap(s, b.cube, normalize(rayDir));
		if(b.mode == RDU_BACKGROUND_ATMOSPHERE)
			return skyEval(s.textures[size_t(b.texture)], rayPos, normalize(rayDir),
				v3(b.sun_direction[0], b.sun_direction[1], b.sun_direction[2]));
		if(b.mode == RDU_BACKGROUND_LATLONG)
			return sampleLatLong(s.textures[size_t(b.texture)], normalize(rayDir));
		return v3(b.color[0], b.color[1], b.color[2]);
	}
	if(b.mode == RDU_BACKGROUND_SKYBOX)
		return sampleCubemap(s, b.cube, normalize(rayDir));
	if(b.mode == RDU_BACKGROUND_ATMOSPHERE)
		return skyEval(s.textures[size_t(b.texture)], rayPos, normalize(rayDir),
			v3(b.sun_direction[0], b.sun_direction[1], b.sun_direction[2]));
	if(b.mode == RDU_BACKGROUND_LATLONG)
		return sampleLatLong(s.textures[size_t(b.texture)], normalize(rayDir));
	return v3(0.0f);
}

// Light::sample for the three analytic lights. Returns the direction towards the light.
v3 lightSample(const rdu_light_desc & L, v3 position, float & dist, float & attenuation) {
	if(L.type == RDU_LIGHT_DIRECTIONAL) { // DirectionalLight.cpp:40-44
		attenuation = 1.0f;
		dist = std::numeric_limits<float>::max();
		return -v3(L.direction[0], L.direction[1], L.direction[2]);
	}
	v3 direction = v3(L.position[0], L.position[1], L.position[2]) - position;
	dist = length(direction);
	attenuation = 0.0f;
	if(dist > L.radius)
		return v3(0.0f);
	if(dist > 0.0f)
		direction = direction / dist;
	float angleAttenuation = 1.0f;
	if(L.type == RDU_LIGHT_SPOT) { // SpotLight.cpp:52-81, including its cos(cos(outer)) cone test
		const v3 ldir(L.direction[0], L.direction[1], L.direction[2]);
		const float currentCos = dot(-direction, ldir);
		const float outerCos = std::cos(L.angles[1]);
		if(currentCos < std::cos(outerCos))
			return v3(0.0f);
		const float innerCos = std::cos(L.angles[0]);
		angleAttenuation = clampf((currentCos - outerCos) / (innerCos - outerCos), 0.0f, 1.0f);
	}
	const float radiusRatio = dist / L.radius;
	const float radiusRatio2 = radiusRatio * radiusRatio;
	const float attenNum = clampf(1.0f - radiusRatio2, 0.0f, 1.0f);
	attenuation = L.type == RDU_LIGHT_SPOT ? angleAttenuation * attenNum * attenNum : attenNum * attenNum;
	return direction;
}

v3 interp3(const std::vector<v3> & a, const std::vector<uint32_t> & idx, const Hit & h) {
	const uint32_t i0 = idx[h.localId], i1 = idx[h.localId + 1], i2 = idx[h.localId + 2];
	return h.w * a[i0] + h.u * a[i1] + h.v * a[i2];
}
v2 interp2(const std::vector<v2> & a, const std::vector<uint32_t> & idx, const Hit & h) {
	const uint32_t i0 = idx[h.localId], i1 = idx[h.localId + 1], i2 = idx[h.localId + 2];
	return {h.w * a[i0].x + h.u * a[i1].x + h.v * a[i2].x, h.w * a[i0].y + h.u * a[i1].y + h.v * a[i2].y};
}

// PathTracer::buildLocalFrame (PathTracer.cpp:86-122).
m3 buildLocalFrame(const Scene & s, const Object & obj, const Hit & hit, v3 rayDir, v2 uv) {
	const MeshData & mesh = s.meshes[obj.mesh];
	const v3 n = normalize(interp3(mesh.nrm, mesh.idx, hit));
	const v3 t = normalize(interp3(mesh.tan, mesh.idx, hit));
	const v3 b = normalize(interp3(mesh.bit, mesh.idx, hit));
	m3 tbn;
	tbn.c[0] = mul(obj.invtp, t);
	tbn.c[1] = mul(obj.invtp, b);
	tbn.c[2] = mul(obj.invtp, n);
	const bool frontFacing = dot(tbn.c[2], rayDir) < 0.0f;
	if(!frontFacing)
		tbn.c[2] = tbn.c[2] * -1.0f;
	const rdu_material_desc & mat = s.materials[obj.mat];
	if(obj.useUV && mat.type != RDU_MATERIAL_EMISSIVE) {
		const v4 img = s.textures[size_t(mat.textures[1])].rgbal(uv.x, uv.y);
		const v3 localNormal = normalize(2.0f * v3(img.x, img.y, img.z) - v3(1.0f));
		tbn.c[2] = mul(tbn, localNormal);
	}
	tbn.c[2] = normalize(tbn.c[2]);
	tbn.c[1] = normalize(cross(tbn.c[2], normalize(tbn.c[0])));
	tbn.c[0] = normalize(cross(tbn.c[1], tbn.c[2]));
	return tbn;
}

// PathTracer::checkVisibility (PathTracer.cpp:124-157). `mini` is Rendu's 0.001; the extension's
// area-light rays pass the 1e-4 of the path's own rays.
bool checkVisibility(const Scene & s, v3 startPos, v3 rayDir, float maxDist, uint64_t & rays, float mini = 0.001f) {
	v3 lpos = startPos;
	while(maxDist > 0.0f) {
		++rays;
		const Hit lhit = s.rc.traverse(lpos, rayDir, mini, maxDist, false, nullptr);
		if(!lhit.hit)
			return true;
		const Object & lobj = s.objects[lhit.meshId];
		const rdu_material_desc & lmat = s.materials[lobj.mat];
		if(!lmat.masked || !lobj.useUV)
			return false;
		const MeshData & lmesh = s.meshes[lobj.mesh];
		const v2 luv = interp2(lmesh.uv, lmesh.idx, lhit);
		const float alpha = s.textures[size_t(lmat.textures[0])].rgbal(luv.x, luv.y).w;
		if(alpha < 0.01f) {
			maxDist = maxDist - lhit.dist;
			lpos = lpos + lhit.dist * rayDir;
		} else {
			return false;
		}
	}
	return true;
}

struct RenderCounters {
	std::atomic<uint64_t> ext{0}, shadow{0};
};

// One path of PathTracer::render (PathTracer.cpp:191-288) for pixel (x, y), sample sid.
v3 tracePath(const Scene & s, const rdu_camera_desc & cam, v3 corner, v3 dx, v3 dy, int cellX, v2 cellSize,
	uint32_t x, uint32_t y, uint32_t w, uint32_t h, uint32_t sid, size_t depth, const rdu_render_opts & opt,
	uint64_t & ext, uint64_t & shadow) {
	const PathRng rng{opt.seed, y * w + x, sid};
	float u[4];
	rng.draw4(0, u);
	// getSamplePosition (PathTracer.cpp:71-84)
	const int sidy = int(sid) / cellX;
	const int sidx = int(sid) % cellX;
	const float gx = float(sidx) + u[0], gy = float(sidy) + u[1];
	const v2 screen{float(x) + gx * cellSize.x, float(y) + gy * cellSize.y};
	const v2 ndc{screen.x / float(w), screen.y / float(h)};
	const v3 eye(cam.position[0], cam.position[1], cam.position[2]);
	const v3 worldPos = corner + ndc.x * dx + ndc.y * dy;
	v3 rayPos = eye;
	v3 rayDir = normalize(worldPos - eye);
	v3 sampleColor(0.0f);
	v3 attenuation(1.0f);
	const size_t nL = s.lights.size();
	// Extension: area lights and multiple importance sampling (extended estimator only).
	const AreaLights & A = s.area;
	const bool extended = opt.estimator == RDU_ESTIMATOR_EXTENDED;
	const float pArea = extended ? A.pArea : 0.0f;
	float prevPdf = 0.0f; // BSDF density of the last scattering direction (0: no MIS weight)
	float misDist = 0.0f; // distance travelled since that scattering vertex, through cut-outs
	for(size_t did = 0; did < depth; ++did) {
		++ext;
		const Hit hit = s.rc.traverse(rayPos, rayDir, 0.0001f, 1e8f, false, nullptr);
		if(!hit.hit) {
			v3 bg = evalBackground(s, rayDir, rayPos, ndc, did == 0);
			if(prevPdf > 0.0f && A.pEnv > 0.0f)
				bg = powerWeight(prevPdf, pArea * A.pEnv * envPdf(s, normalize(rayDir))) * bg;
			sampleColor = sampleColor + attenuation * bg;
			break;
		}
		const Object & obj = s.objects[hit.meshId];
		const MeshData & mesh = s.meshes[obj.mesh];
		const v3 p = rayPos + hit.dist * rayDir;
		const bool noUVs = !obj.useUV;
		const v2 uv = noUVs ? v2{0.5f, 0.5f} : interp2(mesh.uv, mesh.idx, hit);
		const rdu_material_desc & mat = s.materials[obj.mat];
		const v4 bCol = s.textures[size_t(mat.textures[0])].rgbal(uv.x, uv.y);
		if(mat.masked && bCol.w < 0.01f) {
			rayPos = p;
			misDist = misDist + hit.dist;
			continue;
		}
		if(mat.type == RDU_MATERIAL_EMISSIVE) {
			v3 le(bCol.x, bCol.y, bCol.z);
			if(prevPdf > 0.0f && A.pEnv < 1.0f && A.objectBase[hit.meshId] >= 0) {
				const int t = A.objectBase[hit.meshId] + int(hit.localId / 3);
				le = powerWeight(prevPdf, pArea * (1.0f - A.pEnv) * triPdf(A, t, misDist + hit.dist, rayDir)) * le;
			}
			sampleColor = sampleColor + attenuation * le;
			break;
		}
		const m3 tbn = buildLocalFrame(s, obj, hit, rayDir, uv);
		const m3 itbn = transpose(tbn);
		const v3 wo = normalize(mul(itbn, -rayDir));
		const v3 baseColor(std::pow(bCol.x, 2.2f), std::pow(bCol.y, 2.2f), std::pow(bCol.z, 2.2f));
		const v4 rmao = s.textures[size_t(mat.textures[2])].rgbal(uv.x, uv.y);
		rng.draw4(1 + uint32_t(did), u);
		const v3 pShift = p + 0.001f * tbn.c[2];
		// Area lights are sampled at every vertex but the last: there, the reference's path ends
		// without tracing a next segment, so light one segment further never counts (Rendu's own
		// analytic lights are sampled at the last vertex too, as the reference does). The pick
		// is the same at every vertex, so the analytic lights keep their probability.
		const bool areaHere = pArea > 0.0f && did + 1 < depth;
		if(pArea > 0.0f && u[0] < pArea) {
			if(areaHere) {
				// Extension: one sample of the environment or of an emissive triangle, weighted
				// against BSDF sampling by the power heuristic. Its visibility ray leaves from the hit
				// point with the interval of the path's own next ray (1e-4, 1e8), so that a light
				// point counts as visible exactly when a BSDF-sampled ray towards it would reach it
				// (with shading normals, an offset origin would see past surfaces that the path's
				// rays hit, and the two strategies would estimate different integrals).
				float e[4];
				rng.draw4(0x20000u + uint32_t(did), e);
				v3 direction, le;
				float pLight = 0.0f, pMis = 0.0f, maxDist = 0.0f;
				if(e[0] < A.pEnv) {
					direction = envSample(s, e[1], e[2], pLight);
					pLight = pArea * A.pEnv * pLight;
					pMis = pArea * A.pEnv * envPdf(s, direction);
					le = sampleLatLong(s.textures[size_t(s.bg.texture)], normalize(direction));
					maxDist = 1e8f;
				} else {
					const float ut = (e[0] - A.pEnv) / (1.0f - A.pEnv);
					const int t = cdfSearch(A.triCdf.data(), int(A.tris.size()), ut);
					const AreaLights::Tri & T = A.tris[size_t(t)];
					const float su = std::sqrt(e[1]);
					const float bu = su * (1.0f - e[2]), bv = su * e[2], bw = 1.0f - bu - bv;
					const v3 y = T.v0 + bu * T.e1 + bv * T.e2;
					const Object & lo = s.objects[T.object];
					const MeshData & lm = s.meshes[lo.mesh];
					v2 luv{0.5f, 0.5f};
					if(lo.useUV) {
						const uint32_t i0 = lm.idx[T.localId], i1 = lm.idx[T.localId + 1], i2 = lm.idx[T.localId + 2];
						luv = {bw * lm.uv[i0].x + bu * lm.uv[i1].x + bv * lm.uv[i2].x, bw * lm.uv[i0].y + bu * lm.uv[i1].y + bv * lm.uv[i2].y};
					}
					const v4 ec = s.textures[size_t(s.materials[lo.mat].textures[0])].rgbal(luv.x, luv.y);
					le = v3(ec.x, ec.y, ec.z);
					const v3 toLight = y - p;
					const float dist = length(toLight);
					direction = toLight / dist;
					pLight = pArea * (1.0f - A.pEnv) * triPdf(A, t, dist, direction);
					pMis = pLight;
					maxDist = dist * 0.999f;
				}
				if(pLight > 0.0f && pLight < std::numeric_limits<float>::infinity()) {
					const v3 lwi = normalize(mul(itbn, direction));
					const v3 evalLight = ggxEval(wo, baseColor, rmao.x, rmao.y, lwi);
					const float w = powerWeight(pMis, ggxPdf(wo, baseColor, rmao.x, rmao.y, lwi));
					const v3 illumination = evalLight * le * (w / pLight);
					if((illumination.x > 0.0f || illumination.y > 0.0f || illumination.z > 0.0f) &&
						checkVisibility(s, p, direction, maxDist, shadow, 0.0001f))
						sampleColor = sampleColor + attenuation * illumination;
				}
			}
		} else if(nL > 0) {
			float ul = u[0];
			float lightPdf = 1.0f / float(nL);
			if(pArea > 0.0f) { // extension: the analytic lights share the pick with the area lights
				ul = (u[0] - pArea) / (1.0f - pArea);
				lightPdf = (1.0f - pArea) / float(nL);
			}
			const size_t lid = std::min(size_t(ul * float(nL)), nL - 1); // Random::Int(0, nL-1)
			const rdu_light_desc & light = s.lights[lid];
			float maxDist, falloff;
			const v3 direction = lightSample(light, pShift, maxDist, falloff);
			bool visible = falloff > 0.0f;
			if(visible && light.casts_shadow) {
				visible = checkVisibility(s, pShift, direction, maxDist, shadow);
			}
			if(visible) {
				const v3 lwi = normalize(mul(itbn, direction));
				const v3 evalLight = ggxEval(wo, baseColor, rmao.x, rmao.y, lwi);
				const v3 illumination = falloff * evalLight * v3(light.intensity[0], light.intensity[1], light.intensity[2]) / lightPdf;
				sampleColor = sampleColor + attenuation * illumination;
			}
		}
		v3 wi;
		float bsdfPdf = 0.0f;
		const v3 eval = ggxSampleAndEval(wo, baseColor, rmao.x, rmao.y, u + 1, wi, &bsdfPdf);
		const v3 nextRayDir = normalize(mul(tbn, wi));
		attenuation = attenuation * eval;
		prevPdf = pArea > 0.0f ? bsdfPdf : 0.0f;
		misDist = 0.0f;
		if(opt.estimator == RDU_ESTIMATOR_EXTENDED && int(did) + 1 >= opt.rr_start && did + 1 < depth) {
			// Russian roulette (extension): survive with p = min(max component, 0.95), reweight by 1/p.
			const float pmax = gmax(attenuation.x, gmax(attenuation.y, attenuation.z));
			const float ps = gmin(pmax, 0.95f);
			float rr[4];
			rng.draw4(0x10000u + uint32_t(did), rr);
			if(!(rr[0] < ps))
				break;
			attenuation = attenuation / ps;
		}
		if(did < depth - 1) {
			rayPos = p;
			rayDir = normalize(nextRayDir);
		}
	}
	return sampleColor;
}

} // namespace

extern "C" {

void * orc_pt_create() { return new Scene(); }
void orc_pt_destroy(void * h) { delete static_cast<Scene *>(h); }

int orc_pt_add_texture(void * h, const rdu_texture_desc * d) {
	Scene * s = static_cast<Scene *>(h);
	Texture t;
	t.w = int(d->width);
	t.h = int(d->height);
	t.px.assign(d->rgba, d->rgba + size_t(d->width) * d->height * 4);
	s->textures.push_back(std::move(t));
	return int(s->textures.size()) - 1;
}

int orc_pt_add_mesh(void * h, const rdu_mesh_desc * d) {
	Scene * s = static_cast<Scene *>(h);
	MeshData m;
	const size_t nv = d->n_vertices;
	for(size_t i = 0; i < nv; ++i) {
		m.nrm.emplace_back(d->normals[3 * i], d->normals[3 * i + 1], d->normals[3 * i + 2]);
		m.tan.emplace_back(d->tangents[3 * i], d->tangents[3 * i + 1], d->tangents[3 * i + 2]);
		m.bit.emplace_back(d->bitangents[3 * i], d->bitangents[3 * i + 1], d->bitangents[3 * i + 2]);
		if(d->texcoords)
			m.uv.push_back({d->texcoords[2 * i], d->texcoords[2 * i + 1]});
	}
	m.hasUV = d->texcoords != nullptr;
	m.idx.assign(d->indices, d->indices + d->n_indices);
	s->meshes.push_back(std::move(m));
	s->meshPos.emplace_back(d->positions, d->positions + nv * 3);
	return int(s->meshes.size()) - 1;
}

int orc_pt_add_material(void * h, const rdu_material_desc * d) {
	Scene * s = static_cast<Scene *>(h);
	s->materials.push_back(*d);
	return int(s->materials.size()) - 1;
}

int orc_pt_add_object(void * h, const rdu_object_desc * d) {
	Scene * s = static_cast<Scene *>(h);
	Object o;
	o.mesh = d->mesh;
	o.mat = d->material;
	o.useUV = d->use_texcoords != 0 && s->meshes[d->mesh].hasUV;
	std::memcpy(o.model, d->model, sizeof(o.model));
	o.invtp = inverseTransposeUpper(d->model);
	s->objects.push_back(o);
	return 0;
}

int orc_pt_add_light(void * h, const rdu_light_desc * d) {
	static_cast<Scene *>(h)->lights.push_back(*d);
	return 0;
}

int orc_pt_set_background(void * h, const rdu_background_desc * d) {
	static_cast<Scene *>(h)->bg = *d;
	return 0;
}

// PathTracer::PathTracer (PathTracer.cpp:11-21).
int orc_pt_commit(void * h) {
	Scene * s = static_cast<Scene *>(h);
	for(const Object & o : s->objects) {
		const MeshData & m = s->meshes[o.mesh];
		const std::vector<float> & p = s->meshPos[o.mesh];
		s->rc.addMesh(p.data(), p.size() / 3, m.idx.data(), m.idx.size(), o.model);
	}
	for(size_t m = 0; m < s->rc.meshCount; ++m)
		if(s->rc.hierarchy[m].right == 0)
			return -2;
	s->rc.updateHierarchy();
	buildAreaLights(*s);
	s->committed = true;
	return 0;
}

// Camera::pose + Camera::pixelShifts (Camera.cpp:10-18, 42-49).
void orc_camera_shifts(const rdu_camera_desc * c, float * corner, float * dx, float * dy) {
	const v3 eye(c->position[0], c->position[1], c->position[2]);
	const v3 center(c->center[0], c->center[1], c->center[2]);
	v3 up = normalize(v3(c->up[0], c->up[1], c->up[2]));
	const v3 viewDir = normalize(center - eye);
	const v3 right = cross(viewDir, up);
	up = cross(right, viewDir);
	const float heightScale = std::tan(0.5f * c->fov);
	const float widthScale = c->ratio * heightScale;
	const float imageDist = length(center - eye); // glm::distance(eye, center) = length(center - eye)
	const v3 cr = center + imageDist * ((-widthScale) * right + heightScale * up);
	const v3 ddx = ((2.0f * widthScale) * imageDist) * right;
	const v3 ddy = ((-2.0f * heightScale) * imageDist) * up;
	for(int i = 0; i < 3; ++i) {
		corner[i] = cr[i];
		dx[i] = ddx[i];
		dy[i] = ddy[i];
	}
}

// PathTracer::render (PathTracer.cpp:159-299). Returns the spp used, or < 0 on error.
// counters (nullable): [extension rays, shadow rays].
int orc_pt_render(void * h, const rdu_camera_desc * cam, size_t samples, size_t depth, float * image, uint32_t w,
	uint32_t hgt, uint32_t components, const rdu_render_opts * opts, int threads, uint64_t * counters) {
	Scene * s = static_cast<Scene *>(h);
	if(!s->committed || components < 3 || samples == 0)
		return -1;
	rdu_render_opts opt{};
	if(opts)
		opt = *opts;
	samples = size_t(std::pow(2.0, std::round(std::log2(float(samples)))));
	float c3[3], dx3[3], dy3[3];
	orc_camera_shifts(cam, c3, dx3, dy3);
	const v3 corner(c3[0], c3[1], c3[2]), dx(dx3[0], dx3[1], dx3[2]), dy(dy3[0], dy3[1], dy3[2]);
	// getSampleGrid (PathTracer.cpp:57-69)
	const int k = int(std::floor(std::log2(samples)));
	int cx = int(std::pow(2, k / 2)), cy = cx;
	if(k % 2 == 1) {
		cx = int(std::pow(2, (k + 1) / 2));
		cy = int(std::pow(2, (k - 1) / 2));
	}
	const v2 cellSize{1.0f / float(cx), 1.0f / float(cy)};
	std::atomic<uint64_t> ext(0), shadow(0);
	forParallel(0, hgt, threads, [&](size_t y) {
		uint64_t e = 0, sh = 0;
		for(uint32_t x = 0; x < w; ++x) {
			float * px = image + (size_t(y) * w + x) * components;
			for(uint32_t sid = 0; sid < samples; ++sid) {
				const v3 c = tracePath(*s, *cam, corner, dx, dy, cx, cellSize, x, uint32_t(y), w, hgt, sid, depth, opt, e, sh);
				px[0] += gmin(c.x, 5.0f);
				px[1] += gmin(c.y, 5.0f);
				px[2] += gmin(c.z, 5.0f);
			}
			px[0] = px[0] / float(samples);
			px[1] = px[1] / float(samples);
			px[2] = px[2] / float(samples);
		}
		ext += e;
		shadow += sh;
	});
	if(counters) {
		counters[0] = ext.load();
		counters[1] = shadow.load();
	}
	return int(samples);
}

// Single-function probes for the closed-form tests.
void orc_ggx_eval(const float * wo, const float * base, float rough, float metal, const float * wi, float * out) {
	const v3 r = ggxEval(v3(wo[0], wo[1], wo[2]), v3(base[0], base[1], base[2]), rough, metal, v3(wi[0], wi[1], wi[2]));
	out[0] = r.x;
	out[1] = r.y;
	out[2] = r.z;
}
void orc_ggx_sample(const float * wo, const float * base, float rough, float metal, const float * u, float * wi, float * weight) {
	v3 w;
	const v3 r = ggxSampleAndEval(v3(wo[0], wo[1], wo[2]), v3(base[0], base[1], base[2]), rough, metal, u, w);
	wi[0] = w.x;
	wi[1] = w.y;
	wi[2] = w.z;
	weight[0] = r.x;
	weight[1] = r.y;
	weight[2] = r.z;
}
void orc_ggx_pdf(const float * wo, const float * base, float rough, float metal, const float * wi, float * pdf) {
	// Mixture pdf of sampleAndEval (MaterialGGX.cpp:58, 83) for a given direction.
	const v3 b(base[0], base[1], base[2]);
	const float probaSpecular = mixf(1.0f / (dot(b, v3(1.0f)) / 3.0f + 1.0f), 1.0f, metal);
	float pdfSpec = 0.0f;
	const v3 wiv(wi[0], wi[1], wi[2]);
	ggxBrdf(v3(wo[0], wo[1], wo[2]), b, alphaFromRoughness(rough), metal, wiv, &pdfSpec);
	*pdf = mixf(kOneOverPi * gmax(wiv.z, 0.0f), pdfSpec, probaSpecular);
}
void orc_sample_disk(float u0, float u1, float * out) {
	const v2 d = sampleDisk(u0, u1);
	out[0] = d.x;
	out[1] = d.y;
}
void orc_texture_probe(void * h, int tex, int mode, float x, float y, float z, float * out) {
	// mode 0: rgbal(x, y); 1: cubemap(dir) with faces tex..tex+5; 2: lat-long(dir)
	const Scene * s = static_cast<const Scene *>(h);
	if(mode == 0) {
		const v4 c = s->textures[size_t(tex)].rgbal(x, y);
		out[0] = c.x;
		out[1] = c.y;
		out[2] = c.z;
		out[3] = c.w;
		return;
	}
	v3 c;
	if(mode == 1) {
		const int32_t faces[6] = {tex, tex + 1, tex + 2, tex + 3, tex + 4, tex + 5};
		c = sampleCubemap(*s, faces, v3(x, y, z));
	} else {
		c = sampleLatLong(s->textures[size_t(tex)], v3(x, y, z));
	}
	out[0] = c.x;
	out[1] = c.y;
	out[2] = c.z;
	out[3] = 0.0f;
}
// Light::sample probe: returns the direction, dist and attenuation for a light at `position`.
void orc_light_sample(const rdu_light_desc * L, const float * position, float * dir, float * dist, float * atten) {
	float d = 0.0f, a = 0.0f;
	const v3 r = lightSample(*L, v3(position[0], position[1], position[2]), d, a);
	dir[0] = r.x;
	dir[1] = r.y;
	dir[2] = r.z;
	*dist = d;
	*atten = a;
}
// MaterialSky::eval probe with texture `tex` of scene h as the scattering table.
void orc_sky_eval(void * h, int tex, const float * origin, const float * dir, const float * sun, float * out) {
	const Scene * s = static_cast<const Scene *>(h);
	const v3 r = skyEval(s->textures[size_t(tex)], v3(origin[0], origin[1], origin[2]), v3(dir[0], dir[1], dir[2]),
		v3(sun[0], sun[1], sun[2]));
	out[0] = r.x;
	out[1] = r.y;
	out[2] = r.z;
}
void orc_sky_phase(float cosAngle, float * rayleigh, float * mie) {
	*rayleigh = skyRayleighPhase(cosAngle);
	*mie = skyMiePhase(cosAngle);
}
// buildLocalFrame probe on object `object` of a committed scene, at hit (localId, u, v) of its
// mesh; the texture coordinates are interpolated as render does. out: tbn columns t, b, n.
void orc_local_frame(void * h, uint32_t object, uint32_t localId, float u, float v, const float * rayDir, float * out) {
	const Scene * s = static_cast<const Scene *>(h);
	const Object & obj = s->objects[object];
	Hit hit;
	hit.hit = true;
	hit.dist = 1.0f;
	hit.u = u;
	hit.v = v;
	hit.w = 1.0f - u - v;
	hit.localId = localId;
	hit.meshId = object;
	const MeshData & mesh = s->meshes[obj.mesh];
	const v2 uv = obj.useUV ? interp2(mesh.uv, mesh.idx, hit) : v2{0.5f, 0.5f};
	const m3 tbn = buildLocalFrame(*s, obj, hit, v3(rayDir[0], rayDir[1], rayDir[2]), uv);
	for(int c = 0; c < 3; ++c)
		for(int r = 0; r < 3; ++r)
			out[3 * c + r] = tbn.c[c][r];
}
void orc_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t * out) {
	const u4 r = philox4x32_10({c0, c1, c2, c3}, k0, k1);
	out[0] = r.x;
	out[1] = r.y;
	out[2] = r.z;
	out[3] = r.w;
}

} // extern "C"
