// Device-side scene data and shading functions of the path tracer: Rendu's MaterialGGX
// (src/apps/pathtracer/MaterialGGX.cpp:5-101), MaterialSky (MaterialSky.cpp:6-98), Image::rgbl/rgbal
// (src/engine/resources/Image.cpp:64-106), Texture::sampleCubemap (Texture.cpp:141-190), the light
// samplers (src/engine/scene/lights/*.cpp) and PathTracer::buildLocalFrame / evalBackground
// (PathTracer.cpp:23-122), in the same operation order (compiled without contraction).
// Transcendentals (pow, exp, cos, sin, atan2, acos) are the device library's, within a few ulp of
// the host libm.
#pragma once
#include "rdu_vec.hpp"
#include "trace.hpp"
#include "../../include/rdu.h"

namespace rdu {

struct DevTexture {
	uint32_t offset; // first texel in DevScene::texels
	int32_t w, h;
};

struct DevMesh {
	uint32_t vtx; // first vertex in the attribute arrays
	uint32_t idx; // first index in the index array
	int32_t hasUV;
};

struct DevObject {
	uint32_t mesh, mat;
	int32_t useUV;
	float invtp[9]; // inverse(transpose(mat3(model))), column-major
};

struct DevMaterial {
	int32_t type, masked;
	int32_t tex[3];
};

struct DevScene {
	const float4 * texels = nullptr;
	const DevTexture * textures = nullptr;
	const float * nrm = nullptr; // 3 floats per vertex
	const float * tan = nullptr;
	const float * bit = nullptr;
	const float * uv = nullptr; // 2 floats per vertex (0 where a mesh has none)
	const uint32_t * idx = nullptr;
	const DevMesh * meshes = nullptr;
	const DevObject * objects = nullptr;
	const DevMaterial * materials = nullptr;
	const rdu_light_desc * lights = nullptr;
	uint32_t nLights = 0;
	int32_t anyMasked = 0;
	rdu_background_desc bg{};
	// Extension (extended estimator): light sampling of the lat-long environment and of emissive
	// triangles, combined with BSDF sampling by the power heuristic.
	const float * envPmf = nullptr;    // envW x envH texel probabilities
	const float * envRowCdf = nullptr; // envH + 1
	const float * envColCdf = nullptr; // envH rows of envW + 1
	int32_t envW = 0, envH = 0;
	const float4 * emTris = nullptr;   // 4 float4 per emissive triangle: {v0, area} {e1, pmf}
	                                   // {e2, object bits} {n = cross(e1, e2), first index bits}
	const float * emCdf = nullptr;     // nEm + 1
	const int32_t * emBase = nullptr;  // per object: its first emissive triangle, or -1
	uint32_t nEm = 0;
	float pEnv = 0.0f;                 // environment vs triangles within area-light sampling
	float pArea = 0.0f;                // area lights vs analytic lights
};

struct f2 {
	float x, y;
};
struct f4v {
	float x, y, z, w;
};

constexpr float kPi = 3.14159265358979323846264338327950288f;
constexpr float kOneOverPi = 0.318309886183790671537767526745028724f;
constexpr float kTwoPi = 6.28318530717958647692528676655900576f;
constexpr float kHalfPi = 1.57079632679489661923132169163975144f;
constexpr float kQuarterPi = 0.785398163397448309615660845819875721f;

RDU_HD float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
RDU_HD f3 mix3(f3 x, f3 y, float a) { return {mixf(x.x, y.x, a), mixf(x.y, y.y, a), mixf(x.z, y.z, a)}; }
RDU_HD float clampf(float v, float lo, float hi) { return gmin(gmax(v, lo), hi); }
RDU_HD int modPos(int x, int w) { return ((x % w) + w) % w; }

// ---- counter-based random numbers: Philox4x32-10 (Salmon et al., SC'11) -------------------------
struct u4 {
	uint32_t x, y, z, w;
};
RDU_HD u4 philox(u4 c, uint32_t k0, uint32_t k1) {
	for(int r = 0; r < 10; ++r) {
		const uint64_t p0 = uint64_t(0xD2511F53u) * c.x;
		const uint64_t p1 = uint64_t(0xCD9E8D57u) * c.z;
		c = {uint32_t(p1 >> 32) ^ c.y ^ k0, uint32_t(p1), uint32_t(p0 >> 32) ^ c.w ^ k1, uint32_t(p0)};
		k0 += 0x9E3779B9u;
		k1 += 0xBB67AE85u;
	}
	return c;
}
RDU_HD float u01(uint32_t x) { return float(x >> 8) * (1.0f / 16777216.0f); }
// Four uniforms of one path sample: key (seed, pixel), counter (sample, slot). Slot 0 holds the
// pixel jitter, slot 1 + b the draws of bounce b, slot 0x10000 + b Russian roulette.
RDU_HD void draw4(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t slot, float u[4]) {
	const u4 r = philox({sample, slot, 0x52445531u, 0u}, seed, pixel);
	u[0] = u01(r.x);
	u[1] = u01(r.y);
	u[2] = u01(r.z);
	u[3] = u01(r.w);
}

// ---- textures ----------------------------------------------------------------------------------
__device__ __forceinline__ f4v texel(const DevScene & s, const DevTexture & t, int x, int y) {
	const float4 c = s.texels[t.offset + uint32_t(y * t.w + x)];
	return {c.x, c.y, c.z, c.w};
}
// Image::rgbal (Image.cpp:86-106)
__device__ __forceinline__ f4v rgbal(const DevScene & s, int tex, float x, float y) {
	const DevTexture t = s.textures[tex];
	const float xi = x * float(t.w);
	const float yi = y * float(t.h);
	const float xb = floorf(xi);
	const float yb = floorf(yi);
	const float dx = xi - xb;
	const float dy = yi - yb;
	const int x0 = modPos(int(xb), t.w);
	const int y0 = modPos(int(yb), t.h);
	const int x1 = modPos(int(xb) + 1, t.w);
	const int y1 = modPos(int(yb) + 1, t.h);
	const f4v p00 = texel(s, t, x0, y0), p01 = texel(s, t, x0, y1), p10 = texel(s, t, x1, y0), p11 = texel(s, t, x1, y1);
	const float a = 1.0f - dx, b = 1.0f - dy;
	f4v r;
	r.x = a * (b * p00.x + dy * p01.x) + dx * (b * p10.x + dy * p11.x);
	r.y = a * (b * p00.y + dy * p01.y) + dx * (b * p10.y + dy * p11.y);
	r.z = a * (b * p00.z + dy * p01.z) + dx * (b * p10.z + dy * p11.z);
	r.w = a * (b * p00.w + dy * p01.w) + dx * (b * p10.w + dy * p11.w);
	return r;
}
__device__ __forceinline__ f3 rgbl(const DevScene & s, int tex, float x, float y) {
	const f4v c = rgbal(s, tex, x, y);
	return {c.x, c.y, c.z};
}

// Texture::sampleCubemap (Texture.cpp:141-190); faces px, nx, py, ny, pz, nz.
__device__ inline f3 sampleCubemap(const DevScene & s, const int32_t * faces, f3 dir) {
	const f3 a = {fabsf(dir.x), fabsf(dir.y), fabsf(dir.z)};
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
	// Face texture by masks rather than an indexed read, so a descriptor copied into registers is
	// not forced to memory.
	int tex = 0;
	for(int k = 0; k < 6; ++k)
		tex |= faces[k] & -int(side == k);
	const DevTexture t = s.textures[tex];
	const float eps = 1.0f / float(t.w < t.h ? t.w : t.h);
	x = clampf(x, 0.0f + eps, 1.0f - eps);
	y = clampf(y, 0.0f + eps, 1.0f - eps);
	return rgbl(s, tex, x, y);
}

__device__ inline bool sphereRoots(f3 o, f3 d, float radius, f2 & roots) {
	const float a = dot(d, d);
	const float b = dot(o, d);
	const float c = dot(o, o) - radius * radius;
	const float delta = b * b - a * c;
	if(delta < 0.0f)
		return false;
	const float ds = sqrt_rn(delta);
	roots = {(-b - ds) / a, (-b + ds) / a};
	return true;
}

// MaterialSky::eval (MaterialSky.cpp:6-98), Sky::AtmosphereParameters (Sky.hpp:53-65).
__device__ inline f3 skyEval(const DevScene & s, int table, f3 rayOrigin, f3 rayDir, f3 sunDir) {
	const f3 sunColor = {1.474f, 1.8504f, 1.91198f};
	const f3 kRayleigh = {5.5e-6f, 13.0e-6f, 22.4e-6f};
	const float groundRadius = 6371e3f, topRadius = 6471e3f, sunIntensity = 20.0f, kMie = 21e-6f;
	const float heightRayleigh = 8000.0f, heightMie = 1200.0f, gMie = 0.758f, sunRadius = 0.04675f;
	const float sunRadiusCos = 0.998f;
	const unsigned samplesCount = 16;
	const f3 planetPos = rayOrigin + mk(0.0f, groundRadius, 0.0f) + mk(0.0f, 1.0f, 0.0f);
	f2 interTop = {0, 0}, interGround = {0, 0};
	if(!sphereRoots(planetPos, rayDir, topRadius, interTop))
		return splat(0.0f);
	const bool didHitGround = sphereRoots(planetPos, rayDir, groundRadius, interGround);
	const float distanceToInter = gmin(interTop.y, didHitGround ? interGround.x : 0.0f);
	const float stepSize = (distanceToInter - interTop.x) / float(samplesCount);
	const float cosViewSun = dot(rayDir, sunDir);
	float rayleighDist = 0.0f, mieDist = 0.0f;
	f3 rayleighScatt = splat(0.0f), mieScatt = splat(0.0f), transmittance = splat(0.0f);
	for(unsigned i = 0; i < samplesCount; ++i) {
		const f3 currPos = planetPos + ((float(i) + 0.5f) * stepSize) * rayDir;
		float currHeight = length(currPos) - groundRadius;
		if(i == (samplesCount - 1) && currHeight < 0.0f)
			currHeight = 0.0f;
		const float rayleighStep = expf(-currHeight / heightRayleigh) * stepSize;
		const float mieStep = expf(-currHeight / heightMie) * stepSize;
		rayleighDist += rayleighStep;
		mieDist += mieStep;
		const f3 ex = -(mieDist * splat(kMie) + rayleighDist * kRayleigh);
		const f3 directAttenuation = {expf(ex.x), expf(ex.y), expf(ex.z)};
		const float relativeHeight = (length(currPos) - groundRadius) / (topRadius - groundRadius);
		const float relativeCosAngle = -0.5f * sunDir.y + 0.5f;
		float ux = (511.0f / 512.0f) * relativeHeight + 0.5f / 512.0f;
		float uy = (511.0f / 512.0f) * relativeCosAngle + 0.5f / 512.0f;
		ux = clampf(ux, 0.5f / 512.0f, 511.0f / 512.0f);
		uy = clampf(uy, 0.5f / 512.0f, 511.0f / 512.0f);
		const f3 secondary = rgbl(s, table, ux, uy);
		const f3 attenuation = directAttenuation * secondary;
		rayleighScatt = rayleighScatt + rayleighStep * attenuation;
		mieScatt = mieScatt + mieStep * attenuation;
		transmittance = transmittance + directAttenuation;
	}
	const float k = 1.0f / (4.0f * kPi);
	const float rayleighPhase = k * 3.0f / 4.0f * (1.0f + cosViewSun * cosViewSun);
	const float g2 = gMie * gMie;
	const float miePhase = k * 3.0f * (1.0f - g2) / (2.0f * (2.0f + g2)) * (1.0f + cosViewSun * cosViewSun) /
						   powf(1.0f + g2 - 2.0f * gMie * cosViewSun, 3.0f / 2.0f);
	const f3 rayleighPart = rayleighPhase * kRayleigh * rayleighScatt;
	const f3 miePart = kMie * miePhase * mieScatt;
	f3 sunRadiance = splat(0.0f);
	const bool didHitGroundForward = didHitGround && interGround.y > 0.0f;
	if(!didHitGroundForward && dot(rayDir, sunDir) > sunRadiusCos)
		sunRadiance = sunColor / (kPi * sunRadius * sunRadius);
	return sunIntensity * (rayleighPart + miePart) + transmittance * sunRadiance;
}

// Extension: latitude-longitude environment (u along the azimuth from -z towards +x, v from +y).
__device__ __forceinline__ f3 sampleLatLong(const DevScene & s, int tex, f3 d) {
	const float u = 0.5f + atan2f(d.x, -d.z) * (0.5f * kOneOverPi);
	const float v = acosf(clampf(d.y, -1.0f, 1.0f)) * kOneOverPi;
	return rgbl(s, tex, u, v);
}

// ---- extension: area-light sampling ----------------------------------------------------------
constexpr float kTwoPiSq = 19.7392088021787171f; // 2 pi^2

// Largest k in [0, n) with cdf[k] <= u (cdf: n + 1 entries from 0 to 1, u in [0, 1)).
__device__ inline int cdfSearch(const float * cdf, int n, float u) {
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

// Density (solid angle) of the environment strategy for direction d.
__device__ inline float envPdf(const DevScene & s, f3 d) {
	if(s.envW == 0)
		return 0.0f;
	const float u = 0.5f + atan2f(d.x, -d.z) * (0.5f * kOneOverPi);
	const float v = acosf(clampf(d.y, -1.0f, 1.0f)) * kOneOverPi;
	int i = int(u * float(s.envW)), j = int(v * float(s.envH));
	i = i < 0 ? 0 : (i >= s.envW ? s.envW - 1 : i);
	j = j < 0 ? 0 : (j >= s.envH ? s.envH - 1 : j);
	const float sinT = sqrt_rn(gmax(0.0f, 1.0f - d.y * d.y));
	if(!(sinT > 0.0f))
		return 0.0f;
	return s.envPmf[j * s.envW + i] * float(s.envW * s.envH) / (kTwoPiSq * sinT);
}

// One direction towards the environment; pdf = the density it was drawn with.
__device__ inline f3 envSample(const DevScene & s, float u1, float u2, float & pdf) {
	const int j = cdfSearch(s.envRowCdf, s.envH, u1);
	const float a = s.envRowCdf[j], b = s.envRowCdf[j + 1];
	const float dv = (u1 - a) / (b - a);
	const float * cc = s.envColCdf + size_t(j) * (s.envW + 1);
	const int i = cdfSearch(cc, s.envW, u2);
	const float du = (u2 - cc[i]) / (cc[i + 1] - cc[i]);
	const float v = (float(j) + dv) / float(s.envH), u = (float(i) + du) / float(s.envW);
	const float theta = v * kPi, phi = (u - 0.5f) * kTwoPi;
	const float sinT = sinf(theta), cosT = cosf(theta);
	pdf = sinT > 0.0f ? s.envPmf[j * s.envW + i] * float(s.envW * s.envH) / (kTwoPiSq * sinT) : 0.0f;
	return {sinT * sinf(phi), cosT, -sinT * cosf(phi)};
}

// Density (solid angle, at distance dist along direction dir) of the triangle strategy for a point
// on emissive triangle t.
__device__ inline float triPdf(const DevScene & s, uint32_t t, float dist, f3 dir) {
	const float4 a = s.emTris[4 * t], b = s.emTris[4 * t + 1], n4 = s.emTris[4 * t + 3];
	const f3 n = {n4.x, n4.y, n4.z};
	const float cosL = fabsf(dot(n, dir)) / length(n);
	return b.w / a.w * (dist * dist) / cosL;
}

// Power heuristic weight of a strategy with density p (> 0) against one with density q.
RDU_HD float powerWeight(float p, float q) {
	const float r = q / p;
	return 1.0f / (1.0f + r * r);
}

// PathTracer::evalBackground (PathTracer.cpp:23-55).
__device__ inline f3 evalBackground(const DevScene & s, f3 rayDir, f3 rayPos, f2 ndc, bool direct) {
	const rdu_background_desc & b = s.bg;
	const f3 sun = {b.sun_direction[0], b.sun_direction[1], b.sun_direction[2]};
	if(direct && b.mode == RDU_BACKGROUND_IMAGE)
		return rgbl(s, b.texture, ndc.x, ndc.y);
	if(b.mode == RDU_BACKGROUND_SKYBOX)
		return sampleCubemap(s, b.cube, normalize(rayDir));
	if(b.mode == RDU_BACKGROUND_ATMOSPHERE)
		return skyEval(s, b.texture, rayPos, normalize(rayDir), sun);
	if(b.mode == RDU_BACKGROUND_LATLONG)
		return sampleLatLong(s, b.texture, normalize(rayDir));
	if(direct && b.mode == RDU_BACKGROUND_COLOR)
		return {b.color[0], b.color[1], b.color[2]};
	return splat(0.0f);
}

// ---- MaterialGGX --------------------------------------------------------------------------------
RDU_HD f3 ggxF(f3 F0, float VdotH) { return F0 + powf(1.0f - VdotH, 5.0f) * (splat(1.0f) - F0); }
RDU_HD float ggxD(float NdotH, float alpha) {
	const float halfDenum = NdotH * NdotH * (alpha * alpha - 1.0f) + 1.0f;
	const float halfTerm = alpha / gmax(0.0001f, halfDenum);
	return halfTerm * halfTerm * kOneOverPi;
}
RDU_HD float ggxV(float NdotL, float NdotV, float alpha) {
	const float alpha2 = alpha * alpha;
	const float visL = NdotV * sqrt_rn((-NdotL * alpha2 + NdotL) * NdotL + alpha2);
	const float visV = NdotL * sqrt_rn((-NdotV * alpha2 + NdotV) * NdotV + alpha2);
	return 0.5f / gmax(0.0001f, visV + visL);
}
RDU_HD float alphaFromRoughness(float r) {
	const float rc = gmax(0.045f, r);
	return gmax(0.0001f, rc * rc);
}
RDU_HD f3 ggxBrdf(f3 wo, f3 base, float alpha, float metallic, f3 wi, float * pdf) {
	const f3 h = normalize(wi + wo);
	const float NdotH = gmax(h.z, 0.0f);
	const float VdotH = gmax(dot(wi, h), 0.0f);
	const float NdotL = gmax(wo.z, 0.0f);
	const float NdotV = gmax(wi.z, 0.0f);
	const float Dh = ggxD(NdotH, alpha);
	if(pdf) {
		const float hPdf = Dh * NdotH;
		*pdf = hPdf / (4.0f * gmax(0.0001f, VdotH));
	}
	const f3 F0 = mix3(splat(0.04f), base, metallic);
	const f3 specular = Dh * ggxV(NdotL, NdotV, alpha) * ggxF(F0, VdotH);
	const f3 diffuse = (1.0f - metallic) * kOneOverPi * base * (splat(1.0f) - F0);
	const f3 multiAdj = splat(1.0f) + (2.0f * alpha * alpha * NdotV) * F0;
	return (diffuse + specular * multiAdj) * NdotV;
}
RDU_HD f2 sampleDisk(float u0, float u1) {
	const float x = 2.0f * u0 - 1.0f;
	const float y = 2.0f * u1 - 1.0f;
	if(x == 0.0f && y == 0.0f)
		return {0.0f, 0.0f};
	float angle, radius;
	if(fabsf(x) > fabsf(y)) {
		radius = x;
		angle = kQuarterPi * y / x;
	} else {
		radius = y;
		angle = kHalfPi - kQuarterPi * x / y;
	}
	return {radius * cosf(angle), radius * sinf(angle)};
}
// u: [lobe choice, dir0, dir1]; pdfOut (optional) receives the mixture density of wi.
RDU_HD f3 ggxSampleAndEval(f3 wo, f3 base, float roughness, float metallic, const float * u, f3 & wi,
	float * pdfOut = nullptr) {
	const float probaSpecular = mixf(1.0f / (dot(base, splat(1.0f)) / 3.0f + 1.0f), 1.0f, metallic);
	const float alpha = alphaFromRoughness(roughness);
	if(u[0] < probaSpecular) {
		const float a2 = alpha * alpha;
		const float x = u[1];
		const float phiH = u[2] * kTwoPi;
		const float cosThetaHSqr = gmin((1.0f - x) / ((a2 - 1.0f) * x + 1.0f), 1.0f);
		const float cosThetaH = sqrt_rn(cosThetaHSqr);
		const float sinThetaH = sqrt_rn(1.0f - cosThetaHSqr);
		const f3 lh = {sinThetaH * cosf(phiH), sinThetaH * sinf(phiH), cosThetaH};
		wi = (2.0f * dot(wo, lh)) * lh - wo;
	} else {
		const f2 xy = sampleDisk(u[1], u[2]);
		const float z = sqrt_rn(gmax(0.0f, 1.0f - xy.x * xy.x - xy.y * xy.y));
		wi = {xy.x, xy.y, z};
		if(wo.z < 0.0f)
			wi.z *= -1.0f;
	}
	if(wi.z < 0.0f)
		return splat(0.0f);
	float pdfSpec = 0.0f;
	const f3 brdf = ggxBrdf(wo, base, alpha, metallic, wi, &pdfSpec);
	const float pdf = mixf(kOneOverPi * gmax(wi.z, 0.0f), pdfSpec, probaSpecular);
	if(pdfOut)
		*pdfOut = pdf;
	if(pdf == 0.0f)
		return splat(0.0f);
	return brdf / pdf;
}
// Extension: the mixture density of ggxSampleAndEval for a given direction (0 below the horizon).
RDU_HD float ggxPdf(f3 wo, f3 base, float roughness, float metallic, f3 wi) {
	if(wi.z < 0.0f)
		return 0.0f;
	const float probaSpecular = mixf(1.0f / (dot(base, splat(1.0f)) / 3.0f + 1.0f), 1.0f, metallic);
	float pdfSpec = 0.0f;
	(void)ggxBrdf(wo, base, alphaFromRoughness(roughness), metallic, wi, &pdfSpec);
	return mixf(kOneOverPi * gmax(wi.z, 0.0f), pdfSpec, probaSpecular);
}
RDU_HD f3 ggxEval(f3 wo, f3 base, float roughness, float metallic, f3 wi) {
	if(wi.z < 0.0f)
		return splat(0.0f);
	return ggxBrdf(wo, base, alphaFromRoughness(roughness), metallic, wi, nullptr);
}

// ---- lights ------------------------------------------------------------------------------------
__device__ inline f3 lightSample(const rdu_light_desc & L, f3 position, float & dist, float & attenuation) {
	if(L.type == RDU_LIGHT_DIRECTIONAL) {
		attenuation = 1.0f;
		dist = 3.402823466e+38f;
		return -mk(L.direction[0], L.direction[1], L.direction[2]);
	}
	f3 direction = mk(L.position[0], L.position[1], L.position[2]) - position;
	dist = length(direction);
	attenuation = 0.0f;
	if(dist > L.radius)
		return splat(0.0f);
	if(dist > 0.0f)
		direction = direction / dist;
	float angleAttenuation = 1.0f;
	if(L.type == RDU_LIGHT_SPOT) {
		const f3 ldir = {L.direction[0], L.direction[1], L.direction[2]};
		const float currentCos = dot(-direction, ldir);
		const float outerCos = cosf(L.angles[1]);
		if(currentCos < cosf(outerCos))
			return splat(0.0f);
		const float innerCos = cosf(L.angles[0]);
		angleAttenuation = clampf((currentCos - outerCos) / (innerCos - outerCos), 0.0f, 1.0f);
	}
	const float radiusRatio = dist / L.radius;
	const float radiusRatio2 = radiusRatio * radiusRatio;
	const float attenNum = clampf(1.0f - radiusRatio2, 0.0f, 1.0f);
	attenuation = L.type == RDU_LIGHT_SPOT ? angleAttenuation * attenNum * attenNum : attenNum * attenNum;
	return direction;
}

// ---- geometry at a hit ------------------------------------------------------------------------
struct M3 {
	f3 c[3];
};
RDU_HD f3 mul(const M3 & m, f3 v) {
	return {m.c[0].x * v.x + m.c[1].x * v.y + m.c[2].x * v.z, m.c[0].y * v.x + m.c[1].y * v.y + m.c[2].y * v.z,
		m.c[0].z * v.x + m.c[1].z * v.y + m.c[2].z * v.z};
}
// transpose(m) * v
RDU_HD f3 mulT(const M3 & m, f3 v) { return {dot(m.c[0], v), dot(m.c[1], v), dot(m.c[2], v)}; }

__device__ __forceinline__ f3 attr3(const float * a, uint32_t i) { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }

// Raycaster::interpolateAttribute (Raycaster.hpp:90-97): w*A0 + u*A1 + v*A2.
__device__ __forceinline__ f3 interp3(const float * a, uint32_t i0, uint32_t i1, uint32_t i2, float w, float u, float v) {
	return w * attr3(a, i0) + u * attr3(a, i1) + v * attr3(a, i2);
}
__device__ __forceinline__ f2 interp2(const float * a, uint32_t i0, uint32_t i1, uint32_t i2, float w, float u, float v) {
	return {w * a[2 * i0] + u * a[2 * i1] + v * a[2 * i2], w * a[2 * i0 + 1] + u * a[2 * i1 + 1] + v * a[2 * i2 + 1]};
}

struct HitVerts {
	uint32_t i0, i1, i2;
};
__device__ __forceinline__ HitVerts hitVerts(const DevScene & s, const DevObject & obj, uint32_t localId) {
	const DevMesh m = s.meshes[obj.mesh];
	const uint32_t * ix = s.idx + m.idx + localId;
	return {m.vtx + ix[0], m.vtx + ix[1], m.vtx + ix[2]};
}

// PathTracer::buildLocalFrame (PathTracer.cpp:86-122).
__device__ inline M3 buildLocalFrame(const DevScene & s, const DevObject & obj, const DevMaterial & mat, HitVerts hv,
	float w, float u, float v, f3 rayDir, f2 uv) {
	const f3 n = normalize(interp3(s.nrm, hv.i0, hv.i1, hv.i2, w, u, v));
	const f3 t = normalize(interp3(s.tan, hv.i0, hv.i1, hv.i2, w, u, v));
	const f3 b = normalize(interp3(s.bit, hv.i0, hv.i1, hv.i2, w, u, v));
	M3 inv;
	inv.c[0] = {obj.invtp[0], obj.invtp[1], obj.invtp[2]};
	inv.c[1] = {obj.invtp[3], obj.invtp[4], obj.invtp[5]};
	inv.c[2] = {obj.invtp[6], obj.invtp[7], obj.invtp[8]};
	M3 tbn;
	tbn.c[0] = mul(inv, t);
	tbn.c[1] = mul(inv, b);
	tbn.c[2] = mul(inv, n);
	if(!(dot(tbn.c[2], rayDir) < 0.0f))
		tbn.c[2] = tbn.c[2] * -1.0f;
	if(obj.useUV && mat.type != RDU_MATERIAL_EMISSIVE) {
		const f4v img = rgbal(s, mat.tex[1], uv.x, uv.y);
		const f3 localNormal = normalize(2.0f * mk(img.x, img.y, img.z) - splat(1.0f));
		tbn.c[2] = mul(tbn, localNormal);
	}
	tbn.c[2] = normalize(tbn.c[2]);
	tbn.c[1] = normalize(cross(tbn.c[2], normalize(tbn.c[0])));
	tbn.c[0] = normalize(cross(tbn.c[1], tbn.c[2]));
	return tbn;
}

} // namespace rdu
