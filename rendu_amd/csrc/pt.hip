// Wavefront path tracer for gfx950: Rendu's PathTracer::render (src/apps/pathtracer/PathTracer.cpp:
// 159-299) split into kernels over queues of live paths.
//
// Per pass (a batch of pixel samples, path p = s * pixels + pixel):
//   generate   camera rays (stratified jitter, PathTracer.cpp:57-84, 191-198); queue = all paths
//   for each bounce b:
//     extend   closest hit of every queued path (persistent traversal kernel, trace.hip)
//     shade    background / emissive / masked cut-out; local frame; NEE light sample (shadow ray
//              queued with its would-be contribution); BSDF sample; the surviving paths are
//              appended to the next queue
//     shadow   any hit (or, with masked materials, Rendu's closest-hit walk through cut-outs) of
//              the queued shadow rays; visible ones add their contribution to the path
//   finalize   clamp min(c, 5) per path; the pixel sums add the pass's samples in sample order, so
//              every pixel accumulates exactly like Rendu's loop over sid (then / spp).
//
// Queues are compacted with a wavefront ballot + mbcnt prefix and one atomic per wavefront.
// Counts live on the device: the extend/shadow traversals read them there, so a bounce issues no
// host synchronisation.
#include "pt.hpp"

namespace rdu {

namespace {

constexpr int kPtBlock = 256;

// Append `p` to a queue for every lane whose `want` is set: one atomic per wavefront.
__device__ __forceinline__ void appendQueue(bool want, uint32_t p, uint32_t * queue, uint32_t * count) {
	const uint64_t m = __ballot(want);
	if(!m)
		return;
	const uint32_t lane = __lane_id();
	const uint32_t leader = uint32_t(__ffsll((unsigned long long)m) - 1);
	uint32_t base = 0;
	if(lane == leader)
		base = atomicAdd(count, uint32_t(__popcll(m)));
	base = __shfl(base, int(leader), 64);
	if(want) {
		const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
		queue[base + rank] = p;
	}
}

// Reserve one slot of an array for every lane whose `want` is set (one atomic per wavefront);
// returns the lane's slot (meaningless where `want` is false). Called by whole wavefronts.
__device__ __forceinline__ uint32_t appendSlot(bool want, uint32_t * count) {
	const uint64_t m = __ballot(want);
	if(!m)
		return 0;
	const uint32_t lane = __lane_id();
	const uint32_t leader = uint32_t(__ffsll((unsigned long long)m) - 1);
	uint32_t base = 0;
	if(lane == leader)
		base = atomicAdd(count, uint32_t(__popcll(m)));
	base = __shfl(base, int(leader), 64);
	return base + __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

// Trust cell of a position in the cache's box on a res^3 grid.
__device__ __forceinline__ uint32_t cacheCell(const RcGrid & g, f3 p, int res) {
	const float v[3] = {p.x, p.y, p.z};
	uint32_t c[3];
	for(int a = 0; a < 3; ++a) {
		const float t = (v[a] - g.lo[a]) * g.inv[a] * float(res);
		const int i = t > 0.0f ? (t < float(res - 1) ? int(t) : res - 1) : 0;
		c[a] = uint32_t(i);
	}
	return c[0] + uint32_t(res) * (c[1] + uint32_t(res) * c[2]);
}

// |sum(prediction) - sum(target)| within maxBias of the larger sum, plus twice the standard
// deviation of the target sum (targets are one-sample estimates), plus 1e-3 per pair.
__device__ __forceinline__ bool cellPasses(const PtCache & C, float4 s) {
	const float var = gmax(s.w - s.y * s.y / gmax(s.z, 1e-6f), 0.0f);
	return fabsf(s.x - s.y) <= C.maxBias * gmax(s.x, s.y) + 2.0f * sqrtf(var) + 1e-3f * s.z;
}

__device__ __forceinline__ bool cacheTrusted(const PtCache & C, f3 p) {
	if(C.alwaysTrust)
		return true;
	const float4 s = C.cells[cacheCell(C.grid, p, kCacheCellRes)];
	if(s.z >= C.minCount)
		return cellPasses(C, s);
	const float4 c = C.cells[kCacheFineCells + cacheCell(C.grid, p, kCacheCoarseRes)];
	return c.z >= C.minCount && cellPasses(C, c);
}

__device__ __forceinline__ float luminance(float4 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

__device__ __forceinline__ void putRay(float4 * rays, uint32_t p, f3 o, f3 d, float tmin, float tmax) {
	rays[2 * size_t(p)] = make_float4(o.x, o.y, o.z, tmin);
	rays[2 * size_t(p) + 1] = make_float4(d.x, d.y, d.z, tmax);
}

// Screen position of sample sid of pixel (x, y): getSamplePosition (PathTracer.cpp:71-84).
__device__ __forceinline__ f2 samplePos(const PtFrame & f, uint32_t seed, uint32_t x, uint32_t y, uint32_t sid) {
	float u[4];
	draw4(seed, y * f.width + x, sid, 0, u);
	const int sidy = int(sid) / f.cellX;
	const int sidx = int(sid) % f.cellX;
	const float gx = float(sidx) + u[0], gy = float(sidy) + u[1];
	const f2 screen = {float(x) + gx * f.cellW, float(y) + gy * f.cellH};
	return {screen.x / float(f.width), screen.y / float(f.height)};
}

__global__ __launch_bounds__(kPtBlock) void ptGenerate(PtFrame f, PtPaths P, uint32_t nPaths) {
	const uint32_t p = blockIdx.x * kPtBlock + threadIdx.x;
	if(p >= nPaths)
		return;
	const uint32_t sampleBase = P.counts[kCountBase];
	const uint32_t seed = P.counts[kCountSeed];
	const uint32_t pixels = f.width * f.height;
	const uint32_t s = p / pixels, pix = p % pixels;
	const uint32_t x = pix % f.width, y = pix / f.width;
	const f2 ndc = samplePos(f, seed, x, y, sampleBase + s);
	const f3 eye = f.eye;
	const f3 worldPos = f.corner + ndc.x * f.dx + ndc.y * f.dy;
	const f3 dir = normalize(worldPos - eye);
	putRay(P.rays, p, eye, dir, 0.0001f, 1e8f);
	P.atten[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
	P.color[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	P.mis[p] = make_float2(0.0f, 0.0f);
	P.queue[0][p] = p;
}

// Start of a pass: all paths active, empty queues, and this pass's first sample index taken from
// (and advanced in) the device counters.
__global__ void ptResetCounts(uint32_t * counts, uint32_t active, uint32_t samplesInPass) {
	if(threadIdx.x == 0) {
		counts[kCountActive] = active;
		counts[kCountNext] = 0;
		counts[kCountShadow] = 0;
		counts[kCountShadowNext] = 0;
		counts[kCountOverflow] = 0;
		const uint32_t base = counts[kCountBaseNext];
		counts[kCountBase] = base;
		counts[kCountBaseNext] = base + samplesInPass;
	}
}

__global__ void ptSetBase(uint32_t * counts, uint32_t base, uint32_t seed) {
	if(threadIdx.x == 0) {
		counts[kCountBaseNext] = base;
		counts[kCountSeed] = seed;
	}
}

// After a bounce: next queue becomes the active one; totals for the statistics.
__global__ void ptAdvance(uint32_t * counts, unsigned long long * totals) {
	if(threadIdx.x == 0) {
		totals[0] += counts[kCountActive];
		totals[1] += counts[kCountShadow];
		counts[kCountActive] = counts[kCountNext];
		counts[kCountNext] = 0;
		counts[kCountShadow] = 0;
		counts[kCountShadowNext] = 0;
		counts[kCountOverflow] = 0;
	}
}

__global__ void ptShadowSwap(uint32_t * counts, unsigned long long * totals) {
	if(threadIdx.x == 0) {
		totals[1] += counts[kCountShadow];
		counts[kCountShadow] = counts[kCountShadowNext];
		counts[kCountShadowNext] = 0;
	}
}

// One bounce of PathTracer::render's loop body (PathTracer.cpp:202-286) for queue entry i of n.
// The path's sample and pixel come from its id less pathBegin (the first path of its job in a
// batch pass; 0 otherwise). Called by the active lanes of a wavefront together (the queue appends
// are wavefront ballots over the active lanes).
template<bool CACHE>
__device__ __forceinline__ void shadeEntry(const DevScene & S, const PtFrame & f, const PtPaths & P,
	const uint32_t * queueIn, uint32_t * queueOut, uint32_t * counts, uint32_t bounce, const PtOptions & opt,
	const PtCache & C, uint32_t i, uint32_t n, uint32_t sampleBase, uint32_t seed, uint32_t pathBegin) {
	bool cont = false, shadow = false;
	bool doQuery = false, doRecord = false; // radiance-cache mode only
	f3 cPos{0, 0, 0}, cDir{0, 0, 0}, cColor{0, 0, 0}, cAtten{0, 0, 0};
	uint32_t p = 0;
	if(i < n) {
		p = queueIn[i];
		const uint32_t pixels = f.width * f.height;
		const uint32_t local = p - pathBegin;
		const uint32_t s = local / pixels, pix = local % pixels;
		const uint32_t sid = sampleBase + s;
		const float4 ro = P.rays[2 * size_t(p)];
		const float4 rd = P.rays[2 * size_t(p) + 1];
		const f3 rayPos = {ro.x, ro.y, ro.z};
		const f3 rayDir = {rd.x, rd.y, rd.z};
		const rdu_hit hit = P.hits[p];
		const float4 a4 = P.atten[p];
		f3 atten = {a4.x, a4.y, a4.z};
		float4 c4 = P.color[p];
		f3 color = {c4.x, c4.y, c4.z};
		// Extension (extended estimator): area lights and multiple importance sampling.
		const float pArea = opt.estimator == RDU_ESTIMATOR_EXTENDED ? S.pArea : 0.0f;
		const float2 mis = P.mis[p];
		if(!hit.hit) {
			const f2 ndc = samplePos(f, seed, pix % f.width, pix / f.width, sid);
			f3 bg = evalBackground(S, rayDir, rayPos, ndc, bounce == 0);
			if(mis.x > 0.0f && S.pEnv > 0.0f)
				bg = powerWeight(mis.x, pArea * S.pEnv * envPdf(S, normalize(rayDir))) * bg;
			color = color + atten * bg;
		} else {
			const DevObject obj = S.objects[hit.mesh_id];
			const DevMaterial mat = S.materials[obj.mat];
			const f3 pt = rayPos + hit.t * rayDir;
			const float hw = 1.0f - hit.u - hit.v; // Hit::w (Raycaster.cpp:10-12)
			const HitVerts hv = hitVerts(S, obj, hit.local_id);
			const f2 uv = obj.useUV ? interp2(S.uv, hv.i0, hv.i1, hv.i2, hw, hit.u, hit.v) : f2{0.5f, 0.5f};
			const f4v bCol = rgbal(S, mat.tex[0], uv.x, uv.y);
			if(mat.masked && bCol.w < 0.01f) {
				// Cut-out: continue from the hit point in the same direction; this consumes a bounce.
				if(bounce + 1 < opt.depth) {
					putRay(P.rays, p, pt, rayDir, 0.0001f, 1e8f);
					P.mis[p] = make_float2(mis.x, mis.y + hit.t);
					cont = true;
				}
			} else if(mat.type == RDU_MATERIAL_EMISSIVE) {
				f3 le = mk(bCol.x, bCol.y, bCol.z);
				if(mis.x > 0.0f && S.pEnv < 1.0f && S.emBase[hit.mesh_id] >= 0) {
					const uint32_t t = uint32_t(S.emBase[hit.mesh_id]) + hit.local_id / 3u;
					le = powerWeight(mis.x, pArea * (1.0f - S.pEnv) * triPdf(S, t, mis.y + hit.t, rayDir)) * le;
				}
				color = color + atten * le;
			} else {
				bool terminate = false;
				if(CACHE && int(bounce) >= C.terminateBounce) {
					const uint32_t px = pix % f.width, py = pix / f.width;
					const bool trainPixel = ((px + C.trainX) & 3u) == 0 && ((py + C.trainY) & 3u) == 0;
					cPos = pt;
					cDir = -rayDir;
					cColor = color;
					cAtten = atten;
					if(trainPixel && C.records)
						doRecord = true;
					else if(cacheTrusted(C, pt))
						doQuery = terminate = true;
				}
				if(!terminate) {
				const M3 tbn = buildLocalFrame(S, obj, mat, hv, hw, hit.u, hit.v, rayDir, uv);
				const f3 wo = normalize(mulT(tbn, -rayDir));
				const f3 baseColor = {powf(bCol.x, 2.2f), powf(bCol.y, 2.2f), powf(bCol.z, 2.2f)};
				const f4v rmao = rgbal(S, mat.tex[2], uv.x, uv.y);
				float u[4];
				draw4(seed, pix, sid, 1 + bounce, u);
				if(pArea > 0.0f && u[0] < pArea) {
					// Extension: one sample of the environment or of an emissive triangle, at every
					// vertex but the last (the reference's paths end there without a next segment),
					// weighted against BSDF sampling by the power heuristic. The visibility ray
					// leaves from the hit point with the path's own interval (1e-4, 1e8), so a light
					// point is visible exactly when a BSDF ray towards it would reach it.
					if(bounce + 1 < opt.depth) {
						float e[4];
						draw4(seed, pix, sid, 0x20000u + bounce, e);
						f3 direction, le;
						float pLight = 0.0f, pMis = 0.0f, maxDist = 0.0f;
						if(e[0] < S.pEnv) {
							direction = envSample(S, e[1], e[2], pLight);
							pLight = pArea * S.pEnv * pLight;
							pMis = pArea * S.pEnv * envPdf(S, direction);
							le = sampleLatLong(S, S.bg.texture, normalize(direction));
							maxDist = 1e8f;
						} else {
							const float ut = (e[0] - S.pEnv) / (1.0f - S.pEnv);
							const uint32_t t = uint32_t(cdfSearch(S.emCdf, int(S.nEm), ut));
							const float4 t0 = S.emTris[4 * t], t1 = S.emTris[4 * t + 1], t2 = S.emTris[4 * t + 2],
										 t3 = S.emTris[4 * t + 3];
							const float su = sqrt_rn(e[1]);
							const float bu = su * (1.0f - e[2]), bv = su * e[2], bw = 1.0f - bu - bv;
							const f3 y = mk(t0.x, t0.y, t0.z) + bu * mk(t1.x, t1.y, t1.z) + bv * mk(t2.x, t2.y, t2.z);
							const DevObject lo = S.objects[__float_as_uint(t2.w)];
							f2 luv = {0.5f, 0.5f};
							if(lo.useUV) {
								const HitVerts lv = hitVerts(S, lo, __float_as_uint(t3.w));
								luv = interp2(S.uv, lv.i0, lv.i1, lv.i2, bw, bu, bv);
							}
							const f4v ec = rgbal(S, S.materials[lo.mat].tex[0], luv.x, luv.y);
							le = mk(ec.x, ec.y, ec.z);
							const f3 toLight = y - pt;
							const float dist = length(toLight);
							direction = toLight / dist;
							pLight = pArea * (1.0f - S.pEnv) * triPdf(S, t, dist, direction);
							pMis = pLight;
							maxDist = dist * 0.999f;
						}
						if(pLight > 0.0f && pLight < __builtin_inff()) {
							const f3 lwi = normalize(mulT(tbn, direction));
							const f3 evalLight = ggxEval(wo, baseColor, rmao.x, rmao.y, lwi);
							const float w = powerWeight(pMis, ggxPdf(wo, baseColor, rmao.x, rmao.y, lwi));
							const f3 illum = evalLight * le * (w / pLight);
							if(illum.x > 0.0f || illum.y > 0.0f || illum.z > 0.0f) {
								const f3 contrib = atten * illum;
								putRay(P.shadowRays, p, pt, direction, 0.0001f, maxDist);
								P.shadowContrib[p] = make_float4(contrib.x, contrib.y, contrib.z, 0.0f);
								shadow = true;
							}
						}
					}
				} else if(S.nLights > 0) {
					float ul = u[0];
					float lightPdf = 1.0f / float(S.nLights);
					if(pArea > 0.0f) { // extension: the analytic lights share the pick with the area lights
						ul = (u[0] - pArea) / (1.0f - pArea);
						lightPdf = (1.0f - pArea) / float(S.nLights);
					}
					uint32_t lid = uint32_t(ul * float(S.nLights));
					lid = lid < S.nLights ? lid : S.nLights - 1;
					const rdu_light_desc & light = S.lights[lid];
					const f3 pShift = pt + 0.001f * tbn.c[2];
					float maxDist, falloff;
					const f3 direction = lightSample(light, pShift, maxDist, falloff);
					if(falloff > 0.0f) {
						const f3 lwi = normalize(mulT(tbn, direction));
						const f3 evalLight = ggxEval(wo, baseColor, rmao.x, rmao.y, lwi);
						const f3 illum = falloff * evalLight * mk(light.intensity[0], light.intensity[1], light.intensity[2]) / lightPdf;
						const f3 contrib = atten * illum;
						// checkVisibility (PathTracer.cpp:124-157) starts only while maxDist > 0.
						if(light.casts_shadow && maxDist > 0.0f) {
							putRay(P.shadowRays, p, pShift, direction, 0.001f, maxDist);
							P.shadowContrib[p] = make_float4(contrib.x, contrib.y, contrib.z, 0.0f);
							shadow = true;
						} else {
							color = color + contrib;
						}
					}
				}
				f3 wi;
				float bsdfPdf = 0.0f;
				const f3 eval = ggxSampleAndEval(wo, baseColor, rmao.x, rmao.y, u + 1, wi, &bsdfPdf);
				const f3 nextDir = normalize(mul(tbn, wi));
				atten = atten * eval;
				bool alive = bounce + 1 < opt.depth;
				if(alive && opt.estimator == RDU_ESTIMATOR_EXTENDED && int(bounce) + 1 >= opt.rrStart) {
					const float pmax = gmax(atten.x, gmax(atten.y, atten.z));
					const float ps = gmin(pmax, 0.95f);
					float rr[4];
					draw4(seed, pix, sid, 0x10000u + bounce, rr);
					if(!(rr[0] < ps))
						alive = false;
					else
						atten = atten / ps;
				}
				// A path whose throughput is exactly zero can add nothing more: stop tracing it.
				if(alive && !(atten.x == 0.0f && atten.y == 0.0f && atten.z == 0.0f)) {
					putRay(P.rays, p, pt, normalize(nextDir), 0.0001f, 1e8f);
					P.atten[p] = make_float4(atten.x, atten.y, atten.z, 0.0f);
					P.mis[p] = make_float2(pArea > 0.0f ? bsdfPdf : 0.0f, 0.0f);
					cont = true;
				}
				}
			}
		}
		P.color[p] = make_float4(color.x, color.y, color.z, 0.0f);
	}
	appendQueue(shadow, p, P.shadowQueue[0], counts + kCountShadow);
	appendQueue(cont, p, queueOut, counts + kCountNext);
	if(CACHE) {
		const uint32_t qs = appendSlot(doQuery, C.counts + kCacheQueries);
		if(doQuery) {
			C.query[2 * size_t(qs)] = make_float4(cPos.x, cPos.y, cPos.z, 0.0f);
			C.query[2 * size_t(qs) + 1] = make_float4(cDir.x, cDir.y, cDir.z, 0.0f);
			C.queryAtten[qs] = make_float4(cAtten.x, cAtten.y, cAtten.z, 0.0f);
			C.queryPath[qs] = p;
		}
		const uint32_t rs = appendSlot(doRecord, C.counts + kCacheRecords);
		if(doRecord && rs < C.recCapacity) {
			C.rec[4 * size_t(rs)] = make_float4(cPos.x, cPos.y, cPos.z, 0.0f);
			C.rec[4 * size_t(rs) + 1] = make_float4(cDir.x, cDir.y, cDir.z, 0.0f);
			C.rec[4 * size_t(rs) + 2] = make_float4(cColor.x, cColor.y, cColor.z, 0.0f);
			C.rec[4 * size_t(rs) + 3] = make_float4(cAtten.x, cAtten.y, cAtten.z, 0.0f);
			C.recPath[rs] = p;
		}
	}
}

template<bool CACHE>
__global__ __launch_bounds__(kPtBlock) void ptShade(DevScene S, PtFrame f, PtPaths P, const uint32_t * queueIn,
	uint32_t * queueOut, uint32_t * counts, uint32_t bounce, PtOptions opt, PtCache C) {
	shadeEntry<CACHE>(S, f, P, queueIn, queueOut, counts, bounce, opt, C, blockIdx.x * kPtBlock + threadIdx.x,
		counts[kCountActive], counts[kCountBase], counts[kCountSeed], 0u);
}

// Terminated paths add throughput x cache prediction.
__global__ __launch_bounds__(kPtBlock) void ptCacheApply(PtPaths P, PtCache C) {
	const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
	if(i >= C.counts[kCacheQueries])
		return;
	const uint32_t p = C.queryPath[i];
	const float4 a = C.queryAtten[i], l = C.pred[i], c = P.color[p];
	P.color[p] = make_float4(c.x + a.x * gmax(l.x, 0.0f), c.y + a.y * gmax(l.y, 0.0f), c.z + a.z * gmax(l.z, 0.0f), 0.0f);
}

// Records -> (query, target) pairs: target = (min(final colour, 5) - min(colour on arrival, 5)) /
// throughput, per channel (0 where the throughput is 0), clamped to [0, 100]: the radiance the rest
// of the path brought back, as seen through the estimator's per-sample clamp (PathTracer.cpp:288).
// Index % 8 == 7 is held out.
__global__ __launch_bounds__(kPtBlock) void ptCacheTargets(PtPaths P, PtCache C) {
	const uint32_t r = blockIdx.x * kPtBlock + threadIdx.x;
	const uint32_t nRec = C.counts[kCacheRecords] < C.recCapacity ? C.counts[kCacheRecords] : C.recCapacity;
	const bool valid = r < nRec;
	float4 q0 = {}, q1 = {}, t = {};
	if(valid) {
		const float4 cf = P.color[C.recPath[r]];
		const float4 cb = C.rec[4 * size_t(r) + 2], tb = C.rec[4 * size_t(r) + 3];
		const float d[3] = {gmin(cf.x, 5.0f) - gmin(cb.x, 5.0f), gmin(cf.y, 5.0f) - gmin(cb.y, 5.0f),
			gmin(cf.z, 5.0f) - gmin(cb.z, 5.0f)};
		const float w[3] = {tb.x, tb.y, tb.z};
		float o[3];
		for(int k = 0; k < 3; ++k) {
			const float v = w[k] > 0.0f ? d[k] / w[k] : 0.0f;
			o[k] = v > 0.0f ? (v < 100.0f ? v : 100.0f) : 0.0f;
		}
		q0 = C.rec[4 * size_t(r)];
		q1 = C.rec[4 * size_t(r) + 1];
		t = make_float4(o[0], o[1], o[2], 0.0f);
	}
	const bool hold = valid && (r & 7u) == 7u;
	const bool train = valid && !hold;
	const uint32_t ts = appendSlot(train, C.counts + kCacheTrain);
	const uint32_t hs = appendSlot(hold, C.counts + kCacheHold);
	if(train) {
		C.trainQ[2 * size_t(ts)] = q0;
		C.trainQ[2 * size_t(ts) + 1] = q1;
		C.trainT[ts] = t;
	}
	if(hold) {
		C.holdQ[2 * size_t(hs)] = q0;
		C.holdQ[2 * size_t(hs) + 1] = q1;
		C.holdT[hs] = t;
	}
}

__global__ void ptCacheReset(uint32_t * counts) {
	if(threadIdx.x < kCacheSlots)
		counts[threadIdx.x] = 0;
}

// Cell statistics: scale the running sums by kCacheDecay, then add this pass's held-out luminances.
__global__ __launch_bounds__(kPtBlock) void ptCacheDecay(float4 * cells, uint32_t n) {
	const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
	if(i < n) {
		const float4 c = cells[i];
		cells[i] = make_float4(kCacheDecay * c.x, kCacheDecay * c.y, kCacheDecay * c.z, kCacheDecay * c.w);
	}
}

__global__ __launch_bounds__(kPtBlock) void ptCacheHeldOut(PtCache C) {
	const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
	if(i >= C.counts[kCacheHold])
		return;
	const float4 q = C.holdQ[2 * size_t(i)];
	const f3 p = mk(q.x, q.y, q.z);
	const float4 pr = C.pred[i];
	const float lp = gmax(luminance(pr), 0.0f), lt = luminance(C.holdT[i]);
	const uint32_t cells[2] = {cacheCell(C.grid, p, kCacheCellRes),
		uint32_t(kCacheFineCells) + cacheCell(C.grid, p, kCacheCoarseRes)};
	for(uint32_t cell : cells) {
		float * c = reinterpret_cast<float *>(C.cells + cell);
		atomicAdd(c, lp);
		atomicAdd(c + 1, lt);
		atomicAdd(c + 2, 1.0f);
		atomicAdd(c + 3, lt * lt);
	}
}

// Visible shadow rays add their contribution (any-hit result per path).
__global__ __launch_bounds__(kPtBlock) void ptShadowApply(PtPaths P, const uint32_t * counts) {
	const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
	if(i >= counts[kCountShadow])
		return;
	const uint32_t p = P.shadowQueue[0][i];
	if(P.occluded[p])
		return;
	const float4 c = P.color[p], d = P.shadowContrib[p];
	P.color[p] = make_float4(c.x + d.x, c.y + d.y, c.z + d.z, 0.0f);
}

// Rendu's shadow walk through alpha cut-outs (PathTracer.cpp:124-157), one closest hit per step:
// no hit -> visible; an unmasked (or UV-less) occluder -> occluded; a transparent texel ->
// advance past it and query again.
__device__ __forceinline__ void walkEntry(const DevScene & S, const PtPaths & P, const uint32_t * queueIn,
	uint32_t * queueOut, uint32_t * counts, uint32_t i, uint32_t n) {
	bool again = false;
	uint32_t p = 0;
	if(i < n) {
		p = queueIn[i];
		const rdu_hit h = P.shadowHits[p];
		bool visible = false;
		if(!h.hit) {
			visible = true;
		} else {
			const DevObject obj = S.objects[h.mesh_id];
			const DevMaterial mat = S.materials[obj.mat];
			if(mat.masked && obj.useUV) {
				const float hw = 1.0f - h.u - h.v;
				const HitVerts hv = hitVerts(S, obj, h.local_id);
				const f2 luv = interp2(S.uv, hv.i0, hv.i1, hv.i2, hw, h.u, h.v);
				const float alpha = rgbal(S, mat.tex[0], luv.x, luv.y).w;
				if(alpha < 0.01f) {
					const float4 ro = P.shadowRays[2 * size_t(p)];
					const float4 rd = P.shadowRays[2 * size_t(p) + 1];
					const float maxDist = rd.w - h.t;
					const f3 lpos = mk(ro.x, ro.y, ro.z) + h.t * mk(rd.x, rd.y, rd.z);
					if(maxDist > 0.0f) {
						putRay(P.shadowRays, p, lpos, mk(rd.x, rd.y, rd.z), ro.w, maxDist); // the walk keeps its interval start
						again = true;
					} else {
						visible = true;
					}
				}
			}
		}
		if(visible) {
			const float4 c = P.color[p], d = P.shadowContrib[p];
			P.color[p] = make_float4(c.x + d.x, c.y + d.y, c.z + d.z, 0.0f);
		}
	}
	appendQueue(again, p, queueOut, counts + kCountShadowNext);
}

__global__ __launch_bounds__(kPtBlock) void ptShadowWalk(DevScene S, PtPaths P, const uint32_t * queueIn,
	uint32_t * queueOut, uint32_t * counts) {
	walkEntry(S, P, queueIn, queueOut, counts, blockIdx.x * kPtBlock + threadIdx.x, counts[kCountShadow]);
}

// image.rgb += min(color, 5) over the pass's samples in sample order (PathTracer.cpp:288).
__global__ __launch_bounds__(kPtBlock) void ptAccumulate(PtFrame f, PtPaths P, float * image, uint32_t components,
	uint32_t samplesInPass) {
	const uint32_t pix = blockIdx.x * kPtBlock + threadIdx.x;
	const uint32_t pixels = f.width * f.height;
	if(pix >= pixels)
		return;
	float * px = image + size_t(pix) * components;
	float r = px[0], g = px[1], b = px[2];
	for(uint32_t s = 0; s < samplesInPass; ++s) {
		const float4 c = P.color[size_t(s) * pixels + pix];
		r += gmin(c.x, 5.0f);
		g += gmin(c.y, 5.0f);
		b += gmin(c.z, 5.0f);
	}
	px[0] = r;
	px[1] = g;
	px[2] = b;
}

// color = render.rgb / float(samples) (PathTracer.cpp:291-292).
__global__ __launch_bounds__(kPtBlock) void ptNormalize(float * image, uint32_t pixels, uint32_t components, float spp) {
	const uint32_t pix = blockIdx.x * kPtBlock + threadIdx.x;
	if(pix >= pixels)
		return;
	float * px = image + size_t(pix) * components;
	px[0] = px[0] / spp;
	px[1] = px[1] / spp;
	px[2] = px[2] / spp;
}

inline unsigned blocksFor(size_t n) { return unsigned((n + kPtBlock - 1) / kPtBlock); }

// Refill threshold of the path tracer's traversal launches when no tuning sets one.
constexpr uint32_t kPassRefillMin = 24;

// ---- batch passes ---------------------------------------------------------------------------------

// Index of the last entry of a sorted array `begin` (of n >= 1 entries, begin[0] = 0) not above v.
__device__ __forceinline__ uint32_t lastNotAbove(const PtChunk * chunks, uint32_t n, uint32_t v, bool slots) {
	uint32_t lo = 0, hi = n; // answer in [lo, hi)
	while(hi - lo > 1) {
		const uint32_t mid = (lo + hi) >> 1;
		const uint32_t b = slots ? chunks[mid].slotBegin : chunks[mid].pathBegin;
		if(b <= v)
			lo = mid;
		else
			hi = mid;
	}
	return lo;
}

// Camera rays of every chunk's paths (ptGenerate per job), with the path's chunk and job.
__global__ __launch_bounds__(kPtBlock) void ptGenerateBatch(PtBatch B, PtPaths P, uint32_t nPaths) {
	const uint32_t p = blockIdx.x * kPtBlock + threadIdx.x;
	if(p >= nPaths)
		return;
	const uint32_t k = lastNotAbove(B.chunks, B.nChunks, p, false);
	const PtChunk c = B.chunks[k];
	const PtFrame f = B.jobs[c.job].f;
	const uint32_t pixels = f.width * f.height;
	const uint32_t local = p - c.pathBegin;
	const uint32_t s = local / pixels, pix = local % pixels;
	const uint32_t x = pix % f.width, y = pix / f.width;
	const f2 ndc = samplePos(f, c.seed, x, y, c.sampleBase + s);
	const f3 eye = f.eye;
	const f3 worldPos = f.corner + ndc.x * f.dx + ndc.y * f.dy;
	const f3 dir = normalize(worldPos - eye);
	putRay(P.rays, p, eye, dir, 0.0001f, 1e8f);
	P.atten[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
	P.color[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	P.mis[p] = make_float2(0.0f, 0.0f);
	P.queue[0][p] = p;
	B.pathChunk[p] = k;
	B.pathJob[p] = c.job;
}

// A batch job's descriptor read as a wavefront-uniform value: the descriptor is read through the
// constant address space (scalar loads, as a kernel argument would be), and the scene's pointers are
// tagged as global memory, as they are when they arrive as kernel arguments, and kept in scalar
// registers, so the scene's loads take a scalar base address.
template <class T> __device__ inline T * globalPtr(T * p) {
	const uint64_t v = uint64_t(p);
	const uint32_t lo = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(v))));
	const uint32_t hi = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(v >> 32))));
	return (T *)(__attribute__((address_space(1))) T *)((uint64_t(hi) << 32) | lo);
}

template <class T> __device__ inline const T & constantRef(const T * p) {
	return *(const T *)(const __attribute__((address_space(4))) T *)p;
}

// `j` must be the same in every active lane; it is read from the first one so that the compiler
// keeps the descriptor's address (and so its loads) scalar.
__device__ inline PtJob uniformJob(const PtJob * jobs, uint32_t j) {
	PtJob J = constantRef(jobs + __builtin_amdgcn_readfirstlane(j));
	DevScene & S = J.S;
	S.texels = globalPtr(S.texels);
	S.textures = globalPtr(S.textures);
	S.nrm = globalPtr(S.nrm);
	S.tan = globalPtr(S.tan);
	S.bit = globalPtr(S.bit);
	S.uv = globalPtr(S.uv);
	S.idx = globalPtr(S.idx);
	S.meshes = globalPtr(S.meshes);
	S.objects = globalPtr(S.objects);
	S.materials = globalPtr(S.materials);
	S.lights = globalPtr(S.lights);
	S.envPmf = globalPtr(S.envPmf);
	S.envRowCdf = globalPtr(S.envRowCdf);
	S.envColCdf = globalPtr(S.envColCdf);
	S.emTris = globalPtr(S.emTris);
	S.emCdf = globalPtr(S.emCdf);
	S.emBase = globalPtr(S.emBase);
	return J;
}

__device__ inline PtChunk uniformChunk(const PtChunk * chunks, uint32_t k) {
	return constantRef(chunks + __builtin_amdgcn_readfirstlane(k));
}

// ptShade over a batch queue. A wavefront serves the lanes of its first lane's chunk (whose chunk
// and job descriptors are then wavefront-uniform and read with scalar loads); lanes of other chunks
// (where a wavefront's queue slice straddles two chunks) go to the overflow list, which
// ptShadeBatchOverflow serves chunk by chunk. Paths of one chunk are contiguous in the queues, so
// the overflow is small, and the main kernel keeps the registers of a single-scene shading step.
__global__ __launch_bounds__(kPtBlock) void ptShadeBatch(const PtJob * __restrict__ jobs,
	const PtChunk * __restrict__ chunks, const uint32_t * __restrict__ pathChunk, uint32_t * __restrict__ overflow,
	PtPaths P, const uint32_t * queueIn, uint32_t * queueOut, uint32_t * counts, uint32_t bounce) {
	const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
	const uint32_t n = counts[kCountActive];
	const bool todo = i < n;
	const uint32_t k = todo ? pathChunk[queueIn[i]] : 0u;
	const uint64_t m = __ballot(todo);
	if(!m)
		return;
	const uint32_t leader = uint32_t(__ffsll((unsigned long long)m) - 1);
	const uint32_t kk = uint32_t(__builtin_amdgcn_readlane(int(k), int(leader)));
	const bool mine = todo && k == kk;
	appendQueue(todo && !mine, i, overflow, counts + kCountOverflow);
	if(mine) {
		const PtChunk c = uniformChunk(chunks, kk);
		const PtJob J = uniformJob(jobs, c.job);
		shadeEntry<false>(J.S, J.f, P, queueIn, queueOut, counts, bounce, J.opt, PtCache{}, i, n, c.sampleBase, c.seed,
			c.pathBegin);
	}
}

// The overflow entries of a bounce, one chunk at a time per wavefront. The list is usually a small
// fraction of the queue, so a small grid strides over it (the whole grid's loop bound is the
// same device count).
constexpr unsigned kOverflowBlocks = 512;
__global__ __launch_bounds__(kPtBlock) void ptShadeBatchOverflow(const PtJob * __restrict__ jobs,
	const PtChunk * __restrict__ chunks, const uint32_t * __restrict__ pathChunk, const uint32_t * __restrict__ overflow,
	PtPaths P, const uint32_t * queueIn, uint32_t * queueOut, uint32_t * counts, uint32_t bounce) {
	const uint32_t n = counts[kCountActive];
	const uint32_t total = counts[kCountOverflow];
	const uint32_t stride = gridDim.x * kPtBlock;
	for(uint32_t base = blockIdx.x * kPtBlock + (threadIdx.x & ~63u); base < total; base += stride) {
		const uint32_t o = base + (threadIdx.x & 63u);
		bool todo = o < total;
		const uint32_t i = todo ? overflow[o] : 0u;
		const uint32_t k = todo ? pathChunk[queueIn[i]] : 0u;
		for(;;) {
			const uint64_t m = __ballot(todo);
			if(!m)
				break;
			const uint32_t leader = uint32_t(__ffsll((unsigned long long)m) - 1);
			const uint32_t kk = uint32_t(__builtin_amdgcn_readlane(int(k), int(leader)));
			if(todo && k == kk) {
				const PtChunk c = uniformChunk(chunks, kk);
				const PtJob J = uniformJob(jobs, c.job);
				shadeEntry<false>(J.S, J.f, P, queueIn, queueOut, counts, bounce, J.opt, PtCache{}, i, n, c.sampleBase,
					c.seed, c.pathBegin);
				todo = false;
			}
		}
	}
}

// ptShadowWalk over a batch queue, one job at a time per wavefront.
__global__ __launch_bounds__(kPtBlock) void ptShadowWalkBatch(PtBatch B, PtPaths P, const uint32_t * queueIn,
	uint32_t * queueOut, uint32_t * counts) {
	const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
	const uint32_t n = counts[kCountShadow];
	bool todo = i < n;
	const uint32_t j = todo ? B.pathJob[queueIn[i]] : 0u;
	for(;;) {
		const uint64_t m = __ballot(todo);
		if(!m)
			break;
		const uint32_t leader = uint32_t(__ffsll((unsigned long long)m) - 1);
		const uint32_t jj = uint32_t(__builtin_amdgcn_readlane(int(j), int(leader)));
		if(todo && j == jj) {
			walkEntry(uniformJob(B.jobs, jj).S, P, queueIn, queueOut, counts, i, n);
			todo = false;
		}
	}
}

// ptAccumulate for every chunk: slot t is pixel t - slotBegin of its chunk's job image; the chunk's
// samples are added in sample order.
__global__ __launch_bounds__(kPtBlock) void ptAccumulateBatch(PtBatch B, PtPaths P) {
	const uint32_t t = blockIdx.x * kPtBlock + threadIdx.x;
	if(t >= B.nSlots)
		return;
	const uint32_t k = lastNotAbove(B.chunks, B.nChunks, t, true);
	const PtChunk c = B.chunks[k];
	const PtJob & J = B.jobs[c.job];
	const uint32_t pixels = J.f.width * J.f.height;
	const uint32_t pix = t - c.slotBegin;
	float * px = J.image + size_t(pix) * J.components;
	float r = px[0], g = px[1], b = px[2];
	for(uint32_t s = 0; s < c.samples; ++s) {
		const float4 col = P.color[size_t(c.pathBegin) + size_t(s) * pixels + pix];
		r += gmin(col.x, 5.0f);
		g += gmin(col.y, 5.0f);
		b += gmin(col.z, 5.0f);
	}
	px[0] = r;
	px[1] = g;
	px[2] = b;
}

// A tree's 4-wide nodes (8 float4: six plane records, the four child codes, padding) copied into an
// arena: node codes move by nodeOffset, leaf codes ~(first << 1 | count - 1) by triOffset triangles.
__global__ __launch_bounds__(kPtBlock) void ptRebaseWideKernel(float4 * __restrict__ dst, const float4 * __restrict__ src,
	uint32_t nodes, int32_t nodeOffset, int32_t triOffset) {
	const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
	if(i >= nodes * 8u)
		return;
	float4 v = src[i];
	if((i & 7u) == 6u) {
		int32_t c[4] = {__float_as_int(v.x), __float_as_int(v.y), __float_as_int(v.z), __float_as_int(v.w)};
		for(int k = 0; k < 4; ++k) {
			if(c[k] == kWideEmpty)
				continue;
			c[k] = c[k] >= 0 ? c[k] + nodeOffset : c[k] - 2 * triOffset;
		}
		v = make_float4(__int_as_float(c[0]), __int_as_float(c[1]), __int_as_float(c[2]), __int_as_float(c[3]));
	}
	dst[i] = v;
}

} // namespace

int ptSetSampleBase(uint32_t * counts, uint32_t base, uint32_t seed, hipStream_t stream) {
	hipLaunchKernelGGL(ptSetBase, dim3(1), dim3(64), 0, stream, counts, base, seed);
	return int(hipGetLastError());
}

int ptRenderPass(const DevScene & S, const DevBVH & bvh, const PtFrame & f, const PtPaths & P, uint32_t samplesInPass,
	const PtOptions & opt, float * dImage, uint32_t components, PtScratch & scratch, const PtCache * cache,
	const TraceTuning & tune, hipStream_t stream, hipEvent_t accumulateAfter) {
	const uint32_t pixels = f.width * f.height;
	const uint32_t nPaths = pixels * samplesInPass;
	const PtCache C = cache ? *cache : PtCache{};
	// Queue-ordered traversal defaults: 64-ray interleaved runs per wavefront (a die's wavefronts
	// then share one window of neighbouring paths, and so one part of the tree in that die's L2;
	// on the 1080p single frame 5.6 ms per 2-spp pass against 10.1 ms with contiguous runs) and a
	// refill once 24 lanes idle (the frame with three passes in flight: 0.139-0.143 s with any
	// threshold from 8 to 48; the 64-job batch call: 1.03-1.05 s at 24 against 1.12-1.14 s at 8).
	TraceTuning tn = tune;
	if(tn.chunk == 0)
		tn.chunk = 64;
	if(tn.refillMin == 0)
		tn.refillMin = kPassRefillMin;
	hipLaunchKernelGGL(ptResetCounts, dim3(1), dim3(64), 0, stream, P.counts, nPaths, samplesInPass);
	if(cache)
		hipLaunchKernelGGL(ptCacheReset, dim3(1), dim3(64), 0, stream, C.counts);
	hipLaunchKernelGGL(ptGenerate, dim3(blocksFor(nPaths)), dim3(kPtBlock), 0, stream, f, P, nPaths);
	int q = 0;
	for(uint32_t b = 0; b < opt.depth; ++b) {
		TraceOptions to;
		applyTuning(to, tn);
		to.wide = true;
		to.nearFar = b > 0; // camera rays are coherent, the later bounces are not
		to.perm = P.queue[q];
		to.countDev = P.counts + kCountActive;
		int e = launchClosest(bvh, P.rays, nPaths, P.hits, to, scratch.gstack, nullptr, stream);
		if(e)
			return e;
		if(cache)
			hipLaunchKernelGGL(ptShade<true>, dim3(blocksFor(nPaths)), dim3(kPtBlock), 0, stream, S, f, P, P.queue[q],
				P.queue[1 - q], P.counts, b, opt, C);
		else
			hipLaunchKernelGGL(ptShade<false>, dim3(blocksFor(nPaths)), dim3(kPtBlock), 0, stream, S, f, P, P.queue[q],
				P.queue[1 - q], P.counts, b, opt, C);
		// Shadow rays exist when there are analytic lights, or area lights under the extended estimator.
		if(S.nLights > 0 || (opt.estimator == RDU_ESTIMATOR_EXTENDED && S.pArea > 0.0f)) {
			if(!S.anyMasked) {
				TraceOptions so;
				applyTuning(so, tn);
				so.wide = true;
				so.nearFar = true; // shadow rays leave scattered surface points: incoherent
				so.perm = P.shadowQueue[0];
				so.countDev = P.counts + kCountShadow;
				e = launchAny(bvh, P.shadowRays, nPaths, P.occluded, so, scratch.gstack, nullptr, stream);
				if(e)
					return e;
				hipLaunchKernelGGL(ptShadowApply, dim3(blocksFor(nPaths)), dim3(kPtBlock), 0, stream, P, P.counts);
			} else {
				// Cut-out walk: each step is a closest-hit query on the shadow rays still walking
				// (their count stays on the device). Steps run in rounds of maxShadowLayers; after
				// each round the host reads how many rays walk on and stops at none, so a ray
				// crosses any number of transparent layers, as Rendu's loop does.
				int layer = 0;
				for(;;) {
					for(int k = 0; k < opt.maxShadowLayers; ++k, ++layer) {
						TraceOptions so;
						applyTuning(so, tn);
						so.wide = true;
						so.perm = P.shadowQueue[layer & 1];
						so.countDev = P.counts + kCountShadow;
						e = launchClosest(bvh, P.shadowRays, nPaths, P.shadowHits, so, scratch.gstack, nullptr, stream);
						if(e)
							return e;
						hipLaunchKernelGGL(ptShadowWalk, dim3(blocksFor(nPaths)), dim3(kPtBlock), 0, stream, S, P,
							P.shadowQueue[layer & 1], P.shadowQueue[1 - (layer & 1)], P.counts);
						hipLaunchKernelGGL(ptShadowSwap, dim3(1), dim3(64), 0, stream, P.counts, P.totals);
					}
					if(!scratch.walkLeft)
						return int(hipErrorInvalidValue); // masked scenes need the host count
					hipError_t he = hipMemcpyAsync(scratch.walkLeft, P.counts + kCountShadow, sizeof(uint32_t),
						hipMemcpyDeviceToHost, stream);
					if(he == hipSuccess)
						he = hipStreamSynchronize(stream);
					if(he != hipSuccess)
						return int(he);
					if(*scratch.walkLeft == 0)
						break;
				}
			}
		}
		hipLaunchKernelGGL(ptAdvance, dim3(1), dim3(64), 0, stream, P.counts, P.totals);
		q = 1 - q;
	}
	if(cache) {
		// Terminated paths: one inference over the compacted queries (count read on the device).
		int e = rcInfer(C.grid, *C.params, C.query, C.counts + kCacheQueries, nPaths, C.pred, stream);
		if(e)
			return e;
		hipLaunchKernelGGL(ptCacheApply, dim3(blocksFor(nPaths)), dim3(kPtBlock), 0, stream, P, C);
		hipLaunchKernelGGL(ptCacheTargets, dim3(blocksFor(C.recCapacity)), dim3(kPtBlock), 0, stream, P, C);
	}
	if(accumulateAfter) {
		const hipError_t he = hipStreamWaitEvent(stream, accumulateAfter, 0);
		if(he != hipSuccess)
			return int(he);
	}
	hipLaunchKernelGGL(ptAccumulate, dim3(blocksFor(pixels)), dim3(kPtBlock), 0, stream, f, P, dImage, components,
		samplesInPass);
	if(cache) {
		const uint32_t cells = uint32_t(kRcCells);
		const uint32_t holdCap = C.recCapacity / 8u + 1u;
		int e = rcInfer(C.grid, *C.params, C.holdQ, C.counts + kCacheHold, holdCap, C.pred, stream);
		if(e)
			return e;
		hipLaunchKernelGGL(ptCacheDecay, dim3(blocksFor(cells)), dim3(kPtBlock), 0, stream, C.cells, cells);
		hipLaunchKernelGGL(ptCacheHeldOut, dim3(blocksFor(holdCap)), dim3(kPtBlock), 0, stream, C);
	}
	return int(hipGetLastError());
}

int ptRebaseWide(float4 * dst, const float4 * src, uint32_t nodes, int32_t nodeOffset, int32_t triOffset,
	hipStream_t stream) {
	if(nodes == 0)
		return 0;
	hipLaunchKernelGGL(ptRebaseWideKernel, dim3(blocksFor(size_t(nodes) * 8)), dim3(kPtBlock), 0, stream, dst, src, nodes,
		nodeOffset, triOffset);
	return int(hipGetLastError());
}

int ptRenderBatchPass(const PtBatch & B, const DevBVH & arena, const PtPaths & P, uint32_t nPaths, uint32_t maxDepth,
	bool shadows, bool masked, int maxShadowLayers, PtScratch & scratch, const TraceTuning & tune, hipStream_t stream,
	hipEvent_t accumulateAfter) {
	if(nPaths == 0 || B.nChunks == 0)
		return 0;
	TraceTuning tn = tune; // the defaults of ptRenderPass
	if(tn.chunk == 0)
		tn.chunk = 64;
	if(tn.refillMin == 0)
		tn.refillMin = kPassRefillMin;
	hipLaunchKernelGGL(ptResetCounts, dim3(1), dim3(64), 0, stream, P.counts, nPaths, 0u);
	hipLaunchKernelGGL(ptGenerateBatch, dim3(blocksFor(nPaths)), dim3(kPtBlock), 0, stream, B, P, nPaths);
	int q = 0;
	for(uint32_t b = 0; b < maxDepth; ++b) {
		TraceOptions to;
		applyTuning(to, tn);
		to.wide = true;
		to.nearFar = b > 0;
		to.perm = P.queue[q];
		to.countDev = P.counts + kCountActive;
		to.rayJob = B.pathJob;
		to.jobRoots = B.roots;
		int e = launchClosest(arena, P.rays, nPaths, P.hits, to, scratch.gstack, nullptr, stream);
		if(e)
			return e;
		hipLaunchKernelGGL(ptShadeBatch, dim3(blocksFor(nPaths)), dim3(kPtBlock), 0, stream, B.jobs, B.chunks, B.pathChunk,
			B.overflow, P, P.queue[q], P.queue[1 - q], P.counts, b);
		hipLaunchKernelGGL(ptShadeBatchOverflow, dim3(std::min(blocksFor(nPaths), kOverflowBlocks)), dim3(kPtBlock), 0,
			stream, B.jobs, B.chunks, B.pathChunk, B.overflow, P, P.queue[q], P.queue[1 - q], P.counts, b);
		if(shadows) {
			if(!masked) {
				TraceOptions so;
				applyTuning(so, tn);
				so.wide = true;
				so.nearFar = true;
				so.perm = P.shadowQueue[0];
				so.countDev = P.counts + kCountShadow;
				so.rayJob = B.pathJob;
				so.jobRoots = B.roots;
				e = launchAny(arena, P.shadowRays, nPaths, P.occluded, so, scratch.gstack, nullptr, stream);
				if(e)
					return e;
				hipLaunchKernelGGL(ptShadowApply, dim3(blocksFor(nPaths)), dim3(kPtBlock), 0, stream, P, P.counts);
			} else {
				int layer = 0;
				for(;;) {
					for(int k = 0; k < maxShadowLayers; ++k, ++layer) {
						TraceOptions so;
						applyTuning(so, tn);
						so.wide = true;
						so.perm = P.shadowQueue[layer & 1];
						so.countDev = P.counts + kCountShadow;
						so.rayJob = B.pathJob;
						so.jobRoots = B.roots;
						e = launchClosest(arena, P.shadowRays, nPaths, P.shadowHits, so, scratch.gstack, nullptr, stream);
						if(e)
							return e;
						hipLaunchKernelGGL(ptShadowWalkBatch, dim3(blocksFor(nPaths)), dim3(kPtBlock), 0, stream, B, P,
							P.shadowQueue[layer & 1], P.shadowQueue[1 - (layer & 1)], P.counts);
						hipLaunchKernelGGL(ptShadowSwap, dim3(1), dim3(64), 0, stream, P.counts, P.totals);
					}
					if(!scratch.walkLeft)
						return int(hipErrorInvalidValue);
					hipError_t he = hipMemcpyAsync(scratch.walkLeft, P.counts + kCountShadow, sizeof(uint32_t),
						hipMemcpyDeviceToHost, stream);
					if(he == hipSuccess)
						he = hipStreamSynchronize(stream);
					if(he != hipSuccess)
						return int(he);
					if(*scratch.walkLeft == 0)
						break;
				}
			}
		}
		hipLaunchKernelGGL(ptAdvance, dim3(1), dim3(64), 0, stream, P.counts, P.totals);
		q = 1 - q;
	}
	if(accumulateAfter) {
		const hipError_t he = hipStreamWaitEvent(stream, accumulateAfter, 0);
		if(he != hipSuccess)
			return int(he);
	}
	hipLaunchKernelGGL(ptAccumulateBatch, dim3(blocksFor(B.nSlots)), dim3(kPtBlock), 0, stream, B, P);
	return int(hipGetLastError());
}

int ptNormalizeImage(float * dImage, uint32_t pixels, uint32_t components, uint32_t spp, hipStream_t stream) {
	hipLaunchKernelGGL(ptNormalize, dim3(blocksFor(pixels)), dim3(kPtBlock), 0, stream, dImage, pixels, components, float(spp));
	return int(hipGetLastError());
}

} // namespace rdu
