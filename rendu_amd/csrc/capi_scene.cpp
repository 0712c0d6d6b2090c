// C ABI of the path tracer (include/rdu.h): the drop-in boundary for Rendu's PathTracer
// (src/apps/pathtracer/PathTracer.hpp:10-76). Host bookkeeping only; rendering runs in pt.hip.
#include "pt.hpp"
#include "raycaster_impl.hpp"

#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <vector>

namespace rdu {

// Working memory of one render of the wavefront path tracer: path state, ray and shadow queues,
// counters and the device image. It belongs to a render, not to a scene: a render leases a set
// from a per-process pool and gives it back when it returns, so the process holds one set per
// render that ran at the same time (each grown to the largest pass it served and kept for reuse)
// instead of one per scene. Batches of many scenes then reuse memory instead of asking the driver
// for hundreds of megabytes per scene.
struct RenderPaths {
	int device = 0;
	DevBuffer pRays, pHits, pAtten, pColor, pQ0, pQ1, pCounts, pTotals, pSRays, pSContrib, pSQ0, pSQ1, pOcc, pSHits,
		pGstack, pImage, pMis;
	DevBuffer pChunk, pJob, pOverflow; // batch passes: each path's chunk and job, the shading overflow list
	size_t pathCap = 0;             // paths every path buffer above holds (0 while a growth is unfinished)
	uint32_t * walkLeft = nullptr;  // pinned host word for the shadow walk's count (scenes with cut-outs)
};

namespace {
std::mutex pathsMutex;
// Every set ever made. Never destroyed: the sets hold device memory until the process ends, and
// freeing it from a static destructor could run after the HIP runtime has shut down.
std::vector<std::unique_ptr<RenderPaths>> & pathsAll = *new std::vector<std::unique_ptr<RenderPaths>>();
std::vector<RenderPaths *> pathsFree;

// A free set of `device` for passes of P paths: `prefer` (the set the caller's scene used last,
// so that a captured pass finds its buffers again) if it is free and holds P paths, else the
// smallest one that already holds P paths, else the largest (which then grows), else a new one.
RenderPaths * leasePaths(int device, size_t P, const RenderPaths * prefer) {
	std::lock_guard<std::mutex> g(pathsMutex);
	for(size_t i = 0; prefer && i < pathsFree.size(); ++i)
		if(pathsFree[i] == prefer && prefer->device == device && prefer->pathCap >= P) {
			RenderPaths * r = pathsFree[i];
			pathsFree.erase(pathsFree.begin() + std::ptrdiff_t(i));
			return r;
		}
	size_t pick = pathsFree.size();
	for(size_t i = 0; i < pathsFree.size(); ++i) {
		const RenderPaths * c = pathsFree[i];
		if(c->device != device)
			continue;
		if(pick == pathsFree.size()) {
			pick = i;
			continue;
		}
		const RenderPaths * b = pathsFree[pick];
		const bool cFits = c->pathCap >= P, bFits = b->pathCap >= P;
		if((cFits && (!bFits || c->pathCap < b->pathCap)) || (!cFits && !bFits && c->pathCap > b->pathCap))
			pick = i;
	}
	if(pick < pathsFree.size()) {
		RenderPaths * r = pathsFree[pick];
		pathsFree.erase(pathsFree.begin() + std::ptrdiff_t(pick));
		return r;
	}
	pathsAll.push_back(std::make_unique<RenderPaths>());
	pathsAll.back()->device = device;
	return pathsAll.back().get();
}

// Holds a leased set for the duration of a render. On return the render's stream is drained first
// (already the case after a successful render; an error may leave work in flight).
class PathsLease {
public:
	PathsLease(int device, size_t P, hipStream_t stream, const RenderPaths * prefer = nullptr)
		: _r(leasePaths(device, P, prefer)), _stream(stream) {}
	~PathsLease() {
		(void)hipStreamSynchronize(_stream);
		std::lock_guard<std::mutex> g(pathsMutex);
		pathsFree.push_back(_r);
	}
	PathsLease(const PathsLease &) = delete;
	PathsLease & operator=(const PathsLease &) = delete;
	RenderPaths * operator->() const { return _r; }
	RenderPaths & operator*() const { return *_r; }

private:
	RenderPaths * _r;
	hipStream_t _stream;
};

// The kernels' view of a set's path buffers.
PtPaths pathsOf(const RenderPaths & R) {
	PtPaths P;
	P.rays = static_cast<float4 *>(R.pRays.ptr);
	P.hits = static_cast<rdu_hit *>(R.pHits.ptr);
	P.atten = static_cast<float4 *>(R.pAtten.ptr);
	P.color = static_cast<float4 *>(R.pColor.ptr);
	P.queue[0] = static_cast<uint32_t *>(R.pQ0.ptr);
	P.queue[1] = static_cast<uint32_t *>(R.pQ1.ptr);
	P.counts = static_cast<uint32_t *>(R.pCounts.ptr);
	P.totals = static_cast<unsigned long long *>(R.pTotals.ptr);
	P.shadowRays = static_cast<float4 *>(R.pSRays.ptr);
	P.shadowContrib = static_cast<float4 *>(R.pSContrib.ptr);
	P.shadowQueue[0] = static_cast<uint32_t *>(R.pSQ0.ptr);
	P.shadowQueue[1] = static_cast<uint32_t *>(R.pSQ1.ptr);
	P.occluded = static_cast<uint8_t *>(R.pOcc.ptr);
	P.shadowHits = static_cast<rdu_hit *>(R.pSHits.ptr);
	P.mis = static_cast<float2 *>(R.pMis.ptr);
	return P;
}

// Events of one call, destroyed when it returns.
struct EventSet {
	std::vector<hipEvent_t> all;
	int make(hipEvent_t & e, unsigned flags) {
		const hipError_t he = hipEventCreateWithFlags(&e, flags);
		if(he == hipSuccess)
			all.push_back(e);
		return he == hipSuccess ? 0 : int(he);
	}
	~EventSet() {
		for(hipEvent_t e : all)
			(void)hipEventDestroy(e);
	}
};
} // namespace

} // namespace rdu

struct rdu_scene {
	struct Tex {
		uint32_t w, h, offset;
	};
	struct MeshRec {
		uint32_t vtx, idx, nv, ni;
		bool hasUV;
	};
	std::vector<float> texels; // RGBA
	std::vector<Tex> textures;
	std::vector<float> positions, nrm, tan, bit, uv;
	std::vector<uint32_t> indices;
	std::vector<MeshRec> meshes;
	std::vector<rdu_material_desc> materials;
	std::vector<rdu_object_desc> objects;
	std::vector<rdu_light_desc> lights;
	rdu_background_desc bg{};
	bool committed = false;
	bool gpuBuild = false; // commit builds the tree with rdu_raycaster_update_hierarchy_gpu
	rdu_raycaster * rc = nullptr;
	rdu::DevScene dev;
	rdu::DevBuffer dTexels, dTextures, dNrm, dTan, dBit, dUv, dIdx, dMeshes, dObjects, dMaterials, dLights;
	rdu::DevBuffer dEnvPmf, dEnvRow, dEnvCol, dEmTris, dEmCdf, dEmBase; // extended estimator's light sampling
	// Radiance-cache mode buffers, grown on demand.
	rdu::DevBuffer cQuery, cQAtten, cQPath, cPred, cRec, cRecPath, cTrainQ, cTrainT, cHoldQ, cHoldT, cCounts;
	size_t cacheCap = 0;
	rdu_render_stats last{};
	rdu_cache_stats lastCache{};
	size_t lastTrainPairs = 0; // pairs left in cTrainQ / cTrainT by the last cached pass
	hipStream_t stream = nullptr; // renders of this scene run here (created on first render)
	hipStream_t laneStreams[2] = {nullptr, nullptr}; // further passes in flight of a plain render
	// Instanced scenes (rdu_scene_set_instancing): the raycaster holds each mesh once in object space
	// with a tree per mesh; a top level over the objects' boxes and the objects' inverse models make
	// the tree the path tracer traverses.
	bool instancing = false;
	rdu::DevBuffer dInstWide, dInst;
	// Host copies of what dInstWide and dInst hold: batch renders rebase them into their arena.
	std::vector<float> instWideHost;
	std::vector<float4> instTableHost;
	rdu::DevBVH ibvh;
	const rdu::DevBVH & traceBvh() const { return instancing ? ibvh : rc->dev; }
	const rdu::RenderPaths * lastPaths = nullptr;     // path set of the last render (sets live until exit)
	// Graph replay of full passes (rdu_scene_set_graph): the captured pass, the launch
	// configuration it was captured with, and that of the last eager full pass.
	bool useGraph = false;
	hipGraph_t graph = nullptr;
	hipGraphExec_t graphExec = nullptr;
	std::vector<unsigned char> graphKey, lastPassKey;
	// Autotuned traversal parameters of this scene's passes (rdu_scene_autotune).
	rdu::TraceTuning tuned;
	bool tunedSet = false;
	void dropGraph() {
		if(graphExec)
			(void)hipGraphExecDestroy(graphExec);
		if(graph)
			(void)hipGraphDestroy(graph);
		graphExec = nullptr;
		graph = nullptr;
		graphKey.clear();
		lastPassKey.clear();
	}
	~rdu_scene() {
		dropGraph();
		// Renders run on `stream` (uploads are synchronous): once it drains, every block of the
		// scene is idle and is freed without waiting for the rest of the device.
		if(stream)
			(void)hipStreamSynchronize(stream);
		for(hipStream_t ls : laneStreams)
			if(ls)
				(void)hipStreamSynchronize(ls);
		for(rdu::DevBuffer * b : {&dTexels, &dTextures, &dNrm, &dTan, &dBit, &dUv, &dIdx, &dMeshes, &dObjects, &dMaterials,
				 &dLights, &dEnvPmf, &dEnvRow, &dEnvCol, &dEmTris, &dEmCdf, &dEmBase, &dInstWide, &dInst, &cQuery, &cQAtten, &cQPath, &cPred, &cRec,
				 &cRecPath, &cTrainQ, &cTrainT, &cHoldQ, &cHoldT, &cCounts})
			b->releaseOn(nullptr);
		if(rc)
			rdu_raycaster_destroy(rc);
		for(hipStream_t ls : laneStreams)
			if(ls)
				(void)hipStreamDestroy(ls);
		if(stream)
			(void)hipStreamDestroy(stream);
	}
};

using namespace rdu;

namespace {

// glm inverse(mat3) (func_matrix.inl:269-291) of transpose(mat3(model)), column-major out.
void inverseTransposeUpper(const float * m4, float * out) {
	float m[3][3]; // m[c][r] = transpose(mat3(model))[c][r] = model[r][c]
	for(int c = 0; c < 3; ++c)
		for(int r = 0; r < 3; ++r)
			m[c][r] = m4[4 * r + c];
	const float ood = 1.0f / (+m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) - m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
								 m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]));
	float inv[3][3];
	inv[0][0] = +(m[1][1] * m[2][2] - m[2][1] * m[1][2]) * ood;
	inv[1][0] = -(m[1][0] * m[2][2] - m[2][0] * m[1][2]) * ood;
	inv[2][0] = +(m[1][0] * m[2][1] - m[2][0] * m[1][1]) * ood;
	inv[0][1] = -(m[0][1] * m[2][2] - m[2][1] * m[0][2]) * ood;
	inv[1][1] = +(m[0][0] * m[2][2] - m[2][0] * m[0][2]) * ood;
	inv[2][1] = -(m[0][0] * m[2][1] - m[2][0] * m[0][1]) * ood;
	inv[0][2] = +(m[0][1] * m[1][2] - m[1][1] * m[0][2]) * ood;
	inv[1][2] = -(m[0][0] * m[1][2] - m[1][0] * m[0][2]) * ood;
	inv[2][2] = +(m[0][0] * m[1][1] - m[1][0] * m[0][1]) * ood;
	for(int c = 0; c < 3; ++c)
		for(int r = 0; r < 3; ++r)
			out[3 * c + r] = inv[c][r];
}

bool texOk(const rdu_scene * s, int32_t t) { return t >= 0 && size_t(t) < s->textures.size(); }

template<typename T>
int uploadVec(DevBuffer & b, const std::vector<T> & v) {
	return b.upload(v.data(), v.size() * sizeof(T));
}

// Sizes a leased set for passes of P paths over this scene's tree (the spill stack depends on the
// tree's depth, so it is checked on every render). Growth is ordered on the render's stream: the
// set's previous render drained its own stream before the lease, so the old blocks are idle. While
// a growth is unfinished pathCap is 0, so a set whose growth failed half-way (some buffers freed,
// some small) is regrown in full by the next render that leases it instead of being used as is.
int ensurePathsFor(RenderPaths & R, const DevBVH & bvh, bool masked, size_t P, bool batch, hipStream_t stream);
int ensurePaths(RenderPaths & R, const rdu_scene * s, size_t P, hipStream_t stream) {
	return ensurePathsFor(R, s->traceBvh(), s->dev.anyMasked != 0, P, false, stream);
}
int ensurePathsFor(RenderPaths & R, const DevBVH & bvh, bool masked, size_t P, bool batch, hipStream_t stream) {
	int st;
	if(P > R.pathCap) {
		R.pathCap = 0;
		if((st = R.pRays.reserveOn(P * 32, stream)) || (st = R.pHits.reserveOn(P * sizeof(rdu_hit), stream)) ||
			(st = R.pAtten.reserveOn(P * 16, stream)) || (st = R.pColor.reserveOn(P * 16, stream)) ||
			(st = R.pQ0.reserveOn(P * 4, stream)) || (st = R.pQ1.reserveOn(P * 4, stream)) ||
			(st = R.pCounts.reserveOn(kCountSlots * 4, stream)) || (st = R.pTotals.reserveOn(2 * 8, stream)) ||
			(st = R.pSRays.reserveOn(P * 32, stream)) || (st = R.pSContrib.reserveOn(P * 16, stream)) ||
			(st = R.pSQ0.reserveOn(P * 4, stream)) || (st = R.pSQ1.reserveOn(P * 4, stream)) ||
			(st = R.pOcc.reserveOn(P, stream)) || (st = R.pSHits.reserveOn(P * sizeof(rdu_hit), stream)) ||
			(st = R.pMis.reserveOn(P * 8, stream)))
			return st;
		R.pathCap = P;
	}
	if(batch && (R.pChunk.bytes < P * 4 || R.pJob.bytes < P * 4 || R.pOverflow.bytes < P * 4) &&
		((st = R.pChunk.reserveOn(P * 4, stream)) || (st = R.pJob.reserveOn(P * 4, stream)) ||
			(st = R.pOverflow.reserveOn(P * 4, stream))))
		return st;
	if(!R.walkLeft && masked &&
		(st = hipCheck(hipHostMalloc(reinterpret_cast<void **>(&R.walkLeft), sizeof(uint32_t), hipHostMallocDefault),
			 "hipHostMalloc")))
		return st;
	TraceOptions o;
	o.wide = true;
	const size_t g = globalStackInts(bvh, o, P);
	if(g && (st = R.pGstack.reserveOn(g * sizeof(int), stream)))
		return st;
	return RDU_OK;
}

// Cache-mode buffers for passes of up to P paths with `records` training records, on the scene's
// stream (the only one their users run on). cacheCap is 0 while a growth is unfinished.
int ensureCacheBuffers(rdu_scene * s, size_t P, size_t records) {
	if(P <= s->cacheCap && records <= s->cRecPath.bytes / 4)
		return RDU_OK;
	const size_t hold = records / 8 + 1;
	const size_t pred = P > hold ? P : hold;
	hipStream_t q = s->stream;
	s->cacheCap = 0;
	int st;
	if((st = s->cQuery.reserveOn(P * 32, q)) || (st = s->cQAtten.reserveOn(P * 16, q)) || (st = s->cQPath.reserveOn(P * 4, q)) ||
		(st = s->cPred.reserveOn(pred * 16, q)) || (st = s->cRec.reserveOn(records * 64, q)) ||
		(st = s->cRecPath.reserveOn(records * 4, q)) || (st = s->cTrainQ.reserveOn(records * 32, q)) ||
		(st = s->cTrainT.reserveOn(records * 16, q)) || (st = s->cHoldQ.reserveOn(hold * 32, q)) ||
		(st = s->cHoldT.reserveOn(hold * 16, q)) || (st = s->cCounts.reserveOn(kCacheSlots * 4, q)))
		return st;
	s->cacheCap = P;
	return RDU_OK;
}


// Extension (extended estimator): the light-sampling tables, built in double precision in a fixed
// order: lat-long texel weights = the largest luminance of the texel's bilinear footprint times
// sin(theta) at the texel centre (row and column CDFs); emissive triangles (objects in order, world
// vertices as the raycaster transformed them) weighted by area times the mean luminance of their
// emission texture; environment vs triangles 1/2 each when both exist, area lights vs analytic
// lights 1/2 each when both exist.
// A mesh's vertices under an object's model, in the raycaster's own arithmetic (HostScene::addMesh:
// glm mat4 * vec4 order, identity models copied as they are).
std::vector<f3> worldVertices(const float * xyz, size_t nv, const float * model) {
	std::vector<f3> v(nv);
	bool identity = true;
	for(int c = 0; c < 4; ++c)
		for(int r = 0; r < 4; ++r)
			identity = identity && (model[4 * c + r] == (c == r ? 1.0f : 0.0f));
	for(size_t i = 0; i < nv; ++i) {
		const f3 p = mk(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
		if(identity) {
			v[i] = p;
			continue;
		}
		float o[3];
		for(int r = 0; r < 3; ++r) {
			const float a0 = model[r] * p.x + model[4 + r] * p.y;
			const float a1 = model[8 + r] * p.z + model[12 + r] * 1.0f;
			o[r] = a0 + a1;
		}
		v[i] = mk(o[0], o[1], o[2]);
	}
	return v;
}

struct AreaTables {
	std::vector<float> envPmf, envRow, envCol, emCdf;
	std::vector<float4> emTris;
	std::vector<int32_t> emBase;
	int32_t envW = 0, envH = 0;
	float pEnv = 0.0f, pArea = 0.0f;
};

inline double lumd(double r, double g, double b) { return 0.2126 * r + 0.7152 * g + 0.0722 * b; }

AreaTables buildAreaTables(const rdu_scene * s) {
	AreaTables A;
	auto texel = [&](const rdu_scene::Tex & t, int x, int y) {
		return s->texels.data() + 4 * (size_t(t.offset) + size_t(y) * t.w + size_t(x));
	};
	double envPower = 0.0;
	if(s->bg.mode == RDU_BACKGROUND_LATLONG) {
		const rdu_scene::Tex & t = s->textures[size_t(s->bg.texture)];
		const int W = int(t.w), H = int(t.h);
		std::vector<double> f(size_t(W) * size_t(H)), rows(static_cast<size_t>(H));
		double total = 0.0;
		for(int j = 0; j < H; ++j) {
			const double sinT = std::sin(3.14159265358979323846 * (double(j) + 0.5) / double(H));
			double row = 0.0;
			for(int i = 0; i < W; ++i) {
				double m = 0.0;
				for(int dj = 0; dj < 2; ++dj)
					for(int di = 0; di < 2; ++di) {
						const float * c = texel(t, (i + di) % W, (j + dj) % H);
						m = std::max(m, lumd(c[0], c[1], c[2]));
					}
				f[size_t(j) * W + i] = m * sinT;
				row += m * sinT;
			}
			rows[size_t(j)] = row;
			total += row;
		}
		if(total > 0.0) {
			A.envW = W;
			A.envH = H;
			A.envPmf.resize(size_t(W) * H);
			A.envRow.resize(size_t(H) + 1);
			A.envCol.resize(size_t(H) * (W + 1));
			double cum = 0.0;
			A.envRow[0] = 0.0f;
			for(int j = 0; j < H; ++j) {
				cum += rows[size_t(j)];
				A.envRow[size_t(j) + 1] = j + 1 == H ? 1.0f : float(cum / total);
				float * cc = &A.envCol[size_t(j) * (W + 1)];
				cc[0] = 0.0f;
				double c2 = 0.0;
				for(int i = 0; i < W; ++i) {
					c2 += f[size_t(j) * W + i];
					cc[i + 1] = i + 1 == W ? 1.0f : (rows[size_t(j)] > 0.0 ? float(c2 / rows[size_t(j)]) : float(double(i + 1) / W));
					A.envPmf[size_t(j) * W + i] = float(f[size_t(j) * W + i] / total);
				}
			}
			envPower = total;
		}
	}
	A.emBase.assign(s->objects.size(), -1);
	std::vector<double> w;
	const std::vector<f3> & verts = s->rc->scene.vertices;
	std::vector<f3> own; // instanced scenes: an emissive object's world vertices
	size_t vbase = 0;
	for(size_t o = 0; o < s->objects.size(); ++o) {
		const rdu_object_desc & obj = s->objects[o];
		const rdu_scene::MeshRec & m = s->meshes[obj.mesh];
		const rdu_material_desc & mat = s->materials[obj.material];
		if(mat.type == RDU_MATERIAL_EMISSIVE) {
			const rdu_scene::Tex & t = s->textures[size_t(mat.textures[0])];
			double mean = 0.0;
			for(uint32_t y = 0; y < t.h; ++y)
				for(uint32_t x = 0; x < t.w; ++x) {
					const float * c = texel(t, int(x), int(y));
					mean += lumd(c[0], c[1], c[2]);
				}
			mean /= double(t.w) * double(t.h);
			A.emBase[o] = int32_t(w.size());
			const f3 * vp = verts.data() + vbase;
			if(s->instancing) {
				own = worldVertices(s->positions.data() + 3 * size_t(m.vtx), m.nv, obj.model);
				vp = own.data();
			}
			for(uint32_t k = 0; k + 2 < m.ni; k += 3) {
				const uint32_t * ix = s->indices.data() + m.idx + k;
				const f3 v0 = vp[ix[0]], v1 = vp[ix[1]], v2 = vp[ix[2]];
				const f3 e1 = v1 - v0, e2 = v2 - v0;
				const f3 n = cross(e1, e2);
				const double nx = n.x, ny = n.y, nz = n.z;
				const double area = 0.5 * std::sqrt(nx * nx + ny * ny + nz * nz);
				uint32_t obits = uint32_t(o), kbits = k;
				float of, kf;
				std::memcpy(&of, &obits, 4);
				std::memcpy(&kf, &kbits, 4);
				A.emTris.push_back(make_float4(v0.x, v0.y, v0.z, float(area)));
				A.emTris.push_back(make_float4(e1.x, e1.y, e1.z, 0.0f)); // pmf below
				A.emTris.push_back(make_float4(e2.x, e2.y, e2.z, of));
				A.emTris.push_back(make_float4(n.x, n.y, n.z, kf));
				w.push_back(area * mean);
			}
		}
		vbase += m.nv;
	}
	double triPower = 0.0;
	for(double x : w)
		triPower += x;
	if(triPower > 0.0) {
		A.emCdf.resize(w.size() + 1);
		double cum = 0.0;
		A.emCdf[0] = 0.0f;
		for(size_t k = 0; k < w.size(); ++k) {
			cum += w[k];
			A.emTris[4 * k + 1].w = float(w[k] / triPower);
			A.emCdf[k + 1] = k + 1 == w.size() ? 1.0f : float(cum / triPower);
		}
	} else {
		A.emTris.clear();
		A.emBase.assign(s->objects.size(), -1);
	}
	const bool env = envPower > 0.0, tri = !A.emTris.empty();
	A.pEnv = env ? (tri ? 0.5f : 1.0f) : 0.0f;
	A.pArea = (env || tri) ? (s->lights.empty() ? 1.0f : 0.5f) : 0.0f;
	return A;
}

template<typename T>
void appendBytes(std::vector<unsigned char> & k, const T & v) {
	const unsigned char * p = reinterpret_cast<const unsigned char *>(&v);
	k.insert(k.end(), p, p + sizeof(T));
}

// Everything a pass's launches depend on besides the committed scene (fixed after commit): two
// passes with equal keys enqueue identical kernels with identical arguments.
std::vector<unsigned char> passKey(const PtFrame & f, const PtOptions & po, const PtPaths & P, const PtScratch & scratch,
	const float * dImage, uint32_t samplesInPass, uint32_t components, const TraceTuning & tune) {
	std::vector<unsigned char> k;
	appendBytes(k, f);
	appendBytes(k, po);
	appendBytes(k, P);
	appendBytes(k, scratch.gstack);
	appendBytes(k, dImage);
	appendBytes(k, samplesInPass);
	appendBytes(k, components);
	appendBytes(k, tune);
	return k;
}

TuneKey sceneTuneKey(const rdu_scene * s, size_t passPaths) {
	const uint32_t cls = kTunePathPass | (s->rc->gpuBuilt ? kTuneGpuTree : 0u) | (s->instancing ? kTuneInstanced : 0u);
	const rdu::DevBVH & b = s->traceBvh();
	return makeTuneKey(s->rc->scene.tris.size(), b.wideDepth, s->rc->scene.meshCount, passPaths, cls);
}

// Captures what `enqueue` launches on `stream` into the scene's graph (replacing any previous one).
int capturePass(rdu_scene * s, hipStream_t stream, const std::function<int()> & enqueue) {
	s->dropGraph();
	int st = hipCheck(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
	if(st)
		return st;
	const int e = enqueue();
	hipGraph_t g = nullptr;
	const hipError_t ee = hipStreamEndCapture(stream, &g);
	if(e || ee != hipSuccess) {
		if(g)
			(void)hipGraphDestroy(g);
		return hipCheck(e ? hipError_t(e) : ee, "pass capture");
	}
	hipGraphExec_t x = nullptr;
	if((st = hipCheck(hipGraphInstantiate(&x, g, nullptr, nullptr, 0), "hipGraphInstantiate"))) {
		(void)hipGraphDestroy(g);
		return st;
	}
	s->graph = g;
	s->graphExec = x;
	return RDU_OK;
}

} // namespace

extern "C" {

int rdu_scene_create(rdu_scene ** out) {
	clearError();
	if(!out)
		return fail(RDU_ERR_INVALID, "null output handle");
	*out = new(std::nothrow) rdu_scene();
	return *out ? RDU_OK : fail(RDU_ERR_ALLOC, "host allocation failed");
}

void rdu_scene_destroy(rdu_scene * s) { delete s; }

int rdu_scene_set_instancing(rdu_scene * s, int enable) {
	clearError();
	if(!s)
		return fail(RDU_ERR_INVALID, "null scene");
	if(s->committed)
		return fail(RDU_ERR_STATE, "scene already committed");
	s->instancing = enable != 0;
	return RDU_OK;
}

int rdu_scene_set_gpu_build(rdu_scene * s, int enable) {
	clearError();
	if(!s)
		return fail(RDU_ERR_INVALID, "null scene");
	if(s->committed)
		return fail(RDU_ERR_STATE, "scene already committed");
	s->gpuBuild = enable != 0;
	return RDU_OK;
}

int rdu_scene_add_texture(rdu_scene * s, const rdu_texture_desc * d, int32_t * id) {
	clearError();
	if(!s || !d || !id || (!d->rgba && d->width * d->height))
		return fail(RDU_ERR_INVALID, "null argument");
	if(s->committed)
		return fail(RDU_ERR_STATE, "scene already committed");
	if(d->width == 0 || d->height == 0)
		return fail(RDU_ERR_INVALID, "empty texture");
	rdu_scene::Tex t{d->width, d->height, uint32_t(s->texels.size() / 4)};
	s->texels.insert(s->texels.end(), d->rgba, d->rgba + size_t(d->width) * d->height * 4);
	s->textures.push_back(t);
	*id = int32_t(s->textures.size() - 1);
	return RDU_OK;
}

int rdu_scene_add_mesh(rdu_scene * s, const rdu_mesh_desc * d, int32_t * id) {
	clearError();
	if(!s || !d || !id)
		return fail(RDU_ERR_INVALID, "null argument");
	if(s->committed)
		return fail(RDU_ERR_STATE, "scene already committed");
	if(!d->positions || !d->normals || !d->tangents || !d->bitangents || !d->indices)
		return fail(RDU_ERR_INVALID, "the path tracer needs positions, normals, tangent frames and indices");
	if(d->n_indices % 3 != 0 || d->n_indices == 0)
		return fail(RDU_ERR_INVALID, "index count must be a positive multiple of 3");
	for(size_t i = 0; i < d->n_indices; ++i)
		if(d->indices[i] >= d->n_vertices)
			return fail(RDU_ERR_INVALID, "index out of range of the mesh vertices");
	rdu_scene::MeshRec m{uint32_t(s->nrm.size() / 3), uint32_t(s->indices.size()), uint32_t(d->n_vertices),
		uint32_t(d->n_indices), d->texcoords != nullptr};
	const size_t nv = d->n_vertices;
	s->positions.insert(s->positions.end(), d->positions, d->positions + 3 * nv);
	s->nrm.insert(s->nrm.end(), d->normals, d->normals + 3 * nv);
	s->tan.insert(s->tan.end(), d->tangents, d->tangents + 3 * nv);
	s->bit.insert(s->bit.end(), d->bitangents, d->bitangents + 3 * nv);
	if(d->texcoords)
		s->uv.insert(s->uv.end(), d->texcoords, d->texcoords + 2 * nv);
	else
		s->uv.insert(s->uv.end(), 2 * nv, 0.0f);
	s->indices.insert(s->indices.end(), d->indices, d->indices + d->n_indices);
	s->meshes.push_back(m);
	*id = int32_t(s->meshes.size() - 1);
	return RDU_OK;
}

int rdu_scene_add_material(rdu_scene * s, const rdu_material_desc * d, int32_t * id) {
	clearError();
	if(!s || !d || !id)
		return fail(RDU_ERR_INVALID, "null argument");
	if(s->committed)
		return fail(RDU_ERR_STATE, "scene already committed");
	for(int k = 0; k < 3; ++k)
		if(!texOk(s, d->textures[k]))
			return fail(RDU_ERR_INVALID, "material texture id out of range");
	s->materials.push_back(*d);
	*id = int32_t(s->materials.size() - 1);
	return RDU_OK;
}

int rdu_scene_add_object(rdu_scene * s, const rdu_object_desc * d) {
	clearError();
	if(!s || !d)
		return fail(RDU_ERR_INVALID, "null argument");
	if(s->committed)
		return fail(RDU_ERR_STATE, "scene already committed");
	if(d->mesh >= s->meshes.size() || d->material >= s->materials.size())
		return fail(RDU_ERR_INVALID, "object mesh or material id out of range");
	s->objects.push_back(*d);
	return RDU_OK;
}

int rdu_scene_add_light(rdu_scene * s, const rdu_light_desc * d) {
	clearError();
	if(!s || !d)
		return fail(RDU_ERR_INVALID, "null argument");
	if(s->committed)
		return fail(RDU_ERR_STATE, "scene already committed");
	if(d->type < RDU_LIGHT_POINT || d->type > RDU_LIGHT_SPOT)
		return fail(RDU_ERR_INVALID, "unknown light type");
	s->lights.push_back(*d);
	return RDU_OK;
}

int rdu_scene_set_background(rdu_scene * s, const rdu_background_desc * d) {
	clearError();
	if(!s || !d)
		return fail(RDU_ERR_INVALID, "null argument");
	if(s->committed)
		return fail(RDU_ERR_STATE, "scene already committed");
	s->bg = *d;
	return RDU_OK;
}

namespace {
// A 4-wide child code moved into a batch arena whose job starts at the given node, triangle and
// instance offsets: nodes and instances shift by their offsets, a triangle leaf ~(first << 1 |
// count - 1) by twice the triangle offset; empty slots stay empty.
int32_t rebaseCode(int32_t c, uint32_t nodeOffset, uint32_t triOffset, uint32_t instOffset) {
	if(c == kWideEmpty)
		return c;
	if(c >= 0)
		return c + int32_t(nodeOffset);
	if(c < kInstLimit)
		return c + int32_t(instOffset);
	return c - 2 * int32_t(triOffset);
}

bool isIdentity(const float * M) {
	for(int k = 0; k < 16; ++k)
		if(M[k] != ((k % 5 == 0) ? 1.0f : 0.0f))
			return false;
	return true;
}

// Instanced commit. A mesh used by one object is added in world space (the object's model applied
// as Rendu applies it), so its triangles and tree are the flattened scene's, and its root is a leaf
// of the top level; a mesh shared by several objects is added once in object space. Each gets a
// tree of its own (Rendu's build per mesh); a 4-wide top level goes over the objects' world boxes.
// Per object the table holds the inverse model and its tree's root, with a flag for objects whose
// tree is already in world space (identity models): their rays enter untransformed. The world box of a
// transformed object bounds its mesh's transformed vertices (the transformed corners of the mesh
// box once a vertex budget is spent), widened by a relative 1e-5 so that float rounding never lets
// it cull a ray the object-space test keeps.
int commitInstanced(rdu_scene * s) {
	int st;
	std::vector<uint32_t> users(s->meshes.size(), 0), single(s->meshes.size(), 0);
	for(size_t o = 0; o < s->objects.size(); ++o) {
		++users[s->objects[o].mesh];
		single[s->objects[o].mesh] = uint32_t(o);
	}
	std::vector<uint32_t> rcMesh(s->meshes.size(), UINT32_MAX), objectOf;
	uint32_t added = 0;
	for(size_t m = 0; m < s->meshes.size(); ++m) {
		if(!users[m])
			continue;
		const rdu_scene::MeshRec & r = s->meshes[m];
		const float * model = users[m] == 1 ? s->objects[single[m]].model : nullptr;
		if((st = rdu_raycaster_add_mesh(s->rc, s->positions.data() + 3 * size_t(r.vtx), r.nv, s->indices.data() + r.idx, r.ni,
				model)))
			return st;
		rcMesh[m] = added++;
		objectOf.push_back(users[m] == 1 ? single[m] : UINT32_MAX);
	}
	if((st = buildInstanceBase(s->rc, objectOf)))
		return st;
	const DeviceLayout & L = s->rc->layout;
	std::vector<HBox> boxes(s->objects.size());
	std::vector<int32_t> leafCodes(s->objects.size());
	std::vector<float4> table(4 * s->objects.size());
	size_t vertexBudget = size_t(1) << 26; // transformed vertices for tight boxes, over all objects
	for(size_t o = 0; o < s->objects.size(); ++o) {
		const rdu_object_desc & ob = s->objects[o];
		const float * M = ob.model; // column-major: element (row r, column c) at M[4 c + r]
		const uint32_t rm = rcMesh[ob.mesh];
		const HBox & mb = s->rc->scene.nodes[rm].box;
		int32_t root = L.meshWideCodes[rm];
		int32_t oid = int32_t(o);
		const int32_t worldFlag = 1;
		float rootBits, oidBits, worldBits;
		std::memcpy(&rootBits, &root, 4);
		std::memcpy(&oidBits, &oid, 4);
		std::memcpy(&worldBits, &worldFlag, 4);
		// A single-use mesh's tree hangs in the top level as it is (its triangles carry the
		// object's id); other objects are entered through their record.
		leafCodes[o] = users[ob.mesh] == 1 ? root : kInstCode0 + oid;
		if(users[ob.mesh] == 1 || isIdentity(M)) {
			// The tree is in world space: the box is its root's, the ray enters as it is.
			boxes[o] = mb;
			table[4 * o + 0] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
			table[4 * o + 1] = make_float4(0.0f, 1.0f, 0.0f, 0.0f);
			table[4 * o + 2] = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
			table[4 * o + 3] = make_float4(rootBits, oidBits, worldBits, 0.0f);
			continue;
		}
		double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
		const auto extend = [&](const double p[3]) {
			for(int r = 0; r < 3; ++r) {
				const double w = double(M[r]) * p[0] + double(M[4 + r]) * p[1] + double(M[8 + r]) * p[2] + double(M[12 + r]);
				lo[r] = std::min(lo[r], w);
				hi[r] = std::max(hi[r], w);
			}
		};
		const rdu_scene::MeshRec & mr = s->meshes[ob.mesh];
		if(mr.nv <= vertexBudget) {
			vertexBudget -= mr.nv;
			// Every vertex of the mesh, each once: a box that holds the triangles' hull (an
			// unreferenced vertex can only widen it).
			const float * P = s->positions.data() + 3 * size_t(mr.vtx);
			for(size_t v = 0; v < mr.nv; ++v) {
				const double p[3] = {P[3 * v], P[3 * v + 1], P[3 * v + 2]};
				extend(p);
			}
		} else {
			for(int k = 0; k < 8; ++k) {
				const double p[3] = {(k & 1) ? mb.mx.x : mb.mn.x, (k & 2) ? mb.mx.y : mb.mn.y, (k & 4) ? mb.mx.z : mb.mn.z};
				extend(p);
			}
		}
		double ext = 0.0;
		for(int r = 0; r < 3; ++r)
			ext = std::max(ext, hi[r] - lo[r]);
		const double pad = 1e-5 * ext + 1e-7;
		for(int r = 0; r < 3; ++r) {
			lo[r] -= pad + 1e-6 * std::fabs(lo[r]);
			hi[r] += pad + 1e-6 * std::fabs(hi[r]);
		}
		boxes[o] = HBox{mk(float(lo[0]), float(lo[1]), float(lo[2])), mk(float(hi[0]), float(hi[1]), float(hi[2]))};
		// Inverse of the affine model: [A t]^-1 = [A^-1, -A^-1 t], in double.
		double A[3][3];
		for(int r = 0; r < 3; ++r)
			for(int c = 0; c < 3; ++c)
				A[r][c] = M[4 * c + r];
		const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
			A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
		if(!(std::fabs(det) > 0.0) || !std::isfinite(det))
			return fail(RDU_ERR_INVALID, "object with a singular model matrix in an instanced scene");
		double I[3][3];
		I[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / det;
		I[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
		I[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
		I[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / det;
		I[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
		I[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
		I[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / det;
		I[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
		I[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
		const double t[3] = {M[12], M[13], M[14]};
		for(int r = 0; r < 3; ++r) {
			const double tr = -(I[r][0] * t[0] + I[r][1] * t[1] + I[r][2] * t[2]);
			table[4 * o + size_t(r)] = make_float4(float(I[r][0]), float(I[r][1]), float(I[r][2]), float(tr));
		}
		table[4 * o + 3] = make_float4(rootBits, oidBits, 0.0f, 0.0f);
	}
	std::vector<float> wide = L.wide;
	uint32_t wideCount = L.wideCount, topDepth = 0;
	HBox top;
	const int32_t topCode = buildInstanceTop(boxes, leafCodes, wide, wideCount, topDepth, top);
	if((st = uploadVec(s->dInstWide, wide)) || (st = uploadVec(s->dInst, table)))
		return st;
	s->instWideHost = std::move(wide);
	s->instTableHost = std::move(table);
	DevBVH & b = s->ibvh;
	b = s->rc->dev;
	b.wide = static_cast<const float4 *>(s->dInstWide.ptr);
	b.inst = static_cast<const float4 *>(s->dInst.ptr);
	b.wideTopCode = topCode;
	const float tl[3] = {top.mn.x, top.mn.y, top.mn.z}, th[3] = {top.mx.x, top.mx.y, top.mx.z};
	std::memcpy(b.wideTopMin, tl, sizeof(tl));
	std::memcpy(b.wideTopMax, th, sizeof(th));
	std::memcpy(b.sceneMin, tl, sizeof(tl));
	std::memcpy(b.sceneMax, th, sizeof(th));
	// Stack entries: the top level's levels, a mesh tree's, and the return marker.
	b.wideDepth = topDepth + L.wideDepth + 1;
	return RDU_OK;
}
} // namespace

int rdu_scene_commit(rdu_scene * s) {
	clearError();
	if(!s)
		return fail(RDU_ERR_INVALID, "null scene");
	if(s->committed)
		return fail(RDU_ERR_STATE, "scene already committed");
	if(s->objects.empty())
		return fail(RDU_ERR_INVALID, "scene without objects");
	const rdu_background_desc & b = s->bg;
	if(b.mode < RDU_BACKGROUND_COLOR || b.mode > RDU_BACKGROUND_LATLONG)
		return fail(RDU_ERR_INVALID, "unknown background mode");
	if((b.mode == RDU_BACKGROUND_IMAGE || b.mode == RDU_BACKGROUND_ATMOSPHERE || b.mode == RDU_BACKGROUND_LATLONG) &&
		!texOk(s, b.texture))
		return fail(RDU_ERR_INVALID, "background texture id out of range");
	if(b.mode == RDU_BACKGROUND_SKYBOX)
		for(int k = 0; k < 6; ++k)
			if(!texOk(s, b.cube[k]))
				return fail(RDU_ERR_INVALID, "cubemap face id out of range");
	int st = rdu_raycaster_create(&s->rc);
	if(st)
		return st;
	// A failed commit leaves the scene uncommitted and without a raycaster, so it can be retried.
	struct DropOnFailure {
		rdu_scene * s;
		~DropOnFailure() {
			if(!s->committed && s->rc) {
				rdu_raycaster_destroy(s->rc);
				s->rc = nullptr;
			}
		}
	} dropOnFailure{s};
	if(s->instancing) {
		if(s->gpuBuild)
			return fail(RDU_ERR_UNSUPPORTED, "instancing builds its per-mesh trees on the host (GPU build unsupported)");
		if(s->objects.size() > size_t(kInstMax))
			return fail(RDU_ERR_UNSUPPORTED, "too many objects for an instanced scene");
		if((st = commitInstanced(s)))
			return st;
	} else {
	// PathTracer::PathTracer: every object's mesh, with its model, in object order (the raycaster's
	// stores sized once for all of them).
	size_t totalVerts = 0, totalTris = 0;
	for(const rdu_object_desc & o : s->objects) {
		totalVerts += s->meshes[o.mesh].nv;
		totalTris += s->meshes[o.mesh].ni / 3;
	}
	s->rc->scene.reserve(totalVerts, totalTris, s->objects.size());
	for(const rdu_object_desc & o : s->objects) {
		const rdu_scene::MeshRec & m = s->meshes[o.mesh];
		if((st = rdu_raycaster_add_mesh(s->rc, s->positions.data() + 3 * size_t(m.vtx), m.nv, s->indices.data() + m.idx, m.ni,
				o.model)))
			return st;
	}
	if(s->gpuBuild) {
		// Built on the scene's own stream, so that builds of different scenes overlap.
		if(!s->stream && (st = hipCheck(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate")))
			return st;
		if((st = rdu_raycaster_update_hierarchy_gpu(s->rc, s->stream)))
			return st;
	} else if((st = rdu_raycaster_update_hierarchy(s->rc))) {
		return st;
	}
	}
	std::vector<DevTexture> tex;
	for(const auto & t : s->textures)
		tex.push_back({t.offset, int32_t(t.w), int32_t(t.h)});
	std::vector<DevMesh> meshes;
	for(const auto & m : s->meshes)
		meshes.push_back({m.vtx, m.idx, m.hasUV ? 1 : 0});
	std::vector<DevObject> objs;
	for(const auto & o : s->objects) {
		DevObject d;
		d.mesh = o.mesh;
		d.mat = o.material;
		d.useUV = (o.use_texcoords && s->meshes[o.mesh].hasUV) ? 1 : 0; // Object.cpp:42
		inverseTransposeUpper(o.model, d.invtp);
		objs.push_back(d);
	}
	std::vector<DevMaterial> mats;
	int anyMasked = 0;
	for(const auto & m : s->materials) {
		mats.push_back({m.type, m.masked, {m.textures[0], m.textures[1], m.textures[2]}});
		anyMasked |= m.masked ? 1 : 0;
	}
	if((st = uploadVec(s->dTexels, s->texels)) || (st = uploadVec(s->dTextures, tex)) || (st = uploadVec(s->dNrm, s->nrm)) ||
		(st = uploadVec(s->dTan, s->tan)) || (st = uploadVec(s->dBit, s->bit)) || (st = uploadVec(s->dUv, s->uv)) ||
		(st = uploadVec(s->dIdx, s->indices)) || (st = uploadVec(s->dMeshes, meshes)) || (st = uploadVec(s->dObjects, objs)) ||
		(st = uploadVec(s->dMaterials, mats)) || (st = uploadVec(s->dLights, s->lights)))
		return st;
	DevScene & d = s->dev;
	d.texels = static_cast<const float4 *>(s->dTexels.ptr);
	d.textures = static_cast<const DevTexture *>(s->dTextures.ptr);
	d.nrm = static_cast<const float *>(s->dNrm.ptr);
	d.tan = static_cast<const float *>(s->dTan.ptr);
	d.bit = static_cast<const float *>(s->dBit.ptr);
	d.uv = static_cast<const float *>(s->dUv.ptr);
	d.idx = static_cast<const uint32_t *>(s->dIdx.ptr);
	d.meshes = static_cast<const DevMesh *>(s->dMeshes.ptr);
	d.objects = static_cast<const DevObject *>(s->dObjects.ptr);
	d.materials = static_cast<const DevMaterial *>(s->dMaterials.ptr);
	d.lights = static_cast<const rdu_light_desc *>(s->dLights.ptr);
	d.nLights = uint32_t(s->lights.size());
	d.anyMasked = anyMasked;
	d.bg = s->bg;
	const AreaTables A = buildAreaTables(s);
	if((st = uploadVec(s->dEmBase, A.emBase)) || (!A.envPmf.empty() && ((st = uploadVec(s->dEnvPmf, A.envPmf)) ||
			(st = uploadVec(s->dEnvRow, A.envRow)) || (st = uploadVec(s->dEnvCol, A.envCol)))) ||
		(!A.emTris.empty() && ((st = uploadVec(s->dEmTris, A.emTris)) || (st = uploadVec(s->dEmCdf, A.emCdf)))))
		return st;
	d.envPmf = static_cast<const float *>(s->dEnvPmf.ptr);
	d.envRowCdf = static_cast<const float *>(s->dEnvRow.ptr);
	d.envColCdf = static_cast<const float *>(s->dEnvCol.ptr);
	d.envW = A.envW;
	d.envH = A.envH;
	d.emTris = static_cast<const float4 *>(s->dEmTris.ptr);
	d.emCdf = static_cast<const float *>(s->dEmCdf.ptr);
	d.emBase = static_cast<const int32_t *>(s->dEmBase.ptr);
	d.nEm = uint32_t(A.emTris.size() / 4);
	d.pEnv = A.pEnv;
	d.pArea = A.pArea;
	s->committed = true;
	return RDU_OK;
}

const rdu_raycaster * rdu_scene_raycaster(const rdu_scene * s) { return (s && !s->instancing) ? s->rc : nullptr; }


// tuneOut: instead of rendering, time one pass per candidate of the traversal launch parameters
// (image unused) and keep the fastest for the scene.
// camFrame (when not null) gives the camera frame directly instead of `cam`'s pose inputs.
// Camera::pose + Camera::pixelShifts of a camera description (or the caller's own frame) and the
// sample grid of spp samples (PathTracer.cpp:57-69, 176-196).
static PtFrame makeFrame(const rdu_camera_desc * cam, const rdu_camera_frame * camFrame, uint32_t width,
	uint32_t height, size_t spp) {
	PtFrame f;
	f.width = width;
	f.height = height;
	if(camFrame) {
		f.eye = mk(camFrame->position[0], camFrame->position[1], camFrame->position[2]);
		f.corner = mk(camFrame->corner[0], camFrame->corner[1], camFrame->corner[2]);
		f.dx = mk(camFrame->dx[0], camFrame->dx[1], camFrame->dx[2]);
		f.dy = mk(camFrame->dy[0], camFrame->dy[1], camFrame->dy[2]);
	} else {
		// Camera::pose + Camera::pixelShifts (Camera.cpp:10-18, 42-49).
		const f3 eye = mk(cam->position[0], cam->position[1], cam->position[2]);
		const f3 center = mk(cam->center[0], cam->center[1], cam->center[2]);
		f3 up = normalize(mk(cam->up[0], cam->up[1], cam->up[2]));
		const f3 viewDir = normalize(center - eye);
		const f3 right = cross(viewDir, up);
		up = cross(right, viewDir);
		const float heightScale = std::tan(0.5f * cam->fov);
		const float widthScale = cam->ratio * heightScale;
		const float imageDist = length(center - eye);
		f.eye = eye;
		f.corner = center + imageDist * ((-widthScale) * right + heightScale * up);
		f.dx = ((2.0f * widthScale) * imageDist) * right;
		f.dy = ((-2.0f * heightScale) * imageDist) * up;
	}
	// getSampleGrid (PathTracer.cpp:57-69).
	const int k = int(std::floor(std::log2(double(spp))));
	int cx = int(std::pow(2, k / 2)), cy = cx;
	if(k % 2 == 1) {
		cx = int(std::pow(2, (k + 1) / 2));
		cy = int(std::pow(2, (k - 1) / 2));
	}
	f.cellX = cx;
	f.cellW = 1.0f / float(cx);
	f.cellH = 1.0f / float(cy);
	return f;
}

static int renderImpl(rdu_scene * s, const rdu_camera_desc * cam, size_t samples, size_t depth, float * image,
	uint32_t width, uint32_t height, uint32_t components, const rdu_render_opts * opts, rdu_rcache * cache,
	const rdu_cache_opts * copts, TuneResult * tuneOut = nullptr, const rdu_camera_frame * camFrame = nullptr) {
	clearError();
	const auto t0 = std::chrono::steady_clock::now();
	if(!s)
		return fail(RDU_ERR_INVALID, "[PathTracer] No scene available.");
	if(!s->committed)
		return fail(RDU_ERR_STATE, "scene not committed");
	if((!cam && !camFrame) || (!image && !tuneOut))
		return fail(RDU_ERR_INVALID, "null camera or image");
	if(components < 3)
		return fail(RDU_ERR_INVALID, "[PathTracer] Expected a RGB image.");
	if(width == 0 || height == 0 || samples == 0 || depth == 0)
		return fail(RDU_ERR_INVALID, "empty image, zero samples or zero depth");
	// PathTracer.cpp:169-173: round to a power of two.
	const size_t spp = size_t(std::pow(2.0, std::round(std::log2(float(samples)))));
	rdu_render_opts opt{};
	if(opts)
		opt = *opts;
	PtOptions po;
	po.depth = uint32_t(depth);
	po.estimator = opt.estimator;
	po.rrStart = opt.rr_start;
	const PtFrame f = makeFrame(cam, camFrame, width, height, spp);

	const size_t pixels = size_t(width) * height;
	// Passes in flight (plain renders; see the pass loop) and the default pass size: with one pass
	// at a time, 2^25 paths (about 6 GB of path state), so that the sparse late bounces of a pass run
	// over large queues (a 1080p 64-spp frame in 4 passes takes 0.64x the time of 32 passes); with
	// several in flight, 2^24 (the frame: 0.144 s against 0.147 s at 2^25 and 0.172 s with one pass
	// at a time). The image is the same bit for bit either way: samples accumulate in sample order
	// whatever the split. Cached renders train between passes, so they keep passes of 2^22 paths.
	constexpr uint32_t kRenderLanes = 3;
	static_assert(kRenderLanes == 1 + sizeof(s->laneStreams) / sizeof(s->laneStreams[0]), "a stream per further lane");
	const uint32_t wantLanes = (!cache && !s->useGraph && !s->dev.anyMasked)
		? std::min(opt.passes_in_flight ? opt.passes_in_flight : kRenderLanes, kRenderLanes)
		: 1u;
	const size_t maxPaths = opt.max_paths ? opt.max_paths
	                                      : (cache ? (size_t(1) << 22) : (wantLanes > 1 ? size_t(1) << 24 : size_t(1) << 25));
	if(pixels > (size_t(1) << 31))
		return fail(RDU_ERR_UNSUPPORTED, "image larger than 2^31 pixels");
	size_t perPass = maxPaths / pixels;
	perPass = perPass < 1 ? 1 : (perPass > spp ? spp : perPass);
	int st = setDevice(s->rc->device);
	if(st)
		return st;
	// A stream per scene: renders of different scenes from different host threads overlap on the GPU.
	if(!s->stream && (st = hipCheck(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate")))
		return st;
	hipStream_t stream = s->stream;
	PathsLease R(s->rc->device, pixels * perPass, stream, s->lastPaths);
	s->lastPaths = &*R;
	if((st = ensurePaths(*R, s, pixels * perPass, stream)))
		return st;
	const size_t imgBytes = pixels * components * sizeof(float);
	if((st = R->pImage.reserveOn(imgBytes, stream)))
		return st;
	if(!tuneOut && (st = hipCheck(hipMemcpyAsync(R->pImage.ptr, image, imgBytes, hipMemcpyHostToDevice, stream), "image upload")))
		return st;
	if((st = hipCheck(hipMemsetAsync(R->pTotals.ptr, 0, 16, stream), "hipMemsetAsync")))
		return st;
	PtPaths P = pathsOf(*R);
	PtScratch scratch;
	scratch.gstack = static_cast<int *>(R->pGstack.ptr);
	scratch.walkLeft = R->walkLeft;
	// Radiance-cache mode set-up.
	PtCache C;
	rdu_cache_opts co{};
	if(copts)
		co = *copts;
	const uint32_t trainSteps = co.train_steps ? co.train_steps : 4u;
	const uint32_t trainBatch = co.train_batch ? co.train_batch : 16384u;
	rdu_cache_stats cstats{};
	if(cache) {
		if(rcacheDevice(cache) != s->rc->device)
			return fail(RDU_ERR_INVALID, "cache and scene live on different devices");
		C.terminateBounce = co.from_primary ? 0 : int32_t(co.terminate_bounce ? co.terminate_bounce : 2u);
		C.maxBias = co.max_bias > 0.0f ? co.max_bias : 0.2f;
		C.minCount = co.min_count > 0.0f ? co.min_count : 8.0f;
		C.alwaysTrust = co.always_trust ? 1u : 0u;
		C.records = (co.frozen && co.always_trust) ? 0u : 1u;
		const size_t passPaths = pixels * perPass;
		const size_t tail = depth > size_t(C.terminateBounce) ? depth - size_t(C.terminateBounce) : 0;
		// One pixel in 16 trains; the 4 x 4 pattern may hold up to ceil(w/4) ceil(h/4) pixels.
		const size_t trainPaths = size_t((width + 3) / 4) * ((height + 3) / 4) * perPass;
		const size_t records = trainPaths * tail + 64;
		if(records > 0xffffffffull)
			return fail(RDU_ERR_UNSUPPORTED, "too many training records per pass");
		if((st = ensureCacheBuffers(s, passPaths, records)))
			return st;
		if(!rcacheTrain(cache, trainBatch))
			return fail(RDU_ERR_ALLOC, "cache training buffers");
		C.grid = *rcacheGrid(cache);
		C.params = rcacheParams(cache);
		C.counts = static_cast<uint32_t *>(s->cCounts.ptr);
		C.query = static_cast<float4 *>(s->cQuery.ptr);
		C.queryAtten = static_cast<float4 *>(s->cQAtten.ptr);
		C.queryPath = static_cast<uint32_t *>(s->cQPath.ptr);
		C.pred = static_cast<float4 *>(s->cPred.ptr);
		C.rec = static_cast<float4 *>(s->cRec.ptr);
		C.recPath = static_cast<uint32_t *>(s->cRecPath.ptr);
		C.recCapacity = uint32_t(records);
		C.trainQ = static_cast<float4 *>(s->cTrainQ.ptr);
		C.trainT = static_cast<float4 *>(s->cTrainT.ptr);
		C.holdQ = static_cast<float4 *>(s->cHoldQ.ptr);
		C.holdT = static_cast<float4 *>(s->cHoldT.ptr);
		C.cells = rcacheCells(cache);
	}
	const size_t passPaths = pixels * perPass;
	float * dImage = static_cast<float *>(R->pImage.ptr);
	if((st = hipCheck(hipError_t(ptSetSampleBase(P.counts, 0, opt.seed, stream)), "sample base")))
		return st;
	if(tuneOut) {
		const uint32_t n = uint32_t(perPass < spp ? perPass : spp);
		hipEvent_t e0, e1;
		if((st = hipCheck(hipEventCreate(&e0), "hipEventCreate")))
			return st;
		if((st = hipCheck(hipEventCreate(&e1), "hipEventCreate"))) {
			(void)hipEventDestroy(e0);
			return st;
		}
		// One candidate: a warm-up pass, then the mean of two timed ones.
		auto timeOne = [&](const TraceTuning & t, float & ms) -> int {
			ms = 0.0f;
			int e = RDU_OK;
			for(int rep = 0; rep < 3 && !e; ++rep) {
				if((e = hipCheck(hipError_t(ptSetSampleBase(P.counts, 0, opt.seed, stream)), "sample base")))
					break;
				if(rep > 0 && (e = hipCheck(hipEventRecord(e0, stream), "hipEventRecord")))
					break;
				if((e = hipCheck(hipError_t(ptRenderPass(s->dev, s->traceBvh(), f, P, n, po, dImage, components, scratch, nullptr,
									 t, stream)),
						"path tracing pass")))
					break;
				if(rep == 0)
					continue;
				float t1 = 0.0f;
				if(!(e = hipCheck(hipEventRecord(e1, stream), "hipEventRecord")) &&
					!(e = hipCheck(hipEventSynchronize(e1), "hipEventSynchronize")) &&
					!(e = hipCheck(hipEventElapsedTime(&t1, e0, e1), "hipEventElapsedTime")))
					ms += 0.5f * t1;
			}
			return e;
		};
		st = tuneSearch(timeOne, persistentOccupancy(false), *tuneOut);
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
		if(st)
			return st;
		s->tuned = tuneOut->best;
		s->tunedSet = true;
		tuneCacheStore(sceneTuneKey(s, passPaths), tuneOut->best);
		return hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
	}
	// Traversal launch parameters: the scene's own tuning, else the tuning cache's entry for the
	// scene's statistics and pass size, else the defaults.
	TraceTuning tune;
	if(s->tunedSet)
		tune = s->tuned;
	else
		(void)tuneCacheLookup(sceneTuneKey(s, passPaths), tune);
	uint32_t pass = 0, graphPasses = 0, graphCaptures = 0;
	// Plain renders of several passes keep up to wantLanes passes in flight, each lane with its
	// own stream and path state, so that one pass's traversal overlaps another's shading and the
	// sparse late bounces of a pass share the GPU with the dense early ones of the next. Each pass's
	// accumulation waits (by an event) for the previous pass's, so every pixel sums its samples in
	// order: the image is the same bit for bit as with one pass at a time.
	const size_t passCount = (spp + perPass - 1) / perPass;
	const size_t lanes = std::min(size_t(wantLanes), passCount);
	std::vector<std::unique_ptr<PathsLease>> laneLeases;
	EventSet events;
	unsigned long long laneTotals[2] = {0, 0};
	if(lanes > 1) {
		struct Lane {
			hipStream_t s;
			PtPaths P;
			PtScratch scratch;
		};
		std::vector<Lane> lane(lanes);
		lane[0] = Lane{stream, P, scratch};
		hipEvent_t setupDone;
		if((st = hipCheck(hipError_t(events.make(setupDone, hipEventDisableTiming)), "hipEventCreate")) ||
			(st = hipCheck(hipEventRecord(setupDone, stream), "hipEventRecord")))
			return st;
		for(size_t k = 1; k < lanes; ++k) {
			hipStream_t & ls = s->laneStreams[k - 1];
			if(!ls && (st = hipCheck(hipStreamCreateWithFlags(&ls, hipStreamNonBlocking), "hipStreamCreate")))
				return st;
			if((st = hipCheck(hipStreamWaitEvent(ls, setupDone, 0), "hipStreamWaitEvent")))
				return st;
			laneLeases.push_back(std::make_unique<PathsLease>(s->rc->device, passPaths, ls));
			RenderPaths & Rk = **laneLeases.back();
			if((st = ensurePaths(Rk, s, passPaths, ls)) ||
				(st = hipCheck(hipMemsetAsync(Rk.pTotals.ptr, 0, 16, ls), "hipMemsetAsync")))
				return st;
			lane[k].s = ls;
			lane[k].P = pathsOf(Rk);
			lane[k].scratch.gstack = static_cast<int *>(Rk.pGstack.ptr);
			lane[k].scratch.walkLeft = Rk.walkLeft;
		}
		hipEvent_t accumulated = nullptr;
		for(size_t base = 0; base < spp; base += perPass, ++pass) {
			const size_t n = (spp - base) < perPass ? (spp - base) : perPass;
			Lane & ln = lane[pass % lanes];
			hipEvent_t e;
			if((st = hipCheck(hipError_t(ptSetSampleBase(ln.P.counts, uint32_t(base), opt.seed, ln.s)), "sample base")) ||
				(st = hipCheck(hipError_t(ptRenderPass(s->dev, s->traceBvh(), f, ln.P, uint32_t(n), po, dImage, components,
								   ln.scratch, nullptr, tune, ln.s, accumulated)),
						"path tracing pass")) ||
				(st = hipCheck(hipError_t(events.make(e, hipEventDisableTiming)), "hipEventCreate")) ||
				(st = hipCheck(hipEventRecord(e, ln.s), "hipEventRecord")))
				return st;
			accumulated = e;
		}
		if((st = hipCheck(hipStreamWaitEvent(stream, accumulated, 0), "hipStreamWaitEvent")))
			return st;
		for(size_t k = 1; k < lanes; ++k) {
			unsigned long long t[2] = {0, 0};
			if((st = hipCheck(hipMemcpyAsync(t, lane[k].P.totals, 16, hipMemcpyDeviceToHost, lane[k].s), "stats download")) ||
				(st = hipCheck(hipStreamSynchronize(lane[k].s), "hipStreamSynchronize")))
				return st;
			laneTotals[0] += t[0];
			laneTotals[1] += t[1];
		}
	}
	for(size_t base = 0; lanes == 1 && base < spp; base += perPass, ++pass) {
		const size_t n = (spp - base) < perPass ? (spp - base) : perPass;
		if(cache) {
			// Shift the training pattern every pass (and with the seed) so every pixel trains.
			const uint32_t k = pass + opt.seed * 7u;
			C.trainX = k & 3u;
			C.trainY = (k >> 2) & 3u;
		}
		auto enqueue = [&]() -> int {
			return ptRenderPass(s->dev, s->traceBvh(), f, P, uint32_t(n), po, dImage, components, scratch, cache ? &C : nullptr,
				tune, stream);
		};
		// Scenes with cut-outs run eagerly: their shadow walk reads a count on the host.
		if(s->useGraph && !cache && n == perPass && !s->dev.anyMasked) {
			// A configuration seen on the previous full pass is captured once, then replayed.
			std::vector<unsigned char> key = passKey(f, po, P, scratch, dImage, uint32_t(n), components, tune);
			if(!(s->graphExec && key == s->graphKey) && key == s->lastPassKey) {
				if((st = capturePass(s, stream, enqueue)))
					return st;
				s->graphKey = key;
				++graphCaptures;
			}
			if(s->graphExec && key == s->graphKey) {
				if((st = hipCheck(hipGraphLaunch(s->graphExec, stream), "hipGraphLaunch")))
					return st;
				++graphPasses;
			} else {
				if((st = hipCheck(hipError_t(enqueue()), "path tracing pass")))
					return st;
				s->lastPassKey = std::move(key);
			}
		} else if((st = hipCheck(hipError_t(enqueue()), "path tracing pass"))) {
			return st;
		}
		if(cache) {
			uint32_t counts[kCacheSlots] = {0, 0, 0, 0};
			if((st = hipCheck(hipMemcpyAsync(counts, C.counts, sizeof(counts), hipMemcpyDeviceToHost, stream), "counts")) ||
				(st = hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize")))
				return st;
			cstats.cache_queries += counts[kCacheQueries];
			cstats.records += counts[kCacheRecords];
			cstats.train_pairs += counts[kCacheTrain];
			const uint32_t avail = counts[kCacheTrain];
			s->lastTrainPairs = avail;
			if(!co.frozen && avail > 0) {
				RcTrainBuffers * tb = rcacheTrain(cache, trainBatch);
				for(uint32_t k = 0; k < trainSteps; ++k) {
					const uint32_t off = avail > trainBatch ? uint32_t((uint64_t(k) * trainBatch) % (avail - trainBatch + 1)) : 0u;
					const uint32_t cnt = avail < trainBatch ? avail : trainBatch;
					if((st = hipCheck(hipError_t(rcTrainStep(C.grid, *rcacheParams(cache), *tb, *rcacheAdam(cache), C.trainQ + 2 * size_t(off),
										  C.trainT + off, cnt, stream)),
							"cache training")))
						return st;
					cstats.train_steps += 1;
				}
			}
		}
	}
	if(cache) {
		const size_t nc = size_t(kRcCells); // fine and coarse cells passing the trust test
		std::vector<float4> cells(nc);
		if((st = hipCheck(hipMemcpyAsync(cells.data(), C.cells, nc * 16, hipMemcpyDeviceToHost, stream), "cells")) ||
			(st = hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize")))
			return st;
		for(const float4 & c : cells) {
			const float var = std::max(c.w - c.y * c.y / std::max(c.z, 1e-6f), 0.0f);
			if(c.z >= C.minCount &&
				std::fabs(c.x - c.y) <= C.maxBias * std::max(c.x, c.y) + 2.0f * std::sqrt(var) + 1e-3f * c.z)
				cstats.trusted_cells += 1;
		}
	}
	s->lastCache = cstats;
	if((st = hipCheck(hipError_t(ptNormalizeImage(static_cast<float *>(R->pImage.ptr), uint32_t(pixels), components,
							  uint32_t(spp), stream)),
			"normalise")))
		return st;
	if((st = hipCheck(hipMemcpyAsync(image, R->pImage.ptr, imgBytes, hipMemcpyDeviceToHost, stream), "image download")))
		return st;
	unsigned long long totals[2] = {0, 0};
	if((st = hipCheck(hipMemcpyAsync(totals, R->pTotals.ptr, 16, hipMemcpyDeviceToHost, stream), "stats download")) ||
		(st = hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize")))
		return st;
	const auto t1 = std::chrono::steady_clock::now();
	s->last.seconds = std::chrono::duration<double>(t1 - t0).count();
	s->last.paths = uint64_t(pixels) * spp;
	s->last.extension_rays = totals[0] + laneTotals[0];
	s->last.shadow_rays = totals[1] + laneTotals[1];
	s->last.samples = uint32_t(spp);
	s->last.passes = pass;
	s->last.graph_passes = graphPasses;
	s->last.graph_captures = graphCaptures;
	return RDU_OK;
}

int rdu_pathtracer_render(rdu_scene * s, const rdu_camera_desc * cam, size_t samples, size_t depth, float * image,
	uint32_t width, uint32_t height, uint32_t components, const rdu_render_opts * opts) {
	return renderImpl(s, cam, samples, depth, image, width, height, components, opts, nullptr, nullptr);
}

int rdu_pathtracer_render_frame(rdu_scene * s, const rdu_camera_frame * frame, size_t samples, size_t depth,
	float * image, uint32_t width, uint32_t height, uint32_t components, const rdu_render_opts * opts) {
	if(!frame) {
		clearError();
		return fail(RDU_ERR_INVALID, "null camera frame");
	}
	return renderImpl(s, nullptr, samples, depth, image, width, height, components, opts, nullptr, nullptr, nullptr, frame);
}

int rdu_pathtracer_render_batch(const rdu_batch_job * jobs, size_t n, uint32_t max_paths, double * job_seconds,
	rdu_render_stats * stats) {
	clearError();
	const auto t0 = std::chrono::steady_clock::now();
	if(!jobs || n == 0)
		return fail(RDU_ERR_INVALID, "null or empty job list");
	if(n > 65536)
		return fail(RDU_ERR_UNSUPPORTED, "more than 65536 jobs in one batch");
	int device = -1;
	for(size_t j = 0; j < n; ++j) {
		const rdu_batch_job & b = jobs[j];
		if(!b.scene || !b.scene->committed)
			return fail(RDU_ERR_STATE, "batch job with no committed scene");
		if((!b.camera && !b.frame) || !b.image)
			return fail(RDU_ERR_INVALID, "batch job with a null camera or image");
		if(b.components < 3)
			return fail(RDU_ERR_INVALID, "[PathTracer] Expected a RGB image.");
		if(b.width == 0 || b.height == 0 || b.samples == 0 || b.depth == 0 || b.depth > 1024)
			return fail(RDU_ERR_INVALID, "batch job with an empty image, zero samples or a depth outside 1..1024");
		if(size_t(b.width) * b.height > (size_t(1) << 26))
			return fail(RDU_ERR_UNSUPPORTED, "batch job image larger than 2^26 pixels");
		if(device >= 0 && b.scene->rc->device != device)
			return fail(RDU_ERR_INVALID, "batch jobs on different devices");
		device = b.scene->rc->device;
	}
	int st = setDevice(device);
	if(st)
		return st;
	// Paths per pass: with passes in flight on three streams, 2^24 renders the 64-job batch as fast
	// as 2^25 in half the memory (about 3.2 GB of path state per pass).
	const size_t budget = max_paths ? max_paths : (size_t(1) << 24);

	// Per job: sample count (power of two), frame, options; the tree's node and triangle counts.
	struct JobInfo {
		size_t spp, pixels, next = 0, imgOffset = 0;
		uint32_t nodes = 0, tris = 0, nodeOffset = 0, triOffset = 0, insts = 0, instOffset = 0;
		int32_t code = 0;
		float top[6];
	};
	std::vector<JobInfo> info(n);
	std::vector<PtJob> table(n);
	size_t nodesTotal = 0, trisTotal = 0, instsTotal = 0, imgFloats = 0;
	uint32_t maxDepth = 0, wideDepth = 0;
	bool shadows = false, masked = false;
	for(size_t j = 0; j < n; ++j) {
		const rdu_batch_job & b = jobs[j];
		const rdu_raycaster * rc = b.scene->rc;
		JobInfo & I = info[j];
		I.spp = size_t(std::pow(2.0, std::round(std::log2(float(b.samples)))));
		I.pixels = size_t(b.width) * b.height;
		const bool inst = b.scene->instancing;
		const DevBVH & tb = b.scene->traceBvh();
		I.nodes = inst ? uint32_t(b.scene->instWideHost.size() / 32)
					   : rc->gpuBuilt ? rc->gpu.wideCount : uint32_t(rc->layout.wide.size() / 32);
		I.tris = rc->gpuBuilt ? rc->gpu.triangles : uint32_t(rc->layout.tris.size() / 12);
		I.insts = inst ? uint32_t(b.scene->instTableHost.size() / 4) : 0u;
		I.nodeOffset = uint32_t(nodesTotal);
		I.triOffset = uint32_t(trisTotal);
		I.instOffset = uint32_t(instsTotal);
		I.imgOffset = imgFloats;
		nodesTotal += I.nodes;
		trisTotal += I.tris;
		instsTotal += I.insts;
		imgFloats += I.pixels * b.components;
		I.code = rebaseCode(tb.wideTopCode, I.nodeOffset, I.triOffset, I.instOffset);
		for(int a = 0; a < 3; ++a) {
			I.top[a] = tb.wideTopMin[a];
			I.top[3 + a] = tb.wideTopMax[a];
		}
		wideDepth = std::max(wideDepth, tb.wideDepth);
		const rdu_render_opts o = b.opts;
		PtJob & J = table[j];
		J.S = b.scene->dev;
		J.f = makeFrame(b.camera, b.frame, b.width, b.height, I.spp);
		J.opt.depth = uint32_t(b.depth);
		J.opt.estimator = o.estimator;
		J.opt.rrStart = o.rr_start;
		J.components = b.components;
		maxDepth = std::max(maxDepth, uint32_t(b.depth));
		shadows = shadows || J.S.nLights > 0 || (o.estimator == RDU_ESTIMATOR_EXTENDED && J.S.pArea > 0.0f);
		masked = masked || J.S.anyMasked != 0;
	}
	if(nodesTotal >= size_t(0x7fffffff) || trisTotal >= (size_t(1) << (instsTotal ? 29 : 30)) || instsTotal > size_t(kInstMax))
		return fail(RDU_ERR_UNSUPPORTED, "batch trees too large for one arena");
	// Instanced jobs: their nodes (per-mesh trees and top level) and instance records rebased here
	// on the host, every code checked against the arena before anything is uploaded.
	std::vector<std::vector<float>> instNodes(n);
	std::vector<float4> instTable(4 * instsTotal);
	const auto codeInArena = [&](int32_t c) {
		if(c >= 0)
			return c == kWideEmpty || size_t(c) < nodesTotal;
		if(c < kInstLimit)
			return size_t(c - kInstCode0) < instsTotal;
		const uint32_t code = uint32_t(~c);
		return size_t(code >> 1) + (code & 1u) < trisTotal;
	};
	for(size_t j = 0; j < n; ++j) {
		const JobInfo & I = info[j];
		if(!jobs[j].scene->instancing)
			continue;
		if(!codeInArena(I.code))
			return fail(RDU_ERR_STATE, "batch arena: an instanced job's root code out of range");
		std::vector<float> & W = instNodes[j];
		W = jobs[j].scene->instWideHost;
		for(size_t k = 0; k < size_t(I.nodes); ++k)
			for(int slot = 0; slot < 4; ++slot) {
				float & f = W[32 * k + 24 + size_t(slot)];
				const int32_t c = rebaseCode(__builtin_bit_cast(int32_t, f), I.nodeOffset, I.triOffset, I.instOffset);
				if(!codeInArena(c))
					return fail(RDU_ERR_STATE, "batch arena: an instanced job's node code out of range");
				f = __builtin_bit_cast(float, c);
			}
		const std::vector<float4> & T = jobs[j].scene->instTableHost;
		for(size_t i = 0; i < size_t(I.insts); ++i) {
			float4 * r = &instTable[4 * (size_t(I.instOffset) + i)];
			r[0] = T[4 * i];
			r[1] = T[4 * i + 1];
			r[2] = T[4 * i + 2];
			r[3] = T[4 * i + 3];
			const int32_t root = rebaseCode(__builtin_bit_cast(int32_t, r[3].x), I.nodeOffset, I.triOffset, I.instOffset);
			if(root < kInstLimit || !codeInArena(root))
				return fail(RDU_ERR_STATE, "batch arena: an instance's tree root out of range");
			r[3].x = __builtin_bit_cast(float, root);
		}
	}

	hipStream_t stream = nullptr;
	if((st = hipCheck(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate")))
		return st;
	// Device memory of the batch; freed in stream order when the call returns.
	struct Local {
		hipStream_t s;
		std::vector<hipStream_t> more; // the other streams passes run on
		DevBuffer arenaNodes, arenaTris, arenaInst, roots, jobsDev, chunks, images;
		std::vector<hipEvent_t> events;
		~Local() {
			for(hipStream_t m : more)
				(void)hipStreamSynchronize(m);
			(void)hipStreamSynchronize(s);
			for(DevBuffer * b : {&arenaNodes, &arenaTris, &arenaInst, &roots, &jobsDev, &chunks, &images})
				b->releaseOn(s);
			for(hipEvent_t e : events)
				(void)hipEventDestroy(e);
			(void)hipStreamSynchronize(s);
			for(hipStream_t m : more)
				(void)hipStreamDestroy(m);
			(void)hipStreamDestroy(s);
		}
	} L;
	L.s = stream;
	if((st = L.arenaNodes.reserveOn(std::max<size_t>(nodesTotal, 1) * 128, stream)) ||
		(st = L.arenaTris.reserveOn(std::max<size_t>(trisTotal, 1) * 48, stream)) || (st = L.roots.reserveOn(n * 32, stream)) ||
		(st = L.jobsDev.reserveOn(n * sizeof(PtJob), stream)) || (st = L.images.reserveOn(imgFloats * 4, stream)) ||
		(instsTotal && (st = L.arenaInst.reserveOn(instsTotal * 64, stream))))
		return st;
	if(instsTotal && (st = hipCheck(hipMemcpyAsync(L.arenaInst.ptr, instTable.data(), instsTotal * 64, hipMemcpyHostToDevice, stream),
						  "arena instances")))
		return st;
	float4 * arenaNodes = static_cast<float4 *>(L.arenaNodes.ptr);
	float4 * arenaTris = static_cast<float4 *>(L.arenaTris.ptr);
	float * images = static_cast<float *>(L.images.ptr);
	std::vector<float4> roots(2 * n);
	for(size_t j = 0; j < n; ++j) {
		const JobInfo & I = info[j];
		const rdu_raycaster * rc = jobs[j].scene->rc;
		if((jobs[j].scene->instancing
				   ? (st = hipCheck(hipMemcpyAsync(arenaNodes + 8 * size_t(I.nodeOffset), instNodes[j].data(), size_t(I.nodes) * 128,
										hipMemcpyHostToDevice, stream),
						  "arena nodes"))
				   : (st = hipCheck(hipError_t(ptRebaseWide(arenaNodes + 8 * size_t(I.nodeOffset), rc->dev.wide, I.nodes,
										int32_t(I.nodeOffset), int32_t(I.triOffset), stream)),
						  "arena nodes"))) ||
			(I.tris && (st = hipCheck(hipMemcpyAsync(arenaTris + 3 * size_t(I.triOffset), rc->dev.tris, size_t(I.tris) * 48,
								 hipMemcpyDeviceToDevice, stream),
							"arena triangles"))))
			return st;
		roots[2 * j] = make_float4(I.top[0], I.top[1], I.top[2], __builtin_bit_cast(float, I.code));
		roots[2 * j + 1] = make_float4(I.top[3], I.top[4], I.top[5], 0.0f);
		table[j].image = images + I.imgOffset;
		const rdu_batch_job & b = jobs[j];
		if((st = hipCheck(hipMemcpyAsync(table[j].image, b.image, I.pixels * b.components * 4, hipMemcpyHostToDevice, stream),
				"image upload")))
			return st;
	}
	if((st = hipCheck(hipMemcpyAsync(L.roots.ptr, roots.data(), roots.size() * 16, hipMemcpyHostToDevice, stream), "roots")) ||
		(st = hipCheck(hipMemcpyAsync(L.jobsDev.ptr, table.data(), n * sizeof(PtJob), hipMemcpyHostToDevice, stream), "jobs")))
		return st;
	DevBVH arena;
	arena.wide = arenaNodes;
	arena.tris = arenaTris;
	arena.inst = instsTotal ? static_cast<const float4 *>(L.arenaInst.ptr) : nullptr;
	arena.wideDepth = wideDepth;

	// Schedule, made up front: longest jobs first (pixels x samples x depth), each pass taking as
	// many samples of each unfinished job as fit in the path budget; a job's samples run in order
	// over its passes.
	std::vector<size_t> order(n);
	for(size_t j = 0; j < n; ++j)
		order[j] = j;
	std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
		return info[a].pixels * info[a].spp * jobs[a].depth > info[b].pixels * info[b].spp * jobs[b].depth;
	});
	struct Pass {
		size_t chunkBegin = 0, chunkCount = 0, used = 0, slots = 0;
		std::vector<size_t> finished; // jobs whose last samples this pass renders
	};
	std::vector<Pass> schedule;
	std::vector<PtChunk> allChunks;
	for(size_t remaining = n; remaining > 0;) {
		Pass ps;
		ps.chunkBegin = allChunks.size();
		for(size_t j : order) {
			JobInfo & I = info[j];
			if(I.next >= I.spp)
				continue;
			size_t take = (budget - std::min(budget, ps.used)) / I.pixels;
			if(take == 0 && ps.used == 0)
				take = 1; // an image larger than the budget: one sample per pass
			take = std::min(take, I.spp - I.next);
			if(take == 0)
				continue;
			PtChunk c;
			c.job = uint32_t(j);
			c.pathBegin = uint32_t(ps.used);
			c.samples = uint32_t(take);
			c.sampleBase = uint32_t(I.next);
			c.seed = jobs[j].opts.seed;
			c.slotBegin = uint32_t(ps.slots);
			allChunks.push_back(c);
			ps.used += take * I.pixels;
			ps.slots += I.pixels;
			I.next += take;
			if(I.next >= I.spp)
				ps.finished.push_back(j);
		}
		ps.chunkCount = allChunks.size() - ps.chunkBegin;
		remaining -= ps.finished.size();
		schedule.push_back(std::move(ps));
	}
	size_t passCap = 1;
	for(const Pass & ps : schedule)
		passCap = std::max(passCap, ps.used);
	if(passCap > 0xffffffffull)
		return fail(RDU_ERR_UNSUPPORTED, "batch pass above 2^32 paths");
	if((st = L.chunks.reserveOn(allChunks.size() * sizeof(PtChunk), stream)) ||
		(st = hipCheck(hipMemcpyAsync(L.chunks.ptr, allChunks.data(), allChunks.size() * sizeof(PtChunk),
						   hipMemcpyHostToDevice, stream),
				"chunks")))
		return st;

	// Passes in flight: up to kBatchLanes, each lane with its own stream and path state, so that one
	// pass's traversal overlaps another's shading. Passes accumulate in schedule order (each waits
	// for the previous pass's accumulation), so every pixel sums its samples in order.
	constexpr size_t kBatchLanes = 3;
	const size_t lanes = std::min(kBatchLanes, schedule.size());
	struct Lane {
		hipStream_t s = nullptr;
		std::unique_ptr<PathsLease> lease;
		PtPaths P;
		PtScratch scratch;
		PtBatch B;
	};
	std::vector<Lane> lane(lanes);
	hipEvent_t setupDone;
	if((st = hipCheck(hipEventCreateWithFlags(&setupDone, hipEventDisableTiming), "hipEventCreate")))
		return st;
	L.events.push_back(setupDone);
	if((st = hipCheck(hipEventRecord(setupDone, stream), "hipEventRecord")))
		return st;
	for(size_t k = 0; k < lanes; ++k) {
		Lane & ln = lane[k];
		if(k == 0)
			ln.s = stream;
		else {
			if((st = hipCheck(hipStreamCreateWithFlags(&ln.s, hipStreamNonBlocking), "hipStreamCreate")))
				return st;
			L.more.push_back(ln.s);
			if((st = hipCheck(hipStreamWaitEvent(ln.s, setupDone, 0), "hipStreamWaitEvent")))
				return st;
		}
		ln.lease = std::make_unique<PathsLease>(device, passCap, ln.s);
		RenderPaths & R = **ln.lease;
		if((st = ensurePathsFor(R, arena, masked, passCap, true, ln.s)) ||
			(st = hipCheck(hipMemsetAsync(R.pTotals.ptr, 0, 16, ln.s), "hipMemsetAsync")))
			return st;
		ln.P = pathsOf(R);
		ln.scratch.gstack = static_cast<int *>(R.pGstack.ptr);
		ln.scratch.walkLeft = R.walkLeft;
		ln.B.jobs = static_cast<const PtJob *>(L.jobsDev.ptr);
		ln.B.pathChunk = static_cast<uint32_t *>(R.pChunk.ptr);
		ln.B.pathJob = static_cast<uint32_t *>(R.pJob.ptr);
		ln.B.overflow = static_cast<uint32_t *>(R.pOverflow.ptr);
		ln.B.roots = static_cast<const float4 *>(L.roots.ptr);
	}
	hipEvent_t start;
	if((st = hipCheck(hipEventCreate(&start), "hipEventCreate")))
		return st;
	L.events.push_back(start);
	if((st = hipCheck(hipEventRecord(start, stream), "hipEventRecord")))
		return st;
	std::vector<hipEvent_t> doneAt(n, nullptr);
	// A finished job's image: normalised after its last accumulation, copied back, time-stamped.
	// Enqueued one pass late, so that the copy (which waits on the host for its stream) finds the
	// next pass already queued on another stream.
	const auto deliver = [&](const Pass & ps, hipStream_t s) -> int {
		for(size_t j : ps.finished) {
			const JobInfo & I = info[j];
			const rdu_batch_job & b = jobs[j];
			hipEvent_t e;
			int r = hipCheck(hipEventCreate(&e), "hipEventCreate");
			if(r)
				return r;
			L.events.push_back(e);
			if((r = hipCheck(hipError_t(ptNormalizeImage(table[j].image, uint32_t(I.pixels), b.components, uint32_t(I.spp), s)),
					"normalise")) ||
				(r = hipCheck(hipMemcpyAsync(b.image, table[j].image, I.pixels * b.components * 4, hipMemcpyDeviceToHost, s),
						"image download")) ||
				(r = hipCheck(hipEventRecord(e, s), "hipEventRecord")))
				return r;
			doneAt[j] = e;
		}
		return RDU_OK;
	};
	const TraceTuning tune;
	hipEvent_t prevAccumulated = nullptr;
	for(size_t p = 0; p < schedule.size(); ++p) {
		const Pass & ps = schedule[p];
		Lane & ln = lane[p % lanes];
		ln.B.chunks = static_cast<const PtChunk *>(L.chunks.ptr) + ps.chunkBegin;
		ln.B.nChunks = uint32_t(ps.chunkCount);
		ln.B.nSlots = uint32_t(ps.slots);
		if((st = hipCheck(hipError_t(ptRenderBatchPass(ln.B, arena, ln.P, uint32_t(ps.used), maxDepth, shadows, masked, 16,
								 ln.scratch, tune, ln.s, prevAccumulated)),
				"batch pass")))
			return st;
		hipEvent_t acc;
		if((st = hipCheck(hipEventCreateWithFlags(&acc, hipEventDisableTiming), "hipEventCreate")))
			return st;
		L.events.push_back(acc);
		if((st = hipCheck(hipEventRecord(acc, ln.s), "hipEventRecord")))
			return st;
		prevAccumulated = acc;
		if(p > 0 && (st = deliver(schedule[p - 1], lane[(p - 1) % lanes].s)))
			return st;
	}
	if((st = deliver(schedule.back(), lane[(schedule.size() - 1) % lanes].s)))
		return st;
	const uint32_t passes = uint32_t(schedule.size());
	unsigned long long totals[2] = {0, 0};
	for(Lane & ln : lane) {
		unsigned long long t[2] = {0, 0};
		if((st = hipCheck(hipMemcpyAsync(t, (*ln.lease)->pTotals.ptr, 16, hipMemcpyDeviceToHost, ln.s), "stats download")) ||
			(st = hipCheck(hipStreamSynchronize(ln.s), "hipStreamSynchronize")))
			return st;
		totals[0] += t[0];
		totals[1] += t[1];
	}
	// Completion time of each job from the call's start: host time to the first enqueued work plus
	// the device time from there to the job's last copy.
	const double hostLead = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	float lastMs = 0.0f;
	for(size_t j = 0; j < n; ++j) {
		float ms = 0.0f;
		if((st = hipCheck(hipEventElapsedTime(&ms, start, doneAt[j]), "hipEventElapsedTime")))
			return st;
		lastMs = std::max(lastMs, ms);
		if(job_seconds)
			job_seconds[j] = double(ms) * 1e-3;
	}
	if(job_seconds) {
		// Shift to the call's clock: the last job finished when the call's wait returned.
		const double shift = hostLead - double(lastMs) * 1e-3;
		for(size_t j = 0; j < n; ++j)
			job_seconds[j] += shift > 0.0 ? shift : 0.0;
	}
	if(stats) {
		size_t paths = 0;
		for(size_t j = 0; j < n; ++j)
			paths += info[j].pixels * info[j].spp;
		*stats = rdu_render_stats{};
		stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
		stats->paths = paths;
		stats->extension_rays = totals[0];
		stats->shadow_rays = totals[1];
		stats->passes = passes;
	}
	return RDU_OK;
}

int rdu_scene_set_graph(rdu_scene * s, int enable) {
	clearError();
	if(!s)
		return fail(RDU_ERR_INVALID, "null scene");
	s->useGraph = enable != 0;
	if(!s->useGraph)
		s->dropGraph();
	return RDU_OK;
}

int rdu_scene_autotune(rdu_scene * s, const rdu_camera_desc * cam, size_t samples, size_t depth, uint32_t width,
	uint32_t height, const rdu_render_opts * opts, rdu_tune_report * report) {
	TuneResult res;
	const int st = renderImpl(s, cam, samples, depth, nullptr, width, height, 3, opts, nullptr, nullptr, &res);
	if(st)
		return st;
	if(report) {
		*report = rdu_tune_report{};
		report->refill_min = res.best.refillMin;
		report->chunk = res.best.chunk;
		report->blocks_per_cu = res.best.blocksPerCu;
		report->best_ms = res.bestMs;
		report->default_ms = res.defaultMs;
		report->candidates = res.timed;
		report->source = 1;
	}
	return RDU_OK;
}

int rdu_pathtracer_render_cached(rdu_scene * s, rdu_rcache * cache, const rdu_camera_desc * cam, size_t samples,
	size_t depth, float * image, uint32_t width, uint32_t height, uint32_t components, const rdu_render_opts * opts,
	const rdu_cache_opts * copts) {
	if(!cache) {
		clearError();
		return fail(RDU_ERR_INVALID, "null cache");
	}
	return renderImpl(s, cam, samples, depth, image, width, height, components, opts, cache, copts);
}

int rdu_scene_last_render_stats(const rdu_scene * s, rdu_render_stats * out) {
	clearError();
	if(!s || !out)
		return fail(RDU_ERR_INVALID, "null argument");
	*out = s->last;
	return RDU_OK;
}

int rdu_scene_last_cache_stats(const rdu_scene * s, rdu_cache_stats * out) {
	clearError();
	if(!s || !out)
		return fail(RDU_ERR_INVALID, "null argument");
	*out = s->lastCache;
	return RDU_OK;
}

int rdu_scene_last_cache_pairs(const rdu_scene * s, float * queries, float * targets, size_t capacity, size_t * count) {
	clearError();
	if(!s || !count)
		return fail(RDU_ERR_INVALID, "null argument");
	*count = s->lastTrainPairs;
	const size_t n = capacity < s->lastTrainPairs ? capacity : s->lastTrainPairs;
	if(n == 0)
		return RDU_OK;
	if(!queries || !targets)
		return fail(RDU_ERR_INVALID, "null output arrays");
	int st = setDevice(s->rc->device);
	if(st)
		return st;
	if((st = hipCheck(hipMemcpy(queries, s->cTrainQ.ptr, n * 32, hipMemcpyDeviceToHost), "copy")) ||
		(st = hipCheck(hipMemcpy(targets, s->cTrainT.ptr, n * 16, hipMemcpyDeviceToHost), "copy")))
		return st;
	return RDU_OK;
}

} // extern "C"
