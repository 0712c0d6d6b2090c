// rdu_rendu.hpp — header-only C++ drop-in classes with Rendu's signatures over the C ABI of rdu.h.
//
// rdu::Raycaster mirrors src/engine/raycaster/Raycaster.hpp:10-157 and rdu::PathTracer mirrors
// src/apps/pathtracer/PathTracer.hpp:10-76. They are templates over the caller's vector and mesh
// types so that Rendu's glm::vec3 / Mesh plug in unchanged (a Mesh needs `positions`, `indices`,
// `normals`, `tangents`, `bitangents`, `texcoords` with glm-like element types). Errors are
// reported through rdu_last_error() and, as in Rendu, logged rather than thrown: the optional
// `log` callback receives the message (Rendu's Log::Error sink).
#pragma once
#include "rdu.h"

#include <cfloat>
#include <functional>
#include <string>
#include <vector>

namespace rdu {

using LogSink = std::function<void(const std::string &)>;

inline void report(int status, const LogSink & log) {
	if(status != RDU_OK && log)
		log(std::string("[rendu_amd] ") + rdu_last_error());
}

class Raycaster {
public:
	// Raycaster::Hit (Raycaster.hpp:16-42).
	struct Hit {
		bool hit = false;
		float dist = FLT_MAX;
		float u = 0.0f, v = 0.0f, w = 0.0f;
		unsigned long localId = 0, meshId = 0;
	};

	explicit Raycaster(LogSink log = {}) : _log(std::move(log)) { report(rdu_raycaster_create(&_h), _log); }
	~Raycaster() { rdu_raycaster_destroy(_h); }
	Raycaster(const Raycaster &) = delete;
	Raycaster & operator=(const Raycaster &) = delete;
	Raycaster(Raycaster &&) = delete;
	Raycaster & operator=(Raycaster &&) = delete;

	// Raycaster::addMesh(const Mesh &, const glm::mat4 &). Mat4 is column-major like glm::mat4.
	template<typename Mesh, typename Mat4>
	void addMesh(const Mesh & mesh, const Mat4 & model) {
		std::vector<float> xyz;
		xyz.reserve(mesh.positions.size() * 3);
		for(const auto & p : mesh.positions) {
			xyz.push_back(p[0]);
			xyz.push_back(p[1]);
			xyz.push_back(p[2]);
		}
		float m[16];
		for(int c = 0; c < 4; ++c)
			for(int r = 0; r < 4; ++r)
				m[4 * c + r] = model[c][r];
		report(rdu_raycaster_add_mesh(_h, xyz.data(), mesh.positions.size(), mesh.indices.data(), mesh.indices.size(), m),
			_log);
	}

	void updateHierarchy() { report(rdu_raycaster_update_hierarchy(_h), _log); }

	template<typename Vec3>
	Hit intersects(const Vec3 & origin, const Vec3 & direction, float mini = 0.0001f, float maxi = 1e8f) const {
		const float o[3] = {origin[0], origin[1], origin[2]}, d[3] = {direction[0], direction[1], direction[2]};
		rdu_hit h{};
		report(rdu_raycaster_intersect1(_h, o, d, mini, maxi, &h), _log);
		Hit out;
		if(h.hit) {
			out.hit = true;
			out.dist = h.t;
			out.u = h.u;
			out.v = h.v;
			out.w = 1.0f - h.u - h.v;
			out.localId = h.local_id;
			out.meshId = h.mesh_id;
		}
		return out;
	}

	template<typename Vec3>
	bool intersectsAny(const Vec3 & origin, const Vec3 & direction, float mini = 0.0001f, float maxi = 1e8f) const {
		const float o[3] = {origin[0], origin[1], origin[2]}, d[3] = {direction[0], direction[1], direction[2]};
		int occ = 0;
		report(rdu_raycaster_intersect_any1(_h, o, d, mini, maxi, &occ), _log);
		return occ != 0;
	}

	template<typename Vec3>
	bool visible(const Vec3 & p0, const Vec3 & p1) const {
		const float a[3] = {p0[0], p0[1], p0[2]}, b[3] = {p1[0], p1[1], p1[2]};
		int vis = 0;
		report(rdu_raycaster_visible(_h, a, b, &vis), _log);
		return vis != 0;
	}

	// Raycaster::interpolateAttribute (Raycaster.hpp:90-97).
	template<typename T, typename Mesh>
	static T interpolateAttribute(const Hit & hit, const Mesh & geometry, const std::vector<T> & attribute) {
		const unsigned long i0 = geometry.indices[hit.localId];
		const unsigned long i1 = geometry.indices[hit.localId + 1];
		const unsigned long i2 = geometry.indices[hit.localId + 2];
		return hit.w * attribute[i0] + hit.u * attribute[i1] + hit.v * attribute[i2];
	}

	// Batch entry for throughput (no counterpart in Rendu).
	int intersect(const rdu_ray * rays, size_t n, rdu_hit * hits, uint32_t flags = 0, void * stream = nullptr) const {
		return rdu_raycaster_intersect(_h, rays, n, hits, flags, stream);
	}

	const rdu_raycaster * handle() const { return _h; }

private:
	rdu_raycaster * _h = nullptr;
	LogSink _log;
};

// PathTracer (PathTracer.hpp:10-76) over an rdu_scene the caller fills with rdu_scene_* calls
// (or from an adapter that walks Rendu's Scene).
class PathTracer {
public:
	explicit PathTracer(rdu_scene * scene, LogSink log = {}) : _scene(scene), _log(std::move(log)) {
		report(rdu_scene_commit(_scene), _log);
	}
	// Camera exposes position(), center(), up(), fov(), ratio() like Rendu's Camera; Image exposes
	// width, height, components and a float pixel buffer `pixels` (Image.hpp:86-102).
	template<typename Camera, typename Image>
	void render(const Camera & camera, size_t samples, size_t depth, Image & render, const rdu_render_opts * opts = nullptr) {
		rdu_camera_desc c{};
		for(int i = 0; i < 3; ++i) {
			c.position[i] = camera.position()[i];
			c.center[i] = camera.center()[i];
			c.up[i] = camera.up()[i];
		}
		c.fov = camera.fov();
		c.ratio = camera.ratio();
		report(rdu_pathtracer_render(_scene, &c, samples, depth, render.pixels.data(), render.width, render.height,
				   render.components, opts),
			_log);
	}
	const rdu_raycaster * raycaster() const { return rdu_scene_raycaster(_scene); }
	// Extension (no Rendu counterpart): replay repeated full passes from a captured HIP graph.
	void setGraph(bool enable) { report(rdu_scene_set_graph(_scene, enable ? 1 : 0), _log); }

private:
	rdu_scene * _scene;
	LogSink _log;
};

} // namespace rdu
