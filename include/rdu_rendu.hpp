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
#include <memory>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace rdu {

using LogSink = std::function<void(const std::string &)>;

inline int report(int status, const LogSink & log) {
	if(status != RDU_OK && log)
		log(std::string("[rendu_amd] ") + rdu_last_error());
	return status;
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

// The Rendu classes the PathTracer adapter tells apart with dynamic_cast: the three analytic
// lights (src/engine/scene/lights/{Point,Spot,Directional}Light.hpp) and the atmosphere background
// (src/engine/scene/Sky.hpp). A tag: rdu::SceneTypes<PointLight, SpotLight, DirectionalLight, Sky>{}.
template<typename Point, typename Spot, typename Directional, typename SkyT>
struct SceneTypes {};

// Stand-in table type for scenes without an atmosphere background.
struct NoImage {
	unsigned int width = 0, height = 0, components = 0;
	std::vector<float> pixels;
};

// PathTracer (PathTracer.hpp:10-76).
//  - PathTracer(scene, SceneTypes<...>{}, log, table) mirrors PathTracer(const std::shared_ptr<Scene>&)
//    (PathTracer.cpp:11-21): it walks Rendu's Scene as render reads it — objects in order (mesh(),
//    model(), material(), useTexCoords(); the material's type(), masked() and textures 0-2), the
//    analytic lights, backgroundMode and background (its colour parameter, image, cubemap faces
//    or Sky direction) — into a library scene it owns, then builds the raycaster. An atmosphere
//    background also needs Rendu's 512 x 512 "scattering-precomputed" image (`table`), which
//    Rendu's MaterialSky fetches from its resource manager (MaterialSky.cpp:9).
//  - PathTracer(rdu_scene*, log) commits a library scene the caller filled with rdu_scene_* calls
//    (not owned).
class PathTracer {
public:
	template<typename SceneT, typename P, typename S, typename D, typename K, typename TableImage = NoImage>
	explicit PathTracer(const std::shared_ptr<SceneT> & scene, SceneTypes<P, S, D, K>, LogSink log = {},
		const TableImage * table = nullptr)
		: _log(std::move(log)), _owned(true) {
		if(!scene) {
			if(_log)
				_log("[PathTracer] No scene available.");
			return;
		}
		if(report(rdu_scene_create(&_scene), _log))
			return;
		Builder b(_scene, _log);
		for(const auto & obj : scene->objects)
			b.object(obj);
		for(const auto & light : scene->lights)
			b.template light<P, S, D>(light.get());
		b.template background<K>(int(scene->backgroundMode), scene->background.get(), table);
		if(!b.failed)
			report(rdu_scene_commit(_scene), _log);
	}
	explicit PathTracer(rdu_scene * scene, LogSink log = {}) : _scene(scene), _log(std::move(log)) {
		report(rdu_scene_commit(_scene), _log);
	}
	~PathTracer() {
		if(_owned)
			rdu_scene_destroy(_scene);
	}
	PathTracer(const PathTracer &) = delete;
	PathTracer & operator=(const PathTracer &) = delete;

	// PathTracer::render(camera, samples, depth, render) (PathTracer.cpp:159-299). Camera exposes
	// Rendu's position() and pixelShifts(corner, dx, dy) (Camera.hpp:84, Camera.cpp:42-49): the
	// frame render reads. Image exposes width, height, components and the float `pixels`
	// (Image.hpp:143-146), accumulated into and divided by the power-of-two sample count.
	template<typename Camera, typename Image>
	void render(const Camera & camera, size_t samples, size_t depth, Image & render, const rdu_render_opts * opts = nullptr) {
		using V = typename std::decay<decltype(camera.position())>::type;
		V corner, dx, dy;
		camera.pixelShifts(corner, dx, dy);
		const V & eye = camera.position();
		rdu_camera_frame f{};
		for(int i = 0; i < 3; ++i) {
			f.position[i] = eye[i];
			f.corner[i] = corner[i];
			f.dx[i] = dx[i];
			f.dy[i] = dy[i];
		}
		report(rdu_pathtracer_render_frame(_scene, &f, samples, depth, render.pixels.data(), render.width, render.height,
				   render.components, opts),
			_log);
	}
	const rdu_raycaster * raycaster() const { return rdu_scene_raycaster(_scene); }
	// Extension (no Rendu counterpart): replay repeated full passes from a captured HIP graph.
	void setGraph(bool enable) { report(rdu_scene_set_graph(_scene, enable ? 1 : 0), _log); }
	const rdu_scene * handle() const { return _scene; }

private:
	// Fills a library scene from Rendu's objects, lights and background.
	struct Builder {
		rdu_scene * scene;
		const LogSink & log;
		bool failed = false;
		std::vector<std::pair<const void *, int32_t>> textures, meshes, materials;
		Builder(rdu_scene * s, const LogSink & l) : scene(s), log(l) {}

		bool check(int status) {
			if(status != RDU_OK) {
				failed = true;
				report(status, log);
			}
			return status == RDU_OK;
		}
		static int32_t find(const std::vector<std::pair<const void *, int32_t>> & v, const void * key) {
			for(const auto & e : v)
				if(e.first == key)
					return e.second;
			return -1;
		}
		// Image::rgbal reads 4 floats per texel (Image.cpp:29-32): RGBA images are passed as they
		// are; others get their channels and alpha 1.
		template<typename Image>
		int32_t texture(const Image & img) {
			const int32_t known = find(textures, &img);
			if(known >= 0)
				return known;
			std::vector<float> rgba;
			const float * px = img.pixels.data();
			if(img.components != 4) {
				rgba.assign(size_t(img.width) * img.height * 4, 0.0f);
				for(size_t i = 0; i < size_t(img.width) * img.height; ++i) {
					for(unsigned int c = 0; c < img.components && c < 4; ++c)
						rgba[4 * i + c] = img.pixels[img.components * i + c];
					if(img.components < 4)
						rgba[4 * i + 3] = 1.0f;
				}
				px = rgba.data();
			}
			const rdu_texture_desc d{img.width, img.height, px};
			int32_t id = -1;
			if(!check(rdu_scene_add_texture(scene, &d, &id)))
				return 0;
			textures.emplace_back(&img, id);
			return id;
		}
		template<typename V>
		static std::vector<float> flat(const std::vector<V> & v, int n) {
			std::vector<float> out;
			out.reserve(v.size() * size_t(n));
			for(const auto & e : v)
				for(int k = 0; k < n; ++k)
					out.push_back(e[k]);
			return out;
		}
		template<typename Mesh>
		int32_t mesh(const Mesh * m) {
			const int32_t known = find(meshes, m);
			if(known >= 0)
				return known;
			if(m->tangents.empty() && log)
				log("The path tracer requires local tangent frames for all meshes.");
			const std::vector<float> pos = flat(m->positions, 3), nrm = flat(m->normals, 3), tan = flat(m->tangents, 3),
									 bit = flat(m->bitangents, 3), uv = flat(m->texcoords, 2);
			std::vector<uint32_t> idx(m->indices.begin(), m->indices.end());
			const rdu_mesh_desc d{pos.data(), nrm.data(), tan.data(), bit.data(), uv.empty() ? nullptr : uv.data(), idx.data(),
				m->positions.size(), idx.size()};
			int32_t id = -1;
			if(!check(rdu_scene_add_mesh(scene, &d, &id)))
				return 0;
			meshes.emplace_back(m, id);
			return id;
		}
		template<typename Material>
		int32_t material(const Material & mat) {
			const int32_t known = find(materials, &mat);
			if(known >= 0)
				return known;
			rdu_material_desc d{};
			d.type = int32_t(mat.type());
			d.masked = mat.masked() ? 1 : 0;
			for(int k = 0; k < 3; ++k)
				d.textures[k] = texture(mat.textures()[size_t(k)]->images[0]);
			int32_t id = -1;
			if(!check(rdu_scene_add_material(scene, &d, &id)))
				return 0;
			materials.emplace_back(&mat, id);
			return id;
		}
		template<typename Object>
		void object(const Object & obj) {
			rdu_object_desc d{};
			d.mesh = uint32_t(mesh(obj.mesh()));
			d.material = uint32_t(material(obj.material()));
			d.use_texcoords = obj.useTexCoords() ? 1 : 0;
			const auto & m = obj.model();
			for(int c = 0; c < 4; ++c)
				for(int r = 0; r < 4; ++r)
					d.model[4 * c + r] = m[c][r];
			check(rdu_scene_add_object(scene, &d));
		}
		template<typename P, typename S, typename D, typename Light>
		void light(const Light * l) {
			rdu_light_desc d{};
			d.casts_shadow = l->castsShadow() ? 1 : 0;
			for(int k = 0; k < 3; ++k)
				d.intensity[k] = l->intensity()[k];
			if(const P * p = dynamic_cast<const P *>(l)) {
				d.type = RDU_LIGHT_POINT;
				for(int k = 0; k < 3; ++k)
					d.position[k] = p->position()[k];
				d.radius = p->radius();
			} else if(const S * s = dynamic_cast<const S *>(l)) {
				d.type = RDU_LIGHT_SPOT;
				for(int k = 0; k < 3; ++k) {
					d.position[k] = s->position()[k];
					d.direction[k] = s->direction()[k];
				}
				d.radius = s->radius();
				d.angles[0] = s->angles()[0];
				d.angles[1] = s->angles()[1];
			} else if(const D * dl = dynamic_cast<const D *>(l)) {
				d.type = RDU_LIGHT_DIRECTIONAL;
				for(int k = 0; k < 3; ++k)
					d.direction[k] = dl->direction()[k];
			} else {
				if(log)
					log("[rendu_amd] unknown light type: skipped");
				return;
			}
			check(rdu_scene_add_light(scene, &d));
		}
		// Scene::Background COLOR, IMAGE, SKYBOX, ATMOSPHERE are 0..3, as RDU_BACKGROUND_*.
		template<typename K, typename Object, typename TableImage>
		void background(int mode, const Object * bg, const TableImage * table) {
			rdu_background_desc d{};
			d.mode = mode;
			d.texture = -1;
			for(int k = 0; k < 6; ++k)
				d.cube[k] = -1;
			if(bg) {
				const auto & mat = bg->material();
				if(mode == RDU_BACKGROUND_COLOR && !mat.parameters().empty()) {
					for(int k = 0; k < 3; ++k)
						d.color[k] = mat.parameters()[0][k];
				} else if(mode == RDU_BACKGROUND_IMAGE) {
					d.texture = texture(mat.textures()[0]->images[0]);
				} else if(mode == RDU_BACKGROUND_SKYBOX) {
					for(int k = 0; k < 6; ++k)
						d.cube[k] = texture(mat.textures()[0]->images[size_t(k)]);
				} else if(mode == RDU_BACKGROUND_ATMOSPHERE) {
					if(const K * sky = dynamic_cast<const K *>(bg))
						for(int k = 0; k < 3; ++k)
							d.sun_direction[k] = sky->direction()[k];
					if(!table || table->width == 0) {
						if(log)
							log("[rendu_amd] an atmosphere background needs the scattering table image");
						failed = true;
						return;
					}
					d.texture = texture(*table);
				}
			}
			check(rdu_scene_set_background(scene, &d));
		}
	};

	rdu_scene * _scene = nullptr;
	LogSink _log;
	bool _owned = false;
};

} // namespace rdu
