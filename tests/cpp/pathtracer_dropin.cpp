// A Rendu-style caller of rdu::PathTracer: the scene is held in types shaped like Rendu's own
// (Scene, Object, Mesh, Material, Texture, Image, the three lights, Sky, Camera: the members and
// accessors PathTracer.cpp reads), the tracer is built from a std::shared_ptr<Scene> as
// PathTracer::PathTracer(const std::shared_ptr<Scene>&) is (PathTracer.hpp:18-21), and render() takes
// the Camera and a pre-zeroed Image as PathTracer::render does (PathTracer.hpp:26).
//
// Usage: pathtracer_dropin <scene.bin> <image.bin>
//   scene.bin (little endian): u32 textures; per texture u32 w, h, f32 rgba[w h 4]
//     u32 meshes; per mesh u32 nv, ni, has_uv; f32 pos[3 nv], nrm[3 nv], tan[3 nv], bit[3 nv],
//       (uv[2 nv] if has_uv); u32 idx[ni]
//     u32 materials; per material i32 type, masked, tex[3]
//     u32 objects; per object u32 mesh, material, use_uv; f32 model[16] (column-major)
//     u32 lights; per light i32 type, casts_shadow; f32 intensity[3], position[3], direction[3],
//       radius, angles[2]
//     background: i32 mode; f32 color[3]; i32 texture; i32 cube[6]; f32 sun[3]
//     camera: f32 position[3], center[3], up[3], fov, ratio
//     render: u32 width, height, samples, depth, seed
//   image.bin: f32 pixels[height width 3] (written)
#include "rdu_rendu.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

// --- Stand-ins with the shapes of Rendu's types ---------------------------------------------------
struct vec2 {
	float v[2];
	float operator[](int i) const { return v[i]; }
};
struct vec3 {
	float v[3];
	float operator[](int i) const { return v[i]; }
	float & operator[](int i) { return v[i]; }
};
struct vec4 {
	float v[4];
	float operator[](int i) const { return v[i]; }
};
struct mat4 {
	float c[4][4];
	const float * operator[](int i) const { return c[i]; }
};
// glm 0.9.9.7 formulas (func_geometric.inl) for the Camera stand-in.
static vec3 sub(vec3 a, vec3 b) { return {{a[0] - b[0], a[1] - b[1], a[2] - b[2]}}; }
static vec3 add(vec3 a, vec3 b) { return {{a[0] + b[0], a[1] + b[1], a[2] + b[2]}}; }
static vec3 scale(float s, vec3 a) { return {{s * a[0], s * a[1], s * a[2]}}; }
static float dot(vec3 a, vec3 b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static vec3 cross(vec3 x, vec3 y) {
	return {{x[1] * y[2] - y[1] * x[2], x[2] * y[0] - y[2] * x[0], x[0] * y[1] - y[0] * x[1]}};
}
static vec3 normalize(vec3 v) { return scale(1.0f / std::sqrt(dot(v, v)), v); }

struct Image { // Image.hpp:143-146
	unsigned int width = 0, height = 0, components = 0;
	std::vector<float> pixels;
};
struct Texture { // Texture.hpp:96
	std::vector<Image> images;
};
struct Mesh { // Mesh.hpp:126-134
	std::vector<vec3> positions, normals, tangents, bitangents;
	std::vector<vec2> texcoords;
	std::vector<unsigned int> indices;
};
struct Material { // Material.hpp:111-132
	enum Type : int { None = 0, Regular, Parallax, Clearcoat, Anisotropic, Sheen, Iridescent, Subsurface, Emissive };
	Type _type = Regular;
	bool _masked = false;
	std::vector<const Texture *> _textures;
	std::vector<vec4> _parameters;
	const std::vector<const Texture *> & textures() const { return _textures; }
	const std::vector<vec4> & parameters() const { return _parameters; }
	const Type & type() const { return _type; }
	bool masked() const { return _masked; }
};
struct Object { // Object.hpp:54-82
	const Mesh * _mesh = nullptr;
	const Material * _material = nullptr;
	mat4 _model{};
	bool _skipUVs = false;
	virtual ~Object() = default;
	const Mesh * mesh() const { return _mesh; }
	const mat4 & model() const { return _model; }
	const Material & material() const { return *_material; }
	bool useTexCoords() const { return !_skipUVs; }
};
struct Sky final : Object { // Sky.hpp:49
	vec3 _sunDirection{{0.0f, 1.0f, 0.0f}};
	const vec3 & direction() const { return _sunDirection; }
};
struct Light { // Light.hpp:68-78
	vec3 _color{};
	bool _castShadows = true;
	virtual ~Light() = default;
	bool castsShadow() const { return _castShadows; }
	const vec3 & intensity() const { return _color; }
};
struct PointLight : Light { // PointLight.hpp:71-76
	vec3 _pos{};
	float _radius = 1.0f;
	const vec3 & position() const { return _pos; }
	float radius() const { return _radius; }
};
struct SpotLight : Light { // SpotLight.hpp:74-89
	vec3 _pos{}, _dir{};
	vec2 _angles{};
	float _radius = 1.0f;
	const vec3 & position() const { return _pos; }
	const vec3 & direction() const { return _dir; }
	const vec2 & angles() const { return _angles; }
	float radius() const { return _radius; }
};
struct DirectionalLight : Light { // DirectionalLight.hpp:67
	vec3 _dir{};
	const vec3 & direction() const { return _dir; }
};
struct Scene { // Scene.hpp:65-77
	enum class Background { COLOR, IMAGE, SKYBOX, ATMOSPHERE };
	std::vector<Object> objects;
	std::vector<std::shared_ptr<Light>> lights;
	Background backgroundMode = Background::COLOR;
	std::unique_ptr<Object> background;
};
struct Camera { // Camera.cpp:10-18, 42-49
	vec3 _eye{}, _center{}, _up{}, _right{};
	float _fov = 1.0f, _ratio = 1.0f;
	void pose(vec3 position, vec3 center, vec3 up) {
		_eye = position;
		_center = center;
		_up = normalize(up);
		const vec3 viewDir = normalize(sub(_center, _eye));
		_right = cross(viewDir, _up);
		_up = cross(_right, viewDir);
	}
	void projection(float ratio, float fov) {
		_ratio = ratio;
		_fov = fov;
	}
	const vec3 & position() const { return _eye; }
	void pixelShifts(vec3 & corner, vec3 & dx, vec3 & dy) const {
		const float heightScale = std::tan(0.5f * _fov);
		const float widthScale = _ratio * heightScale;
		const vec3 d = sub(_eye, _center);
		const float imageDist = std::sqrt(dot(d, d)); // glm::distance
		corner = add(_center, scale(imageDist, add(scale(-widthScale, _right), scale(heightScale, _up))));
		dx = scale(2.0f * widthScale * imageDist, _right);
		dy = scale(-2.0f * heightScale * imageDist, _up);
	}
};

// --- Reading the dump -------------------------------------------------------------------------------
struct Reader {
	FILE * f;
	bool ok = true;
	template<typename T>
	T get() {
		T v{};
		ok = ok && std::fread(&v, sizeof(T), 1, f) == 1;
		return v;
	}
	template<typename T>
	void fill(T * p, size_t n) {
		ok = ok && std::fread(p, sizeof(T), n, f) == n;
	}
};

int main(int argc, char ** argv) {
	if(argc < 3)
		return 2;
	FILE * f = std::fopen(argv[1], "rb");
	if(!f)
		return 2;
	Reader rd{f};
	std::vector<std::unique_ptr<Texture>> textures;
	const uint32_t nt = rd.get<uint32_t>();
	for(uint32_t i = 0; i < nt && rd.ok; ++i) {
		auto t = std::make_unique<Texture>();
		Image img;
		img.width = rd.get<uint32_t>();
		img.height = rd.get<uint32_t>();
		img.components = 4;
		img.pixels.resize(size_t(img.width) * img.height * 4);
		rd.fill(img.pixels.data(), img.pixels.size());
		t->images.push_back(std::move(img));
		textures.push_back(std::move(t));
	}
	std::vector<std::unique_ptr<Mesh>> meshes;
	const uint32_t nm = rd.get<uint32_t>();
	for(uint32_t i = 0; i < nm && rd.ok; ++i) {
		auto m = std::make_unique<Mesh>();
		const uint32_t nv = rd.get<uint32_t>(), ni = rd.get<uint32_t>(), hasUV = rd.get<uint32_t>();
		for(auto * a : {&m->positions, &m->normals, &m->tangents, &m->bitangents}) {
			a->resize(nv);
			rd.fill(a->data(), nv);
		}
		if(hasUV) {
			m->texcoords.resize(nv);
			rd.fill(m->texcoords.data(), nv);
		}
		m->indices.resize(ni);
		rd.fill(m->indices.data(), ni);
		meshes.push_back(std::move(m));
	}
	std::vector<std::unique_ptr<Material>> materials;
	const uint32_t nmat = rd.get<uint32_t>();
	for(uint32_t i = 0; i < nmat && rd.ok; ++i) {
		auto m = std::make_unique<Material>();
		m->_type = Material::Type(rd.get<int32_t>());
		m->_masked = rd.get<int32_t>() != 0;
		for(int k = 0; k < 3; ++k)
			m->_textures.push_back(textures[size_t(rd.get<int32_t>())].get());
		materials.push_back(std::move(m));
	}
	auto scene = std::make_shared<Scene>();
	const uint32_t no = rd.get<uint32_t>();
	for(uint32_t i = 0; i < no && rd.ok; ++i) {
		Object o;
		o._mesh = meshes[rd.get<uint32_t>()].get();
		o._material = materials[rd.get<uint32_t>()].get();
		o._skipUVs = rd.get<uint32_t>() == 0;
		rd.fill(&o._model.c[0][0], 16);
		scene->objects.push_back(o);
	}
	const uint32_t nl = rd.get<uint32_t>();
	for(uint32_t i = 0; i < nl && rd.ok; ++i) {
		const int32_t type = rd.get<int32_t>();
		const bool shadow = rd.get<int32_t>() != 0;
		vec3 inten, pos, dir;
		rd.fill(inten.v, 3);
		rd.fill(pos.v, 3);
		rd.fill(dir.v, 3);
		const float radius = rd.get<float>();
		vec2 ang;
		rd.fill(ang.v, 2);
		std::shared_ptr<Light> l;
		if(type == RDU_LIGHT_POINT) {
			auto p = std::make_shared<PointLight>();
			p->_pos = pos;
			p->_radius = radius;
			l = p;
		} else if(type == RDU_LIGHT_SPOT) {
			auto s = std::make_shared<SpotLight>();
			s->_pos = pos;
			s->_dir = dir;
			s->_angles = ang;
			s->_radius = radius;
			l = s;
		} else {
			auto d = std::make_shared<DirectionalLight>();
			d->_dir = dir;
			l = d;
		}
		l->_color = inten;
		l->_castShadows = shadow;
		scene->lights.push_back(l);
	}
	const int32_t mode = rd.get<int32_t>();
	vec3 color, sun;
	rd.fill(color.v, 3);
	const int32_t bgTex = rd.get<int32_t>();
	int32_t cube[6];
	rd.fill(cube, 6);
	rd.fill(sun.v, 3);
	Camera cam;
	vec3 cp, cc, cu;
	rd.fill(cp.v, 3);
	rd.fill(cc.v, 3);
	rd.fill(cu.v, 3);
	const float fov = rd.get<float>(), ratio = rd.get<float>();
	cam.pose(cp, cc, cu);
	cam.projection(ratio, fov);
	const uint32_t w = rd.get<uint32_t>(), h = rd.get<uint32_t>(), spp = rd.get<uint32_t>(), depth = rd.get<uint32_t>(),
				   seed = rd.get<uint32_t>();
	std::fclose(f);
	if(!rd.ok) {
		std::printf("truncated scene file\n");
		return 2;
	}
	// Scene::background: an Object whose material carries the colour, image or cubemap; a Sky for
	// the atmosphere, with the scattering table separately (Rendu's resource manager holds it).
	Material bgMat;
	Texture cubemap;
	const Image * table = nullptr;
	scene->backgroundMode = Scene::Background(mode);
	if(mode == RDU_BACKGROUND_ATMOSPHERE) {
		auto sky = std::make_unique<Sky>();
		sky->_sunDirection = sun;
		sky->_material = &bgMat;
		scene->background = std::move(sky);
		table = &textures[size_t(bgTex)]->images[0];
	} else {
		scene->background = std::make_unique<Object>();
		scene->background->_material = &bgMat;
		if(mode == RDU_BACKGROUND_COLOR) {
			bgMat._parameters.push_back({{color[0], color[1], color[2], 1.0f}});
		} else if(mode == RDU_BACKGROUND_IMAGE) {
			bgMat._textures.push_back(textures[size_t(bgTex)].get());
		} else if(mode == RDU_BACKGROUND_SKYBOX) {
			for(int k = 0; k < 6; ++k)
				cubemap.images.push_back(textures[size_t(cube[k])]->images[0]);
			bgMat._textures.push_back(&cubemap);
		}
	}

	std::vector<std::string> log;
	const rdu::LogSink sink = [&](const std::string & m) { log.push_back(m); };
	std::unique_ptr<rdu::PathTracer> pt;
	if(table)
		pt.reset(new rdu::PathTracer(scene, rdu::SceneTypes<PointLight, SpotLight, DirectionalLight, Sky>{}, sink, table));
	else
		pt.reset(new rdu::PathTracer(scene, rdu::SceneTypes<PointLight, SpotLight, DirectionalLight, Sky>{}, sink));
	Image render;
	render.width = w;
	render.height = h;
	render.components = 3;
	render.pixels.assign(size_t(w) * h * 3, 0.0f);
	rdu_render_opts opts{};
	opts.seed = seed;
	pt->render(cam, spp, depth, render, &opts);
	for(const auto & m : log)
		std::printf("logged: %s\n", m.c_str());
	if(!log.empty())
		return 3;
	FILE * o = std::fopen(argv[2], "wb");
	if(!o || std::fwrite(render.pixels.data(), 4, render.pixels.size(), o) != render.pixels.size())
		return 2;
	std::fclose(o);
	std::printf("rendered %ux%u %u spp\n", w, h, spp);
	return 0;
}
