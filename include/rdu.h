/*
 * rdu.h — C ABI of rendu_amd, the MI355X (gfx950) implementation of Rendu's BVH raycaster and
 * CPU path tracer. Every entry point takes plain pointers and sizes; no C++ or torch types cross
 * this boundary. Status codes: RDU_OK (0) on success, a negative RDU_ERR_* otherwise, with a
 * message in rdu_last_error() (thread-local). No exceptions cross the ABI.
 *
 * Ownership follows Rendu's raycaster: the library copies what it is given
 * (Raycaster::addMesh copies vertices and indices, src/engine/raycaster/Raycaster.cpp:14-52);
 * callers own every input and output buffer.
 *
 * Threading: setup entry points (create, add_mesh, update_hierarchy) are single-threaded and must
 * all precede any query; query entry points may be called concurrently afterwards, as Rendu's
 * const intersects/intersectsAny are called from forParallel workers
 * (src/apps/pathtracer/PathTracer.cpp:186, 204).
 */
#ifndef RDU_H
#define RDU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RDU_OK 0
#define RDU_ERR_INVALID -1     /* bad argument (null pointer, size, unknown enum) */
#define RDU_ERR_STATE -2       /* call out of order (query before update_hierarchy, add after) */
#define RDU_ERR_HIP -3         /* HIP runtime error (launch, copy, no device) */
#define RDU_ERR_ALLOC -4       /* device allocation failed */
#define RDU_ERR_UNSUPPORTED -5 /* input the implementation does not handle */

/* One ray. The direction need not be unit length: like Rendu's Ray (Intersection.cpp:3-5) the
 * library uses normalize(d) and 1/normalize(d). tmin/tmax are Rendu's mini/maxi
 * (defaults 1e-4 and 1e8, Raycaster.hpp:65). 32 bytes. */
typedef struct {
	float ox, oy, oz, tmin;
	float dx, dy, dz, tmax;
} rdu_ray;

/* Closest hit, Rendu's Raycaster::Hit (Raycaster.hpp:16-42) without the derived w = 1 - u - v.
 * A miss is {t = FLT_MAX, u = v = 0, hit = 0, mesh_id = local_id = 0}, Hit() (Raycaster.cpp:6-8).
 * local_id is Hit::localId: 3 * (triangle index within its mesh). 24 bytes. */
typedef struct {
	float t, u, v;
	uint32_t hit, mesh_id, local_id;
} rdu_hit;

/* Flags for the batch queries. */
#define RDU_DEVICE_POINTERS 1u /* rays/outputs are device memory: asynchronous on `stream` */
#define RDU_REFERENCE_ORDER 2u /* visit nodes in Rendu's exact depth-first order
                                  (Raycaster.cpp:144-185): also equal hits on exact ties */
#define RDU_INCOHERENT 4u      /* hint: rays are incoherent; the library may process them in
                                  another order (results still land at their input index) */

typedef struct rdu_raycaster rdu_raycaster;

/* Replaces Raycaster::Raycaster() (Raycaster.hpp:45). Copy/move are deleted there: handles are
 * not copyable here either. */
int rdu_raycaster_create(rdu_raycaster** out);
void rdu_raycaster_destroy(rdu_raycaster* rc);

/* Replaces Raycaster::addMesh(const Mesh&, const glm::mat4&) (Raycaster.hpp:51, Raycaster.cpp:14-52).
 * xyz: n_vertices * 3 floats (Mesh::positions); indices: n_indices uint32 (Mesh::indices, a
 * multiple of 3); model: 16 floats, glm column-major, NULL for identity. mesh_id = call order. */
int rdu_raycaster_add_mesh(rdu_raycaster* rc, const float* xyz, size_t n_vertices, const uint32_t* indices,
	size_t n_indices, const float* model_colmajor);

/* Replaces Raycaster::updateHierarchy() (Raycaster.hpp:56, Raycaster.cpp:54-142): builds Rendu's
 * exact tree on the host, flattens it into the device node/triangle layout and uploads it. */
int rdu_raycaster_update_hierarchy(rdu_raycaster* rc);

/* Alternative to updateHierarchy (no reference counterpart): builds a different tree on the GPU
 * (linear BVH over all triangles: Morton codes of the centroids, radix sort, Karras radix tree, one
 * triangle per leaf, boxes bottom-up, 4-wide collapse) on `stream`. Closest hits keep Rendu's t,
 * u, v and triangle except where triangles tie at equal t; any-hit results are identical. The tree
 * has no Rendu visiting order: RDU_REFERENCE_ORDER batches fail with RDU_ERR_STATE, per-ray calls
 * use near-first order, export_nodes / export_triangles are unavailable. Call instead of
 * rdu_raycaster_update_hierarchy, once. */
int rdu_raycaster_update_hierarchy_gpu(rdu_raycaster* rc, void* stream);

/* Batch form of Raycaster::intersects(origin, direction, mini, maxi) (Raycaster.hpp:65,
 * Raycaster.cpp:144-185): hits[i] = closest hit of rays[i]. With RDU_DEVICE_POINTERS the call is
 * asynchronous on `stream` (a hipStream_t, NULL = default stream); otherwise it copies through
 * library staging buffers and returns when done. */
int rdu_raycaster_intersect(const rdu_raycaster* rc, const rdu_ray* rays, size_t n, rdu_hit* hits, uint32_t flags,
	void* stream);

/* Allocation-free variant of rdu_raycaster_intersect. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same results as rdu_raycaster_intersect with the same flags;
 * rays and hits must be device memory (flags must hold RDU_DEVICE_POINTERS) and the call is
 * asynchronous on `stream`. It allocates nothing: the spill stack and the ray-ordering buffers of a
 * batch come from a fixed 64 MiB device array that each device the library runs on holds until the
 * process ends. A batch needing more (about 2^21 RDU_INCOHERENT rays with ordering, fewer on deep
 * trees) is rejected with RDU_ERR_UNSUPPORTED before anything is launched. Every call on a device
 * shares that array: calls that overlap on one device race, so callers serialise them (one
 * stream, or their own ordering). */
int rdu_raycaster_intersect_noalloc(const rdu_raycaster* rc, const rdu_ray* rays, size_t n, rdu_hit* hits,
	uint32_t flags, void* stream);

/* Batch form of Raycaster::intersectsAny (Raycaster.hpp:74, Raycaster.cpp:187-224):
 * occluded[i] = 1 if any triangle is hit in (tmin, tmax), else 0. */
int rdu_raycaster_intersect_any(const rdu_raycaster* rc, const rdu_ray* rays, size_t n, uint8_t* occluded,
	uint32_t flags, void* stream);

/* Per-ray Raycaster::intersects with Rendu's signature and defaults (Raycaster.hpp:65,
 * Raycaster.cpp:144-185), synchronous and thread-safe, in Rendu's visiting order: every field
 * bit-identical to Rendu's and to the RDU_REFERENCE_ORDER batch. One ray is a chain of dependent
 * reads that a host core finishes sooner than a GPU round trip, so on a tree from
 * rdu_raycaster_update_hierarchy the call walks the host copy of the uploaded tree (no lock, no
 * allocation); on a GPU-built tree, or after rdu_raycaster_set_per_ray_device(rc, 1), it is a
 * one-ray GPU launch. Prefer the batch entry for throughput. */
int rdu_raycaster_intersect1(const rdu_raycaster* rc, const float origin[3], const float direction[3], float mini,
	float maxi, rdu_hit* out);

/* Where per-ray calls (intersect1, intersect_any1, visible) run on a host-built tree: 0 (default)
 * the host walk of the uploaded tree, 1 a one-ray GPU launch each (no reference counterpart). */
int rdu_raycaster_set_per_ray_device(rdu_raycaster* rc, int on_device);

/* Per-ray Raycaster::intersectsAny. *occluded = 1 on a hit. */
int rdu_raycaster_intersect_any1(const rdu_raycaster* rc, const float origin[3], const float direction[3],
	float mini, float maxi, int* occluded);

/* Raycaster::visible(p0, p1) (Raycaster.hpp:82, Raycaster.cpp:226-230):
 * !intersectsAny(p0, p1 - p0, 1e-4, |p1 - p0|). */
int rdu_raycaster_visible(const rdu_raycaster* rc, const float p0[3], const float p1[3], int* visible);

/* Size of the built structure (no reference counterpart; the reference logs these counts,
 * Raycaster.cpp:48-49, 141). */
typedef struct {
	uint64_t meshes, triangles, vertices, nodes; /* nodes = Rendu's _hierarchy.size() */
	uint64_t internal_nodes, max_depth;          /* depth counts nodes on the longest root-leaf path */
	uint64_t device_bytes;                       /* every array resident on the device */
	uint64_t wide_nodes, wide_depth;             /* 4-wide collapse used by near-first queries */
	uint64_t binary_bytes;                       /* bytes a reference-order query can touch */
	uint64_t wide_bytes;                         /* bytes a near-first query can touch */
} rdu_bvh_info;
int rdu_raycaster_info(const rdu_raycaster* rc, rdu_bvh_info* out);

/* Rendu's node array after updateHierarchy (Raycaster.hpp:127-132, _hierarchy), for parity
 * checks against a restatement: boxes n*6 floats (min xyz, max xyz), left/right n each (child
 * node ids, or first triangle / count for leaves), leaf n bytes. Arrays sized by info.nodes. */
int rdu_raycaster_export_nodes(const rdu_raycaster* rc, float* boxes, uint64_t* left, uint64_t* right, uint8_t* leaf);

/* Rendu's _triangles after updateHierarchy (Raycaster.hpp:117-124, permuted by the build):
 * mesh_id and local_id of each entry, info.triangles each. Leaves index this list. */
int rdu_raycaster_export_triangles(const rdu_raycaster* rc, uint32_t* mesh_id, uint32_t* local_id);

/* Device-side traversal statistics of the last batch launched with the same flags (diagnostic):
 * counts of internal-node expansions and triangle tests summed over the batch, obtained by
 * re-running the batch through an instrumented kernel. rays/n as for rdu_raycaster_intersect
 * (host memory). */
int rdu_raycaster_trace_stats(const rdu_raycaster* rc, const rdu_ray* rays, size_t n, int any_hit, uint32_t flags,
	uint64_t* node_visits, uint64_t* tri_tests);

/* ------------------------------------------------------------------------------------------------
 * Path tracer: Rendu's PathTracer (src/apps/pathtracer/PathTracer.hpp:10-76) over a scene
 * description that carries what PathTracer::render reads from Rendu's Scene
 * (src/engine/scene/Scene.hpp:65-77): objects in order (object i is raycaster mesh i), their
 * meshes, materials and textures, the analytic lights and the background.
 * ------------------------------------------------------------------------------------------------ */

/* A float RGBA image (Image with 4 components, row-major, y = 0 first row; Image.hpp:86-102).
 * Constant-colour textures are 8x8 images of that colour (ResourcesManager.cpp:489-504). */
typedef struct {
	uint32_t width, height;
	const float* rgba; /* width * height * 4 floats */
} rdu_texture_desc;

/* Mesh fields the path tracer reads (Mesh.hpp:126-134). normals/tangents/bitangents: n_vertices*3
 * floats (required); texcoords: n_vertices*2 floats or NULL (then the object cannot use UVs). */
typedef struct {
	const float* positions;
	const float* normals;
	const float* tangents;
	const float* bitangents;
	const float* texcoords;
	const uint32_t* indices;
	size_t n_vertices, n_indices;
} rdu_mesh_desc;

/* Material::Type values used by the path tracer (Material.hpp:12-88): every non-emissive type
 * is shaded as Lambert + GGX; EMISSIVE adds its texture-0 colour and stops the path. */
#define RDU_MATERIAL_REGULAR 1
#define RDU_MATERIAL_EMISSIVE 8

typedef struct {
	int32_t type;       /* RDU_MATERIAL_* */
	int32_t masked;     /* Material::masked(): alpha cut-out through texture 0's alpha */
	int32_t textures[3]; /* texture ids: 0 albedo+alpha, 1 normal map, 2 roughness/metal/AO */
} rdu_material_desc;

typedef struct {
	uint32_t mesh;          /* mesh id from rdu_scene_add_mesh */
	uint32_t material;      /* material id */
	int32_t use_texcoords;  /* Object::useTexCoords() (Object.hpp:69) */
	float model[16];        /* glm column-major model matrix (Object::model()) */
} rdu_object_desc;

#define RDU_LIGHT_POINT 0       /* PointLight::sample (PointLight.cpp:68-87) */
#define RDU_LIGHT_DIRECTIONAL 1 /* DirectionalLight::sample (DirectionalLight.cpp:40-44) */
#define RDU_LIGHT_SPOT 2        /* SpotLight::sample (SpotLight.cpp:52-81) */

typedef struct {
	int32_t type;
	int32_t casts_shadow;
	float intensity[3];
	float position[3];  /* point, spot */
	float direction[3]; /* directional, spot (the direction the light travels) */
	float radius;       /* point, spot: radius of influence */
	float angles[2];    /* spot: inner, outer half angles (radians) */
} rdu_light_desc;

#define RDU_BACKGROUND_COLOR 0      /* Scene::Background::COLOR */
#define RDU_BACKGROUND_IMAGE 1      /* IMAGE: texture stretched over the screen, direct hits only */
#define RDU_BACKGROUND_SKYBOX 2     /* SKYBOX: cubemap faces px, nx, py, ny, pz, nz */
#define RDU_BACKGROUND_ATMOSPHERE 3 /* ATMOSPHERE: MaterialSky::eval with a 512x512 scattering table */
#define RDU_BACKGROUND_LATLONG 4    /* extension: latitude-longitude environment map */

typedef struct {
	int32_t mode;
	float color[3];     /* COLOR */
	int32_t texture;    /* IMAGE, ATMOSPHERE (scattering table), LATLONG */
	int32_t cube[6];    /* SKYBOX */
	float sun_direction[3]; /* ATMOSPHERE */
} rdu_background_desc;

/* Camera::pose + projection (src/engine/input/Camera.cpp:10-18, 42-49). */
typedef struct {
	float position[3], center[3], up[3];
	float fov;   /* vertical field of view, radians */
	float ratio; /* width / height */
} rdu_camera_desc;

#define RDU_ESTIMATOR_REFERENCE 0 /* Rendu's estimator exactly (PathTracer.cpp:235-375) */
#define RDU_ESTIMATOR_EXTENDED 1  /* adds Russian roulette from rr_start (extension) */

typedef struct {
	uint32_t seed;       /* key of the counter-based random stream */
	int32_t estimator;   /* RDU_ESTIMATOR_* */
	int32_t rr_start;    /* extended: first bounce subject to Russian roulette */
	uint32_t max_paths;  /* paths per pass (0 = library default: 2^24 for plain renders with several
	                      * passes in flight, 2^25 with one, 2^22 for cached renders; the image does
	                      * not depend on it) */
	uint32_t passes_in_flight; /* plain renders: passes kept in flight at once on streams of the
	                            * scene's own, each with its own path state (0 = library default, 3;
	                            * at most 3; cached, graph-replayed and cut-out scenes take one);
	                            * the image does not depend on it */
} rdu_render_opts;

typedef struct rdu_scene rdu_scene;

int rdu_scene_create(rdu_scene** out);
void rdu_scene_destroy(rdu_scene* scene);
int rdu_scene_add_texture(rdu_scene* scene, const rdu_texture_desc* desc, int32_t* id);
int rdu_scene_add_mesh(rdu_scene* scene, const rdu_mesh_desc* desc, int32_t* id);
int rdu_scene_add_material(rdu_scene* scene, const rdu_material_desc* desc, int32_t* id);
int rdu_scene_add_object(rdu_scene* scene, const rdu_object_desc* desc);
int rdu_scene_add_light(rdu_scene* scene, const rdu_light_desc* desc);
int rdu_scene_set_background(rdu_scene* scene, const rdu_background_desc* desc);

/* Replaces PathTracer::PathTracer(scene) (PathTracer.cpp:11-21): adds every object's mesh with its
 * model to a raycaster (mesh id = object index), builds Rendu's hierarchy and uploads scene data. */
int rdu_scene_commit(rdu_scene* scene);
/* Before commit (no reference counterpart): build the scene's tree on the GPU
 * (rdu_raycaster_update_hierarchy_gpu) instead of Rendu's host build. */
int rdu_scene_set_gpu_build(rdu_scene* scene, int enable);
/* Before commit (no reference counterpart, an extension; Rendu bakes every object's model into
 * world-space triangles, Raycaster.cpp:19-26): instancing. A mesh shared by several objects gets one
 * tree of its own in object space (Rendu's build of that mesh alone), shared by those objects; a
 * mesh used by one object is baked into world space as Rendu does it. A 4-wide top level over the
 * objects' world boxes leads rays into the trees, transformed by the inverse model where the tree is
 * in object space (not for an identity model). Hits on shared meshes then differ from Rendu's in
 * the last bits (the transform's rounding), so renders match the flattened scene's statistically,
 * not bit for bit; a scene without shared meshes (or whose shared meshes' objects all have identity
 * models) renders as the flattened scene does. Such a scene has no raycaster handle
 * (rdu_scene_raycaster returns NULL), renders alone or in rdu_pathtracer_render_batch, and builds
 * on the host; the models of objects on shared meshes must be invertible. */
int rdu_scene_set_instancing(rdu_scene* scene, int enable);

/* The raycaster owned by a committed scene (PathTracer::raycaster(), PathTracer.hpp:29); NULL for
 * an instanced scene. */
const rdu_raycaster* rdu_scene_raycaster(const rdu_scene* scene);

/* Replaces PathTracer::render(camera, samples, depth, image) (PathTracer.cpp:159-299):
 * image is the caller's host buffer of width*height*components floats (components >= 3), which
 * like Rendu's Image is accumulated into (+=) and then divided by the sample count; the caller
 * zeroes it first. samples is rounded to a power of two as Rendu does. opts may be NULL. */
int rdu_pathtracer_render(rdu_scene* scene, const rdu_camera_desc* camera, size_t samples, size_t depth,
	float* image, uint32_t width, uint32_t height, uint32_t components, const rdu_render_opts* opts);

/* The camera as PathTracer::render reads it (PathTracer.cpp:176-196): Camera::position() and the
 * outputs of Camera::pixelShifts(corner, dx, dy) (Camera.cpp:42-49). */
typedef struct {
	float position[3], corner[3], dx[3], dy[3];
} rdu_camera_frame;

/* rdu_pathtracer_render with the camera given by its frame: a caller holding Rendu's Camera passes
 * exactly the frame Rendu's render would use (the frame of a rdu_camera_desc is recomputed from
 * its pose inputs and equals it only when they are Camera::pose's own inputs). */
int rdu_pathtracer_render_frame(rdu_scene* scene, const rdu_camera_frame* frame, size_t samples, size_t depth,
	float* image, uint32_t width, uint32_t height, uint32_t components, const rdu_render_opts* opts);

/* Timing and counters of the last render on this scene (no reference counterpart). */
typedef struct {
	double seconds;          /* wall time of the render call */
	uint64_t paths;          /* pixel samples traced */
	uint64_t extension_rays; /* closest-hit queries */
	uint64_t shadow_rays;    /* visibility queries */
	uint32_t samples;        /* spp after rounding */
	uint32_t passes;         /* passes of at most max_paths paths */
	uint32_t graph_passes;   /* passes replayed from a captured HIP graph */
	uint32_t graph_captures; /* graphs captured by this render */
} rdu_render_stats;
int rdu_scene_last_render_stats(const rdu_scene* scene, rdu_render_stats* out);

/* One job of a batch render: a committed scene, its camera (a description, or the frame Rendu's
 * render reads; exactly one of the two non-NULL), samples, depth, and the caller's image as in
 * rdu_pathtracer_render (accumulated into, then divided by the power-of-two sample count). */
typedef struct {
	rdu_scene* scene;
	const rdu_camera_desc* camera;
	const rdu_camera_frame* frame;
	size_t samples, depth;
	float* image;
	uint32_t width, height, components;
	rdu_render_opts opts; /* seed, estimator, rr_start; max_paths is ignored */
} rdu_batch_job;

/* Batched render jobs (no reference counterpart): every job of the list in one set of wavefront
 * launches per bounce. Each pass takes samples of several jobs (longest jobs first) up to max_paths
 * paths (0 = 2^24); their trees are copied into one node arena, rays start at their own job's
 * root, shading reads each wavefront's job descriptor once. Up to three passes are in flight on
 * streams of the call's own, each with its own path state; passes add into the images in order.
 * Every image equals, bit for bit, what
 * rdu_pathtracer_render gives that scene alone with the same options. job_seconds (n entries, or
 * NULL) receives each job's completion time from the call's start; stats (or NULL) the totals
 * (seconds, paths, rays, passes). The scenes must live on one device and not be rendered
 * elsewhere during the call. */
int rdu_pathtracer_render_batch(const rdu_batch_job* jobs, size_t n, uint32_t max_paths, double* job_seconds,
	rdu_render_stats* stats);

/* HIP graph replay of render passes (no reference counterpart). When enabled, a plain render's
 * full passes (max_paths paths each; a shorter last pass runs eagerly) are captured once into a
 * HIP graph and replayed: a pass whose launch configuration (camera frame, size, depth,
 * estimator, buffers, traversal tuning) equals the previous full pass's is captured at that point
 * and replayed for every later pass with the same configuration, within a render and across
 * renders of the scene (also with another seed: progressive renders replay). Each pass reads its
 * first sample index and the seed from device memory, so a replay renders exactly what the eager
 * pass would. Cached renders, and renders of scenes with masked (alpha cut-out) materials, whose
 * shadow walk checks on the host whether rays still walk, always run eagerly. Off by default. */
int rdu_scene_set_graph(rdu_scene* scene, int enable);


/* Autotuner (no reference counterpart: Rendu's traversal has no launch parameters). Times the
 * near-first persistent traversal on the caller's batch over candidate launch parameters by
 * coordinate descent: the wavefront refill threshold (idle lanes a wavefront waits for before
 * fetching new rays: 4..64), the ray tile (rays per interleaved run of a wavefront: 32..512) and
 * the grid size (resident workgroups per CU: 1/4..3/4 of the occupancy limit). Keeps the fastest
 * for later queries of the same class (RDU_INCOHERENT or not, closest or any hit) on this
 * raycaster, and stores it in the process-wide tuning cache under the scene's statistics
 * (triangle count, 4-wide tree depth, mesh count, each as a power-of-two bucket where a count;
 * batch size bucket; class; host- or GPU-built tree). Batches of a class this raycaster has not
 * tuned take the cache's entry for their statistics, if any. flags as for the batch queries; the
 * batch is only read. Not thread-safe against concurrent queries on the same raycaster. */
typedef struct {
	uint32_t refill_min;    /* chosen threshold (0 = default: 32 sorted batches, 16 input order) */
	float best_ms;          /* mean time of one batch with the chosen parameters */
	float default_ms;       /* mean time with the untuned defaults */
	uint32_t candidates;    /* parameter sets timed */
	uint32_t chunk;         /* chosen ray tile (0 = default: 64 sorted, contiguous runs input order) */
	uint32_t blocks_per_cu; /* chosen workgroups per CU (0 = occupancy limit) */
	uint32_t source;        /* rdu_raycaster_tuning: 0 defaults, 1 this raycaster, 2 tuning cache */
} rdu_tune_report;
int rdu_raycaster_autotune(rdu_raycaster* rc, const rdu_ray* rays, size_t n, int any_hit, uint32_t flags, void* stream,
	rdu_tune_report* report);
/* The launch parameters a batch of n rays of this class would use now, and where they come from
 * (refill_min, chunk, blocks_per_cu, source; the times are 0). */
int rdu_raycaster_tuning(const rdu_raycaster* rc, size_t n, int any_hit, uint32_t flags, rdu_tune_report* out);
/* Sets this raycaster's launch parameters for a class without timing anything (refill_min 0..64,
 * chunk 0 or 16..2^20, blocks_per_cu 0..64 from `params`; 0 = default; other fields ignored). */
int rdu_raycaster_set_tuning(rdu_raycaster* rc, int any_hit, uint32_t flags, const rdu_tune_report* params);
/* The process-wide tuning cache: save to / merge from a text file (a header line, then one entry
 * per line), clear, count entries. A malformed file is rejected whole with RDU_ERR_INVALID. */
int rdu_tune_cache_save(const char* path);
int rdu_tune_cache_load(const char* path);
int rdu_tune_cache_clear(void);
int rdu_tune_cache_size(size_t* entries);

/* Autotuner of render passes (no reference counterpart): times one pass of the given render
 * configuration (the image is not touched) for each candidate of the traversal launch
 * parameters, as rdu_raycaster_autotune does for batches, and keeps the fastest for this
 * scene's later renders; it also enters the process-wide tuning cache under the scene's
 * statistics and the pass size, where untuned scenes of the same class find it. */
int rdu_scene_autotune(rdu_scene* scene, const rdu_camera_desc* camera, size_t samples, size_t depth, uint32_t width,
	uint32_t height, const rdu_render_opts* opts, rdu_tune_report* report);

/* Kernel timing (no reference counterpart; for measurement). While enabled, each batch query
 * records a pair of HIP events on its stream around the traversal kernel alone (not the ray
 * ordering or copies), for up to 4096 launches. rdu_raycaster_kernel_time waits for the recorded
 * launches, returns their summed duration and count, and starts a new count. */
int rdu_raycaster_set_kernel_timing(rdu_raycaster* rc, int enable);
int rdu_raycaster_kernel_time(rdu_raycaster* rc, double* total_ms, uint64_t* launches);

/* ------------------------------------------------------------------------------------------------
 * Output step of the path-tracer app: replaces the tonemap loop of src/apps/pathtracer/main.cpp:109-116
 * and the 8-bit conversion of Image::saveAsLDR (src/engine/resources/Image.cpp:126-158). pixels:
 * width x height x components floats, row-major, row 0 on top; out: the same count of bytes.
 * RDU_IMAGE_TONEMAP: c -> 1 - exp(-c) on the colour channels; RDU_IMAGE_SRGB: v -> pow(v, 1/2.2)
 * except alpha; then byte = (uint8)min(255, max(0, 255 v)); RDU_IMAGE_IGNORE_ALPHA: alpha byte 255.
 * flags may add RDU_DEVICE_POINTERS (then `stream` orders the work and nothing is synchronised).
 * ------------------------------------------------------------------------------------------------ */
#define RDU_IMAGE_TONEMAP 0x10u
#define RDU_IMAGE_SRGB 0x20u
#define RDU_IMAGE_IGNORE_ALPHA 0x40u
int rdu_image_to_ldr(const float* pixels, size_t width, size_t height, size_t components, uint8_t* out, uint32_t flags,
	void* stream);

/* ------------------------------------------------------------------------------------------------
 * Learned radiance cache (no Rendu counterpart; an extension of the path tracer): outgoing RGB
 * radiance at a position in a direction, from a 16-level hash grid over the scene box (2^19 entries x
 * 2 fp16 features per level) and degree-3 spherical harmonics of the direction (16 features),
 * followed by a 5 x 64 ReLU MLP, trained online with bf16 MFMA kernels accumulating in fp32 (Adam,
 * relative L2 loss). A query is 8 floats laid out like rdu_ray: position xyz, unused, outgoing
 * direction xyz, unused.
 * ------------------------------------------------------------------------------------------------ */
typedef struct rdu_rcache rdu_rcache;
#define RDU_RCACHE_MLP_PARAMS 20480      /* W0 64x48, W1..W4 64x64, W5 16x64 (rows >= 3 zero) */
#define RDU_RCACHE_TABLE_PARAMS 16777216 /* 16 levels x 2^19 entries x 2 features */

int rdu_rcache_create(rdu_rcache** out, const float scene_min[3], const float scene_max[3], uint32_t seed);
void rdu_rcache_destroy(rdu_rcache* cache);
int rdu_rcache_set_learning_rate(rdu_rcache* cache, float lr);
/* queries: n x 8 floats; radiance: n x 4 floats out (rgb, unused). flags: RDU_DEVICE_POINTERS. */
int rdu_rcache_infer(rdu_rcache* cache, const float* queries, size_t n, float* radiance, uint32_t flags, void* stream);
/* steps optimisation steps on one batch of (query, target radiance) pairs (n x 8 and n x 4 floats):
 * Adam (lr, beta1 0.9, beta2 0.99, eps 1e-15) on the MLP weights with an l2 term of 1e-6, and lazy Adam
 * on the hash tables: only the entries the batch touches update their moments and weights. */
int rdu_rcache_train(rdu_rcache* cache, const float* queries, const float* targets, size_t n, uint32_t steps,
	uint32_t flags, void* stream, float* mean_loss);
/* Gradients of the mean relative L2 loss for one batch, without an update (host arrays):
 * mlp_grad RDU_RCACHE_MLP_PARAMS floats, table_grad RDU_RCACHE_TABLE_PARAMS floats or NULL. */
int rdu_rcache_gradients(rdu_rcache* cache, const float* queries, const float* targets, size_t n, float* mlp_grad,
	float* table_grad);
int rdu_rcache_get_params(const rdu_rcache* cache, float* mlp, float* tables, int32_t* level_resolutions);
int rdu_rcache_set_params(rdu_rcache* cache, const float* mlp, const float* tables);

/* Path tracing with the cache (no Rendu counterpart). As rdu_pathtracer_render, except that each
 * pass of paths:
 *  - traces one pixel in 16 (a 4 x 4 pattern shifted every pass) in full, recording every surface
 *    vertex from bounce `terminate_bounce` on; after the pass each record becomes a training pair
 *    (query at the vertex towards the previous one, target = radiance the rest of the path brought
 *    back / throughput on arrival, clamped to [0, 100]); 1 record in 8 is held out;
 *  - ends every other path at its first surface vertex from `terminate_bounce` on where the cache
 *    is trusted, adding throughput x cache prediction;
 *  - then runs `train_steps` Adam steps of `train_batch` pairs (none when `frozen`), and updates
 *    trust from the held-out pairs, per cell of a fine 32^3 and a coarse 8^3 grid over the cache's
 *    box: a cell is trusted once it has `min_count` held-out pairs (running sums scaled by 0.9
 *    every pass) and |sum(prediction) - sum(target)| in luminance is at most `max_bias` of the
 *    larger sum plus twice the standard deviation of the target sum; a vertex uses its fine cell,
 *    or its coarse cell while the fine one has too few pairs. Untrusted vertices are traced on,
 *    as without the cache.
 * Zero fields take the defaults: terminate_bounce 2, train_steps 4, train_batch 16384, max_bias 0.2,
 * min_count 8. `always_trust` skips the cell test (every eligible vertex ends in the cache); with
 * `frozen` as well nothing is recorded and no pixel is traced in full. The estimator options (RR,
 * seeds) are those of `opts`. */
typedef struct {
	uint32_t terminate_bounce;
	uint32_t train_steps;
	uint32_t train_batch;
	float max_bias;
	float min_count;
	uint32_t frozen;
	uint32_t always_trust;
	uint32_t from_primary; /* 1: terminate from bounce 0 (the primary vertex); overrides terminate_bounce */
} rdu_cache_opts;
typedef struct {
	uint64_t cache_queries; /* paths ended in the cache */
	uint64_t records;       /* training-path vertices recorded */
	uint64_t train_pairs;   /* pairs available for training (records minus held out) */
	uint64_t train_steps;   /* optimisation steps run */
	uint32_t trusted_cells; /* cells trusted at the end of the render */
} rdu_cache_stats;
int rdu_pathtracer_render_cached(rdu_scene* scene, rdu_rcache* cache, const rdu_camera_desc* camera, size_t samples,
	size_t depth, float* image, uint32_t width, uint32_t height, uint32_t components, const rdu_render_opts* opts,
	const rdu_cache_opts* cache_opts);
int rdu_scene_last_cache_stats(const rdu_scene* scene, rdu_cache_stats* out);
/* The training pairs of the last pass of the last cached render (host arrays of `capacity` pairs:
 * queries 8 floats, targets 4 floats each); *count receives the number of pairs available. */
int rdu_scene_last_cache_pairs(const rdu_scene* scene, float* queries, float* targets, size_t capacity, size_t* count);

/* Tangent frames of a triangle mesh as Mesh::computeTangentsAndBitangents(force) makes them
 * (Mesh.cpp:356-508), host code. positions and normals: n_vertices x 3 floats; texcoords:
 * n_vertices x 2 floats, or NULL for the constant (0.5, 0.5) of the forced frame (Mesh.cpp:362-366);
 * indices: n_indices (a multiple of 3) vertex indices, rewritten in place. Per triangle corner,
 * MikkTSpace's tangent (genTangSpaceDefault of the mikktspace library, Mesh.cpp:386-394); then each
 * vertex keeps its first corner's frame, a later corner whose frame differs from the corner before
 * it gets a new vertex, appended after the existing ones. Outputs (capacity >= n_vertices +
 * n_indices entries): source_vertex[v] = the input vertex that vertex v copies (identity for the
 * first n_vertices), tangents and bitangents (3 floats per vertex, unit length, bitangent = -sign *
 * cross(n, t) normalised); *out_vertices = the vertex count after the split. No triangle: every
 * vertex gets tangent +x and bitangent +y (Mesh.cpp:388-392). */
int rdu_mesh_tangent_frames(const float* positions, const float* normals, const float* texcoords, size_t n_vertices,
	uint32_t* indices, size_t n_indices, uint32_t* source_vertex, float* tangents, float* bitangents, size_t capacity,
	size_t* out_vertices);
/* The MikkTSpace step alone (the callbacks of Mesh.cpp:25-91 over an indexed triangle list):
 * corner_tangents receives n_indices x 4 floats, the tangent and the bitangent sign (+1 or -1) of
 * every triangle corner, as the library's setTSpaceBasic callback reports them. */
int rdu_mesh_corner_tangents(const float* positions, const float* normals, const float* texcoords, size_t n_vertices,
	const uint32_t* indices, size_t n_indices, float* corner_tangents);

/* Last error message of the calling thread ("" if none). */
const char* rdu_last_error(void);

/* Library version string. */
const char* rdu_version(void);

#ifdef __cplusplus
}
#endif

#endif /* RDU_H */
