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

/* Batch form of Raycaster::intersects(origin, direction, mini, maxi) (Raycaster.hpp:65,
 * Raycaster.cpp:144-185): hits[i] = closest hit of rays[i]. With RDU_DEVICE_POINTERS the call is
 * asynchronous on `stream` (a hipStream_t, NULL = default stream); otherwise it copies through
 * library staging buffers and returns when done. */
int rdu_raycaster_intersect(const rdu_raycaster* rc, const rdu_ray* rays, size_t n, rdu_hit* hits, uint32_t flags,
	void* stream);

/* Batch form of Raycaster::intersectsAny (Raycaster.hpp:74, Raycaster.cpp:187-224):
 * occluded[i] = 1 if any triangle is hit in (tmin, tmax), else 0. */
int rdu_raycaster_intersect_any(const rdu_raycaster* rc, const rdu_ray* rays, size_t n, uint8_t* occluded,
	uint32_t flags, void* stream);

/* Per-ray Raycaster::intersects with Rendu's signature and defaults. Runs as a batch of one on the
 * GPU in reference order (synchronous); prefer the batch entry for throughput. */
int rdu_raycaster_intersect1(const rdu_raycaster* rc, const float origin[3], const float direction[3], float mini,
	float maxi, rdu_hit* out);

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

/* Last error message of the calling thread ("" if none). */
const char* rdu_last_error(void);

/* Library version string. */
const char* rdu_version(void);

#ifdef __cplusplus
}
#endif

#endif /* RDU_H */
