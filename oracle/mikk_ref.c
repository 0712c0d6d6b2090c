/* ORACLE — test infrastructure only: never linked into or called by the product.
 *
 * Callbacks that drive the reference's own mikktspace library (src/libs/mikktspace/mikktspace.c,
 * compiled from where it lies into oracle/_ref by oracle/Makefile) over an indexed triangle list
 * the way Rendu's Mesh does (Mesh.cpp:25-91, 368-394): three vertices per face, attributes read
 * through the index buffer, constant (0.5, 0.5) texture coordinates when none are given
 * (Mesh.cpp:362-366), one (tangent, sign) per face corner from setTSpaceBasic. */
#include <mikktspace.h>
#include <stddef.h>

typedef struct {
	const float * pos;
	const float * nrm;
	const float * uv;
	const unsigned * idx;
	int faces;
	float * out;
} RefMesh;

static int refFaces(const SMikkTSpaceContext * c) { return ((const RefMesh *)c->m_pUserData)->faces; }
static int refVerts(const SMikkTSpaceContext * c, const int f) { (void)c; (void)f; return 3; }

static void refPos(const SMikkTSpaceContext * c, float o[], const int f, const int v) {
	const RefMesh * m = (const RefMesh *)c->m_pUserData;
	const size_t i = m->idx[3 * f + v];
	o[0] = m->pos[3 * i]; o[1] = m->pos[3 * i + 1]; o[2] = m->pos[3 * i + 2];
}

static void refNrm(const SMikkTSpaceContext * c, float o[], const int f, const int v) {
	const RefMesh * m = (const RefMesh *)c->m_pUserData;
	const size_t i = m->idx[3 * f + v];
	o[0] = m->nrm[3 * i]; o[1] = m->nrm[3 * i + 1]; o[2] = m->nrm[3 * i + 2];
}

static void refUv(const SMikkTSpaceContext * c, float o[], const int f, const int v) {
	const RefMesh * m = (const RefMesh *)c->m_pUserData;
	const size_t i = m->idx[3 * f + v];
	if(m->uv) { o[0] = m->uv[2 * i]; o[1] = m->uv[2 * i + 1]; }
	else { o[0] = 0.5f; o[1] = 0.5f; }
}

static void refSet(const SMikkTSpaceContext * c, const float t[], const float sign, const int f, const int v) {
	float * o = ((RefMesh *)c->m_pUserData)->out + 4 * (3 * (size_t)f + v);
	o[0] = t[0]; o[1] = t[1]; o[2] = t[2]; o[3] = sign;
}

/* Returns genTangSpaceDefault's result (1 on success); out: 4 floats per corner. */
int ref_mikk_corner_tangents(const float * pos, const float * nrm, const float * uv, const unsigned * idx, int faces,
	float * out) {
	SMikkTSpaceInterface itf;
	SMikkTSpaceContext ctx;
	RefMesh m;
	itf.m_getNumFaces = refFaces;
	itf.m_getNumVerticesOfFace = refVerts;
	itf.m_getPosition = refPos;
	itf.m_getNormal = refNrm;
	itf.m_getTexCoord = refUv;
	itf.m_setTSpaceBasic = refSet;
	itf.m_setTSpace = NULL;
	m.pos = pos; m.nrm = nrm; m.uv = uv; m.idx = idx; m.faces = faces; m.out = out;
	ctx.m_pInterface = &itf;
	ctx.m_pUserData = &m;
	return genTangSpaceDefault(&ctx);
}
