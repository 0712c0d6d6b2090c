// A Rendu-style caller of the drop-in classes: the code a Rendu maintainer writes is unchanged
// apart from the include and the namespace (Raycaster.hpp:51-82 call shapes).
#include "rdu_rendu.hpp"

#include <array>
#include <cstdio>
#include <vector>

// Stand-ins with the member names the templates read (Rendu's glm::vec3 / glm::mat4 / Mesh).
struct Vec3 {
	float v[3];
	float operator[](int i) const { return v[i]; }
};
struct Mat4 {
	std::array<std::array<float, 4>, 4> c;
	const std::array<float, 4> & operator[](int i) const { return c[i]; }
};
struct Mesh {
	std::vector<Vec3> positions;
	std::vector<unsigned int> indices;
};

int main() {
	std::vector<std::string> log;
	rdu::Raycaster rc([&](const std::string & m) { log.push_back(m); });
	Mesh quad;
	quad.positions = {{{-1, -1, 0}}, {{1, -1, 0}}, {{1, 1, 0}}, {{-1, 1, 0}}};
	quad.indices = {0, 1, 2, 0, 2, 3};
	Mat4 identity{};
	for(int i = 0; i < 4; ++i)
		identity.c[i][i] = 1.0f;
	rc.addMesh(quad, identity);
	rc.updateHierarchy();
	const rdu::Raycaster::Hit h = rc.intersects(Vec3{{0.5f, -0.25f, 1.0f}}, Vec3{{0.0f, 0.0f, -1.0f}});
	const bool any = rc.intersectsAny(Vec3{{0.5f, -0.25f, 1.0f}}, Vec3{{0.0f, 0.0f, -1.0f}});
	if(!log.empty()) {
		std::printf("logged: %s\n", log.front().c_str());
		return 3;
	}
	std::printf("hit=%d dist=%.6f u=%.6f v=%.6f mesh=%lu local=%lu any=%d\n", h.hit ? 1 : 0, h.dist, h.u, h.v, h.meshId,
		h.localId, any ? 1 : 0);
	return (h.hit && any) ? 0 : 1;
}
