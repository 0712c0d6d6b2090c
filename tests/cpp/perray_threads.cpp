// A Rendu-style per-ray caller: T host threads split a list of rays into contiguous spans, as
// System::forParallel splits image rows (src/engine/system/System.hpp:72-104), and each thread
// calls the per-ray entry points one ray at a time, as PathTracer::render's workers call
// Raycaster::intersects (PathTracer.cpp:186, 204).
//
// Usage: perray_threads <scene.bin> <hits.bin> <threads> [device]
//   scene.bin: uint32 n_vertices, n_indices, n_rays; float xyz[3 n_vertices]; uint32 idx[n_indices];
//              rdu_ray rays[n_rays]
//   hits.bin:  rdu_hit closest[n_rays], then uint8 any[n_rays] (written)
//   device:    1 = one-ray GPU launches (rdu_raycaster_set_per_ray_device), default host walk
// Prints one line: "closest_mrays=<rate> any_mrays=<rate> threads=<T> rays=<n>".
#include "rdu.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

static bool readAll(FILE * f, void * p, size_t bytes) { return std::fread(p, 1, bytes, f) == bytes; }

int main(int argc, char ** argv) {
	if(argc < 4) {
		std::fprintf(stderr, "usage: %s scene.bin hits.bin threads [device]\n", argv[0]);
		return 2;
	}
	FILE * f = std::fopen(argv[1], "rb");
	if(!f)
		return 2;
	uint32_t head[3];
	if(!readAll(f, head, sizeof(head)))
		return 2;
	std::vector<float> xyz(size_t(head[0]) * 3);
	std::vector<uint32_t> idx(head[1]);
	std::vector<rdu_ray> rays(head[2]);
	if(!readAll(f, xyz.data(), xyz.size() * 4) || !readAll(f, idx.data(), idx.size() * 4) ||
		!readAll(f, rays.data(), rays.size() * sizeof(rdu_ray)))
		return 2;
	std::fclose(f);
	const int threads = std::atoi(argv[3]);
	const bool device = argc > 4 && std::atoi(argv[4]) != 0;

	rdu_raycaster * rc = nullptr;
	if(rdu_raycaster_create(&rc) || rdu_raycaster_add_mesh(rc, xyz.data(), head[0], idx.data(), idx.size(), nullptr) ||
		rdu_raycaster_update_hierarchy(rc) || rdu_raycaster_set_per_ray_device(rc, device ? 1 : 0)) {
		std::printf("error: %s\n", rdu_last_error());
		return 3;
	}
	const size_t n = rays.size();
	std::vector<rdu_hit> hits(n);
	std::vector<uint8_t> any(n);
	std::vector<int> status(size_t(threads), 0);
	auto run = [&](bool closest) {
		std::vector<std::thread> pool;
		const size_t span = (n + size_t(threads) - 1) / size_t(threads);
		for(int t = 0; t < threads; ++t) {
			pool.emplace_back([&, t, closest]() {
				const size_t a = size_t(t) * span, b = (a + span < n) ? a + span : n;
				for(size_t i = a; i < b; ++i) {
					const rdu_ray & r = rays[i];
					const float o[3] = {r.ox, r.oy, r.oz}, d[3] = {r.dx, r.dy, r.dz};
					int st;
					if(closest) {
						st = rdu_raycaster_intersect1(rc, o, d, r.tmin, r.tmax, &hits[i]);
					} else {
						int occ = 0;
						st = rdu_raycaster_intersect_any1(rc, o, d, r.tmin, r.tmax, &occ);
						any[i] = uint8_t(occ);
					}
					if(st)
						status[size_t(t)] = st;
				}
			});
		}
		for(auto & th : pool)
			th.join();
	};
	const auto t0 = std::chrono::steady_clock::now();
	run(true);
	const auto t1 = std::chrono::steady_clock::now();
	run(false);
	const auto t2 = std::chrono::steady_clock::now();
	for(int s : status)
		if(s) {
			std::printf("error: %s\n", rdu_last_error());
			return 3;
		}
	rdu_raycaster_destroy(rc);
	FILE * o = std::fopen(argv[2], "wb");
	if(!o || std::fwrite(hits.data(), sizeof(rdu_hit), n, o) != n || std::fwrite(any.data(), 1, n, o) != n)
		return 2;
	std::fclose(o);
	const double sc = std::chrono::duration<double>(t1 - t0).count(), sa = std::chrono::duration<double>(t2 - t1).count();
	std::printf("closest_mrays=%.4f any_mrays=%.4f threads=%d rays=%zu\n", double(n) / sc / 1e6, double(n) / sa / 1e6,
		threads, n);
	return 0;
}
