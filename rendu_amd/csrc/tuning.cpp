// Autotuner search and the process-wide tuning cache (tuning.hpp).
#include "tuning.hpp"
#include "capi_common.hpp"
#include "../../include/rdu.h"

#include <cstdio>
#include <map>
#include <mutex>
#include <vector>

namespace rdu {

namespace {

std::mutex gTuneLock;
std::map<TuneKey, TraceTuning> & tuneTable() {
	static std::map<TuneKey, TraceTuning> table;
	return table;
}

uint32_t floorLog2(size_t v) {
	uint32_t k = 0;
	while(v > 1) {
		v >>= 1;
		++k;
	}
	return k;
}

// Format 2: blocks_per_cu counts 256-thread workgroups of the persistent traversal (format 1
// files counted 128-thread ones and are rejected, so a stale grid size is never loaded).
constexpr const char * kHeader = "rdu-tune 2";

} // namespace

bool TuneKey::operator<(const TuneKey & o) const {
	if(triLog2 != o.triLog2)
		return triLog2 < o.triLog2;
	if(wideDepth != o.wideDepth)
		return wideDepth < o.wideDepth;
	if(meshLog2 != o.meshLog2)
		return meshLog2 < o.meshLog2;
	if(batchLog2 != o.batchLog2)
		return batchLog2 < o.batchLog2;
	return cls < o.cls;
}

TuneKey makeTuneKey(size_t triangles, uint32_t wideDepth, size_t meshes, size_t batch, uint32_t cls) {
	TuneKey k;
	k.triLog2 = floorLog2(triangles);
	k.wideDepth = wideDepth;
	k.meshLog2 = floorLog2(meshes);
	k.batchLog2 = floorLog2(batch);
	k.cls = cls;
	return k;
}

bool tuningValid(const TraceTuning & t) {
	return t.refillMin <= 64 && (t.chunk == 0 || (t.chunk >= 16 && t.chunk <= (1u << 20))) && t.blocksPerCu <= 64;
}

bool tuneCacheLookup(const TuneKey & key, TraceTuning & out) {
	std::lock_guard<std::mutex> g(gTuneLock);
	const auto it = tuneTable().find(key);
	if(it == tuneTable().end())
		return false;
	out = it->second;
	return true;
}

void tuneCacheStore(const TuneKey & key, const TraceTuning & t) {
	std::lock_guard<std::mutex> g(gTuneLock);
	tuneTable()[key] = t;
}

size_t tuneCacheSize() {
	std::lock_guard<std::mutex> g(gTuneLock);
	return tuneTable().size();
}

void tuneCacheClear() {
	std::lock_guard<std::mutex> g(gTuneLock);
	tuneTable().clear();
}

int tuneCacheSave(const char * path) {
	if(!path)
		return fail(RDU_ERR_INVALID, "null path");
	std::vector<std::pair<TuneKey, TraceTuning>> entries;
	{
		std::lock_guard<std::mutex> g(gTuneLock);
		entries.assign(tuneTable().begin(), tuneTable().end());
	}
	FILE * fp = std::fopen(path, "w");
	if(!fp)
		return fail(RDU_ERR_INVALID, std::string("cannot write ") + path);
	bool ok = std::fprintf(fp, "%s\n", kHeader) > 0;
	for(const auto & e : entries) {
		const TuneKey & k = e.first;
		const TraceTuning & t = e.second;
		ok = ok && std::fprintf(fp, "%u %u %u %u %u %u %u %u\n", k.triLog2, k.wideDepth, k.meshLog2, k.batchLog2, k.cls,
					   t.refillMin, t.chunk, t.blocksPerCu) > 0;
	}
	ok = (std::fclose(fp) == 0) && ok;
	return ok ? RDU_OK : fail(RDU_ERR_INVALID, std::string("write failed: ") + path);
}

int tuneCacheLoad(const char * path) {
	if(!path)
		return fail(RDU_ERR_INVALID, "null path");
	FILE * fp = std::fopen(path, "r");
	if(!fp)
		return fail(RDU_ERR_INVALID, std::string("cannot read ") + path);
	char line[256];
	std::vector<std::pair<TuneKey, TraceTuning>> entries;
	bool ok = std::fgets(line, sizeof(line), fp) != nullptr && std::string(line).rfind(kHeader, 0) == 0;
	while(ok && std::fgets(line, sizeof(line), fp)) {
		TuneKey k;
		TraceTuning t;
		char tail = 0;
		const int got = std::sscanf(line, "%u %u %u %u %u %u %u %u %c", &k.triLog2, &k.wideDepth, &k.meshLog2, &k.batchLog2,
			&k.cls, &t.refillMin, &t.chunk, &t.blocksPerCu, &tail);
		if(got == -1) // blank line
			continue;
		ok = got == 8 && tuningValid(t) && k.triLog2 < 64 && k.meshLog2 < 64 && k.batchLog2 < 64 && k.wideDepth < 256 &&
			 k.cls < 16;
		if(ok)
			entries.emplace_back(k, t);
	}
	std::fclose(fp);
	if(!ok)
		return fail(RDU_ERR_INVALID, std::string("malformed tuning file: ") + path);
	std::lock_guard<std::mutex> g(gTuneLock);
	for(const auto & e : entries)
		tuneTable()[e.first] = e.second;
	return RDU_OK;
}

int tuneSearch(const std::function<int(const TraceTuning &, float &)> & time, int occupancy, TuneResult & out) {
	out = TuneResult{};
	int st = RDU_OK;
	TraceTuning cur;
	float ms = 0.0f;
	if((st = time(cur, ms)))
		return st;
	out.timed = 1;
	out.defaultMs = ms;
	out.bestMs = ms;
	out.best = cur;
	auto tryOne = [&](const TraceTuning & t) -> int {
		if(t == out.best)
			return RDU_OK;
		float m = 0.0f;
		const int e = time(t, m);
		if(e)
			return e;
		++out.timed;
		if(m < out.bestMs) {
			out.bestMs = m;
			out.best = t;
		}
		return RDU_OK;
	};
	// Idle lanes a wavefront waits for before refilling.
	for(uint32_t r : {4u, 8u, 16u, 24u, 32u, 40u, 48u, 64u}) {
		TraceTuning t = out.best;
		t.refillMin = r;
		if((st = tryOne(t)))
			return st;
	}
	// Rays per interleaved run (the ray tile of a wavefront).
	for(uint32_t c : {32u, 64u, 128u, 256u, 512u}) {
		TraceTuning t = out.best;
		t.chunk = c;
		if((st = tryOne(t)))
			return st;
	}
	// Resident workgroups per CU: fractions of the occupancy limit (fewer waves contend less for
	// L2 and the stack LDS, more hide latency better).
	const int occ = occupancy < 1 ? 1 : occupancy;
	std::vector<uint32_t> grids;
	for(int num : {3, 2, 1}) {
		const int b = (occ * num) / 4;
		if(b >= 1 && b < occ && (grids.empty() || grids.back() != uint32_t(b)))
			grids.push_back(uint32_t(b));
	}
	for(uint32_t b : grids) {
		TraceTuning t = out.best;
		t.blocksPerCu = b;
		if((st = tryOne(t)))
			return st;
	}
	return RDU_OK;
}

} // namespace rdu
