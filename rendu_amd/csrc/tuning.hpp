// Autotuning of the persistent traversal's launch parameters (no counterpart in Rendu, whose
// traversal has no launch parameters): a coordinate-descent search over measured candidates and
// a process-wide cache of the winners keyed by scene statistics, so that a scene of the same
// class (size, tree depth, mesh count, batch size, kind of query) starts tuned.
#pragma once
#include "trace.hpp"

#include <cstddef>
#include <cstdint>
#include <functional>

namespace rdu {

// Classes of batch; bits of TuneKey::cls.
enum : uint32_t { kTuneIncoherent = 1, kTuneAny = 2, kTuneGpuTree = 4, kTunePathPass = 8, kTuneInstanced = 16 };

struct TuneKey {
	uint32_t triLog2 = 0;   // floor(log2(triangles))
	uint32_t wideDepth = 0; // levels of the 4-wide tree
	uint32_t meshLog2 = 0;  // floor(log2(meshes))
	uint32_t batchLog2 = 0; // floor(log2(rays or paths per launch))
	uint32_t cls = 0;       // kTune* bits
	bool operator<(const TuneKey & o) const;
};

TuneKey makeTuneKey(size_t triangles, uint32_t wideDepth, size_t meshes, size_t batch, uint32_t cls);

// Process-wide cache (thread-safe).
bool tuneCacheLookup(const TuneKey & key, TraceTuning & out);
void tuneCacheStore(const TuneKey & key, const TraceTuning & t);
size_t tuneCacheSize();
void tuneCacheClear();
// Text file: a header line, then one entry per line
// "triLog2 wideDepth meshLog2 batchLog2 cls refillMin chunk blocksPerCu". Load merges into the
// cache (file entries win) and rejects a malformed file without changing the cache. Return RDU codes.
int tuneCacheSave(const char * path);
int tuneCacheLoad(const char * path);
// Whether each field is 0 (default) or within what the launcher accepts.
bool tuningValid(const TraceTuning & t);

// Coordinate descent: the default, then each refill threshold, then each chunk with the best
// threshold, then each grid size with the best pair. `time` runs one candidate and returns its
// mean milliseconds (or a nonzero RDU code, which stops the search). `occupancy` is the
// kernel's resident workgroups per CU (the grid-size candidates are fractions of it).
struct TuneResult {
	TraceTuning best;
	float bestMs = 0.0f, defaultMs = 0.0f;
	uint32_t timed = 0;
};
int tuneSearch(const std::function<int(const TraceTuning &, float &)> & time, int occupancy, TuneResult & out);

} // namespace rdu
