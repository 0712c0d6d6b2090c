#include "capi_common.hpp"
#include "../../include/rdu.h"

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>

namespace rdu {

namespace {
thread_local std::string tlError;

bool allocFaultInjected() {
	const char * v = std::getenv("RDU_FAIL_ALLOC");
	return v && std::strcmp(v, "1") == 0;
}

// Library buffers come from the current device's default stream-ordered pool, whose release
// threshold is raised once per device so that freed blocks stay in the pool for the next
// allocation instead of going back to the driver. Every driver allocation is a kernel call that
// serialises with the other threads' driver calls (and has been seen to take over a second), so
// scenes built and dropped one after another, or from several threads, reuse memory without them.
int poolReady(int device) {
	static std::mutex m;
	static uint64_t ready = 0;
	if(device < 0 || device >= 64)
		return RDU_OK; // no pool set-up: plain pool behaviour
	std::lock_guard<std::mutex> g(m);
	if((ready >> device) & 1u)
		return RDU_OK;
	hipMemPool_t pool = nullptr;
	int st = hipCheck(hipDeviceGetDefaultMemPool(&pool, device), "hipDeviceGetDefaultMemPool");
	if(st)
		return st;
	uint64_t keep = UINT64_MAX;
	if((st = hipCheck(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep), "hipMemPoolSetAttribute")))
		return st;
	ready |= uint64_t(1) << device;
	return RDU_OK;
}
} // namespace

int fail(int code, const std::string & msg) {
	tlError = msg;
	return code;
}

void clearError() { tlError.clear(); }

int hipCheck(hipError_t e, const char * what) {
	if(e == hipSuccess)
		return RDU_OK;
	(void)hipGetLastError(); // clear the sticky launch error
	return fail(e == hipErrorOutOfMemory ? RDU_ERR_ALLOC : RDU_ERR_HIP,
		std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")");
}

int setDevice(int device) { return hipCheck(hipSetDevice(device), "hipSetDevice"); }

int currentDevice(int * device) {
	int count = 0;
	const hipError_t e = hipGetDeviceCount(&count);
	if(e != hipSuccess || count == 0) {
		(void)hipGetLastError();
		return fail(RDU_ERR_HIP, "no HIP device available: rendu_amd runs its queries on the GPU only");
	}
	return hipCheck(hipGetDevice(device), "hipGetDevice");
}

DevBuffer::~DevBuffer() { release(); }

void DevBuffer::release() {
	if(ptr) {
		// As hipFree would: wait until no stream of the block's device can still use it, then
		// return it (to that device's pool, where the next allocation may take it at once).
		int current = -1;
		(void)hipGetDevice(&current);
		if(device >= 0 && device != current)
			(void)hipSetDevice(device);
		(void)hipDeviceSynchronize();
		(void)hipFreeAsync(ptr, nullptr);
		(void)hipStreamSynchronize(nullptr);
		if(device >= 0 && device != current && current >= 0)
			(void)hipSetDevice(current);
	}
	ptr = nullptr;
	bytes = 0;
	device = -1;
}

int DevBuffer::reserve(size_t n) {
	if(n <= bytes && ptr)
		return RDU_OK;
	release();
	if(allocFaultInjected())
		return fail(RDU_ERR_ALLOC, "device allocation failed (RDU_FAIL_ALLOC=1)");
	int dev = 0, st;
	if((st = hipCheck(hipGetDevice(&dev), "hipGetDevice")) || (st = poolReady(dev)))
		return st;
	const size_t want = n < 256 ? 256 : n;
	// Allocated in the null stream's order, which is complete once that stream has drained: the
	// block is then usable from every stream, like a hipMalloc result.
	hipError_t e = hipMallocAsync(&ptr, want, nullptr);
	if(e == hipSuccess)
		e = hipStreamSynchronize(nullptr);
	if(e != hipSuccess) {
		if(ptr)
			(void)hipFreeAsync(ptr, nullptr);
		ptr = nullptr;
		return hipCheck(e, "hipMallocAsync");
	}
	bytes = want;
	device = dev;
	return RDU_OK;
}

int DevBuffer::upload(const void * host, size_t n) {
	int st = reserve(n);
	if(st)
		return st;
	if(n == 0)
		return RDU_OK;
	return hipCheck(hipMemcpy(ptr, host, n, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

} // namespace rdu

extern "C" {

const char * rdu_last_error(void) { return rdu::tlError.c_str(); }

const char * rdu_version(void) { return "rendu_amd 0.1 (gfx950)"; }

} // extern "C"
