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

// The library's pool per device: hipMallocFromPoolAsync / hipFreeAsync with a release threshold at
// its maximum, so freed blocks are reused without driver calls. Every driver allocation is a kernel
// call that serialises with the other threads' driver calls (and has been seen to take over a
// second), so scenes built and dropped one after another, or from several threads, reuse memory
// without them. Pools live until the process ends.
hipMemPool_t pools[64] = {};

int libraryPool(int device, hipMemPool_t * out) {
	static std::mutex m;
	if(device < 0 || device >= 64)
		return fail(RDU_ERR_UNSUPPORTED, "devices 0..63 only");
	std::lock_guard<std::mutex> g(m);
	if(!pools[device]) {
		hipMemPoolProps props{};
		props.allocType = hipMemAllocationTypePinned;
		props.handleTypes = hipMemHandleTypeNone;
		props.location.type = hipMemLocationTypeDevice;
		props.location.id = device;
		hipMemPool_t p = nullptr;
		int st = hipCheck(hipMemPoolCreate(&p, &props), "hipMemPoolCreate");
		if(st)
			return st;
		uint64_t keep = UINT64_MAX;
		if((st = hipCheck(hipMemPoolSetAttribute(p, hipMemPoolAttrReleaseThreshold, &keep), "hipMemPoolSetAttribute"))) {
			(void)hipMemPoolDestroy(p);
			return st;
		}
		pools[device] = p;
	}
	*out = pools[device];
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

int poolAlloc(void ** ptr, size_t bytes, hipStream_t stream) {
	*ptr = nullptr;
	if(allocFaultInjected())
		return fail(RDU_ERR_ALLOC, "device allocation failed (RDU_FAIL_ALLOC=1)");
	int dev = 0, st;
	hipMemPool_t pool = nullptr;
	if((st = hipCheck(hipGetDevice(&dev), "hipGetDevice")) || (st = libraryPool(dev, &pool)))
		return st;
	const hipError_t e = hipMallocFromPoolAsync(ptr, bytes ? bytes : 256, pool, stream);
	if(e != hipSuccess) {
		*ptr = nullptr;
		return hipCheck(e, "hipMallocFromPoolAsync");
	}
	return RDU_OK;
}

DevBuffer::~DevBuffer() { release(); }

void DevBuffer::release() {
	if(ptr) {
		// As hipFree would: wait until no stream of the block's device can still use it, then
		// return it to the pool.
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

void DevBuffer::releaseOn(hipStream_t s) {
	if(ptr)
		(void)hipFreeAsync(ptr, s);
	ptr = nullptr;
	bytes = 0;
	device = -1;
}

int DevBuffer::reserve(size_t n) {
	if(n <= bytes && ptr)
		return RDU_OK;
	release();
	int dev = 0;
	int st = hipCheck(hipGetDevice(&dev), "hipGetDevice");
	if(st)
		return st;
	const size_t want = n < 256 ? 256 : n;
	// Allocated in the null stream's order, which is complete once that stream has drained: the
	// block is then usable from every stream, like a hipMalloc result.
	if((st = poolAlloc(&ptr, want, nullptr)))
		return st;
	const hipError_t e = hipStreamSynchronize(nullptr);
	if(e != hipSuccess) {
		(void)hipFreeAsync(ptr, nullptr);
		ptr = nullptr;
		return hipCheck(e, "hipStreamSynchronize");
	}
	bytes = want;
	device = dev;
	return RDU_OK;
}

int DevBuffer::reserveOn(size_t n, hipStream_t s) {
	if(n <= bytes && ptr)
		return RDU_OK;
	releaseOn(s);
	int dev = 0;
	int st = hipCheck(hipGetDevice(&dev), "hipGetDevice");
	if(st)
		return st;
	const size_t want = n < 256 ? 256 : n;
	if((st = poolAlloc(&ptr, want, s)))
		return st;
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
