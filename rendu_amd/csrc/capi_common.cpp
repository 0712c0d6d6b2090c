#include "capi_common.hpp"
#include "../../include/rdu.h"

#include <cstdlib>
#include <cstring>

namespace rdu {

namespace {
thread_local std::string tlError;

bool allocFaultInjected() {
	const char * v = std::getenv("RDU_FAIL_ALLOC");
	return v && std::strcmp(v, "1") == 0;
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
	if(ptr)
		(void)hipFree(ptr);
	ptr = nullptr;
	bytes = 0;
}

int DevBuffer::reserve(size_t n) {
	if(n <= bytes && ptr)
		return RDU_OK;
	release();
	if(allocFaultInjected())
		return fail(RDU_ERR_ALLOC, "device allocation failed (RDU_FAIL_ALLOC=1)");
	const size_t want = n < 256 ? 256 : n;
	const hipError_t e = hipMalloc(&ptr, want);
	if(e != hipSuccess) {
		ptr = nullptr;
		return hipCheck(e, "hipMalloc");
	}
	bytes = want;
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
