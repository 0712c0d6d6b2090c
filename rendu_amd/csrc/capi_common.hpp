// Shared host plumbing of the C ABI: thread-local error messages, HIP status checks and a
// growable device buffer. Device allocation can be made to fail on purpose by setting
// RDU_FAIL_ALLOC=1 in the environment (error-path tests).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <string>

namespace rdu {

int fail(int code, const std::string & msg);
void clearError();
int hipCheck(hipError_t e, const char * what);
int setDevice(int device);
int currentDevice(int * device);

// Device buffer owned by the library. reserve() keeps the contents only when no growth happens.
struct DevBuffer {
	void * ptr = nullptr;
	size_t bytes = 0;
	int device = -1; // device of `ptr` (the current device when it was allocated)
	DevBuffer() = default;
	DevBuffer(const DevBuffer &) = delete;
	DevBuffer & operator=(const DevBuffer &) = delete;
	~DevBuffer();
	int reserve(size_t n);
	int upload(const void * host, size_t n);
	void release();
};

} // namespace rdu
