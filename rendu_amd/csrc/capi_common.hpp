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

// Stream-ordered allocation from the library's own memory pool of the current device (created on
// first use, release threshold at its maximum so freed blocks stay pooled for the next allocation
// instead of going back to the driver). The device's default pool, which the host application
// may use itself, is left as it is.
int poolAlloc(void ** ptr, size_t bytes, hipStream_t stream);

// Device buffer owned by the library, from its pool. Growth keeps the contents only when no
// growth happens.
//  - reserve(n): the old block is released as release() does; the new one is allocated in the null
//    stream's order and waited for, so any stream may use it.
//  - reserveOn(n, s): the old block is freed and the new one allocated in the order of stream `s`,
//    with no host wait. The caller guarantees that all earlier use of the old block is ordered
//    before `s`'s current position (finished, or on `s`, or waited for by `s`), and uses the new
//    block from work enqueued on `s` afterwards (or after waiting for `s`).
//  - release(): waits until no stream of the block's device can still use it, then frees it.
//  - releaseOn(s): frees in `s`'s order, no host wait (same guarantee as reserveOn).
struct DevBuffer {
	void * ptr = nullptr;
	size_t bytes = 0;
	int device = -1; // device of `ptr` (the current device when it was allocated)
	DevBuffer() = default;
	DevBuffer(const DevBuffer &) = delete;
	DevBuffer & operator=(const DevBuffer &) = delete;
	~DevBuffer();
	int reserve(size_t n);
	int reserveOn(size_t n, hipStream_t s);
	int upload(const void * host, size_t n);
	void release();
	void releaseOn(hipStream_t s);
};

} // namespace rdu
