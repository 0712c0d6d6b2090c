// Output step of the path-tracer app (src/apps/pathtracer/main.cpp:109-119 and
// Image::saveAsLDR, src/engine/resources/Image.cpp:126-158): exponential tonemap 1 - exp(-c) of the
// colour channels, sRGB approximation pow(v, 1/2.2), then min(255, max(0, 255 v)) truncated to a
// byte, alpha forced to 255 when ignored. One thread per pixel; HBM-bound elementwise pass.
#include "capi_common.hpp"
#include "../../include/rdu.h"

#include <cstdint>

namespace rdu {
namespace {

__global__ void ldrKernel(const float * __restrict__ in, size_t pixels, uint32_t comps, uint32_t flags,
	uint8_t * __restrict__ out) {
	const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if(p >= pixels)
		return;
	const bool tonemap = (flags & RDU_IMAGE_TONEMAP) != 0;
	const bool srgb = (flags & RDU_IMAGE_SRGB) != 0;
	const bool ignoreAlpha = (flags & RDU_IMAGE_IGNORE_ALPHA) != 0;
	for(uint32_t c = 0; c < comps; ++c) {
		float v = in[p * comps + c];
		const bool colour = c < 3 && comps >= 3;
		if(tonemap && colour)
			v = 1.0f - expf(-v);
		if(srgb && c != 3)
			v = powf(v, 1.0f / 2.2f);
		// std::max(0, v) then std::min(255, .): a NaN compares false and yields 0.
		float s = 255.0f * v;
		s = 0.0f < s ? s : 0.0f;
		s = s < 255.0f ? s : 255.0f;
		uint8_t q = uint8_t(s);
		if(c == 3 && ignoreAlpha)
			q = 255;
		out[p * comps + c] = q;
	}
}

} // namespace
} // namespace rdu

using namespace rdu;

extern "C" int rdu_image_to_ldr(const float * pixels, size_t width, size_t height, size_t components, uint8_t * out,
	uint32_t flags, void * stream) {
	clearError();
	if(components < 1 || components > 4)
		return fail(RDU_ERR_INVALID, "components must be 1 to 4");
	const size_t n = width * height;
	if(n == 0)
		return RDU_OK;
	if(!pixels || !out)
		return fail(RDU_ERR_INVALID, "null image");
	hipStream_t s = static_cast<hipStream_t>(stream);
	const bool device = (flags & RDU_DEVICE_POINTERS) != 0;
	const float * dIn = pixels;
	uint8_t * dOut = out;
	DevBuffer inBuf, outBuf;
	int st;
	if(!device) {
		if((st = inBuf.reserve(n * components * 4)) || (st = outBuf.reserve(n * components)))
			return st;
		if((st = hipCheck(hipMemcpyAsync(inBuf.ptr, pixels, n * components * 4, hipMemcpyHostToDevice, s), "copy")))
			return st;
		dIn = static_cast<const float *>(inBuf.ptr);
		dOut = static_cast<uint8_t *>(outBuf.ptr);
	}
	const unsigned blocks = unsigned((n + 255) / 256);
	hipLaunchKernelGGL(ldrKernel, dim3(blocks), dim3(256), 0, s, dIn, n, uint32_t(components), flags, dOut);
	if((st = hipCheck(hipGetLastError(), "tonemap launch")))
		return st;
	if(!device) {
		if((st = hipCheck(hipMemcpyAsync(out, dOut, n * components, hipMemcpyDeviceToHost, s), "copy")) ||
			(st = hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize")))
			return st;
	}
	return RDU_OK;
}
