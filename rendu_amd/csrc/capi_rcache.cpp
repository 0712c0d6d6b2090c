// C ABI of the learned radiance cache (include/rdu.h, rdu_rcache_*). No Rendu counterpart: the
// cache is an extension that terminates paths early into a trained estimate of outgoing radiance.
#include "capi_common.hpp"
#include "rcache.hpp"
#include "../../include/rdu.h"

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

struct rdu_rcache {
	rdu::RcGrid grid{};
	rdu::RcParams p;
	rdu::RcTrainBuffers tb;
	rdu::RcAdam adam;
	rdu::DevBuffer w32, w16, gw, mw, vw, t32, t16, gt, mt, vt, tx, tg, tloss, stagePos, stageTgt, stageOut;
	rdu::DevBuffer rKey, rKeyAlt, rUniq, rCount, rVal, rValAlt, rUniqVal, rTemp; // table-gradient records
	rdu::DevBuffer cells; // per coarse cell trust statistics of the path tracer's cache mode
	int device = -1;
};

using namespace rdu;

namespace {

int ensureTrain(rdu_rcache * c, uint32_t n) {
	const uint32_t cap = (n + 1023u) & ~1023u;
	if(cap <= c->tb.capacity)
		return RDU_OK;
	int st;
	const size_t recs = size_t(cap) * kRcRecordsPerSample;
	const size_t tempBytes = rcReduceTempBytes(uint32_t(recs));
	// capacity and the pointers describe the buffers only after every one of them has grown: a
	// growth that fails half-way leaves capacity 0, so the next call grows them all again.
	c->tb.capacity = 0;
	if((st = c->tx.reserve(size_t(kRcIn + 5 * kRcWidth) * cap * 2)) ||
		(st = c->tg.reserve(size_t(5 * kRcWidth + kRcOutPad) * cap * 2)) ||
		(st = c->tloss.reserve(size_t(cap) * 4)) || (st = c->rKey.reserve(recs * 4)) ||
		(st = c->rKeyAlt.reserve(recs * 4)) || (st = c->rUniq.reserve(recs * 4)) || (st = c->rCount.reserve(4)) ||
		(st = c->rVal.reserve(recs * sizeof(RcGradPair))) || (st = c->rValAlt.reserve(recs * sizeof(RcGradPair))) ||
		(st = c->rUniqVal.reserve(recs * sizeof(RcGradPair))) || (st = c->rTemp.reserve(tempBytes)))
		return st;
	c->tb.x = static_cast<uint16_t *>(c->tx.ptr);
	c->tb.g = static_cast<uint16_t *>(c->tg.ptr);
	c->tb.loss = static_cast<float *>(c->tloss.ptr);
	c->tb.recKey = static_cast<uint32_t *>(c->rKey.ptr);
	c->tb.recKeyAlt = static_cast<uint32_t *>(c->rKeyAlt.ptr);
	c->tb.uniqKey = static_cast<uint32_t *>(c->rUniq.ptr);
	c->tb.uniqCount = static_cast<uint32_t *>(c->rCount.ptr);
	c->tb.recVal = static_cast<RcGradPair *>(c->rVal.ptr);
	c->tb.recValAlt = static_cast<RcGradPair *>(c->rValAlt.ptr);
	c->tb.uniqVal = static_cast<RcGradPair *>(c->rUniqVal.ptr);
	c->tb.temp = c->rTemp.ptr;
	c->tb.tempBytes = tempBytes;
	c->tb.capacity = cap;
	return RDU_OK;
}

// Copies host inputs (`bytes` of them) when the caller gives host pointers.
int stage(DevBuffer & b, const float * src, size_t bytes, bool device, const float4 ** out, hipStream_t s) {
	if(device) {
		*out = reinterpret_cast<const float4 *>(src);
		return RDU_OK;
	}
	int st = b.reserve(bytes);
	if(st)
		return st;
	if((st = hipCheck(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D")))
		return st;
	*out = static_cast<const float4 *>(b.ptr);
	return RDU_OK;
}

} // namespace

extern "C" {

int rdu_rcache_create(rdu_rcache ** out, const float scene_min[3], const float scene_max[3], uint32_t seed) {
	clearError();
	if(!out || !scene_min || !scene_max)
		return fail(RDU_ERR_INVALID, "null argument");
	rdu_rcache * c = new(std::nothrow) rdu_rcache();
	if(!c)
		return fail(RDU_ERR_ALLOC, "host allocation failed");
	int st = currentDevice(&c->device);
	if(st) {
		delete c;
		return st;
	}
	for(int a = 0; a < 3; ++a) {
		const float ext = scene_max[a] - scene_min[a];
		c->grid.lo[a] = scene_min[a];
		c->grid.inv[a] = ext > 0.0f ? 1.0f / ext : 1.0f;
	}
	// Resolutions from 16 to 2048 in geometric steps; dense indexing where the grid fits the table.
	const double b = std::exp((std::log(2048.0) - std::log(16.0)) / double(kRcLevels - 1));
	for(int l = 0; l < kRcLevels; ++l) {
		const int res = int(std::floor(16.0 * std::pow(b, double(l))));
		c->grid.res[l] = res;
		const double cells = double(res + 1) * double(res + 1) * double(res + 1);
		c->grid.dense[l] = cells <= double(kRcTable) ? 1 : 0;
	}
	// Initialisation: tables U(-1e-4, 1e-4); MLP weights U(-sqrt(6/fan_in), sqrt(6/fan_in)); padded
	// output rows zero.
	std::mt19937 gen(seed);
	std::uniform_real_distribution<float> ut(-1e-4f, 1e-4f);
	std::vector<float> tab(kRcTableParams);
	for(auto & v : tab)
		v = ut(gen);
	std::vector<float> w(kRcMlpParams, 0.0f);
	size_t off = 0;
	const int ins[6] = {kRcIn, kRcWidth, kRcWidth, kRcWidth, kRcWidth, kRcWidth};
	const int outs[6] = {kRcWidth, kRcWidth, kRcWidth, kRcWidth, kRcWidth, kRcOutPad};
	for(int l = 0; l < 6; ++l) {
		const float a = std::sqrt(6.0f / float(ins[l]));
		std::uniform_real_distribution<float> uw(-a, a);
		for(int r = 0; r < outs[l]; ++r)
			for(int k = 0; k < ins[l]; ++k)
				w[off + size_t(r) * ins[l] + k] = (l == 5 && r >= kRcOut) ? 0.0f : uw(gen);
		off += size_t(outs[l]) * ins[l];
	}
	if((st = c->w32.upload(w.data(), w.size() * 4)) || (st = c->w16.reserve(w.size() * 2)) ||
		(st = c->gw.reserve(w.size() * 4)) || (st = c->mw.reserve(w.size() * 4)) || (st = c->vw.reserve(w.size() * 4)) ||
		(st = c->t32.upload(tab.data(), tab.size() * 4)) || (st = c->t16.reserve(tab.size() * 2)) ||
		(st = c->gt.reserve(tab.size() * 4)) || (st = c->mt.reserve(tab.size() * 4)) || (st = c->vt.reserve(tab.size() * 4)) ||
		(st = c->cells.reserve(size_t(kRcCells) * 16))) {
		delete c;
		return st;
	}
	for(DevBuffer * b : {&c->gw, &c->mw, &c->vw, &c->gt, &c->mt, &c->vt, &c->cells})
		if((st = hipCheck(hipMemset(b->ptr, 0, b->bytes), "hipMemset"))) {
			delete c;
			return st;
		}
	c->p.w32 = static_cast<float *>(c->w32.ptr);
	c->p.w16 = static_cast<uint16_t *>(c->w16.ptr);
	c->p.gw = static_cast<float *>(c->gw.ptr);
	c->p.mw = static_cast<float *>(c->mw.ptr);
	c->p.vw = static_cast<float *>(c->vw.ptr);
	c->p.t32 = static_cast<float *>(c->t32.ptr);
	c->p.t16 = static_cast<uint16_t *>(c->t16.ptr);
	c->p.gt = static_cast<float *>(c->gt.ptr);
	c->p.mt = static_cast<float *>(c->mt.ptr);
	c->p.vt = static_cast<float *>(c->vt.ptr);
	if((st = hipCheck(hipError_t(rcSyncCopies(c->p, nullptr)), "cache copies")) ||
		(st = hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize"))) {
		delete c;
		return st;
	}
	*out = c;
	return RDU_OK;
}

void rdu_rcache_destroy(rdu_rcache * c) { delete c; }

int rdu_rcache_set_learning_rate(rdu_rcache * c, float lr) {
	clearError();
	if(!c || !(lr > 0.0f))
		return fail(RDU_ERR_INVALID, "bad cache or learning rate");
	c->adam.lr = lr;
	return RDU_OK;
}

int rdu_rcache_infer(rdu_rcache * c, const float * queries, size_t n, float * radiance, uint32_t flags, void * stream) {
	clearError();
	if(!c || (n && (!queries || !radiance)))
		return fail(RDU_ERR_INVALID, "null argument");
	if(n == 0)
		return RDU_OK;
	if(n > 0xffffffffull)
		return fail(RDU_ERR_UNSUPPORTED, "more than 2^32 queries");
	int st = setDevice(c->device);
	if(st)
		return st;
	hipStream_t s = static_cast<hipStream_t>(stream);
	const bool dev = (flags & RDU_DEVICE_POINTERS) != 0;
	const float4 * dPos;
	if((st = stage(c->stagePos, queries, n * 32, dev, &dPos, s)))
		return st;
	float4 * dOut = reinterpret_cast<float4 *>(radiance);
	if(!dev) {
		if((st = c->stageOut.reserve(n * 16)))
			return st;
		dOut = static_cast<float4 *>(c->stageOut.ptr);
	}
	if((st = hipCheck(hipError_t(rcInfer(c->grid, c->p, dPos, nullptr, uint32_t(n), dOut, s)), "cache inference")))
		return st;
	if(!dev) {
		if((st = hipCheck(hipMemcpyAsync(radiance, dOut, n * 16, hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H")) ||
			(st = hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize")))
			return st;
	}
	return RDU_OK;
}

int rdu_rcache_train(rdu_rcache * c, const float * queries, const float * targets, size_t n, uint32_t steps,
	uint32_t flags, void * stream, float * mean_loss) {
	clearError();
	if(!c || !queries || !targets || n == 0)
		return fail(RDU_ERR_INVALID, "null argument or empty batch");
	if(n > (size_t(1) << 26))
		return fail(RDU_ERR_UNSUPPORTED, "training batch above 2^26 samples");
	int st = setDevice(c->device);
	if(st)
		return st;
	hipStream_t s = static_cast<hipStream_t>(stream);
	const bool dev = (flags & RDU_DEVICE_POINTERS) != 0;
	const float4 *dPos, *dTgt;
	if((st = stage(c->stagePos, queries, n * 32, dev, &dPos, s)) || (st = stage(c->stageTgt, targets, n * 16, dev, &dTgt, s)) ||
		(st = ensureTrain(c, uint32_t(n))))
		return st;
	for(uint32_t k = 0; k < steps; ++k)
		if((st = hipCheck(hipError_t(rcTrainStep(c->grid, c->p, c->tb, c->adam, dPos, dTgt, uint32_t(n), s)), "cache training")))
			return st;
	if(mean_loss) {
		std::vector<float> l(n);
		if((st = hipCheck(hipMemcpyAsync(l.data(), c->tb.loss, n * 4, hipMemcpyDeviceToHost, s), "loss copy")) ||
			(st = hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize")))
			return st;
		double sum = 0.0;
		for(float v : l)
			sum += v;
		*mean_loss = float(sum / double(n));
	}
	return RDU_OK;
}

int rdu_rcache_gradients(rdu_rcache * c, const float * queries, const float * targets, size_t n, float * mlp_grad,
	float * table_grad) {
	clearError();
	if(!c || !queries || !targets || n == 0 || !mlp_grad)
		return fail(RDU_ERR_INVALID, "null argument or empty batch");
	int st = setDevice(c->device);
	if(st)
		return st;
	const float4 *dPos, *dTgt;
	if((st = stage(c->stagePos, queries, n * 32, false, &dPos, nullptr)) ||
		(st = stage(c->stageTgt, targets, n * 16, false, &dTgt, nullptr)) || (st = ensureTrain(c, uint32_t(n))))
		return st;
	if((st = hipCheck(hipMemset(c->gw.ptr, 0, c->gw.bytes), "hipMemset")) ||
		(st = hipCheck(hipMemset(c->gt.ptr, 0, c->gt.bytes), "hipMemset")) ||
		(st = hipCheck(hipError_t(rcGradients(c->grid, c->p, c->tb, dPos, dTgt, uint32_t(n), nullptr)), "gradients")) ||
		(st = hipCheck(hipMemcpy(mlp_grad, c->gw.ptr, size_t(kRcMlpParams) * 4, hipMemcpyDeviceToHost), "copy")))
		return st;
	if(table_grad && (st = hipCheck(hipMemcpy(table_grad, c->gt.ptr, kRcTableParams * 4, hipMemcpyDeviceToHost), "copy")))
		return st;
	// Leave no stale gradient behind for the next training step.
	if((st = hipCheck(hipMemset(c->gw.ptr, 0, c->gw.bytes), "hipMemset")) ||
		(st = hipCheck(hipMemset(c->gt.ptr, 0, c->gt.bytes), "hipMemset")))
		return st;
	return RDU_OK;
}

int rdu_rcache_get_params(const rdu_rcache * c, float * mlp, float * tables, int32_t * level_res) {
	clearError();
	if(!c)
		return fail(RDU_ERR_INVALID, "null cache");
	int st = setDevice(c->device);
	if(st)
		return st;
	if(mlp && (st = hipCheck(hipMemcpy(mlp, c->w32.ptr, size_t(kRcMlpParams) * 4, hipMemcpyDeviceToHost), "copy")))
		return st;
	if(tables && (st = hipCheck(hipMemcpy(tables, c->t32.ptr, kRcTableParams * 4, hipMemcpyDeviceToHost), "copy")))
		return st;
	if(level_res)
		for(int l = 0; l < kRcLevels; ++l)
			level_res[l] = c->grid.res[l];
	return RDU_OK;
}

int rdu_rcache_set_params(rdu_rcache * c, const float * mlp, const float * tables) {
	clearError();
	if(!c)
		return fail(RDU_ERR_INVALID, "null cache");
	int st = setDevice(c->device);
	if(st)
		return st;
	if(mlp && (st = hipCheck(hipMemcpy(c->w32.ptr, mlp, size_t(kRcMlpParams) * 4, hipMemcpyHostToDevice), "copy")))
		return st;
	if(tables && (st = hipCheck(hipMemcpy(c->t32.ptr, tables, kRcTableParams * 4, hipMemcpyHostToDevice), "copy")))
		return st;
	if((st = hipCheck(hipError_t(rcSyncCopies(c->p, nullptr)), "copies")) ||
		(st = hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize")))
		return st;
	return RDU_OK;
}

} // extern "C"

namespace rdu {
// Internal access for the path tracer.
const RcGrid * rcacheGrid(const rdu_rcache * c) { return &c->grid; }
RcParams * rcacheParams(rdu_rcache * c) { return &c->p; }
RcTrainBuffers * rcacheTrain(rdu_rcache * c, uint32_t n) { return ensureTrain(c, n) ? nullptr : &c->tb; }
RcAdam * rcacheAdam(rdu_rcache * c) { return &c->adam; }
float4 * rcacheCells(rdu_rcache * c) { return static_cast<float4 *>(c->cells.ptr); }
int rcacheDevice(const rdu_rcache * c) { return c->device; }
} // namespace rdu
