// Streamed host transfers.  Large host arrays move through pinned chunk buffers: a chunk is copied
// into a pinned buffer by the host threads, uploaded on one stream, processed on the library
// stream, downloaded on a third, and copied out by the host threads, while the neighbouring
// chunks are in the other stages.  Pageable hipMemcpy of a multi-GB array runs at a fraction of
// the link rate; this path keeps both directions of the link busy at once.
#include "common.hpp"
#include "../gli_device.h"

namespace gli {

namespace {

constexpr int kSlots = 4;
constexpr size_t kChunkBytes = (size_t)32 << 20;
constexpr size_t kRingMinBytes = (size_t)4 << 20;   // smaller copies go directly, without the ring

struct XferRing {
	char* pin_in[kSlots] = {};
	char* pin_out[kSlots] = {};
	char* dev_in[kSlots] = {};
	char* dev_out[kSlots] = {};
	hipStream_t up = nullptr, down = nullptr;   // default priority
	hipEvent_t ev_up[kSlots] = {}, ev_work[kSlots] = {}, ev_done[kSlots] = {};
};

XferRing* g_ring = nullptr;   // created on first use (context lock held), kept until the process ends

XferRing& ring()
{
	if (g_ring == nullptr) {
		XferRing* r = new XferRing();
		for (int b = 0; b < kSlots; ++b) {
			GLI_HIP(hipHostMalloc((void**)&r->pin_in[b], kChunkBytes, hipHostMallocDefault));
			GLI_HIP(hipHostMalloc((void**)&r->pin_out[b], kChunkBytes, hipHostMallocDefault));
			GLI_HIP(hipMalloc((void**)&r->dev_in[b], kChunkBytes));
			GLI_HIP(hipMalloc((void**)&r->dev_out[b], kChunkBytes));
			GLI_HIP(hipEventCreateWithFlags(&r->ev_up[b], hipEventDisableTiming));
			GLI_HIP(hipEventCreateWithFlags(&r->ev_work[b], hipEventDisableTiming));
			GLI_HIP(hipEventCreateWithFlags(&r->ev_done[b], hipEventDisableTiming));
		}
		GLI_HIP(hipStreamCreateWithFlags(&r->up, hipStreamNonBlocking));
		GLI_HIP(hipStreamCreateWithFlags(&r->down, hipStreamNonBlocking));
		g_ring = r;
	}
	return *g_ring;
}

}  // namespace

size_t stream_chunk_bytes()
{
	return kChunkBytes;
}

void stream_chunks(const void* host_in, size_t in_bytes, void* host_out, size_t out_bytes, int64_t num, void* in_copy,
                   const std::function<void(const void*, void*, int64_t)>& work)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	if (num <= 0) { return; }
	GLI_REQUIRE(in_bytes > 0 && in_bytes <= kChunkBytes && out_bytes <= kChunkBytes, "element too large for the transfer ring");
	XferRing& r = ring();
	const int64_t E = (int64_t)(kChunkBytes / std::max(in_bytes, out_bytes));
	const int64_t nc = ceil_div(num, E);
	auto count = [&](int64_t ch) { return std::min(E, num - ch * E); };
	for (int64_t ch = 0; ch < nc + kSlots; ++ch) {
		if (ch >= kSlots) {
			// slot of chunk d = ch - kSlots: wait until it is through every stage, then drain it
			const int64_t d = ch - kSlots;
			const int b = (int)(d % kSlots);
			GLI_HIP(hipEventSynchronize(r.ev_done[b]));
			if (host_out != nullptr) {
				gli_host_copy((char*)host_out + (size_t)d * E * out_bytes, nullptr, r.pin_out[b], (size_t)count(d) * out_bytes);
			}
		}
		if (ch < nc) {
			const int b = (int)(ch % kSlots);
			const int64_t n = count(ch);
			const size_t off = (size_t)ch * E * in_bytes;
			if (host_in != nullptr) {
				gli_host_copy(r.pin_in[b], (in_copy != nullptr) ? (char*)in_copy + off : nullptr, (const char*)host_in + off,
				              (size_t)n * in_bytes);
				GLI_HIP(hipMemcpyAsync(r.dev_in[b], r.pin_in[b], (size_t)n * in_bytes, hipMemcpyHostToDevice, r.up));
				GLI_HIP(hipEventRecord(r.ev_up[b], r.up));
				GLI_HIP(hipStreamWaitEvent(c.stream, r.ev_up[b], 0));
			}
			work(r.dev_in[b], r.dev_out[b], n);
			if (host_out != nullptr) {
				GLI_HIP(hipEventRecord(r.ev_work[b], c.stream));
				GLI_HIP(hipStreamWaitEvent(r.down, r.ev_work[b], 0));
				GLI_HIP(hipMemcpyAsync(r.pin_out[b], r.dev_out[b], (size_t)n * out_bytes, hipMemcpyDeviceToHost, r.down));
				GLI_HIP(hipEventRecord(r.ev_done[b], r.down));
			}
			else {
				GLI_HIP(hipEventRecord(r.ev_done[b], c.stream));
			}
		}
	}
}

void download(void* host_dst, const void* dev_src, size_t bytes)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	if (bytes == 0) { return; }
	XferRing& r = ring();
	// the data is complete once the library stream's earlier work is
	GLI_HIP(hipEventRecord(r.ev_work[0], c.stream));
	GLI_HIP(hipStreamWaitEvent(r.down, r.ev_work[0], 0));
	const int64_t nc = (int64_t)ceil_div((int64_t)bytes, (int64_t)kChunkBytes);
	auto len = [&](int64_t ch) { return std::min(kChunkBytes, bytes - (size_t)ch * kChunkBytes); };
	for (int64_t ch = 0; ch < nc + kSlots; ++ch) {
		if (ch >= kSlots) {
			const int64_t d = ch - kSlots;
			const int b = (int)(d % kSlots);
			GLI_HIP(hipEventSynchronize(r.ev_done[b]));
			gli_host_copy((char*)host_dst + (size_t)d * kChunkBytes, nullptr, r.pin_out[b], len(d));
		}
		if (ch < nc) {
			const int b = (int)(ch % kSlots);
			GLI_HIP(hipMemcpyAsync(r.pin_out[b], (const char*)dev_src + (size_t)ch * kChunkBytes, len(ch), hipMemcpyDeviceToHost, r.down));
			GLI_HIP(hipEventRecord(r.ev_done[b], r.down));
		}
	}
}

void upload(void* dev_dst, const void* host_src, size_t bytes)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	if (bytes == 0) { return; }
	XferRing& r = ring();
	// the destination may still be read by the library stream's queued work
	GLI_HIP(hipEventRecord(r.ev_work[0], c.stream));
	GLI_HIP(hipStreamWaitEvent(r.up, r.ev_work[0], 0));
	const int64_t nc = (int64_t)ceil_div((int64_t)bytes, (int64_t)kChunkBytes);
	auto len = [&](int64_t ch) { return std::min(kChunkBytes, bytes - (size_t)ch * kChunkBytes); };
	for (int64_t ch = 0; ch < nc; ++ch) {
		const int b = (int)(ch % kSlots);
		if (ch >= kSlots) { GLI_HIP(hipEventSynchronize(r.ev_up[b])); }   // pinned buffer b is free again
		gli_host_copy(r.pin_in[b], nullptr, (const char*)host_src + (size_t)ch * kChunkBytes, len(ch));
		GLI_HIP(hipMemcpyAsync((char*)dev_dst + (size_t)ch * kChunkBytes, r.pin_in[b], len(ch), hipMemcpyHostToDevice, r.up));
		GLI_HIP(hipEventRecord(r.ev_up[b], r.up));
	}
	// later work on the library stream sees the data
	GLI_HIP(hipEventRecord(r.ev_work[0], r.up));
	GLI_HIP(hipStreamWaitEvent(c.stream, r.ev_work[0], 0));
	GLI_HIP(hipStreamSynchronize(r.up));   // the source may be reused when this returns
}

void copy_to_host(void* host_dst, const void* dev_src, size_t bytes)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	if (bytes == 0) { return; }
	if (bytes < kRingMinBytes) {
		GLI_HIP(hipMemcpyAsync(host_dst, dev_src, bytes, hipMemcpyDeviceToHost, c.stream));
		GLI_HIP(hipStreamSynchronize(c.stream));
		return;
	}
	download(host_dst, dev_src, bytes);
}

void copy_to_device(void* dev_dst, const void* host_src, size_t bytes)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	if (bytes == 0) { return; }
	if (bytes < kRingMinBytes) {
		GLI_HIP(hipMemcpyAsync(dev_dst, host_src, bytes, hipMemcpyHostToDevice, c.stream));
		GLI_HIP(hipStreamSynchronize(c.stream));
		return;
	}
	upload(dev_dst, host_src, bytes);
}

}  // namespace gli
