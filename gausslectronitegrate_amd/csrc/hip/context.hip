// Process-wide device context, error state and the Coulomb kernel tables.
#include "common.hpp"
#include "../gli_device.h"

#include <cmath>
#include <cstring>

namespace gli {

static thread_local std::string g_last_error;
static thread_local int g_last_status = 0;

void set_error(const char* file, int line, const std::string& msg)
{
	char where[64];
	snprintf(where, sizeof(where), " [%s:%d]", strrchr(file, '/') ? strrchr(file, '/') + 1 : file, line);
	g_last_error = msg + where;
	g_last_status = -1;
}

Context& ctx()
{
	static Context* c = nullptr;
	static std::once_flag once;
	std::call_once(once, [] {
		c = new Context();
		int n = 0;
		if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
			set_error(__FILE__, __LINE__, "no HIP device available (the integral engine runs on an MI355X / gfx950 GPU)");
			throw Failure{"no HIP device"};
		}
		GLI_HIP(hipGetDevice(&c->device));
		// default priority: non-default priorities are not used anywhere in this library
		GLI_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
		for (int i = 0; i <= Context::kMaxPhases; ++i) { GLI_HIP(hipEventCreate(&c->phase_ev[i])); }
	});
	return *c;
}

// T[i] = erf(d / (sqrt(2) w)) / d with d = sqrt(i) / scale, T[0] = sqrt(2/pi) / w, evaluated
// in x87 extended precision and rounded to double, exactly like eri_integrals.c:18-27,569-574
// (scale 6, w = 1/3) and erida_integrals.c:17-26,445-451 (scale 3, w = sqrt(2)/3).
static void extend_table(std::vector<double>& t, int64_t n, long double w, long double scale)
{
	const int64_t old = (int64_t)t.size();
	if (n <= old) { return; }
	t.resize(n);
	for (int64_t i = old; i < n; ++i) {
		const long double d = sqrtl((long double)i) / scale;
		if (d > 0) {
			t[i] = (double)(erfl(d / (sqrtl(2) * w)) / d);
		}
		else {
			t[i] = (double)(0.7978845608028653558798921198687637L / w);
		}
	}
}

static const double* table(std::vector<double>& host, DevBuf<double>& dev, int64_t n, long double w, long double scale)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	if ((int64_t)dev.n < n) {
		// grow with headroom so that a batch of growing grids uploads once
		const int64_t want = std::max<int64_t>(n, (int64_t)(host.size() * 5 / 4));
		extend_table(host, want, w, scale);
		DevBuf<double> fresh(want);
		GLI_HIP(hipMemcpyAsync(fresh.p, host.data(), want * sizeof(double), hipMemcpyHostToDevice, c.stream));
		GLI_HIP(hipStreamSynchronize(c.stream));
		dev = std::move(fresh);
	}
	return dev.p;
}

const double* eri_table(int64_t n)
{
	Context& c = ctx();
	return table(c.eri_table_host, c.eri_table_dev, n, 1.L / 3, 6.L);
}

const double* erida_table(int64_t n)
{
	Context& c = ctx();
	const long double sqrt2_3 = 0.4714045207910316829338962414032327L;
	return table(c.erida_table_host, c.erida_table_dev, n, sqrt2_3, 3.L);
}

void phase_begin()
{
	Context& c = ctx();
	c.num_phases = 0;
	if (c.phases_on) { GLI_HIP(hipEventRecord(c.phase_ev[0], c.stream)); }
}

void phase_mark(const char* name)
{
	Context& c = ctx();
	if (!c.phases_on || c.num_phases >= Context::kMaxPhases) { return; }
	c.phase_name[c.num_phases] = name;
	c.num_phases++;
	GLI_HIP(hipEventRecord(c.phase_ev[c.num_phases], c.stream));
}

}  // namespace gli

using namespace gli;

extern "C" {

const char* gli_last_error(void)
{
	return g_last_error.c_str();
}

int gli_last_status(void)
{
	return g_last_status;
}

void gli_clear_error(void)
{
	g_last_error.clear();
	g_last_status = 0;
}

void* gli_stream(void)
{
	try {
		return (void*)ctx().stream;
	}
	catch (const Failure&) {
		return nullptr;
	}
}

int gli_synchronize(void)
{
	try {
		GLI_HIP(hipStreamSynchronize(ctx().stream));
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_enable_phase_timing(int on)
{
	try {
		ctx().phases_on = (on != 0);
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_phase_times(double* ms, const char** names, int max)
{
	try {
		Context& c = ctx();
		GLI_HIP(hipStreamSynchronize(c.stream));
		const int n = std::min(max, c.num_phases);
		for (int i = 0; i < n; ++i) {
			float t = 0;
			GLI_HIP(hipEventElapsedTime(&t, c.phase_ev[i], c.phase_ev[i + 1]));
			ms[i] = t;
			if (names != nullptr) { names[i] = c.phase_name[i]; }
		}
		return n;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_device_malloc(void** p, int64_t bytes)
{
	try {
		GLI_HIP(hipMalloc(p, (size_t)std::max<int64_t>(bytes, 1)));
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

void gli_device_free(void* p)
{
	if (p != nullptr) { (void)hipFree(p); }
}

int gli_memcpy(void* dst, const void* src, int64_t bytes, int kind)
{
	// kind: 0 = host->device, 1 = device->host, 2 = device->device
	try {
		const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
		GLI_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, k, ctx().stream));
		GLI_HIP(hipStreamSynchronize(ctx().stream));
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

}  // extern "C"
