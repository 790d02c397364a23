// Process-wide device context, error state and the Coulomb kernel tables.
#include "common.hpp"
#include <algorithm>
#include "../gli_device.h"

#include <cmath>
#include <cstring>
#include <string>

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
		// the device address of the allocation-free entry point's scratch on this device
		eri_scratch_init(c->device);
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

const double* eri_table_host(int64_t n)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	extend_table(c.eri_table_host, n, 1.L / 3, 6.L);
	return c.eri_table_host.data();
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

void points_to_indices(const int64_t* points, int64_t num, int np, int64_t g[6], std::vector<int64_t>& idx)
{
	GLI_REQUIRE(num >= 0, "negative number of point tuples");
	GLI_REQUIRE(num == 0 || points != nullptr, "no points given");
	int64_t span[3] = { 0, 0, 0 };
	for (int64_t e = 0; e < num; ++e) {
		for (int a = 0; a < 3; ++a) {
			int64_t mn = points[(e * np) * 3 + a], mx = mn;
			for (int p = 1; p < np; ++p) {
				mn = std::min(mn, points[(e * np + p) * 3 + a]);
				mx = std::max(mx, points[(e * np + p) * 3 + a]);
			}
			GLI_REQUIRE(mx - mn < 1024, "points of one tuple lie 1024 or more grid steps apart");
			span[a] = std::max(span[a], mx - mn);
		}
	}
	for (int a = 0; a < 3; ++a) {
		g[2 * a] = 0;
		g[2 * a + 1] = span[a] + 1;
	}
	const double N = (double)g[1] * g[3] * g[5];
	double lim = 1;
	for (int p = 0; p < np; ++p) { lim *= N; }
	GLI_REQUIRE(lim < 9007199254740992.0, "point tuples too far apart for one batch: indices would reach 2^53");
	const int64_t Ni = g[1] * g[3] * g[5];
	idx.resize((size_t)num);
	for (int64_t e = 0; e < num; ++e) {
		int64_t mn[3];
		for (int a = 0; a < 3; ++a) {
			mn[a] = points[(e * np) * 3 + a];
			for (int p = 1; p < np; ++p) { mn[a] = std::min(mn[a], points[(e * np + p) * 3 + a]); }
		}
		int64_t t = 0;
		for (int p = 0; p < np; ++p) {
			const int64_t* q = points + (e * np + p) * 3;
			t = t * Ni + ((q[0] - mn[0]) * g[3] + (q[1] - mn[1])) * g[5] + (q[2] - mn[2]);
		}
		idx[(size_t)e] = t;
	}
}

void upload_async(void* dst, const void* src, size_t bytes)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	if (bytes == 0) { return; }
	if (bytes > Context::kStageBytes) {
		GLI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c.stream));
		GLI_HIP(hipStreamSynchronize(c.stream));
		return;
	}
	if (c.stage == nullptr) {
		GLI_HIP(hipHostMalloc((void**)&c.stage, Context::kStageSlots * Context::kStageBytes, hipHostMallocDefault));
		for (int i = 0; i < Context::kStageSlots; ++i) { GLI_HIP(hipEventCreateWithFlags(&c.stage_ev[i], hipEventDisableTiming)); }
	}
	const int slot = c.stage_next;
	c.stage_next = (slot + 1) % Context::kStageSlots;
	if (c.stage_used[slot]) { GLI_HIP(hipEventSynchronize(c.stage_ev[slot])); }
	char* buf = c.stage + (size_t)slot * Context::kStageBytes;
	std::memcpy(buf, src, bytes);
	GLI_HIP(hipMemcpyAsync(dst, buf, bytes, hipMemcpyHostToDevice, c.stream));
	GLI_HIP(hipEventRecord(c.stage_ev[slot], c.stream));
	c.stage_used[slot] = true;
}

int tile_order_grouped()
{
	return ctx().tile_grouped;
}

// fold one completed event set into the per-name sums
static void phase_harvest(Context& c, Context::PhaseSet& set)
{
	if (!set.pending) { return; }
	set.pending = false;
	if (set.n == 0) { return; }
	GLI_HIP(hipEventSynchronize(set.ev[set.n]));
	for (int i = 0; i < set.n; ++i) {
		float t = 0;
		GLI_HIP(hipEventElapsedTime(&t, set.ev[i], set.ev[i + 1]));
		int slot = 0;
		while (slot < c.acc_n && std::strcmp(c.acc_name[slot], set.name[i]) != 0) { ++slot; }
		if (slot == c.acc_n) {
			if (c.acc_n == Context::kMaxPhaseNames) { continue; }
			c.acc_name[c.acc_n] = set.name[i];
			c.acc_ms[c.acc_n] = 0;
			c.acc_calls[c.acc_n] = 0;
			c.acc_n++;
		}
		c.acc_ms[slot] += t;
		c.acc_calls[slot] += 1;
	}
}

void phase_begin()
{
	Context& c = ctx();
	if (!c.phases_on) {
		c.phase_cur = -1;
		return;
	}
	Context::PhaseSet& set = c.phase_sets[c.phase_next];
	phase_harvest(c, set);
	set.n = 0;
	set.pending = true;
	GLI_HIP(hipEventRecord(set.ev[0], c.stream));
	c.phase_cur = c.phase_next;
	c.phase_next = (c.phase_next + 1) % Context::kPhaseSets;
}

void phase_mark(const char* name)
{
	Context& c = ctx();
	if (!c.phases_on || c.phase_cur < 0) { return; }
	Context::PhaseSet& set = c.phase_sets[c.phase_cur];
	if (set.n >= Context::kMaxPhases) { return; }
	set.name[set.n] = name;
	set.n++;
	GLI_HIP(hipEventRecord(set.ev[set.n], c.stream));
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

int gli_set_tuning(const char* knob, int value)
{
	try {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		const std::string k = knob ? knob : "";
		if (k == "tile_order") {
			GLI_REQUIRE(value == 0 || value == 1, "tile_order must be 0 or 1");
			c.tile_grouped = value;
		}
		else if (k == "gather_ept") {
			GLI_REQUIRE(value == 2 || value == 4 || value == 8 || value == 16, "gather_ept must be 2, 4, 8 or 16");
			c.gather_ept = value;
		}
		else if (k == "batch_positions") {
			GLI_REQUIRE(value == 0 || value == 1, "batch_positions must be 0 or 1");
			c.batch_positions = value;
		}
		else if (k == "batch_sorted") {
			GLI_REQUIRE(value == 0 || value == 1, "batch_sorted must be 0 or 1");
			c.batch_sorted = value;
		}
		else if (k == "batch_branches") {
			GLI_REQUIRE(value >= 0 && value <= 2, "batch_branches must be 0, 1 or 2");
			c.batch_branches = value;
		}
		else if (k == "mode1_half") {
			GLI_REQUIRE(value == 0 || value == 1, "mode1_half must be 0 or 1");
			c.mode1_half = value;
		}
		else if (k == "gather_threads") {
			GLI_REQUIRE(value == 256 || value == 512 || value == 1024, "gather_threads must be 256, 512 or 1024");
			c.gather_threads = value;
		}
		else if (k == "gather_lds") {
			GLI_REQUIRE(value == 0 || value == 1, "gather_lds must be 0 or 1");
			c.gather_lds = value;
		}
		else if (k == "gemm_shape") {
			GLI_REQUIRE(value >= 0 && value < kGemmShapes, "gemm_shape must be 0..5");
			c.gemm_shape = value;
		}
		else if (k == "gather_order") {
			GLI_REQUIRE(value >= 0 && value <= 2, "gather_order must be 0, 1 or 2");
			c.gather_xcd = value;
		}
		else if (k == "gather_bucketed") {
			GLI_REQUIRE(value >= 0 && value <= 2, "gather_bucketed must be 0, 1 or 2");
			c.gather_bucketed = value;
		}
		else if (k == "gather_classed") {
			GLI_REQUIRE(value == 0 || value == 1, "gather_classed must be 0 or 1");
			c.gather_classed = value;
		}
		else if (k == "rankq_tet") {
			GLI_REQUIRE(value == 0 || value == 1, "rankq_tet must be 0 or 1");
			c.rankq_tet = value;
		}
		else if (k == "project_chunk_kib") {
			GLI_REQUIRE(value >= 1, "project_chunk_kib must be at least 1");
			c.project_chunk_kib = value;
		}
		else {
			GLI_REQUIRE(false, "unknown tuning knob");
		}
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_get_tuning(const char* knob)
{
	try {
		Context& c = ctx();
		const std::string k = knob ? knob : "";
		if (k == "tile_order") { return c.tile_grouped; }
		if (k == "gather_ept") { return c.gather_ept; }
		if (k == "gather_order") { return c.gather_xcd; }
		if (k == "gather_lds") { return c.gather_lds; }
		if (k == "gather_threads") { return c.gather_threads; }
		if (k == "mode1_half") { return c.mode1_half; }
		if (k == "batch_branches") { return c.batch_branches; }
		if (k == "batch_positions") { return c.batch_positions; }
		if (k == "batch_sorted") { return c.batch_sorted; }
		if (k == "gemm_shape") { return c.gemm_shape; }
		if (k == "project_chunk_kib") { return c.project_chunk_kib; }
		if (k == "gather_bucketed") { return c.gather_bucketed; }
		if (k == "gather_classed") { return c.gather_classed; }
		if (k == "rankq_tet") { return c.rankq_tet; }
		GLI_REQUIRE(false, "unknown tuning knob");
	}
	catch (const Failure&) {
	}
	return -1;
}

int gli_enable_phase_timing(int on)
{
	try {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		if (on && c.phase_sets == nullptr) {
			c.phase_sets = new Context::PhaseSet[Context::kPhaseSets];
			for (int s = 0; s < Context::kPhaseSets; ++s) {
				for (int i = 0; i <= Context::kMaxPhases; ++i) { GLI_HIP(hipEventCreate(&c.phase_sets[s].ev[i])); }
			}
		}
		if (on && !c.phases_on) {
			// a fresh measurement: drop what earlier evaluations left
			for (int s = 0; s < Context::kPhaseSets; ++s) { c.phase_sets[s].pending = false; }
			c.acc_n = 0;
		}
		c.phases_on = (on != 0);
		c.phase_cur = -1;
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
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		if (c.phase_sets != nullptr) {
			// harvest oldest first
			for (int k = 0; k < Context::kPhaseSets; ++k) {
				phase_harvest(c, c.phase_sets[(c.phase_next + k) % Context::kPhaseSets]);
			}
		}
		const int n = std::min(max, c.acc_n);
		for (int i = 0; i < n; ++i) {
			ms[i] = c.acc_ms[i] / std::max(c.acc_calls[i], 1);
			if (names != nullptr) { names[i] = c.acc_name[i]; }
		}
		c.acc_n = 0;
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

int gli_memset(void* dst, int value, int64_t bytes)
{
	try {
		GLI_REQUIRE(bytes >= 0 && (dst != nullptr || bytes == 0), "gli_memset needs a device pointer and a byte count >= 0");
		GLI_HIP(hipMemsetAsync(dst, value, (size_t)bytes, ctx().stream));
		GLI_HIP(hipStreamSynchronize(ctx().stream));
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

}  // extern "C"
