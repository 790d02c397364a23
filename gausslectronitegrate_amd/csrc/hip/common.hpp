// Shared device context, error handling and small helpers of the HIP side.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

namespace gli {

// ---------------------------------------------------------------------------
// errors: every failure records a message; the C entry points turn it into a
// status code, the reference-compatible void entry points into an abort.

void set_error(const char* file, int line, const std::string& msg);

struct Failure {
	std::string msg;
};

#define GLI_HIP(call)                                                                          \
	do {                                                                                       \
		hipError_t e_ = (call);                                                                \
		if (e_ != hipSuccess) {                                                                \
			::gli::set_error(__FILE__, __LINE__, std::string(#call) + ": " + hipGetErrorString(e_)); \
			throw ::gli::Failure{hipGetErrorString(e_)};                                       \
		}                                                                                      \
	} while (0)

#define GLI_REQUIRE(cond, text)                                                 \
	do {                                                                        \
		if (!(cond)) {                                                          \
			::gli::set_error(__FILE__, __LINE__, text);                         \
			throw ::gli::Failure{text};                                         \
		}                                                                       \
	} while (0)

// ---------------------------------------------------------------------------
// device buffers

template <typename T>
struct DevBuf {
	T* p = nullptr;
	size_t n = 0;
	DevBuf() = default;
	explicit DevBuf(size_t count) { alloc(count); }
	DevBuf(const DevBuf&) = delete;
	DevBuf& operator=(const DevBuf&) = delete;
	DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
	DevBuf& operator=(DevBuf&& o) noexcept
	{
		if (this != &o) {
			release();
			p = o.p; n = o.n;
			o.p = nullptr; o.n = 0;
		}
		return *this;
	}
	~DevBuf() { release(); }
	void alloc(size_t count)
	{
		release();
		if (count > 0) { GLI_HIP(hipMalloc(&p, count * sizeof(T))); }
		n = count;
	}
	// grow-only reuse
	void reserve(size_t count)
	{
		if (count > n) { alloc(count); }
	}
	void release()
	{
		if (p != nullptr) { (void)hipFree(p); }
		p = nullptr;
		n = 0;
	}
};

// ---------------------------------------------------------------------------
// process-wide context: one stream at default priority, cached Coulomb tables

struct Context {
	int device = 0;
	hipStream_t stream = nullptr;
	std::recursive_mutex mu;
	// ERI kernel table T[rho], w = 1/3, rho = 36 r^2 (grows on demand)
	std::vector<double> eri_table_host;
	DevBuf<double> eri_table_dev;
	// ERIDA kernel table, w = sqrt(2)/3, rho = 9 r^2
	std::vector<double> erida_table_host;
	DevBuf<double> erida_table_dev;
	// phase timing: a ring of event sets, one per evaluation, so that back-to-back evaluations
	// are timed without synchronising between them; durations accumulate per phase name
	static constexpr int kMaxPhases = 16, kPhaseSets = 64, kMaxPhaseNames = 32;
	struct PhaseSet {
		hipEvent_t ev[kMaxPhases + 1] = {};
		const char* name[kMaxPhases] = {};
		int n = 0;
		bool pending = false;
	};
	PhaseSet* phase_sets = nullptr;   // created on first enable
	int phase_cur = -1, phase_next = 0;
	bool phases_on = false;
	const char* acc_name[kMaxPhaseNames] = {};
	double acc_ms[kMaxPhaseNames] = {};
	int acc_calls[kMaxPhaseNames] = {};
	int acc_n = 0;
	// pinned staging ring for small host-to-device uploads (coefficients), so that an upload
	// never waits for earlier work on the stream
	static constexpr int kStageSlots = 32;
	static constexpr size_t kStageBytes = 8192;
	char* stage = nullptr;
	hipEvent_t stage_ev[kStageSlots] = {};
	bool stage_used[kStageSlots] = {};
	int stage_next = 0;
	// kernel tuning knobs (gli_set_tuning)
	int tile_grouped = 0;   // mode-product tile order (the XCD-grouped order measured slower on the ERI shapes)
	int gather_ept = 8;     // gather elements per lane
	int gather_lds = 1;     // gather keeps its key tables in LDS (when they fit)
	int gather_threads = 1024; // workgroup size of the single-grid gather (its waves share one CU's L1)
	int gather_bucketed = 0; // single-grid gather reads the key tensor in bucket order inside windows: 0 never
	                         // (default: with the k1-contiguous layout the list-order gather is faster, 13^3 5.75
	                         // against 7.40 ms), 1 always, 2 for key tensors above 128 MB
	int gather_classed = 0;  // otherwise each wave of the single-grid gather groups its key-tensor reads by line class
	                         // (with the size-ordered key numbering 1.308 against 1.274 ms at 11^3: off)
	int rankq_tet = 1;       // separable key tensor: the tetrahedral rank-Q kernel (0: the 64x64 tile kernel)
	int mode1_half = 1;     // FP64 paths store the first mode's (m2, m3)-symmetric output once
	int batch_branches = 1; // batch plans: 0 one branch, 1 autotuner times a two-branch graph, 2 always two
	int batch_positions = 1; // batch plans cache every element's key-tensor position at creation
	int batch_sorted = 1;   // ... and sort those positions inside windows of 8192 elements
	int gather_xcd = 2;     // gather order: 0 consecutive pieces, 1 one contiguous chunk of the list per XCD,
	                        // 2 (default) chunks unless the key tensor outgrows the shared last-level cache.
	                        // At 11^3 consecutive pieces are 0.7% faster (1.113 against 1.121 ms) but read 49%
	                        // more bytes past the L2 (calibrated traffic 1.63x instead of 1.26x the algorithmic)
	int gemm_shape = 0;     // FP64 mode-product tile shape (0: 64x64, 1: 128x64, 2: 64x128, 3: 128x128)
	int project_chunk_kib = 262144; // host memory per chunk of product states in the ERI projection (256 MB)
};

Context& ctx();

// ERI / ERIDA Coulomb tables of at least n entries, device resident
const double* eri_table(int64_t n);
const double* erida_table(int64_t n);
// the ERI Coulomb table of at least n entries in host memory (kept by the context; a later call
// that needs more entries may move it)
const double* eri_table_host(int64_t n);

// looks up the device address of the allocation-free ERI entry point's scratch on `device` (the
// current device) and keeps it; a failure is recorded and reported by that entry point
void eri_scratch_init(int device);

// Output-tile order of the mode products.  Workgroups are dispatched round-robin over the 8 XCDs
// (workgroup b runs on XCD b % 8), each with its own L2.  With the grouped order the column tiles
// of one row stripe run back to back on the same XCD, so the stripe of the large operand X is
// read from HBM once and then hits that XCD's L2; with the plain order (row tiles fastest) every
// column tile re-reads X from HBM.
struct TileOrder {
	int nr, nc;   // row and column tiles
	int grouped;  // 1: XCD-grouped order
	__host__ __device__ int blocks() const { return grouped ? ((nr + 7) / 8) * 8 * nc : nr * nc; }
	// (row tile, column tile) of workgroup b; false for padding workgroups
	__device__ bool tile(int b, int& rt, int& ct) const
	{
		if (grouped) {
			const int xcd = b & 7, local = b >> 3;
			rt = (local / nc) * 8 + xcd;
			ct = local % nc;
			return rt < nr;
		}
		if (b >= nr * nc) { return false; }
		rt = b % nr;
		ct = b / nr;
		return true;
	}
};

// the tile order used by the mode products (0 = row tiles fastest, the default; 1 = XCD-grouped;
// gli_set_tuning changes it)
int tile_order_grouped();

// asynchronous host-to-device copy of a small host buffer on the library stream (the source may
// be reused as soon as this returns)
void upload_async(void* dst, const void* src, size_t bytes);

// Tuples of np grid points (points[(e * np + p) * 3 + a]) as linear indices on one grid: each
// tuple moves to its own minimum corner (the integrals depend on point differences only), and
// the grid g[6] is the smallest box starting at 0 that holds every moved tuple.  Fails when the
// indices would reach 2^53.
void points_to_indices(const int64_t* points, int64_t num, int np, int64_t g[6], std::vector<int64_t>& idx);

// Streams num elements through the device in chunks: in_bytes per element of host_in go up,
// work(d_in, d_out, count) runs on the library stream for each chunk, and out_bytes per element
// of d_out come back into host_out.  in_copy (or nullptr) receives a host copy of host_in, made in
// the same pass as the staging copy.  Returns when host_out is complete.
void stream_chunks(const void* host_in, size_t in_bytes, void* host_out, size_t out_bytes, int64_t num, void* in_copy,
                   const std::function<void(const void*, void*, int64_t)>& work);
// bytes of one chunk buffer of the transfer ring
size_t stream_chunk_bytes();
// device-to-host copy through the transfer ring, after the library stream's queued work
void download(void* host_dst, const void* dev_src, size_t bytes);
// host-to-device copy through the transfer ring; the library stream's later work sees the data
void upload(void* dev_dst, const void* host_src, size_t bytes);
// synchronous copies of result or input arrays of any size: direct below 4 MB, else through the ring
void copy_to_host(void* host_dst, const void* dev_src, size_t bytes);
void copy_to_device(void* dev_dst, const void* host_src, size_t bytes);

// phase timing (records an event on the library stream when enabled)
void phase_begin();
void phase_mark(const char* name);

// Y[p][k] = sum_{mu < Kd} X[mu * ldx + p] * B[mu * ldb + k] for p < P, k < C (FP64 MFMA);
// with Cp > 0 and tri_cols / tri_rows, tiles outside k <= r / r <= outer (p = outer*Cp + r) are skipped;
// with tet_C1 > 0 the output goes to the tetrahedral layout of the sorted triple (k, r, outer), all < tet_C1
// Position of the sorted key triple k1 <= k2 <= k3 in the tetrahedral layout of a symmetric 3-tensor
// (the combinatorial number system): C(k3 + 2, 3) + C(k2 + 1, 2) + k1.  The smallest key is the
// contiguous one: with the size-ordered key numbering, the entries that neighbouring elements of a
// sorted list read share their (k2, k3) run more often (on the 11^3 list 145 instead of 186 distinct
// 128-byte lines per 512 consecutive elements).  The last product writes output tiles of fixed k3 whose
// rows run over k2 and columns over k1, so its stores are full lines.
__host__ __device__ inline int64_t tet_index(int k1, int k2, int k3)
{
	return (int64_t)k3 * (k3 + 1) * (k3 + 2) / 6 + (int64_t)k2 * (k2 + 1) / 2 + k1;
}

// T3[m1][m2][m3] = T[m1^2 + m2^2 + m3^2] for m <= M, 0 in the padding up to Mp (library stream)
void kernel_tensor(const double* T, int M, int Mp, double* T3);

// one FP64 mode product Y[p][k] = sum_mu X[mu][p] B[mu][k] (see eri_tensor.hip for the fields)
struct ModeArgs {
	const double* X;
	int64_t ldx, P;
	const double* B;
	int64_t ldb;
	int C, Kd;
	double* Y;
	int64_t ldy;
	int Cp, tri_cols, tri_rows;
	int64_t p_base;
	int tet_C1;
	// symmetric first mode (T3[m1][m2][m3] = T3[m1][m3][m2]): with half_n = n > 0 the rows are
	// p = m2 * n + m3 and only those with m3 >= m2 are computed and stored; a following product
	// with sym_x = 1 reads X[m2][m3][k] for m2 > m3 from X[m3][m2][k] (X rows m3 * Cp + k)
	int half_n = 0;
	int sym_x = 0;
	// rank-Q form of the last product (separable method): X[q][outer Cp + r] = cp_u[q Cp + outer] X[q ldx + r],
	// with ldx = Cp; the rows of a tile share outer
	const double* cp_u = nullptr;
};
constexpr int kGemmShapes = 6;
bool gemm_shape_fits(int shape, int Cp, int tri, int64_t p_base);
void mode_product_shaped(const ModeArgs& a, int shape, int grouped);
void mode_product_grouped(const ModeArgs* a, int n, int shape, int grouped);

void mode_product(const double* X, int64_t ldx, int64_t P, const double* B, int64_t ldb, int C, int Kd,
                  double* Y, int64_t ldy, int Cp, int tri_cols, int tri_rows, int64_t p_base, int tet_C1);
// the same with every field of ModeArgs (shape and order from the context's tuning)
void mode_product_args(const ModeArgs& a);

// split-precision (BF16 digit slices, exact FP32 slice sums, FP64 combination) variant:
// split_factor() prepares B once; mode_product_split() then contracts X with it.  Both use the
// library's shared split workspace unless a caller-owned one is given (a captured graph needs
// buffers that no later call can reallocate); tile_order < 0 takes the context knob; ns is the
// number of base-256 digit slices (5: mixed method, 2: BF16 hi/lo comparison method), fixed by
// split_factor for the products that follow it.
struct SplitWorkspace;
SplitWorkspace* split_workspace_new();
void split_workspace_delete(SplitWorkspace* w);
void split_factor(const double* B, int64_t ldb, int C, int Kd, SplitWorkspace* ws = nullptr, int ns = 5);
void mode_product_split(const double* X, int64_t ldx, int64_t P, int C, int Kd, double* Y, int64_t ldy,
                        int Cp, int tri_cols, int tri_rows, int64_t p_base, int tet_C1, bool cache_x,
                        SplitWorkspace* ws = nullptr, int tile_order = -1);
void split_invalidate_cache();

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t round_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }

}  // namespace gli
