// Split-precision mode product on BF16 MFMA with exact FP32 slice sums.
//
// Y[p][k] = sum_{mu < Kd} X[mu][p] B[mu][k] is evaluated from base-256 digit
// slices (the error-free splitting of Ozaki, Ogita, Oishi and Rump, 2012):
// every column p of X and k of B is scaled by a power of two to |x| < 1/2 and
// written with round-to-nearest digits as x = sum_s d_s 256^-(s+1) + r,
// |d_s| <= 128 (exact in BF16), |r| <= 256^-NS / 2.  A slice product
// d_s d'_t is at most 2^14; the products with the same digit sum s + t = d
// (at most NS of them) share one FP32 MFMA accumulator, whose partial sums
// stay integers below 2^24 over a chunk of KC = 192 terms, hence exact.  Each
// chunk is flushed into a compensated FP64 accumulator (TwoSum, hi + lo) with
// weights 256^-(d+2), and the column scales are applied at the end.  With NS = 5 the dropped terms are
// below about 2^-40 of the column scales (near 1e-11 absolute on the ERI key
// tensor of an 11^3 grid with K = 57).
#include "common.hpp"
#include "../gli_device.h"

namespace gli {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// NS = digit slices per operand: 5 for the mixed method (dropped terms below ~2^-40 of the column
// scales), 2 for the BF16 hi/lo comparison method (~2^-16)
constexpr int SBM = 64, SBN = 64, SBK = 32;   // block tile; K step = one MFMA 16x16x32
constexpr int SPAD = 8;                       // LDS row padding (bf16 elements)

// column exponents of a Kd x P matrix (row stride ld): |X[mu][p]| < 2^e[p]; padding columns get 0
__global__ void k_column_exponents(const double* __restrict__ X, int64_t ld, int64_t P, int Kd, int64_t Ppad, int* __restrict__ e)
{
	const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= Ppad) { return; }
	double mx = 0;
	if (p < P) {
		for (int mu = 0; mu < Kd; ++mu) { mx = fmax(mx, fabs(X[(int64_t)mu * ld + p])); }
	}
	int ex = 0;
	if (mx > 0 && isfinite(mx)) { (void)frexp(mx, &ex); }  // mx = f 2^ex, f in [1/2, 1)
	e[p] = ex + 1;  // |X / 2^e| < 1/2
}

// base-256 digits of a 32 (mu) x 64 (p) tile, written as D[(s * Ppad + p) * Kp + mu]
// (mu contiguous per column: the GEMM loads 16-byte runs of 8 digits)
template <int NS>
__global__ __launch_bounds__(256) void k_split_tile(const double* __restrict__ X, int64_t ld, int64_t P, int Kd,
	int64_t Ppad, int Kp, const int* __restrict__ e, uint16_t* __restrict__ D)
{
	__shared__ uint16_t T[NS][64][SBK + SPAD];
	const int64_t p0 = (int64_t)blockIdx.x * 64;
	const int mu0 = blockIdx.y * SBK;
	for (int id = threadIdx.x; id < 64 * SBK; id += 256) {
		const int r = id % 64, c = id / 64;   // consecutive threads: consecutive p (coalesced reads)
		const int64_t p = p0 + r;
		const int mu = mu0 + c;
		double x = (p < P && mu < Kd) ? ldexp(X[(int64_t)mu * ld + p], -e[p]) : 0.0;
		for (int sl = 0; sl < NS; ++sl) {
			x *= 256.0;
			const double d = rint(x);  // |x| < 128 before the first digit, <= 128 after
			x -= d;
			const float f = (float)d;  // |d| <= 128: exact in the top 16 bits of the float
			T[sl][r][c] = (uint16_t)(__float_as_uint(f) >> 16);
		}
	}
	__syncthreads();
	for (int id = threadIdx.x; id < NS * 64 * (SBK / 8); id += 256) {
		const int sl = id / (64 * 4), r = (id / 4) % 64, c8 = (id % 4) * 8;
		*(uint4*)&D[((int64_t)sl * Ppad + p0 + r) * Kp + mu0 + c8] = *(const uint4*)&T[sl][r][c8];
	}
}

// error-free sum s + e = a + b (Knuth's TwoSum)
__device__ inline void two_sum(double a, double b, double& s, double& e)
{
	s = a + b;
	const double bb = s - a;
	e = (a - (s - bb)) + (b - bb);
}

// compensated FP64 combination of the exact group sums of one chunk: sum_d grp[d] 256^-(d+2)
template <int NS>
__device__ inline void combine_groups(const float* g, double& hi, double& lo)
{
	for (int d = NS - 1; d >= 0; --d) {
		double e;
		two_sum(hi, ldexp((double)g[d], -8 * (d + 2)), hi, e);
		lo += e;
	}
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));   // 8 digits; a native vector stays in registers

// Y[p][k] from the digits of X (Ppad columns) and B (Cpad columns)
constexpr int KC = 192;  // exactness bound of an FP32 group accumulator: NS * 2^14 * KC < 2^24 for NS <= 5

template <int NS, bool kOneChunk>  // kOneChunk: the whole sum fits one exact chunk (Kp <= KC)
__global__ __launch_bounds__(256) void k_mode_product_split(
	const uint16_t* __restrict__ Xd, int64_t Ppad, const int* __restrict__ ex,
	const uint16_t* __restrict__ Bd, int64_t Cpad, const int* __restrict__ eb,
	int64_t P, int C, int Kp, double* __restrict__ Y, int64_t ldy, int Cp, int tri_cols, int tri_rows, int64_t p_base,
	int tet_C1, TileOrder order)
{
	__shared__ uint16_t As[NS][SBM][SBK + SPAD];
	__shared__ uint16_t Bs[NS][SBN][SBK + SPAD];
	const int tid = threadIdx.x;
	const int lane = tid & 63, wave = tid >> 6;
	int rt, ct;
	if (!order.tile(blockIdx.x, rt, ct)) { return; }
	const int64_t p0 = (int64_t)rt * SBM;
	const int k0 = ct * SBN;

	if (tri_cols || tri_rows) {
		const int64_t gp = p_base + p0;
		const int64_t outer = gp / Cp;
		const int64_t rlo = gp % Cp, rhi = rlo + SBM - 1;
		if (tri_cols && k0 > rhi) { return; }
		if (tri_rows && rlo > outer) { return; }
	}

	const int wm = wave >> 1, wn = wave & 1;
	f4 grp[NS][2][2];                   // FP32 group accumulators, one per digit sum
	double y[2][2][kOneChunk ? 1 : 4];  // compensated FP64 accumulators (hi, lo) across chunks
	double ylo[2][2][kOneChunk ? 1 : 4];
#pragma unroll
	for (int i = 0; i < 2; ++i) {
#pragma unroll
		for (int j = 0; j < 2; ++j) {
#pragma unroll
			for (int v = 0; v < (kOneChunk ? 1 : 4); ++v) { y[i][j][v] = 0; ylo[i][j][v] = 0; }
#pragma unroll
			for (int d = 0; d < NS; ++d) { grp[d][i][j] = f4{0, 0, 0, 0}; }
		}
	}

	// register prefetch of the next chunk: NS * 64 rows * 4 runs of 8 digits per operand (written
	// out inline: a lambda capturing the arrays sends them to scratch memory)
	constexpr int PER = NS * SBM * (SBK / 8) / 256;
	u32x4 ra[PER], rb[PER];
#pragma unroll
	for (int r = 0; r < PER; ++r) {
		const int id = tid + 256 * r;
		const int sl = id / (SBM * 4), row = (id / 4) % SBM, c8 = (id % 4) * 8;
		ra[r] = *(const u32x4*)&Xd[((int64_t)sl * Ppad + p0 + row) * Kp + c8];
		rb[r] = *(const u32x4*)&Bd[((int64_t)sl * Cpad + k0 + row) * Kp + c8];
	}
	for (int mu0 = 0; mu0 < Kp; mu0 += SBK) {
#pragma unroll
		for (int r = 0; r < PER; ++r) {
			const int id = tid + 256 * r;
			const int sl = id / (SBM * 4), row = (id / 4) % SBM, c8 = (id % 4) * 8;
			*(u32x4*)&As[sl][row][c8] = ra[r];
			*(u32x4*)&Bs[sl][row][c8] = rb[r];
		}
		__syncthreads();
		{
			// unconditional (the last one re-reads chunk 0 and is unused): one basic block
			const int mn = (mu0 + SBK < Kp) ? mu0 + SBK : 0;
#pragma unroll
			for (int r = 0; r < PER; ++r) {
				const int id = tid + 256 * r;
				const int sl = id / (SBM * 4), row = (id / 4) % SBM, c8 = (id % 4) * 8;
				ra[r] = *(const u32x4*)&Xd[((int64_t)sl * Ppad + p0 + row) * Kp + mn + c8];
				rb[r] = *(const u32x4*)&Bd[((int64_t)sl * Cpad + k0 + row) * Kp + mn + c8];
			}
		}
		bf16x8 b[NS][2];
#pragma unroll
		for (int t = 0; t < NS; ++t) {
#pragma unroll
			for (int j = 0; j < 2; ++j) { b[t][j] = *(const bf16x8*)&Bs[t][wn * 32 + j * 16 + lane % 16][8 * (lane / 16)]; }
		}
#pragma unroll
		for (int sl = 0; sl < NS; ++sl) {
			bf16x8 a[2];
#pragma unroll
			for (int i = 0; i < 2; ++i) { a[i] = *(const bf16x8*)&As[sl][wm * 32 + i * 16 + lane % 16][8 * (lane / 16)]; }
#pragma unroll
			for (int t = 0; t < NS - sl; ++t) {
#pragma unroll
				for (int i = 0; i < 2; ++i) {
#pragma unroll
					for (int j = 0; j < 2; ++j) {
						grp[sl + t][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[t][j], grp[sl + t][i][j], 0, 0, 0);
					}
				}
			}
		}
		__syncthreads();
		// flush the exact group sums into FP64 at the end of every KC chunk and at the end
		if (!kOneChunk && (((mu0 + SBK) % KC) == 0 || mu0 + SBK >= Kp)) {
#pragma unroll
			for (int i = 0; i < 2; ++i) {
#pragma unroll
				for (int j = 0; j < 2; ++j) {
#pragma unroll
					for (int v = 0; v < 4; ++v) {
						float g[NS];
#pragma unroll
						for (int d = 0; d < NS; ++d) { g[d] = grp[d][i][j][v]; }
						combine_groups<NS>(g, y[i][j][kOneChunk ? 0 : v], ylo[i][j][kOneChunk ? 0 : v]);
					}
#pragma unroll
					for (int d = 0; d < NS; ++d) { grp[d][i][j] = f4{0, 0, 0, 0}; }
				}
			}
		}
	}

	// accumulator layout of the 16x16 FP32 block: lane holds column lane%16, rows 4*(lane/16)+v
	int64_t tk3 = 0, tk2 = 0;   // the tile's rows share the largest key tk3 (tetrahedral store)
	if (tet_C1 > 0) {
		tk3 = (p_base + p0) / Cp;
		tk2 = p_base + p0 - tk3 * Cp;
	}
#pragma unroll
	for (int i = 0; i < 2; ++i) {
#pragma unroll
		for (int j = 0; j < 2; ++j) {
			const int col = k0 + wn * 32 + j * 16 + lane % 16;
			const int64_t tbase = (tet_C1 > 0) ? tet_index(col, 0, (int)tk3) : 0;
#pragma unroll
			for (int v = 0; v < 4; ++v) {
				const int dr = wm * 32 + i * 16 + 4 * (lane / 16) + v;
				const int64_t row = p0 + dr;
				if (row >= P || col >= C) { continue; }
				double acc;
				if (kOneChunk) {
					float g[NS];
#pragma unroll
					for (int d = 0; d < NS; ++d) { g[d] = grp[d][i][j][v]; }
					double hi = 0, lo = 0;
					combine_groups<NS>(g, hi, lo);
					acc = hi + lo;
				}
				else {
					acc = y[i][j][kOneChunk ? 0 : v] + ylo[i][j][kOneChunk ? 0 : v];
				}
				const double val = ldexp(acc, ex[row] + eb[col]);
				if (tet_C1 > 0) {
					const int64_t k2 = tk2 + dr;
					if (col <= k2 && k2 <= tk3 && tk3 < tet_C1) { Y[tbase + k2 * (k2 + 1) / 2] = val; }
				}
				else {
					Y[row * ldy + col] = val;
				}
			}
		}
	}
}

struct SplitWorkspace {
	DevBuf<uint16_t> xd, bd, t3d;
	DevBuf<int> xe, be, t3e;
	const double* t3_src = nullptr;
	int64_t t3_P = -1;
	int t3_ns = 0;
	int b_ns = 0;   // digit count of the factor held in bd
};

static SplitWorkspace& split_ws()
{
	static SplitWorkspace* w = new SplitWorkspace();
	return *w;
}

SplitWorkspace* split_workspace_new()
{
	return new SplitWorkspace();
}

void split_workspace_delete(SplitWorkspace* w)
{
	delete w;
}

// digits of a Kd x P operand into (D, e) with padding; returns Ppad
static int64_t split_operand(const double* X, int64_t ld, int64_t P, int Kd, int Kp, DevBuf<uint16_t>& D, DevBuf<int>& e,
                             int ns)
{
	GLI_REQUIRE(ns == 2 || ns == 5, "digit slices must be 2 or 5");
	const int64_t Ppad = round_up(P, SBM);
	D.reserve((size_t)ns * Ppad * Kp);
	e.reserve((size_t)Ppad);
	hipStream_t st = ctx().stream;
	k_column_exponents<<<(unsigned)ceil_div(Ppad, 256), 256, 0, st>>>(X, ld, P, Kd, Ppad, e.p);
	const dim3 grid((unsigned)(Ppad / 64), (unsigned)(Kp / SBK));
	if (ns == 2) { k_split_tile<2><<<grid, 256, 0, st>>>(X, ld, P, Kd, Ppad, Kp, e.p, D.p); }
	else { k_split_tile<5><<<grid, 256, 0, st>>>(X, ld, P, Kd, Ppad, Kp, e.p, D.p); }
	GLI_HIP(hipGetLastError());
	return Ppad;
}

// the B operand (factor matrix) is split once per evaluation and reused by all modes
void split_factor(const double* B, int64_t ldb, int C, int Kd, SplitWorkspace* ws, int ns)
{
	SplitWorkspace& w = ws ? *ws : split_ws();
	const int Kp = (int)round_up(Kd, SBK);
	split_operand(B, ldb, C, Kd, Kp, w.bd, w.be, ns);
	w.b_ns = ns;
}

void mode_product_split(const double* X, int64_t ldx, int64_t P, int C, int Kd, double* Y, int64_t ldy,
                        int Cp, int tri_cols, int tri_rows, int64_t p_base, int tet_C1, bool cache_x,
                        SplitWorkspace* ws, int tile_order)
{
	if (P <= 0 || C <= 0) { return; }
	SplitWorkspace& w = ws ? *ws : split_ws();
	const int ns = w.b_ns;
	GLI_REQUIRE(ns == 2 || ns == 5, "split_factor must run before mode_product_split");
	GLI_REQUIRE(!(tri_cols || tri_rows || tet_C1 > 0) || (Cp % SBM == 0 && p_base % SBM == 0),
	            "mode-product tile does not fit the key padding");
	const int Kp = (int)round_up(Kd, SBK);
	const int64_t Cpad = round_up(C, SBN);
	const uint16_t* xd;
	const int* xe;
	int64_t Ppad;
	if (cache_x) {
		// the kernel tensor operand is the same for every coefficient set of a grid
		if (w.t3_src != X || w.t3_P != P || w.t3_ns != ns) {
			split_operand(X, ldx, P, Kd, Kp, w.t3d, w.t3e, ns);
			w.t3_src = X;
			w.t3_P = P;
			w.t3_ns = ns;
		}
		xd = w.t3d.p;
		xe = w.t3e.p;
		Ppad = round_up(P, SBM);
	}
	else {
		Ppad = split_operand(X, ldx, P, Kd, Kp, w.xd, w.xe, ns);
		xd = w.xd.p;
		xe = w.xe.p;
	}
	const TileOrder order{(int)(Ppad / SBM), (int)(Cpad / SBN), tile_order >= 0 ? tile_order : tile_order_grouped()};
	const unsigned grid = (unsigned)order.blocks();
	hipStream_t st = ctx().stream;
#define GLI_SPLIT_LAUNCH(NSV, ONE) \
	k_mode_product_split<NSV, ONE><<<grid, 256, 0, st>>>(xd, Ppad, xe, w.bd.p, Cpad, w.be.p, P, C, Kp, Y, ldy, Cp, tri_cols, \
	                                                     tri_rows, p_base, tet_C1, order)
	if (ns == 2) {
		if (Kp <= KC) { GLI_SPLIT_LAUNCH(2, true); }
		else { GLI_SPLIT_LAUNCH(2, false); }
	}
	else {
		if (Kp <= KC) { GLI_SPLIT_LAUNCH(5, true); }
		else { GLI_SPLIT_LAUNCH(5, false); }
	}
#undef GLI_SPLIT_LAUNCH
	GLI_HIP(hipGetLastError());
}

void split_invalidate_cache()
{
	split_ws().t3_src = nullptr;
}

}  // namespace gli

using namespace gli;

extern "C" int gli_mode_product_split_device(const double* X, int64_t ldx, int64_t P, const double* B, int64_t ldb, int64_t C,
                                             int64_t Kd, double* Y, int64_t ldy)
{
	try {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		GLI_REQUIRE(P >= 0 && C >= 0 && Kd >= 0 && ldx >= P && ldb >= C && ldy >= C, "inconsistent mode-product shapes");
		split_factor(B, ldb, (int)C, (int)Kd);
		mode_product_split(X, ldx, P, (int)C, (int)Kd, Y, ldy, 0, 0, 0, 0, 0, false);
		GLI_HIP(hipStreamSynchronize(c.stream));
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}
