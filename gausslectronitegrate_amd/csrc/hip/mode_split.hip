// Split-precision mode product on BF16 MFMA with exact FP32 slice sums.
//
// Y[p][k] = sum_{mu < Kd} X[mu][p] B[mu][k] is evaluated from base-256 digit
// slices (the error-free splitting of Ozaki, Ogita, Oishi and Rump, 2012):
// every column p of X and k of B is scaled by a power of two 2^e to |x| < 1
// and written as x = sum_s d_s 256^-(s+1) + r with integer digits
// |d_s| <= 255, which BF16 holds exactly.  A slice product sum_mu d_s d'_t
// is an integer below 255^2 * Kd < 2^24 for Kd <= 256, so the FP32 MFMA
// accumulation is exact; the products with s + t <= NS - 1 are combined in
// FP64 with weights 256^-(s+t+2) 2^(e_p + e_k).  With NS = 5 digits the
// dropped terms are below 2^-40 of the column scales (about 1e-11 absolute on
// the ERI key tensor of an 11^3 grid with K = 57).
#include "common.hpp"
#include "../gli_device.h"

namespace gli {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int NS = 5;                         // digit slices per operand
constexpr int NPROD = NS * (NS + 1) / 2;      // slice products with s + t <= NS - 1
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
	e[p] = ex;
}

// base-256 digits of a 32 (mu) x 64 (p) tile, written as D[(s * Ppad + p) * Kp + mu]
// (mu contiguous per column: the GEMM loads 16-byte runs of 8 digits)
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
			const double d = trunc(x);
			x -= d;
			const float f = (float)d;  // |d| <= 255: exact in the top 16 bits of the float
			T[sl][r][c] = (uint16_t)(__float_as_uint(f) >> 16);
		}
	}
	__syncthreads();
	for (int id = threadIdx.x; id < NS * 64 * (SBK / 8); id += 256) {
		const int sl = id / (64 * 4), r = (id / 4) % 64, c8 = (id % 4) * 8;
		*(uint4*)&D[((int64_t)sl * Ppad + p0 + r) * Kp + mu0 + c8] = *(const uint4*)&T[sl][r][c8];
	}
}

// Y[p][k] from the digits of X (Ppad columns) and B (Cpad columns)
__global__ __launch_bounds__(256) void k_mode_product_split(
	const uint16_t* __restrict__ Xd, int64_t Ppad, const int* __restrict__ ex,
	const uint16_t* __restrict__ Bd, int64_t Cpad, const int* __restrict__ eb,
	int64_t P, int C, int Kp, double* __restrict__ Y, int64_t ldy, int Cp, int tri_cols, int tri_rows, int64_t p_base,
	int tet_C1)
{
	__shared__ uint16_t As[NS][SBM][SBK + SPAD];
	__shared__ uint16_t Bs[NS][SBN][SBK + SPAD];
	const int tid = threadIdx.x;
	const int lane = tid & 63, wave = tid >> 6;
	const int64_t p0 = (int64_t)blockIdx.x * SBM;
	const int k0 = blockIdx.y * SBN;

	if (tri_cols || tri_rows) {
		const int64_t gp = p_base + p0;
		const int64_t outer = gp / Cp;
		const int64_t rlo = gp % Cp, rhi = rlo + SBM - 1;
		if (tri_cols && k0 + SBN - 1 < rlo) { return; }
		if (tri_rows && rhi < outer) { return; }
	}

	const int wm = wave >> 1, wn = wave & 1;
	f4 acc[NPROD][2][2];
	for (int q = 0; q < NPROD; ++q) {
		for (int i = 0; i < 2; ++i) {
			for (int j = 0; j < 2; ++j) { acc[q][i][j] = f4{0, 0, 0, 0}; }
		}
	}

	for (int mu0 = 0; mu0 < Kp; mu0 += SBK) {
		// 16-byte loads: NS * 64 rows * 32 digits = NS * 256 chunks of 8 digits per operand
		for (int id = tid; id < NS * SBM * (SBK / 8); id += 256) {
			const int s = id / (SBM * 4), r = (id / 4) % SBM, c8 = (id % 4) * 8;
			const uint4 va = *(const uint4*)&Xd[((int64_t)s * Ppad + p0 + r) * Kp + mu0 + c8];
			*(uint4*)&As[s][r][c8] = va;
			const uint4 vb = *(const uint4*)&Bd[((int64_t)s * Cpad + k0 + r) * Kp + mu0 + c8];
			*(uint4*)&Bs[s][r][c8] = vb;
		}
		__syncthreads();
		bf16x8 b[NS][2];
		for (int t = 0; t < NS; ++t) {
			for (int j = 0; j < 2; ++j) { b[t][j] = *(const bf16x8*)&Bs[t][wn * 32 + j * 16 + lane % 16][8 * (lane / 16)]; }
		}
		int q = 0;
		for (int s = 0; s < NS; ++s) {
			bf16x8 a[2];
			for (int i = 0; i < 2; ++i) { a[i] = *(const bf16x8*)&As[s][wm * 32 + i * 16 + lane % 16][8 * (lane / 16)]; }
			for (int t = 0; t < NS - s; ++t, ++q) {
				for (int i = 0; i < 2; ++i) {
					for (int j = 0; j < 2; ++j) {
						acc[q][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[t][j], acc[q][i][j], 0, 0, 0);
					}
				}
			}
		}
		__syncthreads();
	}

	// FP64 combination, smallest weights first
	for (int i = 0; i < 2; ++i) {
		for (int j = 0; j < 2; ++j) {
			const int col = k0 + wn * 32 + j * 16 + lane % 16;
			for (int v = 0; v < 4; ++v) {
				const int64_t row = p0 + wm * 32 + i * 16 + 4 * (lane / 16) + v;
				if (row >= P || col >= C) { continue; }
				double y = 0;
				for (int d = NS - 1; d >= 0; --d) {
					// products with s + t = d
					double g = 0;
					int qq = 0;
					for (int s = 0; s < NS; ++s) {
						for (int t = 0; t < NS - s; ++t, ++qq) {
							if (s + t == d) { g += (double)acc[qq][i][j][v]; }
						}
					}
					y += ldexp(g, -8 * (d + 2));
				}
				y = ldexp(y, ex[row] + eb[col]);
				if (tet_C1 > 0) {
					const int64_t gp = p_base + row;
					const int k1 = (int)(gp / Cp), k2 = (int)(gp % Cp);
					if (k1 <= k2 && k2 <= col && col < tet_C1) {
						Y[(int64_t)col * (col + 1) * (col + 2) / 6 + (int64_t)k2 * (k2 + 1) / 2 + k1] = y;
					}
				}
				else {
					Y[row * ldy + col] = y;
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
};

static SplitWorkspace& split_ws()
{
	static SplitWorkspace* w = new SplitWorkspace();
	return *w;
}

// digits of a Kd x P operand into (D, e) with padding; returns Ppad
static int64_t split_operand(const double* X, int64_t ld, int64_t P, int Kd, int Kp, DevBuf<uint16_t>& D, DevBuf<int>& e)
{
	const int64_t Ppad = round_up(P, SBM);
	D.reserve((size_t)NS * Ppad * Kp);
	e.reserve((size_t)Ppad);
	hipStream_t st = ctx().stream;
	k_column_exponents<<<(unsigned)ceil_div(Ppad, 256), 256, 0, st>>>(X, ld, P, Kd, Ppad, e.p);
	k_split_tile<<<dim3((unsigned)(Ppad / 64), (unsigned)(Kp / SBK)), 256, 0, st>>>(X, ld, P, Kd, Ppad, Kp, e.p, D.p);
	GLI_HIP(hipGetLastError());
	return Ppad;
}

// the B operand (factor matrix) is split once per evaluation and reused by all modes
void split_factor(const double* B, int64_t ldb, int C, int Kd)
{
	SplitWorkspace& w = split_ws();
	const int Kp = (int)round_up(Kd, SBK);
	split_operand(B, ldb, C, Kd, Kp, w.bd, w.be);
}

void mode_product_split(const double* X, int64_t ldx, int64_t P, int C, int Kd, double* Y, int64_t ldy,
                        int Cp, int tri_cols, int tri_rows, int64_t p_base, int tet_C1, bool cache_x)
{
	GLI_REQUIRE(Kd <= 256, "split-precision contraction needs at most 256 terms per sum");
	if (P <= 0 || C <= 0) { return; }
	SplitWorkspace& w = split_ws();
	const int Kp = (int)round_up(Kd, SBK);
	const int64_t Cpad = round_up(C, SBN);
	const uint16_t* xd;
	const int* xe;
	int64_t Ppad;
	if (cache_x) {
		// the kernel tensor operand is the same for every coefficient set of a grid
		if (w.t3_src != X || w.t3_P != P) {
			split_operand(X, ldx, P, Kd, Kp, w.t3d, w.t3e);
			w.t3_src = X;
			w.t3_P = P;
		}
		xd = w.t3d.p;
		xe = w.t3e.p;
		Ppad = round_up(P, SBM);
	}
	else {
		Ppad = split_operand(X, ldx, P, Kd, Kp, w.xd, w.xe);
		xd = w.xd.p;
		xe = w.xe.p;
	}
	dim3 grid((unsigned)(Ppad / SBM), (unsigned)(Cpad / SBN));
	k_mode_product_split<<<grid, 256, 0, ctx().stream>>>(xd, Ppad, xe, w.bd.p, Cpad, w.be.p, P, C, Kp, Y, ldy, Cp,
	                                                     tri_cols, tri_rows, p_base, tet_C1);
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
