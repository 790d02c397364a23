// ERI values by the key-tensor contraction, plus the per-element FP64 path.
//
// For a four-index (ij|kl) with grid points p_i..p_l the reference evaluates
// (eri_integrals.c:608-669)
//     sum_{tx,ty,tz} Px[tx] Py[ty] Pz[tz] T[dx^2 + dy^2 + dz^2],  d_a = 3 cd_a + t_a,
// where P_a = P_{s01,s23} is the product table of the axis shifts
// s01 = (p_i - p_j)_a, s23 = (p_k - p_l)_a, cd_a = ((p_i + p_j) - (p_k + p_l))_a,
// and entries with |P| <= tol contribute nothing (eri_integrals.c:77-103, 629-658).
//
// T depends on |d_a| only, so each axis reduces to the folded column
//     B[mu][key] = sum_{t : |3 cd + t| = mu} P'_{s01,s23}[t],   mu = 0 .. M,
// with P' = P * [|P| > tol].  Because f_{-s} = f_s for every coefficient set,
// P_{s01,s23} depends on (|s01|, |s23|) only; for mirror-symmetric coefficients
// P is also even in t and symmetric in (s01, s23), so B depends on
// (min, max of |s01|, |s23|; |cd|).  The integral is then
//     V[k1][k2][k3] = sum_{mu1,mu2,mu3} B[mu1][k1] B[mu2][k2] B[mu3][k3] T[mu1^2 + mu2^2 + mu3^2],
// a fully symmetric 3-tensor over the distinct axis keys, built by three mode
// products of the kernel tensor T3 with the one factor matrix B (FP64 MFMA),
// and each element gathers V at its sorted key triple.  This is an exact
// re-association of the reference sum: same terms, different order.
#include "common.hpp"
#include "eri_integrals.h"
#include "../gli_device.h"
#include "gli_amd.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>
#include <type_traits>
#include <cstdio>
#include <cstring>
#include <vector>
#include <string>
#include <memory>

namespace gli {

typedef double d4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// 1D tables

// f[s][m] = (sqrt(pi)/3) sum_{a+b=m} c_a c_b exp(-(0.5 (3 s + a - b))^2), s = 0 .. R-1
// (gausslet_factors.c:12-50; f_{-s} = f_s), same ascending-a summation order, into fs[m].  ex holds the
// 2K - 1 Gaussians exp(-(0.5 (3 s + d))^2), d = a - b = -(K-1) .. K-1, of the shift: each exp is evaluated
// once per distinct argument, with the same argument and function as in the sum.  Called by a whole
// workgroup; cs holds the K coefficients.
__device__ inline void factor_row(const double* cs, int K, int s, double* ex, double* fs)
{
	for (int q = threadIdx.x; q < 2 * K - 1; q += blockDim.x) {
		const double x = 0.5 * (3 * s + (q - (K - 1)));
		ex[q] = exp(-x * x);
	}
	__syncthreads();
	const double prefac = 0.590817950301838675766;
	for (int m = threadIdx.x; m < 2 * K - 1; m += blockDim.x) {
		double acc = 0;
		const int ilo = max(0, m - (K - 1)), ihi = min(K - 1, m);
		for (int i = ilo; i <= ihi; ++i) {
			const int j = m - i;
			acc += prefac * cs[i] * cs[j] * ex[i - j + K - 1];
		}
		fs[m] = acc;
	}
}

// P[a][b][t] = sum_m f_a[m] f_b[m - (t - (2K-2))], zeroed where |P| <= tol
// (eri_integrals.c:59-103 with the tolerance skips of 629/637/658 folded in).  One workgroup per pair
// (a, b): it computes the two factor rows f_a and f_b itself (factor_row; 2 (2K - 1) K terms, cheaper
// than a separate launch) into LDS, then every t.  Dynamic LDS: K coefficients, 2K - 1 Gaussians and the
// two factor rows.
// Coefficient sets of up to kArgCoeffs entries travel as a kernel argument, so a single evaluation needs no
// host-to-device copy before its first kernel; longer sets are read from device memory.
constexpr int kArgCoeffs = 64;
struct CoeffArg {
	double c[kArgCoeffs];
};

__global__ void k_products(const double* __restrict__ c, CoeffArg ca, int K, int R, double tol, double* __restrict__ P)
{
	extern __shared__ double lds[];
	const int L = 2 * K - 1, LP = 4 * K - 3;
	double* cs = lds;
	double* ex = cs + K;
	double* fa = ex + L;
	double* fb = fa + L;
	const int ab = blockIdx.x;
	const int a = ab / R, b = ab % R;
	for (int q = threadIdx.x; q < K; q += blockDim.x) { cs[q] = (c != nullptr) ? c[q] : ca.c[q]; }
	__syncthreads();
	factor_row(cs, K, a, ex, fa);
	__syncthreads();
	factor_row(cs, K, b, ex, fb);
	__syncthreads();
	for (int t = threadIdx.x; t < LP; t += blockDim.x) {
		const int d = t - (2 * K - 2);  // m - m'
		double acc = 0;
		const int mlo = max(0, d), mhi = min(L - 1, L - 1 + d);
		for (int m = mlo; m <= mhi; ++m) { acc += fa[m] * fb[m - d]; }
		P[(int64_t)ab * LP + t] = (fabs(acc) > tol) ? acc : 0.0;
	}
}

// B[mu][q] = P'[a][b][mu - 3cd] + (mu > 0 ? P'[a][b][-mu - 3cd] : 0)
__global__ void k_fold(const double* __restrict__ P, const int4* __restrict__ keys, int C1, int K, int R,
                       int M, int Cp, double* __restrict__ B)
{
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	const int mu = blockIdx.y;
	if (q >= Cp) { return; }
	double v = 0;
	if (q < C1 && mu <= M) {
		const int4 k = keys[q];
		const int LP = 4 * K - 3, off = 2 * K - 2;
		const double* p = P + ((int64_t)k.x * R + k.y) * LP;
		const int t1 = mu - 3 * k.z + off;
		if (t1 >= 0 && t1 < LP) { v += p[t1]; }
		if (mu > 0) {
			const int t2 = -mu - 3 * k.z + off;
			if (t2 >= 0 && t2 < LP) { v += p[t2]; }
		}
	}
	B[(int64_t)mu * Cp + q] = v;
}

// T3[m1][m2][m3] = T[m1^2 + m2^2 + m3^2] for m <= M, zero padding up to Mp
__global__ void k_kernel_tensor(const double* __restrict__ T, int M, int Mp, double* __restrict__ T3)
{
	const int m3 = blockIdx.x * blockDim.x + threadIdx.x;
	const int m2 = blockIdx.y, m1 = blockIdx.z;
	if (m3 >= Mp) { return; }
	double v = 0;
	if (m1 <= M && m2 <= M && m3 <= M) { v = T[m1 * m1 + m2 * m2 + m3 * m3]; }
	T3[((int64_t)m1 * Mp + m2) * Mp + m3] = v;
}

// ---------------------------------------------------------------------------
// mode product  Y[p][k] = sum_{mu < Kd} X[mu][p] B[mu][k]   (FP64 MFMA 16x16x4)
//
// BM x BN output tile per workgroup of 4 waves in a 2 x 2 layout, each wave (BM/2) x (BN/2)
// = (BM/32) x (BN/32) MFMA blocks; mu staged through LDS GBK at a time with the next slice
// prefetched into registers.  Optional triangular skipping for the symmetric tensor: with
// p = outer * Cp + r,
//   tri_cols: only k <= r is needed;  tri_rows: only r <= outer is needed.
// With both and tet_C1 > 0 the tile writes V[k <= r <= outer] in the tetrahedral layout (tet_index).
// Tiles entirely outside the needed region are not computed (never read later).  The skip test
// needs (p_base + p0) / Cp fixed per tile: Cp and p_base multiples of BM.

constexpr int GBK = 16;   // mu slice of the FP32 comparison kernel
constexpr int KEY_ALIGN = 64;   // key columns are padded to a multiple of this

// tile shapes (kGemmShapes of them), BM x BN output tile and BK-deep mu slices:
// 0: 64x64x16, 1: 128x64x16, 2: 64x128x16, 3: 128x128x16, 4: 64x64x32, 5: 128x64x32
__host__ __device__ constexpr int shape_bm(int s) { return (s == 1 || s == 3 || s == 5) ? 128 : 64; }
__host__ __device__ constexpr int shape_bn(int s) { return (s == 2 || s == 3) ? 128 : 64; }


template <int BM, int BN, int GBK, bool kCp = false>
__device__ inline void mode_product_tile(const ModeArgs& a, int rt, int ct, double (*As)[BM], double (*Bs)[BN])
{
	constexpr int TM = BM / 32, TN = BN / 32;
	const int tid = threadIdx.x;
	const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform
	const int64_t p0 = (int64_t)rt * BM;
	const int k0 = ct * BN;

	if (a.tri_cols || a.tri_rows) {
		const int64_t gp = a.p_base + p0;
		const int64_t outer = gp / a.Cp;
		const int64_t rlo = gp % a.Cp, rhi = rlo + BM - 1;
		if (a.tri_cols && k0 > rhi) { return; }
		if (a.tri_rows && rlo > outer) { return; }
	}
	if (a.half_n > 0) {
		// rows p = m2 n + m3; the tile is needed when one of its rows has m3 >= m2
		const int64_t n = a.half_n, phi = (p0 + BM < a.P ? p0 + BM : a.P) - 1;
		bool need = false;
		for (int64_t m2 = p0 / n; m2 <= phi / n && !need; ++m2) {
			const int64_t lo = (p0 > m2 * n + m2) ? p0 : m2 * n + m2;
			const int64_t hi = (phi < m2 * n + n - 1) ? phi : m2 * n + n - 1;
			need = lo <= hi;
		}
		if (!need) { return; }
	}
	// mirrored reads of a half-stored X: the tile's rows share m3 (Cp is a multiple of BM, p_base = 0)
	const int64_t sym_m3 = a.sym_x ? p0 / a.Cp : 0;
	const int64_t sym_k = p0 - sym_m3 * a.Cp;

	const int wm = wave >> 1, wn = wave & 1;
	d4 acc[TM][TN];
#pragma unroll
	for (int i = 0; i < TM; ++i) {
#pragma unroll
		for (int j = 0; j < TN; ++j) { acc[i][j] = d4{0, 0, 0, 0}; }
	}
	// 16 x 16 blocks of this wave that hold no stored entry: with p = outer Cp + r, tri_cols keeps k <= r
	// and tri_rows r <= outer (the tetrahedral store keeps both), so a block whose smallest column exceeds
	// its largest r, or whose smallest r exceeds outer, is never stored and its MFMAs are skipped
	// (wave-uniform; the stored blocks are unchanged)
	bool need[TM][TN];
	{
		const int64_t gp = a.p_base + p0;
		const int64_t outer = (a.tri_cols || a.tri_rows) ? gp / a.Cp : 0;
		const int64_t rlo = (a.tri_cols || a.tri_rows) ? gp - outer * a.Cp : 0;
#pragma unroll
		for (int i = 0; i < TM; ++i) {
			const int64_t rmin = rlo + wm * (BM / 2) + i * 16, rmax = rmin + 15;
#pragma unroll
			for (int j = 0; j < TN; ++j) {
				const int64_t cmin = k0 + wn * (BN / 2) + j * 16;
				need[i][j] = (!a.tri_cols || cmin <= rmax) && (!a.tri_rows || rmin <= outer);
			}
		}
	}

	// software pipeline: the next mu-slice is loaded into registers while the current one
	// (already in LDS) feeds the MFMAs
	constexpr int PA = (GBK * BM) / 256, PB = (GBK * BN) / 256;
	double ra[PA], rb[PB];
	// branch-free loads with 32-bit element offsets (the host keeps both operands below 2^31
	// elements); an out-of-range element reads offset 0 and is replaced by 0.  Divergent branches
	// here make the compiler move the accumulators between AGPRs and VGPRs on every slice.
	// The in-range masks are applied when the values go to LDS (the next slice), so that the loads
	// stay in flight during this slice's MFMAs.
	const int ldx = (int)a.ldx, ldb = (int)a.ldb, p0i = (int)p0, Pi = (int)a.P;
	const int m3i = (int)sym_m3, sym_base = (int)(sym_m3 * a.ldx + sym_k);
	// rank-Q form: the tile's rows p = outer Cp + r share outer; X[q][p] = cp_u[q Cp + outer] X[q Cp + r]
	const int cp_k1 = kCp ? (int)(p0 / a.Cp) : 0;
	const int cp_off = kCp ? (int)(p0 - (int64_t)cp_k1 * a.Cp) : 0;
	double su[kCp ? PA : 1];
	bool oka[PA], okb[PB];
	auto gload = [&](int mu0) {
#pragma unroll
		for (int r = 0; r < PA; ++r) {
			const int id = tid + 256 * r;
			const int mu = id / BM, p = id % BM;
			const int m = mu0 + mu;
			oka[r] = m < a.Kd && p0i + p < Pi;
			if (kCp) {
				ra[r] = a.X[oka[r] ? (unsigned)(m * ldx + cp_off + p) : 0u];
				su[kCp ? r : 0] = a.cp_u[oka[r] ? (unsigned)(m * a.Cp + cp_k1) : 0u];
				continue;
			}
			const bool mirror = a.sym_x && m > m3i;
			const int off = (mirror ? sym_base + m * a.Cp : m * ldx + p0i) + p;
			ra[r] = a.X[oka[r] ? (unsigned)off : 0u];
		}
#pragma unroll
		for (int r = 0; r < PB; ++r) {
			const int id = tid + 256 * r;
			const int mu = id / BN, k = id % BN;
			okb[r] = mu0 + mu < a.Kd && k0 + k < a.C;
			rb[r] = a.B[okb[r] ? (unsigned)((mu0 + mu) * ldb + k0 + k) : 0u];
		}
	};
	// the mu loop in two versions: every block needed (most tiles: the plain unrolled MFMA sequence), or
	// a wave-uniform test per block (tiles on the faces of the stored region)
	auto mu_loop = [&](auto every) {
		gload(0);
		for (int mu0 = 0; mu0 < a.Kd; mu0 += GBK) {
#pragma unroll
			for (int r = 0; r < PA; ++r) {
				const int id = tid + 256 * r;
				As[id / BM][id % BM] = oka[r] ? (kCp ? ra[r] * su[kCp ? r : 0] : ra[r]) : 0.0;
			}
#pragma unroll
			for (int r = 0; r < PB; ++r) {
				const int id = tid + 256 * r;
				Bs[id / BN][id % BN] = okb[r] ? rb[r] : 0.0;
			}
			__syncthreads();
			gload(mu0 + GBK);   // past the last slice every element is masked off (reads offset 0)
#pragma unroll
			for (int kk = 0; kk < GBK; kk += 4) {
				double av[TM], bv[TN];
#pragma unroll
				for (int i = 0; i < TM; ++i) { av[i] = As[kk + lane / 16][wm * (BM / 2) + i * 16 + lane % 16]; }
#pragma unroll
				for (int j = 0; j < TN; ++j) { bv[j] = Bs[kk + lane / 16][wn * (BN / 2) + j * 16 + lane % 16]; }
#pragma unroll
				for (int i = 0; i < TM; ++i) {
#pragma unroll
					for (int j = 0; j < TN; ++j) {
						if (decltype(every)::value || need[i][j]) {
							acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
						}
					}
				}
			}
			__syncthreads();
		}
	};
	bool every = true;
#pragma unroll
	for (int i = 0; i < TM; ++i) {
#pragma unroll
		for (int j = 0; j < TN; ++j) { every = every && need[i][j]; }
	}
	if (every) { mu_loop(std::true_type()); }
	else { mu_loop(std::false_type()); }

	// per-tile constants, so that no output element needs a division: the rows of a tile share the
	// largest key tk3 (tetrahedral store; Cp and p_base are multiples of BM) and start at (m2, m3) =
	// (hm2, hm3) (half-stored first mode)
	int64_t tk3 = 0, tk2 = 0, hm2 = 0, hm3 = 0;
	if (a.tet_C1 > 0) {
		const int64_t gp0 = a.p_base + p0;
		tk3 = gp0 / a.Cp;
		tk2 = gp0 - tk3 * a.Cp;
	}
	if (a.half_n > 0) {
		hm2 = p0 / a.half_n;
		hm3 = p0 - hm2 * a.half_n;
	}
	// accumulator layout of the 16x16 FP64 block (v_mfma_f64_16x16x4f64): lane holds
	// column lane%16 and rows lane/16 + 4 v, v = 0..3
#pragma unroll
	for (int i = 0; i < TM; ++i) {
#pragma unroll
		for (int j = 0; j < TN; ++j) {
			const int col = k0 + wn * (BN / 2) + j * 16 + lane % 16;
			// tet_index(col, k2, tk3) = tet_index(col, 0, tk3) + k2 (k2 + 1) / 2
			const int64_t tbase = (a.tet_C1 > 0) ? tet_index(col, 0, (int)tk3) : 0;
#pragma unroll
			for (int v = 0; v < 4; ++v) {
				const int d = wm * (BM / 2) + i * 16 + (lane / 16) + 4 * v;   // row - p0
				const int64_t row = p0 + d;
				if (row >= a.P || col >= a.C) { continue; }
				if (a.half_n > 0) {
					int64_t m3 = hm3 + d, m2 = hm2;
					while (m3 >= a.half_n) {
						m3 -= a.half_n;
						++m2;
					}
					if (m3 < m2) { continue; }   // never read
				}
				if (a.tet_C1 > 0) {
					const int64_t k2 = tk2 + d;
					if (col <= k2 && k2 <= tk3 && tk3 < a.tet_C1) { a.Y[tbase + k2 * (k2 + 1) / 2] = acc[i][j][v]; }
				}
				else {
					a.Y[row * a.ldy + col] = acc[i][j][v];
				}
			}
		}
	}
}

template <int BM, int BN, int GBK, bool kCp = false>
__global__ __launch_bounds__(256) void k_mode_product(ModeArgs a, TileOrder order)
{
	__shared__ double As[GBK][BM];
	__shared__ double Bs[GBK][BN];
	int rt, ct;
	if (!order.tile(blockIdx.x, rt, ct)) { return; }
	mode_product_tile<BM, BN, GBK, kCp>(a, rt, ct, As, Bs);
}

// FP32 variants of the mode product (v_mfma_f32_16x16x4f32, FP32 accumulation over the whole mu
// range, FP64 output), the comparison paths of the precision study; neither reaches the 1e-10 target.
//   kSplit = false: operands rounded to FP32, one product per term;
//   kSplit = true:  operands split x = hi + lo with hi = fl32(x), lo = fl32(x - hi); three products
//                   hi*hi + hi*lo + lo*hi per term into the same FP32 accumulator, so the operand
//                   rounding is corrected and the FP32 accumulation error remains.
typedef float f4v __attribute__((ext_vector_type(4)));

template <bool kSplit>
__global__ __launch_bounds__(256) void k_mode_product_f32(ModeArgs a, TileOrder order)
{
	__shared__ float As[GBK][64];
	__shared__ float Bs[GBK][64];
	__shared__ float Al[kSplit ? GBK : 1][64];
	__shared__ float Bl[kSplit ? GBK : 1][64];
	int rt, ct;
	if (!order.tile(blockIdx.x, rt, ct)) { return; }
	const int tid = threadIdx.x;
	const int lane = tid & 63, wave = tid >> 6;
	const int64_t p0 = (int64_t)rt * 64;
	const int k0 = ct * 64;
	if (a.tri_cols || a.tri_rows) {
		const int64_t gp = a.p_base + p0;
		const int64_t outer = gp / a.Cp;
		const int64_t rlo = gp % a.Cp, rhi = rlo + 63;
		if (a.tri_cols && k0 > rhi) { return; }
		if (a.tri_rows && rlo > outer) { return; }
	}
	const int wm = wave >> 1, wn = wave & 1;
	f4v acc[2][2];
	for (int i = 0; i < 2; ++i) {
		for (int j = 0; j < 2; ++j) { acc[i][j] = f4v{0, 0, 0, 0}; }
	}
	for (int mu0 = 0; mu0 < a.Kd; mu0 += GBK) {
		for (int r = 0; r < (GBK * 64) / 256; ++r) {
			const int id = tid + 256 * r;
			const int mu = id / 64, q = id % 64;
			const double x = (mu0 + mu < a.Kd && p0 + q < a.P) ? a.X[(int64_t)(mu0 + mu) * a.ldx + p0 + q] : 0.0;
			const double y = (mu0 + mu < a.Kd && k0 + q < a.C) ? a.B[(int64_t)(mu0 + mu) * a.ldb + k0 + q] : 0.0;
			As[mu][q] = (float)x;
			Bs[mu][q] = (float)y;
			if (kSplit) {
				Al[mu][q] = (float)(x - (double)(float)x);
				Bl[mu][q] = (float)(y - (double)(float)y);
			}
		}
		__syncthreads();
		for (int kk = 0; kk < GBK; kk += 4) {
			float av[2], bv[2];
			for (int i = 0; i < 2; ++i) { av[i] = As[kk + lane / 16][wm * 32 + i * 16 + lane % 16]; }
			for (int j = 0; j < 2; ++j) { bv[j] = Bs[kk + lane / 16][wn * 32 + j * 16 + lane % 16]; }
			if (kSplit) {
				float al[2], bl[2];
				for (int i = 0; i < 2; ++i) { al[i] = Al[kk + lane / 16][wm * 32 + i * 16 + lane % 16]; }
				for (int j = 0; j < 2; ++j) { bl[j] = Bl[kk + lane / 16][wn * 32 + j * 16 + lane % 16]; }
				for (int i = 0; i < 2; ++i) {
					for (int j = 0; j < 2; ++j) {
						acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(al[i], bv[j], acc[i][j], 0, 0, 0);
						acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bl[j], acc[i][j], 0, 0, 0);
					}
				}
			}
			for (int i = 0; i < 2; ++i) {
				for (int j = 0; j < 2; ++j) { acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bv[j], acc[i][j], 0, 0, 0); }
			}
		}
		__syncthreads();
	}
	// FP32 16x16 accumulator layout: lane holds column lane%16 and rows 4*(lane/16) + v
	int64_t tk3 = 0, tk2 = 0;   // the tile's rows share the largest key tk3 (tetrahedral store)
	if (a.tet_C1 > 0) {
		tk3 = (a.p_base + p0) / a.Cp;
		tk2 = a.p_base + p0 - tk3 * a.Cp;
	}
	for (int i = 0; i < 2; ++i) {
		for (int j = 0; j < 2; ++j) {
			const int col = k0 + wn * 32 + j * 16 + lane % 16;
			const int64_t tbase = (a.tet_C1 > 0) ? tet_index(col, 0, (int)tk3) : 0;
			for (int v = 0; v < 4; ++v) {
				const int d = wm * 32 + i * 16 + 4 * (lane / 16) + v;
				const int64_t row = p0 + d;
				if (row >= a.P || col >= a.C) { continue; }
				const double val = (double)acc[i][j][v];
				if (a.tet_C1 > 0) {
					const int64_t k2 = tk2 + d;
					if (col <= k2 && k2 <= tk3 && tk3 < a.tet_C1) { a.Y[tbase + k2 * (k2 + 1) / 2] = val; }
				}
				else {
					a.Y[row * a.ldy + col] = val;
				}
			}
		}
	}
}

// several independent mode products (one per grid of a batch) in one launch; each problem's
// workgroups start at a multiple of 8 so that the XCD-grouped order keeps its meaning
constexpr int kMaxGroup = 16;
struct ModeGroup {
	ModeArgs a[kMaxGroup];
	TileOrder order[kMaxGroup];
	int block_begin[kMaxGroup + 1];
	int n;
};

template <int BM, int BN, int GBK, bool kCp = false>
__global__ __launch_bounds__(256) void k_mode_product_grouped(ModeGroup grp)
{
	__shared__ double As[GBK][BM];
	__shared__ double Bs[GBK][BN];
	int g = 0;
	while (g + 1 < grp.n && (int)blockIdx.x >= grp.block_begin[g + 1]) { ++g; }
	int rt, ct;
	if (!grp.order[g].tile((int)blockIdx.x - grp.block_begin[g], rt, ct)) { return; }
	mode_product_tile<BM, BN, GBK, kCp>(grp.a[g], rt, ct, As, Bs);
}

// ---------------------------------------------------------------------------
// Tetrahedral rank-Q product of the separable method, V[k1 <= k2 <= k3] = sum_q G[q][k1] G[q][k2] U[q][k3],
// on the tetrahedron only.
//
// Work unit: one wave per (k2 block, k1 block) pair of 16 x 16 keys with k1 block <= k2 block, for a chunk
// of 16 consecutive k3 (chunk c needs the pairs with k2 block <= c: (c + 1)(c + 2) / 2 of them).  The
// wave keeps its pair's two G fragments in registers for the whole chunk; per k3 only the A operand's
// scaling by U[q][k3] changes (one multiply per MFMA), and two k3 run as independent accumulator chains.
// A workgroup of 4 waves takes 4 pairs of one chunk and stages that chunk's U columns in LDS.  The
// MFMAs run over S = ceil(Q / 4) steps of 4 nodes (rows q >= Q of G and U are zero).
//
// v_mfma_f64_16x16x4f64 operands: lane l holds A[l % 16][l / 16] (row k2, node 4s + l / 16) and
// B[l / 16][l % 16] (node, column k1); the accumulator holds rows l / 16 + 4 v and column l % 16.
// Each value is the sum of the same products in the same order as the 64 x 64 tile kernel's (A = G U
// rounded once, MFMA steps over ascending q, the tile kernel's extra steps add exact zeros), so V is
// bit-identical to it.
struct RankQProblem {
	const double* G;     // [q][key], row stride Cp
	const double* U;     // [q][key], row stride Cp
	double* V;           // tetrahedral key tensor (tet_index)
	int Cp, C1;          // Cp >= 16 ceil(C1 / 16): the chunk's U columns past C1 are zero padding
	int wg_begin;        // first workgroup of this problem in the launch
};
struct RankQGroup {
	RankQProblem p[kMaxGroup];
	int n;
};

__host__ __device__ inline int rankq_pairs(int c) { return (c + 1) * (c + 2) / 2; }
__host__ __device__ inline int rankq_chunk_wgs(int c) { return (rankq_pairs(c) + 3) / 4; }

template <int S>
__global__ __launch_bounds__(256) void k_rankq_tet(RankQGroup grp)
{
	__shared__ double us[4 * S][16];
	int gi = 0;
	while (gi + 1 < grp.n && (int)blockIdx.x >= grp.p[gi + 1].wg_begin) { ++gi; }
	const RankQProblem pr = grp.p[gi];
	int b = (int)blockIdx.x - pr.wg_begin;
	const int nch = (pr.C1 + 15) / 16;
	int c = 0;
	for (; c < nch; ++c) {
		const int nw = rankq_chunk_wgs(c);
		if (b < nw) { break; }
		b -= nw;
	}
	if (c >= nch) { return; }   // whole workgroup past its problem's chunks
	const int k3lo = 16 * c, nk3 = min(16, pr.C1 - k3lo);
	for (int i = threadIdx.x; i < 4 * S * 16; i += 256) {
		const int q = i >> 4, j = i & 15;
		us[q][j] = pr.U[(int64_t)q * pr.Cp + k3lo + j];
	}
	__syncthreads();
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	const int p = 4 * b + wave;
	if (p >= rankq_pairs(c)) { return; }
	int b2 = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
	while (rankq_pairs(b2) <= p) { ++b2; }
	while (b2 > 0 && rankq_pairs(b2 - 1) > p) { --b2; }
	const int b1 = p - b2 * (b2 + 1) / 2;
	const int r = lane & 15, g = lane >> 4;
	double ga[S], gb[S];
#pragma unroll
	for (int s = 0; s < S; ++s) {
		const int64_t row = (int64_t)(4 * s + g) * pr.Cp;
		ga[s] = pr.G[row + 16 * b2 + r];
		gb[s] = pr.G[row + 16 * b1 + r];
	}
	const int k1 = 16 * b1 + r;
	for (int j = 0; j < nk3; j += 2) {
		d4 acc0 = d4{0, 0, 0, 0}, acc1 = d4{0, 0, 0, 0};
#pragma unroll
		for (int s = 0; s < S; ++s) {
			const double u0 = us[4 * s + g][j], u1 = us[4 * s + g][j + 1];
			acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(ga[s] * u0, gb[s], acc0, 0, 0, 0);
			acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ga[s] * u1, gb[s], acc1, 0, 0, 0);
		}
#pragma unroll
		for (int h = 0; h < 2; ++h) {
			const int k3 = k3lo + j + h;
			if (j + h >= nk3) { break; }
			const int64_t t3 = (int64_t)k3 * (k3 + 1) * (k3 + 2) / 6;
#pragma unroll
			for (int v = 0; v < 4; ++v) {
				const int k2 = 16 * b2 + g + 4 * v;
				if (k1 <= k2 && k2 <= k3) { pr.V[t3 + (int64_t)k2 * (k2 + 1) / 2 + k1] = h ? acc1[v] : acc0[v]; }
			}
		}
	}
}

// ---------------------------------------------------------------------------
// per-element decoding

struct GridDev {
	int n[3];       // points per axis
	int N;          // total points
	int R;          // max points per axis (key tables are built for spans < R)
	int64_t N2;     // N * N
	int64_t N4;     // N^4: four-indices at or above it are outside the grid
};

// packed per-point axis offsets: x | y << 10 | z << 20
__global__ void k_point_coords(int nx, int ny, int nz, int* __restrict__ out)
{
	const int idx = blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= nx * ny * nz) { return; }
	const int z = idx % nz, y = (idx / nz) % ny, x = idx / (nz * ny);
	out[idx] = x | (y << 10) | (z << 20);
}

// four-index -> (i, j, k, l); exact for four-indices below 2^53 via a corrected FP64 quotient
__device__ inline void decode_four(const GridDev& g, int64_t t, int& i, int& j, int& k, int& l)
{
	int64_t q = (int64_t)((double)t / (double)g.N2);
	int64_t r = t - q * g.N2;
	if (r < 0) { q -= 1; r += g.N2; }
	else if (r >= g.N2) { q += 1; r -= g.N2; }
	const unsigned N = (unsigned)g.N;
	i = (int)((unsigned)q / N);
	j = (int)((unsigned)q - (unsigned)i * N);
	k = (int)((unsigned)r / N);
	l = (int)((unsigned)r - (unsigned)k * N);
}

// key id of one axis: a = |s01|, b = |s23|, cd = (pi + pj) - (pk + pl)
__device__ inline int axis_key(const int* __restrict__ lut, int R, int pi, int pj, int pk, int pl)
{
	const int a = abs(pi - pj), b = abs(pk - pl), cd = (pi + pj) - (pk + pl);
	return lut[(a * R + b) * (4 * R - 3) + cd + 2 * (R - 1)];
}

// key id of each axis (x, y, z) of a four-index, and the axis shift pair (|s01|, |s23|) of each
__device__ inline void element_axis_keys(const GridDev& g, const int* __restrict__ lut, const int* __restrict__ coords,
                                         int64_t t, int k[3], int ab[3])
{
	int i, j, kk, l;
	decode_four(g, t, i, j, kk, l);
	const int ci = coords[i], cj = coords[j], ck = coords[kk], cl = coords[l];
	for (int a = 0; a < 3; ++a) {
		const int sh = 10 * a;
		const int pi = (ci >> sh) & 1023, pj = (cj >> sh) & 1023, pk = (ck >> sh) & 1023, pl = (cl >> sh) & 1023;
		k[a] = axis_key(lut, g.R, pi, pj, pk, pl);
		ab[a] = abs(pi - pj) * g.R + abs(pk - pl);
	}
}

// quotient a / d for 0 <= a < 2^24 from a float reciprocal and one correction step
__device__ inline int qdiv(int a, int d, float inv)
{
	int q = (int)((float)a * inv);
	const int r = a - q * d;
	if (r < 0) { q -= 1; }
	else if (r >= d) { q += 1; }
	return q;
}

struct DecodeConsts {
	double inv_N2;   // 1 / N^2
	float inv_N;     // 1 / N
	float inv_nz;    // 1 / n_z
	float inv_nyz;   // 1 / (n_y n_z)
	// Reciprocals rounded up by one ulp: inv_up = next_up(fl(1/d)) > 1/d, with inv_up = (1/d)(1 + delta),
	// 0 < delta <= 1.5 ulp (2^-52 in double, 2^-23 in float).  For a = q d + r, 0 <= r < d:
	//  * fl(a inv_up) >= q: the exact product exceeds a/d >= q, q is representable, rounding is monotone;
	//  * fl(a inv_up) < q + 1: a/d <= q + 1 - 1/d, and the product's relative excess (delta plus one
	//    rounding, below 2 ulp) times q + 1 stays under 1/d whenever (a + d) * 2 ulp < 1, i.e. for
	//    a + d < 2^21 in float and a + d < 2^51 in double.
	// So trunc(a * inv_up) is the exact quotient for a < 2^12 with divisors up to 4096 (float),
	// a < 2^24 with divisor N and a < 2^48 with divisor N^2 <= 2^24 (double); a CPU test checks the float
	// case exhaustively and the double cases at their worst numerators.
	double inv_N2_up, inv_N_up;
	float inv_nyz_up, inv_nz_up;
	int uniform_pairs;   // 1: grids with N <= 4096 and tet indices below 2^32 take the fast path
};

// tetrahedral V index of a four-index; exact arithmetic decode (no table loads but the key table)
__device__ inline int64_t gather_index(const GridDev& g, const DecodeConsts& dc, const short* __restrict__ lut, int64_t t)
{
	const int W = 4 * g.R - 3, off = 2 * (g.R - 1);
	if (g.N > 4096) {
		// large grids: integer division (the float quotients are exact only below 2^24)
		int pt[4];
		decode_four(g, t, pt[0], pt[1], pt[2], pt[3]);
		int c[4][3];
		for (int a = 0; a < 4; ++a) {
			c[a][2] = pt[a] % g.n[2];
			c[a][1] = (pt[a] / g.n[2]) % g.n[1];
			c[a][0] = pt[a] / (g.n[2] * g.n[1]);
		}
		int kk[3];
		for (int a = 0; a < 3; ++a) {
			kk[a] = lut[(abs(c[0][a] - c[1][a]) * g.R + abs(c[2][a] - c[3][a])) * W + (c[0][a] + c[1][a]) - (c[2][a] + c[3][a]) + off];
		}
		int x = kk[0], y = kk[1], z = kk[2], sw;
		if (x > y) { sw = x; x = y; y = sw; }
		if (y > z) { sw = y; y = z; z = sw; }
		if (x > y) { sw = x; x = y; y = sw; }
		return tet_index(x, y, z);
	}
	int64_t q = (int64_t)((double)t * dc.inv_N2);
	int64_t r = t - q * g.N2;
	if (r < 0) { q -= 1; r += g.N2; }
	else if (r >= g.N2) { q += 1; r -= g.N2; }
	const int qi = (int)q, ri = (int)r;
	const int i = qdiv(qi, g.N, dc.inv_N), k = qdiv(ri, g.N, dc.inv_N);
	const int p[4] = { i, qi - i * g.N, k, ri - k * g.N };
	const int nyz = g.n[1] * g.n[2];
	int c[4][3];
#pragma unroll
	for (int a = 0; a < 4; ++a) {
		const int x = qdiv(p[a], nyz, dc.inv_nyz);
		const int rem = p[a] - x * nyz;
		const int y = qdiv(rem, g.n[2], dc.inv_nz);
		c[a][0] = x;
		c[a][1] = y;
		c[a][2] = rem - y * g.n[2];
	}
	int kk[3];
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		kk[a] = lut[(abs(c[0][a] - c[1][a]) * g.R + abs(c[2][a] - c[3][a])) * W + (c[0][a] + c[1][a]) - (c[2][a] + c[3][a]) + off];
	}
	int x = kk[0], y = kk[1], z = kk[2], sw;
	if (x > y) { sw = x; x = y; y = sw; }
	if (y > z) { sw = y; y = z; z = sw; }
	if (x > y) { sw = x; x = y; y = sw; }
	return tet_index(x, y, z);
}

// the value of lane 0 in every lane (scalar registers)
__device__ inline int64_t uniform_i64(int64_t v)
{
	const int lo = __builtin_amdgcn_readfirstlane((int)(uint32_t)v);
	const int hi = __builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
	return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// packed coordinates of a point p < 2^12: exact float quotients (see DecodeConsts)
__device__ inline void point_xyz(unsigned p, const GridDev& g, const DecodeConsts& dc, unsigned& x, unsigned& y, unsigned& z)
{
	const unsigned nyz = (unsigned)(g.n[1] * g.n[2]);
	x = (unsigned)((float)p * dc.inv_nyz_up);
	const unsigned rem = p - __umul24(x, nyz);
	y = (unsigned)((float)rem * dc.inv_nz_up);
	z = rem - __umul24(y, (unsigned)g.n[2]);
}

// One gather problem: the unique four-indices of one grid and where their values go.  The
// optional scale (device pointer, one double) multiplies every value: the 1/a of a physical
// grid spacing a, applied on the device instead of by the caller.
struct GatherDesc {
	const int64_t* idx;
	int64_t num;
	GridDev g;
	DecodeConsts dc;
	const short* lut;
	const uint32_t* tet3;
	const double* V;
	double* out;
	const double* scale;
	int64_t piece_begin;   // first workgroup-sized piece of this problem in a grouped launch
	int nlut, C1;          // entries of the key table and of the tetrahedral offsets
	uint32_t* pos;         // key-tensor position of every element (written by k_positions, read by
	                       // k_gather_cached; kNoPosition for indices outside the grid), or nullptr
	const uint16_t* perm;  // with positions sorted inside windows: each position's place in its window
};
constexpr uint32_t kNoPosition = 0xffffffffu;

// Bytes of the key-table image: the tetrahedral offsets (C1 uint32, padded to a multiple of 4), then
// the 16-bit key table (nlut entries), zero-padded to a multiple of 16 bytes.  The image has this
// layout in device memory (16-byte aligned, starting at GatherDesc::tet3) and in LDS, so a workgroup
// copies it with one 16-byte load per lane and step instead of one 2-byte load per table entry.
__host__ __device__ inline int gather_lds_bytes(int nlut, int C1)
{
	return (4 * ((C1 + 3) & ~3) + 2 * nlut + 15) & ~15;
}
constexpr int kGatherLdsMax = 48 * 1024;

// the key-table image into LDS (11^3: 717 pieces of 16 bytes, one step of a 1024-lane workgroup)
__device__ inline void copy_table_image(const GatherDesc& d, void* lds)
{
	const uint4* __restrict__ src = (const uint4*)d.tet3;
	uint4* dst = (uint4*)lds;
	const int nvec = gather_lds_bytes(d.nlut, d.C1) >> 4;
	for (int q = threadIdx.x; q < nvec; q += blockDim.x) { dst[q] = src[q]; }
}

// Key-tensor positions of one wave's GATHER_EPT groups of 64 consecutive elements (-1 past the end of
// the list and for indices outside the grid).  lut and tet3 point to the descriptor's tables in global
// memory, or to their copies in LDS.  In a sorted list the leading pair (i, j) changes only every few
// thousand elements, so it is decoded once per wave; when every group of the wave shares it, each lane
// decodes only (k, l), with exact reciprocal multiplies, and the loads are phased across the groups
// (all key-table reads, then all tetrahedral offsets) so that the in-order load counter never makes
// one group's reads wait for another's.  Any other wave (a pair boundary, the tail, unsorted or
// out-of-range input, grids without 32-bit tetrahedral indices) takes the general per-lane decode.
// The positions go to one of two consumers inside the decode's branch: fast(uint32_t[GATHER_EPT])
// when the whole wave took the fast decode (every position valid and below 2^32), general(int64_t
// [GATHER_EPT]) otherwise (-1 past the end of the list and for indices outside the grid).
template <int GATHER_EPT, typename Fast, typename General>
__device__ inline void wave_positions(const GatherDesc& d, int64_t wave_base, const short* __restrict__ lut,
                                      const uint32_t* __restrict__ tet3, Fast&& fast_use, General&& general_use)
{
	const int lane = threadIdx.x & 63;
	const GridDev& g = d.g;
	const DecodeConsts& dc = d.dc;
	int64_t t[GATHER_EPT];
	if (wave_base + 64 * GATHER_EPT <= d.num) {
		// a wave wholly inside the list loads without per-element bounds (a wave-uniform test instead of a
		// 64-bit compare and an exec mask per load: 11^3 gather 1.128 -> 1.107 ms)
		const int64_t* __restrict__ src = d.idx + wave_base + lane;
#pragma unroll
		for (int u = 0; u < GATHER_EPT; ++u) { t[u] = src[u * 64]; }
	}
	else {
#pragma unroll
		for (int u = 0; u < GATHER_EPT; ++u) {
			const int64_t e = wave_base + u * 64 + lane;
			t[u] = (e < d.num) ? d.idx[e] : -1;
		}
	}
	// the wave's pair (i, j) from its first element
	const int W = 4 * g.R - 3, off = 2 * (g.R - 1);
	const int64_t t0 = uniform_i64(t[0]);
	int64_t pair_base = -1;   // (i N + j) N^2, or -1 when the fast path is off for this wave
	int base_key[3] = { 0, 0, 0 };
	if (dc.uniform_pairs && t0 >= 0 && t0 < g.N4) {
		const int64_t ij = (int64_t)((double)t0 * dc.inv_N2_up);
		const unsigned i = (unsigned)((double)(unsigned)ij * dc.inv_N_up);
		const unsigned j = (unsigned)ij - __umul24(i, (unsigned)g.N);
		unsigned ci[3], cj[3];
		point_xyz(i, g, dc, ci[0], ci[1], ci[2]);
		point_xyz(j, g, dc, cj[0], cj[1], cj[2]);
#pragma unroll
		for (int a = 0; a < 3; ++a) {
			const int dd = abs((int)ci[a] - (int)cj[a]);
			base_key[a] = __builtin_amdgcn_readfirstlane(dd * g.R * W + (int)(ci[a] + cj[a]) + off);
		}
		pair_base = uniform_i64(ij * g.N2);
	}
	bool fast = pair_base >= 0;
#pragma unroll
	for (int u = 0; u < GATHER_EPT; ++u) { fast = fast && (uint64_t)(t[u] - pair_base) < (uint64_t)g.N2; }
	if (__all(fast)) {
		const unsigned short* __restrict__ ulut = (const unsigned short*)lut;
		unsigned kk[GATHER_EPT][3];
#pragma unroll
		for (int u = 0; u < GATHER_EPT; ++u) {
			const unsigned kl = (unsigned)(t[u] - pair_base);
			const unsigned k = (unsigned)((double)kl * dc.inv_N_up);
			const unsigned l = kl - __umul24(k, (unsigned)g.N);
			unsigned ck[3], cl[3];
			point_xyz(k, g, dc, ck[0], ck[1], ck[2]);
			point_xyz(l, g, dc, cl[0], cl[1], cl[2]);
#pragma unroll
			for (int a = 0; a < 3; ++a) {
				// (|ci - cj| R + |ck - cl|) W + (ci + cj) - (ck + cl) + 2(R - 1): the lane part, then the
				// wave-uniform pair part as a scalar base of the table address.  The lane part is negative
				// when ck = cl, so it stays a signed int: an unsigned one would be zero-extended into a
				// 64-bit address when the table is in global memory
				const int lane_part = (int)__umul24(__usad(ck[a], cl[a], 0u), (unsigned)W) - (int)(ck[a] + cl[a]);
				kk[u][a] = (ulut + base_key[a])[lane_part];
			}
		}
		uint32_t v32[GATHER_EPT];
		unsigned k3[GATHER_EPT];
#pragma unroll
		for (int u = 0; u < GATHER_EPT; ++u) {
			// the sorted keys k1 <= k2 <= k3 (minimum, median of three, maximum)
			const unsigned a0 = kk[u][0], a1 = kk[u][1], a2 = kk[u][2];
			const unsigned k1 = min(min(a0, a1), a2), k2 = max(min(a0, a1), min(max(a0, a1), a2));
			k3[u] = max(max(a0, a1), a2);
			// tet_index without its k3 block offset (added from the table below); k2 (k2 + 1) < 2^32 for C1 < 2^15
			v32[u] = (__umul24(k2, k2 + 1u) >> 1) + k1;
		}
#pragma unroll
		for (int u = 0; u < GATHER_EPT; ++u) { v32[u] += tet3[k3[u]]; }
		fast_use(v32);
	}
	else {
		int64_t vi[GATHER_EPT];
#pragma unroll
		for (int u = 0; u < GATHER_EPT; ++u) { vi[u] = (t[u] >= 0 && t[u] < g.N4) ? gather_index(g, dc, lut, t[u]) : -1; }
		general_use(vi);
	}
}

// values[e] = V[k1][k2][k3] from the tetrahedral key tensor for one wave's GATHER_EPT groups of 64
// consecutive elements (coalesced index loads and value stores, several dependent load chains
// index -> key table -> V in flight per lane); NaN for indices outside the grid
template <int GATHER_EPT>
__device__ inline void gather_wave(const GatherDesc& d, int64_t wave_base, const short* __restrict__ lut,
                                   const uint32_t* __restrict__ tet3)
{
	const int lane = threadIdx.x & 63;
	const double* __restrict__ V = d.V;
	double v[GATHER_EPT];
	wave_positions<GATHER_EPT>(d, wave_base, lut, tet3,
		[&](const uint32_t* vi) {
#pragma unroll
			for (int u = 0; u < GATHER_EPT; ++u) { v[u] = V[vi[u]]; }
		},
		[&](const int64_t* vi) {
#pragma unroll
			for (int u = 0; u < GATHER_EPT; ++u) { v[u] = (vi[u] >= 0) ? V[vi[u]] : __builtin_nan(""); }
		});
	if (d.scale != nullptr) {
		const double sc = *d.scale;
#pragma unroll
		for (int u = 0; u < GATHER_EPT; ++u) { v[u] *= sc; }
	}
	if (wave_base + 64 * GATHER_EPT <= d.num) {
		double* __restrict__ dst = d.out + wave_base + lane;
#pragma unroll
		for (int u = 0; u < GATHER_EPT; ++u) { dst[u * 64] = v[u]; }
	}
	else {
#pragma unroll
		for (int u = 0; u < GATHER_EPT; ++u) {
			const int64_t e = wave_base + u * 64 + lane;
			if (e < d.num) { d.out[e] = v[u]; }
		}
	}
}

// Workgroup b runs on XCD b % 8.  With xcd_chunks > 0 each XCD walks its own contiguous chunk of
// xcd_chunks workgroup-sized pieces (of the concatenated lists in a grouped launch), so the V
// entries its resident waves touch (neighbouring elements share the pair (i, j) and with it most
// of their axis keys) stay in that XCD's L2 instead of being spread over all eight.
__device__ inline int64_t gather_piece(int64_t xcd_chunks)
{
	return (xcd_chunks > 0) ? (int64_t)(blockIdx.x & 7) * xcd_chunks + (blockIdx.x >> 3) : (int64_t)blockIdx.x;
}

// With kLds the workgroup first copies the key table and the tetrahedral offsets into LDS: their
// reads then no longer compete with the V reads for the vector L1's miss path (the bound of this
// kernel); the copy pays off from 2048 elements per workgroup on.
template <int GATHER_EPT, bool kLds>
__device__ inline void gather_block(const GatherDesc& d, int64_t wave_base)
{
	extern __shared__ __align__(16) uint32_t gather_lds[];
	const short* lut = d.lut;
	const uint32_t* tet3 = d.tet3;
	if (kLds) {
		uint32_t* st = gather_lds;
		short* sl = (short*)(gather_lds + ((d.C1 + 3) & ~3));
		copy_table_image(d, gather_lds);
		__syncthreads();
		lut = sl;
		tet3 = st;
	}
	if (wave_base >= d.num) { return; }
	gather_wave<GATHER_EPT>(d, wave_base, lut, tet3);
}

template <int GATHER_EPT, bool kLds>
__global__ __launch_bounds__(1024) void k_gather(GatherDesc d, int64_t xcd_chunks)
{
	const int64_t piece = gather_piece(xcd_chunks);
	const int64_t wave_base = (piece * (blockDim.x >> 6) + (threadIdx.x >> 6)) * (64 * GATHER_EPT);
	if (!kLds && wave_base >= d.num) { return; }
	if (kLds && piece * (blockDim.x >> 6) * (64 * GATHER_EPT) >= d.num) { return; }   // whole workgroup idle
	gather_block<GATHER_EPT, kLds>(d, wave_base);
}

// key-tensor positions of one list (the gather's decode, no V reads; kNoPosition for indices
// outside the grid), with the key tables in LDS
__global__ __launch_bounds__(256) void k_positions(GatherDesc d)
{
	extern __shared__ __align__(16) uint32_t gather_lds[];
	const int64_t wave_base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * (64 * 8);
	if ((int64_t)blockIdx.x * (blockDim.x >> 6) * (64 * 8) >= d.num) { return; }   // whole workgroup idle
	uint32_t* st = gather_lds;
	short* sl = (short*)(gather_lds + ((d.C1 + 3) & ~3));
	copy_table_image(d, gather_lds);
	__syncthreads();
	if (wave_base >= d.num) { return; }
	const int lane = threadIdx.x & 63;
	uint32_t pos[8];
	wave_positions<8>(d, wave_base, sl, st,
		[&](const uint32_t* vi) {
#pragma unroll
			for (int u = 0; u < 8; ++u) { pos[u] = vi[u]; }
		},
		[&](const int64_t* vi) {
#pragma unroll
			for (int u = 0; u < 8; ++u) { pos[u] = (vi[u] >= 0) ? (uint32_t)vi[u] : kNoPosition; }
		});
#pragma unroll
	for (int u = 0; u < 8; ++u) {
		const int64_t e = wave_base + u * 64 + lane;
		if (e < d.num) { d.pos[e] = pos[u]; }
	}
}

// values from cached positions: per element a 4-byte position read, the V read, the scale and the
// 8-byte store; all grids of a batch in one launch, pieces as in k_gather_grouped
template <int GATHER_EPT>
__global__ __launch_bounds__(256) void k_gather_cached_grouped(const GatherDesc* __restrict__ descs, int n, int64_t xcd_chunks)
{
	const int64_t piece = gather_piece(xcd_chunks);
	int gi = 0;
	while (gi + 1 < n && piece >= descs[gi + 1].piece_begin) { ++gi; }
	const GatherDesc d = descs[gi];
	const int64_t first = (piece - d.piece_begin) * (blockDim.x >> 6) * (64 * GATHER_EPT);
	if (piece < d.piece_begin || first >= d.num) { return; }
	const int lane = threadIdx.x & 63;
	const int64_t wave_base = first + (threadIdx.x >> 6) * (64 * GATHER_EPT);
	uint32_t q[GATHER_EPT];
#pragma unroll
	for (int u = 0; u < GATHER_EPT; ++u) {
		const int64_t e = wave_base + u * 64 + lane;
		q[u] = (e < d.num) ? d.pos[e] : kNoPosition;
	}
	double v[GATHER_EPT];
#pragma unroll
	for (int u = 0; u < GATHER_EPT; ++u) {
		const bool ok = (q[u] != kNoPosition);   // a select on the address keeps the load out of a branch
		const double x = d.V[ok ? q[u] : 0u];
		v[u] = ok ? x : __builtin_nan("");
	}
	if (d.scale != nullptr) {
		const double sc = *d.scale;
#pragma unroll
		for (int u = 0; u < GATHER_EPT; ++u) { v[u] *= sc; }
	}
#pragma unroll
	for (int u = 0; u < GATHER_EPT; ++u) {
		const int64_t e = wave_base + u * 64 + lane;
		if (e < d.num) { d.out[e] = v[u]; }
	}
}

// Positions sorted inside windows of kSortWin consecutive elements (a plan's one-time pass): the
// gather then visits each window in key-tensor order, so the 64 reads of a wave fall on fewer
// cache lines, and stages the values in LDS (64 KB per workgroup) to store them in list order.
constexpr int kSortThreads = 1024;
constexpr int kSortWin = 8192;   // = 1024 lanes x 8 elements; as gather pieces: 256 lanes x 32 elements
constexpr int kSortEpt = kSortWin / 256;   // the elements-per-lane value whose pieces are the windows

// sort each window's positions ascending (bitonic, in LDS); perm receives each one's place
__global__ __launch_bounds__(kSortThreads) void k_sort_windows(uint32_t* __restrict__ pos, uint16_t* __restrict__ perm, int64_t num)
{
	__shared__ unsigned long long key[kSortWin];
	const int64_t base = (int64_t)blockIdx.x * kSortWin;
	const int cnt = (int)min((int64_t)kSortWin, num - base);
	for (int i = threadIdx.x; i < kSortWin; i += blockDim.x) {
		key[i] = (i < cnt) ? (((unsigned long long)pos[base + i] << 16) | (unsigned)i) : ~0ull;
	}
	__syncthreads();
	for (int k = 2; k <= kSortWin; k <<= 1) {
		for (int j = k >> 1; j > 0; j >>= 1) {
			for (int i = threadIdx.x; i < kSortWin; i += blockDim.x) {
				const int p = i ^ j;
				if (p > i) {
					const unsigned long long a = key[i], b = key[p];
					if ((a > b) == ((i & k) == 0)) { key[i] = b; key[p] = a; }
				}
			}
			__syncthreads();
		}
	}
	for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
		const unsigned long long v = key[i];
		pos[base + i] = (uint32_t)(v >> 16);
		perm[base + i] = (uint16_t)(v & 0xffffu);
	}
}

// values from window-sorted positions; all grids of a batch in one launch (pieces = windows)
__global__ __launch_bounds__(kSortThreads) void k_gather_sorted_grouped(const GatherDesc* __restrict__ descs, int n, int64_t xcd_chunks)
{
	__shared__ double so[kSortWin];
	const int64_t piece = gather_piece(xcd_chunks);
	int gi = 0;
	while (gi + 1 < n && piece >= descs[gi + 1].piece_begin) { ++gi; }
	const GatherDesc d = descs[gi];
	const int64_t base = (piece - d.piece_begin) * kSortWin;
	if (piece < d.piece_begin || base >= d.num) { return; }   // whole workgroup idle
	const int cnt = (int)min((int64_t)kSortWin, d.num - base);
	constexpr int E = kSortWin / kSortThreads;
	// a whole window (every workgroup but a list's last) loads and stores without per-element bounds
	const bool full = (cnt == kSortWin);
	uint32_t q[E];
	uint16_t p[E];
	if (full) {
#pragma unroll
		for (int r = 0; r < E; ++r) {
			q[r] = d.pos[base + r * kSortThreads + threadIdx.x];
			p[r] = d.perm[base + r * kSortThreads + threadIdx.x];
		}
	}
	else {
#pragma unroll
		for (int r = 0; r < E; ++r) {
			const int i = r * kSortThreads + threadIdx.x;
			q[r] = (i < cnt) ? d.pos[base + i] : kNoPosition;
			p[r] = (i < cnt) ? d.perm[base + i] : 0;
		}
	}
	double v[E];
#pragma unroll
	for (int r = 0; r < E; ++r) {
		const bool ok = (q[r] != kNoPosition);   // a select on the address keeps the load out of a branch
		const double x = d.V[ok ? q[r] : 0u];
		v[r] = ok ? x : __builtin_nan("");
	}
	const double sc = (d.scale != nullptr) ? *d.scale : 1.0;
#pragma unroll
	for (int r = 0; r < E; ++r) {
		const int i = r * kSortThreads + threadIdx.x;
		if (i < cnt) { so[p[r]] = (d.scale != nullptr) ? v[r] * sc : v[r]; }
	}
	__syncthreads();
	if (full) {
#pragma unroll
		for (int r = 0; r < E; ++r) { d.out[base + r * kSortThreads + threadIdx.x] = so[r * kSortThreads + threadIdx.x]; }
	}
	else {
#pragma unroll
		for (int r = 0; r < E; ++r) {
			const int i = r * kSortThreads + threadIdx.x;
			if (i < cnt) { d.out[base + i] = so[i]; }
		}
	}
}

template <int GATHER_EPT, bool kLds>
__global__ __launch_bounds__(256) void k_gather_grouped(const GatherDesc* __restrict__ descs, int n, int64_t xcd_chunks)
{
	const int64_t piece = gather_piece(xcd_chunks);
	int gi = 0;
	while (gi + 1 < n && piece >= descs[gi + 1].piece_begin) { ++gi; }
	const GatherDesc d = descs[gi];
	const int64_t first = (piece - d.piece_begin) * (blockDim.x >> 6) * (64 * GATHER_EPT);
	if (piece < d.piece_begin || first >= d.num) { return; }   // whole workgroup idle
	gather_block<GATHER_EPT, kLds>(d, first + (threadIdx.x >> 6) * (64 * GATHER_EPT));
}

// ---------------------------------------------------------------------------
// Single-call gather with the key-tensor reads grouped by cache line inside windows.
//
// In list order, consecutive elements change their z key almost every time, so the 64 lanes of a
// load read 64 scattered key-tensor entries and touch about 45 distinct 128-byte lines: the L1 miss
// path, not HBM, bounds the plain gather.  Here a workgroup of 1024 lanes takes a window of 8192
// consecutive elements, decodes every element's key-tensor position (wave_positions, as k_gather),
// and counting-sorts the window by bucket = position >> shift in LDS: one LDS atomic per element
// gives its rank inside its bucket, a scan of the bucket counts gives the bucket starts.  Reading
// the key tensor in bucket order puts the elements that share a line into the same 64-lane load.
// The values return to their list places through LDS and leave in coalesced stores.  Each value is
// the same key-tensor entry k_gather reads, so the results are bit-identical.
constexpr int kBucketEpt = 8;           // elements per lane: each wave takes 8 groups of 64 consecutive elements
constexpr int kBucketMaxThreads = 1024;

// LDS bytes of the bucketed gather's region for workgroups of `threads` lanes: it holds the values of
// the window (8 bytes per element); before that, the sort buffers (4-byte position and 2-byte place
// per element, one 4-byte bin per 2 elements); before that, the key tables
__host__ __device__ constexpr int bucket_lds_bytes(int threads)
{
	return threads * kBucketEpt * (int)sizeof(double);
}

// One workgroup of kThreads lanes per window of kThreads * 8 elements; bins <= 4 * kThreads buckets.
template <int kThreads>
__global__ __launch_bounds__(kThreads) void k_gather_bucketed(GatherDesc d, int64_t xcd_chunks, int shift, int bins)
{
	constexpr int kWin = kThreads * kBucketEpt;
	__shared__ __align__(16) unsigned char lds[bucket_lds_bytes(kThreads)];
	__shared__ unsigned wave_tot[kThreads / 64];
	const int64_t base = gather_piece(xcd_chunks) * kWin;
	if (base >= d.num) { return; }   // whole workgroup idle
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	// phase 1: key tables into LDS, then every element's position (registers)
	uint32_t* st = (uint32_t*)lds;
	short* sl = (short*)(lds + 4 * ((d.C1 + 3) & ~3));
	copy_table_image(d, lds);
	__syncthreads();
	uint32_t pos[kBucketEpt];
	const int first = wave * (64 * kBucketEpt);   // this wave's first place in the window
	wave_positions<kBucketEpt>(d, base + first, sl, st,
		[&](const uint32_t* vi) {
#pragma unroll
			for (int u = 0; u < kBucketEpt; ++u) { pos[u] = vi[u]; }
		},
		[&](const int64_t* vi) {
#pragma unroll
			for (int u = 0; u < kBucketEpt; ++u) { pos[u] = (vi[u] >= 0) ? (uint32_t)vi[u] : kNoPosition; }
		});
	__syncthreads();   // the decode tables are no longer read: the region becomes the sort buffers
	// phase 2: counting sort of the window's positions by bucket = position >> shift
	uint32_t* spos = (uint32_t*)lds;                         // [kWin] positions in bucket order
	uint16_t* splace = (uint16_t*)(lds + 4 * kWin);         // [kWin] their places in the window
	unsigned* hist = (unsigned*)(lds + 6 * kWin);           // [4 * kThreads] counts, then starts
	for (int q = tid; q < bins; q += kThreads) { hist[q] = 0; }
	__syncthreads();
	unsigned rank[kBucketEpt];
#pragma unroll
	for (int u = 0; u < kBucketEpt; ++u) {
		rank[u] = (pos[u] != kNoPosition) ? atomicAdd(&hist[pos[u] >> shift], 1u) : 0u;
	}
	__syncthreads();
	// exclusive scan of the counts: lane tid owns bins 4 tid .. 4 tid + 3
	unsigned c[4];
#pragma unroll
	for (int q = 0; q < 4; ++q) { c[q] = (4 * tid + q < bins) ? hist[4 * tid + q] : 0u; }
	const unsigned own = c[0] + c[1] + c[2] + c[3];
	unsigned inc = own;
#pragma unroll
	for (int o = 1; o < 64; o <<= 1) {
		const unsigned y = __shfl_up(inc, o, 64);
		if (lane >= o) { inc += y; }
	}
	if (lane == 63) { wave_tot[wave] = inc; }
	__syncthreads();
	unsigned ex = inc - own, nvalid = 0;
#pragma unroll
	for (int w = 0; w < kThreads / 64; ++w) {
		const unsigned wt = wave_tot[w];
		ex += (w < wave) ? wt : 0u;
		nvalid += wt;
	}
#pragma unroll
	for (int q = 0; q < 4; ++q) {
		if (4 * tid + q < bins) { hist[4 * tid + q] = ex; }
		ex += c[q];
	}
	__syncthreads();
#pragma unroll
	for (int u = 0; u < kBucketEpt; ++u) {
		if (pos[u] != kNoPosition) {
			const unsigned slot = hist[pos[u] >> shift] + rank[u];
			spos[slot] = pos[u];
			splace[slot] = (uint16_t)(first + u * 64 + lane);
		}
	}
	__syncthreads();
	// phase 3: key-tensor reads in bucket order, values to their places, coalesced stores
	uint32_t qp[kBucketEpt];
	unsigned ql[kBucketEpt];
#pragma unroll
	for (int r = 0; r < kBucketEpt; ++r) {
		const unsigned sidx = (unsigned)(r * kThreads + tid);
		qp[r] = (sidx < nvalid) ? spos[sidx] : kNoPosition;
		ql[r] = (sidx < nvalid) ? splace[sidx] : 0u;
	}
	__syncthreads();   // the sort buffers are read: the region holds the values now
	double* vals = (double*)lds;
	double v[kBucketEpt];
#pragma unroll
	for (int r = 0; r < kBucketEpt; ++r) { v[r] = (qp[r] != kNoPosition) ? d.V[qp[r]] : 0.0; }
	const double sc = (d.scale != nullptr) ? *d.scale : 1.0;
#pragma unroll
	for (int r = 0; r < kBucketEpt; ++r) {
		if (qp[r] != kNoPosition) { vals[ql[r]] = (d.scale != nullptr) ? v[r] * sc : v[r]; }
	}
#pragma unroll
	for (int u = 0; u < kBucketEpt; ++u) {
		if (pos[u] == kNoPosition) { vals[first + u * 64 + lane] = __builtin_nan(""); }
	}
	__syncthreads();
#pragma unroll
	for (int u = 0; u < kBucketEpt; ++u) {
		const int64_t e = base + first + u * 64 + lane;
		if (e < d.num) { d.out[e] = vals[first + u * 64 + lane]; }
	}
}

// ---------------------------------------------------------------------------
// Single-call gather with each wave's key-tensor reads grouped by line class.
//
// A wave takes 8 groups of 64 consecutive elements (512 elements) and decodes their key-tensor
// positions as k_gather does.  Its 8 key-tensor loads then take the elements class by class, with
// class = (position / 16) mod 8, the residue of the element's 128-byte line: elements that share a
// line share a class, so they fall into one 64-lane load instead of being spread over eight.  Over
// the 9^3 list the 8 loads of a wave touch 45 distinct lines each in list order and 29 grouped by
// class (a full sort of the 512 would give 25).  The permutation is a counting sort inside the
// wave, with no workgroup barrier: per-lane class counts in nibbles, a DPP prefix sum over the
// lanes, and one 8-byte LDS slot per element that holds (position, place in the wave).  The values
// return to their list places through the same slots and leave in coalesced stores.  Each value is
// the same key-tensor entry k_gather reads, so the results are bit-identical.
constexpr int kClassEpt = 8;

// inclusive prefix sum over the 64 lanes (row shifts inside each row of 16, then the row
// broadcasts of lanes 15 and 31); every lane must be active
__device__ inline unsigned wave_inclusive_sum(unsigned v)
{
	v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);   // row_shr:1
	v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);   // row_shr:2
	v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);   // row_shr:4
	v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);   // row_shr:8
	v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);   // row_bcast:15
	v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);   // row_bcast:31
	return v;
}

// LDS bytes of the classed gather for workgroups of `threads` lanes: the key tables (when kept in
// LDS), then 8 bytes per element of every wave
__host__ __device__ inline int classed_table_bytes(int nlut, int C1)
{
	return (gather_lds_bytes(nlut, C1) + 15) & ~15;
}
__host__ __device__ inline int classed_lds_bytes(int threads, int nlut, int C1, bool lds)
{
	return (lds ? classed_table_bytes(nlut, C1) : 0) + threads * kClassEpt * 8;
}

// one wave's 512 elements; slots: the wave's 512 LDS slots
__device__ inline void gather_wave_classed(const GatherDesc& d, int64_t wave_base, const short* __restrict__ lut,
                                           const uint32_t* __restrict__ tet3, uint2* slots)
{
	constexpr int E = kClassEpt;
	const int lane = threadIdx.x & 63;
	const double* __restrict__ V = d.V;
	double v[E];
	wave_positions<E>(d, wave_base, lut, tet3,
		[&](const uint32_t* p) {
			// class of each element, its rank among the lane's earlier elements of that class, and the
			// lane's count per class (nibbles: at most 8 per class)
			unsigned cl[E], rk[E], cnt = 0;
#pragma unroll
			for (int u = 0; u < E; ++u) {
				cl[u] = (p[u] >> 4) & 7u;
				rk[u] = (cnt >> (4 * cl[u])) & 15u;
				cnt += 1u << (4 * cl[u]);
			}
			// counts of classes 2r and 2r + 1 in the 16-bit halves of own[r]; the wave totals stay
			// <= 512, so the halves never carry into each other
			unsigned own[4], ex[4];
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				own[r] = ((cnt >> (8 * r)) & 15u) | (((cnt >> (8 * r + 4)) & 15u) << 16);
				ex[r] = wave_inclusive_sum(own[r]);
			}
			// class starts: exclusive prefix of the wave's class totals (lane 63's inclusive sums)
			unsigned run = 0;
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				const unsigned tot = (unsigned)__builtin_amdgcn_readlane((int)ex[r], 63);
				const unsigned lo = tot & 0xffffu;
				ex[r] = ex[r] - own[r] + (run | ((run + lo) << 16));
				run += lo + (tot >> 16);
			}
#pragma unroll
			for (int u = 0; u < E; ++u) {
				const unsigned c = cl[u];
				const unsigned w01 = (c & 2u) ? ex[1] : ex[0];
				const unsigned w23 = (c & 2u) ? ex[3] : ex[2];
				const unsigned w = (c & 4u) ? w23 : w01;
				const unsigned slot = (((w >> (16 * (c & 1u))) & 0xffffu) + rk[u]) & (unsigned)(E * 64 - 1);
				slots[slot] = make_uint2(p[u], (unsigned)(u * 64 + lane));
			}
			// a wave's LDS operations complete in issue order; the barrier keeps the compiler from
			// moving the reads above the writes of other lanes' slots
			__builtin_amdgcn_wave_barrier();
			uint2 q[E];
#pragma unroll
			for (int u = 0; u < E; ++u) { q[u] = slots[u * 64 + lane]; }
#pragma unroll
			for (int u = 0; u < E; ++u) { v[u] = V[q[u].x]; }
			__builtin_amdgcn_wave_barrier();
			double* vals = (double*)slots;
#pragma unroll
			for (int u = 0; u < E; ++u) { vals[q[u].y & (unsigned)(E * 64 - 1)] = v[u]; }
			__builtin_amdgcn_wave_barrier();
#pragma unroll
			for (int u = 0; u < E; ++u) { v[u] = vals[u * 64 + lane]; }
		},
		[&](const int64_t* vi) {
#pragma unroll
			for (int u = 0; u < E; ++u) { v[u] = (vi[u] >= 0) ? V[vi[u]] : __builtin_nan(""); }
		});
	if (d.scale != nullptr) {
		const double sc = *d.scale;
#pragma unroll
		for (int u = 0; u < E; ++u) { v[u] *= sc; }
	}
	if (wave_base + 64 * E <= d.num) {
		double* __restrict__ dst = d.out + wave_base + lane;
#pragma unroll
		for (int u = 0; u < E; ++u) { dst[u * 64] = v[u]; }
	}
	else {
#pragma unroll
		for (int u = 0; u < E; ++u) {
			const int64_t e = wave_base + u * 64 + lane;
			if (e < d.num) { d.out[e] = v[u]; }
		}
	}
}

// one workgroup of blockDim.x lanes per piece of 8 blockDim.x consecutive elements; the wave slots
// follow the key tables in dynamic LDS
template <bool kLds>
__global__ __launch_bounds__(1024) void k_gather_classed(GatherDesc d, int64_t xcd_chunks)
{
	extern __shared__ __align__(16) uint32_t gather_lds[];
	const int64_t piece = gather_piece(xcd_chunks);
	if (piece * (blockDim.x >> 6) * (64 * kClassEpt) >= d.num) { return; }   // whole workgroup idle
	const int64_t wave_base = (piece * (blockDim.x >> 6) + (threadIdx.x >> 6)) * (64 * kClassEpt);
	const short* lut = d.lut;
	const uint32_t* tet3 = d.tet3;
	int table_bytes = 0;
	if (kLds) {
		uint32_t* st = gather_lds;
		short* sl = (short*)(gather_lds + ((d.C1 + 3) & ~3));
		copy_table_image(d, gather_lds);
		__syncthreads();
		lut = sl;
		tet3 = st;
		table_bytes = classed_table_bytes(d.nlut, d.C1);
	}
	if (wave_base >= d.num) { return; }
	uint2* slots = (uint2*)((unsigned char*)gather_lds + table_bytes) + (threadIdx.x >> 6) * (64 * kClassEpt);
	gather_wave_classed(d, wave_base, lut, tet3, slots);
}

// ---------------------------------------------------------------------------
// per-element FP64 path: the reference's binned sum (eri_integrals.c:608-669), one workgroup per
// element, no reuse across elements.  On the folded columns B[.][kx], B[.][ky], B[.][kz]
// (t -> |t| is an exact regrouping of the reference's terms):
//   C[rho] = sum_{m1^2 + m2^2 = rho} Bx[m1] By[m2]                        (x, y bins)
//   value  = sum_{rho: C[rho] != 0} C[rho] * sum_{m3} Bz[m3] T[rho + m3^2]
// Bins are computed per rho from a table of its representations m1^2 + m2^2 = rho, so no bin is
// written twice and the result does not depend on scheduling; the nonzero bins of a chunk of rho
// are compacted into LDS and the z sums run with lanes over consecutive bins, so the table reads
// of a wave fall on a few lines.

constexpr int DIRECT_THREADS = 256;
constexpr int DIRECT_CHUNK = 2048;   // rho values per compaction round

struct DirectCount {
	unsigned long long* sums;      // [0] reference-form flops, [1] executed flops; nullptr: not counted
	const int* nnz_rows;           // nonzeros of P[a][b] per shift pair (a * R + b)
};

__device__ inline double wave_sum(double v)
{
	for (int o = 32; o > 0; o >>= 1) { v += __shfl_xor(v, o, 64); }
	return v;
}

__global__ __launch_bounds__(DIRECT_THREADS) void k_eri_direct(const int64_t* __restrict__ idx, int64_t num, GridDev g,
	const int* __restrict__ lut, const int* __restrict__ coords, const double* __restrict__ B, int Cp, int M, int Mp,
	const double* __restrict__ T, const int* __restrict__ rep_off, const uint32_t* __restrict__ rep_list,
	DirectCount cnt, double* __restrict__ out)
{
	extern __shared__ double smem[];
	double* col = smem;                                   // 3 x Mp folded columns
	double* list_c = smem + 3 * Mp;                       // compacted nonzero bins: values
	int* list_rho = (int*)(list_c + DIRECT_CHUNK);        //                          and rho
	__shared__ int win[3][2];
	__shared__ int wave_tot[DIRECT_THREADS / 64];
	__shared__ double wave_acc[DIRECT_THREADS / 64];
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	constexpr int J = DIRECT_CHUNK / DIRECT_THREADS;
	for (int64_t e = blockIdx.x; e < num; e += gridDim.x) {
		const int64_t t = idx[e];
		if (t < 0 || t >= g.N4) {
			if (tid == 0) { out[e] = __builtin_nan(""); }
			continue;
		}
		int k[3], ab[3];
		element_axis_keys(g, lut, coords, t, k, ab);
		if (tid < 3) {
			win[tid][0] = M + 1;
			win[tid][1] = -1;
		}
		__syncthreads();
		for (int a = 0; a < 3; ++a) {
			for (int m = tid; m < Mp; m += DIRECT_THREADS) {
				const double v = (m <= M) ? B[(int64_t)m * Cp + k[a]] : 0.0;
				col[a * Mp + m] = v;
				if (v != 0) {
					atomicMin(&win[a][0], m);
					atomicMax(&win[a][1], m);
				}
			}
		}
		__syncthreads();
		const int lo0 = win[0][0], hi0 = win[0][1], lo1 = win[1][0], hi1 = win[1][1], lo2 = win[2][0], hi2 = win[2][1];
		double acc = 0;
		int my_bins = 0;
		if (hi0 >= lo0 && hi1 >= lo1 && hi2 >= lo2) {
			const double* cx = col;
			const double* cy = col + Mp;
			const double* cz = col + 2 * Mp;
			const int rho_lo = lo0 * lo0 + lo1 * lo1, rho_hi = hi0 * hi0 + hi1 * hi1;
			for (int base = rho_lo; base <= rho_hi; base += DIRECT_CHUNK) {
				// bins of this chunk: thread tid takes rho = base + 64 * (J * wave + j) + lane
				double cv[J];
				uint64_t mask[J];
				int count = 0;
#pragma unroll
				for (int j = 0; j < J; ++j) {
					const int rho = base + 64 * (J * wave + j) + lane;
					double c = 0;
					if (rho <= rho_hi) {
						const int r0 = rep_off[rho], r1 = rep_off[rho + 1];
						for (int r = r0; r < r1; ++r) {
							const uint32_t pq = rep_list[r];
							const int m1 = (int)(pq & 0xffffu), m2 = (int)(pq >> 16);
							if (m1 >= lo0 && m1 <= hi0 && m2 >= lo1 && m2 <= hi1) { c += cx[m1] * cy[m2]; }
						}
					}
					cv[j] = c;
					mask[j] = __ballot(c != 0);
					count += __popcll(mask[j]);
				}
				if (lane == 0) { wave_tot[wave] = count; }
				__syncthreads();
				int pos = 0, total = 0;
				for (int w = 0; w < DIRECT_THREADS / 64; ++w) {
					pos += (w < wave) ? wave_tot[w] : 0;
					total += wave_tot[w];
				}
#pragma unroll
				for (int j = 0; j < J; ++j) {
					const uint64_t below = mask[j] & ((1ull << lane) - 1ull);
					if (cv[j] != 0) {
						const int at = pos + __popcll(below);
						list_c[at] = cv[j];
						list_rho[at] = base + 64 * (J * wave + j) + lane;
					}
					pos += __popcll(mask[j]);
				}
				__syncthreads();
				my_bins += (tid == 0) ? total : 0;
				// z sums over the compacted bins, lanes over consecutive bins
				// (zero z entries inside the window add exact zeros: T is finite)
				for (int q = tid; q < total; q += DIRECT_THREADS) {
					const double* Tr = T + list_rho[q];
					double s = 0;
#pragma unroll 4
					for (int m3 = lo2; m3 <= hi2; ++m3) { s += cz[m3] * Tr[m3 * m3]; }
					acc += s * list_c[q];
				}
				__syncthreads();
			}
		}
		acc = wave_sum(acc);
		if (lane == 0) { wave_acc[wave] = acc; }
		__syncthreads();
		if (tid == 0) {
			double v = 0;
			for (int w = 0; w < DIRECT_THREADS / 64; ++w) { v += wave_acc[w]; }
			out[e] = v;
			if (cnt.sums != nullptr && hi0 >= lo0 && hi1 >= lo1 && hi2 >= lo2) {
				const unsigned long long nx = cnt.nnz_rows[ab[0]], ny = cnt.nnz_rows[ab[1]], nz = cnt.nnz_rows[ab[2]];
				// executed: every in-window (m1, m2) pair, and per nonzero bin the z window + 1
				const unsigned long long R = (unsigned long long)my_bins;
				const unsigned long long wx = hi0 - lo0 + 1, wy = hi1 - lo1 + 1, ez = hi2 - lo2 + 1;
				atomicAdd(&cnt.sums[0], 2ull * nx * ny + 2ull * R * (nz + 1ull));
				atomicAdd(&cnt.sums[1], 2ull * wx * wy + 2ull * R * (ez + 1ull));
			}
		}
		__syncthreads();
	}
}

// nonzeros of each P[a][b] row (the reference's n per axis: entries with |P| > tol)
__global__ void k_nnz_rows(const double* __restrict__ P, int rows, int LP, int* __restrict__ nnz)
{
	const int r = blockIdx.x;
	if (r >= rows) { return; }
	int c = 0;
	for (int t = threadIdx.x; t < LP; t += blockDim.x) { c += (P[(int64_t)r * LP + t] != 0) ? 1 : 0; }
	for (int o = 32; o > 0; o >>= 1) { c += __shfl_xor(c, o, 64); }
	__shared__ int part[4];
	if ((threadIdx.x & 63) == 0) { part[threadIdx.x >> 6] = c; }
	__syncthreads();
	if (threadIdx.x == 0) { nnz[r] = part[0] + part[1] + part[2] + part[3]; }
}

// ---------------------------------------------------------------------------
// host side: key tables, workspace, orchestration

struct KeySet {
	int R = 0;
	bool symmetric = false;
	int C1 = 0;
	std::vector<int4> keys;  // (a, b, cd, 0) of each distinct folded column
	std::vector<int> lut;    // (a * R + b) * (4R - 3) + cd + 2(R - 1) -> key id
	std::vector<short> lut16;    // the same table as int16 (key ids stay below 32768)
	std::vector<uint32_t> tet3;  // tet_index(0, 0, k) per key
	std::vector<uint32_t> image; // the gather's key-table image: tet3, then lut16 (gather_lds_bytes long)
	bool tet32 = false;         // every tetrahedral index fits 32 bits
	bool on_device = false;     // the DevBuf copies below exist
	DevBuf<int4> keys_dev;
	DevBuf<int> lut_dev;
	DevBuf<uint32_t> image_dev;  // tet3 and lut16 on the device, as one image
};

// the key set of per-axis length R in host memory, built once per (R, symmetric) and kept
static KeySet& key_tables(int R, bool symmetric)
{
	static std::map<std::pair<int, bool>, KeySet*> cache;
	auto it = cache.find({R, symmetric});
	if (it != cache.end()) { return *it->second; }
	KeySet* ks = new KeySet();
	ks->R = R;
	ks->symmetric = symmetric;
	const int W = 4 * R - 3;
	ks->lut.assign((size_t)R * R * W, -1);
	// every per-axis 4-tuple with span < R (a translation class per key)
	auto canon_of = [&](int i, int j, int k, int l) {
		const int a = std::abs(i - j), b = std::abs(k - l), cd = (i + j) - (k + l);
		return symmetric ? std::make_tuple(std::min(a, b), std::max(a, b), std::abs(cd)) : std::make_tuple(a, b, cd);
	};
	std::map<std::tuple<int, int, int>, int> ids;
	for (int i = 0; i < R; ++i) {
		for (int j = 0; j < R; ++j) {
			for (int k = 0; k < R; ++k) {
				for (int l = 0; l < R; ++l) { ids.emplace(canon_of(i, j, k, l), 0); }
			}
		}
	}
	// Key numbering: ascending total size |s01| + |s23| + |cd|, then the larger shift.  Neighbouring
	// elements of a sorted list differ in one axis by a unit step of l, which changes |s23| and cd by one
	// each: their keys then have close numbers, so the key-tensor entries a wave reads share more
	// 128-byte lines (on the 11^3 list 186 instead of 233 distinct lines per 512 consecutive elements).
	std::vector<std::tuple<int, int, int>> order;
	for (auto& kv : ids) { order.push_back(kv.first); }
	std::stable_sort(order.begin(), order.end(), [](const std::tuple<int, int, int>& x, const std::tuple<int, int, int>& y) {
		const int sx = std::get<0>(x) + std::get<1>(x) + std::abs(std::get<2>(x));
		const int sy = std::get<0>(y) + std::get<1>(y) + std::abs(std::get<2>(y));
		if (sx != sy) { return sx < sy; }
		const int mx = std::max(std::get<0>(x), std::get<1>(x)), my = std::max(std::get<0>(y), std::get<1>(y));
		if (mx != my) { return mx < my; }
		return x < y;
	});
	for (size_t q = 0; q < order.size(); ++q) {
		ids[order[q]] = (int)q;
		ks->keys.push_back(int4{std::get<0>(order[q]), std::get<1>(order[q]), std::get<2>(order[q]), 0});
	}
	for (int i = 0; i < R; ++i) {
		for (int j = 0; j < R; ++j) {
			for (int k = 0; k < R; ++k) {
				for (int l = 0; l < R; ++l) {
					const int a = std::abs(i - j), b = std::abs(k - l), cd = (i + j) - (k + l);
					ks->lut[(a * R + b) * W + cd + 2 * (R - 1)] = ids[canon_of(i, j, k, l)];
				}
			}
		}
	}
	ks->C1 = (int)ks->keys.size();
	GLI_REQUIRE(ks->C1 < 32768, "too many axis keys for the 16-bit key table");
	ks->lut16.assign(ks->lut.begin(), ks->lut.end());
	// C(k3 + 2, 3): tetrahedral offset of the largest key, as long as every index fits 32 bits
	ks->tet32 = tet_index(0, 0, ks->C1) <= (int64_t)UINT32_MAX;
	ks->tet3.resize((size_t)ks->C1);
	for (int k = 0; k < ks->C1; ++k) { ks->tet3[(size_t)k] = ks->tet32 ? (uint32_t)tet_index(0, 0, k) : 0u; }
	ks->image.assign((size_t)gather_lds_bytes((int)ks->lut16.size(), ks->C1) / 4, 0u);
	std::memcpy(ks->image.data(), ks->tet3.data(), ks->tet3.size() * sizeof(uint32_t));
	std::memcpy(ks->image.data() + ((ks->C1 + 3) & ~3), ks->lut16.data(), ks->lut16.size() * sizeof(short));
	cache[{R, symmetric}] = ks;
	return *ks;
}

// the same key set with its device copies (allocated once per key set)
static KeySet& key_set(int R, bool symmetric)
{
	KeySet& ks = key_tables(R, symmetric);
	if (!ks.on_device) {
		ks.keys_dev.alloc(ks.keys.size());
		ks.lut_dev.alloc(ks.lut.size());
		ks.image_dev.alloc(ks.image.size());
		GLI_HIP(hipMemcpy(ks.keys_dev.p, ks.keys.data(), ks.keys.size() * sizeof(int4), hipMemcpyHostToDevice));
		GLI_HIP(hipMemcpy(ks.lut_dev.p, ks.lut.data(), ks.lut.size() * sizeof(int), hipMemcpyHostToDevice));
		GLI_HIP(hipMemcpy(ks.image_dev.p, ks.image.data(), ks.image.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
		ks.on_device = true;
	}
	return ks;
}

struct EriWorkspace {
	DevBuf<double> coeffs, P, B, Y1, Y2, V;
	// separable method: the Gaussian factors [g | W g] of the folded range sep_M (Mp x 2 Qp, cached)
	// and the rank-Q factors [G; U] of the current evaluation (2 Qp x Cp)
	DevBuf<double> sep_gw, GU;
	int sep_M = -1, sep_Qp = 0, sep_Q = 0;
	DevBuf<int> coords;
	DevBuf<double> T3;
	int T3_M = -1;
	// per-element path: representations m1^2 + m2^2 = rho for m1, m2 <= rep_M, grouped by rho
	DevBuf<int> rep_off;
	DevBuf<uint32_t> rep_list;
	int rep_M = -1;
	// per-element path flop counters
	bool count_direct = false;
	DevBuf<unsigned long long> direct_sums;
	DevBuf<int> nnz_rows;
};

static EriWorkspace& eri_ws()
{
	static EriWorkspace* w = new EriWorkspace();
	return *w;
}

static bool mirror_symmetric(const double* c, int64_t K)
{
	for (int64_t i = 0; i < K; ++i) {
		if (c[i] != c[K - 1 - i]) { return false; }
	}
	return true;
}

struct EriSetup {
	int K, R, M, Mp, C1, Cp;
	GridDev g;
	const KeySet* ks;
	// device copies of the Coulomb table and of the key set's tables (the library's cached buffers,
	// or the allocation-free entry point's scratch)
	const double* T;
	const int4* d_keys;
	const short* d_lut16;
	const uint32_t* d_tet3;
};

// shapes, key set and Coulomb table for K coefficients on a grid (no device work but the tables).
// device = false: host tables only, with every device pointer left null for the caller to set.
static EriSetup eri_shape(int64_t K, const int64_t grid[6], bool sym, bool device = true)
{
	GLI_REQUIRE(K >= 3 && (K % 2) == 1, "the number of gausslet coefficients must be odd and at least 3");
	GLI_REQUIRE(grid[1] > 0 && grid[3] > 0 && grid[5] > 0, "grid axes must have a positive number of points");
	GLI_REQUIRE(grid[1] < 1024 && grid[3] < 1024 && grid[5] < 1024, "at most 1023 points per axis");
	{
		const double N = (double)grid[1] * grid[3] * grid[5];
		GLI_REQUIRE(N * N * N * N < 9007199254740992.0, "grid too large: four-indices must stay below 2^53");
	}
	EriSetup s;
	s.K = (int)K;
	s.R = (int)std::max(grid[1], std::max(grid[3], grid[5]));
	s.M = 6 * (s.R - 1) + 2 * (s.K - 1);
	s.Mp = (int)round_up(s.M + 1, 16);
	GLI_REQUIRE(s.Mp <= 2048, "grid too long: the folded kernel range 6(n-1) + 2(K-1) must stay below 2048");
	const KeySet& ks = device ? key_set(s.R, sym) : key_tables(s.R, sym);
	s.ks = &ks;
	s.C1 = ks.C1;
	s.Cp = (int)round_up(s.C1, KEY_ALIGN);
	s.g.n[0] = (int)grid[1];
	s.g.n[1] = (int)grid[3];
	s.g.n[2] = (int)grid[5];
	s.g.N = (int)(grid[1] * grid[3] * grid[5]);
	s.g.R = s.R;
	s.g.N2 = (int64_t)s.g.N * s.g.N;
	s.g.N4 = s.g.N2 * s.g.N2;
	s.T = device ? eri_table(3 * (int64_t)s.M * s.M + 1) : nullptr;
	s.d_keys = device ? ks.keys_dev.p : nullptr;
	s.d_tet3 = device ? ks.image_dev.p : nullptr;   // the start of the key-table image
	s.d_lut16 = device ? (const short*)(ks.image_dev.p + ((ks.C1 + 3) & ~3)) : nullptr;
	return s;
}

// f, P (zeroed where |P| <= tol) and the folded key columns B from device-resident coefficients
// tables of one evaluation from device-resident coefficients (d_coeffs) or, with d_coeffs null, from
// host coefficients h_coeffs of at most kArgCoeffs entries passed as a kernel argument
static void launch_tables(const EriSetup& s, const double* d_coeffs, double tol, double* P, double* B,
                          const double* h_coeffs = nullptr)
{
	GLI_REQUIRE(tol >= 0, "tolerance must be non-negative");
	hipStream_t st = ctx().stream;
	const int L = 2 * s.K - 1;
	CoeffArg ca{};
	if (d_coeffs == nullptr) {
		GLI_REQUIRE(h_coeffs != nullptr && s.K <= kArgCoeffs, "coefficients by argument: at most 64 of them");
		for (int q = 0; q < s.K; ++q) { ca.c[q] = h_coeffs[q]; }
	}
	k_products<<<(unsigned)(s.R * s.R), 256, (size_t)(s.K + 3 * L) * sizeof(double), st>>>(d_coeffs, ca, s.K, s.R, tol, P);
	if (B != nullptr) {   // the separable method needs no folded columns (k_separable_tables)
		k_fold<<<dim3((unsigned)ceil_div(s.Cp, 128), s.Mp), 128, 0, st>>>(P, s.d_keys, s.C1, s.K, s.R, s.M, s.Cp, B);
	}
	GLI_HIP(hipGetLastError());
}

void kernel_tensor(const double* T, int M, int Mp, double* T3)
{
	k_kernel_tensor<<<dim3((unsigned)ceil_div(Mp, 64), (unsigned)Mp, (unsigned)Mp), 64, 0, ctx().stream>>>(T, M, Mp, T3);
	GLI_HIP(hipGetLastError());
}

static void launch_kernel_tensor(const EriSetup& s, double* T3)
{
	k_kernel_tensor<<<dim3((unsigned)ceil_div(s.Mp, 64), (unsigned)s.Mp, (unsigned)s.Mp), 64, 0, ctx().stream>>>(s.T, s.M, s.Mp, T3);
	GLI_HIP(hipGetLastError());
}

static DecodeConsts decode_consts(const EriSetup& s)
{
	DecodeConsts dc;
	dc.inv_N2 = 1.0 / (double)s.g.N2;
	dc.inv_N = 1.0f / (float)s.g.N;
	dc.inv_nz = 1.0f / (float)s.g.n[2];
	dc.inv_nyz = 1.0f / (float)(s.g.n[1] * s.g.n[2]);
	dc.inv_N2_up = std::nextafter(1.0 / (double)s.g.N2, HUGE_VAL);
	dc.inv_N_up = std::nextafter(1.0 / (double)s.g.N, HUGE_VAL);
	dc.inv_nyz_up = std::nextafterf(1.0f / (float)(s.g.n[1] * s.g.n[2]), HUGE_VALF);
	dc.inv_nz_up = std::nextafterf(1.0f / (float)s.g.n[2], HUGE_VALF);
	dc.uniform_pairs = (s.g.N <= 4096 && s.ks->tet32) ? 1 : 0;
	return dc;
}

static GatherDesc gather_desc(const EriSetup& s, const int64_t* d_idx, int64_t num, const double* V, double* out,
                              const double* scale)
{
	GatherDesc d;
	d.idx = d_idx;
	d.num = num;
	d.g = s.g;
	d.dc = decode_consts(s);
	d.lut = s.d_lut16;
	d.tet3 = s.d_tet3;
	d.V = V;
	d.out = out;
	d.scale = scale;
	d.piece_begin = 0;
	d.nlut = s.R * s.R * (4 * s.R - 3);
	d.C1 = s.C1;
	d.pos = nullptr;
	d.perm = nullptr;
	return d;
}

static int64_t gather_pieces(int64_t num, int ept)
{
	return ceil_div(num, 256 * (int64_t)ept);
}

static bool gather_lds_fits(const GatherDesc& d)
{
	return gather_lds_bytes(d.nlut, d.C1) <= kGatherLdsMax;
}

template <typename F>
static void gather_dispatch(int ept, bool lds, F&& f)
{
	switch (ept * 2 + (lds ? 1 : 0)) {
		case 4: f(std::integral_constant<int, 2>(), std::false_type()); break;
		case 5: f(std::integral_constant<int, 2>(), std::true_type()); break;
		case 9: f(std::integral_constant<int, 4>(), std::true_type()); break;
		case 16: f(std::integral_constant<int, 8>(), std::false_type()); break;
		case 17: f(std::integral_constant<int, 8>(), std::true_type()); break;
		case 32: f(std::integral_constant<int, 16>(), std::false_type()); break;
		case 33: f(std::integral_constant<int, 16>(), std::true_type()); break;
		default: f(std::integral_constant<int, 4>(), std::false_type()); break;
	}
}

// Gather order for a key tensor of nV entries (knob value 2, the default).
// One chunk of the list per XCD keeps each XCD's key-tensor lines in its own L2.  Past a threshold of
// 128 MB of key tensor (half of the 256 MB shared last-level cache) consecutive pieces win instead: all
// XCDs then read neighbouring parts of the list and share the cached lines (first-encounter key
// numbering: 13^3, 299 MB: 7.93 against 8.22 ms; 11^3, 71 MB: 1.444 against 1.451 ms, for twice the
// L2-to-fabric bytes).  With the size-ordered numbering, the smallest key contiguous and the 16-byte
// table copy, consecutive pieces are also 0.7% faster at 11^3 (1.113 against 1.121 ms), for 49% more
// L2-to-fabric bytes (calibrated traffic 1.63x the algorithmic bytes instead of 1.26x): the chunks stay.
constexpr int64_t kLargeKeyTensorEntries = ((int64_t)128 << 20) / (int64_t)sizeof(double);

static int resolve_gather_order(int knob, int64_t nV)
{
	if (knob != 2) { return knob; }
	return nV > kLargeKeyTensorEntries ? 0 : 1;
}

// the bucketed gather needs 32-bit positions (the fast decode's grids) and room for the key tables
static bool gather_bucketed_fits(const GatherDesc& d, int threads)
{
	return d.dc.uniform_pairs && gather_lds_bytes(d.nlut, d.C1) <= bucket_lds_bytes(threads);
}

// bucket = position >> shift, with at most max_bins buckets over the nV key-tensor entries
static void bucket_shape(int64_t nV, int max_bins, int& shift, int& bins)
{
	shift = 0;
	while (((nV - 1) >> shift) + 1 > max_bins) { ++shift; }
	bins = (int)(((nV - 1) >> shift) + 1);
}

static void launch_gather_bucketed(const GatherDesc& d, int xcd, int threads)
{
	const int64_t pieces = ceil_div(d.num, (int64_t)threads * kBucketEpt);
	if (pieces == 0) { return; }
	const int64_t chunks = xcd ? ceil_div(pieces, 8) : 0;
	const int64_t blocks = xcd ? 8 * chunks : pieces;
	GLI_REQUIRE(blocks < 2147483647, "too many elements for one gather launch");
	int shift = 0, bins = 1;
	bucket_shape(std::max<int64_t>(tet_index(0, 0, d.C1), 1), 4 * threads, shift, bins);
	hipStream_t st = ctx().stream;
	if (threads == 1024) { k_gather_bucketed<1024><<<(unsigned)blocks, 1024, 0, st>>>(d, chunks, shift, bins); }
	else if (threads == 512) { k_gather_bucketed<512><<<(unsigned)blocks, 512, 0, st>>>(d, chunks, shift, bins); }
	else { k_gather_bucketed<256><<<(unsigned)blocks, 256, 0, st>>>(d, chunks, shift, bins); }
	GLI_HIP(hipGetLastError());
}

// the classed gather's largest workgroup LDS (gfx950: 160 KB per workgroup)
constexpr int kClassedLdsMax = 160 * 1024;

static void launch_gather_classed(const GatherDesc& d, int xcd, int lds)
{
	const int threads = ctx().gather_threads;
	const int64_t pieces = ceil_div(d.num, (int64_t)threads * kClassEpt);
	if (pieces == 0) { return; }
	const int64_t chunks = xcd ? ceil_div(pieces, 8) : 0;
	const int64_t blocks = xcd ? 8 * chunks : pieces;
	GLI_REQUIRE(blocks < 2147483647, "too many elements for one gather launch");
	const bool use_lds = lds && gather_lds_fits(d);
	const int shm = classed_lds_bytes(threads, d.nlut, d.C1, use_lds);
	GLI_REQUIRE(shm <= kClassedLdsMax, "classed gather: workgroup LDS above 160 KB");
	// dynamic LDS above 64 KB needs the kernel's limit raised once
	static bool raised[2] = { false, false };
	const void* fn = use_lds ? (const void*)k_gather_classed<true> : (const void*)k_gather_classed<false>;
	if (shm > 64 * 1024 && !raised[use_lds ? 1 : 0]) {
		GLI_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, kClassedLdsMax));
		raised[use_lds ? 1 : 0] = true;
	}
	hipStream_t st = ctx().stream;
	if (use_lds) { k_gather_classed<true><<<(unsigned)blocks, threads, (size_t)shm, st>>>(d, chunks); }
	else { k_gather_classed<false><<<(unsigned)blocks, threads, (size_t)shm, st>>>(d, chunks); }
	GLI_HIP(hipGetLastError());
}

// returns the phase name of the kernel it launched: "gather" (k_gather), "gather_bucketed" or
// "gather_classed"
static const char* launch_gather(const GatherDesc& d, int ept, int xcd, int lds)
{
	// window-bucketed reads: on with knob 1; with knob 2 (automatic) for key tensors above 128 MB,
	// where they measured faster (13^3: 7.50 against 7.88 ms; 11^3: 1.52 against 1.46 ms)
	const int bk = ctx().gather_bucketed;
	const bool bucketed = bk == 1 || (bk == 2 && tet_index(0, 0, d.C1) > kLargeKeyTensorEntries);
	if (bucketed && gather_bucketed_fits(d, ctx().gather_threads)) {
		launch_gather_bucketed(d, xcd, ctx().gather_threads);
		return "gather_bucketed";
	}
	if (ctx().gather_classed && d.dc.uniform_pairs) {
		launch_gather_classed(d, xcd, lds);
		return "gather_classed";
	}
	const int threads = ctx().gather_threads;
	const int64_t pieces = ceil_div(d.num, (int64_t)threads * ept);
	if (pieces == 0) { return "gather"; }
	const int64_t chunks = xcd ? ceil_div(pieces, 8) : 0;
	const int64_t blocks = xcd ? 8 * chunks : pieces;
	GLI_REQUIRE(blocks < 2147483647, "too many elements for one gather launch");
	const bool use_lds = lds && gather_lds_fits(d);
	const size_t shm = use_lds ? (size_t)gather_lds_bytes(d.nlut, d.C1) : 0;
	hipStream_t st = ctx().stream;
	gather_dispatch(ept, use_lds, [&](auto E, auto Lds) {
		k_gather<decltype(E)::value, decltype(Lds)::value><<<(unsigned)blocks, threads, shm, st>>>(d, chunks);
	});
	GLI_HIP(hipGetLastError());
	return "gather";
}

static void launch_gather_grouped(const GatherDesc* d_descs, int n, int64_t total_pieces, int ept, int xcd, int lds,
                                  int lds_bytes)
{
	if (total_pieces == 0) { return; }
	const int64_t chunks = xcd ? ceil_div(total_pieces, 8) : 0;
	const int64_t blocks = xcd ? 8 * chunks : total_pieces;
	GLI_REQUIRE(blocks < 2147483647, "too many elements for one gather launch");
	const bool use_lds = lds && lds_bytes <= kGatherLdsMax;
	const size_t shm = use_lds ? (size_t)lds_bytes : 0;
	hipStream_t st = ctx().stream;
	gather_dispatch(ept, use_lds, [&](auto E, auto Lds) {
		k_gather_grouped<decltype(E)::value, decltype(Lds)::value><<<(unsigned)blocks, 256, shm, st>>>(d_descs, n, chunks);
	});
	GLI_HIP(hipGetLastError());
}

// key-tensor positions of d's list into d.pos (needs tetrahedral indices below 2^32 and the key
// tables in LDS)
static void launch_positions(const GatherDesc& d)
{
	const int64_t blocks = ceil_div(d.num, 256 * 8);
	if (blocks == 0) { return; }
	GLI_REQUIRE(blocks < 2147483647, "too many elements for one launch");
	const int lds = gather_lds_bytes(d.nlut, d.C1);
	GLI_REQUIRE(lds <= kGatherLdsMax, "key tables too large for the position pass");
	k_positions<<<(unsigned)blocks, 256, (size_t)lds, ctx().stream>>>(d);
	GLI_HIP(hipGetLastError());
}

static void launch_gather_cached_grouped(const GatherDesc* d_descs, int n, int64_t total_pieces, int ept, int xcd)
{
	if (total_pieces == 0) { return; }
	const int64_t chunks = xcd ? ceil_div(total_pieces, 8) : 0;
	const int64_t blocks = xcd ? 8 * chunks : total_pieces;
	GLI_REQUIRE(blocks < 2147483647, "too many elements for one gather launch");
	hipStream_t st = ctx().stream;
	switch (ept) {
		case 16: k_gather_cached_grouped<16><<<(unsigned)blocks, 256, 0, st>>>(d_descs, n, chunks); break;
		case 8: k_gather_cached_grouped<8><<<(unsigned)blocks, 256, 0, st>>>(d_descs, n, chunks); break;
		default: k_gather_cached_grouped<4><<<(unsigned)blocks, 256, 0, st>>>(d_descs, n, chunks); break;
	}
	GLI_HIP(hipGetLastError());
}

static void launch_gather_sorted_grouped(const GatherDesc* d_descs, int n, int64_t total_pieces, int xcd)
{
	if (total_pieces == 0) { return; }
	const int64_t chunks = xcd ? ceil_div(total_pieces, 8) : 0;
	const int64_t blocks = xcd ? 8 * chunks : total_pieces;
	GLI_REQUIRE(blocks < 2147483647, "too many elements for one gather launch");
	k_gather_sorted_grouped<<<(unsigned)blocks, kSortThreads, 0, ctx().stream>>>(d_descs, n, chunks);
	GLI_HIP(hipGetLastError());
}

bool gemm_shape_fits(int shape, int Cp, int tri, int64_t p_base)
{
	if (shape < 0 || shape >= kGemmShapes) { return false; }
	return !tri || (Cp % shape_bm(shape) == 0 && p_base % shape_bm(shape) == 0);
}

static TileOrder tile_order_for(int shape, int64_t P, int C, int grouped)
{
	return TileOrder{(int)ceil_div(P, shape_bm(shape)), (int)ceil_div(C, shape_bn(shape)), grouped};
}

template <int BM, int BN, int BK>
static void launch_mode(const ModeArgs& a, const TileOrder& o, hipStream_t st)
{
	k_mode_product<BM, BN, BK><<<(unsigned)o.blocks(), 256, 0, st>>>(a, o);
}

// the tile kernels address both operands with 32-bit element offsets
static void check_mode_extent(const ModeArgs& a)
{
	GLI_REQUIRE((int64_t)a.Kd * a.ldx <= INT32_MAX && (int64_t)a.Kd * a.ldb <= INT32_MAX && a.P <= INT32_MAX,
	            "mode-product operands above 2^31 elements");
}

void mode_product_shaped(const ModeArgs& a, int shape, int grouped)
{
	if (a.P <= 0 || a.C <= 0) { return; }
	check_mode_extent(a);
	GLI_REQUIRE(gemm_shape_fits(shape, a.Cp, a.tri_cols || a.tri_rows || a.tet_C1 > 0, a.p_base),
	            "mode-product tile shape does not fit the key padding");
	GLI_REQUIRE(ceil_div(a.P, 64) * ceil_div(a.C, 64) < 2147483647 - 8 * ceil_div(a.C, 64), "mode product too large");
	const TileOrder o = tile_order_for(shape, a.P, a.C, grouped);
	hipStream_t st = ctx().stream;
	switch (shape) {
		case 1: launch_mode<128, 64, 16>(a, o, st); break;
		case 2: launch_mode<64, 128, 16>(a, o, st); break;
		case 3: launch_mode<128, 128, 16>(a, o, st); break;
		case 4: launch_mode<64, 64, 32>(a, o, st); break;
		case 5: launch_mode<128, 64, 32>(a, o, st); break;
		default: launch_mode<64, 64, 16>(a, o, st); break;
	}
	GLI_HIP(hipGetLastError());
}

void mode_product_args(const ModeArgs& a)
{
	GLI_REQUIRE(!a.sym_x || (a.p_base == 0 && a.Cp % 64 == 0), "mirrored mode-product input needs rows aligned to the key padding");
	int shape = ctx().gemm_shape;
	if (!gemm_shape_fits(shape, a.Cp, a.tri_cols || a.tri_rows || a.tet_C1 > 0 || a.sym_x, a.p_base)) { shape = 0; }
	mode_product_shaped(a, shape, tile_order_grouped());
}

void mode_product(const double* X, int64_t ldx, int64_t P, const double* B, int64_t ldb, int C, int Kd,
                  double* Y, int64_t ldy, int Cp, int tri_cols, int tri_rows, int64_t p_base, int tet_C1)
{
	mode_product_args(ModeArgs{X, ldx, P, B, ldb, C, Kd, Y, ldy, Cp, tri_cols, tri_rows, p_base, tet_C1});
}

static void mode_product_f32(const ModeArgs& a, bool split)
{
	if (a.P <= 0 || a.C <= 0) { return; }
	GLI_REQUIRE(!(a.tri_cols || a.tri_rows || a.tet_C1 > 0) || (a.Cp % 64 == 0 && a.p_base % 64 == 0),
	            "mode-product tile does not fit the key padding");
	const TileOrder o{(int)ceil_div(a.P, 64), (int)ceil_div(a.C, 64), 0};
	if (split) { k_mode_product_f32<true><<<(unsigned)o.blocks(), 256, 0, ctx().stream>>>(a, o); }
	else { k_mode_product_f32<false><<<(unsigned)o.blocks(), 256, 0, ctx().stream>>>(a, o); }
	GLI_HIP(hipGetLastError());
}

void mode_product_grouped(const ModeArgs* a, int n, int shape, int grouped)
{
	GLI_REQUIRE(n >= 1 && n <= kMaxGroup, "too many problems in one grouped mode product");
	const bool cp = a[0].cp_u != nullptr;
	for (int g = 0; g < n; ++g) {
		check_mode_extent(a[g]);
		GLI_REQUIRE((a[g].cp_u != nullptr) == cp && (!cp || (shape == 0 && a[g].ldx == a[g].Cp && a[g].p_base == 0)),
		            "grouped rank-Q products: every problem rank-Q, 64x64 tiles, rows k1 Cp + k2");
	}
	ModeGroup grp;
	grp.n = n;
	int blocks = 0;
	for (int g = 0; g < n; ++g) {
		GLI_REQUIRE(gemm_shape_fits(shape, a[g].Cp, a[g].tri_cols || a[g].tri_rows || a[g].tet_C1 > 0, a[g].p_base),
		            "mode-product tile shape does not fit the key padding");
		grp.a[g] = a[g];
		grp.order[g] = tile_order_for(shape, std::max<int64_t>(a[g].P, 1), std::max(a[g].C, 1), grouped);
		grp.block_begin[g] = blocks;
		const int64_t nb = (a[g].P > 0 && a[g].C > 0) ? round_up(grp.order[g].blocks(), 8) : 0;
		GLI_REQUIRE(blocks + nb < 2147483647, "grouped mode product too large");
		if (nb == 0) { grp.order[g].nr = 0; }
		blocks += (int)nb;
	}
	grp.block_begin[n] = blocks;
	if (blocks == 0) { return; }
	hipStream_t st = ctx().stream;
	if (cp) {
		k_mode_product_grouped<64, 64, 16, true><<<(unsigned)blocks, 256, 0, st>>>(grp);
		GLI_HIP(hipGetLastError());
		return;
	}
	switch (shape) {
		case 1: k_mode_product_grouped<128, 64, 16><<<(unsigned)blocks, 256, 0, st>>>(grp); break;
		case 2: k_mode_product_grouped<64, 128, 16><<<(unsigned)blocks, 256, 0, st>>>(grp); break;
		case 3: k_mode_product_grouped<128, 128, 16><<<(unsigned)blocks, 256, 0, st>>>(grp); break;
		case 4: k_mode_product_grouped<64, 64, 32><<<(unsigned)blocks, 256, 0, st>>>(grp); break;
		case 5: k_mode_product_grouped<128, 64, 32><<<(unsigned)blocks, 256, 0, st>>>(grp); break;
		default: k_mode_product_grouped<64, 64, 16><<<(unsigned)blocks, 256, 0, st>>>(grp); break;
	}
	GLI_HIP(hipGetLastError());
}

// representations rho = m1^2 + m2^2 (0 <= m1, m2 <= M), grouped by rho with m1 then m2 ascending:
// rep_list[rep_off[rho] .. rep_off[rho + 1]) = m1 | m2 << 16
static void direct_representations(EriWorkspace& w, int M)
{
	if (w.rep_M == M) { return; }
	const int64_t nrho = 2 * (int64_t)M * M + 1;
	std::vector<int> off((size_t)nrho + 1, 0);
	for (int a = 0; a <= M; ++a) {
		for (int b = 0; b <= M; ++b) { off[(size_t)(a * a + b * b) + 1] += 1; }
	}
	for (int64_t r = 0; r < nrho; ++r) { off[(size_t)r + 1] += off[(size_t)r]; }
	std::vector<uint32_t> list((size_t)(M + 1) * (M + 1));
	std::vector<int> fill(off.begin(), off.end() - 1);
	for (int a = 0; a <= M; ++a) {
		for (int b = 0; b <= M; ++b) { list[(size_t)fill[(size_t)(a * a + b * b)]++] = (uint32_t)a | ((uint32_t)b << 16); }
	}
	GLI_HIP(hipStreamSynchronize(ctx().stream));   // the previous tables may still be read
	w.rep_off.alloc(off.size());
	w.rep_list.alloc(list.size());
	GLI_HIP(hipMemcpy(w.rep_off.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
	GLI_HIP(hipMemcpy(w.rep_list.p, list.data(), list.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
	w.rep_M = M;
}

// ---------------------------------------------------------------------------
// Separable method.  With d = sqrt(i) / 6 and w = 1/3 the Coulomb table is
//     T[i] = erf(d / (sqrt(2) w)) / d = 6 erf(sqrt(i / 8)) / sqrt(i) = C int_0^1 exp(-i s^2 / 8) ds,
// C = 6 / sqrt(2 pi) (also at i = 0), because erf(x) / x = (2 / sqrt(pi)) int_0^1 exp(-x^2 s^2) ds.
// For i = m1^2 + m2^2 + m3^2 the integrand factorises, so a quadrature rule (s_q, w_q) gives
//     T3[m1][m2][m3] ~ sum_q W_q g_q[m1] g_q[m2] g_q[m3],  g_q[m] = exp(-m^2 s_q^2 / 8),  W_q = C w_q,
// and the key tensor becomes V[k1][k2][k3] = sum_q U[q][k1] G[q][k2] G[q][k3] with G = g^T B and
// U = (W g)^T B: one rank-Q product in place of the three mode products.  The rule is composite
// Gauss-Legendre on geometric panels [0, r^J], [r^J, r^(J-1)], ..., [r, 1] with r = 0.35 and J the
// smallest with r^J <= 1 / (1.5 sqrt(3 M^2)): every integrand exp(-i s^2 / 8), i <= 3 M^2, then varies
// by a bounded amount per panel.  Counted from s = 0 the panels take 5, 6, 9 and 14 nodes, every
// further one 16 (the panels near 0 hold smooth, nearly flat parts of every integrand).  Q = 82 nodes
// for the folded ranges of 3^3..13^3 grids with K = 57 (M = 124..184); against 6 erf(sqrt(i/8)) / sqrt(i)
// the rule is within 2.2e-15 relative for every i <= 3 M^2 and every M from 0 to 318.
constexpr long double kSepRatio = 0.35L;
constexpr int kSepNodesLow[4] = { 5, 6, 9, 14 };   // the first four panels from s = 0
constexpr int kSepNodesPanel = 16;                 // every further panel

// Gauss-Legendre nodes x and weights w on [-1, 1]: Newton iterations on P_p from the usual cosine
// guesses, in extended precision
static void gauss_legendre(int p, long double* x, long double* w)
{
	const long double pi = 3.141592653589793238462643383279502884L;
	for (int i = 0; i < p; ++i) {
		long double z = cosl(pi * (i + 0.75L) / (p + 0.5L)), dp = 1;
		for (int it = 0; it < 100; ++it) {
			long double p0 = 1, p1 = z;   // P_{k-1}(z), P_k(z)
			for (int k = 2; k <= p; ++k) {
				const long double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
				p0 = p1;
				p1 = p2;
			}
			dp = p * (z * p1 - p0) / (z * z - 1);   // P_p'(z)
			const long double dz = p1 / dp;
			z -= dz;
			if (fabsl(dz) <= 1e-19L) { break; }
		}
		x[i] = z;
		w[i] = 2 / ((1 - z * z) * dp * dp);
	}
}

// nodes s_q in (0, 1) and weights w_q of the composite rule for a folded range M
void separable_rule(int M, std::vector<long double>& s, std::vector<long double>& w)
{
	const long double bound = 1.0L / (1.5L * sqrtl(3.0L * (long double)std::max(M, 1) * std::max(M, 1)));
	int J = 0;
	long double lo = 1.0L;
	while (lo > bound) {
		lo *= kSepRatio;
		++J;
	}
	long double x[kSepNodesPanel], wx[kSepNodesPanel];
	s.clear();
	w.clear();
	long double a = 0, b = lo;
	for (int panel = 0; panel <= J; ++panel) {
		const int p = panel < 4 ? kSepNodesLow[panel] : kSepNodesPanel;
		gauss_legendre(p, x, wx);
		for (int i = 0; i < p; ++i) {
			s.push_back((b - a) / 2 * x[i] + (a + b) / 2);
			w.push_back((b - a) / 2 * wx[i]);
		}
		a = b;
		b = (panel + 1 == J) ? 1.0L : b / kSepRatio;
	}
}

// [g | W g] for a folded range M (Mp rows, 2 Qp columns, zero past m = M and q = Q), host memory
static std::vector<double> separable_gw(int M, int Mp, int& Qp, int& Q)
{
	std::vector<long double> sq, wq;
	separable_rule(M, sq, wq);
	Q = (int)sq.size();
	Qp = (int)ceil_div(Q, 16) * 16;
	const long double C = 6.0L / sqrtl(2.0L * 3.141592653589793238462643383279502884L);
	std::vector<double> gw((size_t)Mp * 2 * Qp, 0.0);
	for (int m = 0; m <= M && m < Mp; ++m) {
		for (int q = 0; q < Q; ++q) {
			const long double g = expl(-(long double)m * m * sq[(size_t)q] * sq[(size_t)q] / 8);
			gw[(size_t)m * 2 * Qp + q] = (double)g;
			gw[(size_t)m * 2 * Qp + Qp + q] = (double)(C * wq[(size_t)q] * g);
		}
	}
	return gw;
}

// [g | W g] for the folded range of s, cached on the device per M
static void separable_factors(EriWorkspace& w, const EriSetup& s)
{
	if (w.sep_M == s.M) { return; }
	int Qp = 0, Q = 0;
	const std::vector<double> gw = separable_gw(s.M, s.Mp, Qp, Q);
	GLI_HIP(hipStreamSynchronize(ctx().stream));   // an earlier evaluation may still read the old factors
	w.sep_gw.alloc(gw.size());
	GLI_HIP(hipMemcpy(w.sep_gw.p, gw.data(), gw.size() * sizeof(double), hipMemcpyHostToDevice));
	w.sep_M = s.M;
	w.sep_Qp = Qp;
	w.sep_Q = Q;
}

// The separable method's tables in one launch: one workgroup per key column (a, b, cd) computes the two
// factor rows f_a and f_b (factor_row), its shift pair's products P'_ab (the k_products sums, zeroed where
// |P| <= tol) and the column's rank-Q factors
//     [G; U][r][key] = sum_{t < 4K-3} P'_ab[t] gw[|t - (2K-2) + 3 cd|][r]   (terms with that index above M vanish),
// all in LDS.  gw = [g | W g] holds the Gaussians of the separable rule per folded index mu; B[mu][key] (k_fold)
// sums the P' entries t with |t - (2K-2) + 3 cd| = mu, so this is [g | W g]^T B without forming B.  The 2R - 1
// or so keys of one shift pair recompute the same P', which costs less than a second launch.  Each row's sum
// issues 16 independent loads per step before its FMAs (four accumulators, added at the end): the launch is
// latency-bound.  Dynamic LDS: K coefficients, 2K - 1 Gaussians, two factor rows, 4K - 3 products.  The
// coefficients come from device memory c, or from ca when c is null.
__device__ inline void separable_tables_column(const double* __restrict__ c, const CoeffArg& ca, int K, int R,
                                               double tol, const int4* __restrict__ keys, int C1, int M,
                                               const double* __restrict__ gw, int rows, int Cp,
                                               double* __restrict__ GU, int key)
{
	extern __shared__ double lds[];
	if (key >= C1) {
		for (int r = threadIdx.x; r < rows; r += blockDim.x) { GU[(int64_t)r * Cp + key] = 0.0; }
		return;
	}
	const int L = 2 * K - 1, LP = 4 * K - 3, off = 2 * K - 2;
	double* cs = lds;
	double* ex = cs + K;
	double* fa = ex + L;
	double* fb = fa + L;
	double* ps = fb + L;
	const int4 k = keys[key];
	for (int q = threadIdx.x; q < K; q += blockDim.x) { cs[q] = (c != nullptr) ? c[q] : ca.c[q]; }
	__syncthreads();
	factor_row(cs, K, k.x, ex, fa);
	__syncthreads();
	factor_row(cs, K, k.y, ex, fb);
	__syncthreads();
	for (int t = threadIdx.x; t < LP; t += blockDim.x) {
		const int d = t - off;
		double acc = 0;
		const int mlo = max(0, d), mhi = min(L - 1, L - 1 + d);
		for (int m = mlo; m <= mhi; ++m) { acc += fa[m] * fb[m - d]; }
		ps[t] = (fabs(acc) > tol) ? acc : 0.0;
	}
	__syncthreads();
	const int shift = 3 * k.z - off;
	const int tlo = max(0, -M - shift), thi = min(LP - 1, M - shift);
	for (int r = threadIdx.x; r < rows; r += blockDim.x) {
		double acc[4] = { 0, 0, 0, 0 };
		int t = tlo;
		for (; t + 16 <= thi + 1; t += 16) {
			double g[16];
#pragma unroll
			for (int u = 0; u < 16; ++u) { g[u] = gw[(int64_t)abs(t + u + shift) * rows + r]; }
#pragma unroll
			for (int u = 0; u < 16; ++u) { acc[u & 3] = fma(ps[t + u], g[u], acc[u & 3]); }
		}
		for (; t <= thi; ++t) { acc[0] = fma(ps[t], gw[(int64_t)abs(t + shift) * rows + r], acc[0]); }
		GU[(int64_t)r * Cp + key] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
	}
}

__global__ __launch_bounds__(256) void k_separable_tables(const double* __restrict__ c, CoeffArg ca, int K, int R,
                                                          double tol, const int4* __restrict__ keys, int C1, int M,
                                                          const double* __restrict__ gw, int rows, int Cp,
                                                          double* __restrict__ GU)
{
	separable_tables_column(c, ca, K, R, tol, keys, C1, M, gw, rows, Cp, GU, (int)blockIdx.x);
}

// The same for every grid of a batch plan in one launch: problem g owns workgroups wg_begin .. wg_begin + Cp - 1
// (one per key column); its coefficients are in device memory.
struct SepTablesProblem {
	const double* c;
	const int4* keys;
	const double* gw;
	double* GU;
	double tol;
	int K, R, C1, M, rows, Cp;
	int wg_begin;
};
struct SepTablesGroup {
	SepTablesProblem p[kMaxGroup];
	int n;
};

__global__ __launch_bounds__(256) void k_separable_tables_grouped(SepTablesGroup grp)
{
	int g = 0;
	while (g + 1 < grp.n && (int)blockIdx.x >= grp.p[g + 1].wg_begin) { ++g; }
	const SepTablesProblem& pr = grp.p[g];
	const CoeffArg none{};
	separable_tables_column(pr.c, none, pr.K, pr.R, pr.tol, pr.keys, pr.C1, pr.M, pr.gw, pr.rows, pr.Cp, pr.GU,
	                        (int)blockIdx.x - pr.wg_begin);
}

// V[k1 <= k2 <= k3] = sum_q U[q][k3] G[q][k2] G[q][k1], tetrahedral store (the rank-Q tile kernel; a tile's
// rows p = k3 Cp + k2 share k3, its columns run over k1)
static void mode_product_cp(const ModeArgs& a)
{
	if (a.P <= 0 || a.C <= 0) { return; }
	check_mode_extent(a);
	int shape = ctx().gemm_shape;
	if (!gemm_shape_fits(shape, a.Cp, 1, 0)) { shape = 0; }
	GLI_REQUIRE(a.cp_u != nullptr && a.ldx == a.Cp && a.p_base == 0 && gemm_shape_fits(shape, a.Cp, 1, 0),
	            "rank-Q product: rows must be k3 Cp + k2 with the key padding a multiple of the tile");
	const TileOrder o = tile_order_for(shape, a.P, a.C, tile_order_grouped());
	hipStream_t st = ctx().stream;
	switch (shape) {
		case 1: k_mode_product<128, 64, 16, true><<<(unsigned)o.blocks(), 256, 0, st>>>(a, o); break;
		case 2: k_mode_product<64, 128, 16, true><<<(unsigned)o.blocks(), 256, 0, st>>>(a, o); break;
		case 3: k_mode_product<128, 128, 16, true><<<(unsigned)o.blocks(), 256, 0, st>>>(a, o); break;
		case 4: k_mode_product<64, 64, 32, true><<<(unsigned)o.blocks(), 256, 0, st>>>(a, o); break;
		case 5: k_mode_product<128, 64, 32, true><<<(unsigned)o.blocks(), 256, 0, st>>>(a, o); break;
		default: k_mode_product<64, 64, 16, true><<<(unsigned)o.blocks(), 256, 0, st>>>(a, o); break;
	}
	GLI_HIP(hipGetLastError());
}

// The tetrahedral rank-Q kernel over n problems with Q nodes (G and U hold Qp >= 4 ceil(Q / 4) rows).
// Returns false, having launched nothing, when the knob selects the tile kernel or no instance covers
// S = ceil(Q / 4) steps; the caller then runs the tile kernel.
// the node counts the tetrahedral kernel has instances for (S = ceil(Q / 4) steps), with the knob on
static bool rankq_tet_supported(int Q)
{
	const int S = (Q + 3) / 4;
	return ctx().rankq_tet != 0 && Q >= 1 &&
	       (S == 2 || S == 3 || S == 5 || S == 9 || S == 13 || S == 17 || S == 21 || S == 25 || S == 29);
}

static bool rankq_tet_launch(const RankQProblem* probs, int n, int Q, int Qp)
{
	if (ctx().rankq_tet == 0 || n < 1 || n > kMaxGroup) { return false; }
	const int S = (Q + 3) / 4;
	GLI_REQUIRE(Q >= 1 && 4 * S <= Qp, "rank-Q product: the factors need 4 ceil(Q / 4) rows");
	RankQGroup grp;
	grp.n = n;
	int wgs = 0;
	for (int g = 0; g < n; ++g) {
		grp.p[g] = probs[g];
		grp.p[g].wg_begin = wgs;
		const int nch = (probs[g].C1 + 15) / 16;
		GLI_REQUIRE(16 * nch <= probs[g].Cp, "rank-Q product: the key padding must cover whole chunks of 16");
		for (int c = 0; c < nch; ++c) { wgs += rankq_chunk_wgs(c); }
	}
	if (wgs == 0) { return true; }
	hipStream_t st = ctx().stream;
	switch (S) {
		case 2: k_rankq_tet<2><<<(unsigned)wgs, 256, 0, st>>>(grp); break;
		case 3: k_rankq_tet<3><<<(unsigned)wgs, 256, 0, st>>>(grp); break;
		case 5: k_rankq_tet<5><<<(unsigned)wgs, 256, 0, st>>>(grp); break;
		case 9: k_rankq_tet<9><<<(unsigned)wgs, 256, 0, st>>>(grp); break;
		case 13: k_rankq_tet<13><<<(unsigned)wgs, 256, 0, st>>>(grp); break;
		case 17: k_rankq_tet<17><<<(unsigned)wgs, 256, 0, st>>>(grp); break;
		case 21: k_rankq_tet<21><<<(unsigned)wgs, 256, 0, st>>>(grp); break;
		case 25: k_rankq_tet<25><<<(unsigned)wgs, 256, 0, st>>>(grp); break;
		case 29: k_rankq_tet<29><<<(unsigned)wgs, 256, 0, st>>>(grp); break;
		default: return false;   // the rule's node counts are 5, 11, 20, 34, 50, 66, 82, 98, 114, 130, ...
	}
	GLI_HIP(hipGetLastError());
	return true;
}

// The key tensor of the separable method, queued on the library stream: the rank-Q factors G and U
// straight from the coefficients (d_coeffs on the device, or h_coeffs of at most kArgCoeffs entries as a
// kernel argument) in one k_separable_tables launch, then the rank-Q product into V.  No pair-product
// launch precedes it.
static void key_tensor_separable(const EriSetup& s, EriWorkspace& w, const double* d_coeffs, const double* h_coeffs,
                                 double tol)
{
	GLI_REQUIRE(tol >= 0, "tolerance must be non-negative");
	separable_factors(w, s);
	const int64_t Qp = w.sep_Qp, Cp = s.Cp;
	w.GU.reserve((size_t)(2 * Qp * Cp));
	CoeffArg ca{};
	if (d_coeffs == nullptr) {
		GLI_REQUIRE(h_coeffs != nullptr && s.K <= kArgCoeffs, "coefficients by argument: at most 64 of them");
		for (int q = 0; q < s.K; ++q) { ca.c[q] = h_coeffs[q]; }
	}
	const size_t shm = (size_t)(s.K + 3 * (2 * s.K - 1) + (4 * s.K - 3)) * sizeof(double);
	k_separable_tables<<<(unsigned)Cp, 256, shm, ctx().stream>>>(d_coeffs, ca, s.K, s.R, tol, s.d_keys, s.C1, s.M,
	                                                             w.sep_gw.p, (int)(2 * Qp), (int)Cp, w.GU.p);
	GLI_HIP(hipGetLastError());
	phase_mark("tables");
	const RankQProblem pr{w.GU.p, w.GU.p + Qp * Cp, w.V.p, (int)Cp, s.C1, 0};
	if (!rankq_tet_launch(&pr, 1, w.sep_Q, (int)Qp)) {
		ModeArgs a{w.GU.p, Cp, Cp * Cp, w.GU.p, Cp, s.Cp, (int)Qp, w.V.p, 0, s.Cp, 1, 1, 0, s.C1};
		a.cp_u = w.GU.p + Qp * Cp;
		mode_product_cp(a);
	}
	phase_mark("separable_tensor");
}

// values of device-resident four-indices into device-resident values
// Everything of one evaluation that does not depend on the index list: tables, and for the
// key-tensor methods the three mode products up to V.  Queued on the library stream.
struct EriPrepared {
	EriSetup s;
	int method = GLI_ERI_TENSOR_FP64;
	DirectCount cnt{nullptr, nullptr};
	size_t direct_shm = 0;
};

// The FP64 key tensor V from the folded columns B and the kernel tensor T3: three MFMA mode products
//   Y1[m2][m3][k3] = sum_m1 T3[m1][m2][m3] B[m1][k3]   (symmetric in (m2, m3): with half, only m3 >= m2
//                                                        is computed, and mode 2 reads the rest mirrored)
//   Y2[m3][k3][k2] = sum_m2 Y1[m2][m3][k3] B[m2][k2]   (k2 <= k3 only)
//   V[k1 <= k2 <= k3] = sum_m3 Y2[m3][k3][k2] B[m3][k1], written in tetrahedral layout
// (V is symmetric, so the first product may carry the largest key)
// queued on the library stream.  Every FP64 entry point runs exactly this sequence.
static void key_tensor_fp64(const EriSetup& s, const double* B, const double* T3, double* Y1, double* Y2, double* V, bool half)
{
	const int64_t Mp = s.Mp, Cp = s.Cp;
	if (half) {
		ModeArgs a{T3, Mp * Mp, Mp * Mp, B, Cp, s.Cp, s.Mp, Y1, Cp, s.Cp, 0, 0, 0, 0};
		a.half_n = (int)Mp;
		mode_product_args(a);
	}
	else { mode_product(T3, Mp * Mp, Mp * Mp, B, Cp, s.Cp, s.Mp, Y1, Cp, s.Cp, 0, 0, 0, 0); }
	phase_mark("mode1");
	if (half) {
		ModeArgs a{Y1, Mp * Cp, Mp * Cp, B, Cp, s.Cp, s.Mp, Y2, Cp, s.Cp, 1, 0, 0, 0};
		a.sym_x = 1;
		mode_product_args(a);
	}
	else { mode_product(Y1, Mp * Cp, Mp * Cp, B, Cp, s.Cp, s.Mp, Y2, Cp, s.Cp, 1, 0, 0, 0); }
	phase_mark("mode2");
	mode_product(Y2, Cp * Cp, Cp * Cp, B, Cp, s.Cp, s.Mp, V, 0, s.Cp, 1, 1, 0, s.C1);
	phase_mark("mode3");
}

static EriPrepared eri_prepare(const double* coeffs, int64_t K, const int64_t grid[6], double tol, int method, bool any)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	GLI_REQUIRE(method >= GLI_ERI_TENSOR_FP64 && method <= GLI_ERI_SEPARABLE_FP64, "unknown ERI method");
	GLI_REQUIRE(coeffs != nullptr && K >= 3, "the number of gausslet coefficients must be odd and at least 3");
	phase_begin();
	EriPrepared ep;
	ep.s = eri_shape(K, grid, mirror_symmetric(coeffs, K));
	ep.method = method;
	const EriSetup& s = ep.s;
	EriWorkspace& w = eri_ws();
	w.P.reserve((size_t)s.R * s.R * (4 * s.K - 3));
	w.B.reserve((size_t)s.Mp * s.Cp);
	const double* d_coeffs = nullptr;   // up to 64 coefficients go as a kernel argument
	if (K > kArgCoeffs) {
		w.coeffs.reserve(K);
		upload_async(w.coeffs.p, coeffs, K * sizeof(double));
		d_coeffs = w.coeffs.p;
	}
	if (method == GLI_ERI_SEPARABLE_FP64 && any) {
		// one launch for the pair products and the rank-Q factors (k_separable_tables), then the rank-Q product
		w.V.reserve((size_t)tet_index(0, 0, s.C1));
		key_tensor_separable(s, w, d_coeffs, coeffs, tol);
		return ep;
	}
	launch_tables(s, d_coeffs, tol, w.P.p, method == GLI_ERI_SEPARABLE_FP64 ? nullptr : w.B.p, coeffs);
	phase_mark("tables");
	if (!any) { return ep; }
	const int64_t Mp = s.Mp, Cp = s.Cp;

	if (method == GLI_ERI_DIRECT_FP64) {
		w.coords.reserve((size_t)s.g.N);
		k_point_coords<<<(unsigned)ceil_div(s.g.N, 256), 256, 0, c.stream>>>(s.g.n[0], s.g.n[1], s.g.n[2], w.coords.p);
		direct_representations(w, s.M);
		if (w.count_direct) {
			const int rows = s.R * s.R;
			w.nnz_rows.reserve((size_t)rows);
			k_nnz_rows<<<(unsigned)rows, 256, 0, c.stream>>>(w.P.p, rows, 4 * s.K - 3, w.nnz_rows.p);
			ep.cnt = DirectCount{w.direct_sums.p, w.nnz_rows.p};
		}
		ep.direct_shm = 3 * (size_t)s.Mp * sizeof(double) + DIRECT_CHUNK * (sizeof(double) + sizeof(int));
		GLI_REQUIRE(ep.direct_shm <= 65536, "grid too long for the per-element path: 6(n-1) + 2(K-1) must stay below 1700");
		return ep;
	}
	const bool mixed = (method == GLI_ERI_TENSOR_MIXED || method == GLI_ERI_TENSOR_BF16X2);
	const int ns = (method == GLI_ERI_TENSOR_BF16X2) ? 2 : 5;
	const bool fp32 = (method == GLI_ERI_TENSOR_FP32 || method == GLI_ERI_TENSOR_FP32X2);
	const bool fp32_split = (method == GLI_ERI_TENSOR_FP32X2);
	GLI_REQUIRE(!mixed || s.Mp <= 256, "the split-precision method needs 6(n-1) + 2(K-1) < 256; use the FP64 method");
	// the mode products of the comparison FP32 path
	auto mode = [&](const double* X, int64_t ldx, int64_t P, double* Y, int64_t ldy, int tri_cols, int tri_rows, int tet_C1) {
		if (fp32) { mode_product_f32(ModeArgs{X, ldx, P, w.B.p, Cp, s.Cp, s.Mp, Y, ldy, s.Cp, tri_cols, tri_rows, 0, tet_C1}, fp32_split); }
		else { mode_product(X, ldx, P, w.B.p, Cp, s.Cp, s.Mp, Y, ldy, s.Cp, tri_cols, tri_rows, 0, tet_C1); }
	};

	// kernel tensor, cached per folded range
	if (w.T3_M != s.M) {
		w.T3.reserve((size_t)Mp * Mp * Mp);
		launch_kernel_tensor(s, w.T3.p);
		w.T3_M = s.M;
		split_invalidate_cache();
	}
	phase_mark("kernel_tensor");
	if (mixed) {
		split_factor(w.B.p, Cp, s.Cp, s.Mp, nullptr, ns);
		phase_mark("split_factor");
	}
	const bool half = !mixed && !fp32 && c.mode1_half;
	w.Y1.reserve((size_t)Mp * Mp * Cp);
	w.Y2.reserve((size_t)Mp * Cp * Cp);
	const int64_t nV = tet_index(0, 0, s.C1);
	w.V.reserve((size_t)nV);
	if (!mixed && !fp32) {
		key_tensor_fp64(s, w.B.p, w.T3.p, w.Y1.p, w.Y2.p, w.V.p, half);
		return ep;
	}
	// split-precision and FP32 comparison methods
	if (mixed) { mode_product_split(w.T3.p, Mp * Mp, Mp * Mp, s.Cp, s.Mp, w.Y1.p, Cp, s.Cp, 0, 0, 0, 0, true); }
	else { mode(w.T3.p, Mp * Mp, Mp * Mp, w.Y1.p, Cp, 0, 0, 0); }
	phase_mark("mode1");
	if (mixed) { mode_product_split(w.Y1.p, Mp * Cp, Mp * Cp, s.Cp, s.Mp, w.Y2.p, Cp, s.Cp, 1, 0, 0, 0, false); }
	else { mode(w.Y1.p, Mp * Cp, Mp * Cp, w.Y2.p, Cp, 1, 0, 0); }
	phase_mark("mode2");
	if (mixed) { mode_product_split(w.Y2.p, Cp * Cp, Cp * Cp, s.Cp, s.Mp, w.V.p, 0, s.Cp, 1, 1, 0, s.C1, false); }
	else { mode(w.Y2.p, Cp * Cp, Cp * Cp, w.V.p, 0, 1, 1, s.C1); }
	phase_mark("mode3");
	return ep;
}

// ---------------------------------------------------------------------------
// Allocation-free evaluation (gli_eri_values_device_noalloc), for callers that cannot allocate device
// memory.  The whole workspace of one FP64 evaluation (coefficients, factor / product / fold tables,
// Coulomb table, kernel tensor, the two intermediate products, the key tensor and the key tables) is
// carved out of one file-scope device array, `scratch`.  Every GPU the library runs on holds its own
// copy of the array from module load to process end.  In host code the array's name is only a
// host-side stand-in: its address on a device comes from hipGetSymbolAddress, once per device at
// context set-up (eri_scratch_init, with that device current), and only that address is passed to
// kernels and copies.  Each buffer is written by its producing kernel before anything reads it, so
// nothing has to start at zero; the array keeps the last call's contents, and calls that overlap on
// one GPU share it and race.
constexpr size_t kEriScratchBytes = (size_t)96 << 20;   // grids up to 7^3 with K <= 57 (82 MB)
constexpr int kScratchMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kEriScratchBytes];
static void* g_scratch_ptr[kScratchMaxDevices] = {};
static hipError_t g_scratch_err[kScratchMaxDevices] = {};

void eri_scratch_init(int device)
{
	if (device < 0 || device >= kScratchMaxDevices) { return; }
	void* p = nullptr;
	const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
	g_scratch_err[device] = e;
	g_scratch_ptr[device] = (e == hipSuccess) ? p : nullptr;
}

// consecutive 256-byte aligned pieces of the scratch
struct ScratchCarve {
	char* base;
	size_t cap;
	size_t used = 0;
	template <typename T>
	T* take(int64_t count)
	{
		const size_t at = (used + 255) & ~(size_t)255;
		used = at + (size_t)std::max<int64_t>(count, 1) * sizeof(T);
		return (T*)(base + at);
	}
};

static void eri_values_noalloc(const double* coeffs, int64_t K, const int64_t grid[6], const int64_t* d_idx, int64_t num,
                               double tol, int method, double* d_vals)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	GLI_REQUIRE(method == GLI_ERI_TENSOR_FP64, "the allocation-free entry point computes the FP64 key-tensor method (0) only");
	GLI_REQUIRE(coeffs != nullptr && K >= 3, "the number of gausslet coefficients must be odd and at least 3");
	GLI_REQUIRE(num >= 0, "negative number of indices");
	GLI_REQUIRE(num == 0 || (d_idx != nullptr && d_vals != nullptr), "null index or value array");
	GLI_REQUIRE(tol >= 0, "tolerance must be non-negative");
	char* base = (c.device >= 0 && c.device < kScratchMaxDevices) ? (char*)g_scratch_ptr[c.device] : nullptr;
	if (base == nullptr) {
		const hipError_t e = (c.device >= 0 && c.device < kScratchMaxDevices) ? g_scratch_err[c.device] : hipErrorInvalidDevice;
		GLI_REQUIRE(false, std::string("no scratch address on this device: ") + hipGetErrorString(e));
	}
	EriSetup s = eri_shape(K, grid, mirror_symmetric(coeffs, K), false);
	GLI_REQUIRE(s.ks->tet32, "key tensor too large for 32-bit tetrahedral offsets");
	const int64_t Mp = s.Mp, Cp = s.Cp, nT = 3 * (int64_t)s.M * s.M + 1;
	ScratchCarve sc{base, kEriScratchBytes};
	double* d_coeffs = sc.take<double>(K);
	double* P = sc.take<double>((int64_t)s.R * s.R * (4 * s.K - 3));
	double* B = sc.take<double>(Mp * Cp);
	double* T = sc.take<double>(nT);
	double* T3 = sc.take<double>(Mp * Mp * Mp);
	double* Y1 = sc.take<double>(Mp * Mp * Cp);
	double* Y2 = sc.take<double>(Mp * Cp * Cp);
	double* V = sc.take<double>(tet_index(0, 0, s.C1));
	int4* keys = sc.take<int4>(s.C1);
	uint32_t* image = sc.take<uint32_t>((int64_t)s.ks->image.size());   // tet3 and lut16, as the gather reads them
	if (sc.used > sc.cap) {
		char msg[256];
		snprintf(msg, sizeof(msg), "the allocation-free entry point needs %.1f MB of scratch for this grid and K, it holds %zu MB "
		         "(grids up to 7^3 with K <= 57); use gli_eri_values_device", sc.used / 1048576.0, kEriScratchBytes >> 20);
		GLI_REQUIRE(false, msg);
	}
	// small host tables into the scratch; the copies are complete before the host sources can move
	const double* hT = eri_table_host(nT);
	hipStream_t st = c.stream;
	GLI_HIP(hipMemcpyAsync(d_coeffs, coeffs, (size_t)K * sizeof(double), hipMemcpyHostToDevice, st));
	GLI_HIP(hipMemcpyAsync(T, hT, (size_t)nT * sizeof(double), hipMemcpyHostToDevice, st));
	GLI_HIP(hipMemcpyAsync(keys, s.ks->keys.data(), (size_t)s.C1 * sizeof(int4), hipMemcpyHostToDevice, st));
	GLI_HIP(hipMemcpyAsync(image, s.ks->image.data(), s.ks->image.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
	GLI_HIP(hipStreamSynchronize(st));
	s.T = T;
	s.d_keys = keys;
	s.d_tet3 = image;
	s.d_lut16 = (const short*)(image + ((s.C1 + 3) & ~3));
	// the same kernels in the same order as gli_eri_values_device with the FP64 method
	phase_begin();
	launch_tables(s, d_coeffs, tol, P, B);
	phase_mark("tables");
	if (num == 0) { return; }
	launch_kernel_tensor(s, T3);
	phase_mark("kernel_tensor");
	key_tensor_fp64(s, B, T3, Y1, Y2, V, c.mode1_half != 0);
	phase_mark(launch_gather(gather_desc(s, d_idx, num, V, d_vals, nullptr), c.gather_ept,
	                         resolve_gather_order(c.gather_xcd, tet_index(0, 0, s.C1)), c.gather_lds));
}

// values of num indices (device arrays) from a prepared evaluation, queued on the library stream
static void eri_apply(const EriPrepared& ep, const int64_t* d_idx, int64_t num, double* d_vals, bool mark)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	if (num <= 0) { return; }
	const EriSetup& s = ep.s;
	EriWorkspace& w = eri_ws();
	if (ep.method == GLI_ERI_DIRECT_FP64) {
		const int64_t blocks = std::min<int64_t>(num, 8192);
		k_eri_direct<<<(unsigned)blocks, DIRECT_THREADS, ep.direct_shm, c.stream>>>(d_idx, num, s.g, s.ks->lut_dev.p, w.coords.p, w.B.p,
			s.Cp, s.M, s.Mp, s.T, w.rep_off.p, w.rep_list.p, ep.cnt, d_vals);
		GLI_HIP(hipGetLastError());
		if (mark) { phase_mark("direct"); }
		return;
	}
	const char* kernel = launch_gather(gather_desc(s, d_idx, num, w.V.p, d_vals, nullptr), c.gather_ept,
	                                   resolve_gather_order(c.gather_xcd, tet_index(0, 0, s.C1)), c.gather_lds);
	if (mark) { phase_mark(kernel); }
}

static void eri_values_dev(const double* coeffs, int64_t K, const int64_t grid[6], const int64_t* d_idx, int64_t num,
                           double tol, int method, double* d_vals)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	GLI_REQUIRE(num >= 0, "negative number of indices");
	const EriPrepared ep = eri_prepare(coeffs, K, grid, tol, method, num > 0);
	eri_apply(ep, d_idx, num, d_vals, true);
}

// ---------------------------------------------------------------------------
// batched evaluation of several grids: a plan owns every buffer, picks its kernel variants by
// timing them on its own shapes, and replays one captured hipGraph per evaluation
//
// graph: parameter upload (coefficients and output scales, pinned -> device), per-grid factor /
// product / fold tables, one grouped launch per mode product over all grids (or per-grid
// split-precision products when those are faster), one grouped gather over the concatenated
// index lists of all grids.

struct BatchGrid {
	EriSetup s;
	int64_t K = 0;
	bool sym = false;
	const int64_t* d_idx = nullptr;
	int64_t num = 0;
	double* d_vals = nullptr;
	DevBuf<double> P, B, Y1, Y2, V;
	DevBuf<double> GU;      // separable plans: the rank-Q factors [G; U] (2 Qp x Cp)
	const double* sep_gw = nullptr;   // ... and the Gaussian factors [g | W g] of the grid's folded range
	int sep_Qp = 0, sep_Q = 0;
	DevBuf<uint32_t> pos;   // key-tensor position of every element (plans with cached positions)
	DevBuf<uint16_t> perm;  // place of each sorted position in its window (plans with sorted windows)
	const double* T3 = nullptr;
	size_t coeff_off = 0, scale_off = 0;
};

struct BatchChoice {
	int gemm_shape[3] = { 0, 0, 0 };
	int gemm_order[3] = { 1, 1, 1 };
	bool split[3] = { false, false, false };
	int gather_ept = 8;
	int gather_order = 1;
	int gather_lds = 1;
};

struct BatchPlan {
	int method = GLI_ERI_TENSOR_FP64;
	double tol = 0;
	std::vector<std::unique_ptr<BatchGrid>> grids;
	std::vector<std::pair<int, std::unique_ptr<DevBuf<double>>>> t3;   // kernel tensor per folded range M
	std::vector<std::pair<int, std::unique_ptr<DevBuf<double>>>> sep_gw;   // separable plans: [g | W g] per M
	DevBuf<double> d_params;
	double* h_params = nullptr;
	size_t n_params = 0;
	DevBuf<GatherDesc> d_gdesc;
	int gdesc_ept = -1;
	int64_t total_pieces = 0;
	int lds_bytes = 0;   // LDS for the largest key table of the batch
	bool half = false;   // half-stored first mode (FP64 plans)
	bool cached = false; // the gather reads positions computed once at creation instead of decoding
	bool sorted = false; // those positions are sorted inside windows (k_sort_windows, k_gather_sorted_grouped)
	// two-branch graph: the grid with the most elements on the library stream, the others on a
	// second stream (default priority) forked after the tables and joined before the end
	bool branches = false;
	int lead = -1;
	hipStream_t side = nullptr;
	hipEvent_t fork = nullptr, join = nullptr;
	DevBuf<GatherDesc> d_gdesc_lead, d_gdesc_rest;
	int64_t pieces_lead = 0, pieces_rest = 0;
	int gdesc2_ept = -1;
	SplitWorkspace* split_ws = nullptr;
	BatchChoice choice;
	std::string report;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	hipEvent_t done = nullptr;
	bool launched = false;

	~BatchPlan()
	{
		if (launched && done) { (void)hipEventSynchronize(done); }
		if (exec) { (void)hipGraphExecDestroy(exec); }
		if (graph) { (void)hipGraphDestroy(graph); }
		if (done) { (void)hipEventDestroy(done); }
		if (fork) { (void)hipEventDestroy(fork); }
		if (join) { (void)hipEventDestroy(join); }
		if (side) { (void)hipStreamDestroy(side); }
		if (h_params) { (void)hipHostFree(h_params); }
		if (split_ws) { split_workspace_delete(split_ws); }
	}
};

// half: the first mode stores only m3 >= m2 and the second reads it mirrored (FP64 plans)
static ModeArgs batch_mode_args(const BatchGrid& bg, int mode, bool half)
{
	const EriSetup& s = bg.s;
	const int64_t Mp = s.Mp, Cp = s.Cp;
	if (mode == 0) {
		ModeArgs a{bg.T3, Mp * Mp, Mp * Mp, bg.B.p, Cp, s.Cp, s.Mp, bg.Y1.p, Cp, s.Cp, 0, 0, 0, 0};
		a.half_n = half ? (int)Mp : 0;
		return a;
	}
	if (mode == 1) {
		ModeArgs a{bg.Y1.p, Mp * Cp, Mp * Cp, bg.B.p, Cp, s.Cp, s.Mp, bg.Y2.p, Cp, s.Cp, 1, 0, 0, 0};
		a.sym_x = half ? 1 : 0;
		return a;
	}
	return ModeArgs{bg.Y2.p, Cp * Cp, Cp * Cp, bg.B.p, Cp, s.Cp, s.Mp, bg.V.p, 0, s.Cp, 1, 1, 0, s.C1};
}

// the tables of one grid of a plan (separable plans need no folded columns)
static void batch_tables(const BatchPlan& plan, BatchGrid& g)
{
	launch_tables(g.s, plan.d_params.p + g.coeff_off, plan.tol, g.P.p,
	              plan.method == GLI_ERI_SEPARABLE_FP64 ? nullptr : g.B.p);
}

// separable plans: the rank-Q product of one grid (V from its factors [G; U])
static ModeArgs batch_sep_args(const BatchGrid& bg)
{
	const EriSetup& s = bg.s;
	const int64_t Cp = s.Cp, Qp = bg.sep_Qp;
	ModeArgs a{bg.GU.p, Cp, Cp * Cp, bg.GU.p, Cp, s.Cp, (int)Qp, bg.V.p, 0, s.Cp, 1, 1, 0, s.C1};
	a.cp_u = bg.GU.p + Qp * Cp;
	return a;
}

// separable plans: every grid's pair products and factors [G; U] in one grouped launch, then one grouped
// rank-Q product over all grids (the single-grid path's sums, so bit-identical to it; the tile order applies
// to the tile kernel only)
static void batch_separable(BatchPlan& plan, int order)
{
	ModeArgs a[kMaxGroup];
	RankQProblem pr[kMaxGroup];
	SepTablesGroup tg;
	int n = 0, wgs = 0;
	size_t shm = 0;
	bool tet = true;   // every grid's node count has a tetrahedral instance
	GLI_REQUIRE(plan.grids.size() <= (size_t)kMaxGroup, "too many grids in one plan");
	for (auto& g : plan.grids) {
		const EriSetup& s = g->s;
		const int64_t Qg = g->sep_Qp;
		tg.p[n] = SepTablesProblem{plan.d_params.p + g->coeff_off, s.d_keys, g->sep_gw, g->GU.p, plan.tol, s.K, s.R, s.C1,
		                           s.M, (int)(2 * Qg), s.Cp, wgs};
		wgs += s.Cp;
		shm = std::max(shm, (size_t)(s.K + 3 * (2 * s.K - 1) + (4 * s.K - 3)) * sizeof(double));
		a[n] = batch_sep_args(*g);
		pr[n] = RankQProblem{g->GU.p, g->GU.p + Qg * s.Cp, g->V.p, s.Cp, s.C1, 0};
		tet = tet && rankq_tet_supported(g->sep_Q);
		++n;
	}
	tg.n = n;
	if (wgs > 0) {
		// every grid's pair products and [G; U] in one launch (the single-grid path's k_separable_tables sums)
		k_separable_tables_grouped<<<(unsigned)wgs, 256, shm, ctx().stream>>>(tg);
		GLI_HIP(hipGetLastError());
	}
	if (!tet) {   // the tile kernel takes any mix of node counts in one launch
		mode_product_grouped(a, n, 0, order);
		return;
	}
	// the tetrahedral kernel: one grouped launch per distinct node count
	bool done[kMaxGroup] = {};
	for (int g0 = 0; g0 < n; ++g0) {
		if (done[g0]) { continue; }
		const int Q = plan.grids[g0]->sep_Q;
		RankQProblem sub[kMaxGroup];
		int m = 0, Qp = 1 << 30;
		for (int g = g0; g < n; ++g) {
			if (done[g] || plan.grids[g]->sep_Q != Q) { continue; }
			sub[m++] = pr[g];
			Qp = std::min(Qp, plan.grids[g]->sep_Qp);
			done[g] = true;
		}
		GLI_REQUIRE(rankq_tet_launch(sub, m, Q, Qp), "rank-Q product: no tetrahedral instance for this node count");
	}
}

// the gather of a batch plan: from the cached positions when the plan holds them
static void batch_launch_gather(const BatchPlan& plan, const GatherDesc* descs, int n, int64_t pieces, int ept, int order,
                                int lds)
{
	if (plan.sorted) { launch_gather_sorted_grouped(descs, n, pieces, order); }   // pieces of kSortEpt elements per lane
	else if (plan.cached) { launch_gather_cached_grouped(descs, n, pieces, ept, order); }
	else { launch_gather_grouped(descs, n, pieces, ept, order, lds, plan.lds_bytes); }
}

// gather descriptors for one elements-per-lane choice (piece offsets depend on it)
static void batch_gather_descs(BatchPlan& plan, int ept)
{
	if (plan.gdesc_ept == ept) { return; }
	std::vector<GatherDesc> h;
	int64_t pieces = 0;
	for (auto& g : plan.grids) {
		GatherDesc d = gather_desc(g->s, g->d_idx, g->num, g->V.p, g->d_vals, plan.d_params.p + g->scale_off);
		d.pos = plan.cached ? g->pos.p : nullptr;
		d.perm = plan.sorted ? g->perm.p : nullptr;
		d.piece_begin = pieces;
		pieces += gather_pieces(g->num, ept);
		h.push_back(d);
	}
	plan.d_gdesc.reserve(h.size());
	GLI_HIP(hipMemcpyAsync(plan.d_gdesc.p, h.data(), h.size() * sizeof(GatherDesc), hipMemcpyHostToDevice, ctx().stream));
	GLI_HIP(hipStreamSynchronize(ctx().stream));
	plan.gdesc_ept = ept;
	plan.total_pieces = pieces;
}

static void batch_mode(BatchPlan& plan, int mode, const BatchChoice& ch)
{
	ModeArgs a[kMaxGroup];
	const int n = (int)plan.grids.size();
	for (int g = 0; g < n; ++g) { a[g] = batch_mode_args(*plan.grids[g], mode, plan.half); }
	mode_product_grouped(a, n, ch.gemm_shape[mode], ch.gemm_order[mode]);
}

static void batch_mode_split(BatchPlan& plan, int g, int mode, const BatchChoice& ch)
{
	const ModeArgs a = batch_mode_args(*plan.grids[g], mode, plan.half);
	mode_product_split(a.X, a.ldx, a.P, a.C, a.Kd, a.Y, a.ldy, a.Cp, a.tri_cols, a.tri_rows, a.p_base, a.tet_C1, false,
	                   plan.split_ws, ch.gemm_order[mode]);
}

// gather descriptors of the two branches (lead grid alone, the others together)
static void batch_gather_descs_branches(BatchPlan& plan, int ept)
{
	if (plan.gdesc2_ept == ept) { return; }
	std::vector<GatherDesc> lead, rest;
	int64_t pl = 0, pr = 0;
	for (size_t gi = 0; gi < plan.grids.size(); ++gi) {
		const BatchGrid& g = *plan.grids[gi];
		GatherDesc d = gather_desc(g.s, g.d_idx, g.num, g.V.p, g.d_vals, plan.d_params.p + g.scale_off);
		d.pos = plan.cached ? g.pos.p : nullptr;
		d.perm = plan.sorted ? g.perm.p : nullptr;
		const bool is_lead = (int)gi == plan.lead;
		int64_t& pieces = is_lead ? pl : pr;
		d.piece_begin = pieces;
		pieces += gather_pieces(g.num, ept);
		(is_lead ? lead : rest).push_back(d);
	}
	plan.d_gdesc_lead.reserve(lead.size());
	plan.d_gdesc_rest.reserve(std::max<size_t>(rest.size(), 1));
	GLI_HIP(hipMemcpyAsync(plan.d_gdesc_lead.p, lead.data(), lead.size() * sizeof(GatherDesc), hipMemcpyHostToDevice, ctx().stream));
	if (!rest.empty()) {
		GLI_HIP(hipMemcpyAsync(plan.d_gdesc_rest.p, rest.data(), rest.size() * sizeof(GatherDesc), hipMemcpyHostToDevice, ctx().stream));
	}
	GLI_HIP(hipStreamSynchronize(ctx().stream));
	plan.gdesc2_ept = ept;
	plan.pieces_lead = pl;
	plan.pieces_rest = pr;
}

// grouped mode product over the lead grid (lead = true) or over all other grids
static void batch_mode_subset(BatchPlan& plan, int mode, const BatchChoice& ch, bool lead)
{
	ModeArgs a[kMaxGroup];
	int n = 0;
	for (size_t g = 0; g < plan.grids.size(); ++g) {
		if (((int)g == plan.lead) == lead) { a[n++] = batch_mode_args(*plan.grids[g], mode, plan.half); }
	}
	if (n > 0) { mode_product_grouped(a, n, ch.gemm_shape[mode], ch.gemm_order[mode]); }
}

// launches of the library's helpers go to another stream for the lifetime of this object (the
// caller holds the context lock)
struct StreamScope {
	hipStream_t saved;
	explicit StreamScope(hipStream_t s) : saved(ctx().stream) { ctx().stream = s; }
	~StreamScope() { ctx().stream = saved; }
};

// the FP64 evaluation as two branches: the other grids' products and gather on the side stream
// overlap the lead grid's products and gather on the library stream
static void batch_compute_branches(BatchPlan& plan, const BatchChoice& ch)
{
	hipStream_t lib_stream = ctx().stream;
	for (auto& g : plan.grids) { batch_tables(plan, *g); }
	batch_gather_descs_branches(plan, ch.gather_ept);
	GLI_HIP(hipEventRecord(plan.fork, lib_stream));
	GLI_HIP(hipStreamWaitEvent(plan.side, plan.fork, 0));
	{
		StreamScope on_side(plan.side);
		for (int m = 0; m < 3; ++m) { batch_mode_subset(plan, m, ch, false); }
		batch_launch_gather(plan, plan.d_gdesc_rest.p, (int)plan.grids.size() - 1, plan.pieces_rest, ch.gather_ept,
		                    ch.gather_order, ch.gather_lds);
	}
	GLI_HIP(hipEventRecord(plan.join, plan.side));
	for (int m = 0; m < 3; ++m) { batch_mode_subset(plan, m, ch, true); }
	batch_launch_gather(plan, plan.d_gdesc_lead.p, 1, plan.pieces_lead, ch.gather_ept, ch.gather_order, ch.gather_lds);
	GLI_HIP(hipStreamWaitEvent(lib_stream, plan.join, 0));
}

// everything after the parameter upload
static void batch_compute(BatchPlan& plan, const BatchChoice& ch)
{
	if (plan.method != GLI_ERI_SEPARABLE_FP64) {   // separable plans build their tables in batch_separable
		for (auto& g : plan.grids) { batch_tables(plan, *g); }
	}
	const bool any_split = ch.split[0] || ch.split[1] || ch.split[2];
	if (plan.method == GLI_ERI_SEPARABLE_FP64) {
		batch_separable(plan, ch.gemm_order[2]);
	}
	else if (!any_split) {
		for (int m = 0; m < 3; ++m) { batch_mode(plan, m, ch); }
	}
	else {
		// the split workspace holds one factor at a time: grid by grid
		for (size_t g = 0; g < plan.grids.size(); ++g) {
			const BatchGrid& bg = *plan.grids[g];
			split_factor(bg.B.p, bg.s.Cp, bg.s.Cp, bg.s.Mp, plan.split_ws);
			for (int m = 0; m < 3; ++m) {
				if (ch.split[m]) { batch_mode_split(plan, (int)g, m, ch); }
				else { mode_product_shaped(batch_mode_args(bg, m, plan.half), ch.gemm_shape[m], ch.gemm_order[m]); }
			}
		}
	}
	batch_gather_descs(plan, ch.gather_ept);
	batch_launch_gather(plan, plan.d_gdesc.p, (int)plan.grids.size(), plan.total_pieces, ch.gather_ept, ch.gather_order,
	                    ch.gather_lds);
}

// median time (ms) of a few repetitions of f on the library stream
template <typename F>
static double time_ms(F f, int reps = 5)
{
	hipStream_t st = ctx().stream;
	hipEvent_t e0, e1;
	GLI_HIP(hipEventCreate(&e0));
	GLI_HIP(hipEventCreate(&e1));
	f();
	std::vector<double> ts;
	for (int r = 0; r < reps; ++r) {
		GLI_HIP(hipEventRecord(e0, st));
		f();
		GLI_HIP(hipEventRecord(e1, st));
		GLI_HIP(hipEventSynchronize(e1));
		float ms = 0;
		GLI_HIP(hipEventElapsedTime(&ms, e0, e1));
		ts.push_back(ms);
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	std::sort(ts.begin(), ts.end());
	return ts[ts.size() / 2];
}

static std::string fmt(const char* f, double v)
{
	char b[64];
	std::snprintf(b, sizeof(b), f, v);
	return b;
}

// the autotuner: every kernel choice is timed on the plan's own shapes and data
static BatchChoice batch_autotune(BatchPlan& plan)
{
	BatchChoice ch;
	std::string rep = "{\"modes\": [";
	if (plan.method != GLI_ERI_SEPARABLE_FP64) {
		for (auto& g : plan.grids) { batch_tables(plan, *g); }
	}
	// a mixed plan always runs the split-precision products (the autotuner picks among their
	// variants only), so that the method a plan reports is the arithmetic it executes
	const bool mixed = plan.method == GLI_ERI_TENSOR_MIXED;
	const bool separable = plan.method == GLI_ERI_SEPARABLE_FP64;
	if (separable) {
		// one rank-Q product per grid in one grouped launch: only its tile order is a choice
		double best = 1e30;
		rep += "{\"rank_q\": [";
		for (int order = 0; order < 2; ++order) {
			const double t = time_ms([&] { batch_separable(plan, order); });
			rep += std::string(order ? ", " : "") + "{\"order\": " + std::to_string(order) + ", \"ms\": " + fmt("%.4f", t) + "}";
			if (t < best) {
				best = t;
				ch.gemm_order[2] = order;
			}
		}
		batch_separable(plan, ch.gemm_order[2]);
		rep += "], \"chosen_order\": " + std::to_string(ch.gemm_order[2]) + "}";
	}
	for (int m = 0; m < 3 && !separable; ++m) {
		double best = 1e30;
		rep += std::string(m ? ", " : "") + "{\"mode\": " + std::to_string(m + 1) + ", \"candidates\": [";
		bool first = true;
		for (int shape = 0; shape < kGemmShapes && !mixed; ++shape) {
			bool fits = true;
			for (auto& g : plan.grids) {
				const ModeArgs a = batch_mode_args(*g, m, plan.half);
				fits = fits && gemm_shape_fits(shape, a.Cp, a.tri_cols || a.tri_rows || a.tet_C1 > 0 || a.sym_x, a.p_base);
			}
			if (!fits) { continue; }
			for (int order = 0; order < 2; ++order) {
				BatchChoice c = ch;
				c.gemm_shape[m] = shape;
				c.gemm_order[m] = order;
				c.split[m] = false;
				const double t = time_ms([&] { batch_mode(plan, m, c); });
				rep += std::string(first ? "" : ", ") + "{\"fp64_shape\": " + std::to_string(shape) + ", \"order\": " +
				       std::to_string(order) + ", \"ms\": " + fmt("%.4f", t) + "}";
				first = false;
				if (t < best) {
					best = t;
					ch.gemm_shape[m] = shape;
					ch.gemm_order[m] = order;
					ch.split[m] = false;
				}
			}
		}
		if (mixed) {
			for (int order = 0; order < 2; ++order) {
				BatchChoice c = ch;
				c.gemm_order[m] = order;
				const double t = time_ms([&] {
					for (size_t g = 0; g < plan.grids.size(); ++g) {
						const BatchGrid& bg = *plan.grids[g];
						split_factor(bg.B.p, bg.s.Cp, bg.s.Cp, bg.s.Mp, plan.split_ws);
						batch_mode_split(plan, (int)g, m, c);
					}
				});
				rep += std::string(first ? "" : ", ") + "{\"split_bf16\": 1, \"order\": " + std::to_string(order) +
				       ", \"ms\": " + fmt("%.4f", t) + "}";
				first = false;
				if (t < best) {
					best = t;
					ch.gemm_order[m] = order;
					ch.split[m] = true;
				}
			}
		}
		// leave this mode's output in place for the next mode's timing
		if (ch.split[m]) {
			for (size_t g = 0; g < plan.grids.size(); ++g) {
				const BatchGrid& bg = *plan.grids[g];
				split_factor(bg.B.p, bg.s.Cp, bg.s.Cp, bg.s.Mp, plan.split_ws);
				batch_mode_split(plan, (int)g, m, ch);
			}
		}
		else {
			batch_mode(plan, m, ch);
		}
		rep += "], \"chosen\": " + std::string(ch.split[m] ? "\"split_bf16\"" : "\"fp64\"") + ", \"shape\": " +
		       std::to_string(ch.gemm_shape[m]) + ", \"order\": " + std::to_string(ch.gemm_order[m]) + "}";
	}
	rep += "], \"gather\": [";
	double best = 1e30;
	bool first = true;
	for (int lds = 0; lds < 2; ++lds) {
		if (lds && (plan.cached || plan.lds_bytes > kGatherLdsMax)) { continue; }   // no key tables with cached positions
		for (int ept : { 2, 4, 8, 16, kSortEpt }) {
			if (plan.cached && ept == 2) { continue; }
			if (plan.sorted != (ept == kSortEpt)) { continue; }   // sorted plans: the windows are the pieces
			for (int order = 0; order < 2; ++order) {
				batch_gather_descs(plan, ept);
				const double t = time_ms([&] {
					batch_launch_gather(plan, plan.d_gdesc.p, (int)plan.grids.size(), plan.total_pieces, ept, order, lds);
				});
				rep += std::string(first ? "" : ", ") + "{\"ept\": " + std::to_string(ept) + ", \"order\": " + std::to_string(order) +
				       ", \"lds\": " + std::to_string(lds) + ", \"ms\": " + fmt("%.4f", t) + "}";
				first = false;
				if (t < best) {
					best = t;
					ch.gather_ept = ept;
					ch.gather_order = order;
					ch.gather_lds = lds;
				}
			}
		}
	}
	rep += "], \"gather_chosen\": {\"ept\": " + std::to_string(ch.gather_ept) + ", \"order\": " + std::to_string(ch.gather_order) +
	       ", \"lds\": " + std::to_string(ch.gather_lds) + ", \"cached_positions\": " + (plan.cached ? "true" : "false") +
	       ", \"sorted_windows\": " + (plan.sorted ? "true" : "false") + "}}";
	plan.report = rep;
	return ch;
}

static BatchPlan* batch_create(int n, const double* const* coeffs, const int64_t* K, const int64_t* grids,
                               const int64_t* const* d_idx, const int64_t* num, double tol, int method, double* const* d_values,
                               int autotune)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	GLI_REQUIRE(n >= 1 && n <= kMaxGroup, "a batch holds 1 to 16 grids");
	GLI_REQUIRE(method == GLI_ERI_TENSOR_FP64 || method == GLI_ERI_TENSOR_MIXED || method == GLI_ERI_SEPARABLE_FP64,
	            "a batch runs the key-tensor methods (FP64, mixed or separable)");
	GLI_REQUIRE(tol >= 0, "tolerance must be non-negative");
	std::unique_ptr<BatchPlan> plan(new BatchPlan());
	plan->method = method;
	plan->tol = tol;
	plan->half = (method == GLI_ERI_TENSOR_FP64) && ctx().mode1_half;
	size_t params = 0;
	for (int g = 0; g < n; ++g) {
		GLI_REQUIRE(coeffs[g] != nullptr && num[g] >= 0, "bad batch entry");
		std::unique_ptr<BatchGrid> bg(new BatchGrid());
		bg->K = K[g];
		bg->sym = mirror_symmetric(coeffs[g], K[g]);
		bg->s = eri_shape(K[g], grids + 6 * g, bg->sym);
		GLI_REQUIRE(method != GLI_ERI_TENSOR_MIXED || bg->s.Mp <= 256,
		            "the split-precision method needs 6(n-1) + 2(K-1) < 256; use the FP64 method");
		bg->d_idx = d_idx[g];
		bg->num = num[g];
		bg->d_vals = d_values[g];
		const EriSetup& s = bg->s;
		bg->P.alloc((size_t)s.R * s.R * (4 * s.K - 3));
		bg->B.alloc((size_t)s.Mp * s.Cp);
		bg->V.alloc((size_t)tet_index(0, 0, s.C1));
		if (method == GLI_ERI_SEPARABLE_FP64) {
			// the Gaussian factors per folded range, the rank-Q factors per grid; no T3, Y1 or Y2
			for (auto& t : plan->sep_gw) {
				if (t.first == s.M) { bg->sep_gw = t.second->p; }
			}
			int Qp = 0, Q = 0;
			const std::vector<double> gw = separable_gw(s.M, s.Mp, Qp, Q);
			if (bg->sep_gw == nullptr) {
				std::unique_ptr<DevBuf<double>> d(new DevBuf<double>(gw.size()));
				GLI_HIP(hipMemcpy(d->p, gw.data(), gw.size() * sizeof(double), hipMemcpyHostToDevice));
				bg->sep_gw = d->p;
				plan->sep_gw.emplace_back(s.M, std::move(d));
			}
			bg->sep_Qp = Qp;
			bg->sep_Q = Q;
			bg->GU.alloc((size_t)2 * Qp * s.Cp);
			bg->coeff_off = params;
			params += (size_t)K[g];
			plan->grids.push_back(std::move(bg));
			continue;
		}
		bg->Y1.alloc((size_t)s.Mp * s.Mp * s.Cp);
		bg->Y2.alloc((size_t)s.Mp * s.Cp * s.Cp);
		bool found = false;
		for (auto& t : plan->t3) {
			if (t.first == s.M) {
				bg->T3 = t.second->p;
				found = true;
			}
		}
		if (!found) {
			std::unique_ptr<DevBuf<double>> t3(new DevBuf<double>((size_t)s.Mp * s.Mp * s.Mp));
			launch_kernel_tensor(s, t3->p);
			bg->T3 = t3->p;
			plan->t3.emplace_back(s.M, std::move(t3));
		}
		bg->coeff_off = params;
		params += (size_t)K[g];
		plan->grids.push_back(std::move(bg));
	}
	for (auto& g : plan->grids) {
		g->scale_off = params++;
		const GatherDesc gd = gather_desc(g->s, nullptr, 0, nullptr, nullptr, nullptr);
		plan->lds_bytes = std::max(plan->lds_bytes, gather_lds_bytes(gd.nlut, gd.C1));
	}
	plan->n_params = params;
	plan->d_params.alloc(params);
	GLI_HIP(hipHostMalloc((void**)&plan->h_params, params * sizeof(double), hipHostMallocDefault));
	for (int g = 0; g < n; ++g) {
		std::memcpy(plan->h_params + plan->grids[g]->coeff_off, coeffs[g], K[g] * sizeof(double));
		plan->h_params[plan->grids[g]->scale_off] = 1.0;
	}
	if (method == GLI_ERI_TENSOR_MIXED) { plan->split_ws = split_workspace_new(); }
	GLI_HIP(hipMemcpyAsync(plan->d_params.p, plan->h_params, params * sizeof(double), hipMemcpyHostToDevice, c.stream));

	// every element's position in its grid's key tensor, once per plan: the index lists are fixed for
	// the plan's lifetime, so each replay reads 4-byte positions instead of decoding 8-byte four-indices
	bool cache_ok = c.batch_positions != 0 && plan->lds_bytes <= kGatherLdsMax;
	for (auto& g : plan->grids) { cache_ok = cache_ok && g->s.ks->tet32; }
	if (cache_ok) {
		for (auto& g : plan->grids) {
			g->pos.alloc((size_t)std::max<int64_t>(g->num, 1));
			GatherDesc d = gather_desc(g->s, g->d_idx, g->num, nullptr, nullptr, nullptr);
			d.pos = g->pos.p;
			launch_positions(d);
			if (c.batch_sorted && g->num > 0) {
				g->perm.alloc((size_t)g->num);
				k_sort_windows<<<(unsigned)ceil_div(g->num, kSortWin), kSortThreads, 0, c.stream>>>(g->pos.p, g->perm.p, g->num);
				GLI_HIP(hipGetLastError());
			}
		}
		GLI_HIP(hipStreamSynchronize(c.stream));
		plan->cached = true;
		plan->sorted = c.batch_sorted != 0;
	}

	BatchChoice ch;
	if (autotune) { ch = batch_autotune(*plan); }
	else {
		for (int m = 0; m < 3; ++m) {
			ch.split[m] = (method == GLI_ERI_TENSOR_MIXED);
			ch.gemm_order[m] = c.tile_grouped;
			ch.gemm_shape[m] = c.gemm_shape;
			for (auto& g : plan->grids) {
				const ModeArgs a = batch_mode_args(*g, m, plan->half);
				if (!gemm_shape_fits(ch.gemm_shape[m], a.Cp, a.tri_cols || a.tri_rows || a.tet_C1 > 0 || a.sym_x, a.p_base)) { ch.gemm_shape[m] = 0; }
			}
		}
		ch.gather_ept = plan->sorted ? kSortEpt : c.gather_ept;
		int64_t nV_max = 0;
		for (auto& g : plan->grids) { nV_max = std::max(nV_max, tet_index(0, 0, g->s.C1)); }
		ch.gather_order = resolve_gather_order(c.gather_xcd, nV_max);
		ch.gather_lds = c.gather_lds;
		plan->report = "{\"autotuned\": false}";
	}
	plan->choice = ch;
	// one eager pass sizes every workspace, then the same sequence is captured
	batch_compute(*plan, ch);
	GLI_HIP(hipStreamSynchronize(c.stream));
	BatchPlan& pl = *plan;
	auto capture = [&](bool branches, hipGraph_t* graph, hipGraphExec_t* exec) {
		GLI_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
		try {
			GLI_HIP(hipMemcpyAsync(pl.d_params.p, pl.h_params, params * sizeof(double), hipMemcpyHostToDevice, c.stream));
			if (branches) { batch_compute_branches(pl, ch); }
			else { batch_compute(pl, ch); }
		}
		catch (...) {
			hipGraph_t broken = nullptr;
			(void)hipStreamEndCapture(c.stream, &broken);
			if (broken) { (void)hipGraphDestroy(broken); }
			throw;
		}
		GLI_HIP(hipStreamEndCapture(c.stream, graph));
		GLI_HIP(hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0));
	};
	capture(false, &plan->graph, &plan->exec);
	// two-branch variant: FP64 plans of two or more grids, timed against the single branch when
	// autotuning (knob batch_branches: 0 never, 1 autotuner decides, 2 always)
	const bool eligible = n >= 2 && !(ch.split[0] || ch.split[1] || ch.split[2]) && method != GLI_ERI_SEPARABLE_FP64;
	if (eligible && c.batch_branches > 0 && (autotune || c.batch_branches == 2)) {
		int64_t most = -1;
		for (int g = 0; g < n; ++g) {
			if (plan->grids[g]->num > most) { most = plan->grids[g]->num; plan->lead = g; }
		}
		GLI_HIP(hipStreamCreateWithFlags(&plan->side, hipStreamNonBlocking));
		GLI_HIP(hipEventCreateWithFlags(&plan->fork, hipEventDisableTiming));
		GLI_HIP(hipEventCreateWithFlags(&plan->join, hipEventDisableTiming));
		batch_compute_branches(*plan, ch);   // eager pass: descriptors and a check that it runs
		GLI_HIP(hipStreamSynchronize(plan->side));
		GLI_HIP(hipStreamSynchronize(c.stream));
		hipGraph_t g2 = nullptr;
		hipGraphExec_t e2 = nullptr;
		capture(true, &g2, &e2);
		bool use = c.batch_branches == 2;
		std::string note = "\"forced\"";
		if (!use) {
			const double t1 = time_ms([&] { GLI_HIP(hipGraphLaunch(plan->exec, c.stream)); });
			const double t2 = time_ms([&] { GLI_HIP(hipGraphLaunch(e2, c.stream)); });
			use = t2 < t1;
			note = "{\"one_branch_ms\": " + fmt("%.4f", t1) + ", \"two_branch_ms\": " + fmt("%.4f", t2) + "}";
		}
		if (use) {
			GLI_HIP(hipGraphExecDestroy(plan->exec));
			GLI_HIP(hipGraphDestroy(plan->graph));
			plan->exec = e2;
			plan->graph = g2;
			plan->branches = true;
		}
		else {
			GLI_HIP(hipGraphExecDestroy(e2));
			GLI_HIP(hipGraphDestroy(g2));
		}
		std::string& r = plan->report;
		if (!r.empty() && r.back() == '}') {
			r.pop_back();
			r += std::string(r.size() > 1 ? ", " : "") + "\"branches\": {\"timing\": " + note + ", \"chosen\": " +
			     (use ? "\"two\"" : "\"one\"") + ", \"lead_grid\": " + std::to_string(plan->lead) + "}}";
		}
	}
	GLI_HIP(hipEventCreateWithFlags(&plan->done, hipEventDisableTiming));
	return plan.release();
}

static void batch_run(BatchPlan& plan, const double* const* coeffs, const double* scale)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	if (plan.launched) { GLI_HIP(hipEventSynchronize(plan.done)); }
	for (size_t g = 0; g < plan.grids.size(); ++g) {
		const BatchGrid& bg = *plan.grids[g];
		if (coeffs != nullptr) {
			GLI_REQUIRE(coeffs[g] != nullptr, "missing coefficients in a batch run");
			GLI_REQUIRE(mirror_symmetric(coeffs[g], bg.K) == bg.sym,
			            "a batch plan is built for one coefficient symmetry per grid: mirror symmetry changed");
			std::memcpy(plan.h_params + bg.coeff_off, coeffs[g], bg.K * sizeof(double));
		}
		plan.h_params[bg.scale_off] = (scale != nullptr) ? scale[g] : 1.0;
	}
	GLI_HIP(hipGraphLaunch(plan.exec, c.stream));
	GLI_HIP(hipEventRecord(plan.done, c.stream));
	plan.launched = true;
}

}  // namespace gli

using namespace gli;

extern "C" {

int gli_eri_values_device(const double* coeffs, int64_t K, const int64_t grid[6], const int64_t* d_idx, int64_t num,
                          double tol, int method, double* d_values)
{
	try {
		eri_values_dev(coeffs, K, grid, d_idx, num, tol, method, d_values);
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_hip_eri_values_copy(const double* coeffs, int64_t K, const int64_t g[6], const int64_t* indices, int64_t num,
                            double tol, int method, double* values, int64_t* indices_copy)
{
	try {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		if (num == 0) { return 0; }
		GLI_REQUIRE(num > 0, "negative number of indices");
		const EriPrepared ep = eri_prepare(coeffs, K, g, tol, method, true);
		if ((size_t)num * sizeof(int64_t) <= stream_chunk_bytes()) {
			// one chunk: plain copies on the library stream
			DevBuf<int64_t> di(num);
			DevBuf<double> dv(num);
			GLI_HIP(hipMemcpyAsync(di.p, indices, num * sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
			if (indices_copy != nullptr) { std::memcpy(indices_copy, indices, num * sizeof(int64_t)); }
			eri_apply(ep, di.p, num, dv.p, false);
			GLI_HIP(hipMemcpyAsync(values, dv.p, num * sizeof(double), hipMemcpyDeviceToHost, c.stream));
			GLI_HIP(hipStreamSynchronize(c.stream));
			return 0;
		}
		// indices up and values down in chunks, each chunk's gather between its two copies
		stream_chunks(indices, sizeof(int64_t), values, sizeof(double), num, indices_copy,
		              [&](const void* d_in, void* d_out, int64_t n) { eri_apply(ep, (const int64_t*)d_in, n, (double*)d_out, false); });
		GLI_HIP(hipStreamSynchronize(c.stream));
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_sparse_eri_compute(const struct gausslet_data* gdata, const struct cartesian_grid_3d* grid, double tol,
                           struct sparse_eri_gausslet_integrals* eri)
{
	int64_t* d_idx = nullptr;
	double* d_val = nullptr;
	int64_t* h_idx = nullptr;
	double* h_val = nullptr;
	try {
		GLI_REQUIRE(gdata != nullptr && grid != nullptr && eri != nullptr, "null argument");
		const int64_t K = gdata->indices.num;
		GLI_REQUIRE(gdata->coefficients != nullptr && K >= 3 && K % 2 == 1 && gdata->indices.istart == -(K - 1) / 2,
		            "the gausslet coefficients need an odd count K >= 3 with indices -(K-1)/2 .. (K-1)/2");
		const int64_t n = grid->coord_range[0].num;
		for (int a = 0; a < 3; ++a) {
			GLI_REQUIRE(grid->coord_range[a].num == n && n % 2 == 1 && grid->coord_range[a].istart == -(n - 1) / 2,
			            "the enumeration needs a cubic grid centred at the origin with an odd number of points per axis");
		}
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		int64_t num = 0;
		if (gli_eri_enumerate_device(n, &d_idx, &num) != 0) { throw Failure{"enumeration"}; }
		int64_t g6[6];
		for (int a = 0; a < 3; ++a) {
			g6[2 * a] = grid->coord_range[a].istart;
			g6[2 * a + 1] = grid->coord_range[a].num;
		}
		GLI_HIP(hipMalloc((void**)&d_val, (size_t)std::max<int64_t>(num, 1) * sizeof(double)));
		eri_values_dev(gdata->coefficients, K, g6, d_idx, num, tol, gli_get_eri_method(), d_val);
		h_idx = (int64_t*)gli_host_alloc((size_t)num * sizeof(int64_t));
		h_val = (double*)gli_host_alloc((size_t)num * sizeof(double));
		GLI_REQUIRE(h_idx != nullptr && h_val != nullptr, "host allocation of the result arrays failed");
		copy_to_host(h_idx, d_idx, (size_t)num * sizeof(int64_t));
		copy_to_host(h_val, d_val, (size_t)num * sizeof(double));
		// each pointer is cleared before its free can fail, so the error path frees it at most once
		int64_t* fi = d_idx;
		d_idx = nullptr;
		GLI_HIP(hipFree(fi));
		double* fv = d_val;
		d_val = nullptr;
		GLI_HIP(hipFree(fv));
		eri->grid = *grid;
		eri->num_entries = num;
		eri->four_indices = h_idx;
		eri->integral_values = h_val;
		return 0;
	}
	catch (const Failure&) {
		if (d_idx != nullptr) { (void)hipFree(d_idx); }
		if (d_val != nullptr) { (void)hipFree(d_val); }
		free(h_idx);
		free(h_val);
		return -1;
	}
}

int gli_eri_values_device_noalloc(const double* coeffs, int64_t K, const int64_t grid[6], const int64_t* d_idx, int64_t num,
                                  double tol, int method, double* d_vals)
{
	try {
		eri_values_noalloc(coeffs, K, grid, d_idx, num, tol, method, d_vals);
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_hip_eri_values(const double* coeffs, int64_t K, const int64_t g[6], const int64_t* indices, int64_t num,
                       double tol, int method, double* values)
{
	return gli_hip_eri_values_copy(coeffs, K, g, indices, num, tol, method, values, nullptr);
}

int gli_eri_integrals_points(const double* coeffs, int64_t K, const int64_t* points, int64_t num, double tol, int method,
                             double* values)
{
	try {
		int64_t g[6];
		std::vector<int64_t> idx;
		points_to_indices(points, num, 4, g, idx);
		GLI_REQUIRE(num == 0 || values != nullptr, "no output array given");
		if (num == 0) { return 0; }
		return gli_hip_eri_values(coeffs, K, g, idx.data(), num, tol, method, values);
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_eri_values_batch_device(int num_grids, const double* const* coeffs, const int64_t* K, const int64_t* grids,
                                const int64_t* const* d_idx, const int64_t* num, double tol, int method, double* const* d_values)
{
	try {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		if (method == GLI_ERI_DIRECT_FP64) {
			for (int b = 0; b < num_grids; ++b) {
				eri_values_dev(coeffs[b], K[b], grids + 6 * b, d_idx[b], num[b], tol, method, d_values[b]);
			}
			return 0;
		}
		std::unique_ptr<BatchPlan> plan(batch_create(num_grids, coeffs, K, grids, d_idx, num, tol, method, d_values, 0));
		batch_run(*plan, nullptr, nullptr);
		GLI_HIP(hipEventSynchronize(plan->done));
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_eri_batch_create(int num_grids, const double* const* coeffs, const int64_t* K, const int64_t* grids,
                         const int64_t* const* d_idx, const int64_t* num, double tol, int method, double* const* d_values,
                         int autotune, gli_eri_batch** plan)
{
	try {
		GLI_REQUIRE(plan != nullptr, "null plan pointer");
		*plan = nullptr;
		*plan = reinterpret_cast<gli_eri_batch*>(batch_create(num_grids, coeffs, K, grids, d_idx, num, tol, method, d_values, autotune));
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_eri_batch_run(gli_eri_batch* plan, const double* const* coeffs, const double* scale)
{
	try {
		GLI_REQUIRE(plan != nullptr, "null plan");
		batch_run(*reinterpret_cast<BatchPlan*>(plan), coeffs, scale);
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_eri_batch_describe(const gli_eri_batch* plan, char* buf, int len)
{
	if (plan == nullptr) { return -1; }
	const std::string& r = reinterpret_cast<const BatchPlan*>(plan)->report;
	if (buf != nullptr && len > 0) {
		const size_t k = std::min(r.size(), (size_t)len - 1);
		std::memcpy(buf, r.data(), k);
		buf[k] = '\0';
	}
	return (int)r.size();
}

void gli_eri_batch_destroy(gli_eri_batch* plan)
{
	if (plan == nullptr) { return; }
	try {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		delete reinterpret_cast<BatchPlan*>(plan);
	}
	catch (const Failure&) {
	}
}

int gli_mode_product_device(const double* X, int64_t ldx, int64_t P, const double* B, int64_t ldb, int64_t C, int64_t Kd,
                            double* Y, int64_t ldy)
{
	try {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		GLI_REQUIRE(P >= 0 && C >= 0 && Kd >= 0 && ldx >= P && ldb >= C && ldy >= C, "inconsistent mode-product shapes");
		mode_product(X, ldx, P, B, ldb, (int)C, (int)Kd, Y, ldy, 0, 0, 0, 0, 0);
		GLI_HIP(hipStreamSynchronize(c.stream));
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

// the separable method's quadrature rule for a folded range M (host arithmetic only)
int gli_eri_separable_rule(int M, double* nodes, double* weights, int max)
{
	if (M < 0) { return -1; }
	std::vector<long double> sq, wq;
	separable_rule(M, sq, wq);
	const int Q = (int)sq.size();
	if (nodes != nullptr && weights != nullptr) {
		for (int q = 0; q < Q && q < max; ++q) {
			nodes[q] = (double)sq[(size_t)q];
			weights[q] = (double)wq[(size_t)q];
		}
	}
	return Q;
}

// number of distinct folded columns (axis keys) the contraction uses on an R-point axis
int gli_eri_key_count(int64_t R, int symmetric)
{
	try {
		return key_set((int)R, symmetric != 0).C1;
	}
	catch (const Failure&) {
		return -1;
	}
}

void gli_release_workspace(void)
{
	std::lock_guard<std::recursive_mutex> lk(ctx().mu);
	EriWorkspace& w = eri_ws();
	w.coeffs.release(); w.P.release(); w.B.release();
	w.Y1.release(); w.Y2.release(); w.V.release(); w.T3.release();
	w.T3_M = -1;
	w.GU.release(); w.sep_gw.release();
	w.sep_M = -1;
	w.rep_off.release(); w.rep_list.release();
	w.rep_M = -1;
}

int gli_enable_direct_flop_count(int on)
{
	try {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		EriWorkspace& w = eri_ws();
		if (on) {
			w.direct_sums.reserve(2);
			GLI_HIP(hipMemsetAsync(w.direct_sums.p, 0, 2 * sizeof(unsigned long long), c.stream));
		}
		w.count_direct = (on != 0);
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_direct_flops(double* reference_form, double* executed)
{
	try {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		EriWorkspace& w = eri_ws();
		GLI_REQUIRE(w.count_direct, "flop counting of the per-element path is not enabled");
		unsigned long long h[2];
		GLI_HIP(hipMemcpyAsync(h, w.direct_sums.p, sizeof(h), hipMemcpyDeviceToHost, c.stream));
		GLI_HIP(hipMemsetAsync(w.direct_sums.p, 0, sizeof(h), c.stream));
		GLI_HIP(hipStreamSynchronize(c.stream));
		if (reference_form) { *reference_form = (double)h[0]; }
		if (executed) { *executed = (double)h[1]; }
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

}  // extern "C"
