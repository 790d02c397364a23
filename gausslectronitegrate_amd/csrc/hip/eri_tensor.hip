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
#include "../gli_device.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>

namespace gli {

typedef double d4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// 1D tables

// f[s][m] = (sqrt(pi)/3) sum_{a+b=m} c_a c_b exp(-(0.5 (3 s + a - b))^2), s = 0 .. R-1
// (gausslet_factors.c:12-50; f_{-s} = f_s).  Same ascending-a summation order.
__global__ void k_factors(const double* __restrict__ c, int K, double* __restrict__ f)
{
	const int L = 2 * K - 1;
	const int s = blockIdx.y;
	const int m = blockIdx.x * blockDim.x + threadIdx.x;
	if (m >= L) { return; }
	const double prefac = 0.590817950301838675766;
	double acc = 0;
	const int ilo = max(0, m - (K - 1)), ihi = min(K - 1, m);
	for (int i = ilo; i <= ihi; ++i) {
		const int j = m - i;
		const double x = 0.5 * (3 * s + (i - j));
		acc += prefac * c[i] * c[j] * exp(-x * x);
	}
	f[(int64_t)s * L + m] = acc;
}

// P[a][b][t] = sum_m f_a[m] f_b[m - (t - (2K-2))], zeroed where |P| <= tol
// (eri_integrals.c:59-103 with the tolerance skips of 629/637/658 folded in)
__global__ void k_products(const double* __restrict__ f, int K, int R, double tol, double* __restrict__ P)
{
	const int L = 2 * K - 1, LP = 4 * K - 3;
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	const int ab = blockIdx.y;
	if (t >= LP) { return; }
	const int a = ab / R, b = ab % R;
	const int d = t - (2 * K - 2);  // m - m'
	const double* fa = f + (int64_t)a * L;
	const double* fb = f + (int64_t)b * L;
	double acc = 0;
	const int mlo = max(0, d), mhi = min(L - 1, L - 1 + d);
	for (int m = mlo; m <= mhi; ++m) { acc += fa[m] * fb[m - d]; }
	P[(int64_t)ab * LP + t] = (fabs(acc) > tol) ? acc : 0.0;
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
// 64x64 output tile per workgroup of 4 waves (each 32x32 = 2x2 MFMA blocks),
// mu staged through LDS 16 at a time.  Optional triangular skipping for the
// symmetric tensor: with p = outer * Cp + r,
//   tri_cols: only k >= r is needed;  tri_rows: only outer <= r is needed.
// Tiles entirely outside the needed region are not computed (never read later).

constexpr int GBM = 64, GBN = 64, GBK = 16;

// position of the sorted key triple k1 <= k2 <= k3 in the tetrahedral layout of a symmetric 3-tensor
__host__ __device__ inline int64_t tet_index(int k1, int k2, int k3)
{
	return (int64_t)k3 * (k3 + 1) * (k3 + 2) / 6 + (int64_t)k2 * (k2 + 1) / 2 + k1;
}

__global__ __launch_bounds__(256) void k_mode_product(
	const double* __restrict__ X, int64_t ldx, int64_t P, const double* __restrict__ B, int64_t ldb, int C, int Kd,
	double* __restrict__ Y, int64_t ldy, int Cp, int tri_cols, int tri_rows, int64_t p_base, int tet_C1, TileOrder order)
{
	__shared__ double As[GBK][GBM];
	__shared__ double Bs[GBK][GBN];
	const int tid = threadIdx.x;
	const int lane = tid & 63, wave = tid >> 6;
	int rt, ct;
	if (!order.tile(blockIdx.x, rt, ct)) { return; }
	const int64_t p0 = (int64_t)rt * GBM;
	const int k0 = ct * GBN;

	if (tri_cols || tri_rows) {
		// global row index (with p_base) decomposes as outer * Cp + r; Cp % GBM == 0 keeps outer fixed per tile
		const int64_t gp = p_base + p0;
		const int64_t outer = gp / Cp;
		const int64_t rlo = gp % Cp, rhi = rlo + GBM - 1;
		if (tri_cols && k0 + GBN - 1 < rlo) { return; }
		if (tri_rows && rhi < outer) { return; }
	}

	const int wm = wave >> 1, wn = wave & 1;
	d4 acc[2][2];
	for (int i = 0; i < 2; ++i) {
		for (int j = 0; j < 2; ++j) { acc[i][j] = d4{0, 0, 0, 0}; }
	}

	// software pipeline: the next mu-tile is loaded into registers while the current one
	// (already in LDS) feeds the MFMAs
	constexpr int PER = (GBK * GBM) / 256;
	double ra[PER], rb[PER];
	auto gload = [&](int mu0) {
#pragma unroll
		for (int r = 0; r < PER; ++r) {
			const int id = tid + 256 * r;
			const int mu = id / GBM, p = id % GBM;
			ra[r] = (mu0 + mu < Kd && p0 + p < P) ? X[(int64_t)(mu0 + mu) * ldx + p0 + p] : 0.0;
			const int k = id % GBN;
			rb[r] = (mu0 + mu < Kd && k0 + k < C) ? B[(int64_t)(mu0 + mu) * ldb + k0 + k] : 0.0;
		}
	};
	gload(0);
	for (int mu0 = 0; mu0 < Kd; mu0 += GBK) {
#pragma unroll
		for (int r = 0; r < PER; ++r) {
			const int id = tid + 256 * r;
			As[id / GBM][id % GBM] = ra[r];
			Bs[id / GBN][id % GBN] = rb[r];
		}
		__syncthreads();
		if (mu0 + GBK < Kd) { gload(mu0 + GBK); }
#pragma unroll
		for (int kk = 0; kk < GBK; kk += 4) {
			double a[2], b[2];
			for (int i = 0; i < 2; ++i) { a[i] = As[kk + lane / 16][wm * 32 + i * 16 + lane % 16]; }
			for (int j = 0; j < 2; ++j) { b[j] = Bs[kk + lane / 16][wn * 32 + j * 16 + lane % 16]; }
			for (int i = 0; i < 2; ++i) {
				for (int j = 0; j < 2; ++j) {
					acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
				}
			}
		}
		__syncthreads();
	}

	// accumulator layout of the 16x16 FP64 block (v_mfma_f64_16x16x4f64): lane holds
	// column lane%16 and rows lane/16 + 4 v, v = 0..3
	for (int i = 0; i < 2; ++i) {
		for (int j = 0; j < 2; ++j) {
			const int col = k0 + wn * 32 + j * 16 + lane % 16;
			for (int v = 0; v < 4; ++v) {
				const int64_t row = p0 + wm * 32 + i * 16 + (lane / 16) + 4 * v;
				if (row >= P || col >= C) { continue; }
				if (tet_C1 > 0) {
					const int64_t gp = p_base + row;
					const int k1 = (int)(gp / Cp), k2 = (int)(gp % Cp);
					if (k1 <= k2 && k2 <= col && col < tet_C1) { Y[tet_index(k1, k2, col)] = acc[i][j][v]; }
				}
				else {
					Y[row * ldy + col] = acc[i][j][v];
				}
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

__device__ inline void element_keys(const GridDev& g, const int* __restrict__ lut, const int* __restrict__ coords,
                                    int64_t t, int& k1, int& k2, int& k3)
{
	int i, j, k, l;
	decode_four(g, t, i, j, k, l);
	const int ci = coords[i], cj = coords[j], ck = coords[k], cl = coords[l];
	int kk[3];
	for (int a = 0; a < 3; ++a) {
		const int sh = 10 * a;
		kk[a] = axis_key(lut, g.R, (ci >> sh) & 1023, (cj >> sh) & 1023, (ck >> sh) & 1023, (cl >> sh) & 1023);
	}
	// sort ascending (V is symmetric in its three arguments)
	int x = kk[0], y = kk[1], z = kk[2], s;
	if (x > y) { s = x; x = y; y = s; }
	if (y > z) { s = y; y = z; z = s; }
	if (x > y) { s = x; x = y; y = s; }
	k1 = x; k2 = y; k3 = z;
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
	// reciprocals rounded up by one ulp: for a < 2^12 (float) and a < 2^24 (double, divisor
	// N) resp. a < 2^48 (divisor N^2), trunc(a * inv_up) is the exact quotient, no correction
	// step needed (exhaustively checked for every divisor up to 4096)
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

// values[e] = V[k1][k2][k3] from the tetrahedral key tensor.  Each wave takes GATHER_EPT groups
// of 64 consecutive elements (coalesced index loads and value stores, several dependent load
// chains index -> key table -> V in flight per lane).  In a sorted list the leading pair (i, j)
// changes only every few thousand elements, so it is decoded once per wave; a group whose lanes
// all share it decodes only (k, l) per lane with exact reciprocal multiplies.  Any other group
// (a pair boundary, the tail, unsorted or out-of-range input) takes the general per-lane decode.
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

template <int GATHER_EPT>
__global__ __launch_bounds__(256) void k_gather(const int64_t* __restrict__ idx, int64_t num, GridDev g, DecodeConsts dc,
	const short* __restrict__ lut, const uint32_t* __restrict__ tet3, const double* __restrict__ V, double* __restrict__ out)
{
	const int lane = threadIdx.x & 63;
	const int64_t wave_base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * (64 * GATHER_EPT);
	int64_t t[GATHER_EPT];
#pragma unroll
	for (int u = 0; u < GATHER_EPT; ++u) {
		const int64_t e = wave_base + u * 64 + lane;
		t[u] = (e < num) ? idx[e] : -1;
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
			const int d = abs((int)ci[a] - (int)cj[a]);
			base_key[a] = __builtin_amdgcn_readfirstlane(d * g.R * W + (int)(ci[a] + cj[a]) + off);
		}
		pair_base = uniform_i64(ij * g.N2);
	}
	bool fast = pair_base >= 0;
#pragma unroll
	for (int u = 0; u < GATHER_EPT; ++u) { fast = fast && (uint64_t)(t[u] - pair_base) < (uint64_t)g.N2; }
	double v[GATHER_EPT];
	if (__all(fast)) {
		// all groups of the wave share the pair: phase the loads across groups so that the
		// in-order load counter never makes one group's V read wait for another's key reads
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
				// (|ci - cj| R + |ck - cl|) W + (ci + cj) - (ck + cl) + 2(R - 1), with the pair part uniform
				const unsigned li = __umul24(__usad(ck[a], cl[a], 0u), (unsigned)W) + ((unsigned)base_key[a] - ck[a] - cl[a]);
				kk[u][a] = ulut[li];
			}
		}
		uint32_t vi[GATHER_EPT];
		unsigned k3[GATHER_EPT];
#pragma unroll
		for (int u = 0; u < GATHER_EPT; ++u) {
			const unsigned lo = min(kk[u][0], kk[u][1]), hi = max(kk[u][0], kk[u][1]);
			const unsigned mid = min(hi, kk[u][2]);
			k3[u] = max(hi, kk[u][2]);
			const unsigned k1 = min(lo, mid), k2 = max(lo, mid);
			vi[u] = (__umul24(k2, k2 + 1u) >> 1) + k1;
		}
#pragma unroll
		for (int u = 0; u < GATHER_EPT; ++u) { vi[u] += tet3[k3[u]]; }
#pragma unroll
		for (int u = 0; u < GATHER_EPT; ++u) { v[u] = V[vi[u]]; }
	}
	else {
		int64_t vi[GATHER_EPT];
#pragma unroll
		for (int u = 0; u < GATHER_EPT; ++u) { vi[u] = (t[u] >= 0 && t[u] < g.N4) ? gather_index(g, dc, lut, t[u]) : -1; }
#pragma unroll
		for (int u = 0; u < GATHER_EPT; ++u) { v[u] = (vi[u] >= 0) ? V[vi[u]] : __builtin_nan(""); }
	}
#pragma unroll
	for (int u = 0; u < GATHER_EPT; ++u) {
		const int64_t e = wave_base + u * 64 + lane;
		if (e < num) { out[e] = v[u]; }
	}
}

// per-element FP64 sum over the folded columns, no reuse across elements:
// value = sum_{m1,m2} B1[m1] B2[m2] sum_{m3} B3[m3] T[m1^2 + m2^2 + m3^2]
__global__ __launch_bounds__(256) void k_eri_direct(const int64_t* __restrict__ idx, int64_t num, GridDev g,
	const int* __restrict__ lut, const int* __restrict__ coords, const double* __restrict__ B, int Cp, int M, int Mp, const double* __restrict__ T,
	double* __restrict__ out)
{
	extern __shared__ double smem[];
	double* col = smem;               // 3 x Mp folded columns
	double* red = smem + 3 * Mp;      // 256 partial sums
	__shared__ int win[3][2];
	const int tid = threadIdx.x;
	for (int64_t e = blockIdx.x; e < num; e += gridDim.x) {
		const int64_t t = idx[e];
		if (t < 0 || t >= g.N4) {
			if (tid == 0) { out[e] = __builtin_nan(""); }
			continue;
		}
		int k[3];
		element_keys(g, lut, coords, t, k[0], k[1], k[2]);
		for (int a = 0; a < 3; ++a) {
			for (int m = tid; m < Mp; m += blockDim.x) { col[a * Mp + m] = (m <= M) ? B[(int64_t)m * Cp + k[a]] : 0.0; }
		}
		__syncthreads();
		if (tid < 3) {
			int lo = M + 1, hi = -1;
			for (int m = 0; m <= M; ++m) {
				if (col[tid * Mp + m] != 0) { lo = min(lo, m); hi = m; }
			}
			win[tid][0] = lo;
			win[tid][1] = hi;
		}
		__syncthreads();
		const int n0 = win[0][1] - win[0][0] + 1, n1 = win[1][1] - win[1][0] + 1;
		double acc = 0;
		if (n0 > 0 && n1 > 0 && win[2][1] >= win[2][0]) {
			const double* c0 = col;
			const double* c1 = col + Mp;
			const double* c2 = col + 2 * Mp;
			for (int pq = tid; pq < n0 * n1; pq += blockDim.x) {
				const int m1 = win[0][0] + pq / n1, m2 = win[1][0] + pq % n1;
				const double w = c0[m1] * c1[m2];
				if (w == 0) { continue; }
				const int base = m1 * m1 + m2 * m2;
				double s = 0;
				for (int m3 = win[2][0]; m3 <= win[2][1]; ++m3) { s += c2[m3] * T[base + m3 * m3]; }
				acc += w * s;
			}
		}
		red[tid] = acc;
		__syncthreads();
		for (int w = 128; w > 0; w >>= 1) {
			if (tid < w) { red[tid] += red[tid + w]; }
			__syncthreads();
		}
		if (tid == 0) { out[e] = red[0]; }
		__syncthreads();
	}
}

// ---------------------------------------------------------------------------
// host side: key tables, workspace, orchestration

struct KeySet {
	int R = 0;
	bool symmetric = false;
	int C1 = 0;
	std::vector<int4> keys;  // (a, b, cd, 0) of each distinct folded column
	std::vector<int> lut;    // (a * R + b) * (4R - 3) + cd + 2(R - 1) -> key id
	DevBuf<int4> keys_dev;
	DevBuf<int> lut_dev;
	DevBuf<short> lut16_dev;  // the same table as int16 (key ids stay below 32768)
	DevBuf<uint32_t> tet3_dev;  // tet_index(0, 0, k) per key
	bool tet32 = false;         // every tetrahedral index fits 32 bits
};

static KeySet& key_set(int R, bool symmetric)
{
	static std::map<std::pair<int, bool>, KeySet*> cache;
	auto it = cache.find({R, symmetric});
	if (it != cache.end()) { return *it->second; }
	KeySet* ks = new KeySet();
	ks->R = R;
	ks->symmetric = symmetric;
	const int W = 4 * R - 3;
	ks->lut.assign((size_t)R * R * W, -1);
	std::map<std::tuple<int, int, int>, int> ids;
	// every per-axis 4-tuple with span < R (a translation class per key)
	for (int i = 0; i < R; ++i) {
		for (int j = 0; j < R; ++j) {
			for (int k = 0; k < R; ++k) {
				for (int l = 0; l < R; ++l) {
					const int a = std::abs(i - j), b = std::abs(k - l), cd = (i + j) - (k + l);
					std::tuple<int, int, int> canon = symmetric ? std::make_tuple(std::min(a, b), std::max(a, b), std::abs(cd))
					                                            : std::make_tuple(a, b, cd);
					auto f = ids.find(canon);
					int id;
					if (f == ids.end()) {
						id = (int)ks->keys.size();
						ids[canon] = id;
						ks->keys.push_back(int4{std::get<0>(canon), std::get<1>(canon), std::get<2>(canon), 0});
					}
					else {
						id = f->second;
					}
					ks->lut[(a * R + b) * W + cd + 2 * (R - 1)] = id;
				}
			}
		}
	}
	ks->C1 = (int)ks->keys.size();
	ks->keys_dev.alloc(ks->keys.size());
	ks->lut_dev.alloc(ks->lut.size());
	GLI_HIP(hipMemcpy(ks->keys_dev.p, ks->keys.data(), ks->keys.size() * sizeof(int4), hipMemcpyHostToDevice));
	GLI_HIP(hipMemcpy(ks->lut_dev.p, ks->lut.data(), ks->lut.size() * sizeof(int), hipMemcpyHostToDevice));
	GLI_REQUIRE(ks->C1 < 32768, "too many axis keys for the 16-bit key table");
	std::vector<short> l16(ks->lut.begin(), ks->lut.end());
	ks->lut16_dev.alloc(l16.size());
	GLI_HIP(hipMemcpy(ks->lut16_dev.p, l16.data(), l16.size() * sizeof(short), hipMemcpyHostToDevice));
	// C(k3 + 2, 3): tetrahedral offset of the largest key, as long as every index fits 32 bits
	ks->tet32 = tet_index(0, 0, ks->C1) <= (int64_t)UINT32_MAX;
	std::vector<uint32_t> t3((size_t)ks->C1);
	for (int k = 0; k < ks->C1; ++k) { t3[k] = ks->tet32 ? (uint32_t)tet_index(0, 0, k) : 0u; }
	ks->tet3_dev.alloc(t3.size());
	GLI_HIP(hipMemcpy(ks->tet3_dev.p, t3.data(), t3.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
	cache[{R, symmetric}] = ks;
	return *ks;
}

struct EriWorkspace {
	DevBuf<double> coeffs, f, P, B, Y1, Y2, V;
	DevBuf<int> coords;
	DevBuf<double> T3;
	int T3_M = -1;
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
	const double* T;
};

// tables f, P, B (and T3) for a coefficient set on a grid; all on the library stream
static EriSetup eri_tables(const double* coeffs, int64_t K, const int64_t grid[6], double tol)
{
	GLI_REQUIRE(K >= 3 && (K % 2) == 1, "the number of gausslet coefficients must be odd and at least 3");
	GLI_REQUIRE(grid[1] > 0 && grid[3] > 0 && grid[5] > 0, "grid axes must have a positive number of points");
	GLI_REQUIRE(grid[1] < 1024 && grid[3] < 1024 && grid[5] < 1024, "at most 1023 points per axis");
	GLI_REQUIRE(tol >= 0, "tolerance must be non-negative");
	Context& c = ctx();
	EriWorkspace& w = eri_ws();
	EriSetup s;
	s.K = (int)K;
	s.R = (int)std::max(grid[1], std::max(grid[3], grid[5]));
	s.M = 6 * (s.R - 1) + 2 * (s.K - 1);
	s.Mp = (int)round_up(s.M + 1, 16);
	GLI_REQUIRE(s.Mp <= 2048, "grid too long: the folded kernel range 6(n-1) + 2(K-1) must stay below 2048");
	const bool sym = mirror_symmetric(coeffs, K);
	const KeySet& ks = key_set(s.R, sym);
	s.ks = &ks;
	s.C1 = ks.C1;
	s.Cp = (int)round_up(s.C1, GBM);
	s.g.n[0] = (int)grid[1];
	s.g.n[1] = (int)grid[3];
	s.g.n[2] = (int)grid[5];
	s.g.N = (int)(grid[1] * grid[3] * grid[5]);
	s.g.R = s.R;
	s.g.N2 = (int64_t)s.g.N * s.g.N;
	s.g.N4 = s.g.N2 * s.g.N2;
	const int64_t nT = 3 * (int64_t)s.M * s.M + 1;
	s.T = eri_table(nT);

	const int L = 2 * s.K - 1, LP = 4 * s.K - 3;
	w.coeffs.reserve(K);
	w.f.reserve((size_t)s.R * L);
	w.P.reserve((size_t)s.R * s.R * LP);
	w.B.reserve((size_t)s.Mp * s.Cp);
	w.coords.reserve((size_t)s.g.N);
	upload_async(w.coeffs.p, coeffs, K * sizeof(double));
	k_factors<<<dim3((unsigned)ceil_div(L, 128), s.R), 128, 0, c.stream>>>(w.coeffs.p, s.K, w.f.p);
	k_products<<<dim3((unsigned)ceil_div(LP, 128), s.R * s.R), 128, 0, c.stream>>>(w.f.p, s.K, s.R, tol, w.P.p);
	k_fold<<<dim3((unsigned)ceil_div(s.Cp, 128), s.Mp), 128, 0, c.stream>>>(w.P.p, ks.keys_dev.p, s.C1, s.K, s.R, s.M, s.Cp, w.B.p);
	k_point_coords<<<(unsigned)ceil_div(s.g.N, 256), 256, 0, c.stream>>>(s.g.n[0], s.g.n[1], s.g.n[2], w.coords.p);
	GLI_HIP(hipGetLastError());
	return s;
}

void mode_product(const double* X, int64_t ldx, int64_t P, const double* B, int64_t ldb, int C, int Kd,
                  double* Y, int64_t ldy, int Cp, int tri_cols, int tri_rows, int64_t p_base, int tet_C1)
{
	if (P <= 0 || C <= 0) { return; }
	GLI_REQUIRE(ceil_div(P, GBM) * ceil_div(C, GBN) < 2147483647 - 8 * ceil_div(C, GBN), "mode product too large");
	TileOrder order{(int)ceil_div(P, GBM), (int)ceil_div(C, GBN), tile_order_grouped()};
	k_mode_product<<<(unsigned)order.blocks(), 256, 0, ctx().stream>>>(X, ldx, P, B, ldb, C, Kd, Y, ldy, Cp, tri_cols, tri_rows,
	                                                                   p_base, tet_C1, order);
	GLI_HIP(hipGetLastError());
}

// values of device-resident four-indices into device-resident values
static void eri_values_dev(const double* coeffs, int64_t K, const int64_t grid[6], const int64_t* d_idx, int64_t num,
                           double tol, int method, double* d_vals)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	GLI_REQUIRE(num >= 0, "negative number of indices");
	GLI_REQUIRE(method == GLI_ERI_TENSOR_FP64 || method == GLI_ERI_DIRECT_FP64 || method == GLI_ERI_TENSOR_MIXED,
	            "unknown ERI method");
	{
		const double N = (double)grid[1] * grid[3] * grid[5];
		GLI_REQUIRE(N * N * N * N < 9007199254740992.0, "grid too large: four-indices must stay below 2^53");
	}
	phase_begin();
	EriSetup s = eri_tables(coeffs, K, grid, tol);
	EriWorkspace& w = eri_ws();
	phase_mark("tables");
	if (num == 0) { return; }
	const int64_t Mp = s.Mp, Cp = s.Cp;

	if (method == GLI_ERI_DIRECT_FP64) {
		const int64_t blocks = std::min<int64_t>(num, 65536);
		const size_t shm = (3 * (size_t)s.Mp + 256) * sizeof(double);
		k_eri_direct<<<(unsigned)blocks, 256, shm, c.stream>>>(d_idx, num, s.g, s.ks->lut_dev.p, w.coords.p, w.B.p, s.Cp, s.M, s.Mp, s.T, d_vals);
		GLI_HIP(hipGetLastError());
		phase_mark("direct");
		return;
	}
	const bool mixed = (method == GLI_ERI_TENSOR_MIXED);
	GLI_REQUIRE(!mixed || s.Mp <= 256, "the split-precision method needs 6(n-1) + 2(K-1) < 256; use the FP64 method");

	// kernel tensor, cached per folded range
	if (w.T3_M != s.M) {
		w.T3.reserve((size_t)Mp * Mp * Mp);
		k_kernel_tensor<<<dim3((unsigned)ceil_div(Mp, 64), (unsigned)Mp, (unsigned)Mp), 64, 0, c.stream>>>(s.T, s.M, s.Mp, w.T3.p);
		GLI_HIP(hipGetLastError());
		w.T3_M = s.M;
		split_invalidate_cache();
	}
	phase_mark("kernel_tensor");
	if (mixed) {
		split_factor(w.B.p, Cp, s.Cp, s.Mp);
		phase_mark("split_factor");
	}
	// Y1[m2][m3][k1] = sum_m1 T3[m1][m2][m3] B[m1][k1]
	w.Y1.reserve((size_t)Mp * Mp * Cp);
	if (mixed) { mode_product_split(w.T3.p, Mp * Mp, Mp * Mp, s.Cp, s.Mp, w.Y1.p, Cp, s.Cp, 0, 0, 0, 0, true); }
	else { mode_product(w.T3.p, Mp * Mp, Mp * Mp, w.B.p, Cp, s.Cp, s.Mp, w.Y1.p, Cp, s.Cp, 0, 0, 0, 0); }
	phase_mark("mode1");
	// Y2[m3][k1][k2] = sum_m2 Y1[m2][m3][k1] B[m2][k2]   (k2 >= k1 only)
	w.Y2.reserve((size_t)Mp * Cp * Cp);
	if (mixed) { mode_product_split(w.Y1.p, Mp * Cp, Mp * Cp, s.Cp, s.Mp, w.Y2.p, Cp, s.Cp, 1, 0, 0, 0, false); }
	else { mode_product(w.Y1.p, Mp * Cp, Mp * Cp, w.B.p, Cp, s.Cp, s.Mp, w.Y2.p, Cp, s.Cp, 1, 0, 0, 0); }
	phase_mark("mode2");
	// V[k1 <= k2 <= k3] = sum_m3 Y2[m3][k1][k2] B[m3][k3], written in tetrahedral layout
	const int64_t nV = tet_index(0, 0, s.C1);
	w.V.reserve((size_t)nV);
	if (mixed) { mode_product_split(w.Y2.p, Cp * Cp, Cp * Cp, s.Cp, s.Mp, w.V.p, 0, s.Cp, 1, 1, 0, s.C1, false); }
	else { mode_product(w.Y2.p, Cp * Cp, Cp * Cp, w.B.p, Cp, s.Cp, s.Mp, w.V.p, 0, s.Cp, 1, 1, 0, s.C1); }
	phase_mark("mode3");
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
		const int ept = c.gather_ept;
		const unsigned blocks = (unsigned)ceil_div(num, 256 * ept);
		if (ept == 2) { k_gather<2><<<blocks, 256, 0, c.stream>>>(d_idx, num, s.g, dc, s.ks->lut16_dev.p, s.ks->tet3_dev.p, w.V.p, d_vals); }
		else if (ept == 8) { k_gather<8><<<blocks, 256, 0, c.stream>>>(d_idx, num, s.g, dc, s.ks->lut16_dev.p, s.ks->tet3_dev.p, w.V.p, d_vals); }
		else { k_gather<4><<<blocks, 256, 0, c.stream>>>(d_idx, num, s.g, dc, s.ks->lut16_dev.p, s.ks->tet3_dev.p, w.V.p, d_vals); }
	}
	GLI_HIP(hipGetLastError());
	phase_mark("gather");
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

int gli_hip_eri_values(const double* coeffs, int64_t K, const int64_t g[6], const int64_t* indices, int64_t num,
                       double tol, int method, double* values)
{
	try {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		if (num == 0) { return 0; }
		DevBuf<int64_t> di(num);
		DevBuf<double> dv(num);
		GLI_HIP(hipMemcpyAsync(di.p, indices, num * sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
		eri_values_dev(coeffs, K, g, di.p, num, tol, method, dv.p);
		GLI_HIP(hipMemcpyAsync(values, dv.p, num * sizeof(double), hipMemcpyDeviceToHost, c.stream));
		GLI_HIP(hipStreamSynchronize(c.stream));
		return 0;
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
		for (int b = 0; b < num_grids; ++b) {
			eri_values_dev(coeffs[b], K[b], grids + 6 * b, d_idx[b], num[b], tol, method, d_values[b]);
		}
		return 0;
	}
	catch (const Failure&) {
		return -1;
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
	EriWorkspace& w = eri_ws();
	w.coeffs.release(); w.f.release(); w.P.release(); w.B.release();
	w.Y1.release(); w.Y2.release(); w.V.release(); w.T3.release();
	w.T3_M = -1;
}

}  // extern "C"
