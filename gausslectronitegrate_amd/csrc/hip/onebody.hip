// Nuclear attraction and kinetic matrices, and ERIDA values.
//
// Nuclear (nuclear_integrals.c:133-266): the reference sums, per orbital pair,
// f_x f_y f_z * sum_n Z_n erf(|r - R_n| / w) / |r - R_n| at r = (p_i + p_j)/2 + m/6.
// Every such r lies on the lattice u/6 (u = 3(p_i + p_j) + m per axis), so the
// potential W[u] is evaluated once per lattice point and the matrix is the
// three mode products  V = W x1 Fx^T x2 Fy^T x3 Fz^T  with
// F_a[u][(i_a, j_a)] = f'_{i_a - j_a}[u - 3(p_i + p_j)_a]  (f' = f * [|f| > tol]).
// The lower triangle mirrors the upper one (nuclear_integrals.c:257-263).
//
// Kinetic (kinetic_integrals.c:14-121): K x I x I + I x K x I + I x I x K from
// the 1D table t(s) = 1/2 sum c_a c_b (3 sqrt(pi)/2)(1 - d^2/2) e^{-d^2/4}, d = 3s + a - b.
//
// ERIDA (erida_integrals.c:44-100, 429-542): per element, the folded columns
// b_a[mu] = F'[mu - 3 s_a] + F'[-mu - 3 s_a] of F[d] = (2 pi / 9) sum_{a-b=d} c_a c_b
// contracted with the ERIDA kernel table, one workgroup per element.
#include "common.hpp"
#include "../gli_device.h"

#include <algorithm>
#include <cmath>

namespace gli {

// Euclidean norm scaled by the largest component (util.h:69-117)
__device__ inline double norm3(double v0, double v1, double v2)
{
	const double a = fabs(v0), b = fabs(v1), c = fabs(v2);
	double big, s0, s1;
	if (b <= c) {
		if (a <= c) {
			if (c == 0) { return 0; }
			big = c; s0 = a; s1 = b;
		}
		else { big = a; s0 = b; s1 = c; }
	}
	else {
		if (a <= b) { big = b; s0 = a; s1 = c; }
		else { big = a; s0 = b; s1 = c; }
	}
	const double r0 = s0 / big, r1 = s1 / big;
	return big * sqrt(1 + r0 * r0 + r1 * r1);
}

// The nuclei live in device memory as (x, y, z, Z) per nucleus, any number of them
// (nuclear_integrals.c:133-143 copies an arbitrary num_nuclei); every lattice point sums them in
// the caller's order, as the reference's loop over nuclei does (nuclear_integrals.c:234-244).
// The nuclei are read in chunks of kNucleiChunk through LDS, so the wave's loads are uniform.
constexpr int kNucleiChunk = 64;

__global__ void k_nuclear_potential(const double4* __restrict__ nuc, int num, int ux0, int uy0, int uz0, int Ux, int Uy,
                                    int Uz, double* __restrict__ W)
{
	__shared__ double4 chunk[kNucleiChunk];
	const int iz = blockIdx.x * blockDim.x + threadIdx.x;
	const int iy = blockIdx.y, ix = blockIdx.z;
	const double rx = (ux0 + ix) / 6.0, ry = (uy0 + iy) / 6.0, rz = (uz0 + iz) / 6.0;
	const double w = 1. / 3;
	double val = 0;
	for (int k0 = 0; k0 < num; k0 += kNucleiChunk) {
		const int cnt = min(kNucleiChunk, num - k0);
		__syncthreads();   // the previous chunk is no longer read
		for (int q = threadIdx.x; q < cnt; q += blockDim.x) { chunk[q] = nuc[k0 + q]; }
		__syncthreads();
		for (int k = 0; k < cnt; ++k) {
			const double4 nk = chunk[k];
			const double d = norm3(rx - nk.x, ry - nk.y, rz - nk.z);
			const double g = (d > 0) ? erf(d / w) / d : 2 / (sqrt(M_PI) * w);
			val += nk.w * g;
		}
	}
	if (iz < Uz) { W[((int64_t)ix * Uy + iy) * Uz + iz] = val; }
}

// f'[s][m] for s = 0..S-1, zeroed where |f| <= tol (gausslet_factors.c:12-77 with the skips of
// nuclear_integrals.c:210-228 folded in); the K coefficients staged in LDS (dynamic shared memory)
__global__ void k_factors_tol(const double* __restrict__ c, int K, double tol, double* __restrict__ f)
{
	extern __shared__ double cs[];
	for (int q = threadIdx.x; q < K; q += blockDim.x) { cs[q] = c[q]; }
	__syncthreads();
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
		acc += prefac * cs[i] * cs[j] * exp(-x * x);
	}
	f[(int64_t)s * L + m] = (fabs(acc) > tol) ? acc : 0.0;
}

// F_a[u][q], q = i*n + j over axis offsets, u over the lattice of the axis
__global__ void k_pair_factors(const double* __restrict__ f, int K, int n, int g0, int u0, int U, double* __restrict__ F)
{
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	const int u = blockIdx.y;
	const int C = n * n;
	if (q >= C) { return; }
	const int i = q / n, j = q % n;
	const int m = (u0 + u) - 3 * ((g0 + i) + (g0 + j));
	double v = 0;
	if (m >= -(K - 1) && m <= K - 1) { v = f[(int64_t)abs(i - j) * (2 * K - 1) + m + (K - 1)]; }
	F[(int64_t)u * C + q] = v;
}

// ngi[I][J] from V[(qx*Cy + qy)*Cz + qz], mirrored from the upper triangle
__global__ void k_nuclear_scatter(const double* __restrict__ V, int nx, int ny, int nz, double* __restrict__ out)
{
	const int64_t N = (int64_t)nx * ny * nz;
	const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= N * N) { return; }
	int64_t I = e / N, J = e % N;
	if (I > J) { const int64_t t = I; I = J; J = t; }
	const int iz = (int)(I % nz), iy = (int)((I / nz) % ny), ix = (int)(I / ((int64_t)nz * ny));
	const int jz = (int)(J % nz), jy = (int)((J / nz) % ny), jx = (int)(J / ((int64_t)nz * ny));
	const int64_t qx = ix * nx + jx, qy = iy * ny + jy, qz = iz * nz + jz;
	out[e] = V[(qx * ((int64_t)ny * ny) + qy) * ((int64_t)nz * nz) + qz];
}

// work buffers of the one-body assemblies, grown on demand and kept for later calls (the context
// lock is held while they are used)
struct OneBodyWorkspace {
	DevBuf<double> dc, f, W, F0, F1, F2, Y1, Y2, V, M, t;
	DevBuf<double4> nuc;            // (x, y, z, Z) of every nucleus
	std::vector<double4> h_nuc;     // host copy of the same, uploaded on the library stream
};

static OneBodyWorkspace& onebody_ws()
{
	static OneBodyWorkspace* w = new OneBodyWorkspace();
	return *w;
}

static void nuclear_matrix(const double* coeffs, int64_t K, const int64_t g[6], const double* pos, const int* charges,
                           int num_nuclei, double tol, double* out)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	GLI_REQUIRE(K >= 3 && (K % 2) == 1, "the number of gausslet coefficients must be odd and at least 3");
	GLI_REQUIRE(K <= 4095, "at most 4095 gausslet coefficients (the factor kernel stages them in LDS)");
	GLI_REQUIRE(g[1] > 0 && g[3] > 0 && g[5] > 0, "grid axes must have a positive number of points");
	GLI_REQUIRE(num_nuclei >= 0, "the number of nuclei must not be negative");
	GLI_REQUIRE(num_nuclei == 0 || (pos != nullptr && charges != nullptr), "nuclei without positions or charges");
	hipStream_t st = c.stream;
	const int n[3] = { (int)g[1], (int)g[3], (int)g[5] };
	const int g0[3] = { (int)g[0], (int)g[2], (int)g[4] };
	const int R = std::max(n[0], std::max(n[1], n[2]));
	const int L = 2 * (int)K - 1;
	int u0[3], U[3], C[3];
	for (int a = 0; a < 3; ++a) {
		u0[a] = 6 * g0[a] - ((int)K - 1);
		U[a] = 6 * (n[a] - 1) + 2 * ((int)K - 1) + 1;
		C[a] = n[a] * n[a];
	}
	OneBodyWorkspace& w = onebody_ws();
	w.h_nuc.resize((size_t)std::max(num_nuclei, 1));
	for (int k = 0; k < num_nuclei; ++k) {
		w.h_nuc[(size_t)k] = make_double4(pos[3 * k], pos[3 * k + 1], pos[3 * k + 2], (double)charges[k]);
	}
	w.nuc.reserve(w.h_nuc.size());
	DevBuf<double>& dc = w.dc;
	DevBuf<double>& f = w.f;
	DevBuf<double>& W = w.W;
	DevBuf<double>& F0 = w.F0;
	DevBuf<double>& F1 = w.F1;
	DevBuf<double>& F2 = w.F2;
	DevBuf<double>& Y1 = w.Y1;
	DevBuf<double>& Y2 = w.Y2;
	DevBuf<double>& V = w.V;
	DevBuf<double>& M = w.M;
	const int64_t N = (int64_t)n[0] * n[1] * n[2];
	dc.reserve(K);
	f.reserve((size_t)R * L);
	W.reserve((size_t)U[0] * U[1] * U[2]);
	F0.reserve((size_t)U[0] * C[0]);
	F1.reserve((size_t)U[1] * C[1]);
	F2.reserve((size_t)U[2] * C[2]);
	Y1.reserve((size_t)U[1] * U[2] * C[0]);
	Y2.reserve((size_t)U[2] * C[0] * C[1]);
	V.reserve((size_t)C[0] * C[1] * C[2]);
	M.reserve((size_t)(N * N));
	GLI_HIP(hipMemcpyAsync(dc.p, coeffs, K * sizeof(double), hipMemcpyHostToDevice, st));
	upload_async(w.nuc.p, w.h_nuc.data(), (size_t)num_nuclei * sizeof(double4));
	k_factors_tol<<<dim3((unsigned)ceil_div(L, 128), R), 128, (size_t)K * sizeof(double), st>>>(dc.p, (int)K, tol, f.p);
	k_nuclear_potential<<<dim3((unsigned)ceil_div(U[2], 64), U[1], U[0]), 64, 0, st>>>(w.nuc.p, num_nuclei, u0[0], u0[1], u0[2],
	                                                                                    U[0], U[1], U[2], W.p);
	double* Fs[3] = { F0.p, F1.p, F2.p };
	for (int a = 0; a < 3; ++a) {
		k_pair_factors<<<dim3((unsigned)ceil_div(C[a], 64), U[a]), 64, 0, st>>>(f.p, (int)K, n[a], g0[a], u0[a], U[a], Fs[a]);
	}
	GLI_HIP(hipGetLastError());
	// W[ux][uy][uz] -> Y1[uy][uz][qx] -> Y2[uz][qx][qy] -> V[qx][qy][qz]
	mode_product(W.p, (int64_t)U[1] * U[2], (int64_t)U[1] * U[2], F0.p, C[0], C[0], U[0], Y1.p, C[0], 0, 0, 0, 0, 0);
	mode_product(Y1.p, (int64_t)U[2] * C[0], (int64_t)U[2] * C[0], F1.p, C[1], C[1], U[1], Y2.p, C[1], 0, 0, 0, 0, 0);
	mode_product(Y2.p, (int64_t)C[0] * C[1], (int64_t)C[0] * C[1], F2.p, C[2], C[2], U[2], V.p, C[2], 0, 0, 0, 0, 0);
	k_nuclear_scatter<<<(unsigned)ceil_div(N * N, 256), 256, 0, st>>>(V.p, n[0], n[1], n[2], M.p);
	GLI_HIP(hipGetLastError());
	copy_to_host(out, M.p, (size_t)(N * N) * sizeof(double));
}

// ---------------------------------------------------------------------------
// kinetic

__global__ void k_kinetic_1d(const double* __restrict__ c, int K, double* __restrict__ t)
{
	__shared__ double red[256];
	const int s = blockIdx.x;
	double acc = 0;
	for (int q = threadIdx.x; q < K * K; q += blockDim.x) {
		const int i = q / K, j = q % K;
		const int d = 3 * s + (i - j);
		const double d2 = (double)d * d;
		acc += 0.5 * c[i] * c[j] * (1.5 * sqrt(M_PI) * (1 - 0.5 * d2) * exp(-0.25 * d2));
	}
	red[threadIdx.x] = acc;
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if (threadIdx.x < (unsigned)w) { red[threadIdx.x] += red[threadIdx.x + w]; }
		__syncthreads();
	}
	if (threadIdx.x == 0) { t[s] = red[0]; }
}

__global__ void k_kinetic_fill(const double* __restrict__ t, int nx, int ny, int nz, double* __restrict__ out)
{
	const int64_t N = (int64_t)nx * ny * nz;
	const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= N * N) { return; }
	const int64_t I = e / N, J = e % N;
	const int iz = (int)(I % nz), iy = (int)((I / nz) % ny), ix = (int)(I / ((int64_t)nz * ny));
	const int jz = (int)(J % nz), jy = (int)((J / nz) % ny), jx = (int)(J / ((int64_t)nz * ny));
	const double ex = (ix == jx) ? 1 : 0, ey = (iy == jy) ? 1 : 0, ez = (iz == jz) ? 1 : 0;
	out[e] = t[abs(ix - jx)] * ey * ez + ex * t[abs(iy - jy)] * ez + ex * ey * t[abs(iz - jz)];
}

static void kinetic_matrix(const double* coeffs, int64_t K, const int64_t g[6], double* out)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	GLI_REQUIRE(K >= 1, "at least one gausslet coefficient is required");
	GLI_REQUIRE(g[1] > 0 && g[3] > 0 && g[5] > 0, "grid axes must have a positive number of points");
	hipStream_t st = c.stream;
	const int R = (int)std::max(g[1], std::max(g[3], g[5]));
	const int64_t N = g[1] * g[3] * g[5];
	OneBodyWorkspace& w = onebody_ws();
	DevBuf<double>& dc = w.dc;
	DevBuf<double>& t = w.t;
	DevBuf<double>& M = w.M;
	dc.reserve(K);
	t.reserve(R);
	M.reserve((size_t)(N * N));
	GLI_HIP(hipMemcpyAsync(dc.p, coeffs, K * sizeof(double), hipMemcpyHostToDevice, st));
	k_kinetic_1d<<<R, 256, 0, st>>>(dc.p, (int)K, t.p);
	k_kinetic_fill<<<(unsigned)ceil_div(N * N, 256), 256, 0, st>>>(t.p, (int)g[1], (int)g[3], (int)g[5], M.p);
	GLI_HIP(hipGetLastError());
	copy_to_host(out, M.p, (size_t)(N * N) * sizeof(double));
}

// ---------------------------------------------------------------------------
// ERIDA

// F[d + K - 1] = (2 pi / 9) sum_{i - j = d} c_i c_j, zeroed where |F| <= tol
__global__ void k_erida_factors(const double* __restrict__ c, int K, double tol, double* __restrict__ F)
{
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= 2 * K - 1) { return; }
	const int d = q - (K - 1);
	const double prefac = 0.69813170079773183077;
	double acc = 0;
	for (int i = max(0, d); i <= min(K - 1, K - 1 + d); ++i) { acc += prefac * c[i] * c[i - d]; }
	F[q] = (fabs(acc) > tol) ? acc : 0.0;
}

// Folded ERI-DA columns: per axis the reference sums F[d] T'[(3s + d)^2 + ...] over d
// (erida_integrals.c:482-535); grouping d by mu = |3s + d| gives B[mu][|s|] (F[-d] = F[d] for every
// coefficient set, so the column depends on |s| only).
__global__ void k_erida_fold(const double* __restrict__ F, int K, int R, int M, int Cp, double* __restrict__ B)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	const int mu = blockIdx.y;
	if (k >= Cp) { return; }
	double v = 0;
	if (k < R && mu <= M) {
		int d = mu - 3 * k;
		if (d >= -(K - 1) && d <= K - 1) { v += F[d + K - 1]; }
		if (mu > 0) {
			d = -mu - 3 * k;
			if (d >= -(K - 1) && d <= K - 1) { v += F[d + K - 1]; }
		}
	}
	B[(int64_t)mu * Cp + k] = v;
}

// value of pair t = i N + j: the key tensor at the sorted |displacements| (NaN outside the grid)
__global__ void k_erida_gather(const int64_t* __restrict__ idx, int64_t num, int nx, int ny, int nz,
                               const double* __restrict__ V, double* __restrict__ out)
{
	const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= num) { return; }
	const int64_t N = (int64_t)nx * ny * nz;
	const int64_t t = idx[e];
	if (t < 0 || t >= N * N) {
		out[e] = __builtin_nan("");
		return;
	}
	const int64_t i = t / N, j = t % N;
	int k[3];
	k[2] = abs((int)(i % nz) - (int)(j % nz));
	k[1] = abs((int)((i / nz) % ny) - (int)((j / nz) % ny));
	k[0] = abs((int)(i / ((int64_t)nz * ny)) - (int)(j / ((int64_t)nz * ny)));
	const int lo = min(k[0], k[1]), hi = max(k[0], k[1]);
	const int k3 = max(hi, k[2]), mid = min(hi, k[2]);
	out[e] = V[tet_index(min(lo, mid), max(lo, mid), k3)];
}

struct EridaWorkspace {
	DevBuf<double> coeffs, F, B, Y1, Y2, V, T3;
	int T3_M = -1;
};

// ERI-DA values as a key-tensor contraction (the ERI method with one key per axis: |s|):
// V[k1][k2][k3] = sum_mu T'[mu1^2 + mu2^2 + mu3^2] B[mu1][k1] B[mu2][k2] B[mu3][k3], then one read per
// pair; the summed terms are the reference's (erida_integrals.c:429-542) in another order
static void erida_values_dev(const double* coeffs, int64_t K, const int64_t g[6], const int64_t* d_idx, int64_t num, double tol, double* d_vals)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	GLI_REQUIRE(K >= 1 && (K % 2) == 1, "the number of gausslet coefficients must be odd");
	GLI_REQUIRE(g[1] > 0 && g[3] > 0 && g[5] > 0, "grid axes must have a positive number of points");
	GLI_REQUIRE(g[1] < 1024 && g[3] < 1024 && g[5] < 1024, "at most 1023 points per axis");
	GLI_REQUIRE(tol >= 0, "tolerance must be non-negative");
	static EridaWorkspace* ws = new EridaWorkspace();
	EridaWorkspace& w = *ws;
	hipStream_t st = c.stream;
	const int R = (int)std::max(g[1], std::max(g[3], g[5]));
	const int M = 3 * (R - 1) + ((int)K - 1);
	const int64_t Mp = round_up(M + 1, 16), Cp = round_up(R, 64);
	GLI_REQUIRE(Mp <= 2048, "grid too long for the ERI-DA key tensor");
	const double* T = erida_table(3 * (int64_t)M * M + 1);
	w.coeffs.reserve(K);
	w.F.reserve(2 * K - 1);
	w.B.reserve(Mp * Cp);
	upload_async(w.coeffs.p, coeffs, K * sizeof(double));
	k_erida_factors<<<(unsigned)ceil_div(2 * K - 1, 128), 128, 0, st>>>(w.coeffs.p, (int)K, tol, w.F.p);
	k_erida_fold<<<dim3((unsigned)ceil_div(Cp, 128), (unsigned)Mp), 128, 0, st>>>(w.F.p, (int)K, R, M, (int)Cp, w.B.p);
	GLI_HIP(hipGetLastError());
	if (w.T3_M != M) {
		w.T3.reserve(Mp * Mp * Mp);
		kernel_tensor(T, M, (int)Mp, w.T3.p);
		w.T3_M = M;
	}
	w.Y1.reserve(Mp * Mp * Cp);
	w.Y2.reserve(Mp * Cp * Cp);
	w.V.reserve(tet_index(0, 0, R));
	mode_product(w.T3.p, Mp * Mp, Mp * Mp, w.B.p, Cp, (int)Cp, (int)Mp, w.Y1.p, Cp, (int)Cp, 0, 0, 0, 0);
	mode_product(w.Y1.p, Mp * Cp, Mp * Cp, w.B.p, Cp, (int)Cp, (int)Mp, w.Y2.p, Cp, (int)Cp, 1, 0, 0, 0);
	mode_product(w.Y2.p, Cp * Cp, Cp * Cp, w.B.p, Cp, (int)Cp, (int)Mp, w.V.p, 0, (int)Cp, 1, 1, 0, R);
	if (num > 0) {
		k_erida_gather<<<(unsigned)ceil_div(num, 256), 256, 0, st>>>(d_idx, num, (int)g[1], (int)g[3], (int)g[5], w.V.p, d_vals);
	}
	GLI_HIP(hipGetLastError());
	GLI_HIP(hipStreamSynchronize(st));
}

}  // namespace gli

using namespace gli;

extern "C" {

int gli_hip_nuclear_matrix(const double* coeffs, int64_t K, const int64_t g[6], const double* pos, const int* charges,
                           int num_nuclei, double tol, double* out)
{
	try {
		nuclear_matrix(coeffs, K, g, pos, charges, num_nuclei, tol, out);
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_hip_kinetic_matrix(const double* coeffs, int64_t K, const int64_t g[6], double* out)
{
	try {
		kinetic_matrix(coeffs, K, g, out);
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_kinetic_matrix(const double* coeffs, int64_t K, const int64_t grid[6], double* out)
{
	try {
		GLI_REQUIRE(coeffs != nullptr && grid != nullptr && out != nullptr, "null argument");
		kinetic_matrix(coeffs, K, grid, out);
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_nuclear_matrix(const double* coeffs, int64_t K, const int64_t grid[6], const double* positions, const int* charges,
                       int num_nuclei, double tol, double* out)
{
	try {
		GLI_REQUIRE(coeffs != nullptr && grid != nullptr && out != nullptr, "null argument");
		GLI_REQUIRE(num_nuclei == 0 || (positions != nullptr && charges != nullptr), "null nuclei arrays");
		nuclear_matrix(coeffs, K, grid, positions, charges, num_nuclei, tol, out);
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_erida_values_device(const double* coeffs, int64_t K, const int64_t g[6], const int64_t* d_idx, int64_t num,
                            double tol, double* d_values)
{
	try {
		erida_values_dev(coeffs, K, g, d_idx, num, tol, d_values);
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_hip_erida_values(const double* coeffs, int64_t K, const int64_t g[6], const int64_t* indices, int64_t num,
                         double tol, double* values)
{
	try {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		DevBuf<int64_t> di(std::max<int64_t>(num, 1));
		DevBuf<double> dv(std::max<int64_t>(num, 1));
		if (num > 0) { GLI_HIP(hipMemcpyAsync(di.p, indices, num * sizeof(int64_t), hipMemcpyHostToDevice, c.stream)); }
		erida_values_dev(coeffs, K, g, di.p, num, tol, dv.p);
		if (num > 0) {
			GLI_HIP(hipMemcpyAsync(values, dv.p, num * sizeof(double), hipMemcpyDeviceToHost, c.stream));
			GLI_HIP(hipStreamSynchronize(c.stream));
		}
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_erida_integrals_points(const double* coeffs, int64_t K, const int64_t* points, int64_t num, double tol,
                               double* values)
{
	try {
		int64_t g[6];
		std::vector<int64_t> idx;
		points_to_indices(points, num, 2, g, idx);
		GLI_REQUIRE(num == 0 || values != nullptr, "no output array given");
		if (num == 0) { return 0; }
		return gli_hip_erida_values(coeffs, K, g, idx.data(), num, tol, values);
	}
	catch (const Failure&) {
		return -1;
	}
}

}  // extern "C"
