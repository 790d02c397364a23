// Symmetry-unique ERI and ERIDA index enumeration on the GPU, bit-exact with
// enumerate_symmetry_reduced_eri_indices (eri_integrals.c:382-526) and
// enumerate_symmetry_reduced_erida_indices (erida_integrals.c:300-410).
//
// The reference keeps a four-index t = ((i N + j) N + k) N + l of a grid
// centred at the origin iff its bounding box is centred (2 * centre in
// {-1, 0, 1} per axis) and, for every bias placement of an odd-span box and
// every O_h element, the label-sorted image index is not below t
// (eri_integrals.c:300-352, 463-487).  The identity image forces i <= j,
// k <= l, (i, j) <= (k, l), and the first image index of every element is at
// least i, so the canonical first point i is its own orbit minimum and every
// point p of the tuple has min_g lin(g p) >= i.
//
// Work decomposition: a unit is a pair (i, j) with i an orbit-minimal point;
// a row is a unit plus one k; one thread per row walks the l's that centre
// the box (per-axis bit masks, ascending l) and applies the full test.  Rows
// are numbered in lexicographic (i, j, k) order, so a count pass, an exclusive
// scan and a write pass emit the list already sorted.
#include "common.hpp"
#include "../gli_device.h"

#include <algorithm>
#include <cstring>

namespace gli {

// O_h as signed axis permutations: image coordinate r = sgn[g][r] * p[col[g][r]]
// (the element order of the exported matrices; symmetry.c:9-58)
#include "oh_tables.inc"

__constant__ signed char c_oh_col[48][3] = GLI_OH_COL;
__constant__ signed char c_oh_sgn[48][3] = GLI_OH_SGN;

struct EnumGrid {
	int n, h, N;
};

__device__ inline int lin_c(const EnumGrid& g, int x, int y, int z)
{
	return ((x + g.h) * g.n + (y + g.h)) * g.n + (z + g.h);
}

// Is the centred 4-tuple P (coordinates relative to the grid centre) the minimum of its orbit?
// i = t's first index (an orbit-minimal point), IJ = i*N + j, KL = k*N + l.
//
// An image's four-index starts with the smallest of its four image points.  For a
// placement whose points q all have orbit minimum >= i, every image point is >= i, so an
// image can only fall below t when some g maps a point q onto i itself; then
// minimg[q] == i and g belongs to canon[q], the elements taking q to its orbit minimum.
// A placement containing a point with orbit minimum < i always has an image below t.
// This is the same decision as testing all 48 elements (eri_integrals.c:300-352, 463-487).
__device__ bool eri_is_canonical(const EnumGrid& g, const int P[4][3], int i, int IJ, int KL,
                                 const int* __restrict__ minimg, const unsigned long long* __restrict__ canon)
{
	int c[3];
	for (int a = 0; a < 3; ++a) {
		const int mn = min(min(P[0][a], P[1][a]), min(P[2][a], P[3][a]));
		const int mx = max(max(P[0][a], P[1][a]), max(P[2][a], P[3][a]));
		c[a] = mn + mx;
	}
	for (int bx = 0; bx <= abs(c[0]); ++bx) {
		for (int by = 0; by <= abs(c[1]); ++by) {
			for (int bz = 0; bz <= abs(c[2]); ++bz) {
				int Q[4][3], lq[4];
				for (int p = 0; p < 4; ++p) {
					Q[p][0] = P[p][0] - bx * c[0];
					Q[p][1] = P[p][1] - by * c[1];
					Q[p][2] = P[p][2] - bz * c[2];
					lq[p] = lin_c(g, Q[p][0], Q[p][1], Q[p][2]);
					if (minimg[lq[p]] < i) { return false; }
				}
				for (int p = 0; p < 4; ++p) {
					if (minimg[lq[p]] != i) { continue; }
					for (unsigned long long m = canon[lq[p]]; m; m &= m - 1) {
						const int e = __builtin_ctzll(m);
						const int c0 = c_oh_col[e][0], c1 = c_oh_col[e][1], c2 = c_oh_col[e][2];
						const int s0 = c_oh_sgn[e][0], s1 = c_oh_sgn[e][1], s2 = c_oh_sgn[e][2];
						int L[4];
						for (int r = 0; r < 4; ++r) { L[r] = lin_c(g, s0 * Q[r][c0], s1 * Q[r][c1], s2 * Q[r][c2]); }
						const int a = min(L[0], L[1]), b = max(L[0], L[1]);
						const int cc = min(L[2], L[3]), d = max(L[2], L[3]);
						int ab = a * g.N + b, cd = cc * g.N + d;
						if (ab > cd) { const int t = ab; ab = cd; cd = t; }
						if (ab < IJ || (ab == IJ && cd < KL)) { return false; }
					}
				}
			}
		}
	}
	return true;
}

// bit mask over l_a in [-h, h] of the l coordinates that centre the box on one axis
__device__ inline unsigned centre_mask(const EnumGrid& g, int pi, int pj, int pk)
{
	const int lo = min(min(pi, pj), pk), hi = max(max(pi, pj), pk);
	unsigned m = 0;
	for (int l = -g.h; l <= g.h; ++l) {
		const int s = min(lo, l) + max(hi, l);
		if (s >= -1 && s <= 1) { m |= 1u << (l + g.h); }
	}
	return m;
}

// one row = unit (i, j) and k; kind 0 counts, kind 1 writes
template <int kWrite>
__global__ __launch_bounds__(256) void k_eri_rows(EnumGrid g, const int2* __restrict__ units, const int64_t* __restrict__ row_base,
	const int* __restrict__ klist, const int* __restrict__ kstart, const int* __restrict__ klen,
	const int* __restrict__ unit_icand, const int* __restrict__ minimg, const unsigned long long* __restrict__ canon,
	uint32_t* __restrict__ counts, const int64_t* __restrict__ offsets, int64_t* __restrict__ out)
{
	const int u = blockIdx.x;
	const int2 ij = units[u];
	const int ic = unit_icand[u];
	const int i = ij.x, j = ij.y;
	const int nk = klen[ic];
	const int* ks = klist + kstart[ic];
	int pi[3], pj[3];
	pi[0] = i / (g.n * g.n) - g.h; pi[1] = (i / g.n) % g.n - g.h; pi[2] = i % g.n - g.h;
	pj[0] = j / (g.n * g.n) - g.h; pj[1] = (j / g.n) % g.n - g.h; pj[2] = j % g.n - g.h;
	const int IJ = i * g.N + j;
	for (int kp = threadIdx.x; kp < nk; kp += blockDim.x) {
		const int k = ks[kp];
		const int64_t row = row_base[u] + kp;
		int pk[3];
		pk[0] = k / (g.n * g.n) - g.h; pk[1] = (k / g.n) % g.n - g.h; pk[2] = k % g.n - g.h;
		const unsigned mx = centre_mask(g, pi[0], pj[0], pk[0]);
		const unsigned my = centre_mask(g, pi[1], pj[1], pk[1]);
		const unsigned mz = centre_mask(g, pi[2], pj[2], pk[2]);
		const int lmin = (k == i) ? j : k;
		uint32_t cnt = 0;
		int64_t pos = kWrite ? offsets[row] : 0;
		if (mx && my && mz) {
			int P[4][3];
			for (int a = 0; a < 3; ++a) { P[0][a] = pi[a]; P[1][a] = pj[a]; P[2][a] = pk[a]; }
			for (unsigned bxm = mx; bxm; bxm &= bxm - 1) {
				const int lx = __builtin_ctz(bxm) - g.h;
				for (unsigned bym = my; bym; bym &= bym - 1) {
					const int ly = __builtin_ctz(bym) - g.h;
					for (unsigned bzm = mz; bzm; bzm &= bzm - 1) {
						const int lz = __builtin_ctz(bzm) - g.h;
						const int l = lin_c(g, lx, ly, lz);
						if (l < lmin || minimg[l] < i) { continue; }
						P[3][0] = lx; P[3][1] = ly; P[3][2] = lz;
						if (eri_is_canonical(g, P, i, IJ, k * g.N + l, minimg, canon)) {
							if (kWrite) {
								out[pos++] = (((int64_t)i * g.N + j) * g.N + k) * g.N + l;
							}
							else {
								cnt++;
							}
						}
					}
				}
			}
		}
		if (!kWrite) { counts[row] = cnt; }
	}
}

// ---------------------------------------------------------------------------
// exclusive scan uint32 -> int64 (block sums, recursive scan of the sums, add back)

constexpr int SCAN_T = 256, SCAN_PER = 8, SCAN_TILE = SCAN_T * SCAN_PER;

template <typename In>
__global__ __launch_bounds__(SCAN_T) void k_scan_tiles(const In* __restrict__ in, int64_t n, int64_t* __restrict__ out, int64_t* __restrict__ tile_sums)
{
	__shared__ int64_t s[SCAN_T];
	const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_PER;
	int64_t v[SCAN_PER];
	int64_t run = 0;
	for (int q = 0; q < SCAN_PER; ++q) {
		const int64_t x = (base + q < n) ? (int64_t)in[base + q] : 0;
		v[q] = run;
		run += x;
	}
	s[threadIdx.x] = run;
	__syncthreads();
	for (int off = 1; off < SCAN_T; off <<= 1) {
		int64_t t = (threadIdx.x >= (unsigned)off) ? s[threadIdx.x - off] : 0;
		__syncthreads();
		s[threadIdx.x] += t;
		__syncthreads();
	}
	const int64_t excl = s[threadIdx.x] - run;
	for (int q = 0; q < SCAN_PER; ++q) {
		if (base + q < n) { out[base + q] = v[q] + excl; }
	}
	if (threadIdx.x == SCAN_T - 1) { tile_sums[blockIdx.x] = s[SCAN_T - 1]; }
}

__global__ void k_scan_add(int64_t* __restrict__ out, int64_t n, const int64_t* __restrict__ tile_off)
{
	const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (e < n) { out[e] += tile_off[e / SCAN_TILE]; }
}

// exclusive scan of n counts into out; returns the total
template <typename In>
static int64_t exclusive_scan(const In* in, int64_t n, int64_t* out)
{
	hipStream_t st = ctx().stream;
	if (n == 0) { return 0; }
	const int64_t tiles = ceil_div(n, SCAN_TILE);
	DevBuf<int64_t> sums(tiles), sums_scan(tiles);
	k_scan_tiles<In><<<(unsigned)tiles, SCAN_T, 0, st>>>(in, n, out, sums.p);
	GLI_HIP(hipGetLastError());
	int64_t total = 0;
	if (tiles > 1) {
		total = exclusive_scan<int64_t>(sums.p, tiles, sums_scan.p);
		k_scan_add<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(out, n, sums_scan.p);
		GLI_HIP(hipGetLastError());
	}
	else {
		GLI_HIP(hipMemcpyAsync(&total, sums.p, sizeof(int64_t), hipMemcpyDeviceToHost, st));
		GLI_HIP(hipStreamSynchronize(st));
	}
	return total;
}

template int64_t exclusive_scan<uint32_t>(const uint32_t*, int64_t, int64_t*);

// ---------------------------------------------------------------------------
// host planning for a symmetric n^3 grid

struct EnumPlan {
	EnumGrid g;
	std::vector<int> minimg;       // per point: smallest linear index over its O_h orbit
	std::vector<unsigned long long> canon;  // per point: the elements g with lin(g p) == minimg
	std::vector<int2> units;       // (i, j)
	std::vector<int> unit_icand;   // index of i in the orbit-minimal list
	std::vector<int64_t> row_base; // first row of each unit
	std::vector<int> klist, kstart, klen;
	int64_t rows = 0;
};

static void plan_enumeration(int n, EnumPlan& p)
{
	static const signed char col[48][3] = GLI_OH_COL;
	static const signed char sgn[48][3] = GLI_OH_SGN;
	p.g.n = n;
	p.g.h = (n - 1) / 2;
	p.g.N = n * n * n;
	const int N = p.g.N, h = p.g.h;
	p.minimg.assign(N, 0);
	p.canon.assign(N, 0);
	for (int idx = 0; idx < N; ++idx) {
		const int q[3] = { idx / (n * n) - h, (idx / n) % n - h, idx % n - h };
		int img[48];
		int best = N;
		for (int e = 0; e < 48; ++e) {
			const int x = sgn[e][0] * q[col[e][0]], y = sgn[e][1] * q[col[e][1]], z = sgn[e][2] * q[col[e][2]];
			img[e] = ((x + h) * n + (y + h)) * n + (z + h);
			best = std::min(best, img[e]);
		}
		p.minimg[idx] = best;
		for (int e = 0; e < 48; ++e) {
			if (img[e] == best) { p.canon[idx] |= 1ull << e; }
		}
	}
	std::vector<int> icand;
	for (int idx = 0; idx < N; ++idx) {
		if (p.minimg[idx] == idx) { icand.push_back(idx); }
	}
	for (size_t c = 0; c < icand.size(); ++c) {
		const int i = icand[c];
		p.kstart.push_back((int)p.klist.size());
		int cnt = 0;
		for (int k = i; k < N; ++k) {
			if (p.minimg[k] >= i) { p.klist.push_back(k); cnt++; }
		}
		p.klen.push_back(cnt);
	}
	for (size_t c = 0; c < icand.size(); ++c) {
		const int i = icand[c];
		const int* ks = &p.klist[p.kstart[c]];
		for (int q = 0; q < p.klen[c]; ++q) {
			// j runs over the same admissible set as k (j >= i, orbit minimum >= i)
			p.units.push_back(int2{i, ks[q]});
			p.unit_icand.push_back((int)c);
			p.row_base.push_back(p.rows);
			p.rows += p.klen[c];
		}
	}
}

static void eri_enumerate_dev(int64_t n, int64_t** d_out, int64_t* num)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	GLI_REQUIRE(n >= 1 && (n % 2) == 1, "enumeration needs a symmetric grid with an odd number of points per axis");
	GLI_REQUIRE(n <= 31, "grid too large for the enumeration (at most 31 points per axis)");
	EnumPlan p;
	plan_enumeration((int)n, p);
	DevBuf<int2> units(p.units.size());
	DevBuf<int> unit_ic(p.unit_icand.size()), klist(p.klist.size()), kstart(p.kstart.size()), klen(p.klen.size()), minimg(p.minimg.size());
	DevBuf<int64_t> row_base(p.row_base.size());
	hipStream_t st = c.stream;
	GLI_HIP(hipMemcpyAsync(units.p, p.units.data(), p.units.size() * sizeof(int2), hipMemcpyHostToDevice, st));
	GLI_HIP(hipMemcpyAsync(unit_ic.p, p.unit_icand.data(), p.unit_icand.size() * sizeof(int), hipMemcpyHostToDevice, st));
	GLI_HIP(hipMemcpyAsync(klist.p, p.klist.data(), p.klist.size() * sizeof(int), hipMemcpyHostToDevice, st));
	GLI_HIP(hipMemcpyAsync(kstart.p, p.kstart.data(), p.kstart.size() * sizeof(int), hipMemcpyHostToDevice, st));
	GLI_HIP(hipMemcpyAsync(klen.p, p.klen.data(), p.klen.size() * sizeof(int), hipMemcpyHostToDevice, st));
	GLI_HIP(hipMemcpyAsync(minimg.p, p.minimg.data(), p.minimg.size() * sizeof(int), hipMemcpyHostToDevice, st));
	DevBuf<unsigned long long> canon(p.canon.size());
	GLI_HIP(hipMemcpyAsync(canon.p, p.canon.data(), p.canon.size() * sizeof(unsigned long long), hipMemcpyHostToDevice, st));
	GLI_HIP(hipMemcpyAsync(row_base.p, p.row_base.data(), p.row_base.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
	DevBuf<uint32_t> counts(p.rows);
	DevBuf<int64_t> offsets(p.rows);
	const unsigned nu = (unsigned)p.units.size();
	k_eri_rows<0><<<nu, 256, 0, st>>>(p.g, units.p, row_base.p, klist.p, kstart.p, klen.p, unit_ic.p, minimg.p, canon.p, counts.p, nullptr, nullptr);
	GLI_HIP(hipGetLastError());
	const int64_t total = exclusive_scan<uint32_t>(counts.p, p.rows, offsets.p);
	int64_t* out = nullptr;
	GLI_HIP(hipMalloc(&out, (size_t)std::max<int64_t>(total, 1) * sizeof(int64_t)));
	k_eri_rows<1><<<nu, 256, 0, st>>>(p.g, units.p, row_base.p, klist.p, kstart.p, klen.p, unit_ic.p, minimg.p, canon.p, nullptr, offsets.p, out);
	GLI_HIP(hipGetLastError());
	GLI_HIP(hipStreamSynchronize(st));
	*d_out = out;
	*num = total;
}

// ---------------------------------------------------------------------------
// ERIDA: pairs (i, j), i <= j, centred box, canonical over biases x O_h x (i <-> j)

__device__ bool erida_is_canonical(const EnumGrid& g, const int P[2][3], int t)
{
	int c[3];
	for (int a = 0; a < 3; ++a) { c[a] = P[0][a] + P[1][a]; }
	for (int bx = 0; bx <= abs(c[0]); ++bx) {
		for (int by = 0; by <= abs(c[1]); ++by) {
			for (int bz = 0; bz <= abs(c[2]); ++bz) {
				int Q[2][3];
				for (int p = 0; p < 2; ++p) {
					Q[p][0] = P[p][0] - bx * c[0];
					Q[p][1] = P[p][1] - by * c[1];
					Q[p][2] = P[p][2] - bz * c[2];
				}
				for (int e = 0; e < 48; ++e) {
					int L[2];
					for (int p = 0; p < 2; ++p) {
						L[p] = lin_c(g, c_oh_sgn[e][0] * Q[p][c_oh_col[e][0]], c_oh_sgn[e][1] * Q[p][c_oh_col[e][1]],
						             c_oh_sgn[e][2] * Q[p][c_oh_col[e][2]]);
					}
					if (min(L[0], L[1]) * g.N + max(L[0], L[1]) < t) { return false; }
				}
			}
		}
	}
	return true;
}

// one workgroup per first index i, lanes over j in rounds of 256 consecutive j; kind 0 counts, kind 1
// writes each round's kept pairs in ascending j (wave ballots + a 4-wave prefix in LDS)
template <int kWrite>
__global__ __launch_bounds__(256) void k_erida_rows(EnumGrid g, uint32_t* __restrict__ counts, const int64_t* __restrict__ offsets,
                                                     int64_t* __restrict__ out)
{
	__shared__ uint32_t wave_cnt[4];
	const int i = blockIdx.x;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	int P[2][3];
	P[0][0] = i / (g.n * g.n) - g.h; P[0][1] = (i / g.n) % g.n - g.h; P[0][2] = i % g.n - g.h;
	int64_t base = kWrite ? offsets[i] : 0;
	uint32_t total = 0;
	for (int j0 = i; j0 < g.N; j0 += 256) {
		const int j = j0 + (int)threadIdx.x;
		bool keep = false;
		if (j < g.N) {
			P[1][0] = j / (g.n * g.n) - g.h; P[1][1] = (j / g.n) % g.n - g.h; P[1][2] = j % g.n - g.h;
			bool centred = true;
			for (int a = 0; a < 3; ++a) { centred = centred && abs(P[0][a] + P[1][a]) <= 1; }
			keep = centred && erida_is_canonical(g, P, i * g.N + j);
		}
		const unsigned long long m = __ballot(keep);
		if (lane == 0) { wave_cnt[wave] = (uint32_t)__popcll(m); }
		__syncthreads();
		uint32_t before = 0, round = 0;
		for (int w = 0; w < 4; ++w) {
			before += (w < wave) ? wave_cnt[w] : 0u;
			round += wave_cnt[w];
		}
		if (kWrite && keep) {
			const uint32_t rank = before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
			out[base + total + rank] = (int64_t)i * g.N + j;
		}
		total += round;
		__syncthreads();
	}
	if (!kWrite && threadIdx.x == 0) { counts[i] = total; }
}

static void erida_enumerate_dev(int64_t n, int64_t** d_out, int64_t* num)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	GLI_REQUIRE(n >= 1 && (n % 2) == 1, "enumeration needs a symmetric grid with an odd number of points per axis");
	GLI_REQUIRE(n <= 31, "grid too large for the enumeration (at most 31 points per axis)");
	EnumGrid g;
	g.n = (int)n;
	g.h = (int)(n - 1) / 2;
	g.N = (int)(n * n * n);
	hipStream_t st = c.stream;
	DevBuf<uint32_t> counts(g.N);
	DevBuf<int64_t> offsets(g.N);
	k_erida_rows<0><<<(unsigned)g.N, 256, 0, st>>>(g, counts.p, nullptr, nullptr);
	GLI_HIP(hipGetLastError());
	const int64_t total = exclusive_scan<uint32_t>(counts.p, g.N, offsets.p);
	int64_t* out = nullptr;
	GLI_HIP(hipMalloc(&out, (size_t)std::max<int64_t>(total, 1) * sizeof(int64_t)));
	k_erida_rows<1><<<(unsigned)g.N, 256, 0, st>>>(g, nullptr, offsets.p, out);
	GLI_HIP(hipGetLastError());
	GLI_HIP(hipStreamSynchronize(st));
	*d_out = out;
	*num = total;
}

static int to_host(void (*dev)(int64_t, int64_t**, int64_t*), int64_t n, int64_t** out, int64_t* num)
{
	int64_t* d = nullptr;
	try {
		dev(n, &d, num);
		const size_t bytes = (size_t)*num * sizeof(int64_t);
		int64_t* h = (int64_t*)aligned_alloc(64, (bytes + 63) & ~(size_t)63);
		GLI_REQUIRE(h != nullptr || bytes == 0, "host allocation of the index list failed");
		if (bytes > 0) { GLI_HIP(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost)); }
		(void)hipFree(d);
		*out = h;
		return 0;
	}
	catch (const Failure&) {
		if (d != nullptr) { (void)hipFree(d); }
		return -1;
	}
}

// ---------------------------------------------------------------------------
// dense views of a sparse ERI set (fill_dense_eri_tensor, apply_sparse_eri_tensor and
// project_sparse_eri_gausslet_integrals, eri_integrals.c:714-1100): every tuple of a grid is mapped
// to the canonical four-index of its orbit, whose value is found by binary search in the sorted
// sparse list (0 when absent, as sparse_eri_gausslet_integrals_get_value).

// smallest image and the elements reaching it, per point of a symmetric n^3 grid
static void point_tables(int n, std::vector<int>& minimg, std::vector<unsigned long long>& canon)
{
	static const signed char col[48][3] = GLI_OH_COL;
	static const signed char sgn[48][3] = GLI_OH_SGN;
	const int h = (n - 1) / 2, N = n * n * n;
	minimg.assign(N, 0);
	canon.assign(N, 0);
	for (int idx = 0; idx < N; ++idx) {
		const int q[3] = { idx / (n * n) - h, (idx / n) % n - h, idx % n - h };
		int img[48];
		int best = N;
		for (int e = 0; e < 48; ++e) {
			const int x = sgn[e][0] * q[col[e][0]], y = sgn[e][1] * q[col[e][1]], z = sgn[e][2] * q[col[e][2]];
			img[e] = ((x + h) * n + (y + h)) * n + (z + h);
			best = std::min(best, img[e]);
		}
		minimg[idx] = best;
		for (int e = 0; e < 48; ++e) {
			if (img[e] == best) { canon[idx] |= 1ull << e; }
		}
	}
}

// The canonical four-index of a tuple with absolute coordinates A (relative to the centre of the
// symmetric sparse grid): the box is translated so that twice its centre lies in {-1, 0, 1} per
// axis (trans = centre / 2 truncated, bias = centre - 2 trans, eri_integrals.c:727-733), every bias
// placement is tried, and the minimum over O_h and the 8 label swaps is taken.  Only the images
// whose smallest point is the overall smallest orbit minimum m1 can reach the minimum; those are
// the elements in canon[q] of the points q with minimg[q] == m1.  -1 when a placed point leaves
// the sparse grid.
__device__ int64_t eri_canonical_index(const EnumGrid& g, const int A[4][3], const int* __restrict__ minimg,
                                       const unsigned long long* __restrict__ canon)
{
	int trans[3], bias[3];
	for (int a = 0; a < 3; ++a) {
		const int mn = min(min(A[0][a], A[1][a]), min(A[2][a], A[3][a]));
		const int mx = max(max(A[0][a], A[1][a]), max(A[2][a], A[3][a]));
		const int centre = mn + mx;
		trans[a] = centre / 2;
		bias[a] = centre - 2 * trans[a];
	}
	int m1 = 0x7fffffff;
	for (int bx = 0; bx <= abs(bias[0]); ++bx) {
		for (int by = 0; by <= abs(bias[1]); ++by) {
			for (int bz = 0; bz <= abs(bias[2]); ++bz) {
				const int b[3] = { bx, by, bz };
				for (int p = 0; p < 4; ++p) {
					int Q[3];
					for (int a = 0; a < 3; ++a) {
						Q[a] = A[p][a] - trans[a] - b[a] * bias[a];
						if (Q[a] < -g.h || Q[a] > g.h) { return -1; }
					}
					m1 = min(m1, minimg[lin_c(g, Q[0], Q[1], Q[2])]);
				}
			}
		}
	}
	const int64_t N = g.N, N2 = (int64_t)g.N * g.N;
	int64_t best = INT64_MAX;
	for (int bx = 0; bx <= abs(bias[0]); ++bx) {
		for (int by = 0; by <= abs(bias[1]); ++by) {
			for (int bz = 0; bz <= abs(bias[2]); ++bz) {
				const int b[3] = { bx, by, bz };
				int Q[4][3];
				for (int p = 0; p < 4; ++p) {
					for (int a = 0; a < 3; ++a) { Q[p][a] = A[p][a] - trans[a] - b[a] * bias[a]; }
				}
				for (int p = 0; p < 4; ++p) {
					const int lq = lin_c(g, Q[p][0], Q[p][1], Q[p][2]);
					if (minimg[lq] != m1) { continue; }
					for (unsigned long long m = canon[lq]; m; m &= m - 1) {
						const int e = __builtin_ctzll(m);
						const int c0 = c_oh_col[e][0], c1 = c_oh_col[e][1], c2 = c_oh_col[e][2];
						const int s0 = c_oh_sgn[e][0], s1 = c_oh_sgn[e][1], s2 = c_oh_sgn[e][2];
						int L[4];
						for (int r = 0; r < 4; ++r) { L[r] = lin_c(g, s0 * Q[r][c0], s1 * Q[r][c1], s2 * Q[r][c2]); }
						const int64_t x0 = min(L[0], L[1]), x1 = max(L[0], L[1]);
						const int64_t y0 = min(L[2], L[3]), y1 = max(L[2], L[3]);
						int64_t ab = x0 * N + x1, cd = y0 * N + y1;
						if (ab > cd) { const int64_t t = ab; ab = cd; cd = t; }
						best = min(best, ab * N2 + cd);
					}
				}
			}
		}
	}
	return best;
}

__device__ inline double sparse_lookup(const int64_t* __restrict__ keys, const double* __restrict__ vals, int64_t num, int64_t key)
{
	if (key < 0) { return 0.0; }
	int64_t lo = 0, hi = num;
	while (lo < hi) {
		const int64_t m = (lo + hi) >> 1;
		const int64_t v = keys[m];
		if (key < v) { hi = m; }
		else if (key > v) { lo = m + 1; }
		else { return vals[m]; }
	}
	return 0.0;
}

struct DenseGrid {
	int istart[3];
	int n[3];
	int64_t N;
};

__device__ inline void dense_point(const DenseGrid& d, int64_t q, int A[3])
{
	const int z = (int)(q % d.n[2]);
	const int64_t r = q / d.n[2];
	const int y = (int)(r % d.n[1]);
	const int x = (int)(r / d.n[1]);
	A[0] = d.istart[0] + x;
	A[1] = d.istart[1] + y;
	A[2] = d.istart[2] + z;
}

// out[q] = value of dense tuple t0 + q, for q < count (grid-stride loop)
__global__ __launch_bounds__(256) void k_eri_fill_dense(EnumGrid gs, DenseGrid gd, const int* __restrict__ minimg,
	const unsigned long long* __restrict__ canon, const int64_t* __restrict__ keys, const double* __restrict__ vals, int64_t num,
	int64_t t0, int64_t count, double* __restrict__ out)
{
	for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < count; q += (int64_t)gridDim.x * blockDim.x) {
		int64_t t = t0 + q;
		int A[4][3];
		for (int p = 3; p >= 0; --p) {
			dense_point(gd, t % gd.N, A[p]);
			t /= gd.N;
		}
		out[q] = sparse_lookup(keys, vals, num, eri_canonical_index(gs, A, minimg, canon));
	}
}

// ret[(i N + k) S + s0 + s] = sum_{j,l} (ij|kl) states[(j N + l) S + s0 + s] for s < sn <= APPLY_S:
// one workgroup per (i, k), threads over (j, l), tree reduction in LDS
constexpr int APPLY_S = 8;

__global__ __launch_bounds__(256) void k_eri_apply(EnumGrid g, const int* __restrict__ minimg, const unsigned long long* __restrict__ canon,
	const int64_t* __restrict__ keys, const double* __restrict__ vals, int64_t num, const double* __restrict__ states,
	int64_t S, int s0, int sn, int64_t ik0, double* __restrict__ ret)
{
	__shared__ double red[APPLY_S][256];
	const int64_t N = g.N, N2 = (int64_t)g.N * g.N;
	const int64_t ik = ik0 + blockIdx.x;
	const int i = (int)(ik / N), k = (int)(ik % N);
	int A[4][3];
	A[0][0] = i / (g.n * g.n) - g.h; A[0][1] = (i / g.n) % g.n - g.h; A[0][2] = i % g.n - g.h;
	A[2][0] = k / (g.n * g.n) - g.h; A[2][1] = (k / g.n) % g.n - g.h; A[2][2] = k % g.n - g.h;
	double acc[APPLY_S];
	for (int s = 0; s < APPLY_S; ++s) { acc[s] = 0; }
	for (int64_t jl = threadIdx.x; jl < N2; jl += blockDim.x) {
		const int j = (int)(jl / N), l = (int)(jl % N);
		A[1][0] = j / (g.n * g.n) - g.h; A[1][1] = (j / g.n) % g.n - g.h; A[1][2] = j % g.n - g.h;
		A[3][0] = l / (g.n * g.n) - g.h; A[3][1] = (l / g.n) % g.n - g.h; A[3][2] = l % g.n - g.h;
		const double v = sparse_lookup(keys, vals, num, eri_canonical_index(g, A, minimg, canon));
		const double* x = states + jl * S + s0;
		for (int s = 0; s < APPLY_S; ++s) {
			if (s < sn) { acc[s] += v * x[s]; }
		}
	}
	for (int s = 0; s < APPLY_S; ++s) { red[s][threadIdx.x] = acc[s]; }
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if ((int)threadIdx.x < w) {
			for (int s = 0; s < APPLY_S; ++s) { red[s][threadIdx.x] += red[s][threadIdx.x + w]; }
		}
		__syncthreads();
	}
	if ((int)threadIdx.x < sn) { ret[ik * S + s0 + threadIdx.x] = red[threadIdx.x][0]; }
}

struct SparseDev {
	EnumGrid g;
	DevBuf<int> minimg;
	DevBuf<unsigned long long> canon;
	DevBuf<int64_t> keys;
	DevBuf<double> vals;
};

static void sparse_to_device(const int64_t gs[6], const int64_t* keys, const double* vals, int64_t num, SparseDev& d)
{
	GLI_REQUIRE(gs[1] == gs[3] && gs[1] == gs[5] && (gs[1] % 2) == 1 && gs[0] == -(gs[1] - 1) / 2 && gs[2] == gs[0] && gs[4] == gs[0],
	            "dense views need the sparse set of a symmetric grid with an odd number of points per axis");
	GLI_REQUIRE(gs[1] <= 31, "grid too large for the dense views (at most 31 points per axis)");
	GLI_REQUIRE(num >= 0, "negative number of entries");
	const int n = (int)gs[1];
	d.g.n = n;
	d.g.h = (n - 1) / 2;
	d.g.N = n * n * n;
	std::vector<int> minimg;
	std::vector<unsigned long long> canon;
	point_tables(n, minimg, canon);
	hipStream_t st = ctx().stream;
	d.minimg.alloc(minimg.size());
	d.canon.alloc(canon.size());
	d.keys.alloc(std::max<int64_t>(num, 1));
	d.vals.alloc(std::max<int64_t>(num, 1));
	GLI_HIP(hipMemcpyAsync(d.minimg.p, minimg.data(), minimg.size() * sizeof(int), hipMemcpyHostToDevice, st));
	GLI_HIP(hipMemcpyAsync(d.canon.p, canon.data(), canon.size() * sizeof(unsigned long long), hipMemcpyHostToDevice, st));
	if (num > 0) {
		GLI_HIP(hipMemcpyAsync(d.keys.p, keys, num * sizeof(int64_t), hipMemcpyHostToDevice, st));
		GLI_HIP(hipMemcpyAsync(d.vals.p, vals, num * sizeof(double), hipMemcpyHostToDevice, st));
	}
	GLI_HIP(hipStreamSynchronize(st));
}

}  // namespace gli

using namespace gli;

extern "C" {

int gli_hip_eri_fill_dense(const int64_t gs[6], const int64_t* keys, const double* vals, int64_t num, const int64_t gd[6], double* out)
{
	try {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		SparseDev sd;
		sparse_to_device(gs, keys, vals, num, sd);
		DenseGrid dg;
		for (int a = 0; a < 3; ++a) {
			GLI_REQUIRE(gd[2 * a + 1] > 0 && gd[2 * a + 1] < 1024, "dense grid axes need 1 to 1023 points");
			dg.istart[a] = (int)gd[2 * a];
			dg.n[a] = (int)gd[2 * a + 1];
		}
		dg.N = (int64_t)dg.n[0] * dg.n[1] * dg.n[2];
		GLI_REQUIRE((double)dg.N * dg.N * dg.N * dg.N < 9.0e15, "dense tensor too large");
		const int64_t total = dg.N * dg.N * dg.N * dg.N;
		const int64_t chunk = std::min<int64_t>(total, (int64_t)1 << 26);
		DevBuf<double> buf(std::max<int64_t>(chunk, 1));
		for (int64_t t0 = 0; t0 < total; t0 += chunk) {
			const int64_t cnt = std::min(chunk, total - t0);
			const unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(cnt, 256), 65536);
			k_eri_fill_dense<<<blocks, 256, 0, c.stream>>>(sd.g, dg, sd.minimg.p, sd.canon.p, sd.keys.p, sd.vals.p, num, t0, cnt, buf.p);
			GLI_HIP(hipGetLastError());
			GLI_HIP(hipMemcpyAsync(out + t0, buf.p, cnt * sizeof(double), hipMemcpyDeviceToHost, c.stream));
			GLI_HIP(hipStreamSynchronize(c.stream));
		}
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_hip_eri_apply(const int64_t gs[6], const int64_t* keys, const double* vals, int64_t num, const double* states,
                      int64_t S, double* ret)
{
	try {
		Context& c = ctx();
		std::lock_guard<std::recursive_mutex> lk(c.mu);
		GLI_REQUIRE(S >= 0, "negative number of states");
		SparseDev sd;
		sparse_to_device(gs, keys, vals, num, sd);
		const int64_t N2 = (int64_t)sd.g.N * sd.g.N;
		if (S == 0) { return 0; }
		DevBuf<double> dst(N2 * S), dret(N2 * S);
		GLI_HIP(hipMemcpyAsync(dst.p, states, N2 * S * sizeof(double), hipMemcpyHostToDevice, c.stream));
		const int64_t rows_per_launch = 1 << 20;
		for (int64_t s0 = 0; s0 < S; s0 += APPLY_S) {
			const int sn = (int)std::min<int64_t>(APPLY_S, S - s0);
			for (int64_t ik0 = 0; ik0 < N2; ik0 += rows_per_launch) {
				const unsigned blocks = (unsigned)std::min(rows_per_launch, N2 - ik0);
				k_eri_apply<<<blocks, 256, 0, c.stream>>>(sd.g, sd.minimg.p, sd.canon.p, sd.keys.p, sd.vals.p, num, dst.p, S, (int)s0, sn,
				                                          ik0, dret.p);
				GLI_HIP(hipGetLastError());
			}
		}
		GLI_HIP(hipMemcpyAsync(ret, dret.p, N2 * S * sizeof(double), hipMemcpyDeviceToHost, c.stream));
		GLI_HIP(hipStreamSynchronize(c.stream));
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}


int gli_eri_enumerate_device(int64_t n, int64_t** d_out, int64_t* num)
{
	try {
		eri_enumerate_dev(n, d_out, num);
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_erida_enumerate_device(int64_t n, int64_t** d_out, int64_t* num)
{
	try {
		erida_enumerate_dev(n, d_out, num);
		return 0;
	}
	catch (const Failure&) {
		return -1;
	}
}

int gli_hip_eri_enumerate(int64_t n, int64_t** out, int64_t* num)
{
	return to_host(eri_enumerate_dev, n, out, num);
}

int gli_hip_erida_enumerate(int64_t n, int64_t** out, int64_t* num)
{
	return to_host(erida_enumerate_dev, n, out, num);
}

}  // extern "C"
