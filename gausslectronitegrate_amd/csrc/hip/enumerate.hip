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
// a row is a unit plus one k; a work item (one workgroup) is a block of
// consecutive rows of one unit.  The l's of a row that centre the box form a
// product of three intervals, so an item's candidates (row by row, ascending
// l) are numbered in closed form and the lanes of a wave test 64 consecutive
// candidates: every lane does the same kind of work, whatever the rows'
// candidate counts.  Each candidate is tested once; the results go to an
// acceptance bitmap, and an expand pass writes the accepted four-indices at
// offsets from a scan of the items' counts, already sorted.
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

// Is the centred 4-tuple (i, j, k, l) (coordinates relative to the grid centre) the minimum of its
// orbit?  i = t's first index (an orbit-minimal point), IJ = i*N + j, KL = k*N + l, lp = the
// linear indices (i, j, k, l) of the four points.  A bias placement shifts every point by the same
// vector, which stays inside the grid, so its linear indices are lp minus one offset.
//
// An image's four-index starts with the smallest of its four image points.  For a
// placement whose points q all have orbit minimum >= i, every image point is >= i, so an
// image can only fall below t when some g maps a point q onto i itself; then
// minimg[q] == i and g belongs to canon[q], the elements taking q to its orbit minimum.
// A placement containing a point with orbit minimum < i always has an image below t.
// This is the same decision as testing all 48 elements (eri_integrals.c:300-352, 463-487).
//
// The elements to test are the union of canon[q] over the placement's points in i's orbit;
// the identity image of the unshifted tuple is t itself and is skipped (element 0 is the
// identity).  W[g] holds the signed weights of the original axes in the image's linear
// index, lin(g q) = W[g] . q + H.
//
// mc[q] packs minimg[q] (bits 48-63) and canon[q] (bits 0-47), so a placement's four points
// cost four loads and no branch.  The placements are the subsets of the axes whose box sum c
// is odd (c = +-1).
//
// The unshifted placement needs no loads: i is orbit-minimal and j, k, l all have orbit minimum
// >= i (the unit, row and candidate filters), so its minimum test always passes, and its element
// set e0 is the union of canon over the points among (i, j, k, l) in i's orbit, identity removed,
// which the caller assembles from a per-row mask of (i, j, k) and l's own table entry.

// Per workgroup: i and j are fixed, so their images lin(g i) = W[g] . i + H and lin(g j) are
// tabulated once per element (LP), as are the per-axis extents of (i, j) for the box sums.
struct CanonUnit {
	int i, IJ, H;
	int mn01[3], mx01[3];   // per axis min and max of the coordinates of i and j
};

// Does an image of the placement (i, j, k, l) - s under one of the elements in emask fall below
// t = (IJ, KL)?  lin(g (q - s)) = lin(g q) - W[g] . s, with s the bias shift (kShift false: the
// unshifted placement); K and Lc are the coordinates of k and l.  An image point costs three
// 24-bit multiply-adds, i's and j's come from the table.
template <bool kShift>
__device__ inline bool image_below(const EnumGrid& g, const CanonUnit& u, const int K[3], const int Lc[3], const int sh[3],
                                   unsigned long long emask, int KL, const int4* __restrict__ W, const int2* __restrict__ LP)
{
	for (; emask; emask &= emask - 1) {
		const int e = __builtin_ctzll(emask);
		const int4 w = W[e];
		const int2 lij = LP[e];
		const int S = kShift ? __mul24(w.x, sh[0]) + __mul24(w.y, sh[1]) + __mul24(w.z, sh[2]) : 0;
		const int HS = u.H - S;
		const int L0 = lij.x - S, L1 = lij.y - S;
		const int L2 = __mul24(w.x, K[0]) + __mul24(w.y, K[1]) + __mul24(w.z, K[2]) + HS;
		const int L3 = __mul24(w.x, Lc[0]) + __mul24(w.y, Lc[1]) + __mul24(w.z, Lc[2]) + HS;
		const int a = min(L0, L1), bb = max(L0, L1);
		const int cc = min(L2, L3), d = max(L2, L3);
		int ab = __mul24(a, g.N) + bb, cd = __mul24(cc, g.N) + d;
		if (ab > cd) { const int t = ab; ab = cd; cd = t; }
		if (ab < u.IJ || (ab == u.IJ && cd < KL)) { return true; }
	}
	return false;
}

__device__ bool eri_is_canonical(const EnumGrid& g, const CanonUnit& u, const int K[3], const int Lc[3], const int lp[4], int KL,
                                 unsigned long long e0, const unsigned long long* __restrict__ mc, const int4* __restrict__ W,
                                 const int2* __restrict__ LP)
{
	const int i = u.i;
	if (image_below<false>(g, u, K, Lc, K, e0, KL, W, LP)) { return false; }
	int c[3];
	for (int a = 0; a < 3; ++a) { c[a] = min(u.mn01[a], min(K[a], Lc[a])) + max(u.mx01[a], max(K[a], Lc[a])); }
	const int n2 = g.n * g.n;
	const int dx = __mul24(c[0], n2), dy = __mul24(c[1], g.n), dz = c[2];
	const unsigned act = (c[0] != 0 ? 1u : 0u) | (c[1] != 0 ? 2u : 0u) | (c[2] != 0 ? 4u : 0u);
	// the shifted placements P - b*c, b a nonempty subset of the odd axes
	for (unsigned b = act; b != 0u; b = (b - 1u) & act) {
		// every point's linear index moves by the same offset
		const int off = ((b & 1u) ? dx : 0) + ((b & 2u) ? dy : 0) + ((b & 4u) ? dz : 0);
		unsigned long long m[4];
		for (int p = 0; p < 4; ++p) { m[p] = mc[lp[p] - off]; }
		const int mi0 = (int)(m[0] >> 48), mi1 = (int)(m[1] >> 48), mi2 = (int)(m[2] >> 48), mi3 = (int)(m[3] >> 48);
		const int mn = min(min(mi0, mi1), min(mi2, mi3));
		if (mn < i) { return false; }
		// a point of the placement reaches i (and canon of such a point is never empty)
		if (mn == i) {
			const unsigned long long emask = (((mi0 == i) ? m[0] : 0ull) | ((mi1 == i) ? m[1] : 0ull) | ((mi2 == i) ? m[2] : 0ull) |
			                                  ((mi3 == i) ? m[3] : 0ull)) & 0xffffffffffffull;
			const int sh[3] = { (b & 1u) ? c[0] : 0, (b & 2u) ? c[1] : 0, (b & 4u) ? c[2] : 0 };
			if (image_below<true>(g, u, K, Lc, sh, emask, KL, W, LP)) { return false; }
		}
	}
	return true;
}

// ---------------------------------------------------------------------------
// Candidates.  For a row (unit (i, j) plus k) the l coordinates that centre the box on axis a form
// one interval [A_a, A_a + L_a): with lo, hi the extent of (i, j, k) on that axis and s = lo + hi,
//   s >  1:  l in [-1 - hi, 1 - hi]              (l becomes the new minimum)
//   s < -1:  l in [-1 - lo, 1 - lo]              (l becomes the new maximum)
//   else:    l in [min(lo, -1 - hi), max(hi, 1 - lo)]
// clipped to [-h, h].  The row's candidates are the product of the three intervals in ascending l
// (lexicographic (lx, ly, lz)), minus the prefix below lmin (k, or j when k == i), which the
// identity image rules out.  Candidates are numbered row by row, so testing them in that order and
// keeping the accepted ones yields the sorted list.
//
// A work item is a block of ENUM_ROWS consecutive rows of one unit (units differ in size by a
// factor of 16 and more, so whole units would leave a long tail).  Each item's candidates occupy
// a whole number of 64-bit words of the acceptance bitmap, one bit per candidate.

constexpr int ENUM_T = 256, ENUM_ROWS = 64;

__device__ inline void axis_interval(int h, int pi, int pj, int pk, int& start, int& len)
{
	const int lo = min(min(pi, pj), pk), hi = max(max(pi, pj), pk);
	const int s = lo + hi;
	int a, b;
	if (s > 1) { a = -1 - hi; b = 1 - hi; }
	else if (s < -1) { a = -1 - lo; b = 1 - lo; }
	else { a = min(lo, -1 - hi); b = max(hi, 1 - lo); }
	a = max(a, -h);
	b = min(b, h);
	start = a;
	len = max(0, b - a + 1);
}

// v / d for v < 2^15 and 1 <= d <= 31 (a candidate's number inside its row, an interval length):
// the high word of (v 2^9) * m with m = ceil(2^23 / d), one 24-bit multiply.  Exact: the product
// over 2^32 is v/d + e with 0 <= e < 2^15 2^9 / 2^32 = 2^-8 <= 1/d, and frac(v/d) <= 1 - 1/d.
__host__ __device__ constexpr uint32_t enum_magic(uint32_t d)
{
	return ((1u << 23) + d - 1u) / d;
}

__device__ inline uint32_t udiv_magic(uint32_t v, uint32_t m)
{
	const uint32_t a = (v & 0x7fffu) << 9;
	return (uint32_t)(((uint64_t)a * (uint64_t)(m & 0xffffffu)) >> 32);
}

// x unchanged, with its value range hidden from the optimiser (an empty asm statement, no
// instruction).  A quotient known to be small turns __umul24 of it into a plain multiply, which
// is lowered to the quarter-rate v_mul_lo_u32; hidden, it stays a full-rate v_mul_u32_u24.
__device__ inline uint32_t range_opaque(uint32_t x)
{
	asm("" : "+v"(x));
	return x;
}

// row index of every 64th candidate, for items of up to 64 * ENUM_ROW_AT candidates (every item of a
// 13^3 grid: 64 rows of at most 2197 candidates)
constexpr int ENUM_ROW_AT = 4096;

// the rows of one work item in LDS
struct EnumRows {
	uint32_t pre[ENUM_ROWS + 1];   // exclusive prefix of the rows' candidate counts
	uint32_t iv[ENUM_ROWS];        // per axis 5 bits start + h and 5 bits length (x in bits 0-9, y 10-19, z 20-29)
	uint32_t ks[ENUM_ROWS];        // k (bits 0-15) and the candidates skipped below lmin (bits 16-31)
	uint32_t kc[ENUM_ROWS];        // k's coordinates + h, 5 bits per axis
	uint32_t mz[ENUM_ROWS], my[ENUM_ROWS];   // enum_magic of the z and y interval lengths
	uint32_t wsum[ENUM_T / 64];
	uint8_t row_at[ENUM_ROW_AT];   // row of candidate 64 w (when the item has at most 64 * ENUM_ROW_AT candidates)
};

struct EnumItem {
	int i, j, IJ, r0, rows;
	int pi[3], pj[3];
	const int* ks;
};

__device__ inline EnumItem enum_item(const EnumGrid& g, const int2* __restrict__ items, const int2* __restrict__ units,
                                     const int* __restrict__ unit_icand, const int* __restrict__ klist,
                                     const int* __restrict__ kstart, const int* __restrict__ klen)
{
	EnumItem it;
	const int2 ur = items[blockIdx.x];
	const int2 ij = units[ur.x];
	const int ic = unit_icand[ur.x];
	it.i = ij.x;
	it.j = ij.y;
	it.IJ = it.i * g.N + it.j;
	it.r0 = ur.y;
	it.rows = min(ENUM_ROWS, klen[ic] - ur.y);
	it.ks = klist + kstart[ic];
	// coordinates below n <= 31 (the masks tell the compiler so), minus h <= 15
	const int h = g.h & 15;
	it.pi[0] = ((it.i / (g.n * g.n)) & 31) - h; it.pi[1] = (((it.i / g.n) % g.n) & 31) - h; it.pi[2] = ((it.i % g.n) & 31) - h;
	it.pj[0] = ((it.j / (g.n * g.n)) & 31) - h; it.pj[1] = (((it.j / g.n) % g.n) & 31) - h; it.pj[2] = ((it.j % g.n) & 31) - h;
	return it;
}

// exclusive prefix of v over the workgroup; total receives the sum
__device__ inline uint32_t enum_block_scan(uint32_t v, uint32_t* wsum, uint32_t& total)
{
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	uint32_t x = v;
	for (int off = 1; off < 64; off <<= 1) {
		const uint32_t y = __shfl_up(x, off, 64);
		if (lane >= off) { x += y; }
	}
	if (lane == 63) { wsum[w] = x; }
	__syncthreads();
	uint32_t before = 0;
	total = 0;
	for (int q = 0; q < ENUM_T / 64; ++q) {
		const uint32_t t = wsum[q];
		before += (q < w) ? t : 0u;
		total += t;
	}
	__syncthreads();
	return before + x - v;
}

// the item's rows into LDS (thread r < rows takes row r); returns the item's candidate count
__device__ uint32_t enum_load_rows(const EnumGrid& g, const EnumItem& it, EnumRows& s)
{
	const int r = threadIdx.x;
	uint32_t c = 0;
	if (r < it.rows) {
		const int k = it.ks[it.r0 + r];
		const int pk[3] = { k / (g.n * g.n) - g.h, (k / g.n) % g.n - g.h, k % g.n - g.h };
		int st[3], len[3];
		for (int a = 0; a < 3; ++a) { axis_interval(g.h, it.pi[a], it.pj[a], pk[a], st[a], len[a]); }
		// candidates below lmin: lexicographically smaller (lx, ly, lz) than lmin's coordinates
		const int lmin = (k == it.i) ? it.j : k;
		const int m[3] = { lmin / (g.n * g.n) - g.h, (lmin / g.n) % g.n - g.h, lmin % g.n - g.h };
		int lt[3];
		bool in[3];
		for (int a = 0; a < 3; ++a) {
			lt[a] = min(max(m[a] - st[a], 0), len[a]);
			in[a] = m[a] >= st[a] && m[a] < st[a] + len[a];
		}
		const uint32_t total = (uint32_t)(len[0] * len[1] * len[2]);
		const uint32_t below = (uint32_t)(lt[0] * len[1] * len[2] + (in[0] ? lt[1] * len[2] + (in[1] ? lt[2] : 0) : 0));
		c = total - below;
		uint32_t iv = 0, kc = 0;
		for (int a = 0; a < 3; ++a) {
			iv |= (uint32_t)((st[a] + g.h) | (len[a] << 5)) << (10 * a);
			kc |= (uint32_t)(pk[a] + g.h) << (5 * a);
		}
		s.iv[r] = iv;
		s.ks[r] = (uint32_t)k | (below << 16);
		s.kc[r] = kc;
		s.mz[r] = (len[2] > 0) ? enum_magic((uint32_t)len[2]) : 0u;
		s.my[r] = (len[1] > 0) ? enum_magic((uint32_t)len[1]) : 0u;
	}
	uint32_t total;
	const uint32_t before = enum_block_scan(c, s.wsum, total);
	if (r < it.rows) {
		s.pre[r] = before;
		// the 64-candidate word starts inside this row
		if (total <= 64u * ENUM_ROW_AT) {
			for (uint32_t w = (before + 63) / 64; w * 64 < before + c; ++w) { s.row_at[w] = (uint8_t)r; }
		}
	}
	__syncthreads();
	return total;
}

// the row of candidate q < total of the loaded item: the last row r with pre[r] <= q (it has candidates)
__device__ inline int enum_row(const EnumRows& s, int rows, uint32_t q)
{
	int lo = 0, hi = rows;
	while (hi - lo > 1) {
		const int mid = (lo + hi) >> 1;
		if (s.pre[mid] <= q) { lo = mid; }
		else { hi = mid; }
	}
	return lo;
}

// the same from the word index: the row lies between the rows of the word's first candidate and of
// the next word's first one (usually equal or adjacent), for items of up to 64 * ENUM_ROW_AT candidates
__device__ inline int enum_row_indexed(const EnumRows& s, int rows, uint32_t total, uint32_t q)
{
	const uint32_t w = q >> 6;
	int lo = s.row_at[w];
	int hi = ((w + 1) * 64 < total) ? (int)s.row_at[w + 1] + 1 : rows;
	while (hi - lo > 1) {
		const int mid = (lo + hi) >> 1;
		if (s.pre[mid] <= q) { lo = mid; }
		else { hi = mid; }
	}
	return lo;
}

// candidate q of row r -> k, its coordinates K, the point L (relative to the centre) and l's linear
// index.  Coordinates are at most 31 + h in magnitude (h <= 15), and the masks below let the compiler
// see that they fit the 24-bit multiplies of the canonicality test.
__device__ inline void enum_decode(const EnumGrid& g, const EnumRows& s, int r, uint32_t q, int& k, int K[3], int L[3], int& l)
{
	const uint32_t ksw = s.ks[r], iv = s.iv[r], kc = s.kc[r];
	const int h = g.h & 15;
	k = (int)(ksw & 0xffffu);
	for (int a = 0; a < 3; ++a) { K[a] = (int)((kc >> (5 * a)) & 31u) - h; }
	const uint32_t v = q - s.pre[r] + (ksw >> 16);
	const uint32_t lz = (iv >> 25) & 31u, ly = (iv >> 15) & 31u;
	const uint32_t vz = range_opaque(udiv_magic(v, s.mz[r]));
	const uint32_t vy = range_opaque(udiv_magic(vz, s.my[r]));
	// l's coordinates + h: interval start + offset, each below n <= 31 (the mask outside the
	// barrier, so that it does not narrow the remainder's multiply)
	const uint32_t cz = range_opaque(((iv >> 20) & 31u) + (v - __umul24(vz, lz))) & 31u;
	const uint32_t cy = range_opaque(((iv >> 10) & 31u) + (vz - __umul24(vy, ly))) & 31u;
	const uint32_t cx = ((iv & 31u) + vy) & 31u;
	L[0] = (int)cx - h;
	L[1] = (int)cy - h;
	L[2] = (int)cz - h;
	const uint32_t n = (uint32_t)g.n & 31u;
	l = (int)(__umul24(__umul24(cx, n) + cy, n) + cz);
}

// pass 1: bitmap words of each item
__global__ __launch_bounds__(ENUM_T) void k_eri_item_words(EnumGrid g, const int2* __restrict__ items, const int2* __restrict__ units,
	const int* __restrict__ unit_icand, const int* __restrict__ klist, const int* __restrict__ kstart, const int* __restrict__ klen,
	int64_t* __restrict__ words)
{
	__shared__ EnumRows s;
	const EnumItem it = enum_item(g, items, units, unit_icand, klist, kstart, klen);
	const uint32_t tot = enum_load_rows(g, it, s);
	if (threadIdx.x == 0) { words[blockIdx.x] = (tot + 63) / 64; }
}

// point tables of grids with up to this many points are read from LDS (32 KB at most)
constexpr int ENUM_LDS_POINTS = 4096;

// pass 2: test every candidate once; one bit per candidate, accepted count per item.  With
// kLdsMc the workgroup first copies the point table mc (g.N entries, dynamic LDS) and reads it there.
template <bool kLdsMc>
__global__ __launch_bounds__(ENUM_T) void k_eri_test(EnumGrid g, const int2* __restrict__ items, const int2* __restrict__ units,
	const int* __restrict__ unit_icand, const int* __restrict__ klist, const int* __restrict__ kstart, const int* __restrict__ klen,
	const unsigned long long* __restrict__ mc, const int64_t* __restrict__ word_base, unsigned long long* __restrict__ bits,
	uint32_t* __restrict__ accepted)
{
	extern __shared__ unsigned long long smc[];
	__shared__ EnumRows s;
	__shared__ int4 W[48];
	__shared__ int2 LP[48];                            // per element: lin(g i), lin(g j)
	__shared__ unsigned long long rmask[ENUM_ROWS];   // per row: the unshifted placement's elements from (i, j, k)
	const EnumItem it = enum_item(g, items, units, unit_icand, klist, kstart, klen);
	CanonUnit u;
	u.i = it.i;
	u.IJ = it.IJ;
	u.H = g.h * (g.n * g.n + g.n + 1);
	for (int a = 0; a < 3; ++a) {
		u.mn01[a] = min(it.pi[a], it.pj[a]);
		u.mx01[a] = max(it.pi[a], it.pj[a]);
	}
	if (threadIdx.x < 48) {
		const int e = threadIdx.x;
		int w[3];
		w[c_oh_col[e][0]] = c_oh_sgn[e][0] * g.n * g.n;
		w[c_oh_col[e][1]] = c_oh_sgn[e][1] * g.n;
		w[c_oh_col[e][2]] = c_oh_sgn[e][2];
		W[e] = make_int4(w[0], w[1], w[2], 0);
		LP[e] = make_int2(w[0] * it.pi[0] + w[1] * it.pi[1] + w[2] * it.pi[2] + u.H,
		                  w[0] * it.pj[0] + w[1] * it.pj[1] + w[2] * it.pj[2] + u.H);
	}
	if (kLdsMc) {
		for (int q = threadIdx.x; q < g.N; q += ENUM_T) { smc[q] = mc[q]; }
	}
	const unsigned long long* __restrict__ tmc = kLdsMc ? smc : mc;
	const int lane = threadIdx.x & 63;
	const int64_t word = word_base[blockIdx.x];
	const uint32_t tot = enum_load_rows(g, it, s);   // its barriers also publish W, LP and the point table
	if ((int)threadIdx.x < it.rows) {
		const unsigned long long mi = tmc[it.i], mj = tmc[it.j], mk = tmc[s.ks[threadIdx.x] & 0xffffu];
		const unsigned long long e = mi | (((int)(mj >> 48) == it.i) ? mj : 0ull) | (((int)(mk >> 48) == it.i) ? mk : 0ull);
		rmask[threadIdx.x] = e & 0xfffffffffffeull;   // canon bits without the identity
	}
	__syncthreads();
	const bool indexed = tot <= 64u * ENUM_ROW_AT;
	uint32_t acc = 0;
	for (uint32_t q0 = 0; q0 < tot; q0 += ENUM_T) {
		const uint32_t q = q0 + threadIdx.x;
		bool keep = false;
		if (q < tot) {
			int k, K[3], L[3], l;
			const int r = indexed ? enum_row_indexed(s, it.rows, tot, q) : enum_row(s, it.rows, q);
			enum_decode(g, s, r, q, k, K, L, l);
			const unsigned long long ml = tmc[l];
			const int mil = (int)(ml >> 48);
			if (mil >= it.i) {
				const int lp[4] = { it.i, it.j, k, l };
				const unsigned long long e0 = rmask[r] | ((mil == it.i) ? (ml & 0xfffffffffffeull) : 0ull);
				keep = eri_is_canonical(g, u, K, L, lp, __mul24(k, g.N) + l, e0, tmc, W, LP);
			}
		}
		const unsigned long long m = __ballot(keep);
		const uint32_t wq = q0 + (threadIdx.x & ~63u);
		if (lane == 0 && wq < tot) {
			bits[word + wq / 64] = m;
			acc += (uint32_t)__popcll(m);
		}
	}
	if (lane == 0) { s.wsum[threadIdx.x >> 6] = acc; }
	__syncthreads();
	if (threadIdx.x == 0) {
		uint32_t t = 0;
		for (int w = 0; w < ENUM_T / 64; ++w) { t += s.wsum[w]; }
		accepted[blockIdx.x] = t;
	}
}

// pass 3: the accepted candidates of each item in candidate order, from its output offset on.  Each
// thread takes one bitmap word per round (coalesced loads, one block scan per 256 words) and writes
// the candidates of its set bits.
__global__ __launch_bounds__(ENUM_T) void k_eri_expand(EnumGrid g, const int2* __restrict__ items, const int2* __restrict__ units,
	const int* __restrict__ unit_icand, const int* __restrict__ klist, const int* __restrict__ kstart, const int* __restrict__ klen,
	const int64_t* __restrict__ word_base, const unsigned long long* __restrict__ bits, const int64_t* __restrict__ out_base,
	int64_t* __restrict__ out)
{
	__shared__ EnumRows s;
	const EnumItem it = enum_item(g, items, units, unit_icand, klist, kstart, klen);
	const int64_t word = word_base[blockIdx.x];
	int64_t pos = out_base[blockIdx.x];
	const uint32_t tot = enum_load_rows(g, it, s);
	const uint32_t nw = (tot + 63) / 64;
	const int64_t ij = (int64_t)it.IJ;
	for (uint32_t w0 = 0; w0 < nw; w0 += ENUM_T) {
		const uint32_t w = w0 + threadIdx.x;
		unsigned long long m = (w < nw) ? bits[word + w] : 0ull;
		uint32_t total;
		uint32_t rank = enum_block_scan((uint32_t)__popcll(m), s.wsum, total);
		// the bits of one word are 64 consecutive candidates: one search, then forward steps
		const uint32_t q1 = w * 64 + (uint32_t)__builtin_ctzll(m);
		int r = (m == 0ull) ? 0 : (tot <= 64u * ENUM_ROW_AT) ? enum_row_indexed(s, it.rows, tot, q1) : enum_row(s, it.rows, q1);
		for (; m; m &= m - 1) {
			const uint32_t q = w * 64 + (uint32_t)__builtin_ctzll(m);
			while (r + 1 < it.rows && s.pre[r + 1] <= q) { ++r; }
			int k, K[3], L[3];
			int l;
			enum_decode(g, s, r, q, k, K, L, l);
			out[pos + rank++] = (ij * g.N + k) * g.N + l;
		}
		pos += total;
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
	std::vector<int> klist, kstart, klen;
};

static void plan_enumeration(int n, EnumPlan& p)
{
	static const signed char col[48][3] = GLI_OH_COL;
	static const signed char sgn[48][3] = GLI_OH_SGN;
	GLI_REQUIRE(col[0][0] == 0 && col[0][1] == 1 && col[0][2] == 2 && sgn[0][0] == 1 && sgn[0][1] == 1 && sgn[0][2] == 1,
	            "O_h element 0 must be the identity");
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
		}
	}
}

// The device form of one grid's enumeration plan: unit and row tables, the point table, the work
// items, the items' bitmap word offsets and their per-call count and offset arrays.  It depends on
// the grid size only, so a repeated enumeration of one grid skips the host planning, the uploads
// and the word-count pass.  The last grid's plan is kept (context lock held) until the process ends
// or another grid is enumerated: at most a few tens of MB (13^3: 1.9M items).
struct EnumDevPlan {
	int64_t n = 0;   // 0: no plan
	EnumGrid g{};
	unsigned ni = 0;
	int64_t nwords = 0;
	DevBuf<int2> units, items;
	DevBuf<int> unit_ic, klist, kstart, klen;
	DevBuf<unsigned long long> mc;
	DevBuf<int64_t> word_base, out_base;
	DevBuf<uint32_t> accepted;
};

static EnumDevPlan& enum_dev_plan(int64_t n)
{
	static EnumDevPlan* cached = nullptr;
	if (cached == nullptr) { cached = new EnumDevPlan(); }
	EnumDevPlan& dp = *cached;
	if (dp.n == n) { return dp; }
	dp.n = 0;   // stays empty if the build below fails
	EnumPlan p;
	plan_enumeration((int)n, p);
	hipStream_t st = ctx().stream;
	dp.g = p.g;
	dp.units.alloc(p.units.size());
	dp.unit_ic.alloc(p.unit_icand.size());
	dp.klist.alloc(p.klist.size());
	dp.kstart.alloc(p.kstart.size());
	dp.klen.alloc(p.klen.size());
	GLI_HIP(hipMemcpyAsync(dp.units.p, p.units.data(), p.units.size() * sizeof(int2), hipMemcpyHostToDevice, st));
	GLI_HIP(hipMemcpyAsync(dp.unit_ic.p, p.unit_icand.data(), p.unit_icand.size() * sizeof(int), hipMemcpyHostToDevice, st));
	GLI_HIP(hipMemcpyAsync(dp.klist.p, p.klist.data(), p.klist.size() * sizeof(int), hipMemcpyHostToDevice, st));
	GLI_HIP(hipMemcpyAsync(dp.kstart.p, p.kstart.data(), p.kstart.size() * sizeof(int), hipMemcpyHostToDevice, st));
	GLI_HIP(hipMemcpyAsync(dp.klen.p, p.klen.data(), p.klen.size() * sizeof(int), hipMemcpyHostToDevice, st));
	// minimg (bits 48-63) and canon (bits 0-47) of every point in one table for the candidate test
	std::vector<unsigned long long> hmc(p.canon.size());
	for (size_t q = 0; q < hmc.size(); ++q) {
		GLI_REQUIRE((p.canon[q] >> 48) == 0 && p.minimg[q] >= 0 && p.minimg[q] < 65536, "point tables do not fit the packed form");
		hmc[q] = p.canon[q] | ((unsigned long long)p.minimg[q] << 48);
	}
	dp.mc.alloc(hmc.size());
	GLI_HIP(hipMemcpyAsync(dp.mc.p, hmc.data(), hmc.size() * sizeof(unsigned long long), hipMemcpyHostToDevice, st));
	// work items: blocks of ENUM_ROWS rows of each unit, in (unit, row) order
	std::vector<int2> hitems;
	for (size_t u = 0; u < p.units.size(); ++u) {
		const int nk = p.klen[p.unit_icand[u]];
		for (int r0 = 0; r0 < nk; r0 += ENUM_ROWS) { hitems.push_back(int2{(int)u, r0}); }
	}
	GLI_REQUIRE(hitems.size() < (size_t)INT32_MAX, "enumeration has too many work items");
	const unsigned ni = (unsigned)hitems.size();
	dp.items.alloc(std::max<size_t>(hitems.size(), 1));
	if (ni > 0) { GLI_HIP(hipMemcpyAsync(dp.items.p, hitems.data(), hitems.size() * sizeof(int2), hipMemcpyHostToDevice, st)); }
	dp.word_base.alloc(ni);
	dp.out_base.alloc(ni);
	dp.accepted.alloc(ni);
	dp.nwords = 0;
	if (ni > 0) {
		DevBuf<int64_t> words(ni);
		k_eri_item_words<<<ni, ENUM_T, 0, st>>>(dp.g, dp.items.p, dp.units.p, dp.unit_ic.p, dp.klist.p, dp.kstart.p, dp.klen.p,
		                                        words.p);
		GLI_HIP(hipGetLastError());
		dp.nwords = exclusive_scan<int64_t>(words.p, ni, dp.word_base.p);
	}
	GLI_HIP(hipStreamSynchronize(st));   // the host tables above are read by the copies until here
	dp.ni = ni;
	dp.n = n;
	return dp;
}

static void eri_enumerate_dev(int64_t n, int64_t** d_out, int64_t* num)
{
	Context& c = ctx();
	std::lock_guard<std::recursive_mutex> lk(c.mu);
	GLI_REQUIRE(n >= 1 && (n % 2) == 1, "enumeration needs a symmetric grid with an odd number of points per axis");
	GLI_REQUIRE(n <= 31, "grid too large for the enumeration (at most 31 points per axis)");
	EnumDevPlan& dp = enum_dev_plan(n);
	hipStream_t st = c.stream;
	const unsigned ni = dp.ni;
	int64_t total = 0;
	int64_t* out = nullptr;
	if (ni > 0) {
		// the bitmap is written whole by the test (every word of every item), so it needs no clearing
		DevBuf<unsigned long long> bits((size_t)std::max<int64_t>(dp.nwords, 1));
		if (dp.g.N <= ENUM_LDS_POINTS) {
			k_eri_test<true><<<ni, ENUM_T, (size_t)dp.g.N * sizeof(unsigned long long), st>>>(dp.g, dp.items.p, dp.units.p,
				dp.unit_ic.p, dp.klist.p, dp.kstart.p, dp.klen.p, dp.mc.p, dp.word_base.p, bits.p, dp.accepted.p);
		}
		else {
			k_eri_test<false><<<ni, ENUM_T, 0, st>>>(dp.g, dp.items.p, dp.units.p, dp.unit_ic.p, dp.klist.p, dp.kstart.p, dp.klen.p,
			                                         dp.mc.p, dp.word_base.p, bits.p, dp.accepted.p);
		}
		GLI_HIP(hipGetLastError());
		total = exclusive_scan<uint32_t>(dp.accepted.p, ni, dp.out_base.p);
		GLI_HIP(hipMalloc(&out, (size_t)std::max<int64_t>(total, 1) * sizeof(int64_t)));
		k_eri_expand<<<ni, ENUM_T, 0, st>>>(dp.g, dp.items.p, dp.units.p, dp.unit_ic.p, dp.klist.p, dp.kstart.p, dp.klen.p,
		                                    dp.word_base.p, bits.p, dp.out_base.p, out);
		GLI_HIP(hipGetLastError());
		GLI_HIP(hipStreamSynchronize(st));   // before the bitmap is released
	}
	else {
		GLI_HIP(hipMalloc(&out, sizeof(int64_t)));
	}
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
	int64_t* h = nullptr;
	try {
		dev(n, &d, num);
		const size_t bytes = (size_t)*num * sizeof(int64_t);
		h = (int64_t*)gli_host_alloc(bytes);
		GLI_REQUIRE(h != nullptr, "host allocation of the index list failed");
		if (bytes > 0) { download(h, d, bytes); }
		(void)hipFree(d);
		*out = h;
		return 0;
	}
	catch (const Failure&) {
		if (d != nullptr) { (void)hipFree(d); }
		free(h);
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
		copy_to_device(d.keys.p, keys, (size_t)num * sizeof(int64_t));
		copy_to_device(d.vals.p, vals, (size_t)num * sizeof(double));
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
		copy_to_device(dst.p, states, (size_t)(N2 * S) * sizeof(double));
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
		copy_to_host(ret, dret.p, (size_t)(N2 * S) * sizeof(double));
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
