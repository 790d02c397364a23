/*
 * Oracle: a CPU restatement of the reference integral engine, written for
 * checking only.  Nothing outside tests/, the smoke check and the CPU baseline
 * of the benchmark may load it; the product library never calls into it.
 *
 * Every routine follows the cited reference loop order so that its floating
 * point results agree with the reference build to the last bits where the
 * reference is deterministic (everything except the OpenMP-ordered linked list
 * in the enumeration, which is sorted afterwards anyway).
 *
 * Grids are passed as glong[6] = {istart_x, num_x, istart_y, num_y, istart_z, num_z}.
 * Coefficients are passed as (c, K) with Gaussian indices -(K-1)/2 ... (K-1)/2.
 *
 * Pins (see tests/test_oracle.py): the HDF5 fixtures of the reference tests,
 * the stored notebook values, the reference build under oracle/_ref, closed
 * forms (kinetic/ERI at coincident points, O_h group axioms) and brute-force
 * orbit canonicalisation on small grids.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t glong;

static inline glong lsq(glong x) { return x * x; }
static inline glong lmx(glong a, glong b) { return a >= b ? a : b; }
static inline glong lmn(glong a, glong b) { return a <= b ? a : b; }
static inline double dsq(double x) { return x * x; }

/* ------------------------------------------------------------------------- */
/* grid index maps (grid.h:45-82)                                             */

static inline glong g_num(const glong* g) { return g[1] * g[3] * g[5]; }

static inline glong g_lin(const glong* g, glong ix, glong iy, glong iz)
{
	return (ix * g[3] + iy) * g[5] + iz;
}

static inline void g_pt(const glong* g, glong idx, glong* p)
{
	p[2] = g[4] + idx % g[5];
	p[1] = g[2] + (idx / g[5]) % g[3];
	p[0] = g[0] + (idx / g[5]) / g[3];
}

/* ------------------------------------------------------------------------- */
/* octahedral group (symmetry.c:9-58, 120-150)                                */
/* Each element is a signed axis permutation: row r of the matrix has its
 * single non-zero entry sign[r] in column col[r].  The table is the O_h list
 * in the order of the reference's exported array.                           */

static const signed char oh_code[48][6] = {
	/* col0 col1 col2 sgn0 sgn1 sgn2 */
	{0,1,2, 1, 1, 1}, {1,0,2,-1,-1,-1}, {2,1,0, 1,-1, 1}, {2,0,1,-1, 1,-1},
	{1,2,0, 1,-1,-1}, {0,2,1,-1, 1, 1}, {0,1,2, 1, 1,-1}, {1,0,2,-1,-1, 1},
	{2,1,0, 1,-1,-1}, {2,0,1,-1, 1, 1}, {1,2,0, 1,-1, 1}, {0,2,1,-1, 1,-1},
	{2,1,0,-1,-1,-1}, {2,0,1, 1, 1, 1}, {0,1,2,-1, 1,-1}, {1,0,2, 1,-1, 1},
	{0,2,1, 1, 1,-1}, {1,2,0,-1,-1, 1}, {2,1,0,-1,-1, 1}, {2,0,1, 1, 1,-1},
	{0,1,2,-1, 1, 1}, {1,0,2, 1,-1,-1}, {0,2,1, 1, 1, 1}, {1,2,0,-1,-1,-1},
	{1,2,0, 1, 1,-1}, {0,2,1,-1,-1, 1}, {1,0,2,-1, 1,-1}, {0,1,2, 1,-1, 1},
	{2,0,1,-1,-1,-1}, {2,1,0, 1, 1, 1}, {1,2,0, 1, 1, 1}, {0,2,1,-1,-1,-1},
	{1,0,2,-1, 1, 1}, {0,1,2, 1,-1,-1}, {2,0,1,-1,-1, 1}, {2,1,0, 1, 1,-1},
	{0,2,1, 1,-1,-1}, {1,2,0,-1, 1, 1}, {2,0,1, 1,-1, 1}, {2,1,0,-1, 1,-1},
	{1,0,2, 1, 1, 1}, {0,1,2,-1,-1,-1}, {0,2,1, 1,-1, 1}, {1,2,0,-1, 1,-1},
	{2,0,1, 1,-1,-1}, {2,1,0,-1, 1, 1}, {1,0,2, 1, 1,-1}, {0,1,2,-1,-1, 1},
};

/* matrices as int64 3x3, row-major: out[48*9] */
void orc_octahedral_matrices(glong* out)
{
	for (int e = 0; e < 48; ++e) {
		for (int r = 0; r < 3; ++r) {
			for (int c = 0; c < 3; ++c) {
				out[e * 9 + r * 3 + c] = (oh_code[e][r] == c) ? oh_code[e][3 + r] : 0;
			}
		}
	}
}

/* perm[e*N + idx] = linear index of element e applied to point idx (symmetry.c:120-150) */
void orc_octahedral_perms(const glong* g, glong* perm)
{
	const glong N = g_num(g);
	for (int e = 0; e < 48; ++e) {
		for (glong idx = 0; idx < N; ++idx) {
			glong p[3], q[3];
			g_pt(g, idx, p);
			for (int r = 0; r < 3; ++r) {
				q[r] = oh_code[e][3 + r] * p[(int)oh_code[e][r]];
			}
			perm[e * N + idx] = g_lin(g, q[0] - g[0], q[1] - g[2], q[2] - g[4]);
		}
	}
}

/* ------------------------------------------------------------------------- */
/* gausslet factors (gausslet_factors.c:12-81)                                */
/* fac is ns x (2K-1) (zero padded); fstart/flen give the trimmed window.     */

void orc_factors(const double* c, glong K, glong s0, glong ns, double tol, double* fac, glong* fstart, glong* flen)
{
	const glong L = 2 * K - 1;
	const glong istart = -(K - 1) / 2;
	const double prefac = 0.590817950301838675766;  /* sqrt(pi) / 3 */
	double* all = malloc(L * sizeof(double));
	for (glong k = 0; k < ns; ++k) {
		const glong s = s0 + k;
		memset(all, 0, L * sizeof(double));
		for (glong i = 0; i < K; ++i) {
			for (glong j = 0; j < K; ++j) {
				all[i + j] += prefac * c[i] * c[j] * exp(-dsq(0.5 * (3 * s + (i - j))));
			}
		}
		glong lo = -1, hi = -1;
		for (glong m = 0; m < L; ++m) { if (fabs(all[m]) > tol) { lo = m; break; } }
		for (glong m = L - 1; m >= 0; --m) { if (fabs(all[m]) > tol) { hi = m; break; } }
		memset(&fac[k * L], 0, L * sizeof(double));
		if (lo >= 0) {
			fstart[k] = 2 * istart + lo;
			flen[k] = hi - lo + 1;
			memcpy(&fac[k * L], &all[lo], flen[k] * sizeof(double));
		}
		else {
			fstart[k] = 0;
			flen[k] = 0;
		}
	}
	free(all);
}

/* ------------------------------------------------------------------------- */
/* ERI pair products (eri_integrals.c:46-110)                                 */
/* For shift range [s0, s0+ns): prod is (ns*ns) x LP with LP = 4K-3,
 * pstart/plen the trimmed windows, pair index k*ns + l.                      */

void orc_eri_products(const double* c, glong K, glong s0, glong ns, double tol, double* prod, glong* pstart, glong* plen)
{
	const glong L = 2 * K - 1;
	const glong LP = 4 * K - 3;
	double* fac = malloc(ns * L * sizeof(double));
	glong* fst = malloc(ns * sizeof(glong));
	glong* fln = malloc(ns * sizeof(glong));
	orc_factors(c, K, s0, ns, 0., fac, fst, fln);
	double* all = malloc(LP * sizeof(double));
	for (glong k = 0; k < ns; ++k) {
		for (glong l = 0; l < ns; ++l) {
			const glong a0 = fst[k] - (fst[l] + fln[l] - 1);
			const glong an = fln[k] + fln[l] - 1;
			memset(all, 0, LP * sizeof(double));
			for (glong i = 0; i < fln[k]; ++i) {
				for (glong j = 0; j < fln[l]; ++j) {
					all[(fst[k] + i) - (fst[l] + j) - a0] += fac[k * L + i] * fac[l * L + j];
				}
			}
			glong lo = -1, hi = -1;
			for (glong m = 0; m < an; ++m) { if (fabs(all[m]) > tol) { lo = m; break; } }
			for (glong m = an - 1; m >= 0; --m) { if (fabs(all[m]) > tol) { hi = m; break; } }
			const glong kl = k * ns + l;
			memset(&prod[kl * LP], 0, LP * sizeof(double));
			if (lo >= 0) {
				pstart[kl] = a0 + lo;
				plen[kl] = hi - lo + 1;
				memcpy(&prod[kl * LP], &all[lo], plen[kl] * sizeof(double));
			}
			else {
				pstart[kl] = 0;
				plen[kl] = 0;
			}
		}
	}
	free(all); free(fln); free(fst); free(fac);
}

/* ------------------------------------------------------------------------- */
/* Coulomb kernel tables (eri_integrals.c:18-27, 569-574; erida_integrals.c:17-26, 445-451) */

static long double coulomb_ld(long double w, long double dist)
{
	if (dist > 0) {
		return erfl(dist / (sqrtl(2) * w)) / dist;
	}
	return 0.7978845608028653558798921198687637L / w;  /* sqrt(2/pi) / w */
}

/* ERI table: T[i] for squared distances i/36 (w = 1/3) */
void orc_eri_coulomb_table(glong n, double* T)
{
	for (glong i = 0; i < n; ++i) {
		T[i] = (double)coulomb_ld(1.L / 3, sqrtl(i) / 6);
	}
}

/* ERIDA table: T[i] for squared distances i/9 (w = sqrt(2)/3) */
void orc_erida_coulomb_table(glong n, double* T)
{
	const long double sqrt2_3 = 0.4714045207910316829338962414032327L;
	for (glong i = 0; i < n; ++i) {
		T[i] = (double)coulomb_ld(sqrt2_3, sqrtl(i) / 3);
	}
}

glong orc_eri_table_len(glong max_range, glong K)
{
	return 3 * lsq(6 * (max_range - 1) + 2 * (K - 1)) + 1;
}

glong orc_erida_table_len(glong max_range, glong K)
{
	return 3 * lsq(3 * (max_range - 1) + (K - 1)) + 1;
}

/* ------------------------------------------------------------------------- */
/* one ERI value from per-axis products (body of eri_integrals.c:608-669)     */

/* circular-products histogram bins[rho] = sum_{dx^2+dy^2=rho} Px Py (eri_integrals.c:624-646) */
static void eri_bins(const glong cd6[3], const double* pr[3], const glong st[3], const glong nm[3], double tol, double* bins)
{
	const glong nb = lmx(lsq(cd6[0] + st[0]), lsq(cd6[0] + st[0] + nm[0] - 1))
	               + lmx(lsq(cd6[1] + st[1]), lsq(cd6[1] + st[1] + nm[1] - 1)) + 1;
	memset(bins, 0, nb * sizeof(double));
	for (glong sx = 0; sx < nm[0]; ++sx) {
		if (fabs(pr[0][sx]) <= tol) { continue; }
		const glong qx = lsq(cd6[0] + (st[0] + sx));
		for (glong sy = 0; sy < nm[1]; ++sy) {
			if (fabs(pr[1][sy]) <= tol) { continue; }
			const glong qy = lsq(cd6[1] + (st[1] + sy));
			bins[qx + qy] += pr[0][sx] * pr[1][sy];
		}
	}
}

/* val = sum_rho bins[rho] sum_z Pz T[rho + dz^2] (eri_integrals.c:648-669) */
static double eri_contract(const glong cd6[3], const double* pr[3], const glong st[3], const glong nm[3], double tol, const double* bins, const double* T)
{
	const glong nb = lmx(lsq(cd6[0] + st[0]), lsq(cd6[0] + st[0] + nm[0] - 1))
	               + lmx(lsq(cd6[1] + st[1]), lsq(cd6[1] + st[1] + nm[1] - 1)) + 1;
	double val = 0;
	for (glong rho = 0; rho < nb; ++rho) {
		if (bins[rho] == 0) { continue; }
		double t = 0;
		for (glong sz = 0; sz < nm[2]; ++sz) {
			if (fabs(pr[2][sz]) <= tol) { continue; }
			t += pr[2][sz] * T[rho + lsq(cd6[2] + (st[2] + sz))];
		}
		t *= bins[rho];
		val += t;
	}
	return val;
}

/* ------------------------------------------------------------------------- */
/* sparse ERI values (eri_integrals.c:544-678)                                */

void orc_eri_sparse(const double* c, glong K, const glong* g, const glong* idx, glong num, double tol, double* out)
{
	const glong R = lmx(lmx(g[1], g[3]), g[5]);
	const glong s0 = -R + 1, ns = 2 * R - 1, LP = 4 * K - 3;
	double* prod = malloc(ns * ns * LP * sizeof(double));
	glong* pst = malloc(ns * ns * sizeof(glong));
	glong* pln = malloc(ns * ns * sizeof(glong));
	orc_eri_products(c, K, s0, ns, tol, prod, pst, pln);
	const glong nT = orc_eri_table_len(R, K);
	double* T = malloc(nT * sizeof(double));
	orc_eri_coulomb_table(nT, T);
	const glong N = g_num(g);
	const glong M = 6 * (R - 1) + 2 * (K - 1);
	const glong cap = 2 * lsq(M) + 2;

	#pragma omp parallel
	{
		double* bins = malloc(cap * sizeof(double));
		#pragma omp for schedule(static)
		for (glong n = 0; n < num; ++n) {
			glong r = idx[n];
			const glong il = r % N; r /= N;
			const glong ik = r % N; r /= N;
			const glong ij = r % N; r /= N;
			const glong ii = r;
			glong pi[3], pj[3], pk[3], pl[3];
			g_pt(g, ii, pi); g_pt(g, ij, pj); g_pt(g, ik, pk); g_pt(g, il, pl);
			glong cd6[3], st[3], nm[3];
			const double* pr[3];
			for (int a = 0; a < 3; ++a) {
				cd6[a] = 3 * ((pi[a] + pj[a]) - (pk[a] + pl[a]));
				const glong kl = ((pi[a] - pj[a]) - s0) * ns + ((pk[a] - pl[a]) - s0);
				pr[a] = &prod[kl * LP];
				st[a] = pst[kl];
				nm[a] = pln[kl];
			}
			eri_bins(cd6, pr, st, nm, tol, bins);
			out[n] = eri_contract(cd6, pr, st, nm, tol, bins, T);
		}
		free(bins);
	}
	free(T); free(pln); free(pst); free(prod);
}

/* flops the reference's loops execute per element (eri_integrals.c:624-669): 2 n_x n_y for the bins
 * (n_a = entries with |P| > tol), then per nonzero bin 2 n_z for the z sum and 2 for the scaling and
 * the accumulation: W = 2 n_x n_y + 2 R_xy (n_z + 1) */
void orc_eri_flops(const double* c, glong K, const glong* g, const glong* idx, glong num, double tol, double* out)
{
	const glong R = lmx(lmx(g[1], g[3]), g[5]);
	const glong s0 = -R + 1, ns = 2 * R - 1, LP = 4 * K - 3;
	double* prod = malloc(ns * ns * LP * sizeof(double));
	glong* pst = malloc(ns * ns * sizeof(glong));
	glong* pln = malloc(ns * ns * sizeof(glong));
	orc_eri_products(c, K, s0, ns, tol, prod, pst, pln);
	const glong N = g_num(g);
	const glong M = 6 * (R - 1) + 2 * (K - 1);
	const glong cap = 2 * lsq(M) + 2;
	#pragma omp parallel
	{
		double* bins = malloc(cap * sizeof(double));
		#pragma omp for schedule(static)
		for (glong n = 0; n < num; ++n) {
			glong r = idx[n];
			const glong il = r % N; r /= N;
			const glong ik = r % N; r /= N;
			const glong ij = r % N; r /= N;
			const glong ii = r;
			glong pi[3], pj[3], pk[3], pl[3];
			g_pt(g, ii, pi); g_pt(g, ij, pj); g_pt(g, ik, pk); g_pt(g, il, pl);
			glong cd6[3], st[3], nm[3], cnt[3];
			const double* pr[3];
			for (int a = 0; a < 3; ++a) {
				cd6[a] = 3 * ((pi[a] + pj[a]) - (pk[a] + pl[a]));
				const glong kl = ((pi[a] - pj[a]) - s0) * ns + ((pk[a] - pl[a]) - s0);
				pr[a] = &prod[kl * LP];
				st[a] = pst[kl];
				nm[a] = pln[kl];
				cnt[a] = 0;
				for (glong q = 0; q < nm[a]; ++q) { cnt[a] += fabs(pr[a][q]) > tol; }
			}
			eri_bins(cd6, pr, st, nm, tol, bins);
			const glong nb = lmx(lsq(cd6[0] + st[0]), lsq(cd6[0] + st[0] + nm[0] - 1))
			               + lmx(lsq(cd6[1] + st[1]), lsq(cd6[1] + st[1] + nm[1] - 1)) + 1;
			glong rxy = 0;
			for (glong q = 0; q < nb; ++q) { rxy += bins[q] != 0; }
			out[n] = 2.0 * (double)cnt[0] * (double)cnt[1] + 2.0 * (double)rxy * (double)(cnt[2] + 1);
		}
		free(bins);
	}
	free(pln); free(pst); free(prod);
}

/* one ERI for four points pts[4][3] (eri_integrals.c:133-232).
 * Deviation: the reference sizes its Coulomb table from the pair spreads |s01|, |s23| only
 * (eri_integrals.c:134-159) and reads past it when the two pair centres lie further apart; the
 * table here covers the box holding the four points, as the sparse path's does
 * (eri_integrals.c:551-574).  Wherever the reference stays inside its table the values agree. */
double orc_eri_single(const double* c, glong K, const glong* pts, double tol)
{
	glong r01 = 0, r23 = 0, span = 0;
	for (int a = 0; a < 3; ++a) {
		r01 = lmx(r01, labs(pts[0 * 3 + a] - pts[1 * 3 + a]));
		r23 = lmx(r23, labs(pts[2 * 3 + a] - pts[3 * 3 + a]));
		const glong mn = lmn(lmn(pts[a], pts[3 + a]), lmn(pts[6 + a], pts[9 + a]));
		const glong mx = lmx(lmx(pts[a], pts[3 + a]), lmx(pts[6 + a], pts[9 + a]));
		span = lmx(span, mx - mn);
	}
	const glong R = lmx(r01 + 1, r23 + 1);
	const glong s0 = -R + 1, ns = 2 * R - 1, LP = 4 * K - 3;
	double* prod = malloc(ns * ns * LP * sizeof(double));
	glong* pst = malloc(ns * ns * sizeof(glong));
	glong* pln = malloc(ns * ns * sizeof(glong));
	orc_eri_products(c, K, s0, ns, tol, prod, pst, pln);
	const glong nT = orc_eri_table_len(lmx(R, span + 1), K);
	double* T = malloc(nT * sizeof(double));
	orc_eri_coulomb_table(nT, T);
	glong cd6[3], st[3], nm[3];
	const double* pr[3];
	for (int a = 0; a < 3; ++a) {
		cd6[a] = 3 * ((pts[0 * 3 + a] + pts[1 * 3 + a]) - (pts[2 * 3 + a] + pts[3 * 3 + a]));
		const glong kl = ((pts[0 * 3 + a] - pts[1 * 3 + a]) - s0) * ns + ((pts[2 * 3 + a] - pts[3 * 3 + a]) - s0);
		pr[a] = &prod[kl * LP];
		st[a] = pst[kl];
		nm[a] = pln[kl];
	}
	const glong nb = lmx(lsq(cd6[0] + st[0]), lsq(cd6[0] + st[0] + nm[0] - 1))
	               + lmx(lsq(cd6[1] + st[1]), lsq(cd6[1] + st[1] + nm[1] - 1)) + 1;
	double* bins = malloc(nb * sizeof(double));
	eri_bins(cd6, pr, st, nm, tol, bins);
	const double v = eri_contract(cd6, pr, st, nm, tol, bins, T);
	free(bins); free(T); free(pln); free(pst); free(prod);
	return v;
}

/* ------------------------------------------------------------------------- */
/* canonical four-index (eri_integrals.c:240-352)                             */

glong orc_min_orbit_eri(glong N, const glong* perm, glong i, glong j, glong k, glong l)
{
	glong best = INT64_MAX;
	for (int e = 0; e < 48; ++e) {
		const glong* p = &perm[e * N];
		glong a = p[i], b = p[j], cc = p[k], d = p[l];
		if (a > b) { glong t = a; a = b; b = t; }
		if (cc > d) { glong t = cc; cc = d; d = t; }
		glong ab = a * N + b, cdv = cc * N + d;
		if (ab > cdv) { glong t = ab; ab = cdv; cdv = t; }
		best = lmn(best, ab * (N * N) + cdv);
	}
	return best;
}

static int orbit_not_below(glong N, const glong* perm, glong i, glong j, glong k, glong l, glong index)
{
	for (int e = 0; e < 48; ++e) {
		const glong* p = &perm[e * N];
		glong a = p[i], b = p[j], cc = p[k], d = p[l];
		if (a > b) { glong t = a; a = b; b = t; }
		if (cc > d) { glong t = cc; cc = d; d = t; }
		glong ab = a * N + b, cdv = cc * N + d;
		if (ab > cdv) { glong t = ab; ab = cdv; cdv = t; }
		if (ab * (N * N) + cdv < index) { return 0; }
	}
	return 1;
}

static int cmp_glong(const void* a, const void* b)
{
	const glong x = *(const glong*)a, y = *(const glong*)b;
	return (x > y) - (x < y);
}

struct glist { glong* v; glong n, cap; };

static void glist_push(struct glist* L, glong x)
{
	if (L->n == L->cap) {
		L->cap = L->cap ? 2 * L->cap : 1024;
		L->v = realloc(L->v, L->cap * sizeof(glong));
	}
	L->v[L->n++] = x;
}

/* Enumerate the symmetry-unique four-indices (eri_integrals.c:382-526).
 * With out == NULL only the count is returned.  The caller sizes 'out'.     */
glong orc_eri_enumerate(const glong* g, glong* out)
{
	const glong N = g_num(g);
	glong* perm = malloc(48 * N * sizeof(glong));
	orc_octahedral_perms(g, perm);
	const glong nx = g[1], ny = g[3], nz = g[5];
	struct glist total = {0};

	#pragma omp parallel
	{
		struct glist mine = {0};
		#pragma omp for schedule(dynamic) collapse(4)
		for (glong icx = 0; icx < nx; ++icx)
		for (glong jcx = 0; jcx < nx; ++jcx)
		for (glong kcx = 0; kcx < nx; ++kcx)
		for (glong lcx = 0; lcx < nx; ++lcx) {
			if (icx > jcx || kcx > lcx || icx > kcx) { continue; }
			const glong cx = 2 * g[0] + lmn(lmn(lmn(icx, jcx), kcx), lcx) + lmx(lmx(lmx(icx, jcx), kcx), lcx);
			if (cx < -1 || cx > 1) { continue; }
			for (glong icy = 0; icy < ny; ++icy)
			for (glong jcy = 0; jcy < ny; ++jcy)
			for (glong kcy = 0; kcy < ny; ++kcy)
			for (glong lcy = 0; lcy < ny; ++lcy) {
				const glong cy = 2 * g[2] + lmn(lmn(lmn(icy, jcy), kcy), lcy) + lmx(lmx(lmx(icy, jcy), kcy), lcy);
				if (cy < -1 || cy > 1) { continue; }
				for (glong icz = 0; icz < nz; ++icz)
				for (glong jcz = 0; jcz < nz; ++jcz)
				for (glong kcz = 0; kcz < nz; ++kcz)
				for (glong lcz = 0; lcz < nz; ++lcz) {
					const glong cz = 2 * g[4] + lmn(lmn(lmn(icz, jcz), kcz), lcz) + lmx(lmx(lmx(icz, jcz), kcz), lcz);
					if (cz < -1 || cz > 1) { continue; }
					const glong ii = g_lin(g, icx, icy, icz), ij = g_lin(g, jcx, jcy, jcz);
					const glong ik = g_lin(g, kcx, kcy, kcz), il = g_lin(g, lcx, lcy, lcz);
					const glong t = ((ii * N + ij) * N + ik) * N + il;
					int ok = 1;
					for (glong bx = 0; bx <= labs(cx) && ok; ++bx)
					for (glong by = 0; by <= labs(cy) && ok; ++by)
					for (glong bz = 0; bz <= labs(cz) && ok; ++bz) {
						ok = orbit_not_below(N, perm,
							g_lin(g, icx - bx * cx, icy - by * cy, icz - bz * cz),
							g_lin(g, jcx - bx * cx, jcy - by * cy, jcz - bz * cz),
							g_lin(g, kcx - bx * cx, kcy - by * cy, kcz - bz * cz),
							g_lin(g, lcx - bx * cx, lcy - by * cy, lcz - bz * cz), t);
					}
					if (ok) { glist_push(&mine, t); }
				}
			}
		}
		#pragma omp critical
		{
			for (glong q = 0; q < mine.n; ++q) { glist_push(&total, mine.v[q]); }
		}
		free(mine.v);
	}
	free(perm);
	qsort(total.v, total.n, sizeof(glong), cmp_glong);
	const glong n = total.n;
	if (out != NULL) { memcpy(out, total.v, n * sizeof(glong)); }
	free(total.v);
	return n;
}

/* ------------------------------------------------------------------------- */
/* ERIDA (erida_integrals.c:44-100, 119-201, 209-270, 300-410, 429-542)      */

/* F[d] for d in [-(K-1), K-1], trimmed; returns window via fst/fln; F sized 2K-1 */
void orc_erida_products(const double* c, glong K, double tol, double* F, glong* fst, glong* fln)
{
	const glong L = 2 * K - 1;
	const double prefac = 0.69813170079773183077;  /* 2 pi / 9 */
	double* all = calloc(L, sizeof(double));
	for (glong i = 0; i < K; ++i) {
		for (glong j = 0; j < K; ++j) {
			all[i - j + (K - 1)] += prefac * c[i] * c[j];
		}
	}
	glong lo = -1, hi = -1;
	for (glong m = 0; m < L; ++m) { if (fabs(all[m]) > tol) { lo = m; break; } }
	for (glong m = L - 1; m >= 0; --m) { if (fabs(all[m]) > tol) { hi = m; break; } }
	memset(F, 0, L * sizeof(double));
	if (lo >= 0) {
		*fst = -(K - 1) + lo;
		*fln = hi - lo + 1;
		memcpy(F, &all[lo], *fln * sizeof(double));
	}
	else {
		*fst = 0;
		*fln = 0;
	}
	free(all);
}

static double erida_value(const glong cd3[3], const double* F, glong fst, glong fln, double tol, const double* T, double* bins)
{
	const glong nb = lmx(lsq(cd3[0] + fst), lsq(cd3[0] + fst + fln - 1))
	               + lmx(lsq(cd3[1] + fst), lsq(cd3[1] + fst + fln - 1)) + 1;
	memset(bins, 0, nb * sizeof(double));
	for (glong sx = 0; sx < fln; ++sx) {
		if (fabs(F[sx]) <= tol) { continue; }
		const glong qx = lsq(cd3[0] + (fst + sx));
		for (glong sy = 0; sy < fln; ++sy) {
			if (fabs(F[sy]) <= tol) { continue; }
			bins[qx + lsq(cd3[1] + (fst + sy))] += F[sx] * F[sy];
		}
	}
	double val = 0;
	for (glong rho = 0; rho < nb; ++rho) {
		if (bins[rho] == 0) { continue; }
		double t = 0;
		for (glong sz = 0; sz < fln; ++sz) {
			if (fabs(F[sz]) <= tol) { continue; }
			t += F[sz] * T[rho + lsq(cd3[2] + (fst + sz))];
		}
		t *= bins[rho];
		val += t;
	}
	return val;
}

void orc_erida_sparse(const double* c, glong K, const glong* g, const glong* idx, glong num, double tol, double* out)
{
	double* F = malloc((2 * K - 1) * sizeof(double));
	glong fst, fln;
	orc_erida_products(c, K, tol, F, &fst, &fln);
	const glong R = lmx(lmx(g[1], g[3]), g[5]);
	const glong nT = orc_erida_table_len(R, K);
	double* T = malloc(nT * sizeof(double));
	orc_erida_coulomb_table(nT, T);
	const glong N = g_num(g);
	const glong cap = 2 * lsq(3 * (R - 1) + (K - 1)) + 2;
	#pragma omp parallel
	{
		double* bins = malloc(cap * sizeof(double));
		#pragma omp for schedule(static)
		for (glong n = 0; n < num; ++n) {
			const glong ij = idx[n] % N, ii = idx[n] / N;
			glong pi[3], pj[3], cd3[3];
			g_pt(g, ii, pi); g_pt(g, ij, pj);
			for (int a = 0; a < 3; ++a) { cd3[a] = 3 * (pi[a] - pj[a]); }
			out[n] = erida_value(cd3, F, fst, fln, tol, T, bins);
		}
		free(bins);
	}
	free(T); free(F);
}

double orc_erida_single(const double* c, glong K, const glong* pts, double tol)
{
	double* F = malloc((2 * K - 1) * sizeof(double));
	glong fst, fln;
	orc_erida_products(c, K, tol, F, &fst, &fln);
	glong R = 0, cd3[3];
	for (int a = 0; a < 3; ++a) {
		R = lmx(R, labs(pts[a] - pts[3 + a]));
		cd3[a] = 3 * (pts[a] - pts[3 + a]);
	}
	R += 1;
	const glong nT = orc_erida_table_len(R, K);
	double* T = malloc(nT * sizeof(double));
	orc_erida_coulomb_table(nT, T);
	const glong nb = lmx(lsq(cd3[0] + fst), lsq(cd3[0] + fst + fln - 1))
	               + lmx(lsq(cd3[1] + fst), lsq(cd3[1] + fst + fln - 1)) + 1;
	double* bins = malloc((nb > 0 ? nb : 1) * sizeof(double));
	const double v = erida_value(cd3, F, fst, fln, tol, T, bins);
	free(bins); free(T); free(F);
	return v;
}

glong orc_min_orbit_erida(glong N, const glong* perm, glong i, glong j)
{
	glong best = INT64_MAX;
	for (int e = 0; e < 48; ++e) {
		glong a = perm[e * N + i], b = perm[e * N + j];
		if (a > b) { glong t = a; a = b; b = t; }
		best = lmn(best, a * N + b);
	}
	return best;
}

glong orc_erida_enumerate(const glong* g, glong* out)
{
	const glong N = g_num(g);
	glong* perm = malloc(48 * N * sizeof(glong));
	orc_octahedral_perms(g, perm);
	struct glist L = {0};
	for (glong icx = 0; icx < g[1]; ++icx)
	for (glong jcx = icx; jcx < g[1]; ++jcx) {
		const glong cx = 2 * g[0] + icx + jcx;
		if (cx < -1 || cx > 1) { continue; }
		for (glong icy = 0; icy < g[3]; ++icy)
		for (glong jcy = 0; jcy < g[3]; ++jcy) {
			const glong cy = 2 * g[2] + icy + jcy;
			if (cy < -1 || cy > 1) { continue; }
			for (glong icz = 0; icz < g[5]; ++icz)
			for (glong jcz = 0; jcz < g[5]; ++jcz) {
				const glong cz = 2 * g[4] + icz + jcz;
				if (cz < -1 || cz > 1) { continue; }
				const glong t = g_lin(g, icx, icy, icz) * N + g_lin(g, jcx, jcy, jcz);
				int ok = 1;
				for (glong bx = 0; bx <= labs(cx) && ok; ++bx)
				for (glong by = 0; by <= labs(cy) && ok; ++by)
				for (glong bz = 0; bz <= labs(cz) && ok; ++bz) {
					const glong bi = g_lin(g, icx - bx * cx, icy - by * cy, icz - bz * cz);
					const glong bj = g_lin(g, jcx - bx * cx, jcy - by * cy, jcz - bz * cz);
					for (int e = 0; e < 48 && ok; ++e) {
						glong a = perm[e * N + bi], b = perm[e * N + bj];
						if (a > b) { glong tt = a; a = b; b = tt; }
						if (a * N + b < t) { ok = 0; }
					}
				}
				if (ok) { glist_push(&L, t); }
			}
		}
	}
	free(perm);
	qsort(L.v, L.n, sizeof(glong), cmp_glong);
	const glong n = L.n;
	if (out != NULL) { memcpy(out, L.v, n * sizeof(glong)); }
	free(L.v);
	return n;
}

/* ------------------------------------------------------------------------- */
/* nuclear attraction (nuclear_integrals.c:16-24, 31-126, 133-266)            */

static double nuc_kernel(double w, double dist)
{
	if (dist > 0) { return erf(dist / w) / dist; }
	return 2 / (sqrt(M_PI) * w);
}

static double norm3(const double v[3])
{
	const double a = fabs(v[0]), b = fabs(v[1]), c = fabs(v[2]);
	double big, s0, s1;
	if (b <= c) {
		if (a <= c) { if (c == 0) { return 0; } big = c; s0 = a; s1 = b; }
		else        { big = a; s0 = b; s1 = c; }
	}
	else {
		if (a <= b) { big = b; s0 = a; s1 = c; }
		else        { big = a; s0 = b; s1 = c; }
	}
	const double r0 = s0 / big, r1 = s1 / big;
	return big * sqrt(1 + r0 * r0 + r1 * r1);
}

/* nuclei: pos[num][3], charge[num] */
static double nuc_pair(const double* fx, glong stx, glong nx_, const double* fy, glong sty, glong ny_,
	const double* fz, glong stz, glong nz_, const double ctr[3], const double* pos, const int* chg, int nn, double tol)
{
	double val = 0;
	for (glong sx = 0; sx < nx_; ++sx) {
		if (fabs(fx[sx]) <= tol) { continue; }
		const double px = ctr[0] + (stx + sx) / 6.0;
		for (glong sy = 0; sy < ny_; ++sy) {
			if (fabs(fy[sy]) <= tol) { continue; }
			const double py = ctr[1] + (sty + sy) / 6.0;
			for (glong sz = 0; sz < nz_; ++sz) {
				if (fabs(fz[sz]) <= tol) { continue; }
				const double pz = ctr[2] + (stz + sz) / 6.0;
				const double f = fx[sx] * fy[sy] * fz[sz];
				for (int k = 0; k < nn; ++k) {
					const double d[3] = { px - pos[3 * k], py - pos[3 * k + 1], pz - pos[3 * k + 2] };
					val += chg[k] * f * nuc_kernel(1. / 3, norm3(d));
				}
			}
		}
	}
	return val;
}

void orc_nuclear_matrix(const double* c, glong K, const glong* g, const double* pos, const int* chg, int nn, double tol, double* out)
{
	const glong R = lmx(lmx(g[1], g[3]), g[5]);
	const glong s0 = -R + 1, ns = 2 * R - 1, L = 2 * K - 1;
	double* fac = malloc(ns * L * sizeof(double));
	glong* fst = malloc(ns * sizeof(glong));
	glong* fln = malloc(ns * sizeof(glong));
	orc_factors(c, K, s0, ns, tol, fac, fst, fln);
	const glong N = g_num(g);
	memset(out, 0, N * N * sizeof(double));
	#pragma omp parallel for schedule(dynamic) collapse(2)
	for (glong icx = 0; icx < g[1]; ++icx)
	for (glong jcx = 0; jcx < g[1]; ++jcx) {
		const glong kx = (icx - jcx) - s0;
		for (glong icy = 0; icy < g[3]; ++icy)
		for (glong jcy = 0; jcy < g[3]; ++jcy) {
			const glong ky = (icy - jcy) - s0;
			for (glong icz = 0; icz < g[5]; ++icz)
			for (glong jcz = 0; jcz < g[5]; ++jcz) {
				const glong kz = (icz - jcz) - s0;
				const glong ii = g_lin(g, icx, icy, icz), jj = g_lin(g, jcx, jcy, jcz);
				if (ii > jj) { continue; }
				const double ctr[3] = { g[0] + 0.5 * (icx + jcx), g[2] + 0.5 * (icy + jcy), g[4] + 0.5 * (icz + jcz) };
				out[ii * N + jj] = nuc_pair(&fac[kx * L], fst[kx], fln[kx], &fac[ky * L], fst[ky], fln[ky],
					&fac[kz * L], fst[kz], fln[kz], ctr, pos, chg, nn, tol);
			}
		}
	}
	for (glong i = 0; i < N; ++i) {
		for (glong j = 0; j < i; ++j) { out[i * N + j] = out[j * N + i]; }
	}
	free(fln); free(fst); free(fac);
}

double orc_nuclear_single(const double* c, glong K, const glong* pts, const double* pos, const int* chg, int nn, double tol)
{
	glong R = 0;
	for (int a = 0; a < 3; ++a) { R = lmx(R, labs(pts[a] - pts[3 + a])); }
	R += 1;
	const glong s0 = -R + 1, ns = 2 * R - 1, L = 2 * K - 1;
	double* fac = malloc(ns * L * sizeof(double));
	glong* fst = malloc(ns * sizeof(glong));
	glong* fln = malloc(ns * sizeof(glong));
	orc_factors(c, K, s0, ns, tol, fac, fst, fln);
	const double ctr[3] = { 0.5 * (pts[0] + pts[3]), 0.5 * (pts[1] + pts[4]), 0.5 * (pts[2] + pts[5]) };
	glong k3[3];
	for (int a = 0; a < 3; ++a) { k3[a] = (pts[a] - pts[3 + a]) - s0; }
	const double v = nuc_pair(&fac[k3[0] * L], fst[k3[0]], fln[k3[0]], &fac[k3[1] * L], fst[k3[1]], fln[k3[1]],
		&fac[k3[2] * L], fst[k3[2]], fln[k3[2]], ctr, pos, chg, nn, tol);
	free(fln); free(fst); free(fac);
	return v;
}

/* ------------------------------------------------------------------------- */
/* kinetic (kinetic_integrals.c:14-37, 44-121)                                */

static double kin1d(const double* c, glong K, glong shift)
{
	double val = 0;
	for (glong i = 0; i < K; ++i) {
		for (glong j = 0; j < K; ++j) {
			const glong d = 3 * shift + (i - j);
			const double d2 = d * d;
			val += 0.5 * c[i] * c[j] * (1.5 * sqrt(M_PI) * (1 - 0.5 * d2) * exp(-0.25 * d2));
		}
	}
	return val;
}

void orc_kinetic_1d(const double* c, glong K, glong nshift, double* out)
{
	for (glong s = 0; s < nshift; ++s) { out[s] = kin1d(c, K, s); }
}

void orc_kinetic_matrix(const double* c, glong K, const glong* g, double* out)
{
	const glong R = lmx(lmx(g[1], g[3]), g[5]);
	double* t = malloc(R * sizeof(double));
	orc_kinetic_1d(c, K, R, t);
	const glong N = g_num(g);
	for (glong icx = 0; icx < g[1]; ++icx)
	for (glong jcx = 0; jcx < g[1]; ++jcx)
	for (glong icy = 0; icy < g[3]; ++icy)
	for (glong jcy = 0; jcy < g[3]; ++jcy)
	for (glong icz = 0; icz < g[5]; ++icz)
	for (glong jcz = 0; jcz < g[5]; ++jcz) {
		const glong ii = g_lin(g, icx, icy, icz), jj = g_lin(g, jcx, jcy, jcz);
		out[ii * N + jj] =
			t[labs(icx - jcx)] * (icy == jcy ? 1 : 0) * (icz == jcz ? 1 : 0) +
			(icx == jcx ? 1 : 0) * t[labs(icy - jcy)] * (icz == jcz ? 1 : 0) +
			(icx == jcx ? 1 : 0) * (icy == jcy ? 1 : 0) * t[labs(icz - jcz)];
	}
	free(t);
}

/* ------------------------------------------------------------------------- */
/* lookups and dense reconstruction (eri_integrals.c:685-707, 850-967;
 * erida_integrals.c:549-571, 683-765)                                        */

double orc_lookup(const glong* keys, const double* vals, glong num, glong index)
{
	glong lo = 0, hi = num;
	while (lo < hi) {
		const glong m = (lo + hi) / 2;
		if (index < keys[m]) { hi = m; }
		else if (index > keys[m]) { lo = m + 1; }
		else { return vals[m]; }
	}
	return 0;
}

/* canonical sparse key of a dense four-tuple given by grid offsets on g_dense */
static glong eri_dense_key(const glong* gs, glong Ns, const glong* perm, const glong* gd, const glong off[4][3])
{
	glong ctr[3], tr[3], bi[3];
	for (int a = 0; a < 3; ++a) {
		const glong mn = lmn(lmn(lmn(off[0][a], off[1][a]), off[2][a]), off[3][a]);
		const glong mx = lmx(lmx(lmx(off[0][a], off[1][a]), off[2][a]), off[3][a]);
		ctr[a] = 2 * gd[2 * a] + mn + mx;
		tr[a] = ctr[a] / 2;
		bi[a] = ctr[a] - 2 * tr[a];
	}
	glong best = INT64_MAX;
	for (glong bx = 0; bx <= labs(bi[0]); ++bx)
	for (glong by = 0; by <= labs(bi[1]); ++by)
	for (glong bz = 0; bz <= labs(bi[2]); ++bz) {
		const glong b[3] = { bx, by, bz };
		glong id[4];
		for (int p = 0; p < 4; ++p) {
			glong q[3];
			for (int a = 0; a < 3; ++a) { q[a] = gd[2 * a] + off[p][a] - tr[a] - b[a] * bi[a] - gs[2 * a]; }
			id[p] = g_lin(gs, q[0], q[1], q[2]);
		}
		best = lmn(best, orc_min_orbit_eri(Ns, perm, id[0], id[1], id[2], id[3]));
	}
	return best;
}

void orc_fill_dense_eri(const glong* gs, const glong* keys, const double* vals, glong num, const glong* gd, double* out)
{
	const glong Ns = g_num(gs), Nd = g_num(gd);
	glong* perm = malloc(48 * Ns * sizeof(glong));
	orc_octahedral_perms(gs, perm);
	#pragma omp parallel for schedule(dynamic)
	for (glong i = 0; i < Nd; ++i) {
		for (glong j = 0; j < Nd; ++j)
		for (glong k = 0; k < Nd; ++k)
		for (glong l = 0; l < Nd; ++l) {
			const glong lin[4] = { i, j, k, l };
			glong off[4][3];
			for (int p = 0; p < 4; ++p) {
				glong pt[3];
				g_pt(gd, lin[p], pt);
				for (int a = 0; a < 3; ++a) { off[p][a] = pt[a] - gd[2 * a]; }
			}
			out[((i * Nd + j) * Nd + k) * Nd + l] = orc_lookup(keys, vals, num, eri_dense_key(gs, Ns, perm, gd, off));
		}
	}
	free(perm);
}

void orc_fill_dense_erida(const glong* gs, const glong* keys, const double* vals, glong num, const glong* gd, double* out)
{
	const glong Ns = g_num(gs), Nd = g_num(gd);
	glong* perm = malloc(48 * Ns * sizeof(glong));
	orc_octahedral_perms(gs, perm);
	for (glong i = 0; i < Nd; ++i) {
		for (glong j = 0; j < Nd; ++j) {
			glong pi[3], pj[3], tr[3], bi[3];
			g_pt(gd, i, pi); g_pt(gd, j, pj);
			for (int a = 0; a < 3; ++a) {
				const glong ctr = pi[a] + pj[a];  /* = 2*istart + ic + jc */
				tr[a] = ctr / 2;
				bi[a] = ctr - 2 * tr[a];
			}
			glong best = INT64_MAX;
			for (glong bx = 0; bx <= labs(bi[0]); ++bx)
			for (glong by = 0; by <= labs(bi[1]); ++by)
			for (glong bz = 0; bz <= labs(bi[2]); ++bz) {
				const glong b[3] = { bx, by, bz };
				glong qi[3], qj[3];
				for (int a = 0; a < 3; ++a) {
					qi[a] = pi[a] - tr[a] - b[a] * bi[a] - gs[2 * a];
					qj[a] = pj[a] - tr[a] - b[a] * bi[a] - gs[2 * a];
				}
				best = lmn(best, orc_min_orbit_erida(Ns, perm, g_lin(gs, qi[0], qi[1], qi[2]), g_lin(gs, qj[0], qj[1], qj[2])));
			}
			out[i * Nd + j] = orc_lookup(keys, vals, num, best);
		}
	}
	free(perm);
}

int orc_num_threads(void)
{
#ifdef _OPENMP
	return omp_get_max_threads();
#else
	return 1;
#endif
}

/* ------------------------------------------------------------------------- */
/* canonical key of each four-index of a symmetric grid: minimum over the bias
 * placements of its (centred) box and the O_h x label orbit
 * (pymodule.c:700-780 lookup route, eri_integrals.c:240-291)                 */

void orc_eri_canonical_keys(const glong* g, const glong* idx, glong num, glong* out)
{
	const glong N = g_num(g);
	glong* perm = malloc(48 * N * sizeof(glong));
	orc_octahedral_perms(g, perm);
	#pragma omp parallel for schedule(static)
	for (glong e = 0; e < num; ++e) {
		glong r = idx[e];
		const glong lin[4] = { r / (N * N * N), (r / (N * N)) % N, (r / N) % N, r % N };
		glong pt[4][3];
		for (int p = 0; p < 4; ++p) { g_pt(g, lin[p], pt[p]); }
		glong tr[3], bi[3];
		for (int a = 0; a < 3; ++a) {
			const glong c = lmn(lmn(pt[0][a], pt[1][a]), lmn(pt[2][a], pt[3][a])) + lmx(lmx(pt[0][a], pt[1][a]), lmx(pt[2][a], pt[3][a]));
			tr[a] = c / 2;
			bi[a] = c - 2 * tr[a];
		}
		glong best = INT64_MAX;
		for (glong bx = 0; bx <= labs(bi[0]); ++bx)
		for (glong by = 0; by <= labs(bi[1]); ++by)
		for (glong bz = 0; bz <= labs(bi[2]); ++bz) {
			const glong b[3] = { bx, by, bz };
			glong q[4];
			for (int p = 0; p < 4; ++p) {
				glong o[3];
				for (int a = 0; a < 3; ++a) { o[a] = pt[p][a] - tr[a] - b[a] * bi[a] - g[2 * a]; }
				q[p] = g_lin(g, o[0], o[1], o[2]);
			}
			best = lmn(best, orc_min_orbit_eri(N, perm, q[0], q[1], q[2], q[3]));
		}
		out[e] = best;
	}
	free(perm);
}
