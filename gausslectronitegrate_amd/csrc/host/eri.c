/* Reference-compatible ERI entry points (eri_integrals.h) over the GPU path.
 * Enumeration and values run on the GPU; lookups and the dense/contracted
 * views of a sparse set are host restatements (eri_integrals.c:685-1100). */
#include <stdlib.h>
#include <string.h>
#include "eri_integrals.h"
#include "symmetry.h"
#include "aligned_memory.h"
#include "gli_amd.h"
#include "host_internal.h"

/* ------------------------------------------------------------------------- */
/* canonicalisation (eri_integrals.c:240-291)                                 */

static inline glong eri_image_index(const glong N, const glong* p, glong i, glong j, glong k, glong l)
{
	glong a = p[i], b = p[j], c = p[k], d = p[l];
	if (a > b) { const glong t = a; a = b; b = t; }
	if (c > d) { const glong t = c; c = d; d = t; }
	glong ab = a * N + b, cd = c * N + d;
	if (ab > cd) { const glong t = ab; ab = cd; cd = t; }
	return ab * (N * N) + cd;
}

glong minimum_octahedral_orbit_eri_tensor_index(const glong num_points, const glong* octahedral_perm[48],
	const glong i, const glong j, const glong k, const glong l)
{
	glong best = GLONG_MAX;
	for (int e = 0; e < 48; ++e) {
		best = lmin(best, eri_image_index(num_points, octahedral_perm[e], i, j, k, l));
	}
	return best;
}

/* centre the tuple (translation by half the bounding-box centre, truncated), try both
 * placements of odd spans and take the orbit minimum (eri_integrals.c:733-803) */
glong gli_eri_sparse_key(const struct cartesian_grid_3d* gs, const glong* const perm[48], const union cartesian_grid_point_3d pts[4])
{
	const glong N = cartesian_grid_3d_num_points(gs);
	glong tr[3], bi[3];
	for (int a = 0; a < 3; ++a) {
		const glong mn = lmin(lmin(pts[0].c[a], pts[1].c[a]), lmin(pts[2].c[a], pts[3].c[a]));
		const glong mx = lmax(lmax(pts[0].c[a], pts[1].c[a]), lmax(pts[2].c[a], pts[3].c[a]));
		const glong centre = mn + mx;
		tr[a] = centre / 2;
		bi[a] = centre - 2 * tr[a];
	}
	glong best = GLONG_MAX;
	for (glong bx = 0; bx <= labs(bi[0]); ++bx) {
		for (glong by = 0; by <= labs(bi[1]); ++by) {
			for (glong bz = 0; bz <= labs(bi[2]); ++bz) {
				const glong b[3] = { bx, by, bz };
				glong lin[4];
				for (int p = 0; p < 4; ++p) {
					glong off[3];
					for (int a = 0; a < 3; ++a) {
						off[a] = pts[p].c[a] - tr[a] - b[a] * bi[a] - gs->coord_range[a].istart;
						if (off[a] < 0 || off[a] >= gs->coord_range[a].num) { return -1; }
					}
					lin[p] = (off[0] * gs->coord_range[1].num + off[1]) * gs->coord_range[2].num + off[2];
				}
				for (int e = 0; e < 48; ++e) {
					best = lmin(best, eri_image_index(N, perm[e], lin[0], lin[1], lin[2], lin[3]));
				}
			}
		}
	}
	return best;
}

/* ------------------------------------------------------------------------- */
/* enumeration and values                                                     */

void enumerate_symmetry_reduced_eri_indices(const struct cartesian_grid_3d* grid, struct sparse_eri_indices* ret)
{
	gli_require_symmetric_cubic(grid, "enumerate_symmetry_reduced_eri_indices");
	ret->grid = *grid;
	int64_t* out = NULL;
	int64_t num = 0;
	gli_check(gli_hip_eri_enumerate(grid->coord_range[0].num, &out, &num), "enumerate_symmetry_reduced_eri_indices");
	ret->four_indices = out;
	ret->num = num;
}

void delete_sparse_eri_indices(struct sparse_eri_indices* indices)
{
	aligned_free(indices->four_indices);
	indices->four_indices = NULL;
	indices->num = 0;
}

void compute_sparse_eri_gausslet_integrals(const struct gausslet_data* gdata,
	const struct cartesian_grid_3d* grid, const glong* four_indices, const glong num_indices,
	const double tol, struct sparse_eri_gausslet_integrals* eri)
{
	gli_require_coefficients(gdata, "compute_sparse_eri_gausslet_integrals");
	eri->grid = *grid;
	eri->num_entries = num_indices;
	eri->four_indices = aligned_malloc(num_indices * sizeof(glong));
	eri->integral_values = aligned_malloc(num_indices * sizeof(double));
	if (num_indices > 0) { memcpy(eri->four_indices, four_indices, num_indices * sizeof(glong)); }
	int64_t g[6];
	gli_grid_array(grid, g);
	gli_check(gli_hip_eri_values(gdata->coefficients, gdata->indices.num, g, four_indices, num_indices, tol,
	                             gli_get_eri_method(), eri->integral_values),
	          "compute_sparse_eri_gausslet_integrals");
}

double compute_eri_gausslet_integral(const struct gausslet_data* gdata, const union cartesian_grid_point_3d points[4], const double tol)
{
	gli_require_coefficients(gdata, "compute_eri_gausslet_integral");
	/* the smallest grid holding the four points; the value does not depend on the grid */
	int64_t g[6];
	for (int a = 0; a < 3; ++a) {
		const glong mn = lmin(lmin(points[0].c[a], points[1].c[a]), lmin(points[2].c[a], points[3].c[a]));
		const glong mx = lmax(lmax(points[0].c[a], points[1].c[a]), lmax(points[2].c[a], points[3].c[a]));
		g[2 * a] = mn;
		g[2 * a + 1] = mx - mn + 1;
	}
	const glong N = g[1] * g[3] * g[5];
	glong lin[4];
	for (int p = 0; p < 4; ++p) {
		lin[p] = ((points[p].x - g[0]) * g[3] + (points[p].y - g[2])) * g[5] + (points[p].z - g[4]);
	}
	const glong t = ((lin[0] * N + lin[1]) * N + lin[2]) * N + lin[3];
	double v = 0;
	gli_check(gli_hip_eri_values(gdata->coefficients, gdata->indices.num, g, &t, 1, tol, gli_get_eri_method(), &v),
	          "compute_eri_gausslet_integral");
	return v;
}

/* binary search (eri_integrals.c:685-707) */
double sparse_eri_gausslet_integrals_get_value(const struct sparse_eri_gausslet_integrals* eri, const glong index)
{
	glong lo = 0, hi = eri->num_entries;
	while (lo < hi) {
		const glong m = (lo + hi) / 2;
		if (index < eri->four_indices[m]) { hi = m; }
		else if (index > eri->four_indices[m]) { lo = m + 1; }
		else { return eri->integral_values[m]; }
	}
	return 0;
}

void delete_sparse_eri_gausslet_integrals(struct sparse_eri_gausslet_integrals* eri)
{
	aligned_free(eri->four_indices);
	aligned_free(eri->integral_values);
	eri->four_indices = NULL;
	eri->integral_values = NULL;
	eri->num_entries = 0;
}

/* ------------------------------------------------------------------------- */
/* dense views (eri_integrals.c:714-1100)                                     */

static void perms_of(const struct cartesian_grid_3d* g, glong* perm[48])
{
	evaluate_octahedral_grid_permutations(g, perm);
}

static void free_perms(glong* perm[48])
{
	for (int e = 0; e < 48; ++e) { aligned_free(perm[e]); }
}

/* value of the dense tuple (i, j, k, l) of grid gd, looked up in the sparse set */
static double dense_value(const struct sparse_eri_gausslet_integrals* eri, const glong* const perm[48],
	const struct cartesian_grid_3d* gd, glong i, glong j, glong k, glong l)
{
	union cartesian_grid_point_3d pts[4];
	linear_index_to_cartesian_grid_point_3d(gd, i, &pts[0]);
	linear_index_to_cartesian_grid_point_3d(gd, j, &pts[1]);
	linear_index_to_cartesian_grid_point_3d(gd, k, &pts[2]);
	linear_index_to_cartesian_grid_point_3d(gd, l, &pts[3]);
	const glong key = gli_eri_sparse_key(&eri->grid, perm, pts);
	return key < 0 ? 0 : sparse_eri_gausslet_integrals_get_value(eri, key);
}

void fill_dense_eri_tensor(const struct sparse_eri_gausslet_integrals* eri_sparse, const struct cartesian_grid_3d* grid_dense, double* eri_tensor)
{
	const glong Nd = cartesian_grid_3d_num_points(grid_dense);
	glong* perm[48];
	perms_of(&eri_sparse->grid, perm);
	#pragma omp parallel for schedule(dynamic) collapse(2)
	for (glong i = 0; i < Nd; ++i) {
		for (glong j = 0; j < Nd; ++j) {
			for (glong k = 0; k < Nd; ++k) {
				for (glong l = 0; l < Nd; ++l) {
					eri_tensor[((i * Nd + j) * Nd + k) * Nd + l] = dense_value(eri_sparse, (const glong* const*)perm, grid_dense, i, j, k, l);
				}
			}
		}
	}
	free_perms(perm);
}

/* eri_proj[p,q,r,s] = sum_{ijkl} b[i,p] b[j,q] b[k,r] b[l,s] (ij|kl) over the sparse grid,
 * contracted one (i, j) block at a time */
void project_sparse_eri_gausslet_integrals(const struct sparse_eri_gausslet_integrals* eri, const double* basis, const glong num_states, double* eri_proj)
{
	const glong N = cartesian_grid_3d_num_points(&eri->grid);
	const glong S = num_states;
	const glong S4 = S * S * S * S;
	glong* perm[48];
	perms_of(&eri->grid, perm);
	memset(eri_proj, 0, S4 * sizeof(double));
	#pragma omp parallel
	{
		double* acc = aligned_calloc(S4 * sizeof(double));
		double* V = aligned_malloc(N * N * sizeof(double));
		double* VB = aligned_malloc(N * S * sizeof(double));
		double* W = aligned_malloc(S * S * sizeof(double));
		#pragma omp for schedule(dynamic) collapse(2)
		for (glong i = 0; i < N; ++i) {
			for (glong j = 0; j < N; ++j) {
				for (glong k = 0; k < N; ++k) {
					for (glong l = 0; l < N; ++l) {
						V[k * N + l] = dense_value(eri, (const glong* const*)perm, &eri->grid, i, j, k, l);
					}
				}
				/* VB[k,s] = sum_l V[k,l] b[l,s];  W[r,s] = sum_k b[k,r] VB[k,s] */
				for (glong k = 0; k < N; ++k) {
					for (glong s = 0; s < S; ++s) {
						double t = 0;
						for (glong l = 0; l < N; ++l) { t += V[k * N + l] * basis[l * S + s]; }
						VB[k * S + s] = t;
					}
				}
				for (glong r = 0; r < S; ++r) {
					for (glong s = 0; s < S; ++s) {
						double t = 0;
						for (glong k = 0; k < N; ++k) { t += basis[k * S + r] * VB[k * S + s]; }
						W[r * S + s] = t;
					}
				}
				for (glong p = 0; p < S; ++p) {
					for (glong q = 0; q < S; ++q) {
						const double bpq = basis[i * S + p] * basis[j * S + q];
						for (glong rs = 0; rs < S * S; ++rs) { acc[(p * S + q) * S * S + rs] += bpq * W[rs]; }
					}
				}
			}
		}
		#pragma omp critical
		{
			for (glong x = 0; x < S4; ++x) { eri_proj[x] += acc[x]; }
		}
		aligned_free(W); aligned_free(VB); aligned_free(V); aligned_free(acc);
	}
	free_perms(perm);
}

/* ret[(i,k), s] = sum_{j,l} (ij|kl) states[(j,l), s]  (physicist's grouping) */
void apply_sparse_eri_tensor(const struct sparse_eri_gausslet_integrals* eri, const double* states, const glong num_states, double* ret)
{
	const glong N = cartesian_grid_3d_num_points(&eri->grid);
	const glong S = num_states;
	glong* perm[48];
	perms_of(&eri->grid, perm);
	#pragma omp parallel for schedule(dynamic) collapse(2)
	for (glong i = 0; i < N; ++i) {
		for (glong k = 0; k < N; ++k) {
			double* r = &ret[(i * N + k) * S];
			for (glong s = 0; s < S; ++s) { r[s] = 0; }
			for (glong j = 0; j < N; ++j) {
				for (glong l = 0; l < N; ++l) {
					const double v = dense_value(eri, (const glong* const*)perm, &eri->grid, i, j, k, l);
					const double* x = &states[(j * N + l) * S];
					for (glong s = 0; s < S; ++s) { r[s] += v * x[s]; }
				}
			}
		}
	}
	free_perms(perm);
}
