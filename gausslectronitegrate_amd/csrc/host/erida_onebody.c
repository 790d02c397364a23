/* Reference-compatible ERIDA, nuclear and kinetic entry points
 * (erida_integrals.h, nuclear_integrals.h, kinetic_integrals.h) over the GPU path. */
#include <stdlib.h>
#include <string.h>
#include "erida_integrals.h"
#include "nuclear_integrals.h"
#include "kinetic_integrals.h"
#include "symmetry.h"
#include "aligned_memory.h"
#include "gli_amd.h"
#include "host_internal.h"

/* ------------------------------------------------------------------------- */
/* ERIDA                                                                      */

glong minimum_octahedral_orbit_erida_tensor_index(const glong num_points, const glong* octahedral_perm[48], const glong i, const glong j)
{
	glong best = GLONG_MAX;
	for (int e = 0; e < 48; ++e) {
		glong a = octahedral_perm[e][i], b = octahedral_perm[e][j];
		if (a > b) { const glong t = a; a = b; b = t; }
		best = lmin(best, a * num_points + b);
	}
	return best;
}

glong gli_erida_sparse_key(const struct cartesian_grid_3d* gs, const glong* const perm[48], const union cartesian_grid_point_3d pts[2])
{
	const glong N = cartesian_grid_3d_num_points(gs);
	glong tr[3], bi[3];
	for (int a = 0; a < 3; ++a) {
		const glong centre = pts[0].c[a] + pts[1].c[a];
		tr[a] = centre / 2;
		bi[a] = centre - 2 * tr[a];
	}
	glong best = GLONG_MAX;
	for (glong bx = 0; bx <= labs(bi[0]); ++bx) {
		for (glong by = 0; by <= labs(bi[1]); ++by) {
			for (glong bz = 0; bz <= labs(bi[2]); ++bz) {
				const glong b[3] = { bx, by, bz };
				glong lin[2];
				for (int p = 0; p < 2; ++p) {
					glong off[3];
					for (int a = 0; a < 3; ++a) {
						off[a] = pts[p].c[a] - tr[a] - b[a] * bi[a] - gs->coord_range[a].istart;
						if (off[a] < 0 || off[a] >= gs->coord_range[a].num) { return -1; }
					}
					lin[p] = (off[0] * gs->coord_range[1].num + off[1]) * gs->coord_range[2].num + off[2];
				}
				for (int e = 0; e < 48; ++e) {
					glong x = perm[e][lin[0]], y = perm[e][lin[1]];
					if (x > y) { const glong t = x; x = y; y = t; }
					best = lmin(best, x * N + y);
				}
			}
		}
	}
	return best;
}

void enumerate_symmetry_reduced_erida_indices(const struct cartesian_grid_3d* grid, struct sparse_erida_indices* ret)
{
	gli_require_symmetric_cubic(grid, "enumerate_symmetry_reduced_erida_indices");
	ret->grid = *grid;
	int64_t* out = NULL;
	int64_t num = 0;
	gli_check(gli_hip_erida_enumerate(grid->coord_range[0].num, &out, &num), "enumerate_symmetry_reduced_erida_indices");
	ret->two_indices = out;
	ret->num = num;
}

void delete_sparse_erida_indices(struct sparse_erida_indices* indices)
{
	aligned_free(indices->two_indices);
	indices->two_indices = NULL;
	indices->num = 0;
}

void compute_sparse_erida_gausslet_integrals(const struct gausslet_data* gdata,
	const struct cartesian_grid_3d* grid, const glong* two_indices, const glong num_indices,
	const double tol, struct sparse_erida_gausslet_integrals* erida)
{
	gli_require_coefficients(gdata, "compute_sparse_erida_gausslet_integrals");
	erida->grid = *grid;
	erida->num_entries = num_indices;
	erida->two_indices = gli_require_host(aligned_malloc(num_indices * sizeof(glong)), num_indices * sizeof(glong),
	                                      "compute_sparse_erida_gausslet_integrals");
	erida->integral_values = gli_require_host(aligned_malloc(num_indices * sizeof(double)), num_indices * sizeof(double),
	                                          "compute_sparse_erida_gausslet_integrals");
	if (num_indices > 0) { memcpy(erida->two_indices, two_indices, num_indices * sizeof(glong)); }
	int64_t g[6];
	gli_grid_array(grid, g);
	gli_check(gli_hip_erida_values(gdata->coefficients, gdata->indices.num, g, two_indices, num_indices, tol, erida->integral_values),
	          "compute_sparse_erida_gausslet_integrals");
}

double compute_erida_gausslet_integral(const struct gausslet_data* gdata, const union cartesian_grid_point_3d points[2], const double tol)
{
	gli_require_coefficients(gdata, "compute_erida_gausslet_integral");
	int64_t g[6];
	for (int a = 0; a < 3; ++a) {
		const glong mn = lmin(points[0].c[a], points[1].c[a]), mx = lmax(points[0].c[a], points[1].c[a]);
		g[2 * a] = mn;
		g[2 * a + 1] = mx - mn + 1;
	}
	const glong N = g[1] * g[3] * g[5];
	glong lin[2];
	for (int p = 0; p < 2; ++p) {
		lin[p] = ((points[p].x - g[0]) * g[3] + (points[p].y - g[2])) * g[5] + (points[p].z - g[4]);
	}
	const glong t = lin[0] * N + lin[1];
	double v = 0;
	gli_check(gli_hip_erida_values(gdata->coefficients, gdata->indices.num, g, &t, 1, tol, &v), "compute_erida_gausslet_integral");
	return v;
}

double sparse_erida_gausslet_integrals_get_value(const struct sparse_erida_gausslet_integrals* erida, const glong index)
{
	glong lo = 0, hi = erida->num_entries;
	while (lo < hi) {
		const glong m = (lo + hi) / 2;
		if (index < erida->two_indices[m]) { hi = m; }
		else if (index > erida->two_indices[m]) { lo = m + 1; }
		else { return erida->integral_values[m]; }
	}
	return 0;
}

static double erida_dense_value(const struct sparse_erida_gausslet_integrals* erida, const glong* const perm[48],
	const struct cartesian_grid_3d* gd, glong i, glong j)
{
	union cartesian_grid_point_3d pts[2];
	linear_index_to_cartesian_grid_point_3d(gd, i, &pts[0]);
	linear_index_to_cartesian_grid_point_3d(gd, j, &pts[1]);
	const glong key = gli_erida_sparse_key(&erida->grid, perm, pts);
	return key < 0 ? 0 : sparse_erida_gausslet_integrals_get_value(erida, key);
}

void fill_dense_erida_matrix(const struct sparse_erida_gausslet_integrals* erida_sparse, const struct cartesian_grid_3d* grid_dense, double* erida_matrix)
{
	const glong Nd = cartesian_grid_3d_num_points(grid_dense);
	glong* perm[48];
	evaluate_octahedral_grid_permutations(&erida_sparse->grid, perm);
	#pragma omp parallel for schedule(dynamic)
	for (glong i = 0; i < Nd; ++i) {
		for (glong j = 0; j < Nd; ++j) {
			erida_matrix[i * Nd + j] = erida_dense_value(erida_sparse, (const glong* const*)perm, grid_dense, i, j);
		}
	}
	for (int e = 0; e < 48; ++e) { aligned_free(perm[e]); }
}

/* eri_proj[p,q,r,s] = sum_{i,j} b[i,p] b[i,q] b[j,r] b[j,s] v_ij  (erida_integrals.c:580-676) */
void project_sparse_erida_gausslet_integrals(const struct sparse_erida_gausslet_integrals* erida, const double* restrict basis, const glong num_states, double* restrict eri_proj)
{
	const glong N = cartesian_grid_3d_num_points(&erida->grid);
	const glong S = num_states, S2 = S * S;
	glong* perm[48];
	evaluate_octahedral_grid_permutations(&erida->grid, perm);
	double* V = gli_require_host(aligned_malloc(N * N * sizeof(double)), N * N * sizeof(double), "project_sparse_erida_gausslet_integrals");
	#pragma omp parallel for schedule(dynamic)
	for (glong i = 0; i < N; ++i) {
		for (glong j = 0; j < N; ++j) { V[i * N + j] = erida_dense_value(erida, (const glong* const*)perm, &erida->grid, i, j); }
	}
	/* Khatri-Rao rows KR[i, (p,q)] = b[i,p] b[i,q];  proj = KR^T V KR */
	double* KR = gli_require_host(aligned_malloc(N * S2 * sizeof(double)), N * S2 * sizeof(double), "project_sparse_erida_gausslet_integrals");
	for (glong i = 0; i < N; ++i) {
		for (glong p = 0; p < S; ++p) {
			for (glong q = 0; q < S; ++q) { KR[i * S2 + p * S + q] = basis[i * S + p] * basis[i * S + q]; }
		}
	}
	double* VK = gli_require_host(aligned_malloc(N * S2 * sizeof(double)), N * S2 * sizeof(double), "project_sparse_erida_gausslet_integrals");
	#pragma omp parallel for
	for (glong i = 0; i < N; ++i) {
		for (glong rs = 0; rs < S2; ++rs) {
			double t = 0;
			for (glong j = 0; j < N; ++j) { t += V[i * N + j] * KR[j * S2 + rs]; }
			VK[i * S2 + rs] = t;
		}
	}
	for (glong pq = 0; pq < S2; ++pq) {
		for (glong rs = 0; rs < S2; ++rs) {
			double t = 0;
			for (glong i = 0; i < N; ++i) { t += KR[i * S2 + pq] * VK[i * S2 + rs]; }
			eri_proj[pq * S2 + rs] = t;
		}
	}
	aligned_free(VK); aligned_free(KR); aligned_free(V);
	for (int e = 0; e < 48; ++e) { aligned_free(perm[e]); }
}

void delete_sparse_erida_gausslet_integrals(struct sparse_erida_gausslet_integrals* erida)
{
	aligned_free(erida->two_indices);
	aligned_free(erida->integral_values);
	erida->two_indices = NULL;
	erida->integral_values = NULL;
	erida->num_entries = 0;
}

/* ------------------------------------------------------------------------- */
/* nuclear attraction                                                         */

static void nucleus_arrays(const struct atomic_nucleus* nuclei, int num, double* pos, int* chg)
{
	for (int k = 0; k < num; ++k) {
		for (int a = 0; a < 3; ++a) { pos[3 * k + a] = nuclei[k].pos[a]; }
		chg[k] = nuclei[k].charge;
	}
}

void compute_nuclear_gausslet_integrals(const struct gausslet_data* gdata, const struct cartesian_grid_3d* grid,
	const struct atomic_nucleus* nuclei, const int num_nuclei, const double tol, struct nuclear_gausslet_integrals* ngi)
{
	gli_require_coefficients(gdata, "compute_nuclear_gausslet_integrals");
	ngi->grid = *grid;
	ngi->nuclei = gli_require_host(aligned_malloc(num_nuclei * sizeof(struct atomic_nucleus)), num_nuclei * sizeof(struct atomic_nucleus),
	                               "compute_nuclear_gausslet_integrals");
	if (num_nuclei > 0) { memcpy(ngi->nuclei, nuclei, num_nuclei * sizeof(struct atomic_nucleus)); }
	ngi->num_nuclei = num_nuclei;
	const glong N = cartesian_grid_3d_num_points(grid);
	ngi->integral_values = gli_require_host(gli_host_alloc(N * N * sizeof(double)), N * N * sizeof(double),
	                                        "compute_nuclear_gausslet_integrals");
	double* pos = gli_require_host(malloc((3 * num_nuclei + 1) * sizeof(double)), (3 * num_nuclei + 1) * sizeof(double),
	                               "compute_nuclear_gausslet_integrals");
	int* chg = gli_require_host(malloc((num_nuclei + 1) * sizeof(int)), (num_nuclei + 1) * sizeof(int),
	                            "compute_nuclear_gausslet_integrals");
	nucleus_arrays(nuclei, num_nuclei, pos, chg);
	int64_t g[6];
	gli_grid_array(grid, g);
	gli_check(gli_hip_nuclear_matrix(gdata->coefficients, gdata->indices.num, g, pos, chg, num_nuclei, tol, ngi->integral_values),
	          "compute_nuclear_gausslet_integrals");
	free(chg);
	free(pos);
}

double compute_nuclear_gausslet_integral(const struct gausslet_data* gdata, const union cartesian_grid_point_3d points[2],
	const struct atomic_nucleus* nuclei, const int num_nuclei, const double tol)
{
	gli_require_coefficients(gdata, "compute_nuclear_gausslet_integral");
	/* the two points as the corners of a small grid: entry (first, second) of its matrix */
	int64_t g[6];
	for (int a = 0; a < 3; ++a) {
		const glong mn = lmin(points[0].c[a], points[1].c[a]), mx = lmax(points[0].c[a], points[1].c[a]);
		g[2 * a] = mn;
		g[2 * a + 1] = mx - mn + 1;
	}
	const glong N = g[1] * g[3] * g[5];
	glong lin[2];
	for (int p = 0; p < 2; ++p) {
		lin[p] = ((points[p].x - g[0]) * g[3] + (points[p].y - g[2])) * g[5] + (points[p].z - g[4]);
	}
	double* M = gli_require_host(aligned_malloc(N * N * sizeof(double)), N * N * sizeof(double), "compute_nuclear_gausslet_integral");
	double* pos = gli_require_host(malloc((3 * num_nuclei + 1) * sizeof(double)), (3 * num_nuclei + 1) * sizeof(double),
	                               "compute_nuclear_gausslet_integral");
	int* chg = gli_require_host(malloc((num_nuclei + 1) * sizeof(int)), (num_nuclei + 1) * sizeof(int),
	                            "compute_nuclear_gausslet_integral");
	nucleus_arrays(nuclei, num_nuclei, pos, chg);
	gli_check(gli_hip_nuclear_matrix(gdata->coefficients, gdata->indices.num, g, pos, chg, num_nuclei, tol, M),
	          "compute_nuclear_gausslet_integral");
	/* the matrix is filled from its upper triangle, so read (min, max) */
	const double v = M[lmin(lin[0], lin[1]) * N + lmax(lin[0], lin[1])];
	free(chg);
	free(pos);
	aligned_free(M);
	return v;
}

void delete_nuclear_gausslet_integrals(struct nuclear_gausslet_integrals* ngi)
{
	aligned_free(ngi->integral_values);
	aligned_free(ngi->nuclei);
	ngi->integral_values = NULL;
	ngi->nuclei = NULL;
	ngi->num_nuclei = 0;
}

/* ------------------------------------------------------------------------- */
/* kinetic                                                                    */

void compute_kinetic_gausslet_integrals(const struct gausslet_data* gdata, const struct cartesian_grid_3d* grid, struct kinetic_gausslet_integrals* kgi)
{
	kgi->grid = *grid;
	const glong N = cartesian_grid_3d_num_points(grid);
	kgi->integral_values = gli_require_host(gli_host_alloc(N * N * sizeof(double)), N * N * sizeof(double),
	                                        "compute_kinetic_gausslet_integrals");
	int64_t g[6];
	gli_grid_array(grid, g);
	gli_check(gli_hip_kinetic_matrix(gdata->coefficients, gdata->indices.num, g, kgi->integral_values), "compute_kinetic_gausslet_integrals");
}

double compute_kinetic_gausslet_integral(const struct gausslet_data* gdata, const union cartesian_grid_point_3d points[2])
{
	int64_t g[6];
	for (int a = 0; a < 3; ++a) {
		const glong mn = lmin(points[0].c[a], points[1].c[a]), mx = lmax(points[0].c[a], points[1].c[a]);
		g[2 * a] = mn;
		g[2 * a + 1] = mx - mn + 1;
	}
	const glong N = g[1] * g[3] * g[5];
	glong lin[2];
	for (int p = 0; p < 2; ++p) {
		lin[p] = ((points[p].x - g[0]) * g[3] + (points[p].y - g[2])) * g[5] + (points[p].z - g[4]);
	}
	double* M = gli_require_host(aligned_malloc(N * N * sizeof(double)), N * N * sizeof(double), "compute_kinetic_gausslet_integral");
	gli_check(gli_hip_kinetic_matrix(gdata->coefficients, gdata->indices.num, g, M), "compute_kinetic_gausslet_integral");
	const double v = M[lin[0] * N + lin[1]];
	aligned_free(M);
	return v;
}

void delete_kinetic_gausslet_integrals(struct kinetic_gausslet_integrals* kgi)
{
	aligned_free(kgi->integral_values);
	kgi->integral_values = NULL;
}
