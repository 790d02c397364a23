/* Reference-compatible ERI entry points (eri_integrals.h) over the GPU path.
 * Enumeration, values and the dense / applied / projected views of a sparse
 * set run on the GPU; single lookups are host binary searches
 * (eri_integrals.c:685-707). */
#include <stdio.h>
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
	eri->four_indices = gli_require_host(gli_host_alloc(num_indices * sizeof(glong)), num_indices * sizeof(glong),
	                                     "compute_sparse_eri_gausslet_integrals");
	eri->integral_values = gli_require_host(gli_host_alloc(num_indices * sizeof(double)), num_indices * sizeof(double),
	                                        "compute_sparse_eri_gausslet_integrals");
	int64_t g[6];
	gli_grid_array(grid, g);
	/* the index copy is made while the indices stream to the device */
	gli_check(gli_hip_eri_values_copy(gdata->coefficients, gdata->indices.num, g, four_indices, num_indices, tol,
	                                  gli_get_eri_method(), eri->integral_values, eri->four_indices),
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
/* dense views (eri_integrals.c:714-1100): on the GPU, every tuple canonicalised there and looked  */
/* up by binary search in the sorted sparse list                                                    */

void fill_dense_eri_tensor(const struct sparse_eri_gausslet_integrals* eri_sparse, const struct cartesian_grid_3d* grid_dense, double* eri_tensor)
{
	int64_t gs[6], gd[6];
	gli_require_symmetric_cubic(&eri_sparse->grid, "fill_dense_eri_tensor");
	gli_grid_array(&eri_sparse->grid, gs);
	gli_grid_array(grid_dense, gd);
	gli_check(gli_hip_eri_fill_dense(gs, eri_sparse->four_indices, eri_sparse->integral_values, eri_sparse->num_entries, gd, eri_tensor),
	          "fill_dense_eri_tensor");
}

/* ret[(i,k), s] = sum_{j,l} (ij|kl) states[(j,l), s]  (physicist's grouping, eri_integrals.c:974-1100) */
void apply_sparse_eri_tensor(const struct sparse_eri_gausslet_integrals* eri, const double* states, const glong num_states, double* ret)
{
	int64_t gs[6];
	gli_require_symmetric_cubic(&eri->grid, "apply_sparse_eri_tensor");
	gli_grid_array(&eri->grid, gs);
	gli_check(gli_hip_eri_apply(gs, eri->four_indices, eri->integral_values, eri->num_entries, states, num_states, ret),
	          "apply_sparse_eri_tensor");
}

/* eri_proj[p,q,r,s] = sum_{ijkl} b[i,p] b[j,q] b[k,r] b[l,s] (ij|kl)  (eri_integrals.c:714-830):
 * the GPU applies the tensor to the product states x[(j,l),(q,s)] = b[j,q] b[l,s], which gives
 * R[(i,k),(q,s)]; the two remaining modes i, k are contracted here.  The S^2 product states go
 * through the GPU in chunks of at most the "project_chunk_kib" tuning knob (256 MB by default) of
 * states, and as many of results, so the host memory of a call stays bounded by the chunk, not by
 * 2 N^2 S^2 doubles (11 GB at 11^3 with 20 states).  Each state column is applied and contracted
 * on its own, so the chunking does not change any value. */

void project_sparse_eri_gausslet_integrals(const struct sparse_eri_gausslet_integrals* eri, const double* basis, const glong num_states, double* eri_proj)
{
	const glong N = cartesian_grid_3d_num_points(&eri->grid);
	const glong S = num_states, S2 = S * S;
	memset(eri_proj, 0, S2 * S2 * sizeof(double));
	if (S == 0) { return; }
	const size_t chunk_bytes = (size_t)gli_get_tuning("project_chunk_kib") << 10;
	glong cs = (glong)(chunk_bytes / ((size_t)N * N * sizeof(double)));
	cs = cs < 1 ? 1 : (cs > S2 ? S2 : cs);
	const char* what = "project_sparse_eri_gausslet_integrals";
	double* x = gli_require_host(gli_host_alloc(N * N * cs * sizeof(double)), N * N * cs * sizeof(double), what);
	double* R = gli_require_host(gli_host_alloc(N * N * cs * sizeof(double)), N * N * cs * sizeof(double), what);
	double* T = gli_require_host(aligned_calloc(N * S * S2 * sizeof(double)), N * S * S2 * sizeof(double), what);
	for (glong c0 = 0; c0 < S2; c0 += cs) {
		const glong cn = (S2 - c0 < cs) ? S2 - c0 : cs;
		#pragma omp parallel for
		for (glong j = 0; j < N; ++j) {
			for (glong l = 0; l < N; ++l) {
				for (glong c = 0; c < cn; ++c) {
					const glong q = (c0 + c) / S, s = (c0 + c) % S;
					x[(j * N + l) * cn + c] = basis[j * S + q] * basis[l * S + s];
				}
			}
		}
		apply_sparse_eri_tensor(eri, x, cn, R);
		/* T[i, r, qs] = sum_k b[k, r] R[(i, k), qs] */
		#pragma omp parallel for
		for (glong i = 0; i < N; ++i) {
			for (glong k = 0; k < N; ++k) {
				const double* Rik = &R[(i * N + k) * cn];
				for (glong r = 0; r < S; ++r) {
					const double b = basis[k * S + r];
					double* Tir = &T[(i * S + r) * S2 + c0];
					for (glong c = 0; c < cn; ++c) { Tir[c] += b * Rik[c]; }
				}
			}
		}
	}
	/* proj[p, q, r, s] = sum_i b[i, p] T[i, r, qs] */
	for (glong i = 0; i < N; ++i) {
		for (glong p = 0; p < S; ++p) {
			const double b = basis[i * S + p];
			for (glong r = 0; r < S; ++r) {
				for (glong q = 0; q < S; ++q) {
					for (glong s = 0; s < S; ++s) { eri_proj[((p * S + q) * S + r) * S + s] += b * T[(i * S + r) * S2 + q * S + s]; }
				}
			}
		}
	}
	aligned_free(T);
	aligned_free(R);
	aligned_free(x);
}
