#pragma once
/* Electron repulsion integrals in the diagonal approximation (ERIDA), indexed by
 * orbital pairs (i, j). Same entry points and layouts as the reference
 * erida_integrals.h:7-60. */

#include "grid.h"
#include "gausslet.h"

double compute_erida_gausslet_integral(const struct gausslet_data* gdata, const union cartesian_grid_point_3d points[2], const double tol);

glong minimum_octahedral_orbit_erida_tensor_index(const glong num_points, const glong* octahedral_perm[48], const glong i, const glong j);

struct sparse_erida_indices
{
	glong* two_indices;             /* i*N + j, strictly ascending */
	glong num;
	struct cartesian_grid_3d grid;
};

void enumerate_symmetry_reduced_erida_indices(const struct cartesian_grid_3d* grid, struct sparse_erida_indices* ret);

void delete_sparse_erida_indices(struct sparse_erida_indices* indices);

struct sparse_erida_gausslet_integrals
{
	double* integral_values;
	glong* two_indices;
	glong num_entries;
	struct cartesian_grid_3d grid;
};

void compute_sparse_erida_gausslet_integrals(const struct gausslet_data* gdata,
	const struct cartesian_grid_3d* grid, const glong* two_indices, const glong num_indices,
	const double tol, struct sparse_erida_gausslet_integrals* erida);

double sparse_erida_gausslet_integrals_get_value(const struct sparse_erida_gausslet_integrals* erida, const glong index);

void project_sparse_erida_gausslet_integrals(const struct sparse_erida_gausslet_integrals* erida, const double* restrict basis, const glong num_states, double* restrict eri_proj);

void fill_dense_erida_matrix(const struct sparse_erida_gausslet_integrals* erida_sparse, const struct cartesian_grid_3d* grid_dense, double* erida_matrix);

void delete_sparse_erida_gausslet_integrals(struct sparse_erida_gausslet_integrals* erida);
