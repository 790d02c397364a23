#pragma once
/* Electron repulsion integrals (ij|kl) of gausslet orbitals on a cubic lattice.
 * Same entry points, argument meaning and struct layouts as the reference
 * eri_integrals.h:7-64; every returned array is 64-byte aligned host memory
 * released by the matching delete_* call. */

#include "grid.h"
#include "gausslet.h"

/* one integral for four grid points (eri_integrals.h:7) */
double compute_eri_gausslet_integral(const struct gausslet_data* gdata, const union cartesian_grid_point_3d points[4], const double tol);

/* canonical (smallest) linear four-index over the O_h orbit combined with the
 * i<->j, k<->l, (ij)<->(kl) label swaps (eri_integrals.h:14-16) */
glong minimum_octahedral_orbit_eri_tensor_index(
	const glong num_points, const glong* octahedral_perm[48],
	const glong i, const glong j, const glong k, const glong l);

struct sparse_eri_indices
{
	glong* four_indices;            /* ((i*N + j)*N + k)*N + l, strictly ascending */
	glong num;                      /* number of entries */
	struct cartesian_grid_3d grid;  /* grid the indices refer to */
};

/* symmetry-unique four-indices of a grid centred at the origin (eri_integrals.h:31) */
void enumerate_symmetry_reduced_eri_indices(const struct cartesian_grid_3d* grid, struct sparse_eri_indices* ret);

void delete_sparse_eri_indices(struct sparse_eri_indices* indices);

struct sparse_eri_gausslet_integrals
{
	double* integral_values;        /* one value per four-index */
	glong* four_indices;            /* copy of the requested four-indices (sorted for lookups) */
	glong num_entries;              /* number of entries */
	struct cartesian_grid_3d grid;  /* grid the indices refer to */
};

/* values for an arbitrary list of four-indices (eri_integrals.h:50-52) */
void compute_sparse_eri_gausslet_integrals(const struct gausslet_data* gdata,
	const struct cartesian_grid_3d* grid, const glong* four_indices, const glong num_indices,
	const double tol, struct sparse_eri_gausslet_integrals* eri);

/* value stored for 'index', 0 if absent (binary search; eri_integrals.h:55) */
double sparse_eri_gausslet_integrals_get_value(const struct sparse_eri_gausslet_integrals* eri, const glong index);

/* eri_proj[pqrs] = sum_{ijkl} basis[i,p] basis[j,q] basis[k,r] basis[l,s] (ij|kl) (eri_integrals.h:57) */
void project_sparse_eri_gausslet_integrals(const struct sparse_eri_gausslet_integrals* eri, const double* basis, const glong num_states, double* eri_proj);

/* dense (ij|kl) on 'grid_dense' reconstructed from the sparse set (eri_integrals.h:59) */
void fill_dense_eri_tensor(const struct sparse_eri_gausslet_integrals* eri_sparse, const struct cartesian_grid_3d* grid_dense, double* eri_tensor);

/* ret[(i,k), s] = sum_{j,l} (ij|kl) states[(j,l), s] (physicist's grouping; eri_integrals.h:61) */
void apply_sparse_eri_tensor(const struct sparse_eri_gausslet_integrals* eri, const double* states, const glong num_states, double* ret);

void delete_sparse_eri_gausslet_integrals(struct sparse_eri_gausslet_integrals* eri);
