#pragma once
/* Octahedral point group O_h as 48 signed permutation matrices and the grid
 * permutations they induce; ABI counterpart of symmetry.h:13-24. */

#include "grid.h"
#include "matrix.h"
#include "util.h"

/* element order is part of the ABI (same order as symmetry.c:9-58) */
extern const struct imat3x3 octahedral_matrep[48];

/* octahedral_multiplication_table[i][j] = index of matrep[i] * matrep[j] */
extern const int octahedral_multiplication_table[48][48];

/* perm[idx] = linear index of trans * point(idx); the grid must be closed under 'trans' */
void compute_cartesian_grid_3d_permutation(const struct cartesian_grid_3d* grid, const struct imat3x3* trans, glong* perm);

/* allocates perm[0..47] (N entries each, release with aligned_free) */
void evaluate_octahedral_grid_permutations(const struct cartesian_grid_3d* grid, glong* perm[48]);
