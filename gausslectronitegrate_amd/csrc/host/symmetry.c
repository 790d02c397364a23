/* O_h group data and grid permutations (symmetry.h; reference symmetry.c:9-150).
 * The matrices and their Cayley table are generated by gen_symmetry.py from
 * the signed-axis-permutation codes of the group elements. */
#include "symmetry.h"
#include "aligned_memory.h"
#include "host_internal.h"
#include "oh_tables.inc"

const struct imat3x3 octahedral_matrep[48] = { GLI_OH_MATRICES };

const int octahedral_multiplication_table[48][48] = { GLI_OH_CAYLEY };

void compute_cartesian_grid_3d_permutation(const struct cartesian_grid_3d* grid, const struct imat3x3* trans, glong* perm)
{
	const glong n = cartesian_grid_3d_num_points(grid);
	for (glong idx = 0; idx < n; ++idx) {
		union cartesian_grid_point_3d p, q;
		linear_index_to_cartesian_grid_point_3d(grid, idx, &p);
		transform_cartesian_grid_point_3d(trans, &p, &q);
		perm[idx] = cartesian_grid_point_3d_to_linear_index(grid, &q);
	}
}

void evaluate_octahedral_grid_permutations(const struct cartesian_grid_3d* grid, glong* perm[48])
{
	const glong n = cartesian_grid_3d_num_points(grid);
	for (int e = 0; e < 48; ++e) {
		perm[e] = gli_require_host(aligned_malloc(n * sizeof(glong)), n * sizeof(glong), "evaluate_octahedral_grid_permutations");
		compute_cartesian_grid_3d_permutation(grid, &octahedral_matrep[e], perm[e]);
	}
}
