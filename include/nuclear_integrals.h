#pragma once
/* Nuclear attraction matrix of gausslet orbitals; same entry points and
 * layouts as the reference nuclear_integrals.h:11-46. */

#include "grid.h"
#include "gausslet.h"

struct atomic_nucleus
{
	double pos[3];  /* position in grid units */
	int charge;     /* nuclear charge */
};

double compute_nuclear_gausslet_integral(
	const struct gausslet_data* gdata, const union cartesian_grid_point_3d points[2],
	const struct atomic_nucleus* nuclei, const int num_nuclei, const double tol);

struct nuclear_gausslet_integrals
{
	double* integral_values;        /* N x N, row-major */
	struct cartesian_grid_3d grid;
	struct atomic_nucleus* nuclei;  /* copy of the nuclei */
	int num_nuclei;
};

void compute_nuclear_gausslet_integrals(
	const struct gausslet_data* gdata, const struct cartesian_grid_3d* grid,
	const struct atomic_nucleus* nuclei, const int num_nuclei,
	const double tol, struct nuclear_gausslet_integrals* ngi);

void delete_nuclear_gausslet_integrals(struct nuclear_gausslet_integrals* ngi);
