#pragma once
/* Kinetic energy matrix of gausslet orbitals; same entry points and layouts
 * as the reference kinetic_integrals.h:7-24. */

#include "grid.h"
#include "gausslet.h"

double compute_kinetic_gausslet_integral(const struct gausslet_data* gdata, const union cartesian_grid_point_3d points[2]);

struct kinetic_gausslet_integrals
{
	double* integral_values;        /* N x N, row-major */
	struct cartesian_grid_3d grid;
};

void compute_kinetic_gausslet_integrals(const struct gausslet_data* gdata, const struct cartesian_grid_3d* grid, struct kinetic_gausslet_integrals* kgi);

void delete_kinetic_gausslet_integrals(struct kinetic_gausslet_integrals* kgi);
