#pragma once
/* Cartesian grid points and grids with row-major (x slowest, z fastest) linear
 * indexing; ABI counterpart of grid.h:10-82. */

#include "range.h"

union cartesian_grid_point_3d
{
	struct {
		glong x, y, z;
	};
	glong c[3];
};

struct cartesian_grid_3d
{
	struct range coord_range[3];  /* x, y and z coordinate ranges */
};

static inline glong cartesian_grid_3d_num_points(const struct cartesian_grid_3d* grid)
{
	return grid->coord_range[0].num * grid->coord_range[1].num * grid->coord_range[2].num;
}

/* offsets (ix, iy, iz) from the range starts -> linear index */
static inline glong cartesian_grid_3d_cartesian_to_linear_index(const struct cartesian_grid_3d* grid, const glong ix, const glong iy, const glong iz)
{
	assert(0 <= ix && ix < grid->coord_range[0].num);
	assert(0 <= iy && iy < grid->coord_range[1].num);
	assert(0 <= iz && iz < grid->coord_range[2].num);
	return (ix * grid->coord_range[1].num + iy) * grid->coord_range[2].num + iz;
}

static inline glong cartesian_grid_point_3d_to_linear_index(const struct cartesian_grid_3d* grid, const union cartesian_grid_point_3d* pt)
{
	return cartesian_grid_3d_cartesian_to_linear_index(grid,
		pt->x - grid->coord_range[0].istart,
		pt->y - grid->coord_range[1].istart,
		pt->z - grid->coord_range[2].istart);
}

static inline void linear_index_to_cartesian_grid_point_3d(const struct cartesian_grid_3d* grid, const glong idx, union cartesian_grid_point_3d* pt)
{
	const glong nz = grid->coord_range[2].num;
	const glong ny = grid->coord_range[1].num;
	pt->z = grid->coord_range[2].istart + idx % nz;
	pt->y = grid->coord_range[1].istart + (idx / nz) % ny;
	pt->x = grid->coord_range[0].istart + (idx / nz) / ny;
}
