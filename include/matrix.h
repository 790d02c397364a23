#pragma once
/* Integer 3x3 matrices acting on grid points; ABI counterpart of matrix.h:12-31. */

#include <stdbool.h>
#include "grid.h"
#include "util.h"

struct imat3x3
{
	glong data[3][3];
};

void imat3x3_multiply(const struct imat3x3* s, const struct imat3x3* t, struct imat3x3* ret);

bool imat3x3_equal(const struct imat3x3* s, const struct imat3x3* t);

static inline void transform_cartesian_grid_point_3d(const struct imat3x3* trans, const union cartesian_grid_point_3d* pt, union cartesian_grid_point_3d* ret)
{
	for (int r = 0; r < 3; ++r) {
		ret->c[r] = trans->data[r][0] * pt->x + trans->data[r][1] * pt->y + trans->data[r][2] * pt->z;
	}
}
