/* helpers shared by the host C sources (not part of the public API) */
#pragma once
#include <stddef.h>
#include <stdint.h>
#include "grid.h"
#include "gausslet.h"
#include "gli_device.h"

void gli_check(int status, const char* what);
void gli_grid_array(const struct cartesian_grid_3d* grid, int64_t g[6]);
void gli_require_symmetric_cubic(const struct cartesian_grid_3d* grid, const char* what);
void gli_require_coefficients(const struct gausslet_data* gdata, const char* what);
/* p itself, or a message naming 'what' and abort() when the allocation of 'bytes' failed (the
 * reference API has no status channel) */
void* gli_require_host(void* p, size_t bytes, const char* what);

/* canonical sparse key of a tuple of absolute grid points against a symmetric sparse grid
 * (centring, bias placements, O_h orbit minimum); -1 if the tuple does not fit the grid */
glong gli_eri_sparse_key(const struct cartesian_grid_3d* gs, const glong* const perm[48], const union cartesian_grid_point_3d pts[4]);
glong gli_erida_sparse_key(const struct cartesian_grid_3d* gs, const glong* const perm[48], const union cartesian_grid_point_3d pts[2]);
