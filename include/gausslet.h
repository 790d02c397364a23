#pragma once
/* A 1D gausslet as coefficients of elementary Gaussians exp(-(3x - j)^2 / 2)
 * for j in 'indices'; ABI counterpart of gausslet.h:10-14. */

#include "range.h"

struct gausslet_data
{
	double* coefficients;   /* one coefficient per elementary Gaussian */
	struct range indices;   /* Gaussian indices j (K odd, istart = -(K-1)/2) */
};
