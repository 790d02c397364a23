/* Method selection and the fail-loudly policy of the reference-compatible entry points. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "gli_amd.h"
#include "host_internal.h"

static int g_method = -1;

static int method_from_env(void)
{
	const char* s = getenv("GLI_ERI_METHOD");
	if (s == NULL || s[0] == 0 || strcmp(s, "tensor") == 0) { return GLI_ERI_TENSOR_FP64; }
	if (strcmp(s, "mixed") == 0) { return GLI_ERI_TENSOR_MIXED; }
	if (strcmp(s, "direct") == 0) { return GLI_ERI_DIRECT_FP64; }
	if (strcmp(s, "bf16x2") == 0) { return GLI_ERI_TENSOR_BF16X2; }
	if (strcmp(s, "fp32") == 0) { return GLI_ERI_TENSOR_FP32; }
	if (strcmp(s, "fp32x2") == 0) { return GLI_ERI_TENSOR_FP32X2; }
	if (strcmp(s, "separable") == 0) { return GLI_ERI_SEPARABLE_FP64; }
	fprintf(stderr, "gausslectronitegrate: unknown GLI_ERI_METHOD '%s' (tensor | mixed | direct | bf16x2 | fp32 | fp32x2 | separable)\n", s);
	abort();
}

int gli_set_eri_method(int method)
{
	if (method < GLI_ERI_TENSOR_FP64 || method > GLI_ERI_SEPARABLE_FP64) { return -1; }
	g_method = method;
	return 0;
}

int gli_get_eri_method(void)
{
	if (g_method < 0) { g_method = method_from_env(); }
	return g_method;
}

void gli_check(int status, const char* what)
{
	if (status != 0) {
		fprintf(stderr, "gausslectronitegrate: %s failed on the GPU path: %s\n", what, gli_last_error());
		fflush(stderr);
		abort();
	}
}

void gli_grid_array(const struct cartesian_grid_3d* grid, int64_t g[6])
{
	for (int a = 0; a < 3; ++a) {
		g[2 * a] = grid->coord_range[a].istart;
		g[2 * a + 1] = grid->coord_range[a].num;
	}
}

void gli_require_symmetric_cubic(const struct cartesian_grid_3d* grid, const char* what)
{
	const glong n = grid->coord_range[0].num;
	for (int a = 0; a < 3; ++a) {
		if (grid->coord_range[a].num != n || n % 2 != 1 || grid->coord_range[a].istart != -(n - 1) / 2) {
			fprintf(stderr, "gausslectronitegrate: %s needs a cubic grid centred at the origin with an odd number of points per axis\n", what);
			abort();
		}
	}
}

void gli_require_coefficients(const struct gausslet_data* gdata, const char* what)
{
	if (gdata->indices.num < 1 || gdata->indices.num % 2 != 1 || gdata->indices.istart != -(gdata->indices.num - 1) / 2) {
		fprintf(stderr, "gausslectronitegrate: %s needs an odd number of coefficients with indices -(K-1)/2 .. (K-1)/2\n", what);
		abort();
	}
}

void* gli_require_host(void* p, size_t bytes, const char* what)
{
	if (p == NULL) {
		fprintf(stderr, "gausslectronitegrate: %s: host allocation of %zu bytes failed\n", what, bytes);
		abort();
	}
	return p;
}
