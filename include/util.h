#pragma once
/* Core scalar helpers and the 64-bit index type of the public C API.
 * ABI counterpart: util.h:10-144 of the reference (same type, same inline semantics). */

#include <stdint.h>
#include <stdbool.h>
#include <math.h>
#include <assert.h>

/* signed 64-bit index type used for every count and linear index */
typedef int64_t glong;

#define GLONG_MAX INT64_MAX

static inline double square(const double x) { return x * x; }

static inline glong lsquare(const glong x) { return x * x; }

static inline double cubic_power(const double x) { return x * x * x; }

static inline glong lmin(const glong a, const glong b) { return (a <= b) ? a : b; }

static inline glong lmax(const glong a, const glong b) { return (a >= b) ? a : b; }

/* Euclidean length of a 3-vector, scaled by its largest component to avoid
 * overflow; returns exactly 0 for the zero vector (util.h:69-117). */
static inline double vec3_norm(const double v[3])
{
	const double a = fabs(v[0]), b = fabs(v[1]), c = fabs(v[2]);
	/* the reference picks the largest magnitude with ties resolved in favour of z, then y */
	double big, s0, s1;
	if (b <= c) {
		if (a <= c) { if (c == 0) { return 0; } big = c; s0 = a; s1 = b; }
		else        { big = a; s0 = b; s1 = c; }
	}
	else {
		if (a <= b) { big = b; s0 = a; s1 = c; }
		else        { big = a; s0 = b; s1 = c; }
	}
	const double r0 = s0 / big, r1 = s1 / big;
	return big * sqrt(1 + r0*r0 + r1*r1);
}

/* infinity-norm distance between two length-n vectors */
static inline double uniform_distance(const glong n, const double* x, const double* y)
{
	double d = 0;
	for (glong i = 0; i < n; ++i) { d = fmax(d, fabs(x[i] - y[i])); }
	return d;
}

bool is_permutation(const glong* map, const glong n);

bool is_identity_permutation(const glong* perm, const glong n);

void compose_permutations(const glong n, const glong* p, const glong* q, glong* ret);
