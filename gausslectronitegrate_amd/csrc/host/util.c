/* Permutation, integer-matrix and timing helpers (util.h:139-144, matrix.h:17-19,
 * timing.h:6-8 of the reference API). */
#include <time.h>
#include "util.h"
#include "matrix.h"
#include "timing.h"
#include "aligned_memory.h"

bool is_permutation(const glong* map, const glong n)
{
	bool* seen = aligned_calloc(n > 0 ? n : 1);
	bool ok = true;
	for (glong i = 0; i < n && ok; ++i) {
		if (map[i] < 0 || map[i] >= n || seen[map[i]]) { ok = false; }
		else { seen[map[i]] = true; }
	}
	aligned_free(seen);
	return ok;
}

bool is_identity_permutation(const glong* perm, const glong n)
{
	for (glong i = 0; i < n; ++i) {
		if (perm[i] != i) { return false; }
	}
	return true;
}

/* ret = p o q, i.e. ret[i] = p[q[i]] */
void compose_permutations(const glong n, const glong* p, const glong* q, glong* ret)
{
	for (glong i = 0; i < n; ++i) { ret[i] = p[q[i]]; }
}

void imat3x3_multiply(const struct imat3x3* s, const struct imat3x3* t, struct imat3x3* ret)
{
	struct imat3x3 r;
	for (int i = 0; i < 3; ++i) {
		for (int k = 0; k < 3; ++k) {
			glong acc = 0;
			for (int j = 0; j < 3; ++j) { acc += s->data[i][j] * t->data[j][k]; }
			r.data[i][k] = acc;
		}
	}
	*ret = r;
}

bool imat3x3_equal(const struct imat3x3* s, const struct imat3x3* t)
{
	for (int i = 0; i < 3; ++i) {
		for (int j = 0; j < 3; ++j) {
			if (s->data[i][j] != t->data[i][j]) { return false; }
		}
	}
	return true;
}

int64_t get_time_ticks(void)
{
	struct timespec t;
	clock_gettime(CLOCK_MONOTONIC, &t);
	return 1000000000LL * t.tv_sec + t.tv_nsec;
}

int64_t get_tick_resolution(void)
{
	return 1000000000LL;
}
