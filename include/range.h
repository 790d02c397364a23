#pragma once
/* Integer range {first value, count}; ABI counterpart of range.h:10-28. */

#include "util.h"

struct range
{
	glong istart;  /* first integer of the range */
	glong num;     /* number of consecutive integers */
};

/* range of all differences r_i - s_j */
static inline struct range range_minus(const struct range r, const struct range s)
{
	struct range d;
	d.istart = r.istart - (s.istart + s.num - 1);
	d.num    = r.num + s.num - 1;
	return d;
}
