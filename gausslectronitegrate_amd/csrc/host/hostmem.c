/* Large host arrays: allocation on 2 MB boundaries with transparent huge pages requested, so that
 * the first write of a multi-GB result array takes a page fault per 2 MB instead of per 4 KB, and
 * copies split over the OpenMP threads.  Every array from gli_host_alloc is released with free(),
 * like the arrays of aligned_malloc. */
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include "aligned_memory.h"
#include "gli_device.h"

#define GLI_HUGE_PAGE ((size_t)2 << 20)
#define GLI_HUGE_MIN ((size_t)4 << 20)
#define GLI_COPY_BLOCK ((size_t)1 << 20)

void* gli_host_alloc(size_t bytes)
{
	if (bytes < GLI_HUGE_MIN) { return aligned_malloc(bytes > 0 ? bytes : 1); }
	const size_t len = (bytes + GLI_HUGE_PAGE - 1) & ~(GLI_HUGE_PAGE - 1);
	void* p = NULL;
	if (posix_memalign(&p, GLI_HUGE_PAGE, len) != 0) { return NULL; }
	(void)madvise(p, len, MADV_HUGEPAGE);   /* advisory: without huge pages the array works the same */
	return p;
}

void gli_host_copy(void* dst, void* dst2, const void* src, size_t bytes)
{
	const int64_t nb = (int64_t)((bytes + GLI_COPY_BLOCK - 1) / GLI_COPY_BLOCK);
#pragma omp parallel for schedule(static) if (nb > 1)
	for (int64_t b = 0; b < nb; ++b) {
		const size_t off = (size_t)b * GLI_COPY_BLOCK;
		const size_t len = (off + GLI_COPY_BLOCK <= bytes) ? GLI_COPY_BLOCK : bytes - off;
		memcpy((char*)dst + off, (const char*)src + off, len);
		if (dst2 != NULL) { memcpy((char*)dst2 + off, (const char*)src + off, len); }
	}
}
