#pragma once
/* 64-byte aligned host allocation; every array the library hands back is
 * allocated here and released with free() (aligned_memory.h:7-50). */

#include <stdlib.h>
#include <string.h>

#define MEM_DATA_ALIGN 64

static inline void* aligned_malloc(size_t size)
{
	/* aligned_alloc wants a multiple of the alignment */
	return aligned_alloc(MEM_DATA_ALIGN, (size + MEM_DATA_ALIGN - 1) & ~((size_t)MEM_DATA_ALIGN - 1));
}

static inline void* aligned_calloc(size_t size)
{
	void* p = aligned_malloc(size);
	if (p != NULL) { memset(p, 0, size); }
	return p;
}

static inline void aligned_free(void* memblock)
{
	free(memblock);
}
