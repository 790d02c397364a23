/* Large host arrays: allocation on 2 MB boundaries with transparent huge pages requested, so that
 * the first write of a multi-GB result array takes a page fault per 2 MB instead of per 4 KB, and
 * copies split over the OpenMP threads with streaming stores.  Every array from gli_host_alloc is released with free(),
 * like the arrays of aligned_malloc. */
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <emmintrin.h>
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

/* One block with non-temporal stores: the destinations are written once and not read again by this
 * thread (pinned transfer buffers read by the DMA engine, result arrays handed to the caller), so
 * streaming stores skip the read-for-ownership of every destination line that cached stores pay.
 * Needs 16-byte aligned destinations; the unaligned head and the tail go through memcpy. */
static void copy_block_nt(char* dst, char* dst2, const char* src, size_t len)
{
	if ((((uintptr_t)dst | (uintptr_t)(dst2 != NULL ? dst2 : dst)) & 15u) != 0) {
		memcpy(dst, src, len);
		if (dst2 != NULL) { memcpy(dst2, src, len); }
		return;
	}
	size_t i = 0;
	for (; i + 64 <= len; i += 64) {
		const __m128i a = _mm_loadu_si128((const __m128i*)(src + i));
		const __m128i b = _mm_loadu_si128((const __m128i*)(src + i + 16));
		const __m128i c = _mm_loadu_si128((const __m128i*)(src + i + 32));
		const __m128i d = _mm_loadu_si128((const __m128i*)(src + i + 48));
		_mm_stream_si128((__m128i*)(dst + i), a);
		_mm_stream_si128((__m128i*)(dst + i + 16), b);
		_mm_stream_si128((__m128i*)(dst + i + 32), c);
		_mm_stream_si128((__m128i*)(dst + i + 48), d);
		if (dst2 != NULL) {
			_mm_stream_si128((__m128i*)(dst2 + i), a);
			_mm_stream_si128((__m128i*)(dst2 + i + 16), b);
			_mm_stream_si128((__m128i*)(dst2 + i + 32), c);
			_mm_stream_si128((__m128i*)(dst2 + i + 48), d);
		}
	}
	if (i < len) {
		memcpy(dst + i, src + i, len - i);
		if (dst2 != NULL) { memcpy(dst2 + i, src + i, len - i); }
	}
	_mm_sfence();   /* the streaming stores are globally visible before the caller hands the buffers on */
}

void gli_host_copy(void* dst, void* dst2, const void* src, size_t bytes)
{
	const int64_t nb = (int64_t)((bytes + GLI_COPY_BLOCK - 1) / GLI_COPY_BLOCK);
#pragma omp parallel for schedule(static) if (nb > 1)
	for (int64_t b = 0; b < nb; ++b) {
		const size_t off = (size_t)b * GLI_COPY_BLOCK;
		const size_t len = (off + GLI_COPY_BLOCK <= bytes) ? GLI_COPY_BLOCK : bytes - off;
		copy_block_nt((char*)dst + off, dst2 != NULL ? (char*)dst2 + off : NULL, (const char*)src + off, len);
	}
}
