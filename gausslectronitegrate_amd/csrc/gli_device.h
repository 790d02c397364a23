/*
 * Thin C-ABI between the host C library (the reference-compatible API) and
 * the HIP kernels.  Plain C types only; every function returns 0 on success
 * and a negative status on failure (message: gli_last_error()).
 * Host arrays returned through int64_t** are 64-byte aligned and released
 * with free(), like every array of the public API.
 * A grid is passed as g[6] = {istart_x, num_x, istart_y, num_y, istart_z, num_z}.
 */
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ERI value methods */
#define GLI_ERI_TENSOR_FP64  0  /* key-tensor contraction, FP64 throughout (default) */
#define GLI_ERI_TENSOR_MIXED 1  /* key-tensor contraction, split-precision MFMA bulk, FP64 final sums */
#define GLI_ERI_DIRECT_FP64  2  /* per-element FP64 sum, no cross-element reuse */
#define GLI_ERI_TENSOR_BF16X2 3 /* comparison only: BF16 hi/lo two-term split, FP32 accumulation (~1e-6) */
#define GLI_ERI_TENSOR_FP32  4  /* comparison only: FP32 MFMA products and accumulation (~1e-7) */
#define GLI_ERI_TENSOR_FP32X2 5 /* comparison only: FP32 hi/lo split operands, FP32 accumulation (~1e-7) */

int gli_hip_eri_enumerate(int64_t n, int64_t** out, int64_t* num);
int gli_hip_erida_enumerate(int64_t n, int64_t** out, int64_t* num);

int gli_hip_eri_values(const double* coeffs, int64_t K, const int64_t g[6],
                       const int64_t* indices, int64_t num, double tol, int method, double* values);
/* the same, and a host copy of the indices into indices_copy (fused with their upload) */
int gli_hip_eri_values_copy(const double* coeffs, int64_t K, const int64_t g[6], const int64_t* indices, int64_t num,
                            double tol, int method, double* values, int64_t* indices_copy);

/* host arrays of any size, released with free(); large ones on 2 MB pages (hostmem.c) */
void* gli_host_alloc(size_t bytes);
/* dst[0, bytes) = src, and dst2 too unless NULL, split over the OpenMP threads */
void gli_host_copy(void* dst, void* dst2, const void* src, size_t bytes);
int gli_hip_erida_values(const double* coeffs, int64_t K, const int64_t g[6],
                         const int64_t* indices, int64_t num, double tol, double* values);

int gli_hip_nuclear_matrix(const double* coeffs, int64_t K, const int64_t g[6],
                           const double* pos, const int* charges, int num_nuclei, double tol, double* out);
int gli_hip_kinetic_matrix(const double* coeffs, int64_t K, const int64_t g[6], double* out);

/* dense reconstruction and contractions of a sparse ERI / ERIDA set */
int gli_hip_eri_fill_dense(const int64_t gs[6], const int64_t* keys, const double* vals, int64_t num,
                           const int64_t gd[6], double* out);
/* ret[(i N + k) S + s] = sum_{j,l} (ij|kl) states[(j N + l) S + s] over the sparse grid */
int gli_hip_eri_apply(const int64_t gs[6], const int64_t* keys, const double* vals, int64_t num, const double* states,
                      int64_t S, double* ret);

#ifdef __cplusplus
}
#endif
