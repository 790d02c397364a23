/* Sparse ERI sets in the reference's HDF5 layout (libgli_amd_hdf5.so; additive API).
 *
 * File layout (cli/gli_eri_integrals.c:119-156 writes it, pymodule.c:1218-1326 reads it):
 *   attribute  grid_length  int64   points per axis of the centred cubic grid
 *   attribute  tol          double  truncation tolerance of the computation
 *   dataset    gausslet_coeffs  double[K]
 *   dataset    integral_values  double[num]
 *   dataset    four_indices     int64[num]   (sorted, canonical)
 * Return values: 0 on success, -1 on failure (message on stderr). */
#pragma once

#include <stdint.h>
#include "eri_integrals.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Write a sparse set.  With on_device != 0, four_indices and values are HIP device pointers and
 * are streamed to the file in slices of at most 2^25 entries (no host copy of the whole set). */
int gli_eri_save_hdf5(const char* filename, int64_t grid_length, double tol, const double* coeffs, int64_t K,
                      const int64_t* four_indices, const double* values, int64_t num, int on_device);

/* Read a sparse set into *eri (arrays from aligned_malloc; release with
 * delete_sparse_eri_gausslet_integrals) and its coefficients into *coeffs (release with free). */
int gli_eri_load_hdf5(const char* filename, struct sparse_eri_gausslet_integrals* eri, double** coeffs, int64_t* K,
                      double* tol);

#ifdef __cplusplus
}
#endif
