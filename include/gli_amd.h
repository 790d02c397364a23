#pragma once
/*
 * Additive API of the MI355X build.  None of these entry points exists in the
 * reference; they select the evaluation method, report errors of the GPU path
 * and expose device-resident variants for callers that keep their index lists
 * and values in HBM.  The reference entry points keep their behaviour: they
 * use FP64 by default and abort with a message when the GPU path fails.
 *
 * Device pointers below are HIP device memory of the current device; the
 * library works on one internal stream at default priority (gli_stream()).
 */
#include <stdint.h>
#include "grid.h"
#include "gausslet.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ERI value methods */
#define GLI_ERI_TENSOR_FP64  0  /* key-tensor contraction in FP64 (default, exact re-association) */
#define GLI_ERI_TENSOR_MIXED 1  /* key-tensor contraction, split-precision MFMA bulk, FP64 final sums */
#define GLI_ERI_DIRECT_FP64  2  /* per-element FP64 evaluation, no reuse across elements */
#define GLI_ERI_TENSOR_BF16X2 3 /* comparison only: BF16 hi/lo two-term split, FP32 accumulation (~1e-6) */
#define GLI_ERI_TENSOR_FP32  4  /* comparison only: FP32 MFMA products and accumulation (~1e-7) */
#define GLI_ERI_TENSOR_FP32X2 5 /* comparison only: FP32 hi/lo split operands (three FP32 MFMA products), FP32 accumulation */
#define GLI_ERI_SEPARABLE_FP64 6 /* key tensor from a Gauss-Legendre expansion of the Coulomb kernel in separable
                                  * Gaussians: one rank-Q FP64 MFMA product instead of three mode products
                                  * (kernel reproduced to ~1e-15 relative; values within ~1e-14 of the default) */

/* method used by compute_sparse_eri_gausslet_integrals / compute_eri_gausslet_integral;
 * the environment variable GLI_ERI_METHOD (tensor | mixed | direct | bf16x2 | fp32 | fp32x2 | separable) sets the
 * initial value */
int gli_set_eri_method(int method);
int gli_get_eri_method(void);

/* error reporting of the GPU path: status of the last failing call on this thread */
const char* gli_last_error(void);
int gli_last_status(void);
void gli_clear_error(void);

/* the library stream (hipStream_t) and a full synchronisation of it */
void* gli_stream(void);
int gli_synchronize(void);

/* Kernel tuning knobs (set per workload by an autotuner; the defaults are tuned for the 11^3 grid):
 *   "tile_order"  output-tile order of the mode products: 0 row tiles fastest (default), 1 XCD-grouped
 *   "gemm_shape"  FP64 mode-product tile (rows x columns x mu slice): 0 64x64x16 (default), 1 128x64x16,
 *                 2 64x128x16, 3 128x128x16, 4 64x64x32, 5 128x64x32 (a shape
 *                 that does not divide the key padding of a triangular product falls back to 0)
 *   "gather_ept"  elements per lane in the per-element gather: 2, 4, 8 (default) or 16
 *   "gather_lds"  1 (default): each gather workgroup copies the key tables into LDS first; 0: global reads
 *   "gather_threads" workgroup size of the single-grid gather: 256, 512 or 1024 (default; its 16 waves take
 *                 8192 consecutive elements on one CU, whose L1 then serves more of their key-tensor reads)
 *   "mode1_half"  1 (default): the FP64 paths compute the first mode product's (m2, m3)-symmetric output
 *                 once (m3 >= m2) and the second reads it mirrored; 0: full output
 *   "batch_branches" batch plans of two or more grids with FP64 products: 1 (default) the autotuner also
 *                 times a graph whose largest grid runs beside the others on a second stream and keeps
 *                 the faster; 0 never; 2 always
 *   "batch_positions" 1 (default): a batch plan computes every element's position in its key tensor once,
 *                 at creation, and each replay reads those 4-byte positions instead of decoding the
 *                 8-byte four-indices; 0: every replay decodes
 *   "batch_sorted" 1 (default): with batch_positions, the plan also sorts the positions inside windows of
 *                 8192 consecutive elements (keeping each one's place), so every replay reads the key tensor
 *                 in ascending order within a window and stores the values back in list order through LDS;
 *                 0: positions in list order
 *   "gather_order" 0: workgroups take consecutive pieces of the list; 1: each XCD takes one contiguous
 *                 chunk of the list (its L2 then holds the key-tensor entries it reuses); 2 (default): 1
 *                 while the key tensor is at most 128 MB (11^3: 71 MB), else 0 (13^3: 299 MB)
 *   "gather_bucketed" the single-grid gather decodes a window of 8 x gather_threads elements, counting-sorts
 *                 the window's key-tensor positions by bucket in LDS and reads the key tensor in that order
 *                 (the elements of one cache line share a load), storing the values back in list order:
 *                 0 (default) never, 1 always, 2 when the key tensor exceeds 128 MB (13^3: 299 MB)
 *   "gather_classed" 1: where the bucketed form is not chosen, each wave of the single-grid gather
 *                 (8 groups of 64 elements) issues its 8 key-tensor loads class by class, class = the
 *                 element's 128-byte line mod 8, so the elements of one line share a load; 0 (default):
 *                 list order
 *   "rankq_tet" 1 (default): the separable method's key tensor comes from the tetrahedral rank-Q kernel
 *                 (16 x 16 key blocks per wave, only blocks that meet k1 <= k2 <= k3); 0: from the 64 x 64
 *                 tile kernel; both give the same bits
 *   "project_chunk_kib" host memory (KiB, default 262144 = 256 MB) of one chunk of product states in
 *                 project_sparse_eri_gausslet_integrals; the states and as many results are held per chunk
 * gli_set_tuning returns 0, or -1 for an unknown knob or value; gli_get_tuning returns the value or -1. */
int gli_set_tuning(const char* knob, int value);
int gli_get_tuning(const char* knob);

/* Phase timing of ERI evaluations with HIP events on the library stream.  While enabled, every
 * evaluation records its phase boundaries (no synchronisation between evaluations);
 * gli_phase_times returns, per phase name, the average duration in ms over the evaluations since
 * timing was enabled or last queried, and starts a new average.  The key-tensor method records
 * "tables", "kernel_tensor", "mode1", "mode2", "mode3" and the gather as "gather" (k_gather),
 * "gather_classed" (line classes per wave) or "gather_bucketed" (the window-bucketed form); the
 * per-element method records "direct". */
int gli_enable_phase_timing(int on);
int gli_phase_times(double* ms, const char** names, int max);

/* FP64 flops of the per-element path (GLI_ERI_DIRECT_FP64).  While enabled, each evaluation adds per
 * element, on the device:
 *   reference_form  2 n_x n_y + 2 R_xy (n_z + 1), the terms the reference's loops execute
 *                   (eri_integrals.c:624-669: n_a entries with |P| > tol per axis, R_xy nonzero bins);
 *   executed        the same count on the folded columns this path sums over (every pair of the x and
 *                   y windows, and per nonzero bin the z window + 1).
 * gli_direct_flops returns the sums since counting was enabled or last queried (synchronises). */
int gli_enable_direct_flop_count(int on);
int gli_direct_flops(double* reference_form, double* executed);

/* device memory helpers; kind: 0 host->device, 1 device->host, 2 device->device */
int gli_device_malloc(void** p, int64_t bytes);
void gli_device_free(void* p);
int gli_memcpy(void* dst, const void* src, int64_t bytes, int kind);
/* every byte of a device range set to 'value' (library stream, synchronous) */
int gli_memset(void* dst, int value, int64_t bytes);

/* device-resident enumeration of a symmetric n^3 grid; *d_out is released with gli_device_free */
int gli_eri_enumerate_device(int64_t n, int64_t** d_out, int64_t* num);
int gli_erida_enumerate_device(int64_t n, int64_t** d_out, int64_t* num);

/* device-resident values; grid = {istart_x, num_x, istart_y, num_y, istart_z, num_z};
 * coeffs is host memory (K values); indices outside the grid produce NaN */
int gli_eri_values_device(const double* coeffs, int64_t K, const int64_t grid[6], const int64_t* d_indices,
                          int64_t num, double tol, int method, double* d_values);
int gli_erida_values_device(const double* coeffs, int64_t K, const int64_t grid[6], const int64_t* d_indices,
                            int64_t num, double tol, double* d_values);

/* gli_eri_values_device without device allocation.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  It allocates nothing on the device: the whole workspace comes from one static
 * 96 MB scratch array per GPU, held from library load to process end, which serves grids up to 7^3 with
 * K <= 57 (larger ones return -1 with the size they need).  Method GLI_ERI_TENSOR_FP64 only; the values
 * are bit-identical to gli_eri_values_device's.  Synchronises the library stream once (after copying
 * the coefficients and small host tables in), then queues the kernels on it.  Every call on one GPU
 * shares that GPU's scratch: calls that overlap (from several host threads without the library's
 * serialisation, or with a previous call's kernels still reading it from another stream) race. */
int gli_eri_values_device_noalloc(const double* coeffs, int64_t K, const int64_t grid[6], const int64_t* d_indices,
                                  int64_t num, double tol, int method, double* d_values);

/* Many single integrals in one GPU evaluation: the batched forms of compute_eri_gausslet_integral
 * (eri_integrals.c:133-232) and compute_erida_gausslet_integral (erida_integrals.c:119-201).
 * points[(e * 4 + p) * 3 + a] is axis a of point p (i, j, k, l) of element e (two points i, j for
 * ERIDA); values[e] receives (ij|kl) resp. (ii|jj); host memory, synchronous.  The ERI method is
 * one of GLI_ERI_*.  Each tuple is moved to its own minimum corner (the values depend on point
 * differences only) and all tuples share the smallest grid that holds them. */
int gli_eri_integrals_points(const double* coeffs, int64_t K, const int64_t* points, int64_t num, double tol, int method,
                             double* values);
int gli_erida_integrals_points(const double* coeffs, int64_t K, const int64_t* points, int64_t num, double tol,
                               double* values);

/* enumerate_symmetry_reduced_eri_indices (eri_integrals.c:382-526) followed by
 * compute_sparse_eri_gausslet_integrals (eri_integrals.c:544-678) on the same symmetric grid, with the
 * index list kept on the GPU in between: one download of indices and values instead of a download, an
 * upload and a host copy.  Fills *eri exactly like the two calls (release with
 * delete_sparse_eri_gausslet_integrals); the method is gli_get_eri_method().  Returns 0, or -1 with
 * gli_last_error() set (for a grid that is not cubic, centred and odd, or coefficients that are not). */
struct sparse_eri_gausslet_integrals;
int gli_sparse_eri_compute(const struct gausslet_data* gdata, const struct cartesian_grid_3d* grid, double tol,
                           struct sparse_eri_gausslet_integrals* eri);

/* The one-body matrices of compute_kinetic_gausslet_integrals (kinetic_integrals.c:44-121) and
 * compute_nuclear_gausslet_integrals (nuclear_integrals.c:133-266) written into a caller's host array
 * out[N * N] (row-major, N = the grid's point count) instead of a result struct; for callers that own
 * their output buffer.  positions[3 * n + a] is axis a of nucleus n, charges[n] its charge; synchronous. */
int gli_kinetic_matrix(const double* coeffs, int64_t K, const int64_t grid[6], double* out);
int gli_nuclear_matrix(const double* coeffs, int64_t K, const int64_t grid[6], const double* positions, const int* charges,
                       int num_nuclei, double tol, double* out);

/* several grids in one call (for example 5^3, 7^3, 9^3 and 11^3 with different K): a batch plan
 * without autotuning, run once; synchronous */
int gli_eri_values_batch_device(int num_grids, const double* const* coeffs, const int64_t* K, const int64_t* grids,
                                const int64_t* const* d_indices, const int64_t* num, double tol, int method,
                                double* const* d_values);

/* Batch plan for repeated evaluation of up to 16 grids (grids[6*g .. 6*g+5], K[g] coefficients,
 * d_indices[g] / num[g] device-resident four-indices, d_values[g] device outputs; method
 * GLI_ERI_TENSOR_FP64 or GLI_ERI_TENSOR_MIXED).  The plan owns every intermediate buffer.  With
 * autotune != 0 it times the candidate kernels on its own shapes (mode-product tile shape and
 * order, FP64 or split-precision products under the mixed method, gather variant) and keeps the
 * fastest; then it captures the whole evaluation (parameter upload, tables, three grouped mode
 * products, one grouped gather over all grids) into one hipGraph.
 * gli_eri_batch_run replays it with new coefficients (NULL: those of the previous run, which must
 * keep each grid's mirror symmetry) and per-grid output scales (NULL: 1; e.g. 1/a for a grid
 * spacing a); it is asynchronous on gli_stream().  gli_eri_batch_describe writes the autotuner's
 * timings and choices as JSON (returns the full length).  The index and value arrays must stay
 * allocated, and the index lists unchanged, until the plan is destroyed. */
typedef struct gli_eri_batch gli_eri_batch;
int gli_eri_batch_create(int num_grids, const double* const* coeffs, const int64_t* K, const int64_t* grids,
                         const int64_t* const* d_indices, const int64_t* num, double tol, int method,
                         double* const* d_values, int autotune, gli_eri_batch** plan);
int gli_eri_batch_run(gli_eri_batch* plan, const double* const* coeffs, const double* scale);
int gli_eri_batch_describe(const gli_eri_batch* plan, char* buf, int len);
void gli_eri_batch_destroy(gli_eri_batch* plan);

/* Y[p][k] = sum_{mu < Kd} X[mu*ldx + p] * B[mu*ldb + k] (FP64 MFMA; device pointers, synchronous) */
int gli_mode_product_device(const double* X, int64_t ldx, int64_t P, const double* B, int64_t ldb, int64_t C, int64_t Kd,
                            double* Y, int64_t ldy);

/* the same contraction from BF16 digit slices with exact FP32 slice sums and an FP64 combination
 * (needs Kd <= 256; error below 2^-40 of the column scales) */
int gli_mode_product_split_device(const double* X, int64_t ldx, int64_t P, const double* B, int64_t ldb, int64_t C,
                                  int64_t Kd, double* Y, int64_t ldy);

/* number of distinct folded axis columns of an R-point axis (symmetric: mirror-symmetric coefficients) */
int gli_eri_key_count(int64_t R, int symmetric);

/* Quadrature rule of the separable method (GLI_ERI_SEPARABLE_FP64) for a folded range M = 6(R-1) + 2(K-1):
 * nodes s_q in (0, 1) and weights w_q with 6 erf(sqrt(i/8)) / sqrt(i) ~ (6 / sqrt(2 pi)) sum_q w_q exp(-i s_q^2 / 8)
 * for 0 <= i <= 3 M^2.  Writes at most max entries (nothing when an array is null) and returns the node
 * count Q, or -1 for M < 0.  Host arithmetic only. */
int gli_eri_separable_rule(int M, double* nodes, double* weights, int max);

/* release cached device workspace */
void gli_release_workspace(void);

/* Canonical sparse four-index (two-index for ERIDA) of grid points pts on the symmetric grid gs,
 * with the O_h permutations of evaluate_octahedral_grid_permutations: the box is centred and the
 * minimum over bias placements, O_h and label swaps is taken (eri_integrals.c:240-291); -1 when a
 * centred point leaves the grid.  The Python surface uses them for single-element lookups. */
glong gli_eri_sparse_key(const struct cartesian_grid_3d* gs, const glong* const perm[48], const union cartesian_grid_point_3d pts[4]);
glong gli_erida_sparse_key(const struct cartesian_grid_3d* gs, const glong* const perm[48], const union cartesian_grid_point_3d pts[2]);

#ifdef __cplusplus
}
#endif
