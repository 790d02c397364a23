/* HDF5 helpers (hdf5_util.h) and the sparse ERI file layout of the reference (gli_hdf5.h). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "hdf5_util.h"
#include "gli_hdf5.h"
#include "gli_amd.h"
#include "aligned_memory.h"

static herr_t fail(const char* what, const char* name)
{
	fprintf(stderr, "gausslectronitegrate: %s failed for '%s'\n", what, name);
	return -1;
}

/* shape of a dataset: rank (and the extents when dims != NULL) */
static herr_t dataset_shape(hid_t file, const char* name, int* ndims, hsize_t* dims)
{
	const hid_t dset = H5Dopen(file, name, H5P_DEFAULT);
	if (dset < 0) { return fail("H5Dopen", name); }
	const hid_t space = H5Dget_space(dset);
	if (space < 0) {
		H5Dclose(dset);
		return fail("H5Dget_space", name);
	}
	herr_t st = 0;
	const int r = H5Sget_simple_extent_ndims(space);
	if (r < 0) { st = fail("H5Sget_simple_extent_ndims", name); }
	else {
		if (ndims) { *ndims = r; }
		if (dims && H5Sget_simple_extent_dims(space, dims, NULL) < 0) { st = fail("H5Sget_simple_extent_dims", name); }
	}
	H5Sclose(space);
	H5Dclose(dset);
	return st;
}

herr_t get_hdf5_dataset_ndims(hid_t file, const char* name, int* ndims)
{
	return dataset_shape(file, name, ndims, NULL);
}

herr_t get_hdf5_dataset_dims(hid_t file, const char* name, hsize_t* dims)
{
	return dataset_shape(file, name, NULL, dims);
}

herr_t read_hdf5_dataset(hid_t file, const char* name, hid_t mem_type, void* data)
{
	const hid_t dset = H5Dopen(file, name, H5P_DEFAULT);
	if (dset < 0) { return fail("H5Dopen", name); }
	const herr_t st = H5Dread(dset, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
	H5Dclose(dset);
	return st < 0 ? fail("H5Dread", name) : 0;
}

herr_t write_hdf5_dataset(hid_t file, const char* name, int degree, const hsize_t dims[], hid_t mem_type_store, hid_t mem_type_input, const void* data)
{
	const hid_t space = H5Screate_simple(degree, dims, NULL);
	if (space < 0) { return fail("H5Screate_simple", name); }
	const hid_t dset = H5Dcreate(file, name, mem_type_store, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
	if (dset < 0) {
		H5Sclose(space);
		return fail("H5Dcreate", name);
	}
	const herr_t st = H5Dwrite(dset, mem_type_input, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
	H5Dclose(dset);
	H5Sclose(space);
	return st < 0 ? fail("H5Dwrite", name) : 0;
}

herr_t get_hdf5_attribute_dims(hid_t file, const char* name, hsize_t* dims)
{
	const hid_t attr = H5Aopen(file, name, H5P_DEFAULT);
	if (attr < 0) { return fail("H5Aopen", name); }
	const hid_t space = H5Aget_space(attr);
	herr_t st = 0;
	if (space < 0) { st = fail("H5Aget_space", name); }
	else {
		if (H5Sget_simple_extent_dims(space, dims, NULL) < 0) { st = fail("H5Sget_simple_extent_dims", name); }
		H5Sclose(space);
	}
	H5Aclose(attr);
	return st;
}

herr_t read_hdf5_attribute(hid_t file, const char* name, hid_t mem_type, void* data)
{
	const hid_t attr = H5Aopen(file, name, H5P_DEFAULT);
	if (attr < 0) { return fail("H5Aopen", name); }
	const herr_t st = H5Aread(attr, mem_type, data);
	H5Aclose(attr);
	return st < 0 ? fail("H5Aread", name) : 0;
}

static herr_t write_attribute(hid_t file, const char* name, hid_t space, hid_t mem_type_store, hid_t mem_type_input, const void* data)
{
	if (space < 0) { return fail("H5Screate", name); }
	const hid_t attr = H5Acreate(file, name, mem_type_store, space, H5P_DEFAULT, H5P_DEFAULT);
	if (attr < 0) {
		H5Sclose(space);
		return fail("H5Acreate", name);
	}
	const herr_t st = H5Awrite(attr, mem_type_input, data);
	H5Aclose(attr);
	H5Sclose(space);
	return st < 0 ? fail("H5Awrite", name) : 0;
}

herr_t write_hdf5_scalar_attribute(hid_t file, const char* name, hid_t mem_type_store, hid_t mem_type_input, const void* data)
{
	return write_attribute(file, name, H5Screate(H5S_SCALAR), mem_type_store, mem_type_input, data);
}

herr_t write_hdf5_vector_attribute(hid_t file, const char* name, hid_t mem_type_store, hid_t mem_type_input, const glong length, const void* data)
{
	const hsize_t dims[1] = { (hsize_t)length };
	return write_attribute(file, name, H5Screate_simple(1, dims, NULL), mem_type_store, mem_type_input, data);
}

/* ------------------------------------------------------------------------- */
/* sparse ERI files                                                           */

/* one 1D dataset written slice by slice; with on_device the slices are staged through host memory */
static int write_streamed(hid_t file, const char* name, hid_t store, hid_t input, size_t elem, const void* data, int64_t num,
                          int on_device)
{
	const hsize_t dims[1] = { (hsize_t)num };
	const hid_t space = H5Screate_simple(1, dims, NULL);
	if (space < 0) { return fail("H5Screate_simple", name); }
	const hid_t dset = H5Dcreate(file, name, store, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
	if (dset < 0) {
		H5Sclose(space);
		return fail("H5Dcreate", name);
	}
	const int64_t slice = (int64_t)1 << 25;
	void* stage = NULL;
	int st = 0;
	if (on_device && num > 0) {
		stage = malloc((size_t)(num < slice ? num : slice) * elem);
		if (stage == NULL) { st = fail("staging allocation", name); }
	}
	for (int64_t o = 0; st == 0 && o < num; o += slice) {
		const int64_t n = (num - o < slice) ? num - o : slice;
		const void* src = (const char*)data + (size_t)o * elem;
		if (on_device) {
			if (gli_memcpy(stage, src, n * (int64_t)elem, 1) != 0) {
				st = fail("device copy", name);
				break;
			}
			src = stage;
		}
		const hsize_t off[1] = { (hsize_t)o }, cnt[1] = { (hsize_t)n };
		const hid_t mspace = H5Screate_simple(1, cnt, NULL);
		if (mspace < 0 || H5Sselect_hyperslab(space, H5S_SELECT_SET, off, NULL, cnt, NULL) < 0 ||
		    H5Dwrite(dset, input, mspace, space, H5P_DEFAULT, src) < 0) {
			st = fail("H5Dwrite", name);
		}
		if (mspace >= 0) { H5Sclose(mspace); }
	}
	free(stage);
	H5Dclose(dset);
	H5Sclose(space);
	return st;
}

int gli_eri_save_hdf5(const char* filename, int64_t grid_length, double tol, const double* coeffs, int64_t K,
                      const int64_t* four_indices, const double* values, int64_t num, int on_device)
{
	if (filename == NULL || grid_length <= 0 || K <= 0 || num < 0) { return fail("argument check", filename ? filename : "(null)"); }
	const hid_t file = H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
	if (file < 0) { return fail("H5Fcreate", filename); }
	const long long gl = grid_length;
	const hsize_t dk[1] = { (hsize_t)K };
	int st = 0;
	if (write_hdf5_scalar_attribute(file, "grid_length", H5T_STD_I64LE, H5T_NATIVE_LLONG, &gl) < 0 ||
	    write_hdf5_scalar_attribute(file, "tol", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &tol) < 0 ||
	    write_hdf5_dataset(file, "gausslet_coeffs", 1, dk, H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, coeffs) < 0 ||
	    write_streamed(file, "integral_values", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, sizeof(double), values, num, on_device) < 0 ||
	    write_streamed(file, "four_indices", H5T_STD_I64LE, H5T_NATIVE_LLONG, sizeof(int64_t), four_indices, num, on_device) < 0) {
		st = -1;
	}
	if (H5Fclose(file) < 0) { st = fail("H5Fclose", filename); }
	return st;
}

int gli_eri_load_hdf5(const char* filename, struct sparse_eri_gausslet_integrals* eri, double** coeffs, int64_t* K, double* tol)
{
	if (filename == NULL || eri == NULL || coeffs == NULL || K == NULL || tol == NULL) { return fail("argument check", "(null)"); }
	memset(eri, 0, sizeof(*eri));
	*coeffs = NULL;
	const hid_t file = H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT);
	if (file < 0) { return fail("H5Fopen", filename); }
	long long gl = 0;
	hsize_t dc[1] = { 0 }, dv[1] = { 0 }, di[1] = { 0 };
	int st = 0;
	if (read_hdf5_attribute(file, "grid_length", H5T_NATIVE_LLONG, &gl) < 0 || gl <= 0 ||
	    read_hdf5_attribute(file, "tol", H5T_NATIVE_DOUBLE, tol) < 0 ||
	    get_hdf5_dataset_dims(file, "gausslet_coeffs", dc) < 0 || (dc[0] % 2) != 1 ||
	    get_hdf5_dataset_dims(file, "integral_values", dv) < 0 ||
	    get_hdf5_dataset_dims(file, "four_indices", di) < 0 || di[0] != dv[0]) {
		st = fail("reading the sparse ERI layout", filename);
	}
	if (st == 0) {
		*K = (int64_t)dc[0];
		*coeffs = aligned_malloc(dc[0] * sizeof(double));
		eri->num_entries = (glong)dv[0];
		eri->integral_values = aligned_malloc((dv[0] ? dv[0] : 1) * sizeof(double));
		eri->four_indices = aligned_malloc((di[0] ? di[0] : 1) * sizeof(glong));
		for (int a = 0; a < 3; ++a) {
			eri->grid.coord_range[a].istart = -(gl - 1) / 2;
			eri->grid.coord_range[a].num = gl;
		}
		if (read_hdf5_dataset(file, "gausslet_coeffs", H5T_NATIVE_DOUBLE, *coeffs) < 0 ||
		    (dv[0] > 0 && read_hdf5_dataset(file, "integral_values", H5T_NATIVE_DOUBLE, eri->integral_values) < 0) ||
		    (di[0] > 0 && read_hdf5_dataset(file, "four_indices", H5T_NATIVE_LLONG, eri->four_indices) < 0)) {
			st = -1;
		}
	}
	H5Fclose(file);
	if (st != 0) {
		aligned_free(*coeffs);
		*coeffs = NULL;
		aligned_free(eri->integral_values);
		aligned_free(eri->four_indices);
		memset(eri, 0, sizeof(*eri));
	}
	return st;
}
