// Thin C-ABI layer over the HDF5 C library: file create/open, integer attributes, float64 datasets.
// Built into its own shared library (libchemtensor_h5.so) only where the HDF5 headers and library
// exist; libchemtensor_gpu.so loads it on the first save_mps / load_mps call. The calls mirror the
// reference's helpers (src/util/hdf5_util.c:97-260): little-endian 32-bit integer attributes,
// IEEE float64 little-endian datasets written without object time stamps.

#include <hdf5.h>
#include <cstdint>
#include <cstdio>

extern "C" {

// create (truncate) a file; returns a handle > 0 or -1
int64_t ct_h5_create(const char* filename)
{
	const hid_t f = H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
	return f < 0 ? -1 : (int64_t)f;
}

// open a file read-only; returns a handle > 0 or -1
int64_t ct_h5_open(const char* filename)
{
	const hid_t f = H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT);
	return f < 0 ? -1 : (int64_t)f;
}

int ct_h5_close(int64_t file) { return H5Fclose((hid_t)file) < 0 ? -1 : 0; }

// root attribute of len 32-bit integers (scalar dataspace if scalar != 0, len must then be 1)
int ct_h5_write_int_attr(int64_t file, const char* name, const int32_t* data, int64_t len, int scalar)
{
	hsize_t dims[1] = { (hsize_t)len };
	const hid_t space = scalar ? H5Screate(H5S_SCALAR) : H5Screate_simple(1, dims, nullptr);
	if (space < 0) return -1;
	const hid_t attr = H5Acreate2((hid_t)file, name, H5T_STD_I32LE, space, H5P_DEFAULT, H5P_DEFAULT);
	if (attr < 0) { H5Sclose(space); return -1; }
	const herr_t st = H5Awrite(attr, H5T_NATIVE_INT32, data);
	H5Aclose(attr);
	H5Sclose(space);
	return st < 0 ? -1 : 0;
}

// number of elements of a root attribute (1 for a scalar), -1 if absent
int64_t ct_h5_attr_len(int64_t file, const char* name)
{
	if (H5Aexists((hid_t)file, name) <= 0) return -1;
	const hid_t attr = H5Aopen((hid_t)file, name, H5P_DEFAULT);
	if (attr < 0) return -1;
	const hid_t space = H5Aget_space(attr);
	const hssize_t n = space < 0 ? -1 : H5Sget_simple_extent_npoints(space);
	if (space >= 0) H5Sclose(space);
	H5Aclose(attr);
	return (int64_t)n;
}

// read a root attribute as 32-bit integers into data (cap entries)
int ct_h5_read_int_attr(int64_t file, const char* name, int32_t* data, int64_t cap)
{
	const int64_t n = ct_h5_attr_len(file, name);
	if (n < 0 || n > cap) return -1;
	const hid_t attr = H5Aopen((hid_t)file, name, H5P_DEFAULT);
	if (attr < 0) return -1;
	const herr_t st = H5Aread(attr, H5T_NATIVE_INT32, data);
	H5Aclose(attr);
	return st < 0 ? -1 : 0;
}

int ct_h5_write_f64_dataset(int64_t file, const char* name, int ndim, const int64_t* dims, const double* data)
{
	hsize_t d[32];
	if (ndim < 0 || ndim > 32) return -1;
	for (int i = 0; i < ndim; i++) d[i] = (hsize_t)dims[i];
	const hid_t space = H5Screate_simple(ndim, d, nullptr);
	if (space < 0) return -1;
	const hid_t cpl = H5Pcreate(H5P_DATASET_CREATE);
	H5Pset_obj_track_times(cpl, 0);
	const hid_t dset = H5Dcreate2((hid_t)file, name, H5T_IEEE_F64LE, space, H5P_DEFAULT, cpl, H5P_DEFAULT);
	herr_t st = -1;
	if (dset >= 0) {
		st = H5Dwrite(dset, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
		H5Dclose(dset);
	}
	H5Pclose(cpl);
	H5Sclose(space);
	return st < 0 ? -1 : 0;
}

// shape of a dataset: returns its rank (dims filled up to cap) and its type class in *is_float
// (1: IEEE float of 8 bytes, 0: anything else), -1 if absent
int ct_h5_dataset_shape(int64_t file, const char* name, int64_t* dims, int cap, int* is_f64)
{
	if (H5Lexists((hid_t)file, name, H5P_DEFAULT) <= 0) return -1;
	const hid_t dset = H5Dopen2((hid_t)file, name, H5P_DEFAULT);
	if (dset < 0) return -1;
	const hid_t space = H5Dget_space(dset);
	const hid_t type = H5Dget_type(dset);
	const int nd = space < 0 ? -1 : H5Sget_simple_extent_ndims(space);
	if (nd >= 0 && nd <= cap) {
		hsize_t d[32];
		H5Sget_simple_extent_dims(space, d, nullptr);
		for (int i = 0; i < nd; i++) dims[i] = (int64_t)d[i];
	}
	*is_f64 = (type >= 0 && H5Tget_class(type) == H5T_FLOAT && H5Tget_size(type) == 8) ? 1 : 0;
	if (type >= 0) H5Tclose(type);
	if (space >= 0) H5Sclose(space);
	H5Dclose(dset);
	return (nd > cap) ? -1 : nd;
}

int ct_h5_read_f64_dataset(int64_t file, const char* name, double* data)
{
	const hid_t dset = H5Dopen2((hid_t)file, name, H5P_DEFAULT);
	if (dset < 0) return -1;
	const herr_t st = H5Dread(dset, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
	H5Dclose(dset);
	return st < 0 ? -1 : 0;
}

}  // extern "C"
