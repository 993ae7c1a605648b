// h5_extract: dump the root-level datasets and attributes of an HDF5 fixture file into raw
// little-endian binary files plus a manifest, so that make_fixtures.py can turn them into
// committed .npz fixtures. Built and run only where the HDF5 C library is available (conda);
// the tests read the .npz fixtures and never need it.
//
// usage: h5_extract <file.hdf5> <outdir>
// manifest lines: <kind:D|A> <name> <type:f8|i8|c16> <ndim> <dim0> <dim1> ...

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <hdf5.h>

static FILE* manifest;
static const char* outdir;

static void dump(const char* kind, const char* name, hid_t type, hid_t space, int is_attr, hid_t obj)
{
	int nd = H5Sget_simple_extent_ndims(space);
	hsize_t dims[32];
	H5Sget_simple_extent_dims(space, dims, NULL);
	hsize_t n = 1;
	for (int i = 0; i < nd; i++) n *= dims[i];
	H5T_class_t cls = H5Tget_class(type);
	const char* tname;
	hid_t mem;
	size_t esize;
	if (cls == H5T_FLOAT) { tname = "f8"; mem = H5Tcopy(H5T_NATIVE_DOUBLE); esize = 8; }
	else if (cls == H5T_INTEGER) { tname = "i8"; mem = H5Tcopy(H5T_NATIVE_INT64); esize = 8; }
	else if (cls == H5T_COMPOUND) {
		tname = "c16";
		mem = H5Tcreate(H5T_COMPOUND, 16);
		H5Tinsert(mem, "r", 0, H5T_NATIVE_DOUBLE);
		H5Tinsert(mem, "i", 8, H5T_NATIVE_DOUBLE);
		esize = 16;
	}
	else { fprintf(stderr, "skipping %s (unsupported class)\n", name); return; }
	void* buf = malloc(n * esize + 16);
	herr_t st = is_attr ? H5Aread(obj, mem, buf) : H5Dread(obj, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
	if (st < 0) { fprintf(stderr, "read failed: %s\n", name); exit(1); }
	char path[4096];
	snprintf(path, sizeof(path), "%s/%s_%s.bin", outdir, kind, name);
	FILE* f = fopen(path, "wb");
	fwrite(buf, esize, n, f);
	fclose(f);
	fprintf(manifest, "%s %s %s %d", kind, name, tname, nd);
	for (int i = 0; i < nd; i++) fprintf(manifest, " %llu", (unsigned long long)dims[i]);
	fprintf(manifest, "\n");
	free(buf);
	H5Tclose(mem);
}

static herr_t attr_cb(hid_t loc, const char* name, const H5A_info_t* info, void* op)
{
	(void)info; (void)op;
	hid_t a = H5Aopen(loc, name, H5P_DEFAULT);
	hid_t t = H5Aget_type(a), s = H5Aget_space(a);
	dump("A", name, t, s, 1, a);
	H5Sclose(s); H5Tclose(t); H5Aclose(a);
	return 0;
}

static herr_t link_cb(hid_t loc, const char* name, const H5L_info_t* info, void* op)
{
	(void)info; (void)op;
	H5O_info_t oi;
	H5Oget_info_by_name(loc, name, &oi, H5P_DEFAULT);
	if (oi.type != H5O_TYPE_DATASET) return 0;
	hid_t d = H5Dopen2(loc, name, H5P_DEFAULT);
	hid_t t = H5Dget_type(d), s = H5Dget_space(d);
	dump("D", name, t, s, 0, d);
	H5Aiterate2(d, H5_INDEX_NAME, H5_ITER_INC, NULL, attr_cb, NULL);
	H5Sclose(s); H5Tclose(t); H5Dclose(d);
	return 0;
}

int main(int argc, char** argv)
{
	if (argc != 3) { fprintf(stderr, "usage: %s file.hdf5 outdir\n", argv[0]); return 2; }
	outdir = argv[2];
	hid_t f = H5Fopen(argv[1], H5F_ACC_RDONLY, H5P_DEFAULT);
	if (f < 0) return 1;
	char mpath[4096];
	snprintf(mpath, sizeof(mpath), "%s/manifest.txt", outdir);
	manifest = fopen(mpath, "w");
	H5Aiterate2(f, H5_INDEX_NAME, H5_ITER_INC, NULL, attr_cb, NULL);
	H5Literate(f, H5_INDEX_NAME, H5_ITER_INC, NULL, link_cb, NULL);
	fclose(manifest);
	H5Fclose(f);
	return 0;
}
