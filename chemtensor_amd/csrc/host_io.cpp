// save_mps / load_mps (src/state/mps.c:1219-1461): an MPS in an HDF5 file with the reference's
// layout, so files move between the two libraries:
//   root attributes  nsites (scalar, 32-bit integer), qsite (d entries), qbond_i for i = 0 .. nsites
//                    (the quantum numbers of virtual bond i: left bond of site i, right bond of the
//                    last site for i = nsites)
//   datasets         tensor_i, the dense site tensor i [D_i, d, D_{i+1}], IEEE float64 little-endian
// The HDF5 calls live in libchemtensor_h5.so (h5_shim.cpp), built only where the HDF5 library exists
// and loaded from this library's directory on first use; without it both calls report the failure
// and return -1, as the reference does for any HDF5 error. Only real double MPS are supported here
// (the reference also stores complex types as compound datasets).

#include <dlfcn.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ct_internal.h"

namespace ctg {

namespace {

struct H5Api
{
	int64_t (*create)(const char*) = nullptr;
	int64_t (*open)(const char*) = nullptr;
	int (*close)(int64_t) = nullptr;
	int (*write_int_attr)(int64_t, const char*, const int32_t*, int64_t, int) = nullptr;
	int64_t (*attr_len)(int64_t, const char*) = nullptr;
	int (*read_int_attr)(int64_t, const char*, int32_t*, int64_t) = nullptr;
	int (*write_f64)(int64_t, const char*, int, const int64_t*, const double*) = nullptr;
	int (*shape)(int64_t, const char*, int64_t*, int, int*) = nullptr;
	int (*read_f64)(int64_t, const char*, double*) = nullptr;
	std::string error;
};

void h5_anchor() {}

H5Api load_h5()
{
	H5Api a;
	std::string path;
	if (const char* e = getenv("CT_H5_LIB")) path = e;
	else {
		Dl_info info;
		if (dladdr((void*)&h5_anchor, &info) == 0 || info.dli_fname == nullptr) { a.error = "cannot locate libchemtensor_gpu.so"; return a; }
		path = info.dli_fname;
		const size_t slash = path.rfind('/');
		path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/libchemtensor_h5.so";
	}
	void* h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
	if (!h) { a.error = std::string("HDF5 support not built (") + dlerror() + ")"; return a; }
	bool ok = true;
	auto sym = [&](const char* name) { void* p = dlsym(h, name); if (!p) ok = false; return p; };
	a.create = (int64_t (*)(const char*))sym("ct_h5_create");
	a.open = (int64_t (*)(const char*))sym("ct_h5_open");
	a.close = (int (*)(int64_t))sym("ct_h5_close");
	a.write_int_attr = (int (*)(int64_t, const char*, const int32_t*, int64_t, int))sym("ct_h5_write_int_attr");
	a.attr_len = (int64_t (*)(int64_t, const char*))sym("ct_h5_attr_len");
	a.read_int_attr = (int (*)(int64_t, const char*, int32_t*, int64_t))sym("ct_h5_read_int_attr");
	a.write_f64 = (int (*)(int64_t, const char*, int, const int64_t*, const double*))sym("ct_h5_write_f64_dataset");
	a.shape = (int (*)(int64_t, const char*, int64_t*, int, int*))sym("ct_h5_dataset_shape");
	a.read_f64 = (int (*)(int64_t, const char*, double*))sym("ct_h5_read_f64_dataset");
	if (!ok) { a.error = "libchemtensor_h5.so lacks an entry point"; a.create = nullptr; }
	return a;
}

const H5Api& h5()
{
	static std::once_flag once;
	static H5Api api;
	std::call_once(once, [] { api = load_h5(); });
	return api;
}

// Visit every stored entry of t: f(logical flat index in the dense tensor, packed entry pointer).
// The dense tensor is row-major over the logical dimensions; within a block the entries follow the
// logical indices carrying the block's quantum number, in order (bst.c:891-915).
template <class F>
void for_each_entry(const block_sparse_tensor* t, F f)
{
	const Meta m = meta_of(t);
	const int nd = m.ndim;
	std::vector<ct_long> idx(nd), bd(nd), pos(nd), stride(nd);
	ct_long s = 1;
	for (int a = nd - 1; a >= 0; a--) { stride[a] = s; s *= m.dim_logical[a]; }
	std::vector<std::vector<ct_long>> fan(nd);
	for (ct_long k = 0; k < m.nblocks; k++) {
		if (!m.present(k)) continue;
		m.block_index(k, idx.data());
		m.block_dims(k, bd.data());
		for (int a = 0; a < nd; a++) fan[a] = fanout(m, a, m.qnums_blocks[a][idx[a]]);
		double* data = (double*)t->blocks[k]->data;
		const ct_long n = m.block_size(k);
		std::fill(pos.begin(), pos.end(), 0);
		for (ct_long e = 0; e < n; e++) {
			ct_long flat = 0;
			for (int a = 0; a < nd; a++) flat += fan[a][pos[a]] * stride[a];
			f(flat, data + e);
			for (int a = nd - 1; a >= 0; a--) { if (++pos[a] < bd[a]) break; pos[a] = 0; }
		}
	}
}

}  // namespace

}  // namespace ctg

extern "C" {

int save_mps(const char* filename, const struct mps* mps)
{
	using namespace ctg;
	const H5Api& a = h5();
	if (!a.create) { fprintf(stderr, "save_mps: %s\n", a.error.c_str()); return -1; }
	if (mps == nullptr || mps->nsites <= 0) { fprintf(stderr, "save_mps: empty MPS\n"); return -1; }
	for (int i = 0; i < mps->nsites; i++)
		if (mps->a[i].dtype != CT_DOUBLE_REAL || mps->a[i].ndim != 3) { fprintf(stderr, "save_mps: only real double MPS tensors are supported\n"); return -1; }
	const int64_t f = a.create(filename);
	if (f < 0) { fprintf(stderr, "'H5Fcreate' failed for '%s'\n", filename); return -1; }
	int ret = 0;
	const int32_t ns = mps->nsites;
	std::vector<int32_t> buf(mps->qsite, mps->qsite + mps->d);
	if (a.write_int_attr(f, "nsites", &ns, 1, 1) < 0 || a.write_int_attr(f, "qsite", buf.data(), (int64_t)mps->d, 0) < 0) ret = -1;
	for (int i = 0; ret == 0 && i <= mps->nsites; i++) {
		const block_sparse_tensor& t = (i < mps->nsites) ? mps->a[i] : mps->a[mps->nsites - 1];
		const int ax = (i < mps->nsites) ? 0 : 2;
		buf.assign(t.qnums_logical[ax], t.qnums_logical[ax] + t.dim_logical[ax]);
		const std::string name = "qbond_" + std::to_string(i);
		if (a.write_int_attr(f, name.c_str(), buf.data(), (int64_t)t.dim_logical[ax], 0) < 0) ret = -1;
	}
	for (int i = 0; ret == 0 && i < mps->nsites; i++) {
		const block_sparse_tensor* t = &mps->a[i];
		const int64_t dims[3] = { t->dim_logical[0], t->dim_logical[1], t->dim_logical[2] };
		std::vector<double> dense((size_t)(dims[0] * dims[1] * dims[2]), 0.0);
		for_each_entry(t, [&](ct_long flat, double* v) { dense[(size_t)flat] = *v; });
		const std::string name = "tensor_" + std::to_string(i);
		if (a.write_f64(f, name.c_str(), 3, dims, dense.data()) < 0) {
			fprintf(stderr, "writing '%s' tensor data to HDF5 file '%s' failed\n", name.c_str(), filename);
			ret = -1;
		}
	}
	if (a.close(f) < 0) { fprintf(stderr, "'H5Fclose' failed for '%s'\n", filename); ret = -1; }
	return ret;
}

int load_mps(const char* filename, struct mps* mps)
{
	using namespace ctg;
	const H5Api& a = h5();
	if (!a.create) { fprintf(stderr, "load_mps: %s\n", a.error.c_str()); return -1; }
	const int64_t f = a.open(filename);
	if (f < 0) { fprintf(stderr, "'H5Fopen' failed for '%s'\n", filename); return -1; }
	auto fail = [&](const std::string& msg) { fprintf(stderr, "load_mps: %s in HDF5 file '%s'\n", msg.c_str(), filename); a.close(f); return -1; };
	int32_t nsites = 0;
	if (a.attr_len(f, "nsites") != 1 || a.read_int_attr(f, "nsites", &nsites, 1) < 0) return fail("reading 'nsites' failed");
	if (nsites <= 0) return fail("'nsites' must be positive");
	const int64_t d = a.attr_len(f, "qsite");
	if (d <= 0) return fail("reading 'qsite' failed");
	std::vector<int32_t> qsite((size_t)d);
	if (a.read_int_attr(f, "qsite", qsite.data(), d) < 0) return fail("reading 'qsite' failed");
	std::vector<std::vector<int32_t>> qb((size_t)nsites + 1);
	for (int i = 0; i <= nsites; i++) {
		const std::string name = "qbond_" + std::to_string(i);
		const int64_t n = a.attr_len(f, name.c_str());
		if (n <= 0) return fail("reading '" + name + "' failed");
		qb[(size_t)i].resize((size_t)n);
		if (a.read_int_attr(f, name.c_str(), qb[(size_t)i].data(), n) < 0) return fail("reading '" + name + "' failed");
	}
	// every site tensor must be a real float64 array of the bond and site dimensions
	for (int i = 0; i < nsites; i++) {
		const std::string name = "tensor_" + std::to_string(i);
		int64_t dims[8];
		int is_f64 = 0;
		if (a.shape(f, name.c_str(), dims, 8, &is_f64) != 3) return fail("'" + name + "' is missing or not of rank 3");
		if (!is_f64) return fail("'" + name + "' is not real float64 (only real double MPS are supported)");
		if (dims[0] != (int64_t)qb[(size_t)i].size() || dims[1] != d || dims[2] != (int64_t)qb[(size_t)i + 1].size())
			return fail("'" + name + "' has dimensions that do not match the bond quantum numbers");
	}
	struct mps out;
	out.nsites = nsites;
	out.d = d;
	out.qsite = (qnumber*)aligned_alloc(16, ((size_t)d * sizeof(qnumber) + 15) / 16 * 16);
	memcpy(out.qsite, qsite.data(), (size_t)d * sizeof(qnumber));
	out.a = (block_sparse_tensor*)calloc((size_t)nsites, sizeof(block_sparse_tensor));
	for (int i = 0; i < nsites; i++) {
		const ct_long dim[3] = { (ct_long)qb[(size_t)i].size(), d, (ct_long)qb[(size_t)i + 1].size() };
		const tensor_axis_direction ad[3] = { TENSOR_AXIS_OUT, TENSOR_AXIS_OUT, TENSOR_AXIS_IN };
		const qnumber* qn[3] = { qb[(size_t)i].data(), out.qsite, qb[(size_t)i + 1].data() };
		allocate_block_sparse_tensor(CT_DOUBLE_REAL, 3, dim, ad, qn, &out.a[i]);
	}
	for (int i = 0; i < nsites; i++) {
		const std::string name = "tensor_" + std::to_string(i);
		block_sparse_tensor* t = &out.a[i];
		std::vector<double> dense((size_t)(t->dim_logical[0] * t->dim_logical[1] * t->dim_logical[2]));
		if (a.read_f64(f, name.c_str(), dense.data()) < 0) {
			delete_mps(&out);
			return fail("could not read the entries of '" + name + "'");
		}
		for_each_entry(t, [&](ct_long flat, double* v) { *v = dense[(size_t)flat]; });
	}
	if (a.close(f) < 0) { delete_mps(&out); fprintf(stderr, "'H5Fclose' failed for '%s'\n", filename); return -1; }
	*mps = out;
	return 0;
}

}  // extern "C"
