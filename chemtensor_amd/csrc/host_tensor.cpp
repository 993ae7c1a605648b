// Host-side block bookkeeping and the reference-compatible host entry points of
// libchemtensor_gpu (allocation, copy, (de)serialisation, block lookup).
//
// The block table follows src/tensor/block_sparse_tensor.c:37-150: the distinct quantum
// numbers of every axis are sorted ascending, a block exists iff the signed sum of its block
// quantum numbers vanishes, blocks are stored in row-major order over the block index, and
// within a block the entries follow the first-appearance order of their logical indices.

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstring>

#include "ct_internal.h"

namespace ctg {

void Meta::block_index(ct_long k, ct_long* idx) const
{
	for (int i = ndim - 1; i >= 0; i--) { idx[i] = k % dim_blocks[i]; k /= dim_blocks[i]; }
}

ct_long Meta::flat_index(const ct_long* idx) const
{
	ct_long k = 0;
	for (int i = 0; i < ndim; i++) k = k * dim_blocks[i] + idx[i];
	return k;
}

void Meta::block_dims(ct_long k, ct_long* dims) const
{
	ct_long idx[16];
	block_index(k, idx);
	for (int i = 0; i < ndim; i++) dims[i] = block_count[i][idx[i]];
}

ct_long Meta::block_size(ct_long k) const
{
	ct_long idx[16];
	block_index(k, idx);
	ct_long n = 1;
	for (int i = 0; i < ndim; i++) n *= block_count[i][idx[i]];
	return n;
}

ct_long Meta::find_qnum(int ax, qnumber q) const
{
	const auto& v = qnums_blocks[ax];
	auto it = std::lower_bound(v.begin(), v.end(), q);
	if (it == v.end() || *it != q) return -1;
	return (ct_long)(it - v.begin());
}

Meta make_meta(int ndim, const ct_long* dim, const int* axis_dir, const qnumber* const* qnums)
{
	Meta m;
	m.ndim = ndim;
	m.dim_logical.assign(dim, dim + ndim);
	m.axis_dir.assign(axis_dir, axis_dir + ndim);
	m.qnums_logical.resize(ndim);
	m.qnums_blocks.resize(ndim);
	m.block_count.resize(ndim);
	m.dim_blocks.resize(ndim);
	for (int i = 0; i < ndim; i++) {
		m.qnums_logical[i].assign(qnums[i], qnums[i] + dim[i]);
		std::vector<qnumber> q(m.qnums_logical[i]);
		std::sort(q.begin(), q.end());
		q.erase(std::unique(q.begin(), q.end()), q.end());
		m.qnums_blocks[i] = q;
		m.block_count[i].assign(q.size(), 0);
		for (qnumber x : m.qnums_logical[i]) m.block_count[i][std::lower_bound(q.begin(), q.end(), x) - q.begin()]++;
		m.dim_blocks[i] = (ct_long)q.size();
	}
	m.nblocks = 1;
	for (int i = 0; i < ndim; i++) m.nblocks *= m.dim_blocks[i];
	m.offset.assign(m.nblocks, -1);
	ct_long off = 0;
	if (ndim == 0) {
		m.offset[0] = 0;
		m.nelem = 1;
		return m;
	}
	// the blocks in row-major order (bst.c:119-143): for every index of the leading ndim - 1 axes the
	// quantum number of the last axis is fixed by the vanishing sum, so at most one block of the
	// last axis exists; it is found by binary search instead of scanning the last axis
	if (m.nblocks == 0) {
		m.nelem = 0;
		return m;
	}
	const int last = ndim - 1;
	const std::vector<qnumber>& qlast = m.qnums_blocks[last];
	std::vector<ct_long> idx((size_t)ndim, 0);
	ct_long prefix = 0;
	const ct_long nprefix = m.nblocks / m.dim_blocks[last];
	for (ct_long p = 0; p < nprefix; p++, prefix++) {
		qnumber qsum = 0;
		ct_long sz = 1;
		for (int i = 0; i < last; i++) { qsum += axis_dir[i] * m.qnums_blocks[i][idx[i]]; sz *= m.block_count[i][idx[i]]; }
		// axis_dir[last] * q + qsum = 0 with axis_dir = +-1
		const qnumber q = -axis_dir[last] * qsum;
		const auto it = std::lower_bound(qlast.begin(), qlast.end(), q);
		if (it != qlast.end() && *it == q) {
			const ct_long j = (ct_long)(it - qlast.begin());
			m.offset[prefix * m.dim_blocks[last] + j] = off;
			off += sz * m.block_count[last][j];
		}
		for (int i = last - 1; i >= 0; i--) { if (++idx[i] < m.dim_blocks[i]) break; idx[i] = 0; }
	}
	m.nelem = off;
	return m;
}

Meta meta_of(const block_sparse_tensor* t)
{
	std::vector<int> ad(t->ndim);
	for (int i = 0; i < t->ndim; i++) ad[i] = t->axis_dir[i];
	return make_meta(t->ndim, t->dim_logical, ad.data(), (const qnumber* const*)t->qnums_logical);
}

std::vector<ct_long> fanout(const Meta& m, int ax, qnumber q)
{
	std::vector<ct_long> r;
	for (ct_long j = 0; j < m.dim_logical[ax]; j++) if (m.qnums_logical[ax][j] == q) r.push_back(j);
	return r;
}

// allocation compatible with the reference's ct_malloc / ct_free (aligned_memory.h:18-26)
static void* ct_malloc(size_t n)
{
	const size_t sz = ((n > 0 ? n : 1) + 15) / 16 * 16;
	void* p = aligned_alloc(16, sz);
	if (!p) { fprintf(stderr, "chemtensor_gpu: host allocation of %zu bytes failed\n", n); abort(); }
	return p;
}

static void* ct_calloc(size_t n, size_t s)
{
	void* p = ct_malloc(n * s);
	memset(p, 0, n * s > 0 ? n * s : 1);
	return p;
}

}  // namespace ctg

using namespace ctg;

static void alloc_dense(int ndim, const ct_long* dim, dense_tensor* t)
{
	t->dtype = CT_DOUBLE_REAL;
	t->ndim = ndim;
	t->dim = (ct_long*)ctg::ct_malloc((size_t)(ndim > 0 ? ndim : 1) * sizeof(ct_long));
	ct_long n = 1;
	for (int i = 0; i < ndim; i++) { t->dim[i] = dim[i]; n *= dim[i]; }
	t->data = ctg::ct_calloc((size_t)n, sizeof(double));
}

static ct_long dense_nelem(const dense_tensor* t)
{
	ct_long n = 1;
	for (int i = 0; i < t->ndim; i++) n *= t->dim[i];
	return n;
}

extern "C" {

void allocate_block_sparse_tensor(const enum numeric_type dtype, const int ndim, const ct_long* dim, const enum tensor_axis_direction* axis_dir, const qnumber** qnums, struct block_sparse_tensor* t)
{
	if (dtype != CT_DOUBLE_REAL) {
		fprintf(stderr, "chemtensor_gpu: only CT_DOUBLE_REAL tensors are supported\n");
		abort();
	}
	std::vector<int> ad(ndim);
	for (int i = 0; i < ndim; i++) ad[i] = axis_dir[i];
	const Meta m = make_meta(ndim, dim, ad.data(), qnums);
	t->dtype = dtype;
	t->ndim = ndim;
	if (ndim == 0) {
		t->dim_logical = nullptr; t->dim_blocks = nullptr; t->axis_dir = nullptr;
		t->qnums_logical = nullptr; t->qnums_blocks = nullptr;
		t->blocks = (dense_tensor**)ct_calloc(1, sizeof(dense_tensor*));
		t->blocks[0] = (dense_tensor*)ct_calloc(1, sizeof(dense_tensor));
		alloc_dense(0, nullptr, t->blocks[0]);
		return;
	}
	t->dim_logical = (ct_long*)ct_malloc(ndim * sizeof(ct_long));
	t->dim_blocks = (ct_long*)ct_malloc(ndim * sizeof(ct_long));
	t->axis_dir = (tensor_axis_direction*)ct_malloc(ndim * sizeof(tensor_axis_direction));
	t->qnums_logical = (qnumber**)ct_calloc(ndim, sizeof(qnumber*));
	t->qnums_blocks = (qnumber**)ct_calloc(ndim, sizeof(qnumber*));
	for (int i = 0; i < ndim; i++) {
		t->dim_logical[i] = dim[i];
		t->dim_blocks[i] = m.dim_blocks[i];
		t->axis_dir[i] = axis_dir[i];
		t->qnums_logical[i] = (qnumber*)ct_malloc((size_t)dim[i] * sizeof(qnumber));
		memcpy(t->qnums_logical[i], qnums[i], (size_t)dim[i] * sizeof(qnumber));
		t->qnums_blocks[i] = (qnumber*)ct_malloc((size_t)m.dim_blocks[i] * sizeof(qnumber));
		memcpy(t->qnums_blocks[i], m.qnums_blocks[i].data(), (size_t)m.dim_blocks[i] * sizeof(qnumber));
	}
	t->blocks = (dense_tensor**)ct_calloc((size_t)m.nblocks, sizeof(dense_tensor*));
	ct_long bd[16];
	for (ct_long k = 0; k < m.nblocks; k++) {
		if (!m.present(k)) continue;
		m.block_dims(k, bd);
		t->blocks[k] = (dense_tensor*)ct_calloc(1, sizeof(dense_tensor));
		alloc_dense(ndim, bd, t->blocks[k]);
	}
}

void allocate_block_sparse_tensor_like(const struct block_sparse_tensor* s, struct block_sparse_tensor* t)
{
	allocate_block_sparse_tensor(s->dtype, s->ndim, s->dim_logical, s->axis_dir, (const qnumber**)s->qnums_logical, t);
}

void delete_block_sparse_tensor(struct block_sparse_tensor* t)
{
	if (t->blocks == nullptr) return;
	ct_long nb = 1;
	for (int i = 0; i < t->ndim; i++) nb *= t->dim_blocks[i];
	for (ct_long k = 0; k < nb; k++) {
		if (t->blocks[k] != nullptr) {
			free(t->blocks[k]->data);
			free(t->blocks[k]->dim);
			free(t->blocks[k]);
		}
	}
	free(t->blocks);
	t->blocks = nullptr;
	for (int i = 0; i < t->ndim; i++) { free(t->qnums_blocks[i]); free(t->qnums_logical[i]); }
	free(t->qnums_blocks); free(t->qnums_logical);
	free(t->axis_dir); free(t->dim_blocks); free(t->dim_logical);
	t->qnums_blocks = nullptr; t->qnums_logical = nullptr; t->axis_dir = nullptr; t->dim_blocks = nullptr; t->dim_logical = nullptr;
}

void copy_block_sparse_tensor(const struct block_sparse_tensor* src, struct block_sparse_tensor* dst)
{
	allocate_block_sparse_tensor_like(src, dst);
	ct_long nb = 1;
	for (int i = 0; i < src->ndim; i++) nb *= src->dim_blocks[i];
	for (ct_long k = 0; k < nb; k++) {
		if (src->blocks[k] == nullptr) continue;
		memcpy(dst->blocks[k]->data, src->blocks[k]->data, (size_t)dense_nelem(src->blocks[k]) * sizeof(double));
	}
}

struct dense_tensor* block_sparse_tensor_get_block(const struct block_sparse_tensor* t, const qnumber* qnums)
{
	ct_long k = 0;
	for (int i = 0; i < t->ndim; i++) {
		ct_long j;
		for (j = 0; j < t->dim_blocks[i]; j++) if (t->qnums_blocks[i][j] == qnums[i]) break;
		if (j == t->dim_blocks[i]) return nullptr;
		k = k * t->dim_blocks[i] + j;
	}
	return t->blocks[k];
}

double block_sparse_tensor_norm2(const struct block_sparse_tensor* t)
{
	ct_long nb = 1;
	for (int i = 0; i < t->ndim; i++) nb *= t->dim_blocks[i];
	double s = 0;
	for (ct_long k = 0; k < nb; k++) {
		const dense_tensor* b = t->blocks[k];
		if (!b) continue;
		const double* d = (const double*)b->data;
		const ct_long n = dense_nelem(b);
		for (ct_long j = 0; j < n; j++) s += d[j] * d[j];
	}
	return std::sqrt(s);
}

ct_long block_sparse_tensor_num_elements_blocks(const struct block_sparse_tensor* t)
{
	ct_long nb = 1;
	for (int i = 0; i < t->ndim; i++) nb *= t->dim_blocks[i];
	ct_long n = 0;
	for (ct_long k = 0; k < nb; k++) if (t->blocks[k]) n += dense_nelem(t->blocks[k]);
	return n;
}

void block_sparse_tensor_serialize_entries(const struct block_sparse_tensor* t, void* entries)
{
	ct_long nb = 1;
	for (int i = 0; i < t->ndim; i++) nb *= t->dim_blocks[i];
	double* out = (double*)entries;
	ct_long c = 0;
	for (ct_long k = 0; k < nb; k++) {
		const dense_tensor* b = t->blocks[k];
		if (!b) continue;
		const ct_long n = dense_nelem(b);
		memcpy(out + c, b->data, (size_t)n * sizeof(double));
		c += n;
	}
}

void block_sparse_tensor_deserialize_entries(struct block_sparse_tensor* t, const void* entries)
{
	ct_long nb = 1;
	for (int i = 0; i < t->ndim; i++) nb *= t->dim_blocks[i];
	const double* in = (const double*)entries;
	ct_long c = 0;
	for (ct_long k = 0; k < nb; k++) {
		dense_tensor* b = t->blocks[k];
		if (!b) continue;
		const ct_long n = dense_nelem(b);
		memcpy(b->data, in + c, (size_t)n * sizeof(double));
		c += n;
	}
}

}  // extern "C"
