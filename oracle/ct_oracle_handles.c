/// \file ct_oracle_handles.c
/// \brief Heap-handle helpers that let the Python test harness drive the CPU oracle
/// through ctypes (TEST INFRASTRUCTURE ONLY).
///
/// Tensors, MPS and MPO objects are returned as pointers to heap-allocated structs with
/// the reference memory layout, so they can also be handed unchanged to the product
/// library's drop-in entry points.

#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "ct_oracle.h"

#define ORC_EXPORT __attribute__((visibility("default")))

static void* hmalloc(size_t n) { void* p = malloc(n > 0 ? n : 1); if (!p) abort(); return p; }

ORC_EXPORT struct block_sparse_tensor* orc_h_bst_new(int ndim, const ct_long* dim, const int* axis_dir, const qnumber* qnums_concat)
{
	struct block_sparse_tensor* t = hmalloc(sizeof(*t));
	enum tensor_axis_direction ad[16];
	const qnumber* qn[16];
	ct_long off = 0;
	for (int i = 0; i < ndim; i++) { ad[i] = (enum tensor_axis_direction)axis_dir[i]; qn[i] = qnums_concat + off; off += dim[i]; }
	orc_allocate_block_sparse_tensor(ndim, dim, ad, qn, t);
	return t;
}

ORC_EXPORT void orc_h_bst_free(struct block_sparse_tensor* t)
{
	if (t == NULL) return;
	orc_delete_block_sparse_tensor(t);
	free(t);
}

ORC_EXPORT int orc_h_bst_ndim(const struct block_sparse_tensor* t) { return t->ndim; }

ORC_EXPORT void orc_h_bst_dims(const struct block_sparse_tensor* t, ct_long* dim_logical, int* axis_dir)
{
	for (int i = 0; i < t->ndim; i++) { dim_logical[i] = t->dim_logical[i]; axis_dir[i] = t->axis_dir[i]; }
}

ORC_EXPORT void orc_h_bst_qnums(const struct block_sparse_tensor* t, int ax, qnumber* out)
{
	memcpy(out, t->qnums_logical[ax], (size_t)t->dim_logical[ax] * sizeof(qnumber));
}

ORC_EXPORT ct_long orc_h_bst_num_blocks_present(const struct block_sparse_tensor* t)
{
	const ct_long nb = (t->ndim == 0 ? 1 : orc_bst_num_blocks(t));
	ct_long c = 0;
	for (ct_long k = 0; k < nb; k++) if (t->blocks[k] != NULL) c++;
	return c;
}

ORC_EXPORT void orc_h_bst_set_dense(struct block_sparse_tensor* t, const double* dense)
{
	struct dense_tensor d;
	d.data = (void*)dense; d.dim = t->dim_logical; d.ndim = t->ndim; d.dtype = CT_DOUBLE_REAL;
	orc_dense_to_block_sparse_tensor_entries(&d, t);
}

ORC_EXPORT void orc_h_bst_get_dense(const struct block_sparse_tensor* t, double* dense)
{
	struct dense_tensor d;
	orc_block_sparse_to_dense_tensor(t, &d);
	memcpy(dense, d.data, (size_t)orc_dense_num_elements(&d) * sizeof(double));
	orc_delete_dense_tensor(&d);
}

ORC_EXPORT void orc_h_bst_fill_random(struct block_sparse_tensor* t, uint64_t seed, double alpha)
{
	struct orc_rng_state rng;
	orc_seed_rng_state(seed, &rng);
	orc_block_sparse_tensor_fill_random_normal(alpha, 0, &rng, t);
}

ORC_EXPORT struct block_sparse_tensor* orc_h_dot(const struct block_sparse_tensor* s, int axs, const struct block_sparse_tensor* t, int axt, int ndim_mult)
{
	struct block_sparse_tensor* r = hmalloc(sizeof(*r));
	orc_block_sparse_tensor_dot(s, (enum tensor_axis_range)axs, t, (enum tensor_axis_range)axt, ndim_mult, r);
	return r;
}

ORC_EXPORT struct block_sparse_tensor* orc_h_transpose(const int* perm, const struct block_sparse_tensor* t)
{
	struct block_sparse_tensor* r = hmalloc(sizeof(*r));
	orc_block_sparse_tensor_transpose(perm, t, r);
	return r;
}

ORC_EXPORT struct block_sparse_tensor* orc_h_flatten(const struct block_sparse_tensor* t, int i_ax, int new_dir)
{
	struct block_sparse_tensor* r = hmalloc(sizeof(*r));
	orc_block_sparse_tensor_flatten_axes(t, i_ax, (enum tensor_axis_direction)new_dir, r);
	return r;
}

ORC_EXPORT struct block_sparse_tensor* orc_h_split(const struct block_sparse_tensor* t, int i_ax, ct_long d0, ct_long d1, int dir0, int dir1, const qnumber* q0, const qnumber* q1)
{
	struct block_sparse_tensor* r = hmalloc(sizeof(*r));
	const ct_long nd[2] = { d0, d1 };
	const enum tensor_axis_direction ad[2] = { (enum tensor_axis_direction)dir0, (enum tensor_axis_direction)dir1 };
	const qnumber* qn[2] = { q0, q1 };
	orc_block_sparse_tensor_split_axis(t, i_ax, nd, ad, qn, r);
	return r;
}

ORC_EXPORT struct block_sparse_tensor* orc_h_trace(const struct block_sparse_tensor* t, int ndim_trace)
{
	struct block_sparse_tensor* r = hmalloc(sizeof(*r));
	orc_block_sparse_tensor_cyclic_partial_trace(t, ndim_trace, r);
	return r;
}

ORC_EXPORT struct block_sparse_tensor* orc_h_apply_local_hamiltonian(const struct block_sparse_tensor* a, const struct block_sparse_tensor* w, const struct block_sparse_tensor* l, const struct block_sparse_tensor* r)
{
	struct block_sparse_tensor* b = hmalloc(sizeof(*b));
	orc_apply_local_hamiltonian(a, w, l, r, b);
	return b;
}

ORC_EXPORT struct block_sparse_tensor* orc_h_copy(const struct block_sparse_tensor* t)
{
	struct block_sparse_tensor* r = hmalloc(sizeof(*r));
	orc_copy_block_sparse_tensor(t, r);
	return r;
}

ORC_EXPORT struct mpo* orc_h_mpo_xxz(int nsites, double J, double D, double h)
{
	struct mpo* m = hmalloc(sizeof(*m));
	orc_construct_heisenberg_xxz_1d_mpo(nsites, J, D, h, m);
	return m;
}

ORC_EXPORT struct mpo* orc_h_mpo_hubbard(int nsites, double t, double u, double mu)
{
	struct mpo* m = hmalloc(sizeof(*m));
	orc_construct_fermi_hubbard_1d_mpo(nsites, t, u, mu, m);
	return m;
}

ORC_EXPORT void orc_h_mpo_free(struct mpo* m) { if (m) { orc_delete_mpo(m); free(m); } }
ORC_EXPORT struct block_sparse_tensor* orc_h_mpo_site(struct mpo* m, int i) { return &m->a[i]; }
ORC_EXPORT int orc_h_mpo_nsites(const struct mpo* m) { return m->nsites; }

ORC_EXPORT ct_long orc_h_mpo_matrix(const struct mpo* m, double* out, ct_long nmax)
{
	double* mat; ct_long n;
	orc_mpo_to_matrix(m, &mat, &n);
	if (n * n <= nmax) memcpy(out, mat, (size_t)(n * n) * sizeof(double));
	free(mat);
	return n;
}

ORC_EXPORT struct block_sparse_tensor* orc_h_mpo_merge(const struct block_sparse_tensor* a0, const struct block_sparse_tensor* a1)
{
	struct block_sparse_tensor* r = hmalloc(sizeof(*r));
	orc_mpo_merge_tensor_pair(a0, a1, r);
	return r;
}

ORC_EXPORT struct mps* orc_h_mps_random(int nsites, ct_long d, const qnumber* qsite, qnumber sector, ct_long max_vdim, uint64_t seed)
{
	struct mps* m = hmalloc(sizeof(*m));
	struct orc_rng_state rng;
	orc_seed_rng_state(seed, &rng);
	orc_construct_random_mps(nsites, d, qsite, sector, max_vdim, &rng, m);
	return m;
}

ORC_EXPORT void orc_h_mps_free(struct mps* m) { if (m) { orc_delete_mps(m); free(m); } }
ORC_EXPORT struct block_sparse_tensor* orc_h_mps_site(struct mps* m, int i) { return &m->a[i]; }

ORC_EXPORT ct_long orc_h_mps_vector(const struct mps* m, double* out, ct_long nmax)
{
	double* v; ct_long n;
	orc_mps_to_statevector(m, &v, &n);
	if (n <= nmax) memcpy(out, v, (size_t)n * sizeof(double));
	free(v);
	return n;
}

ORC_EXPORT double orc_h_mps_orthonormalize(struct mps* m, int mode) { return orc_mps_orthonormalize_qr(m, (enum mps_orthonormalization_mode)mode); }

ORC_EXPORT int orc_h_dmrg(const struct mpo* H, struct mps* psi, int num_sweeps, int maxiter, double tol, ct_long max_vdim, double* en_sweeps, double* entropy)
{
	return orc_dmrg_twosite(H, num_sweeps, maxiter, tol, max_vdim, psi, en_sweeps, entropy);
}

// environments of a two-site problem at bond pair (i, i+1) of a right-normalised MPS:
// out[0] = theta (merged sites i, i+1), out[1] = h2, out[2] = left block, out[3] = right block
ORC_EXPORT void orc_h_local_problem(const struct mpo* H, struct mps* psi, int i, struct block_sparse_tensor** out)
{
	const int L = H->nsites;
	orc_mps_orthonormalize_qr(psi, MPS_ORTHONORMAL_RIGHT);
	// left-normalise sites 0..i-1 so that both environments are isometric
	for (int j = 0; j < i; j++) orc_mps_local_orthonormalize_qr(&psi->a[j], &psi->a[j + 1]);
	struct block_sparse_tensor* rb = hmalloc((size_t)L * sizeof(struct block_sparse_tensor));
	orc_compute_right_operator_blocks(psi, psi, H, rb);
	struct block_sparse_tensor lcur;
	orc_create_dummy_operator_block_left(&psi->a[0], &psi->a[0], &H->a[0], &lcur);
	for (int j = 0; j < i; j++) {
		struct block_sparse_tensor ln;
		orc_contraction_operator_step_left(&psi->a[j], &psi->a[j], &H->a[j], &lcur, &ln);
		orc_delete_block_sparse_tensor(&lcur);
		lcur = ln;
	}
	for (int k = 0; k < 4; k++) out[k] = hmalloc(sizeof(struct block_sparse_tensor));
	orc_mps_merge_tensor_pair(&psi->a[i], &psi->a[i + 1], out[0]);
	orc_mpo_merge_tensor_pair(&H->a[i], &H->a[i + 1], out[1]);
	*out[2] = lcur;
	orc_copy_block_sparse_tensor(&rb[i + 1], out[3]);
	for (int j = 0; j < L; j++) orc_delete_block_sparse_tensor(&rb[j]);
	free(rb);
}

// bond quantum numbers exactly as construct_random_mps draws them (mps.c:93-150);
// qbonds_out receives the concatenation of all L+1 bond lists, dims_out their lengths
ORC_EXPORT void orc_h_random_bond_qnums(int nsites, ct_long d, const qnumber* qsite, qnumber sector, ct_long max_vdim, uint64_t seed, ct_long* dims_out, qnumber* qbonds_out)
{
	struct mps m;
	struct orc_rng_state rng;
	orc_seed_rng_state(seed, &rng);
	orc_construct_random_mps(nsites, d, qsite, sector, max_vdim, &rng, &m);
	ct_long off = 0;
	for (int l = 0; l <= nsites; l++) {
		const struct block_sparse_tensor* a = &m.a[l < nsites ? l : nsites - 1];
		const int ax = (l < nsites ? 0 : 2);
		dims_out[l] = a->dim_logical[ax];
		memcpy(qbonds_out + off, a->qnums_logical[ax], (size_t)dims_out[l] * sizeof(qnumber));
		off += dims_out[l];
	}
	orc_delete_mps(&m);
}

// Krylov fixtures: dense symmetric matrix a (n x n) applied through the matrix-free interface
ORC_EXPORT void orc_h_lanczos_dense(ct_long n, const double* a, const double* vstart, int maxiter, double* alpha, double* beta, double* v, int* numiter)
{
	char* data = hmalloc(sizeof(ct_long) + (size_t)(n * n) * sizeof(double));
	memcpy(data, &n, sizeof(ct_long));
	memcpy(data + sizeof(ct_long), a, (size_t)(n * n) * sizeof(double));
	orc_lanczos_iteration_d(n, orc_dense_matvec, data, vstart, maxiter, alpha, beta, v, numiter);
	free(data);
}

ORC_EXPORT int orc_h_eigensystem_dense(ct_long n, const double* a, const double* vstart, int maxiter, int numeig, double* lambda, double* u_ritz)
{
	char* data = hmalloc(sizeof(ct_long) + (size_t)(n * n) * sizeof(double));
	memcpy(data, &n, sizeof(ct_long));
	memcpy(data + sizeof(ct_long), a, (size_t)(n * n) * sizeof(double));
	int ret = orc_eigensystem_krylov_symmetric(n, orc_dense_matvec, data, vstart, maxiter, numeig, lambda, u_ritz);
	free(data);
	return ret;
}

// mps_split_tensor_svd on a tensor with given split; outputs are new handles
ORC_EXPORT int orc_h_mps_split_svd(const struct block_sparse_tensor* a, ct_long d0, ct_long d1, const qnumber* q0, const qnumber* q1, double tol, ct_long max_vdim, int renormalize, int distr,
	struct block_sparse_tensor** a0, struct block_sparse_tensor** a1, double* info3)
{
	*a0 = hmalloc(sizeof(struct block_sparse_tensor));
	*a1 = hmalloc(sizeof(struct block_sparse_tensor));
	const ct_long d[2] = { d0, d1 };
	const qnumber* qs[2] = { q0, q1 };
	struct trunc_info info;
	int ret = orc_mps_split_tensor_svd(a, d, qs, tol, max_vdim, renormalize != 0, (enum singular_value_distr)distr, *a0, *a1, &info);
	info3[0] = info.norm_sigma; info3[1] = info.entropy; info3[2] = info.tol_eff;
	return ret;
}

ORC_EXPORT struct block_sparse_tensor* orc_h_mps_merge(const struct block_sparse_tensor* a0, const struct block_sparse_tensor* a1)
{
	struct block_sparse_tensor* r = hmalloc(sizeof(*r));
	orc_mps_merge_tensor_pair(a0, a1, r);
	return r;
}

ORC_EXPORT int orc_h_svd_dense(ct_long m, ct_long n, const double* a, double* u, double* s, double* vh)
{
	struct dense_tensor ta, tu, tv;
	const ct_long da[2] = { m, n };
	const ct_long k = m < n ? m : n;
	const ct_long du[2] = { m, k }, dv[2] = { k, n };
	orc_allocate_dense_tensor(2, da, &ta);
	memcpy(ta.data, a, (size_t)(m * n) * sizeof(double));
	orc_allocate_dense_tensor(2, du, &tu);
	orc_allocate_dense_tensor(2, dv, &tv);
	int ret = orc_dense_tensor_svd_fill(&ta, &tu, s, &tv);
	memcpy(u, tu.data, (size_t)(m * k) * sizeof(double));
	memcpy(vh, tv.data, (size_t)(k * n) * sizeof(double));
	orc_delete_dense_tensor(&ta); orc_delete_dense_tensor(&tu); orc_delete_dense_tensor(&tv);
	return ret;
}

ORC_EXPORT int orc_h_qr_dense(ct_long m, ct_long n, const double* a, double* q, double* r, int rq)
{
	struct dense_tensor ta, tq, tr;
	const ct_long da[2] = { m, n };
	const ct_long k = m < n ? m : n;
	orc_allocate_dense_tensor(2, da, &ta);
	memcpy(ta.data, a, (size_t)(m * n) * sizeof(double));
	int ret;
	if (!rq) {
		const ct_long dq[2] = { m, k }, dr[2] = { k, n };
		orc_allocate_dense_tensor(2, dq, &tq); orc_allocate_dense_tensor(2, dr, &tr);
		ret = orc_dense_tensor_qr_fill(&ta, &tq, &tr);
		memcpy(q, tq.data, (size_t)(m * k) * sizeof(double));
		memcpy(r, tr.data, (size_t)(k * n) * sizeof(double));
	}
	else {
		const ct_long dr[2] = { m, k }, dq[2] = { k, n };
		orc_allocate_dense_tensor(2, dq, &tq); orc_allocate_dense_tensor(2, dr, &tr);
		ret = orc_dense_tensor_rq_fill(&ta, &tr, &tq);
		memcpy(q, tq.data, (size_t)(k * n) * sizeof(double));
		memcpy(r, tr.data, (size_t)(m * k) * sizeof(double));
	}
	orc_delete_dense_tensor(&ta); orc_delete_dense_tensor(&tq); orc_delete_dense_tensor(&tr);
	return ret;
}

ORC_EXPORT int orc_h_tridiag_eigh(int n, double* diag, double* offdiag, double* z) { return orc_symmetric_tridiag_eigh(n, diag, offdiag, z); }

ORC_EXPORT ct_long orc_h_retained_bond_indices(const double* sigma, ct_long n, double tol, int relative, ct_long max_vdim, ct_long* ind, double* info3)
{
	struct trunc_info info;
	ct_long num = orc_retained_bond_indices(sigma, n, tol, relative != 0, max_vdim, ind, &info);
	info3[0] = info.norm_sigma; info3[1] = info.entropy; info3[2] = info.tol_eff;
	return num;
}

ORC_EXPORT void orc_h_rng_draws(uint64_t seed, int kind, ct_long n, uint64_t bound, double* out_d, uint64_t* out_u)
{
	struct orc_rng_state rng;
	orc_seed_rng_state(seed, &rng);
	if (kind == 0) for (ct_long i = 0; i < n; i++) out_u[i] = orc_rand_uint64(&rng);
	else if (kind == 1) for (ct_long i = 0; i < n; i++) out_d[i] = orc_randn(&rng);
	else if (kind == 2) orc_rand_choice(bound, (uint64_t)n, &rng, out_u);
	else for (ct_long i = 0; i < n; i++) out_u[i] = orc_rand_interval(bound, &rng);
}
