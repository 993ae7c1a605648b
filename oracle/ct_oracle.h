/// \file ct_oracle.h
/// \brief CPU FP64 oracle: a plain-C restatement of chemtensor's two-site DMRG path.
///
/// TEST INFRASTRUCTURE ONLY. This library is the checker for the GPU product in
/// chemtensor_amd/. Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
/// bench.py may load it. The product library never links, loads or calls it.
///
/// Every function cites the reference file:line whose algorithm it restates
/// (paths relative to the chemtensor source tree). All symbols carry the prefix
/// 'orc_' so that the oracle and the product can be loaded into one process.
///
/// Data structures have the same memory layout as the reference's
/// (include/tensor/dense_tensor.h:17-23, include/tensor/block_sparse_tensor.h:18-28).
/// Only real double precision (CT_DOUBLE_REAL) entries are supported.

#pragma once

#include <stdint.h>
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int64_t ct_long;   // include/util/util.h:12
typedef int32_t qnumber;   // include/tensor/qnumber.h:17

enum numeric_type { CT_SINGLE_REAL = 0, CT_DOUBLE_REAL = 1, CT_SINGLE_COMPLEX = 2, CT_DOUBLE_COMPLEX = 3 };  // include/numeric.h:15-22
enum tensor_axis_direction { TENSOR_AXIS_IN = -1, TENSOR_AXIS_OUT = 1 };                                   // include/tensor/qnumber.h:76-80
enum tensor_axis_range { TENSOR_AXIS_RANGE_LEADING = 0, TENSOR_AXIS_RANGE_TRAILING = 1 };                  // include/tensor/dense_tensor.h:180-185
enum singular_value_distr { SVD_DISTR_LEFT = 0, SVD_DISTR_RIGHT = 1 };                                     // include/algorithm/bond_ops.h:15-19
enum mps_orthonormalization_mode { MPS_ORTHONORMAL_LEFT = 0, MPS_ORTHONORMAL_RIGHT = 1 };                  // include/state/mps.h:77-81

struct dense_tensor
{
	void* data;
	ct_long* dim;
	enum numeric_type dtype;
	int ndim;
};

struct block_sparse_tensor
{
	struct dense_tensor** blocks;
	ct_long* dim_blocks;
	ct_long* dim_logical;
	enum tensor_axis_direction* axis_dir;
	qnumber** qnums_blocks;
	qnumber** qnums_logical;
	enum numeric_type dtype;
	int ndim;
};

struct mps { struct block_sparse_tensor* a; qnumber* qsite; ct_long d; int nsites; };  // include/state/mps.h:14-20
struct mpo { struct block_sparse_tensor* a; qnumber* qsite; ct_long d; int nsites; };  // include/operator/mpo.h:34-40
struct trunc_info { double norm_sigma; double entropy; double tol_eff; };            // include/algorithm/truncation.h:19-24

struct pcg32_state { uint64_t state; uint64_t inc; };
struct orc_rng_state { struct pcg32_state gen[2]; };

// random numbers (src/util/rng.c, src/util/pcg_basic.c)
void orc_seed_rng_state(uint64_t seed, struct orc_rng_state* state);
uint64_t orc_rand_uint64(struct orc_rng_state* state);
uint64_t orc_rand_interval(uint64_t bound, struct orc_rng_state* state);
void orc_rand_choice(uint64_t bound, uint64_t num_samples, struct orc_rng_state* state, uint64_t* ret);
double orc_randu(struct orc_rng_state* state);
double orc_randn(struct orc_rng_state* state);

// dense tensors (src/tensor/dense_tensor.c)
void orc_allocate_dense_tensor(int ndim, const ct_long* dim, struct dense_tensor* t);
void orc_delete_dense_tensor(struct dense_tensor* t);
ct_long orc_dense_num_elements(const struct dense_tensor* t);
void orc_dense_tensor_transpose(const int* perm, const struct dense_tensor* t, struct dense_tensor* r);
void orc_dense_tensor_dot_update(const struct dense_tensor* s, enum tensor_axis_range axrange_s, const struct dense_tensor* t, enum tensor_axis_range axrange_t, int ndim_mult, struct dense_tensor* r);
int orc_dense_tensor_qr_fill(const struct dense_tensor* a, struct dense_tensor* q, struct dense_tensor* r);
int orc_dense_tensor_rq_fill(const struct dense_tensor* a, struct dense_tensor* r, struct dense_tensor* q);
int orc_dense_tensor_svd_fill(const struct dense_tensor* a, struct dense_tensor* u, double* s, struct dense_tensor* vh);
int orc_symmetric_tridiag_eigh(int n, double* diag, double* offdiag, double* z);

// block-sparse tensors (src/tensor/block_sparse_tensor.c)
void orc_allocate_block_sparse_tensor(int ndim, const ct_long* dim, const enum tensor_axis_direction* axis_dir, const qnumber** qnums, struct block_sparse_tensor* t);
void orc_delete_block_sparse_tensor(struct block_sparse_tensor* t);
void orc_copy_block_sparse_tensor(const struct block_sparse_tensor* src, struct block_sparse_tensor* dst);
ct_long orc_bst_num_blocks(const struct block_sparse_tensor* t);
struct dense_tensor* orc_block_sparse_tensor_get_block(const struct block_sparse_tensor* t, const qnumber* qnums);
double orc_block_sparse_tensor_norm2(const struct block_sparse_tensor* t);
void orc_block_sparse_tensor_reverse_axis_directions(struct block_sparse_tensor* t);
void orc_scale_block_sparse_tensor(double alpha, struct block_sparse_tensor* t);
void orc_block_sparse_tensor_fill_random_normal(double alpha, double shift, struct orc_rng_state* rng, struct block_sparse_tensor* t);
void orc_block_sparse_to_dense_tensor(const struct block_sparse_tensor* s, struct dense_tensor* t);
void orc_dense_to_block_sparse_tensor_entries(const struct dense_tensor* t, struct block_sparse_tensor* s);
void orc_block_sparse_tensor_transpose(const int* perm, const struct block_sparse_tensor* t, struct block_sparse_tensor* r);
void orc_block_sparse_tensor_flatten_axes(const struct block_sparse_tensor* t, int i_ax, enum tensor_axis_direction new_axis_dir, struct block_sparse_tensor* r);
void orc_block_sparse_tensor_split_axis(const struct block_sparse_tensor* t, int i_ax, const ct_long new_dim_logical[2], const enum tensor_axis_direction new_axis_dir[2], const qnumber* new_qnums_logical[2], struct block_sparse_tensor* r);
void orc_block_sparse_tensor_slice(const struct block_sparse_tensor* t, int i_ax, const ct_long* ind, ct_long nind, struct block_sparse_tensor* r);
void orc_block_sparse_tensor_cyclic_partial_trace(const struct block_sparse_tensor* t, int ndim_trace, struct block_sparse_tensor* r);
void orc_block_sparse_tensor_multiply_pointwise_vector(const struct block_sparse_tensor* s, const double* vec, ct_long nvec, enum tensor_axis_range axrange, struct block_sparse_tensor* r);
void orc_block_sparse_tensor_dot(const struct block_sparse_tensor* s, enum tensor_axis_range axrange_s, const struct block_sparse_tensor* t, enum tensor_axis_range axrange_t, int ndim_mult, struct block_sparse_tensor* r);
ct_long orc_block_sparse_tensor_num_elements_blocks(const struct block_sparse_tensor* t);
void orc_block_sparse_tensor_serialize_entries(const struct block_sparse_tensor* t, double* entries);
void orc_block_sparse_tensor_deserialize_entries(struct block_sparse_tensor* t, const double* entries);
int orc_block_sparse_tensor_qr(const struct block_sparse_tensor* a, struct block_sparse_tensor* q, struct block_sparse_tensor* r);
int orc_block_sparse_tensor_rq(const struct block_sparse_tensor* a, struct block_sparse_tensor* r, struct block_sparse_tensor* q);
int orc_block_sparse_tensor_svd(const struct block_sparse_tensor* a, struct block_sparse_tensor* u, double** s, ct_long* ns, struct block_sparse_tensor* vh);

// truncation and bond operations (src/algorithm/truncation.c, src/algorithm/bond_ops.c)
double orc_von_neumann_entropy(const double* sigma, ct_long n);
ct_long orc_retained_bond_indices(const double* sigma, ct_long n, double tol, bool relative_thresh, ct_long max_vdim, ct_long* ind, struct trunc_info* info);
int orc_split_block_sparse_matrix_svd(const struct block_sparse_tensor* a, double tol, bool relative_thresh, ct_long max_vdim, bool renormalize, enum singular_value_distr svd_distr,
	struct block_sparse_tensor* a0, struct block_sparse_tensor* a1, struct trunc_info* info);

// Krylov (src/util/krylov.c)
typedef void orc_linear_func_d(ct_long n, const void* data, const double* v, double* ret);
void orc_lanczos_iteration_d(ct_long n, orc_linear_func_d afunc, const void* adata, const double* vstart, int maxiter, double* alpha, double* beta, double* v, int* numiter);
int orc_eigensystem_krylov_symmetric(ct_long n, orc_linear_func_d afunc, const void* adata, const double* vstart, int maxiter, int numeig, double* lambda, double* u_ritz);
void orc_dense_matvec(ct_long n, const void* data, const double* v, double* ret);

// MPS / MPO (src/state/mps.c, src/operator/mpo.c)
void orc_allocate_mps(int nsites, ct_long d, const qnumber* qsite, const ct_long* dim_bonds, const qnumber** qbonds, struct mps* mps);
void orc_delete_mps(struct mps* mps);
void orc_construct_random_mps(int nsites, ct_long d, const qnumber* qsite, qnumber qnum_sector, ct_long max_vdim, struct orc_rng_state* rng, struct mps* mps);
void orc_mps_local_orthonormalize_qr(struct block_sparse_tensor* a, struct block_sparse_tensor* a_next);
void orc_mps_local_orthonormalize_rq(struct block_sparse_tensor* a, struct block_sparse_tensor* a_prev);
double orc_mps_orthonormalize_qr(struct mps* mps, enum mps_orthonormalization_mode mode);
void orc_mps_merge_tensor_pair(const struct block_sparse_tensor* a0, const struct block_sparse_tensor* a1, struct block_sparse_tensor* a);
int orc_mps_split_tensor_svd(const struct block_sparse_tensor* a, const ct_long d[2], const qnumber* new_qsite[2], double tol, ct_long max_vdim, bool renormalize,
	enum singular_value_distr svd_distr, struct block_sparse_tensor* a0, struct block_sparse_tensor* a1, struct trunc_info* info);
void orc_mps_to_statevector(const struct mps* mps, double** vec, ct_long* n);
void orc_delete_mpo(struct mpo* mpo);
void orc_mpo_merge_tensor_pair(const struct block_sparse_tensor* a0, const struct block_sparse_tensor* a1, struct block_sparse_tensor* a);
void orc_mpo_to_matrix(const struct mpo* mpo, double** mat, ct_long* n);
void orc_construct_heisenberg_xxz_1d_mpo(int nsites, double J, double D, double h, struct mpo* mpo);
void orc_construct_fermi_hubbard_1d_mpo(int nsites, double t, double u, double mu, struct mpo* mpo);

// chain operations and DMRG (src/algorithm/chain_ops.c, src/algorithm/dmrg.c)
void orc_create_dummy_operator_block_right(const struct block_sparse_tensor* a, const struct block_sparse_tensor* b, const struct block_sparse_tensor* w, struct block_sparse_tensor* r);
void orc_create_dummy_operator_block_left(const struct block_sparse_tensor* a, const struct block_sparse_tensor* b, const struct block_sparse_tensor* w, struct block_sparse_tensor* l);
void orc_contraction_operator_step_right(const struct block_sparse_tensor* a, const struct block_sparse_tensor* b, const struct block_sparse_tensor* w, const struct block_sparse_tensor* r, struct block_sparse_tensor* r_next);
void orc_contraction_operator_step_left(const struct block_sparse_tensor* a, const struct block_sparse_tensor* b, const struct block_sparse_tensor* w, const struct block_sparse_tensor* l, struct block_sparse_tensor* l_next);
void orc_compute_right_operator_blocks(const struct mps* psi, const struct mps* chi, const struct mpo* op, struct block_sparse_tensor* r_list);
void orc_apply_local_hamiltonian(const struct block_sparse_tensor* a, const struct block_sparse_tensor* w, const struct block_sparse_tensor* l, const struct block_sparse_tensor* r, struct block_sparse_tensor* b);
int orc_dmrg_twosite(const struct mpo* hamiltonian, int num_sweeps, int maxiter_lanczos, double tol_split, ct_long max_vdim, struct mps* psi, double* en_sweeps, double* entropy);

// flat (pointer + size) helpers used by the Python test harness
int orc_bst_get_meta(const struct block_sparse_tensor* t, ct_long* dim_logical, ct_long* dim_blocks, int* axis_dir);

#ifdef __cplusplus
}
#endif
