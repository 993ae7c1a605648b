/// \file chemtensor_gpu.h
/// \brief C ABI of libchemtensor_gpu.so: an MI355X (gfx950) drop-in for chemtensor's two-site
/// DMRG path (apply_local_hamiltonian over abelian block-sparse tensors inside the Lanczos
/// ground-state iteration of dmrg_twosite).
///
/// Data structures have the reference's exact memory layout; functions named like the
/// reference's keep its signature, argument meaning, ownership and error convention:
/// results are allocated by the callee (free()-compatible, 16-byte aligned), inputs are const,
/// int returns are 0 on success and negative on failure. Only CT_DOUBLE_REAL entries are
/// supported. Errors of the HIP runtime are reported as CT_ERR_HIP (and on stderr).
///
/// Paths in citations are relative to the chemtensor source tree.

#pragma once

#include <stdint.h>
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

//________________________________________________________________________________________________________________________
//
// types (byte-identical to the reference)

typedef int64_t ct_long;   // include/util/util.h:12
typedef int32_t qnumber;   // include/tensor/qnumber.h:17

enum numeric_type { CT_SINGLE_REAL = 0, CT_DOUBLE_REAL = 1, CT_SINGLE_COMPLEX = 2, CT_DOUBLE_COMPLEX = 3, CT_NUM_NUMERIC_TYPES = 4 };  // include/numeric.h:15-22
enum tensor_axis_direction { TENSOR_AXIS_IN = -1, TENSOR_AXIS_OUT = 1 };                                                           // include/tensor/qnumber.h:76-80
enum tensor_axis_range { TENSOR_AXIS_RANGE_LEADING = 0, TENSOR_AXIS_RANGE_TRAILING = 1, TENSOR_AXIS_RANGE_NUM = 2 };                // include/tensor/dense_tensor.h:180-185
enum singular_value_distr { SVD_DISTR_LEFT = 0, SVD_DISTR_RIGHT = 1 };                                                             // include/algorithm/bond_ops.h:15-19
enum mps_orthonormalization_mode { MPS_ORTHONORMAL_LEFT = 0, MPS_ORTHONORMAL_RIGHT = 1 };                                          // include/state/mps.h:77-81

/// include/tensor/dense_tensor.h:17-23
struct dense_tensor
{
	void* data;
	ct_long* dim;
	enum numeric_type dtype;
	int ndim;
};

/// include/tensor/block_sparse_tensor.h:18-28
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

/// include/state/mps.h:14-20
struct mps { struct block_sparse_tensor* a; qnumber* qsite; ct_long d; int nsites; };
/// include/operator/mpo.h:34-40
struct mpo { struct block_sparse_tensor* a; qnumber* qsite; ct_long d; int nsites; };
/// include/algorithm/truncation.h:19-24
struct trunc_info { double norm_sigma; double entropy; double tol_eff; };
/// include/util/pcg_basic.h (struct pcg32_random) and include/util/rng.h:13-16
struct pcg32_random { uint64_t state; uint64_t inc; };
struct rng_state { struct pcg32_random gen[2]; };

/// error codes beyond the reference's (-1 Krylov breakdown / factorisation failure, -2 eigensolver failure)
#define CT_ERR_HIP        (-100)   ///< a HIP runtime call failed
#define CT_ERR_UNSUPPORTED (-101)  ///< input outside the supported domain (dtype, outer bond dimension > 1, ...)
#define CT_ERR_NOMEM      (-102)   ///< device memory exhausted or above the admission budget
#define CT_ERR_ARG        (-103)   ///< invalid argument

//________________________________________________________________________________________________________________________
//
// block-sparse tensors (host side; replaces src/tensor/block_sparse_tensor.c entry points)

/// bst.c:37-150 (declared include/tensor/block_sparse_tensor.h:33)
void allocate_block_sparse_tensor(const enum numeric_type dtype, const int ndim, const ct_long* dim, const enum tensor_axis_direction* axis_dir, const qnumber** qnums, struct block_sparse_tensor* t);
/// bst.c:158-162 (block_sparse_tensor.h:34)
void allocate_block_sparse_tensor_like(const struct block_sparse_tensor* s, struct block_sparse_tensor* t);
/// bst.c:168-205 (block_sparse_tensor.h:35)
void delete_block_sparse_tensor(struct block_sparse_tensor* t);
/// bst.c:213-270 (block_sparse_tensor.h:36)
void copy_block_sparse_tensor(const struct block_sparse_tensor* src, struct block_sparse_tensor* dst);
/// bst.c:278-307 (block_sparse_tensor.h:40)
struct dense_tensor* block_sparse_tensor_get_block(const struct block_sparse_tensor* t, const qnumber* qnums);
/// bst.c:316-334 (block_sparse_tensor.h:46)
double block_sparse_tensor_norm2(const struct block_sparse_tensor* t);
/// bst.c:3111-3124 (block_sparse_tensor.h:219)
ct_long block_sparse_tensor_num_elements_blocks(const struct block_sparse_tensor* t);
/// bst.c:3131-3151 (block_sparse_tensor.h:220)
void block_sparse_tensor_serialize_entries(const struct block_sparse_tensor* t, void* entries);
/// bst.c:3157-3180 (block_sparse_tensor.h:221)
void block_sparse_tensor_deserialize_entries(struct block_sparse_tensor* t, const void* entries);

/// bst.c:1826-2000 (block_sparse_tensor.h:167): r = s . t, computed on the GPU by one grouped
/// FP64 MFMA launch over all (output block, contracted block) pairs. Allocates r.
void block_sparse_tensor_dot(const struct block_sparse_tensor* s, const enum tensor_axis_range axrange_s, const struct block_sparse_tensor* t, const enum tensor_axis_range axrange_t, const int ndim_mult, struct block_sparse_tensor* r);
/// bst.c:785-870 (block_sparse_tensor.h:99), on the GPU. Allocates r.
void block_sparse_tensor_transpose(const int* perm, const struct block_sparse_tensor* t, struct block_sparse_tensor* r);
/// bst.c:1560-1612 (block_sparse_tensor.h:57): trace of the ndim_trace leading axes against the
/// ndim_trace trailing ones (equal dimensions and quantum numbers, opposite directions), on the GPU
/// (one workgroup per output block sums the diagonal entries of every contributing block). A trace
/// over every axis (degree-0 result) is not supported. Allocates r.
void block_sparse_tensor_cyclic_partial_trace(const struct block_sparse_tensor* t, const int ndim_trace, struct block_sparse_tensor* r);

//________________________________________________________________________________________________________________________
//
// chain operations and DMRG (replaces src/algorithm/chain_ops.c, src/algorithm/dmrg.c)

/// chain_ops.c:353-390 (include/algorithm/chain_ops.h:31): b = H_eff a on the GPU. Requires
/// outer (dummy) bond dimension 1 on l and r, and equal bra and ket axes (l axis 3 = a axis 0,
/// w axis 1 = a axis 1, r axis 2 = a axis 2: b has the block structure of a, as in every DMRG
/// call); other operands abort like the reference's assertions. b is allocated. Host buffers in, host buffers out
/// (the input tensors are copied to the device per call: use ct_matvec_plan_* to keep them resident).
void apply_local_hamiltonian(const struct block_sparse_tensor* a, const struct block_sparse_tensor* w, const struct block_sparse_tensor* l, const struct block_sparse_tensor* r, struct block_sparse_tensor* b);

/// dmrg.c:262-399 (include/algorithm/dmrg.h:12): two-site DMRG with device-resident MPS,
/// environments and Lanczos; truncating block SVD/QR in FP64 on the host.
int dmrg_twosite(const struct mpo* hamiltonian, const int num_sweeps, const int maxiter_lanczos, const double tol_split, const ct_long max_vdim,
	struct mps* psi, double* en_sweeps, double* entropy);

//________________________________________________________________________________________________________________________
//
// MPS, MPO and random numbers (replaces src/state/mps.c, src/operator/mpo.c, src/util/rng.c entry points)

/// rng.c:15-27 (include/util/rng.h:20)
void seed_rng_state(const uint64_t seed, struct rng_state* state);
/// rng.c:116-125 (rng.h:50)
double randn(struct rng_state* state);
/// mps.c:93-178 (include/state/mps.h:36)
void construct_random_mps(const enum numeric_type dtype, const int nsites, const ct_long d, const qnumber* qsite, const qnumber qnum_sector, const ct_long max_vdim, struct rng_state* rng_state, struct mps* mps);
/// mps.c:62-74 (mps.h:32)
void delete_mps(struct mps* mps);
/// mps.c:1219-1299 (mps.h:129): the MPS in an HDF5 file with the reference's layout (root attributes
/// nsites, qsite, qbond_0 .. qbond_nsites; datasets tensor_i = dense site tensors, float64). Real
/// double MPS only. Returns 0, or -1 (message on stderr) on any failure, including a build without
/// HDF5 support (libchemtensor_h5.so absent).
int save_mps(const char* filename, const struct mps* mps);
/// mps.c:1309-1461 (mps.h:131): read an MPS saved by save_mps (either library's); the caller frees it
/// with delete_mps. Returns 0, or -1 (message on stderr).
int load_mps(const char* filename, struct mps* mps);
/// mpo.c:182-196 (include/operator/mpo.h:52)
void delete_mpo(struct mpo* mpo);
/// The operator of construct_heisenberg_xxz_1d_mpo_assembly + mpo_from_assembly
/// (src/operator/hamiltonian.c:102-155, mpo.c:59-175) built directly as an MPO (bond dim 5).
void ct_construct_heisenberg_xxz_1d_mpo(const int nsites, const double J, const double D, const double h, struct mpo* mpo);
/// The operator of construct_fermi_hubbard_1d_mpo_assembly + mpo_from_assembly
/// (src/operator/hamiltonian.c:240-366, mpo.c:59-175) built directly as an MPO (bond dim 6).
void ct_construct_fermi_hubbard_1d_mpo(const int nsites, const double t, const double u, const double mu, struct mpo* mpo);
/// The operator of construct_spin_molecular_hamiltonian_mpo_assembly + mpo_from_assembly
/// (src/operator/hamiltonian.c:2755-2990, mpo.c:59-175): H = sum t_ij a+_{i s} a_{j s}
/// + 1/2 sum v_ijkl a+_{i s} a+_{j u} a_{l u} a_{k s} over nsites spatial orbitals, tkin (n x n) and
/// vint (n x n x n x n, physicists' convention) row-major; built directly as a finite-state-machine
/// MPO over the terms (bond states: left parts of one or two operators, three-operator sums indexed
/// by the fourth operator). Local basis and quantum numbers as the Fermi-Hubbard MPO.
void ct_construct_spin_molecular_hamiltonian_mpo(const int nsites, const double* tkin, const double* vint, struct mpo* mpo);
/// mpo.c:255-282 (mpo.h:56), on the GPU
void mpo_merge_tensor_pair(const struct block_sparse_tensor* a0, const struct block_sparse_tensor* a1, struct block_sparse_tensor* a);

//________________________________________________________________________________________________________________________
//
// device-resident matvec plans (no counterpart in the reference: the reference re-plans and
// re-allocates inside every block_sparse_tensor_dot call, bst.c:1826-2000)

typedef struct ct_matvec_plan ct_matvec_plan;

/// Precision / launch form of a plan.
enum ct_matvec_mode
{
	CT_MATVEC_PLAIN_FP64   = 0,  ///< reference structure: per-block FP64 GEMM launches + transpose / trace kernels
	CT_MATVEC_GROUPED_FP64 = 1,  ///< three persistent grouped FP64 MFMA launches, transposes folded into addressing
	CT_MATVEC_GROUPED_SPLIT = 2, ///< as GROUPED_FP64, operands split into FP32 terms on FP32 MFMA with FP64 accumulation
	CT_MATVEC_FUSED_FP64   = 3,  ///< one launch per matvec: intermediates of a right-bond column chunk stay in LDS (FP64)
	CT_MATVEC_PAIR_FP64    = 4,  ///< one launch for the three contractions of every (left sector, right sector) block pair in LDS (FP64 MFMA) + one partial-sum launch
};

/// Build a plan for H_eff = apply_local_hamiltonian(., w, l, r) on vectors with the block
/// structure of 'a' (its entries are ignored). w, l and r are uploaded and stay resident.
int ct_matvec_plan_create(const struct block_sparse_tensor* a, const struct block_sparse_tensor* w, const struct block_sparse_tensor* l, const struct block_sparse_tensor* r,
	const int mode, ct_matvec_plan** plan);
int ct_matvec_plan_destroy(ct_matvec_plan* plan);
/// Host-only check of the tables a grouped plan would launch (no device needed): every operand index
/// inside its buffer and every entry of the intermediates and of the result written by exactly one
/// output, for the grouped stages, the column-stream stage 2 and the fused stages 1 and 2.
/// Returns 0 when valid, a positive check number otherwise, or a CT_ERR_* code.
int ct_matvec_plan_check(const struct block_sparse_tensor* a, const struct block_sparse_tensor* w, const struct block_sparse_tensor* l,
	const struct block_sparse_tensor* r);
/// Host-only replay of the sector-pair tables (CT_MATVEC_PAIR_FP64) with the kernels' index
/// arithmetic and tile order (no device needed): operand, LDS and partial indices in bounds, LDS
/// operands read only after they were written, every intermediate, partial and result entry written
/// exactly once. mfma_cap (<= 0: the default) bounds the MFMAs of one unit. y (optional, length
/// block_sparse_tensor_num_elements_blocks(a)) receives H_eff a as computed by the replay.
/// Returns 0 when valid, a positive check number otherwise, or a CT_ERR_* code.
int ct_matvec_pair_check(const struct block_sparse_tensor* a, const struct block_sparse_tensor* w, const struct block_sparse_tensor* l,
	const struct block_sparse_tensor* r, const int mfma_cap, double* y);
/// Host-only statistics of the sector-pair tables: [0] units, [1] partial-block doubles, [2] sum and
/// [3] largest unit cost (MFMAs + staging), [4] largest unit LDS (doubles), [5] largest record (words)
int ct_matvec_pair_stats(const struct block_sparse_tensor* a, const struct block_sparse_tensor* w, const struct block_sparse_tensor* l,
	const struct block_sparse_tensor* r, const int mfma_cap, double* stats6);
/// SVD of one dense matrix a (m x n, row-major, host memory) on the device, with the kernel that
/// factorises the blocks of the DMRG split (one-sided Jacobi in the LDS of one workgroup, the
/// device counterpart of the block loop of block_sparse_tensor_svd, block_sparse_tensor.c:2686-2848):
/// u (m x k), s (k, descending), vh (k x n), k = min(m, n), row-major host arrays.
/// Returns 0, -1 when the sweeps did not converge, CT_ERR_UNSUPPORTED when the matrix does not fit
/// the LDS of one workgroup ((max(m,n) + min(m,n)) min(m,n) doubles).
int ct_dense_svd_device(const ct_long m, const ct_long n, const double* a, double* u, double* s, double* vh);
/// One dense matrix through the default device factorisation of the DMRG split (the Gram matrix of
/// the isometric side diagonalised by Jacobi rotations in LDS, the other factor by projection): with
/// left != 0, a = u vh with u (m x k) orthonormal and vh (k x n) = s Vh; with left == 0, vh (k x n)
/// orthonormal and u = U s. s (k) descending, k = min(m, n); row-major host arrays. Returns 0, -1 when
/// the sweeps did not converge, CT_ERR_UNSUPPORTED when the Gram matrix (m x m or n x n) does not fit
/// the LDS of one workgroup.
int ct_dense_gram_factor_device(const ct_long m, const ct_long n, const double* a, const int left, double* u, double* s, double* vh);
/// length of the packed vectors (= block_sparse_tensor_num_elements_blocks(a))
ct_long ct_matvec_plan_dim(const ct_matvec_plan* plan);
/// algorithmic FP64 flops of one matvec: sum over the reference's GEMM calls of 2 m n k
double ct_matvec_plan_flops(const ct_matvec_plan* plan);
/// algorithmic bytes of one matvec: 8 (|a| + |b| + |w| + |l| + |r|)
double ct_matvec_plan_bytes(const ct_matvec_plan* plan);
/// plan statistics: [0] GEMM terms (stage 1,2,3 summed), [1] output tasks, [2] work items, [3] intermediate doubles
int ct_matvec_plan_stats(const ct_matvec_plan* plan, double* stats4);
/// per-stage algorithmic flops (stage 1: theta.R, 2: W.T1, 3: L^T.T2) of the grouped forms
int ct_matvec_plan_stage_flops(const ct_matvec_plan* plan, double* flops3);
/// per-stage algorithmic bytes: 8 (|inputs| + |output|) of each stage's launch
int ct_matvec_plan_stage_bytes(const ct_matvec_plan* plan, double* bytes3);
/// launch form chosen per stage (0 wave/simple, 1 wave/stream, 2 workgroup/chunk, 3 workgroup/stream,
/// 4 column stream of stage 2, 5 stages 1 and 2 fused in one launch, 6 LDS-tiled 64 x 64 workgroup
/// tiles for the large tasks and the best of 0-3 for the others)
/// and the autotuner's time (ms) of every form, [stage * 4 + form]; 0 where not timed
int ct_matvec_plan_forms(const ct_matvec_plan* plan, int* forms3, float* tuned_ms12);
/// average device time (ms) of one stage (0, 1, 2; the fused form has only stage 0) over 'reps'
/// back-to-back launches on the given vectors: the per-launch duration of that stage's kernel
int ct_matvec_plan_stage_time(ct_matvec_plan* plan, const int stage, const double* d_x, double* d_y, const int reps, float* ms);
/// y = H_eff x on device vectors (packed in serialize order) on the plan's stream
int ct_matvec_apply(ct_matvec_plan* plan, const double* d_x, double* d_y);
/// As ct_matvec_apply, without any allocation (no counterpart in the reference: for callers that
/// cannot allocate device memory, such as code under HIP graph capture or tight serving loops). Only
/// CT_MATVEC_PAIR_FP64 plans (CT_ERR_UNSUPPORTED otherwise): the partial blocks go to a file-scope
/// device array of 2^22 doubles (32 MiB) that each GPU holds until the process ends, instead of the
/// plan's own buffer; a plan needing more returns CT_ERR_NOMEM. Calls that overlap on one GPU share
/// the array and race: serialise them (for example on one stream). Nothing in the array needs
/// clearing: each call writes every partial entry before reading it.
int ct_matvec_apply_noalloc(ct_matvec_plan* plan, const double* d_x, double* d_y);
/// as ct_matvec_apply, recording the device time (ms) of each of the three stages
int ct_matvec_apply_timed(ct_matvec_plan* plan, const double* d_x, double* d_y, float* stage_ms3);
/// the plan's HIP stream (hipStream_t) as an opaque pointer
void* ct_matvec_plan_stream(ct_matvec_plan* plan);

//________________________________________________________________________________________________________________________
//
// batches of independent H_eff problems (no counterpart in the reference, which applies one H_eff
// at a time): one fused launch applies all of them to the concatenation of their packed vectors,
// and the batched Lanczos runs one recurrence per problem with the matvecs of a step in that launch

typedef struct ct_matvec_batch ct_matvec_batch;

/// problems p = 0 .. nprob-1 as in ct_matvec_plan_create (fused FP64 form); operands are uploaded
int ct_matvec_batch_create(const int nprob, const struct block_sparse_tensor* const* a, const struct block_sparse_tensor* const* w,
	const struct block_sparse_tensor* const* l, const struct block_sparse_tensor* const* r, ct_matvec_batch** batch);
/// the same with a precision form: CT_MATVEC_GROUPED_FP64 (as ct_matvec_batch_create: grouped or
/// fused FP64, the faster timed at creation) or CT_MATVEC_GROUPED_SPLIT (stages 1 and 3 on FP32 MFMA
/// with exactly sliced FP64 operands, stage 2 the FP64 column stream); other modes: CT_ERR_UNSUPPORTED
int ct_matvec_batch_create_mode(const int nprob, const struct block_sparse_tensor* const* a, const struct block_sparse_tensor* const* w,
	const struct block_sparse_tensor* const* l, const struct block_sparse_tensor* const* r, const int mode, ct_matvec_batch** batch);
int ct_matvec_batch_destroy(ct_matvec_batch* batch);
/// total length of the concatenated vectors
ct_long ct_matvec_batch_dim(const ct_matvec_batch* batch);
/// offsets[0 .. nprob]: problem p occupies [offsets[p], offsets[p+1]) of the concatenated vectors
int ct_matvec_batch_offsets(const ct_matvec_batch* batch, ct_long* offsets);
/// algorithmic flops (reference GEMM count) and bytes of one batched matvec, summed over problems
double ct_matvec_batch_flops(const ct_matvec_batch* batch);
double ct_matvec_batch_bytes(const ct_matvec_batch* batch);
/// form chosen by timing both at creation (1: one fused launch, 0: three grouped launches) and
/// the measured ms per batched matvec of each ([0] grouped, [1] fused; 0 if not available)
int ct_matvec_batch_form(const ct_matvec_batch* batch, int* fused, float* ms2);
void* ct_matvec_batch_stream(ct_matvec_batch* batch);
/// y_p = H_eff,p x_p for every problem, one launch
int ct_matvec_batch_apply(ct_matvec_batch* batch, const double* d_x, double* d_y);
/// Lanczos ground state of every problem (krylov.c:172-251, numeig = 1): lambda[nprob],
/// status[nprob] (0, -1 breakdown before one iteration, -2 eigensolver failure; may be NULL),
/// d_u the concatenated Ritz vectors. Returns 0 or the first nonzero status.
int ct_lanczos_ground_state_batch(ct_matvec_batch* batch, const double* d_vstart, const int maxiter, double* lambda, int* status, double* d_u);

//________________________________________________________________________________________________________________________
//
// serving loop (no counterpart in the reference): independent ground-state solves are admitted by
// an HBM budget, stay device-resident, and the matvecs of all running solves of a Lanczos step are
// applied in one batched launch (continuous batching)

typedef struct ct_serve ct_serve;

struct ct_serve_request
{
	int model;              ///< 0: spin-1/2 XXZ chain (param J, D, h), 1: Fermi-Hubbard chain (param t, U, mu)
	int nsites;
	double param[3];
	qnumber sector;         ///< quantum number sector of the state (XXZ: 2 Sz; Hubbard: N * 65536 + 2 Sz, as encode_quantum_number_pair)
	int pad;
	ct_long max_vdim;       ///< maximum virtual bond dimension
	int num_sweeps;         ///< maximum number of sweeps
	int maxiter_lanczos;
	double tol_split;       ///< SVD truncation tolerance
	double tol_energy;      ///< stop after a sweep whose energy changed by less than this (0: always num_sweeps)
	uint64_t seed;          ///< seed of the start state (construct_random_mps)
};

struct ct_serve_result
{
	int status;             ///< 0 done, 1 queued or running, < 0 failed
	int sweeps;             ///< sweeps performed
	double energy;          ///< energy of the last sweep
	double energy_change;   ///< |E(last sweep) - E(previous sweep)| (NaN after one sweep)
	double t_submit, t_admit, t_done;  ///< seconds since the server was created (-1: not yet)
};

/// start a server with a device-memory budget (bytes) for admitted requests and at most
/// max_active requests in flight; the server owns a worker thread and a HIP stream
int ct_serve_create(const double hbm_budget_bytes, const int max_active, ct_serve** server);
/// queue a request; returns its ticket
int ct_serve_submit(ct_serve* server, const struct ct_serve_request* req, int64_t* ticket);
/// state of a request: returns 0 when it has finished (res->status 0 or < 0), 1 while queued or
/// running; with wait != 0 blocks until it has finished
int ct_serve_poll(ct_serve* server, const int64_t ticket, const int wait, struct ct_serve_result* res);
/// energies of the finished sweeps of a request (up to capacity); returns their number
int ct_serve_energies(ct_serve* server, const int64_t ticket, double* en_sweeps, const int capacity);
/// [0] ticks, [1] batched Lanczos solves, [2] finished requests, [3] queued requests, [4] reserved
/// bytes, [5] largest batch, [6] busy seconds of the worker, [7] seconds since creation; seconds
/// spent in [8] starting requests, [9] batch plans, [10] batched Lanczos, [13] the splits and
/// environment updates (wall time; they run in parallel host threads), of which per request summed
/// [11] device work and [12] host block SVDs
int ct_serve_stats(ct_serve* server, double* stats14);
/// stop the worker (queued and running requests are dropped) and free everything
int ct_serve_destroy(ct_serve* server);

/// include/util/krylov.h:10
typedef void lanczos_linear_func_d(const ct_long n, const void* data, const double* v, double* ret);

/// include/util/krylov.h:24 (src/util/krylov.c:172-251): Lanczos eigensystem of the symmetric linear
/// map 'afunc' (host callback on host vectors). The Krylov vectors and the recurrence live on the
/// GPU; every step copies v_j to the host for 'afunc' and the result back. lambda[numeig] ascending,
/// u_ritz (n x numeig, row-major) the Ritz vectors. Returns 0, -1 (fewer iterations than numeig),
/// -2 (tridiagonal eigensolver failure) or a CT_ERR_* code.
int eigensystem_krylov_symmetric(const ct_long n, lanczos_linear_func_d afunc, const void* adata,
	const double* vstart, const int maxiter, const int numeig, double* lambda, double* u_ritz);

/// Lanczos ground state of the plan's H_eff (krylov.c:172-251 with numeig = 1) on the device:
/// d_vstart / d_u are device vectors of length ct_matvec_plan_dim. Returns 0, -1 (breakdown
/// before one iteration) or -2 (tridiagonal eigensolver failure).
int ct_lanczos_ground_state(ct_matvec_plan* plan, const double* d_vstart, const int maxiter, double* lambda, double* d_u);

//________________________________________________________________________________________________________________________
//
// device utilities for callers without their own HIP runtime bindings

int ct_device_alloc(void** ptr, const ct_long nbytes);
int ct_device_free(void* ptr);
int ct_memcpy_h2d(void* dst, const void* src, const ct_long nbytes);
int ct_memcpy_d2h(void* dst, const void* src, const ct_long nbytes);
int ct_device_synchronize(void);
/// number of visible GPUs (0 when none)
int ct_device_count(void);
/// dst[0:n] = src[0:n] (device to device) on the library's stream
int ct_device_copy(double* dst, const double* src, const ct_long n);
int ct_stream_synchronize(void* stream);
/// HIP events (opaque hipEvent_t) for timing work on a plan's stream
int ct_event_create(void** ev);
int ct_event_record(void* ev, void* stream);
/// waits for ev1, then returns the time between the two events in ms
int ct_event_elapsed_ms(void* ev0, void* ev1, float* ms);
int ct_event_destroy(void* ev);

//________________________________________________________________________________________________________________________
//
// seeded synthetic inputs (bit-identical to the CPU restatement used by the tests)

/// Fill every block of t with alpha * N(0,1) draws, blocks in row-major block order (the order
/// block_sparse_tensor_fill_random_normal of src/tensor/block_sparse_tensor.c:498 draws in).
void ct_fill_random_normal(struct block_sparse_tensor* t, double alpha, struct rng_state* rng);
/// Virtual-bond dimensions and quantum numbers exactly as construct_random_mps (src/state/mps.c)
/// draws them for this seed; dims_out has nsites+1 entries, qbonds_out the concatenated bonds.
void ct_random_bond_qnums(const int nsites, const ct_long d, const qnumber* qsite, const qnumber sector, const ct_long max_vdim, const uint64_t seed, ct_long* dims_out, qnumber* qbonds_out);
/// One H_eff problem at the centre bond i = nsites/2 - 1 of a chain (model 0: XXZ J=1 D=0.5 h=0,
/// sector 0; model 1: Fermi-Hubbard t=1 U=4 mu=0, half filling): bond quantum numbers as
/// construct_random_mps(seed), theta / left / right N(0,1) with seeds seed+1 / +2 / +3, h2 the merged
/// MPO pair. The caller frees the four tensors with delete_block_sparse_tensor.
int ct_generate_matvec_problem(const int model, const int nsites, const ct_long chi, const uint64_t seed,
	struct block_sparse_tensor* theta, struct block_sparse_tensor* h2, struct block_sparse_tensor* left, struct block_sparse_tensor* right);
/// The same synthetic H_eff problem for any MPO H (the spin molecular Hamiltonian of
/// ct_construct_spin_molecular_hamiltonian_mpo, for instance): centre bond i = nsites/2 - 1, bond
/// quantum numbers as construct_random_mps(H's site quantum numbers, sector, chi, seed) draws them,
/// theta / left / right N(0,1) with seeds seed+1 / +2 / +3, h2 = H.a[i] merged with H.a[i+1].
int ct_generate_matvec_problem_mpo(const struct mpo* H, const qnumber sector, const ct_long chi, const uint64_t seed,
	struct block_sparse_tensor* theta, struct block_sparse_tensor* h2, struct block_sparse_tensor* left, struct block_sparse_tensor* right);

#ifdef __cplusplus
}
#endif
