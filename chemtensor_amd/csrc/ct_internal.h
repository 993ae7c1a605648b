// Internal declarations of libchemtensor_gpu: host-side block bookkeeping, device tensors,
// grouped-GEMM descriptors and the matvec plan. Host code is C++17; device code is HIP for
// gfx950 (64-wide wavefronts, FP64/FP32 MFMA).

#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../include/chemtensor_gpu.h"

namespace ctg {

//________________________________________________________________________________________________________________________
//
// errors

struct HipError
{
	hipError_t code;
	std::string where;
};

void hip_check(hipError_t e, const char* what, const char* file, int line);
#define CTG_HIP(x) ::ctg::hip_check((x), #x, __FILE__, __LINE__)

//________________________________________________________________________________________________________________________
//
// block bookkeeping (the reference's allocate_block_sparse_tensor, bst.c:37-150)

struct Meta
{
	int ndim = 0;
	std::vector<ct_long> dim_logical;
	std::vector<int> axis_dir;
	std::vector<std::vector<qnumber>> qnums_logical;
	std::vector<std::vector<qnumber>> qnums_blocks;   // sorted ascending, distinct
	std::vector<std::vector<ct_long>> block_count;    // multiplicity of each block quantum number
	std::vector<ct_long> dim_blocks;
	ct_long nblocks = 1;                               // prod(dim_blocks) (1 for ndim 0)
	std::vector<ct_long> offset;                       // per flat block index: packed offset or -1
	ct_long nelem = 0;                                 // number of stored entries

	bool present(ct_long k) const { return offset[k] >= 0; }
	void block_index(ct_long k, ct_long* idx) const;
	ct_long flat_index(const ct_long* idx) const;
	void block_dims(ct_long k, ct_long* dims) const;
	ct_long block_size(ct_long k) const;
	ct_long find_qnum(int ax, qnumber q) const;
};

Meta make_meta(int ndim, const ct_long* dim, const int* axis_dir, const qnumber* const* qnums);
Meta meta_of(const block_sparse_tensor* t);
/// logical indices (in order) whose quantum number equals the block quantum number 'q' on axis 'ax'
std::vector<ct_long> fanout(const Meta& m, int ax, qnumber q);

//________________________________________________________________________________________________________________________
//
// device memory: caching allocator and device tensors

void* dev_alloc(size_t nbytes);
// enqueue a host-to-device copy of host data on s; the host data may be freed on return
void copy_h2d_async(void* dst, const void* src, size_t nbytes, hipStream_t s);
// several host arrays into one device buffer with one transfer: part i lands at dst + dst_off
struct H2DPart { const void* src; size_t nbytes; size_t dst_off; };
void copy_h2d_parts(void* dst, const H2DPart* parts, int nparts, size_t total, hipStream_t s);
void dev_free(void* p);
void dev_trim();
// after the calling thread has synchronised its stream: its cached blocks become available to
// every thread
void dev_cache_to_shared();
// marks a stream capture on the calling thread: the allocator does not trim its cache (a device
// synchronisation) meanwhile
struct CaptureScope
{
	CaptureScope();
	~CaptureScope();
};
// release the calling thread's pinned staging blocks and its cached device blocks (a long-lived
// helper thread calls it before it exits, once its stream has completed)
void release_thread_resources();
size_t dev_bytes_in_use();

struct DevTensor
{
	Meta meta;
	double* d = nullptr;   // packed blocks in serialize order
	bool owner = true;     // false for views sharing another tensor's entries

	DevTensor() = default;
	explicit DevTensor(Meta m);
	/// non-owning view of 'o' (same entries) with its own metadata
	static DevTensor view(const DevTensor& o, Meta m);
	DevTensor(const DevTensor&) = delete;
	DevTensor& operator=(const DevTensor&) = delete;
	DevTensor(DevTensor&& o) noexcept;
	DevTensor& operator=(DevTensor&& o) noexcept;
	~DevTensor();
	void zero(hipStream_t s);
};

hipStream_t default_stream();
DevTensor upload(const block_sparse_tensor* t, hipStream_t s);
void download(const DevTensor& t, block_sparse_tensor* out, hipStream_t s);

/// a block-sparse tensor in packed form in host memory
struct HostPacked
{
	Meta meta;
	std::vector<double> v;
};
HostPacked download_packed(const DevTensor& t, hipStream_t s);
DevTensor upload_packed(const HostPacked& h, hipStream_t s);

// layout of the block SVD of a packed block-sparse matrix (block_sparse_tensor.c:2686-2848): the new
// bond's quantum numbers (one per singular value, blocks in row-major order), the metas of u (m x ni)
// and vh (ni x n) and one job per present block (offsets in doubles; s_off: the block's first
// singular value); 'dummy' when no block is present
struct SvdLayout
{
	std::vector<qnumber> qi;
	bool dummy = false;
	Meta um, vhm;
	struct Job { ct_long a_off, u_off, vh_off, s_off, r, c; };
	std::vector<Job> jobs;
};
SvdLayout plan_block_svd(const Meta& m);

// device block SVD (split.hip): one workgroup per block, one-sided Jacobi in LDS
struct DevSvdJob { int64_t a, u, vh, s; int32_t r, c; };
// per-block selection of columns (axis 1) or rows (axis 0) of a packed matrix, optionally scaled
struct SliceBlock { int64_t src, dst; int32_t rows, cols, src_ld, axis, idx0, pad; };
size_t block_svd_lds_bytes(const SvdLayout& lay);
size_t block_svd_lds_limit();
// u, vh: packed as lay.um / lay.vhm; s: one value per new-bond index, in device column order within
// each block (not sorted); status: one int per block (0 converged, -1 not). False (nothing launched)
// when a block does not fit the LDS of one workgroup.
bool launch_block_svd(const SvdLayout& lay, const double* a, double* u, double* vh, double* s, int* status, hipStream_t st);
// device block factorisation through the Gram matrix of the isometric side (rows when 'left'): u, vh
// packed as lay.um / lay.vhm with the singular values already in the non-isometric factor, s the norms
// of its rows / columns (device column order, not sorted); status as launch_block_svd. False when a
// Gram matrix does not fit.
size_t block_eig_lds_bytes(const SvdLayout& lay, bool left);
size_t block_eig_lds_limit();
bool launch_block_eig(const SvdLayout& lay, const double* a, double* u, double* vh, double* s, int* status, bool left, hipStream_t st);
// the slice of a packed matrix keeping logical indices 'ind' of axis i_ax (as host_slice), reading
// logical index x from within-block position srcpos[x]; scale (indexed like ind) multiplies the kept
// rows / columns when given
void plan_slice_2d(const Meta& m, int i_ax, const std::vector<ct_long>& ind, const std::vector<ct_long>& srcpos, const std::vector<double>* scale,
	Meta& out, std::vector<SliceBlock>& blocks, std::vector<int>& idx, std::vector<double>& sc);
DevTensor device_slice_2d(const DevTensor& src, const Meta& out, const std::vector<SliceBlock>& blocks, const std::vector<int>& idx,
	const std::vector<double>& sc, bool scaled, hipStream_t st);

// host block factorisations (host_bond.cpp)
// block SVD of a packed block-sparse matrix; iso < 0: one-sided Jacobi (u, vh isometric), iso 0 / 1:
// host_svd_factor per block (vh = diag(s) Vh, resp. u = U diag(s))
int host_block_svd(const HostPacked& a, HostPacked& u, std::vector<double>& s, HostPacked& vh, int iso = -1);
// blocks of the calling thread's splits without helper threads (a caller whose threads already
// run independent splits on every core)
void host_svd_serial_on_this_thread(bool on);
ct_long host_retained_bond_indices(const std::vector<double>& sigma, double tol, bool relative, ct_long max_vdim, std::vector<ct_long>& ind, trunc_info& info);
HostPacked host_slice(const HostPacked& t, int i_ax, const std::vector<ct_long>& ind);
void host_scale_axis(HostPacked& t, bool leading, const std::vector<double>& vec);
int host_split_svd(const HostPacked& a, double tol, ct_long max_vdim, bool renormalize, int distr, HostPacked& a0, HostPacked& a1, trunc_info& info);
int host_block_qr(const HostPacked& a, HostPacked& q, HostPacked& r);
int host_block_rq(const HostPacked& a, HostPacked& r, HostPacked& q);

//________________________________________________________________________________________________________________________
//
// grouped GEMM descriptors
//
// A task is one output matrix C (M x N) = sum over its terms of A_t (M x K_t) B_t (K_t x N)
// (or C += ... when 'accumulate'). Column j of a task is split as (jb, jc) = (j / nci, j % nci);
// addresses:  A(i, k) = A[i sa_i + k sa_k],  B(k, j) = B[jb sb_batch + jc sb_j + k sb_k],
//             C(i, j) = C[jb sc_batch + jc + i sc_i].

struct GemmTerm
{
	const double* A;
	const double* B;
	int64_t sb_batch;
	int32_t K;
	int32_t sa_i;
	int32_t sa_k;
	int32_t sb_k;
	int32_t sb_j;
	int32_t pad;
};

struct GemmTask
{
	double* C;
	int64_t sc_batch;
	int32_t M;
	int32_t N;
	int32_t nci;
	int32_t sc_i;
	int32_t t0;
	int32_t t1;
	int32_t accumulate;
	int32_t pad;
};

struct WorkItem
{
	GemmTask task;         // the task descriptor travels with the item (one scalar load per item)
	int32_t m0;
	int32_t n0;
	int32_t cost;
	int32_t task_index;
	GemmTerm first;        // the task's first term: its operand loads need no descriptor round trip
};

// tile edge of one wavefront's output tile (2 x 2 MFMA 16x16 sub-tiles)
constexpr int TILE = 32;

/// host-built list of tasks/terms for one grouped launch
struct GemmBatch
{
	std::vector<GemmTask> tasks;
	std::vector<GemmTerm> terms;
	double flops = 0;
	void add_task(double* C, int M, int N, int nci, int sc_i, int64_t sc_batch, bool accumulate);
	void add_term(const double* A, const double* B, int K, int sa_i, int sa_k, int sb_k, int sb_j, int64_t sb_batch);
	std::vector<WorkItem> work_items() const;
};

/// launch forms of the grouped FP64 GEMM (chosen per stage by the plan-time autotuner)
enum GemmForm
{
	GEMM_WAVE_SIMPLE = 0,  ///< one wavefront per 32x32 tile, one k-step per load round trip
	GEMM_WAVE_STREAM = 1,  ///< one wavefront per tile, loads of 8 k-steps batched across terms
	GEMM_WG_CHUNK    = 2,  ///< one workgroup per tile, chunks of 4 k-steps dealt round-robin to its wavefronts
	GEMM_WG_STREAM   = 3,  ///< one workgroup per tile, batches of 8 k-steps dealt round-robin
	GEMM_NUM_FORMS   = 4,
};

/// a GemmBatch resident on the device
struct DevGemm
{
	char* buf = nullptr;   // one device allocation holding the three tables
	GemmTask* tasks = nullptr;
	GemmTerm* terms = nullptr;
	WorkItem* items = nullptr;
	int ntasks = 0, nterms = 0, nitems = 0;
	int grid = 0;
	int form = 1;          // one of GemmForm
	double avg_steps = 0;  // mean k-steps per work item
	double flops = 0;
	// LDS-tiled form for the large tasks (both edges >= LDS_TASK_MIN): 64 x 64 tiles, one workgroup
	// each (litems), the other tasks' 32 x 32 tiles (ritems) in the chosen form; used when use_lds
	WorkItem* litems = nullptr;
	WorkItem* ritems = nullptr;
	int nlitems = 0, nritems = 0;
	bool use_lds = false;
	void upload(const GemmBatch& b, hipStream_t s);
	void release();
};
constexpr int LDS_TILE = 64;       // output tile edge of the LDS-tiled form
void split_lds_items(const GemmBatch& b, const std::vector<WorkItem>& it, std::vector<WorkItem>& lit, std::vector<WorkItem>& rit);
constexpr int LDS_TASK_MIN = 48;   // tasks with M and N at least this take the LDS-tiled form

// Operand pointers in tasks/terms are byte offsets added to per-launch base pointers
// (absolute addresses when the base is null), so one descriptor list serves any input vector.
struct GemmBases { const double* A; const double* B; double* C; };

// kernel launchers (kernels.hip)
void launch_grouped_gemm_f64(const DevGemm& g, GemmBases b, hipStream_t s);
void launch_grouped_gemm_split(const DevGemm& g, GemmBases b, hipStream_t s);
void launch_plain_gemm_f64(const GemmTask& task, const GemmTerm& term, GemmBases b, hipStream_t s);
void launch_permute(const double* src, double* dst, int ndim, const ct_long* dims, const int* perm, hipStream_t s);
struct PermBlock { int64_t src; int64_t dst; int64_t n; int64_t ddim[6]; int64_t sstride[6]; int32_t ndim; int32_t pad; };
void launch_permute_blocks(const double* src, double* dst, const PermBlock* d_blocks, int nblocks, hipStream_t s);
struct CopySeg { int64_t src; int64_t dst; int64_t len; };
void launch_copy_segments(const double* src, double* dst, const CopySeg* d_segs, int nsegs, hipStream_t s);
void launch_fill(double* p, double v, ct_long n, hipStream_t s);
void launch_copy(const double* src, double* dst, ct_long n, hipStream_t s);
int num_compute_units();

//________________________________________________________________________________________________________________________
//
// device tensor algebra (engine.hip)

enum class DotForm { Grouped, Plain };

/// r = s . t (bst.c:1826-2000) with r's block table planned on the host
DevTensor dev_dot(const DevTensor& s, int axrange_s, const DevTensor& t, int axrange_t, int ndim_mult, DotForm form, hipStream_t st, double* flops = nullptr);
DevTensor dev_transpose(const int* perm, const DevTensor& t, hipStream_t st);
DevTensor dev_cyclic_partial_trace(const DevTensor& t, int ndim_trace, hipStream_t st);
DevTensor dev_flatten_axes(const DevTensor& t, int i_ax, int new_dir, hipStream_t st);
DevTensor dev_split_axis(const DevTensor& t, int i_ax, const ct_long* new_dims, const int* new_dirs, const qnumber* const* new_qnums, hipStream_t st);

//________________________________________________________________________________________________________________________
//
// matvec plan

// Stage 2 as a column stream: for a block pair (A,C) the contraction T2(A,M,Q,C) = sum W T1(A,P,N,C)
// applies one small matrix W_D (rows (M,Q,m,q) with qM + qQ = D, columns (P,N,p,n) with
// qP + qN = D, D = qC + qo - qA) to every column (a,c) of the pair. A wavefront takes 64 columns:
// it loads their K inputs (T1 read once), and writes their R outputs (T2 written once); W_D comes
// from a dense, zero-padded copy through scalar loads.
constexpr int WA_KMAX = 32;
struct WaUnit { int64_t kb, rb, wrow, wcol; int dA, dC, K, R; };   // kb / rb index kbase / rbase; wrow indexes wmat rows, wcol wmatT columns
struct WaItem { WaUnit u; int unit, col0; };                 // the unit travels with the item (one scalar load)
// Stages 1 and 2 fused (no T1 in HBM): a wavefront takes a tile (a0..a0+ta, c0..c0+tc) of a block
// pair, stages the rows of theta(A,P,B) and the columns of R(B,N,C) of every column group (P,N) of
// the pair's class in LDS, forms each T1 value on the fly and applies W_D as in the column stream.
struct F12Group { int64_t th_off, r_off; int dP, dN, dB, pad; };   // th_off / r_off -1: zero T1 block
struct F12Item { WaUnit u; int g0, ng, a0, ta, c0, tc, lds, pad; };
constexpr int F12_LDS = 1024;    // doubles of LDS staging per wavefront (8 KB: 20 wavefronts per CU)
constexpr int F12_RMAX = 16;     // largest number of W_D rows

struct WApplyHost
{
	std::vector<WaUnit> units;
	std::vector<WaItem> items;
	std::vector<F12Group> groups;
	std::vector<F12Item> items12;
	bool f12_ok = true;                  // every tile fits F12_LDS and every class has at most F12_RMAX rows
	std::vector<int64_t> kbase, rbase;   // element offset of column k / row r at a = 0, c = 0 (-1: absent block)
	std::vector<int> kstride, rstride;   // element stride of a
	std::vector<int64_t> widx;           // W_D entries as W element indices (-1: zero), WA_KMAX per row
	std::vector<int64_t> widxT;          // the same, transposed: F12_RMAX per column (k), WA_KMAX columns per class
	int kmax = 0;                        // largest K of any class
};
struct WApply
{
	WaUnit* units = nullptr;
	WaItem* items = nullptr;
	F12Group* groups = nullptr;
	F12Item* items12 = nullptr;
	int nitems12 = 0;
	bool f12_ok = false;
	int64_t* kbase = nullptr;
	int64_t* rbase = nullptr;
	int* kstride = nullptr;
	int* rstride = nullptr;
	double* wmat = nullptr;
	double* wmatT = nullptr;             // W_D columns (F12_RMAX entries each) for the fused stages
	int nitems = 0;
	int kmax = 0;
	bool ok = false;                     // every class fits WA_KMAX columns
	void upload(const WApplyHost& h, const double* w, hipStream_t s);
	void release();
};
void launch_wapply(const WApply& wa, const double* t1, double* t2, hipStream_t s);
// stages 1 and 2 fused: T2 from theta (x) and R directly
void launch_f12(const WApply& wa, const double* x, const double* r, double* t2, hipStream_t s);

// Fused single-launch FP64 matvec. One workgroup owns a "unit": k consecutive columns c of one
// right-bond sector C of the output. It stages the R columns it needs in LDS, computes
// T1(A,P,N,C)[:,:,:,c] = theta(A,P,B) R(B,N,C)[:,:,c] and T2(A,M,Q,C)[:,:,:,c] = W T1 into LDS and
// writes y(X,Q,C)[:,:,c] = sum L^T T2. LDS offsets are stored per column (multiplied by k at run time).
struct FzCopy  { int64_t goff; int rowstart; int dB, dN, dC, rb, nb, j0; };         // R(B,N,C) -> LDS panel of B
struct FzTask1 { int64_t th_off; int rowstart, dB, nb, nJ, rb, t1; };               // T1 rows (a p) of (A,P), j-chunks of 4
struct FzTask2 { int rowstart, dMQ, t2, u0, u1, pad; };                              // T2 rows (a m q) of (A,M,Q)
struct FzTerm2 { int64_t w_off; int dP, dN, t1, nb, j0, pad; };                      // W(M,Q,P,N) x T1(A,P)[.., j0 + n]
struct FzTask3 { int64_t y_off; int rowstart, dX, dQ, dC, u0, u1; };                 // y rows (x q) of (X,Q,C)
struct FzTerm3 { int64_t l_off; int dAM, t2; };                                      // L(0,A,M,X)^T x T2(A,M,Q)
struct FzSector { int m0, n0, m1, rows1, m2, rows2, m3, rows3, split, lds_pc, dC, pad; };
struct FzUnit  { int sector, c0, k, pad; };

struct FusedPlan
{
	FzSector* sectors = nullptr;
	FzUnit* units = nullptr;
	int* order = nullptr;       // blockIdx -> unit (-1: idle), XCD-aware
	int* map = nullptr;         // row -> task of every phase, per sector
	FzCopy* copies = nullptr;
	FzTask1* t1 = nullptr;
	FzTask2* t2 = nullptr;
	FzTerm2* u2 = nullptr;
	FzTask3* t3 = nullptr;
	FzTerm3* u3 = nullptr;
	int grid = 0, nunits = 0;
	size_t lds_bytes = 0;
	double flops = 0;
	void release();
};

// Plan the fused form; returns false when a sector's per-column working set exceeds the LDS budget.
bool build_fused_plan(FusedPlan& f, const Meta& theta, const Meta& w, const Meta& l, const Meta& r, hipStream_t s);
// one problem of a batched fused plan: its metadata and the element offsets of its operands in the
// concatenated theta / y, W, L and R buffers
struct FusedProblem { const Meta* theta; const Meta* w; const Meta* l; const Meta* r; int64_t x_off, w_off, l_off, r_off; };
bool build_fused_batch(FusedPlan& f, const std::vector<FusedProblem>& probs, hipStream_t s);
void launch_fused_matvec(const FusedPlan& f, const double* x, const double* w, const double* l, const double* r, double* y, hipStream_t s);

// Sector-pair form (pair.hip): one workgroup per chunk (rows a0..a0+ta, columns c0..c0+tc) of a
// block pair (A, C) of the left and right bond sectors runs all three contractions in LDS and writes
// its contribution to each output block y(X,Q,C) as a partial block; a reduction launch sums the
// partial blocks of every y block in (A, row chunk) order.
constexpr int PR_THREADS = 256;
constexpr int PR_XCDS = 8;                            // XCDs of the device: workgroup i runs on XCD i mod 8
constexpr int PR_LDS = 4608;                          // doubles of operand / intermediate LDS per workgroup (36 KB)
constexpr int PR_REC_WORDS = 512;                     // largest unit record (4-byte words)
constexpr int PR_RMAX = 32, PR_KMAX = 32;             // largest class W_D (rows x columns)
struct PrRed { int64_t y, p; int32_t n, size, nslots, pad; };   // up to 64 entries of one y block
struct PairHost
{
	std::vector<std::vector<uint32_t>> recs;   // one record per unit
	std::vector<double> cost;                  // estimated time per unit (tiles of the busiest wavefront per stage, staging)
	std::vector<int> xcd;                      // per unit: the XCD it is placed on
	std::vector<int64_t> wd_idx;               // class W_D tables (R16 x K4 each) as W element indices, -1: zero
	std::vector<PrRed> red;
	int64_t part_elems = 0;
	double flops = 0;
	int max_lds = 0, max_words = 0;
	int mfma_cap = 96;                         // units are cut into chunks of at most this many MFMAs
	int slots = 1024;                          // workgroups of the unit launch resident at once (chunks are packed into units only to fit these)
	int lds = PR_LDS;                          // LDS budget of a unit (doubles): 2048, 3072 or 4608
};
struct PairPlan
{
	uint32_t* recs = nullptr;
	int rec_words = 0, nunits = 0;
	double* wd = nullptr;
	PrRed* red = nullptr;
	int nred = 0;
	double* part = nullptr;
	int64_t part_elems = 0;
	double flops = 0;
	int max_lds = 0;
	int lds = PR_LDS;
	void upload(const PairHost& h, const double* w, hipStream_t s);
	void release();
};
// append one problem (operands at the given element offsets of the concatenated buffers); false
// when a class W_D exceeds PR_RMAX x PR_KMAX or a single row and column of a pair exceeds the LDS
bool plan_pair_tables(PairHost& h, const Meta& th, const Meta& W, const Meta& L, const Meta& R, int64_t x_off, int64_t w_off, int64_t l_off, int64_t r_off);
// lds: the unit's LDS budget in doubles (2048, 3072 or 4608; 0: pair_default_lds())
bool build_pair_plan(PairPlan& pp, const Meta& th, const Meta& W, const Meta& L, const Meta& R, const double* w, hipStream_t s, int lds = 0);
bool build_pair_batch(PairPlan& pp, const std::vector<FusedProblem>& probs, const double* w, hipStream_t s);
// host replay of the tables with the kernels' index arithmetic (no device); 0 when every check holds
int pair_replay(const PairHost& h, const double* x, int64_t nx, const double* r, int64_t nr, const double* l, int64_t nl, const double* w, int64_t nw, double* y);
int pair_default_cap();
int pair_default_slots(int device_slots);
int pair_default_lds();
// udot (optional, pp.nunits entries): per unit, its partial entries times x at their y indices
void launch_pair_units(const PairPlan& pp, const double* x, const double* r, const double* l, hipStream_t s, double* udot = nullptr);
void launch_pair_reduce(const PairPlan& pp, double* y, hipStream_t s);
void launch_pair_matvec(const PairPlan& pp, const double* x, const double* r, const double* l, double* y, hipStream_t s);
// the matvec with the Lanczos alpha partials of <y, x> from its partial-sum launch: dpart[0 .. G),
// G = pair_dot_partials(pp)
int pair_dot_partials(const PairPlan& pp);
void launch_pair_matvec_dot(const PairPlan& pp, const double* x, const double* r, const double* l, double* y, double* dpart, hipStream_t s);
// one deferred-normalisation Lanczos step on a sector-pair plan after launch_pair_units(.., dpart):
// beta from npart (the previous step's workgroup partials), alpha from the unit partials dpart, y
// summed from the partial blocks, v_j and u_{j+1} written, npart_out[0 .. G) the partials of
// |u_{j+1}|^2; buffers of pair_step_partials(pp) entries hold either partial set
int pair_step_partials(const PairPlan& pp);
void launch_pair_lz_step(const PairPlan& pp, const double* uj, double* unext, double* vj_out, const double* vjm1, int j, int maxiter,
	const double* dpart, const double* npart, double thresh, double* npart_out, double* scal, hipStream_t s);
// the same launches with the partial blocks in a file-scope device array of PR_SCRATCH_DOUBLES
// (no allocation; concurrent calls on one GPU share the array and race)
constexpr int64_t PR_SCRATCH_DOUBLES = (int64_t)1 << 22;
void launch_pair_matvec_noalloc(const PairPlan& pp, const double* x, const double* r, const double* l, double* y, hipStream_t s);

// Device workspace of the Lanczos iteration: maxiter Krylov vectors, w, the alpha / beta scalars,
// reduction partials and the Ritz coefficients. Grows on demand; freed with its owner.
struct LanczosWork
{
	double* v = nullptr;
	double* w = nullptr;
	double* scal = nullptr;
	double* part = nullptr;            // reduction partials: [0, pcap) alpha / norms, [pcap, 3 pcap) norms of w (two, alternating)
	double* u = nullptr;               // the unnormalised next vector of the recurrence with deferred normalisation
	double* z = nullptr;
	double* vs = nullptr;              // staged start vector (input of the captured recurrence)
	int64_t n = 0;
	int cap = 0, zcap = 0, pcap = 0;
	hipGraphExec_t graph = nullptr;    // captured recurrence for 'graph_maxiter' steps
	int graph_maxiter = 0;
	void reserve(int64_t n, int maxiter, int numeig, int partials = 1024);
	void drop_graph();
	~LanczosWork();
};

// y = A x on device vectors of the Lanczos length, enqueued on the Lanczos stream
using DevMatvec = std::function<void(const double* x, double* y)>;
// y = A x that also writes partial sums of <y, x> to part (in a fixed order); returns their number
struct DevMatvecDot
{
	std::function<void(const double* x, double* y, double* part)> apply;
	int partials = 0;
	// optional (deferred recurrence): the matvec's first launch writing the alpha partials to part,
	// then a step that forms y itself (units / step replace apply and the generic step kernel)
	std::function<void(const double* x, double* part)> units;
	int step_partials = 0;   // entries of either partial set of units / step
	std::function<void(const double* uj, double* unext, double* vj_out, const double* vjm1, int j, int maxiter, const double* dpart,
		const double* npart, double thresh, double* npart_out, double* scal)> step;
};

// Lanczos eigensystem (src/util/krylov.c:24-89 and :172-251): lambda[0:numeig] ascending,
// u (device, n x numeig row-major) the Ritz vectors. Returns 0, -1 (fewer Lanczos iterations than
// numeig) or -2 (tridiagonal eigensolver failure).
// With use_graph, the recurrence (maxiter matvecs and the vector updates) is captured once into a
// HIP graph held by lw and replayed on later calls with the same maxiter; matvec must then only
// enqueue kernels on s (no host synchronisation or copies). With sync_stop (eager only) the host
// reads the breakdown flag before each matvec and stops there, so matvec runs numiter times.
// With matvec_dot (partials > 0) each step's matvec also yields the alpha partials, and the step
// is the matvec plus two vector launches. With deferred (not with sync_stop) the matvec runs on the
// unnormalised next vector and one launch per step finalises beta, alpha and the next vector.
int lanczos_eigensystem(LanczosWork& lw, int64_t n, hipStream_t s, const DevMatvec& matvec, const double* vstart,
	int maxiter, int numeig, double* lambda, double* u, bool use_graph = false, bool sync_stop = false, const DevMatvecDot* matvec_dot = nullptr,
	bool deferred = false);

struct MatvecPlan
{
	int mode = 1;
	hipStream_t stream = nullptr;
	Meta theta_meta;
	DevTensor w, l, r;                          // resident operands
	std::vector<DevTensor> interm;              // intermediates of the chosen form
	// grouped forms: three launches; stage 1 reads x, stage 3 writes y
	DevGemm stage[3];
	GemmBases stage_bases[3] = {};
	WApply wapply;                              // column-stream form of stage 2
	bool use_wapply = false;
	bool use_f12 = false;                       // stages 1 and 2 fused (no T1 in HBM)
	// fused form: one launch
	FusedPlan fused;
	// sector-pair form: one unit launch and one reduction launch
	PairPlan pair;
	// plain form: the reference's operation sequence replayed as individual launches
	std::vector<std::function<void(const double* x, double* y)>> plain_ops;
	int plain_stage_end[3] = { 0, 0, 0 };       // index one past the last op of each dot stage
	std::vector<void*> owned;                   // device descriptor buffers of the plain form
	// bookkeeping
	double flops = 0, bytes = 0;
	double stage_flops[3] = { 0, 0, 0 };        // executed FP64 flops of each stage
	double stage_bytes[3] = { 0, 0, 0 };        // algorithmic bytes of each stage (operands read once, result written once)
	float tuned_ms[3][GEMM_NUM_FORMS] = {};     // autotuner timings (ms per launch) of each stage and form
	double nterms = 0, ntasks = 0, nitems = 0, interm_elems = 0;
	int nunits = 0;                             // fused form: workgroup units
	hipEvent_t ev[4] = { nullptr, nullptr, nullptr, nullptr };
	LanczosWork lz;                             // Lanczos workspace reused across solves
	~MatvecPlan();
};

// grouped three-stage tables of one or several problems (operand offsets relative to the stage bases)
struct GroupedTables
{
	GemmBatch s1, s2, s3;
	WApplyHost wa;
	int64_t t1_elems = 0, t2_elems = 0;
	double stage_bytes[3] = { 0, 0, 0 };
};
// lean: only the tables of the fixed-form launches (stage 2 as the column stream): no MFMA-tile
// stage-2 tables and no fused-stage tables (the caller plans again without it when the column
// stream does not apply, wa.kmax > WA_KMAX)
void plan_grouped_tables(GroupedTables& g, const Meta& th, const Meta& W, const Meta& L, const Meta& R,
	int64_t x_off, int64_t w_off, int64_t l_off, int64_t r_off, bool lean = false);
// host check of the tables (bounds, each intermediate / output entry written once); 0 when valid
int check_grouped_tables(const Meta& th, const Meta& W, const Meta& L, const Meta& R);
// append src, planned on its own (intermediates from offset 0), behind the intermediates of dst;
// src's operand offsets (theta / y, W, L, R) move by x_off, w_off, l_off, r_off
void merge_grouped_tables(GroupedTables& dst, const GroupedTables& src, int64_t x_off = 0, int64_t w_off = 0, int64_t l_off = 0,
	int64_t r_off = 0);

// contract checks of apply_local_hamiltonian (degrees, unit outer bonds, matching axes); throws HipError
void validate_matvec_operands(const Meta& theta, const Meta& w, const Meta& l, const Meta& r);
// time every launch form of each grouped stage (and the column-stream stage 2 when available) and
// keep the fastest; sets *use_wapply
void autotune_grouped(DevGemm* stage, const GemmBases* bases, const WApply* wa, bool* use_wapply, bool* use_f12, int64_t nx, hipStream_t s,
	float (*tuned_ms)[GEMM_NUM_FORMS]);
// one grouped matvec: stage 1, stage 2 (MFMA tiles or column stream), stage 3
void launch_grouped_stages(const DevGemm* stage, const GemmBases* bases, const WApply* wa, bool use_wapply, bool use_f12, const double* x, double* y,
	hipStream_t s, hipEvent_t* ev);

// FP64 flops of the reference's unfused operation sequence: sum of 2 m n k over its GEMM calls
double reference_flops(const Meta& th, const Meta& w, const Meta& l, const Meta& r);

//________________________________________________________________________________________________________________________
//
// batch of independent H_eff problems: one fused launch applies all of them to the concatenated
// vector; the batched Lanczos keeps one recurrence per problem (segments of the concatenated vectors)

struct LzSeg  { int64_t begin, end; int prob, pad; };   // element range of one reduction block
struct LzProb { int b0, b1; int64_t n; };               // reduction blocks and length of a problem

struct BatchLanczosWork
{
	double* v = nullptr;      // maxiter x ntot
	double* w = nullptr;
	double* vs = nullptr;
	double* scal = nullptr;   // per problem: alpha[maxiter], beta[maxiter], numiter, stop
	double* part = nullptr;   // 2 x nblocks
	double* z = nullptr;      // per problem: maxiter Ritz coefficients
	int* nit = nullptr;       // per problem: Lanczos iterations used by the Ritz vector
	int cap = 0;
	hipGraphExec_t graph = nullptr;
	int graph_maxiter = 0;
	void drop_graph();
	~BatchLanczosWork();
};

struct BatchPlan
{
	hipStream_t stream = nullptr;
	int nprob = 0;
	std::vector<int64_t> off;                   // nprob + 1 offsets of each problem in the concatenated vectors
	double* w = nullptr;                        // concatenated operands
	double* l = nullptr;
	double* r = nullptr;
	FusedPlan fused;
	bool fused_ok = false;
	// grouped three-stage form over the concatenated operands
	DevGemm stage[3];
	GemmBases stage_bases[3] = {};
	WApply wapply;
	bool use_wapply = false;
	bool use_f12 = false;
	double* t1 = nullptr;
	double* t2 = nullptr;
	bool use_fused = false;                     // the faster of the two forms, timed at plan creation
	bool split = false;                         // split-FP32 stages 1 and 3 (stage 2 the FP64 column stream)
	float form_ms[2] = { 0, 0 };                // ms per batched matvec: [0] grouped, [1] fused
	LzSeg* segs = nullptr;
	LzProb* probs = nullptr;
	int nblocks = 0;
	double flops = 0, bytes = 0;
	// host seconds of the plan's parts: [0] operand buffers and copies, [1] grouped tables,
	// [2] work items and uploads, [3] Lanczos reduction blocks
	double host_s[4] = { 0, 0, 0, 0 };
	BatchLanczosWork lz;
	~BatchPlan();
};

// pre: the problem's lean grouped tables planned ahead with zero operand offsets (or null)
struct BatchInput { Meta theta; DevTensor w, l, r; const GroupedTables* pre = nullptr; };
// tune: time the grouped launch forms and the fused form and keep the fastest; otherwise the
// grouped form with heuristic launch forms (no timing launches, for plans rebuilt often)
// par(n, f): run f(0 .. n-1), possibly in parallel host threads (host planning only, no HIP calls)
using ParallelFor = std::function<void(int, const std::function<void(int)>&)>;
// split: the split-FP32 form (the grouped tables, stages 1 and 3 on sliced FP32 MFMA, no timing)
BatchPlan* build_batch_plan(std::vector<BatchInput>& in, hipStream_t s, bool tune = true, const ParallelFor* par = nullptr, bool split = false);
void batch_apply(BatchPlan& b, const double* x, double* y);
// per-problem ground states: lambda[nprob], status[nprob] (0, -1 breakdown, -2 eigensolver), u concatenated
int lanczos_ground_state_batch(BatchPlan& b, const double* vstart, int maxiter, double* lambda, int* status, double* u, bool use_graph);

// Two-site DMRG sweep (src/algorithm/dmrg.c:262-399) as a resumable state: after init() the
// current local eigenproblem is (h2[bond], lb[bond], rb[bond + 1]) with start vector 'theta';
// advance() consumes its ground state (energy, eigenvector) and moves to the next bond. With
// tol_energy > 0 the run also ends after a sweep whose energy changed by less than tol_energy.
struct DmrgState
{
	hipStream_t st = nullptr;
	int L = 0, d = 0, maxiter = 0, num_sweeps = 0;
	double tol_split = 0, tol_energy = 0;
	ct_long max_vdim = 0;
	std::vector<qnumber> qsite;
	std::vector<DevTensor> A, W, lb, rb, h2;
	std::vector<double> en_sweeps, entropy;
	int sweep = 0, dir = 0, bond = 0;
	bool done = false;
	DevTensor theta;
	// split of the current bond: advance_begin (device, downloads the matrix), advance_host (host
	// block SVD only, no HIP calls: callers may run it for several states in parallel threads),
	// advance_end (device); advance() runs the three in sequence
	Meta split_meta;
	HostPacked split_in, split_m0, split_m1;
	trunc_info split_info = {};
	int split_ret = 0;
	// device split of the current bond (block SVD on the GPU, truncation on the host): the matrix,
	// its factors, the singular values in the host layout (descending within each block) and the
	// device column of each new-bond index
	bool split_on_device = false;
	SvdLayout split_lay;
	DevTensor split_am, split_u, split_vh;
	std::vector<double> split_s;
	std::vector<ct_long> split_pos, split_ind;
	ct_long split_num = 0;
	// phase timing (environment variable CT_DMRG_PHASES=1): the stream is synchronised at every
	// phase boundary and the wall time since the previous boundary is added to the phase
	enum { PH_PLAN, PH_LANCZOS, PH_SPLIT_BEGIN, PH_SPLIT_HOST, PH_SPLIT_UPLOAD, PH_ENVIRONMENT, PH_MERGE, PH_COUNT };
	bool trace = false;
	double phase_s[PH_COUNT] = {};
	double phase_t0 = 0;
	// plan_ahead: after init and after every split the lean grouped tables of the next local
	// eigenproblem are planned on the calling thread (zero operand offsets), for a batch plan that
	// then only concatenates them (the serving loop)
	bool plan_ahead = false;
	std::unique_ptr<GroupedTables> ahead;
	void plan_ahead_tables();
	void phase_mark();
	void phase_end(int ph);
	void phase_report() const;
	int init(const mpo* H, const mps* psi, int num_sweeps, int maxiter, double tol_split, ct_long max_vdim, double tol_energy);
	void advance_begin(const DevTensor& a_opt);
	void advance_host();
	int advance_end(double en);
	int advance(double en, DevTensor& a_opt);
	enum { SPLIT_HOST = 0, SPLIT_DEVICE_EIG = 1, SPLIT_DEVICE_SVD = 2 };
	int split_choice = -1;                     // a split mode fixed for this state (-1: CT_SPLIT decides)
	bool split_prescaled = false;              // the device factors carry the singular values already
	int split_mode() const;
	void download_mps(mps* psi) const;
	size_t device_bytes() const;
};

MatvecPlan* build_matvec_plan(const block_sparse_tensor* a, const block_sparse_tensor* w, const block_sparse_tensor* l, const block_sparse_tensor* r, int mode, bool autotune);
MatvecPlan* build_matvec_plan_dev(const Meta& theta, DevTensor&& w, DevTensor&& l, DevTensor&& r, int mode, hipStream_t s, bool autotune,
	bool reference_count = true);
void matvec_apply(MatvecPlan& p, const double* x, double* y, float* stage_ms);
// average duration (ms) of one stage's launches, 'reps' launched back to back (stage 0 only for the fused form)
float matvec_stage_time(MatvecPlan& p, int stage, const double* x, double* y, int reps);
int lanczos_ground_state(MatvecPlan& p, const double* vstart, int maxiter, double* lambda, double* u);

// host FP64 linear algebra for the truncating split (host_linalg.cpp)
int host_svd(ct_long m, ct_long n, const double* a, double* u, double* s, double* vh);
// a (m x n, row-major) = left (m x k) * right (k x n), k = min(m, n), with s the singular values in
// descending order; iso 0: left has orthonormal columns and right = diag(s) Vh; iso 1: right has
// orthonormal rows and left = U diag(s). Eigenvectors of the Gram matrix of the shorter side
// (after a QR of a tall matrix), the other factor by projection.
int host_svd_factor(ct_long m, ct_long n, const double* a, int iso, double* left, double* s, double* right);
int host_qr(ct_long m, ct_long n, const double* a, double* q, double* r);
int host_rq(ct_long m, ct_long n, const double* a, double* r, double* q);
int host_tridiag_eigh(int n, double* d, double* e, double* z);

}  // namespace ctg
