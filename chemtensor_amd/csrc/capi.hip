// extern "C" entry points of libchemtensor_gpu (declared in include/chemtensor_gpu.h).

#include <cstring>
#include <memory>

#include "ct_internal.h"

using namespace ctg;

struct ct_matvec_plan
{
	std::unique_ptr<MatvecPlan> p;
};

struct ct_matvec_batch
{
	std::unique_ptr<BatchPlan> b;
};

static int map_error(const HipError& e)
{
	return (e.code == hipErrorNotSupported) ? CT_ERR_UNSUPPORTED : (e.code == hipErrorInvalidValue ? CT_ERR_ARG : CT_ERR_HIP);
}

// void entry points mirror the reference's contract checks (assert): a failure is fatal
[[noreturn]] static void fatal(const char* fn, const char* what)
{
	fprintf(stderr, "chemtensor_gpu: %s failed: %s\n", fn, what);
	abort();
}

extern "C" {

void apply_local_hamiltonian(const struct block_sparse_tensor* a, const struct block_sparse_tensor* w, const struct block_sparse_tensor* l, const struct block_sparse_tensor* r, struct block_sparse_tensor* b)
{
	try {
		std::unique_ptr<MatvecPlan> p(build_matvec_plan(a, w, l, r, CT_MATVEC_GROUPED_FP64, false));
		DevTensor x = upload(a, p->stream);
		// result axes: l axis 3, w axis 1, r axis 2 (chain_ops.c:388); equal to a's axes here
		DevTensor y(p->theta_meta);
		matvec_apply(*p, x.d, y.d, nullptr);
		download(y, b, p->stream);
	}
	catch (const HipError& e) {
		fatal("apply_local_hamiltonian", e.where.c_str());
	}
}

void block_sparse_tensor_dot(const struct block_sparse_tensor* s, const enum tensor_axis_range axrange_s, const struct block_sparse_tensor* t, const enum tensor_axis_range axrange_t, const int ndim_mult, struct block_sparse_tensor* r)
{
	try {
		hipStream_t st = default_stream();
		DevTensor ds = upload(s, st), dt = upload(t, st);
		DevTensor dr = dev_dot(ds, axrange_s, dt, axrange_t, ndim_mult, DotForm::Grouped, st);
		download(dr, r, st);
	}
	catch (const HipError& e) {
		fatal("block_sparse_tensor_dot", e.where.c_str());
	}
}

void block_sparse_tensor_transpose(const int* perm, const struct block_sparse_tensor* t, struct block_sparse_tensor* r)
{
	try {
		hipStream_t st = default_stream();
		DevTensor dt = upload(t, st);
		DevTensor dr = dev_transpose(perm, dt, st);
		download(dr, r, st);
	}
	catch (const HipError& e) {
		fatal("block_sparse_tensor_transpose", e.where.c_str());
	}
}

void block_sparse_tensor_cyclic_partial_trace(const struct block_sparse_tensor* t, const int ndim_trace, struct block_sparse_tensor* r)
{
	try {
		if (!t || !r || ndim_trace < 0 || t->ndim < 2 * ndim_trace) throw HipError{ hipErrorInvalidValue, "invalid arguments" };
		if (t->ndim == 2 * ndim_trace && ndim_trace > 0) throw HipError{ hipErrorNotSupported, "a trace over every axis (degree-0 result)" };
		for (int i = 0; i < ndim_trace; i++) {
			const int k = t->ndim - ndim_trace + i;
			if (t->dim_logical[i] != t->dim_logical[k] || t->axis_dir[i] != -t->axis_dir[k])
				throw HipError{ hipErrorInvalidValue, "traced axes differ in dimension or do not have opposite directions" };
			for (ct_long j = 0; j < t->dim_logical[i]; j++)
				if (t->qnums_logical[i][j] != t->qnums_logical[k][j]) throw HipError{ hipErrorInvalidValue, "traced axes differ in quantum numbers" };
		}
		hipStream_t st = default_stream();
		DevTensor dt = upload(t, st);
		if (ndim_trace == 0) {
			download(dt, r, st);
			return;
		}
		DevTensor dr = dev_cyclic_partial_trace(dt, ndim_trace, st);
		download(dr, r, st);
	}
	catch (const HipError& e) {
		fatal("block_sparse_tensor_cyclic_partial_trace", e.where.c_str());
	}
}

void mpo_merge_tensor_pair(const struct block_sparse_tensor* a0, const struct block_sparse_tensor* a1, struct block_sparse_tensor* a)
{
	try {
		hipStream_t st = default_stream();
		DevTensor d0 = upload(a0, st), d1 = upload(a1, st);
		DevTensor t = dev_dot(d0, TENSOR_AXIS_RANGE_TRAILING, d1, TENSOR_AXIS_RANGE_LEADING, 1, DotForm::Grouped, st);
		const int perm[6] = { 0, 1, 3, 2, 4, 5 };
		DevTensor tp = dev_transpose(perm, t, st);
		DevTensor f1 = dev_flatten_axes(tp, 1, TENSOR_AXIS_OUT, st);
		DevTensor f2 = dev_flatten_axes(f1, 2, TENSOR_AXIS_IN, st);
		download(f2, a, st);
	}
	catch (const HipError& e) {
		fatal("mpo_merge_tensor_pair", e.where.c_str());
	}
}

int ct_matvec_plan_create(const struct block_sparse_tensor* a, const struct block_sparse_tensor* w, const struct block_sparse_tensor* l, const struct block_sparse_tensor* r,
	const int mode, ct_matvec_plan** plan)
{
	if (!a || !w || !l || !r || !plan || mode < 0 || mode > 4) return CT_ERR_ARG;
	try {
		std::unique_ptr<ct_matvec_plan> h(new ct_matvec_plan());
		const char* env = getenv("CT_AUTOTUNE");
		const bool tune = !(env && env[0] == '0');
		h->p.reset(build_matvec_plan(a, w, l, r, mode, tune));
		*plan = h.release();
		return 0;
	}
	catch (const HipError& e) {
		return (e.code == hipErrorNotSupported) ? CT_ERR_UNSUPPORTED : (e.code == hipErrorInvalidValue ? CT_ERR_ARG : CT_ERR_HIP);
	}
}

int ct_matvec_plan_destroy(ct_matvec_plan* plan)
{
	delete plan;
	return 0;
}

ct_long ct_matvec_plan_dim(const ct_matvec_plan* plan) { return plan->p->theta_meta.nelem; }
double ct_matvec_plan_flops(const ct_matvec_plan* plan) { return plan->p->flops; }
double ct_matvec_plan_bytes(const ct_matvec_plan* plan) { return plan->p->bytes; }

int ct_matvec_plan_stage_flops(const ct_matvec_plan* plan, double* flops3)
{
	const MatvecPlan& p = *plan->p;
	for (int i = 0; i < 3; i++) flops3[i] = p.stage_flops[i];
	return 0;
}

int ct_matvec_plan_stage_bytes(const ct_matvec_plan* plan, double* bytes3)
{
	const MatvecPlan& p = *plan->p;
	for (int i = 0; i < 3; i++) bytes3[i] = p.stage_bytes[i];
	return 0;
}

int ct_matvec_plan_forms(const ct_matvec_plan* plan, int* forms3, float* tuned_ms12)
{
	const MatvecPlan& p = *plan->p;
	for (int i = 0; i < 3; i++) {
		forms3[i] = (i < 2 && p.use_f12 && p.wapply.f12_ok) ? GEMM_NUM_FORMS + 1 : (i == 1 && p.use_wapply) ? GEMM_NUM_FORMS
			: (p.stage[i].use_lds && p.stage[i].nlitems > 0) ? GEMM_NUM_FORMS + 2 : p.stage[i].form;
		for (int f = 0; f < GEMM_NUM_FORMS; f++) tuned_ms12[i * GEMM_NUM_FORMS + f] = p.tuned_ms[i][f];
	}
	return 0;
}

int ct_matvec_plan_stage_time(ct_matvec_plan* plan, const int stage, const double* d_x, double* d_y, const int reps, float* ms)
{
	if (!plan || stage < 0 || stage > 2 || reps < 1 || !ms) return CT_ERR_ARG;
	try {
		*ms = matvec_stage_time(*plan->p, stage, d_x, d_y, reps);
		return 0;
	}
	catch (const HipError& e) {
		return map_error(e);
	}
}

int ct_matvec_plan_stats(const ct_matvec_plan* plan, double* stats4)
{
	const MatvecPlan& p = *plan->p;
	stats4[0] = p.nterms;
	stats4[1] = p.ntasks;
	stats4[2] = p.nitems;
	stats4[3] = p.interm_elems;
	return 0;
}

int ct_matvec_apply(ct_matvec_plan* plan, const double* d_x, double* d_y)
{
	try {
		matvec_apply(*plan->p, d_x, d_y, nullptr);
		return 0;
	}
	catch (const HipError&) {
		return CT_ERR_HIP;
	}
}

int ct_matvec_apply_noalloc(ct_matvec_plan* plan, const double* d_x, double* d_y)
{
	if (!plan || !d_x || !d_y) return CT_ERR_ARG;
	const MatvecPlan& p = *plan->p;
	if (p.mode != CT_MATVEC_PAIR_FP64) return CT_ERR_UNSUPPORTED;
	if (p.pair.part_elems + 64 > PR_SCRATCH_DOUBLES) {
		fprintf(stderr, "ct_matvec_apply_noalloc: the plan needs %lld doubles of partial blocks, the scratch array holds %lld\n",
			(long long)(p.pair.part_elems + 64), (long long)PR_SCRATCH_DOUBLES);
		return CT_ERR_NOMEM;
	}
	try {
		launch_pair_matvec_noalloc(p.pair, d_x, p.r.d, p.l.d, d_y, p.stream);
		return 0;
	}
	catch (const HipError&) {
		return CT_ERR_HIP;
	}
}

int ct_matvec_apply_timed(ct_matvec_plan* plan, const double* d_x, double* d_y, float* stage_ms3)
{
	try {
		matvec_apply(*plan->p, d_x, d_y, stage_ms3);
		return 0;
	}
	catch (const HipError&) {
		return CT_ERR_HIP;
	}
}

void* ct_matvec_plan_stream(ct_matvec_plan* plan) { return (void*)plan->p->stream; }

int ct_lanczos_ground_state(ct_matvec_plan* plan, const double* d_vstart, const int maxiter, double* lambda, double* d_u)
{
	try {
		return lanczos_ground_state(*plan->p, d_vstart, maxiter, lambda, d_u);
	}
	catch (const HipError&) {
		return CT_ERR_HIP;
	}
}

// src/util/krylov.c:172-251 with a caller-supplied host matvec: the Krylov vectors, the recurrence
// and the Ritz vectors live on the device; each step copies v_j to the host for 'afunc' and the
// result back. lambda[numeig] ascending, u_ritz n x numeig row-major (host memory).
int eigensystem_krylov_symmetric(const ct_long n, lanczos_linear_func_d afunc, const void* adata,
	const double* vstart, const int maxiter, const int numeig, double* lambda, double* u_ritz)
{
	if (n < 1 || maxiter < 1 || numeig < 1 || numeig > maxiter) return CT_ERR_ARG;
	try {
		hipStream_t s = default_stream();
		LanczosWork lw;
		const auto dfree = [](double* q) { dev_free(q); };
		std::unique_ptr<double, decltype(dfree)> dv((double*)dev_alloc((size_t)n * sizeof(double)), dfree);
		std::unique_ptr<double, decltype(dfree)> du((double*)dev_alloc((size_t)n * numeig * sizeof(double)), dfree);
		std::vector<double> hx((size_t)n), hy((size_t)n);
		CTG_HIP(hipMemcpyAsync(dv.get(), vstart, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
		const DevMatvec mv = [&](const double* x, double* y) {
			CTG_HIP(hipMemcpyAsync(hx.data(), x, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s));
			CTG_HIP(hipStreamSynchronize(s));
			afunc(n, adata, hx.data(), hy.data());
			CTG_HIP(hipMemcpyAsync(y, hy.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
		};
		const int ret = lanczos_eigensystem(lw, n, s, mv, dv.get(), maxiter, numeig, lambda, du.get(), false, true);
		if (ret != 0) return ret;
		CTG_HIP(hipMemcpyAsync(u_ritz, du.get(), (size_t)n * numeig * sizeof(double), hipMemcpyDeviceToHost, s));
		CTG_HIP(hipStreamSynchronize(s));
		return 0;
	}
	catch (const HipError&) {
		return CT_ERR_HIP;
	}
}

int ct_matvec_plan_check(const struct block_sparse_tensor* a, const struct block_sparse_tensor* w, const struct block_sparse_tensor* l,
	const struct block_sparse_tensor* r)
{
	if (!a || !w || !l || !r) return CT_ERR_ARG;
	try {
		const Meta ma = meta_of(a), mw = meta_of(w), ml = meta_of(l), mr = meta_of(r);
		validate_matvec_operands(ma, mw, ml, mr);
		return check_grouped_tables(ma, mw, ml, mr);
	}
	catch (const HipError& e) {
		return map_error(e);
	}
}

int ct_matvec_pair_check(const struct block_sparse_tensor* a, const struct block_sparse_tensor* w, const struct block_sparse_tensor* l,
	const struct block_sparse_tensor* r, const int mfma_cap, double* y)
{
	if (!a || !w || !l || !r) return CT_ERR_ARG;
	try {
		const Meta ma = meta_of(a), mw = meta_of(w), ml = meta_of(l), mr = meta_of(r);
		validate_matvec_operands(ma, mw, ml, mr);
		PairHost h;
		h.mfma_cap = mfma_cap > 0 ? mfma_cap : pair_default_cap();
		h.lds = pair_default_lds();
		h.slots = pair_default_slots(h.slots);
		if (!plan_pair_tables(h, ma, mw, ml, mr, 0, 0, 0, 0)) return CT_ERR_UNSUPPORTED;
		std::vector<double> x((size_t)ma.nelem), vw((size_t)mw.nelem), vl((size_t)ml.nelem), vr((size_t)mr.nelem);
		block_sparse_tensor_serialize_entries(a, x.data());
		block_sparse_tensor_serialize_entries(w, vw.data());
		block_sparse_tensor_serialize_entries(l, vl.data());
		block_sparse_tensor_serialize_entries(r, vr.data());
		return pair_replay(h, x.data(), ma.nelem, vr.data(), mr.nelem, vl.data(), ml.nelem, vw.data(), mw.nelem, y);
	}
	catch (const HipError& e) {
		return map_error(e);
	}
}

int ct_matvec_pair_stats(const struct block_sparse_tensor* a, const struct block_sparse_tensor* w, const struct block_sparse_tensor* l,
	const struct block_sparse_tensor* r, const int mfma_cap, double* stats6)
{
	if (!a || !w || !l || !r || !stats6) return CT_ERR_ARG;
	try {
		const Meta ma = meta_of(a), mw = meta_of(w), ml = meta_of(l), mr = meta_of(r);
		validate_matvec_operands(ma, mw, ml, mr);
		PairHost h;
		h.mfma_cap = mfma_cap > 0 ? mfma_cap : pair_default_cap();
		h.lds = pair_default_lds();
		h.slots = pair_default_slots(h.slots);
		if (!plan_pair_tables(h, ma, mw, ml, mr, 0, 0, 0, 0)) return CT_ERR_UNSUPPORTED;
		double cmax = 0, csum = 0;
		for (double c : h.cost) { cmax = std::max(cmax, c); csum += c; }
		stats6[0] = (double)h.recs.size();
		stats6[1] = (double)h.part_elems;
		stats6[2] = csum;
		stats6[3] = cmax;
		stats6[4] = (double)h.max_lds;
		stats6[5] = (double)h.max_words;
		return 0;
	}
	catch (const HipError& e) {
		return map_error(e);
	}
}

int ct_matvec_batch_create(const int nprob, const struct block_sparse_tensor* const* a, const struct block_sparse_tensor* const* w,
	const struct block_sparse_tensor* const* l, const struct block_sparse_tensor* const* r, ct_matvec_batch** batch)
{
	return ct_matvec_batch_create_mode(nprob, a, w, l, r, CT_MATVEC_GROUPED_FP64, batch);
}

int ct_matvec_batch_create_mode(const int nprob, const struct block_sparse_tensor* const* a, const struct block_sparse_tensor* const* w,
	const struct block_sparse_tensor* const* l, const struct block_sparse_tensor* const* r, const int mode, ct_matvec_batch** batch)
{
	if (nprob < 1 || !a || !w || !l || !r || !batch) return CT_ERR_ARG;
	if (mode != CT_MATVEC_GROUPED_FP64 && mode != CT_MATVEC_GROUPED_SPLIT) return CT_ERR_UNSUPPORTED;
	try {
		hipStream_t s = default_stream();
		std::vector<BatchInput> in((size_t)nprob);
		for (int p = 0; p < nprob; p++) {
			if (!a[p] || !w[p] || !l[p] || !r[p]) return CT_ERR_ARG;
			in[p].theta = meta_of(a[p]);
			in[p].w = upload(w[p], s);
			in[p].l = upload(l[p], s);
			in[p].r = upload(r[p], s);
		}
		std::unique_ptr<ct_matvec_batch> h(new ct_matvec_batch());
		h->b.reset(build_batch_plan(in, s, true, nullptr, mode == CT_MATVEC_GROUPED_SPLIT));
		*batch = h.release();
		return 0;
	}
	catch (const HipError& e) {
		return map_error(e);
	}
}

int ct_matvec_batch_destroy(ct_matvec_batch* batch)
{
	delete batch;
	return 0;
}

ct_long ct_matvec_batch_dim(const ct_matvec_batch* batch) { return batch->b->off.back(); }

int ct_matvec_batch_offsets(const ct_matvec_batch* batch, ct_long* offsets)
{
	for (size_t i = 0; i < batch->b->off.size(); i++) offsets[i] = batch->b->off[i];
	return 0;
}

double ct_matvec_batch_flops(const ct_matvec_batch* batch) { return batch->b->flops; }

int ct_matvec_batch_form(const ct_matvec_batch* batch, int* fused, float* ms2)
{
	if (fused) *fused = batch->b->use_fused ? 1 : 0;
	if (ms2) { ms2[0] = batch->b->form_ms[0]; ms2[1] = batch->b->form_ms[1]; }
	return 0;
}
double ct_matvec_batch_bytes(const ct_matvec_batch* batch) { return batch->b->bytes; }
void* ct_matvec_batch_stream(ct_matvec_batch* batch) { return (void*)batch->b->stream; }

int ct_matvec_batch_apply(ct_matvec_batch* batch, const double* d_x, double* d_y)
{
	try {
		batch_apply(*batch->b, d_x, d_y);
		return 0;
	}
	catch (const HipError& e) {
		return map_error(e);
	}
}

int ct_lanczos_ground_state_batch(ct_matvec_batch* batch, const double* d_vstart, const int maxiter, double* lambda, int* status, double* d_u)
{
	try {
		const char* env = getenv("CT_LANCZOS_GRAPH");
		return lanczos_ground_state_batch(*batch->b, d_vstart, maxiter, lambda, status, d_u, !(env && env[0] == '0'));
	}
	catch (const HipError& e) {
		return map_error(e);
	}
}

int ct_device_alloc(void** ptr, const ct_long nbytes) { return hipMalloc(ptr, (size_t)nbytes) == hipSuccess ? 0 : CT_ERR_NOMEM; }
int ct_device_free(void* ptr) { return hipFree(ptr) == hipSuccess ? 0 : CT_ERR_HIP; }
int ct_memcpy_h2d(void* dst, const void* src, const ct_long nbytes) { return hipMemcpy(dst, src, (size_t)nbytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : CT_ERR_HIP; }
int ct_memcpy_d2h(void* dst, const void* src, const ct_long nbytes) { return hipMemcpy(dst, src, (size_t)nbytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : CT_ERR_HIP; }
int ct_device_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : CT_ERR_HIP; }

int ct_device_count(void)
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess) { (void)hipGetLastError(); return 0; }
	return n;
}

int ct_event_create(void** ev)
{
	hipEvent_t e;
	if (hipEventCreate(&e) != hipSuccess) return CT_ERR_HIP;
	*ev = (void*)e;
	return 0;
}

int ct_event_record(void* ev, void* stream) { return hipEventRecord((hipEvent_t)ev, (hipStream_t)stream) == hipSuccess ? 0 : CT_ERR_HIP; }

int ct_event_elapsed_ms(void* ev0, void* ev1, float* ms)
{
	if (hipEventSynchronize((hipEvent_t)ev1) != hipSuccess) return CT_ERR_HIP;
	return hipEventElapsedTime(ms, (hipEvent_t)ev0, (hipEvent_t)ev1) == hipSuccess ? 0 : CT_ERR_HIP;
}

int ct_event_destroy(void* ev) { return hipEventDestroy((hipEvent_t)ev) == hipSuccess ? 0 : CT_ERR_HIP; }

int ct_device_copy(double* dst, const double* src, const ct_long n)
{
	try {
		launch_copy(src, dst, n, default_stream());
		CTG_HIP(hipStreamSynchronize(default_stream()));
		return 0;
	}
	catch (const HipError&) {
		return CT_ERR_HIP;
	}
}

int ct_stream_synchronize(void* stream) { return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? 0 : CT_ERR_HIP; }

}  // extern "C"

extern "C" {

void ct_fill_random_normal(struct block_sparse_tensor* t, double alpha, struct rng_state* rng);
void ct_random_bond_qnums(const int nsites, const ct_long d, const qnumber* qsite, const qnumber sector, const ct_long max_vdim, const uint64_t seed, ct_long* dims_out, qnumber* qbonds_out);

int ct_generate_matvec_problem_mpo(const struct mpo* Hp, const qnumber sector, const ct_long chi, const uint64_t seed,
	struct block_sparse_tensor* theta, struct block_sparse_tensor* h2, struct block_sparse_tensor* left, struct block_sparse_tensor* right)
{
	if (Hp == nullptr || Hp->nsites < 4 || chi < 1) return CT_ERR_ARG;
	const struct mpo& H = *Hp;
	const int nsites = H.nsites;
	const ct_long d = H.d;
	std::vector<ct_long> dims((size_t)nsites + 1);
	std::vector<qnumber> qall((size_t)(nsites + 1) * (size_t)chi + 8);
	ct_random_bond_qnums(nsites, d, H.qsite, sector, chi, seed, dims.data(), qall.data());
	std::vector<std::vector<qnumber>> qb((size_t)nsites + 1);
	ct_long off = 0;
	for (int l = 0; l <= nsites; l++) { qb[l].assign(qall.begin() + off, qall.begin() + off + dims[l]); off += dims[l]; }
	const int i = nsites / 2 - 1;
	std::vector<qnumber> qp((size_t)(d * d));
	for (ct_long a = 0; a < d; a++) for (ct_long b = 0; b < d; b++) qp[a * d + b] = H.qsite[a] + H.qsite[b];
	struct rng_state rng;
	{
		const ct_long dim[3] = { dims[i], d * d, dims[i + 2] };
		const tensor_axis_direction ad[3] = { TENSOR_AXIS_OUT, TENSOR_AXIS_OUT, TENSOR_AXIS_IN };
		const qnumber* qn[3] = { qb[i].data(), qp.data(), qb[i + 2].data() };
		allocate_block_sparse_tensor(CT_DOUBLE_REAL, 3, dim, ad, qn, theta);
		seed_rng_state(seed + 1, &rng);
		ct_fill_random_normal(theta, 1.0, &rng);
	}
	mpo_merge_tensor_pair(&H.a[i], &H.a[i + 1], h2);
	const qnumber zero[1] = { 0 };
	{
		const block_sparse_tensor* wl = &H.a[i];
		const ct_long dim[4] = { 1, dims[i], wl->dim_logical[0], dims[i] };
		const tensor_axis_direction ad[4] = { TENSOR_AXIS_OUT, TENSOR_AXIS_IN, TENSOR_AXIS_IN, TENSOR_AXIS_OUT };
		const qnumber* qn[4] = { zero, qb[i].data(), wl->qnums_logical[0], qb[i].data() };
		allocate_block_sparse_tensor(CT_DOUBLE_REAL, 4, dim, ad, qn, left);
		seed_rng_state(seed + 2, &rng);
		ct_fill_random_normal(left, 1.0, &rng);
	}
	{
		const block_sparse_tensor* wr = &H.a[i + 1];
		const ct_long dim[4] = { dims[i + 2], wr->dim_logical[3], dims[i + 2], 1 };
		const tensor_axis_direction ad[4] = { TENSOR_AXIS_OUT, TENSOR_AXIS_OUT, TENSOR_AXIS_IN, TENSOR_AXIS_IN };
		const qnumber* qn[4] = { qb[i + 2].data(), wr->qnums_logical[3], qb[i + 2].data(), zero };
		allocate_block_sparse_tensor(CT_DOUBLE_REAL, 4, dim, ad, qn, right);
		seed_rng_state(seed + 3, &rng);
		ct_fill_random_normal(right, 1.0, &rng);
	}
	return 0;
}

int ct_generate_matvec_problem(const int model, const int nsites, const ct_long chi, const uint64_t seed,
	struct block_sparse_tensor* theta, struct block_sparse_tensor* h2, struct block_sparse_tensor* left, struct block_sparse_tensor* right)
{
	if (nsites < 4 || chi < 1 || (model != 0 && model != 1)) return CT_ERR_ARG;
	struct mpo H;
	qnumber sector;
	if (model == 0) { ct_construct_heisenberg_xxz_1d_mpo(nsites, 1.0, 0.5, 0.0, &H); sector = 0; }
	else { ct_construct_fermi_hubbard_1d_mpo(nsites, 1.0, 4.0, 0.0, &H); sector = nsites * 65536; }
	const int ret = ct_generate_matvec_problem_mpo(&H, sector, chi, seed, theta, h2, left, right);
	delete_mpo(&H);
	return ret;
}

}  // extern "C"
