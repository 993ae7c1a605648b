// Two-site DMRG with a device-resident sweep (src/algorithm/dmrg.c:262-399): MPS site tensors,
// MPO tensors, merged two-site MPO tensors and all environment blocks live on the GPU as packed
// block-sparse tensors; contractions, transposes and reshapes run on the GPU; the Lanczos
// iteration runs on the GPU in FP64 with the grouped matvec plan of each bond. The truncating
// block SVD and the QR/RQ orthonormalisation run in FP64 on the host (they are small next to
// the maxiter matvecs of every bond).

#include <cmath>
#include <cstring>
#include <memory>

#include <chrono>
#include <cstdlib>
#include <numeric>
#include <string>

#include "ct_internal.h"

namespace ctg {

__global__ __launch_bounds__(256) void k_scale(double* __restrict__ p, double a, int64_t n)
{
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] *= a;
}

static void dev_scale(DevTensor& t, double a, hipStream_t s)
{
	if (t.meta.nelem == 0) return;
	const int64_t nb = (t.meta.nelem + 255) / 256;
	hipLaunchKernelGGL(k_scale, dim3((unsigned)std::min<int64_t>(nb, 4096)), dim3(256), 0, s, t.d, a, (int64_t)t.meta.nelem);
	CTG_HIP(hipGetLastError());
}

static Meta meta_with_dirs(const Meta& m, const std::vector<int>& dirs)
{
	std::vector<const qnumber*> qn(m.ndim);
	for (int i = 0; i < m.ndim; i++) qn[i] = m.qnums_logical[i].data();
	return make_meta(m.ndim, m.dim_logical.data(), dirs.data(), qn.data());
}

// a 6-axis tensor of unit dimensions with its single entry 1 (if the block exists)
static DevTensor unit6(const qnumber* const* qn, hipStream_t s)
{
	const ct_long dim[6] = { 1, 1, 1, 1, 1, 1 };
	const int ad[6] = { TENSOR_AXIS_OUT, TENSOR_AXIS_OUT, TENSOR_AXIS_IN, TENSOR_AXIS_IN, TENSOR_AXIS_IN, TENSOR_AXIS_OUT };
	HostPacked h;
	h.meta = make_meta(6, dim, ad, qn);
	h.v.assign((size_t)h.meta.nelem, 1.0);
	return upload_packed(h, s);
}

// chain_ops.c:14-46
static DevTensor dummy_right(const Meta& a, const Meta& b, const Meta& w, hipStream_t s)
{
	const qnumber* qn[6] = { a.qnums_logical[2].data(), w.qnums_logical[3].data(), b.qnums_logical[2].data(),
		a.qnums_logical[2].data(), w.qnums_logical[3].data(), b.qnums_logical[2].data() };
	DevTensor u = unit6(qn, s);
	DevTensor t = dev_flatten_axes(u, 4, TENSOR_AXIS_IN, s);
	return dev_flatten_axes(t, 3, TENSOR_AXIS_IN, s);
}

// chain_ops.c:53-85
static DevTensor dummy_left(const Meta& a, const Meta& b, const Meta& w, hipStream_t s)
{
	const qnumber* qn[6] = { a.qnums_logical[0].data(), w.qnums_logical[0].data(), b.qnums_logical[0].data(),
		a.qnums_logical[0].data(), w.qnums_logical[0].data(), b.qnums_logical[0].data() };
	DevTensor u = unit6(qn, s);
	DevTensor t = dev_flatten_axes(u, 0, TENSOR_AXIS_OUT, s);
	return dev_flatten_axes(t, 0, TENSOR_AXIS_OUT, s);
}

static DevTensor reversed(const DevTensor& b)
{
	std::vector<int> dirs(b.meta.axis_dir);
	for (int& d : dirs) d = -d;
	return DevTensor::view(b, meta_with_dirs(b.meta, dirs));
}

// chain_ops.c:116-165
static DevTensor step_right(const DevTensor& a, const DevTensor& b, const DevTensor& w, const DevTensor& r, hipStream_t st)
{
	DevTensor s = dev_dot(a, TENSOR_AXIS_RANGE_TRAILING, r, TENSOR_AXIS_RANGE_LEADING, 1, DotForm::Grouped, st);
	const int p0[5] = { 1, 2, 0, 3, 4 };
	DevTensor t = dev_transpose(p0, s, st);
	s = dev_dot(w, TENSOR_AXIS_RANGE_TRAILING, t, TENSOR_AXIS_RANGE_LEADING, 2, DotForm::Grouped, st);
	const int p1[5] = { 4, 2, 0, 1, 3 };
	t = dev_transpose(p1, s, st);
	DevTensor br = reversed(b);
	s = dev_dot(t, TENSOR_AXIS_RANGE_TRAILING, br, TENSOR_AXIS_RANGE_TRAILING, 2, DotForm::Grouped, st);
	const int p2[4] = { 1, 2, 3, 0 };
	return dev_transpose(p2, s, st);
}

// chain_ops.c:196-245
static DevTensor step_left(const DevTensor& a, const DevTensor& b, const DevTensor& w, const DevTensor& l, hipStream_t st)
{
	DevTensor br = reversed(b);
	DevTensor s = dev_dot(l, TENSOR_AXIS_RANGE_TRAILING, br, TENSOR_AXIS_RANGE_LEADING, 1, DotForm::Grouped, st);
	const int p0[5] = { 0, 1, 4, 2, 3 };
	DevTensor t = dev_transpose(p0, s, st);
	s = dev_dot(t, TENSOR_AXIS_RANGE_TRAILING, w, TENSOR_AXIS_RANGE_LEADING, 2, DotForm::Grouped, st);
	const int p1[5] = { 1, 3, 4, 2, 0 };
	t = dev_transpose(p1, s, st);
	s = dev_dot(a, TENSOR_AXIS_RANGE_LEADING, t, TENSOR_AXIS_RANGE_LEADING, 2, DotForm::Grouped, st);
	const int p2[4] = { 3, 0, 1, 2 };
	return dev_transpose(p2, s, st);
}

// mps.c:1166-1180
static DevTensor merge_mps(const DevTensor& a0, const DevTensor& a1, hipStream_t st)
{
	DevTensor t = dev_dot(a0, TENSOR_AXIS_RANGE_TRAILING, a1, TENSOR_AXIS_RANGE_LEADING, 1, DotForm::Grouped, st);
	return dev_flatten_axes(t, 1, TENSOR_AXIS_OUT, st);
}

// mpo.c:255-282
static DevTensor merge_mpo(const DevTensor& a0, const DevTensor& a1, hipStream_t st)
{
	DevTensor t = dev_dot(a0, TENSOR_AXIS_RANGE_TRAILING, a1, TENSOR_AXIS_RANGE_LEADING, 1, DotForm::Grouped, st);
	const int perm[6] = { 0, 1, 3, 2, 4, 5 };
	DevTensor tp = dev_transpose(perm, t, st);
	DevTensor f = dev_flatten_axes(tp, 1, TENSOR_AXIS_OUT, st);
	return dev_flatten_axes(f, 2, TENSOR_AXIS_IN, st);
}

// mps.c:560-601: right-orthonormalise 'a' by RQ and multiply R into 'a_prev'
static int local_orthonormalize_rq(DevTensor& a, DevTensor& a_prev, hipStream_t st)
{
	const ct_long dr[2] = { a.meta.dim_logical[1], a.meta.dim_logical[2] };
	const std::vector<qnumber> q1 = a.meta.qnums_logical[1], q2 = a.meta.qnums_logical[2];
	const int ad[2] = { TENSOR_AXIS_OUT, TENSOR_AXIS_IN };
	DevTensor am = dev_flatten_axes(a, 1, TENSOR_AXIS_IN, st);
	HostPacked hm = download_packed(am, st), hr, hq;
	if (host_block_rq(hm, hr, hq) != 0) return -1;
	DevTensor dq = upload_packed(hq, st), drr = upload_packed(hr, st);
	const qnumber* qn[2] = { q1.data(), q2.data() };
	a = dev_split_axis(dq, 1, dr, ad, qn, st);
	a_prev = dev_dot(a_prev, TENSOR_AXIS_RANGE_TRAILING, drr, TENSOR_AXIS_RANGE_LEADING, 1, DotForm::Grouped, st);
	return 0;
}

// the dummy tensor of dimensions (D, 1, D) with the outer bond quantum numbers and entry 1
static DevTensor dummy_site(const Meta& a, int ax, hipStream_t st)
{
	const ct_long dim[3] = { a.dim_logical[ax], 1, a.dim_logical[ax] };
	const int ad[3] = { TENSOR_AXIS_OUT, TENSOR_AXIS_OUT, TENSOR_AXIS_IN };
	const qnumber qz[1] = { 0 };
	const qnumber* qn[3] = { a.qnums_logical[ax].data(), qz, a.qnums_logical[ax].data() };
	HostPacked h;
	h.meta = make_meta(3, dim, ad, qn);
	h.v.assign((size_t)h.meta.nelem, 0.0);
	if (h.meta.nelem > 0 && h.meta.present(0)) h.v[0] = 1.0;
	return upload_packed(h, st);
}

int DmrgState::init(const mpo* H, const mps* psi, int num_sweeps_, int maxiter_, double tol_split_, ct_long max_vdim_, double tol_energy_)
{
	L = H->nsites;
	if (L < 2 || psi->nsites != L || num_sweeps_ < 1 || maxiter_ < 1) return CT_ERR_ARG;
	num_sweeps = num_sweeps_;
	maxiter = maxiter_;
	tol_split = tol_split_;
	max_vdim = max_vdim_;
	tol_energy = tol_energy_;
	d = (int)psi->d;
	qsite.assign(psi->qsite, psi->qsite + psi->d);
	A.clear(); W.clear();
	for (int i = 0; i < L; i++) { A.push_back(upload(&psi->a[i], st)); W.push_back(upload(&H->a[i], st)); }

	// mps_orthonormalize_qr(psi, MPS_ORTHONORMAL_RIGHT) (mps.c:609-760)
	for (int i = L - 1; i > 0; i--) if (local_orthonormalize_rq(A[i], A[i - 1], st) != 0) return -1;
	double nrm = 0;
	{
		DevTensor head = dummy_site(A[0].meta, 0, st);
		if (local_orthonormalize_rq(A[0], head, st) != 0) return -1;
		HostPacked hh = download_packed(head, st);
		if (hh.meta.nelem > 0) {
			nrm = hh.v[0];
			if (nrm < 0) { dev_scale(A[0], -1.0, st); nrm = -nrm; }
		}
	}
	if (nrm == 0) printf("Warning: in 'dmrg_twosite': initial MPS has norm zero (possibly due to mismatching quantum numbers)\n");

	lb.clear();
	rb.clear();
	lb.resize(L);
	rb.resize(L);
	rb[L - 1] = dummy_right(A[L - 1].meta, A[L - 1].meta, W[L - 1].meta, st);
	for (int i = L - 1; i > 0; i--) rb[i - 1] = step_right(A[i], A[i], W[i], rb[i], st);
	lb[0] = dummy_left(A[0].meta, A[0].meta, W[0].meta, st);
	h2.clear();
	for (int i = 0; i < L - 1; i++) h2.push_back(merge_mpo(W[i], W[i + 1], st));
	en_sweeps.clear();
	entropy.assign(L - 1, 0.0);
	sweep = 0;
	// dmrg.c:296-357: left-to-right over bonds 0 .. L-3, then right-to-left over L-2 .. 0
	dir = (L > 2) ? 0 : 1;
	bond = (L > 2) ? 0 : L - 2;
	done = false;
	theta = merge_mps(A[bond], A[bond + 1], st);
	plan_ahead_tables();
	CTG_HIP(hipStreamSynchronize(st));
	return 0;
}

void DmrgState::plan_ahead_tables()
{
	ahead.reset();
	if (!plan_ahead || done) return;
	const Meta& w = h2[bond].meta;
	const Meta& l = lb[bond].meta;
	const Meta& r = rb[bond + 1].meta;
	validate_matvec_operands(theta.meta, w, l, r);
	ahead.reset(new GroupedTables());
	plan_grouped_tables(*ahead, theta.meta, w, l, r, 0, 0, 0, 0, true);
	if (ahead->wa.kmax > WA_KMAX) {
		// the column stream does not apply: the full tables (a batch plan then replans anyway)
		ahead.reset(new GroupedTables());
		plan_grouped_tables(*ahead, theta.meta, w, l, r, 0, 0, 0, 0, false);
	}
}

// split of the optimised two-site tensor (mps_split_tensor_svd, mps.c:1119-1160) in three parts:
// device reshape and download, host block SVD with truncation, upload and device reshape
static double wall_now()
{
	return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void DmrgState::phase_mark()
{
	if (!trace) return;
	CTG_HIP(hipStreamSynchronize(st));
	phase_t0 = wall_now();
}

void DmrgState::phase_end(int ph)
{
	if (!trace) return;
	CTG_HIP(hipStreamSynchronize(st));
	const double t = wall_now();
	phase_s[ph] += t - phase_t0;
	phase_t0 = t;
}

void DmrgState::phase_report() const
{
	if (!trace) return;
	static const char* names[PH_COUNT] = { "plan", "lanczos", "split_begin", "split_host", "split_upload", "environment", "merge" };
	fprintf(stderr, "chemtensor_gpu: dmrg phases (L %d, max_vdim %lld, %d sweeps):", L, (long long)max_vdim, sweep);
	for (int i = 0; i < PH_COUNT; i++) fprintf(stderr, " %s %.4f s", names[i], phase_s[i]);
	fprintf(stderr, "\n");
}

void DmrgState::advance_begin(const DevTensor& a_opt)
{
	phase_mark();
	const ct_long dd[2] = { d, d };
	const int ad[2] = { TENSOR_AXIS_OUT, TENSOR_AXIS_OUT };
	const qnumber* qs[2] = { qsite.data(), qsite.data() };
	DevTensor a2 = dev_split_axis(a_opt, 1, dd, ad, qs, st);
	DevTensor tmp = dev_flatten_axes(a2, 0, TENSOR_AXIS_OUT, st);
	DevTensor am = dev_flatten_axes(tmp, 1, TENSOR_AXIS_IN, st);
	split_meta = a2.meta;
	split_on_device = false;
	split_prescaled = false;
	const int mode = split_mode();
	// the isometric factor: U when sweeping left-to-right (the singular values go right), else Vh
	const bool left_iso = (dir == 0);
	if (mode != SPLIT_HOST) {
		SvdLayout lay = plan_block_svd(am.meta);
		const bool fits = !lay.dummy && (mode == SPLIT_DEVICE_EIG ? block_eig_lds_bytes(lay, left_iso) <= block_eig_lds_limit()
			: block_svd_lds_bytes(lay) <= block_svd_lds_limit());

		if (fits) {
			const size_t ni = lay.qi.size(), nj = lay.jobs.size();
			DevTensor u(lay.um), vh(lay.vhm);
			double* sd = (double*)dev_alloc(ni * sizeof(double));
			int* stat = (int*)dev_alloc(nj * sizeof(int));
			std::vector<double> sv(ni);
			std::vector<int> status(nj);
			bool ok = (mode == SPLIT_DEVICE_EIG) ? launch_block_eig(lay, am.d, u.d, vh.d, sd, stat, left_iso, st)
				: launch_block_svd(lay, am.d, u.d, vh.d, sd, stat, st);
			if (ok) {
				CTG_HIP(hipMemcpyAsync(sv.data(), sd, ni * sizeof(double), hipMemcpyDeviceToHost, st));
				CTG_HIP(hipMemcpyAsync(status.data(), stat, nj * sizeof(int), hipMemcpyDeviceToHost, st));
				CTG_HIP(hipStreamSynchronize(st));
				for (int x : status) ok = ok && (x == 0);
			}
			dev_free(sd);
			dev_free(stat);
			if (ok) {
				// descending order within each block (stable, as the host SVD returns them)
				split_s.assign(ni, 0.0);
				split_pos.assign(ni, 0);
				for (const SvdLayout::Job& j : lay.jobs) {
					const ct_long k = std::min(j.r, j.c);
					std::vector<ct_long> ord((size_t)k);
					std::iota(ord.begin(), ord.end(), 0);
					std::stable_sort(ord.begin(), ord.end(), [&](ct_long x, ct_long y) { return sv[(size_t)(j.s_off + x)] > sv[(size_t)(j.s_off + y)]; });
					for (ct_long p = 0; p < k; p++) {
						split_s[(size_t)(j.s_off + p)] = sv[(size_t)(j.s_off + ord[p])];
						split_pos[(size_t)(j.s_off + p)] = ord[p];
					}
				}
				split_lay = std::move(lay);
				split_u = std::move(u);
				split_vh = std::move(vh);
				split_am = std::move(am);
				split_on_device = true;
				split_prescaled = (mode == SPLIT_DEVICE_EIG);
			}
		}
	}
	if (!split_on_device) split_in = download_packed(am, st);
	phase_end(PH_SPLIT_BEGIN);
}

void DmrgState::advance_host()
{
	const double t0 = trace ? wall_now() : 0;
	if (split_on_device) {
		split_num = host_retained_bond_indices(split_s, tol_split, true, max_vdim, split_ind, split_info);
		split_ret = 0;
	}
	else {
		split_ret = host_split_svd(split_in, tol_split, max_vdim, false, dir == 0 ? SVD_DISTR_RIGHT : SVD_DISTR_LEFT, split_m0, split_m1, split_info);
		split_in = HostPacked();
	}
	if (trace) phase_s[PH_SPLIT_HOST] += wall_now() - t0;
}

// how the split is factorised (CT_SPLIT, read at every split; the state's own choice when set):
// "device" (the default): Gram eigenvectors on the device (k_block_eig), the host factorisation for
// splits whose Gram matrices exceed the LDS of a workgroup or that do not converge; "svd": the
// one-sided Jacobi SVD kernel (k_block_svd), same fallback; "host": the host factorisation
int DmrgState::split_mode() const
{
	if (split_choice >= 0) return split_choice;
	const char* e = getenv("CT_SPLIT");
	if (e && std::string(e) == "host") return SPLIT_HOST;
	if (e && std::string(e) == "svd") return SPLIT_DEVICE_SVD;
	return SPLIT_DEVICE_EIG;
}

int DmrgState::advance_end(double en)
{
	if (split_ret < 0) return split_ret;
	phase_mark();
	if (split_on_device && split_num == 0) {
		// every singular value zero: the reference's path (a unit vector on the isometric side)
		split_in = download_packed(split_am, st);
		split_ret = host_split_svd(split_in, tol_split, max_vdim, false, dir == 0 ? SVD_DISTR_RIGHT : SVD_DISTR_LEFT, split_m0, split_m1, split_info);
		split_in = HostPacked();
		split_on_device = false;
		if (split_ret < 0) return split_ret;
	}
	{
		DevTensor d0, d1;
		if (split_on_device) {
			// retained columns of U and rows of Vh; the singular values go to the left factor when
			// sweeping right-to-left, to the right one otherwise (mps_split_tensor_svd)
			const bool left = (dir != 0);
			std::vector<double> sr(split_ind.size());
			for (size_t i = 0; i < sr.size(); i++) sr[i] = split_s[(size_t)split_ind[i]];
			Meta m0, m1;
			std::vector<SliceBlock> b;
			std::vector<int> ix;
			std::vector<double> sc;
			// (the Gram factorisation's non-isometric factor carries them already)
			const bool sc0 = left && !split_prescaled, sc1 = !left && !split_prescaled;
			plan_slice_2d(split_lay.um, 1, split_ind, split_pos, sc0 ? &sr : nullptr, m0, b, ix, sc);
			d0 = device_slice_2d(split_u, m0, b, ix, sc, sc0, st);
			plan_slice_2d(split_lay.vhm, 0, split_ind, split_pos, sc1 ? &sr : nullptr, m1, b, ix, sc);
			d1 = device_slice_2d(split_vh, m1, b, ix, sc, sc1, st);
			split_u = DevTensor();
			split_vh = DevTensor();
			split_am = DevTensor();
			split_on_device = false;
		}
		else {
			d0 = upload_packed(split_m0, st);
			d1 = upload_packed(split_m1, st);
		}
		const Meta& m = split_meta;
		const ct_long dl0[2] = { m.dim_logical[0], m.dim_logical[1] }, dl1[2] = { m.dim_logical[2], m.dim_logical[3] };
		const int ad0[2] = { m.axis_dir[0], m.axis_dir[1] }, ad1[2] = { m.axis_dir[2], m.axis_dir[3] };
		const qnumber* qn0[2] = { m.qnums_logical[0].data(), m.qnums_logical[1].data() };
		const qnumber* qn1[2] = { m.qnums_logical[2].data(), m.qnums_logical[3].data() };
		A[bond] = dev_split_axis(d0, 0, dl0, ad0, qn0, st);
		A[bond + 1] = dev_split_axis(d1, 1, dl1, ad1, qn1, st);
		split_m0 = HostPacked();
		split_m1 = HostPacked();
	}
	phase_end(PH_SPLIT_UPLOAD);
	const int i = bond;
	if (dir == 0) {
		lb[i + 1] = step_left(A[i], A[i], W[i], lb[i], st);
		if (i + 1 < L - 2) bond = i + 1;
		else { dir = 1; bond = L - 2; }
	}
	else {
		entropy[i] = split_info.entropy;
		rb[i] = step_right(A[i + 1], A[i + 1], W[i + 1], rb[i + 1], st);
		if (i > 0) bond = i - 1;
		else {
			DevTensor t = dummy_site(A[0].meta, 0, st);
			if (local_orthonormalize_rq(A[0], t, st) != 0) return -1;
			en_sweeps.push_back(en);
			sweep++;
			const int ns = (int)en_sweeps.size();
			const bool converged = tol_energy > 0 && ns >= 2 && std::fabs(en_sweeps[ns - 1] - en_sweeps[ns - 2]) < tol_energy;
			if (sweep >= num_sweeps || converged) done = true;
			dir = (L > 2) ? 0 : 1;
			bond = (L > 2) ? 0 : L - 2;
		}
	}
	phase_end(PH_ENVIRONMENT);
	if (!done) theta = merge_mps(A[bond], A[bond + 1], st);
	else theta = DevTensor();
	plan_ahead_tables();
	CTG_HIP(hipStreamSynchronize(st));
	phase_end(PH_MERGE);
	return 0;
}

int DmrgState::advance(double en, DevTensor& a_opt)
{
	advance_begin(a_opt);
	advance_host();
	return advance_end(en);
}

void DmrgState::download_mps(mps* psi) const
{
	for (int i = 0; i < L; i++) {
		delete_block_sparse_tensor(&psi->a[i]);
		download(A[i], &psi->a[i], st);
	}
}

size_t DmrgState::device_bytes() const
{
	size_t n = 0;
	for (const auto* v : { &A, &W, &lb, &rb, &h2 })
		for (const DevTensor& t : *v) n += (size_t)t.meta.nelem;
	return n * sizeof(double);
}

// the matvec form of the sweep's Lanczos: grouped FP64, or with CT_DMRG_MATVEC=split the grouped
// form on FP32 MFMA with split operands (the Lanczos scalars, the tridiagonal eigensolver and the
// split stay FP64), to measure the precision option's energy error over whole sweeps
static int dmrg_matvec_mode()
{
	const char* e = getenv("CT_DMRG_MATVEC");
	return (e && std::string(e) == "split") ? CT_MATVEC_GROUPED_SPLIT : CT_MATVEC_GROUPED_FP64;
}

int dmrg_twosite_device(const mpo* H, int num_sweeps, int maxiter, double tol_split, ct_long max_vdim, mps* psi, double* en_sweeps, double* entropy)
{
	DmrgState S;
	S.st = default_stream();
	const char* tr = getenv("CT_DMRG_PHASES");
	S.trace = tr && tr[0] == '1';
	int ret = S.init(H, psi, num_sweeps, maxiter, tol_split, max_vdim, 0.0);
	if (ret != 0) return ret;
	while (!S.done) {
		DevTensor ao;
		double en = 0;
		// minimize_local_energy (dmrg.c:86-147) with the Lanczos iteration on the device
		S.phase_mark();
		std::unique_ptr<MatvecPlan> p(build_matvec_plan_dev(S.theta.meta, DevTensor::view(S.h2[S.bond], S.h2[S.bond].meta),
			DevTensor::view(S.lb[S.bond], S.lb[S.bond].meta), DevTensor::view(S.rb[S.bond + 1], S.rb[S.bond + 1].meta), dmrg_matvec_mode(), S.st, false, false));
		S.phase_end(DmrgState::PH_PLAN);
		ao = DevTensor(S.theta.meta);
		ret = lanczos_ground_state(*p, S.theta.d, maxiter, &en, ao.d);
		S.phase_end(DmrgState::PH_LANCZOS);
		if (ret < 0) return ret;
		ret = S.advance(en, ao);
		if (ret < 0) return ret;
	}
	S.phase_report();
	for (int n = 0; n < num_sweeps; n++) en_sweeps[n] = S.en_sweeps[n];
	for (int i = 0; i < S.L - 1; i++) entropy[i] = S.entropy[i];
	S.download_mps(psi);
	return 0;
}

}  // namespace ctg

extern "C" int dmrg_twosite(const struct mpo* hamiltonian, const int num_sweeps, const int maxiter_lanczos, const double tol_split, const ct_long max_vdim,
	struct mps* psi, double* en_sweeps, double* entropy)
{
	try {
		return ctg::dmrg_twosite_device(hamiltonian, num_sweeps, maxiter_lanczos, tol_split, max_vdim, psi, en_sweeps, entropy);
	}
	catch (const ctg::HipError& e) {
		fprintf(stderr, "chemtensor_gpu: dmrg_twosite failed: %s\n", e.where.c_str());
		return (e.code == hipErrorNotSupported) ? CT_ERR_UNSUPPORTED : CT_ERR_HIP;
	}
}
