// Host FP64 block factorisations of the DMRG sweep on packed block-sparse matrices: block SVD
// with truncation (src/tensor/block_sparse_tensor.c:2686-2848, src/algorithm/truncation.c:110-222,
// src/algorithm/bond_ops.c:15-138), block QR / RQ (block_sparse_tensor.c:2402-2678), slicing
// (:1446-1552) and pointwise scaling (:1654-1705). The bond bookkeeping (order of the new
// bond's quantum numbers, dummy bonds of empty matrices) follows the reference exactly.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "ct_internal.h"

#include <atomic>
#include <numeric>
#include <thread>

namespace ctg {

// host threads lent to block SVDs across all concurrent callers (at most the core count in use)
static std::atomic<int> svd_threads_busy(0);

static int host_svd_threads(int want)
{
	const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
	const int cap = std::min(16, hw);
	int busy = svd_threads_busy.load();
	for (;;) {
		const int n = std::min(want, cap - busy);
		if (n <= 1) return 1;   // nothing taken: the caller runs the blocks itself
		if (svd_threads_busy.compare_exchange_weak(busy, busy + n)) return n;
	}
}

static void host_svd_threads_release(int n) { svd_threads_busy -= n; }

// set by threads of a caller that already keeps every core busy with independent splits (the
// serving pool): their blocks are factorised one after the other, without helper threads
static thread_local bool tl_svd_serial = false;

void host_svd_serial_on_this_thread(bool on) { tl_svd_serial = on; }

}  // namespace ctg

namespace ctg {

static Meta matrix_meta(ct_long d0, const qnumber* q0, ct_long d1, const qnumber* q1, int ad0, int ad1)
{
	const ct_long dim[2] = { d0, d1 };
	const int ad[2] = { ad0, ad1 };
	const qnumber* qn[2] = { q0, q1 };
	return make_meta(2, dim, ad, qn);
}

SvdLayout plan_block_svd(const Meta& m)
{
	SvdLayout L;
	for (ct_long j = 0; j < m.dim_blocks[1]; j++) for (ct_long i = 0; i < m.dim_blocks[0]; i++) {
		if (m.axis_dir[0] * m.qnums_blocks[0][i] + m.axis_dir[1] * m.qnums_blocks[1][j] != 0) continue;
		const ct_long k = std::min(m.block_count[0][i], m.block_count[1][j]);
		for (ct_long l = 0; l < k; l++) L.qi.push_back(m.qnums_blocks[1][j]);
	}
	L.dummy = L.qi.empty();
	if (L.dummy) L.qi.push_back(-m.axis_dir[0] * m.axis_dir[1] * m.qnums_blocks[0][0]);
	const ct_long ni = (ct_long)L.qi.size();
	L.um = matrix_meta(m.dim_logical[0], m.qnums_logical[0].data(), ni, L.qi.data(), m.axis_dir[0], m.axis_dir[1]);
	L.vhm = matrix_meta(ni, L.qi.data(), m.dim_logical[1], m.qnums_logical[1].data(), -m.axis_dir[1], m.axis_dir[1]);
	if (L.dummy) return L;
	for (ct_long i = 0; i < m.dim_blocks[0]; i++) for (ct_long j = 0; j < m.dim_blocks[1]; j++) {
		const ct_long ib[2] = { i, j };
		const ct_long ka = m.flat_index(ib);
		if (!m.present(ka)) continue;
		const ct_long r = m.block_count[0][i], c = m.block_count[1][j];
		const ct_long kq = L.um.find_qnum(1, m.qnums_blocks[1][j]);
		const ct_long iu[2] = { i, kq }, iv[2] = { kq, j };
		ct_long l0 = 0;
		while (l0 < ni && L.qi[l0] != m.qnums_blocks[1][j]) l0++;
		L.jobs.push_back(SvdLayout::Job{ m.offset[ka], L.um.offset[L.um.flat_index(iu)], L.vhm.offset[L.vhm.flat_index(iv)], l0, r, c });
	}
	return L;
}

int host_block_svd(const HostPacked& a, HostPacked& u, std::vector<double>& s, HostPacked& vh, int iso)
{
	const SvdLayout lay = plan_block_svd(a.meta);
	const ct_long ni = (ct_long)lay.qi.size();
	u.meta = lay.um;
	u.v.assign((size_t)u.meta.nelem, 0.0);
	vh.meta = lay.vhm;
	vh.v.assign((size_t)vh.meta.nelem, 0.0);
	s.assign((size_t)ni, 0.0);
	if (lay.dummy) {
		const ct_long ib[2] = { 0, u.meta.find_qnum(1, lay.qi[0]) };
		u.v[(size_t)u.meta.offset[u.meta.flat_index(ib)]] = 1;
		return 0;
	}
	// the blocks are independent: large splits run them in parallel host threads
	struct Job { const double* a; double* ub; double* vb; double* sb; ct_long r, c; double work; };
	std::vector<Job> jobs;
	double total = 0;
	for (const SvdLayout::Job& j : lay.jobs) {
		const ct_long k = std::min(j.r, j.c);
		const double w = (double)std::max(j.r, j.c) * (double)k * (double)k;
		jobs.push_back(Job{ a.v.data() + j.a_off, u.v.data() + j.u_off, vh.v.data() + j.vh_off, s.data() + j.s_off, j.r, j.c, w });
		total += w;
	}
	std::atomic<int> ret(0);
	auto run = [&](size_t i) {
		const Job& j = jobs[i];
		const int r = (iso < 0) ? host_svd(j.r, j.c, j.a, j.ub, j.sb, j.vb) : host_svd_factor(j.r, j.c, j.a, iso, j.ub, j.sb, j.vb);
		if (r != 0) ret = -1;
	};
	const int nt = (total > 4e6 && !tl_svd_serial) ? host_svd_threads((int)jobs.size()) : 1;
	if (nt <= 1) {
		for (size_t i = 0; i < jobs.size(); i++) run(i);
	}
	else {
		// largest blocks first, dealt dynamically
		std::vector<size_t> order(jobs.size());
		std::iota(order.begin(), order.end(), 0);
		std::sort(order.begin(), order.end(), [&](size_t x, size_t y) { return jobs[x].work > jobs[y].work; });
		std::atomic<size_t> next(0);
		std::vector<std::thread> th;
		for (int t = 0; t < nt; t++)
			th.emplace_back([&] { for (size_t q; (q = next++) < order.size();) run(order[q]); });
		for (auto& x : th) x.join();
		host_svd_threads_release(nt);
	}
	return ret;
}

ct_long host_retained_bond_indices(const std::vector<double>& sigma, double tol, bool relative, ct_long max_vdim, std::vector<ct_long>& ind, trunc_info& info)
{
	const ct_long n = (ct_long)sigma.size();
	info.tol_eff = tol;
	std::vector<std::pair<double, ct_long>> ss((size_t)n);
	for (ct_long i = 0; i < n; i++) ss[i] = { sigma[i], i };
	// stable order of equal values, as glibc's merge-sort qsort in the reference
	std::stable_sort(ss.begin(), ss.end(), [](const std::pair<double, ct_long>& x, const std::pair<double, ct_long>& y) { return x.first < y.first; });
	double sq = 0;
	for (auto& p : ss) { p.first = p.first * p.first; sq += p.first; }
	ind.clear();
	if (sq == 0) { info.norm_sigma = 0; info.entropy = 0; return 0; }
	if (relative) for (auto& p : ss) p.first /= sq;
	for (ct_long i = 1; i < n; i++) ss[i].first += ss[i - 1].first;
	if (max_vdim < n) {
		info.tol_eff = std::fmax(tol, ss[n - max_vdim - 1].first);
		for (ct_long i = 0; i < n - max_vdim; i++) ss[i].first = 0;
	}
	std::vector<double> acc((size_t)n);
	for (auto& p : ss) acc[p.second] = p.first;
	for (ct_long i = 0; i < n; i++) if (acc[i] > tol) ind.push_back(i);
	if (ind.empty()) { info.norm_sigma = 0; info.entropy = 0; return 0; }
	double scale = 0, ssq = 1;
	for (ct_long i : ind) {
		const double v = std::fabs(sigma[i]);
		if (v == 0) continue;
		if (scale < v) { ssq = 1 + ssq * (scale / v) * (scale / v); scale = v; }
		else ssq += (v / scale) * (v / scale);
	}
	info.norm_sigma = scale * std::sqrt(ssq);
	double ent = 0;
	for (ct_long i : ind) {
		const double x = sigma[i] / info.norm_sigma;
		if (x > 0) { const double q = x * x; ent -= q * std::log(q); }
	}
	info.entropy = ent;
	return (ct_long)ind.size();
}

HostPacked host_slice(const HostPacked& t, int i_ax, const std::vector<ct_long>& ind)
{
	const Meta& m = t.meta;
	const int nd = m.ndim;
	std::vector<qnumber> qs(ind.size());
	for (size_t j = 0; j < ind.size(); j++) qs[j] = m.qnums_logical[i_ax][ind[j]];
	std::vector<ct_long> dl(m.dim_logical);
	dl[i_ax] = (ct_long)ind.size();
	std::vector<const qnumber*> qn(nd);
	for (int i = 0; i < nd; i++) qn[i] = (i == i_ax ? qs.data() : m.qnums_logical[i].data());
	HostPacked r;
	r.meta = make_meta(nd, dl.data(), m.axis_dir.data(), qn.data());
	r.v.assign((size_t)r.meta.nelem, 0.0);
	// once per axis: the position of every source index within its block, and the (new) logical
	// indices of every block of the result axis in order
	std::vector<ct_long> fin((size_t)m.dim_logical[i_ax]);
	{
		std::vector<ct_long> cnt((size_t)m.dim_blocks[i_ax], 0);
		for (ct_long j = 0; j < m.dim_logical[i_ax]; j++) fin[(size_t)j] = cnt[(size_t)m.find_qnum(i_ax, m.qnums_logical[i_ax][j])]++;
	}
	std::vector<std::vector<ct_long>> fo((size_t)r.meta.dim_blocks[i_ax]);
	for (ct_long j = 0; j < dl[i_ax]; j++) fo[(size_t)r.meta.find_qnum(i_ax, qs[(size_t)j])].push_back(j);
	ct_long ir[16], it[16], br[16], bt[16];
	for (ct_long k = 0; k < r.meta.nblocks; k++) {
		if (!r.meta.present(k)) continue;
		r.meta.block_index(k, ir);
		r.meta.block_dims(k, br);
		const qnumber q = r.meta.qnums_blocks[i_ax][ir[i_ax]];
		std::memcpy(it, ir, sizeof(ct_long) * nd);
		it[i_ax] = m.find_qnum(i_ax, q);
		const ct_long kt = m.flat_index(it);
		m.block_dims(kt, bt);
		const std::vector<ct_long>& fob = fo[(size_t)ir[i_ax]];
		ct_long outer = 1, inner = 1;
		for (int i = 0; i < i_ax; i++) outer *= bt[i];
		for (int i = i_ax + 1; i < nd; i++) inner *= bt[i];
		const double* src = t.v.data() + m.offset[kt];
		double* dst = r.v.data() + r.meta.offset[k];
		for (ct_long o = 0; o < outer; o++) for (ct_long j = 0; j < br[i_ax]; j++) {
			const ct_long js = fin[(size_t)ind[(size_t)fob[(size_t)j]]];
			const double* sp = src + (o * bt[i_ax] + js) * inner;
			double* dp = dst + (o * br[i_ax] + j) * inner;
			if (inner == 1) *dp = *sp;
			else std::memcpy(dp, sp, (size_t)inner * sizeof(double));
		}
	}
	return r;
}

void host_scale_axis(HostPacked& t, bool leading, const std::vector<double>& vec)
{
	const Meta& m = t.meta;
	const int ax = leading ? 0 : m.ndim - 1;
	ct_long ib[16], bd[16];
	for (ct_long k = 0; k < m.nblocks; k++) {
		if (!m.present(k)) continue;
		m.block_index(k, ib);
		m.block_dims(k, bd);
		const std::vector<ct_long> map = fanout(m, ax, m.qnums_blocks[ax][ib[ax]]);
		const ct_long n = m.block_size(k), nt = bd[ax];
		double* d = t.v.data() + m.offset[k];
		if (leading) { const ct_long inner = n / nt; for (ct_long j = 0; j < nt; j++) for (ct_long i = 0; i < inner; i++) d[j * inner + i] *= vec[map[j]]; }
		else { const ct_long outer = n / nt; for (ct_long i = 0; i < outer; i++) for (ct_long j = 0; j < nt; j++) d[i * nt + j] *= vec[map[j]]; }
	}
}

int host_split_svd(const HostPacked& a, double tol, ct_long max_vdim, bool renormalize, int distr, HostPacked& a0, HostPacked& a1, trunc_info& info)
{
	// the factor that carries the singular values comes out of the factorisation already scaled
	// (projection onto the isometric factor), so no division by small singular values is needed
	const int iso = (distr == SVD_DISTR_LEFT) ? 1 : 0;
	HostPacked u, vh;
	std::vector<double> s;
	int ret = host_block_svd(a, u, s, vh, iso);
	if (ret < 0) return ret;
	std::vector<ct_long> ind;
	const ct_long num = host_retained_bond_indices(s, tol, true, max_vdim, ind, info);
	if (num == 0) {
		const std::vector<ct_long> i0 = { 0 };
		a0 = host_slice(u, 1, i0);
		a1 = host_slice(vh, 0, i0);
		const std::vector<double> zero = { 0.0 };
		if (distr == SVD_DISTR_LEFT) host_scale_axis(a0, false, zero);
		else host_scale_axis(a1, true, zero);
		return 0;
	}
	a0 = host_slice(u, 1, ind);
	a1 = host_slice(vh, 0, ind);
	if (renormalize) {
		double all = 0;
		for (double x : s) all += x * x;
		const double scale = std::sqrt(all) / info.norm_sigma;
		HostPacked& t = (distr == SVD_DISTR_LEFT) ? a0 : a1;
		for (double& x : t.v) x *= scale;
	}
	return 0;
}

int host_block_qr(const HostPacked& a, HostPacked& q, HostPacked& r)
{
	const Meta& m = a.meta;
	std::vector<qnumber> qi;
	for (ct_long j = 0; j < m.dim_blocks[1]; j++) for (ct_long i = 0; i < m.dim_blocks[0]; i++) {
		if (m.axis_dir[0] * m.qnums_blocks[0][i] + m.axis_dir[1] * m.qnums_blocks[1][j] != 0) continue;
		const ct_long k = std::min(m.block_count[0][i], m.block_count[1][j]);
		for (ct_long l = 0; l < k; l++) qi.push_back(m.qnums_blocks[1][j]);
	}
	const bool dummy = qi.empty();
	if (dummy) qi.push_back(-m.axis_dir[0] * m.axis_dir[1] * m.qnums_blocks[0][0]);
	const ct_long ni = (ct_long)qi.size();
	q.meta = matrix_meta(m.dim_logical[0], m.qnums_logical[0].data(), ni, qi.data(), m.axis_dir[0], m.axis_dir[1]);
	q.v.assign((size_t)q.meta.nelem, 0.0);
	r.meta = matrix_meta(ni, qi.data(), m.dim_logical[1], m.qnums_logical[1].data(), -m.axis_dir[1], m.axis_dir[1]);
	r.v.assign((size_t)r.meta.nelem, 0.0);
	if (dummy) {
		const ct_long ib[2] = { 0, q.meta.find_qnum(1, qi[0]) };
		q.v[(size_t)q.meta.offset[q.meta.flat_index(ib)]] = 1;
		return 0;
	}
	for (ct_long i = 0; i < m.dim_blocks[0]; i++) for (ct_long j = 0; j < m.dim_blocks[1]; j++) {
		const ct_long ib[2] = { i, j };
		const ct_long ka = m.flat_index(ib);
		if (!m.present(ka)) continue;
		const ct_long kq = q.meta.find_qnum(1, m.qnums_blocks[1][j]);
		const ct_long iq[2] = { i, kq }, ir[2] = { kq, j };
		host_qr(m.block_count[0][i], m.block_count[1][j], a.v.data() + m.offset[ka],
			q.v.data() + q.meta.offset[q.meta.flat_index(iq)], r.v.data() + r.meta.offset[r.meta.flat_index(ir)]);
	}
	return 0;
}

int host_block_rq(const HostPacked& a, HostPacked& r, HostPacked& q)
{
	const Meta& m = a.meta;
	std::vector<qnumber> qi;
	for (ct_long i = 0; i < m.dim_blocks[0]; i++) for (ct_long j = 0; j < m.dim_blocks[1]; j++) {
		if (m.axis_dir[0] * m.qnums_blocks[0][i] + m.axis_dir[1] * m.qnums_blocks[1][j] != 0) continue;
		const ct_long k = std::min(m.block_count[0][i], m.block_count[1][j]);
		for (ct_long l = 0; l < k; l++) qi.push_back(m.qnums_blocks[0][i]);
	}
	const bool dummy = qi.empty();
	if (dummy) qi.push_back(-m.axis_dir[0] * m.axis_dir[1] * m.qnums_blocks[1][0]);
	const ct_long ni = (ct_long)qi.size();
	r.meta = matrix_meta(m.dim_logical[0], m.qnums_logical[0].data(), ni, qi.data(), m.axis_dir[0], -m.axis_dir[0]);
	r.v.assign((size_t)r.meta.nelem, 0.0);
	q.meta = matrix_meta(ni, qi.data(), m.dim_logical[1], m.qnums_logical[1].data(), m.axis_dir[0], m.axis_dir[1]);
	q.v.assign((size_t)q.meta.nelem, 0.0);
	if (dummy) {
		const ct_long ib[2] = { q.meta.find_qnum(0, qi[0]), 0 };
		q.v[(size_t)q.meta.offset[q.meta.flat_index(ib)]] = 1;
		return 0;
	}
	for (ct_long i = 0; i < m.dim_blocks[0]; i++) for (ct_long j = 0; j < m.dim_blocks[1]; j++) {
		const ct_long ib[2] = { i, j };
		const ct_long ka = m.flat_index(ib);
		if (!m.present(ka)) continue;
		const ct_long kq = q.meta.find_qnum(0, m.qnums_blocks[0][i]);
		const ct_long ir[2] = { i, kq }, iq[2] = { kq, j };
		host_rq(m.block_count[0][i], m.block_count[1][j], a.v.data() + m.offset[ka],
			r.v.data() + r.meta.offset[r.meta.flat_index(ir)], q.v.data() + q.meta.offset[q.meta.flat_index(iq)]);
	}
	return 0;
}

}  // namespace ctg
