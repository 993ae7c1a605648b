// Device tensor algebra of libchemtensor_gpu: host-side contraction planning from the
// quantum-number tables (the block pairing of src/tensor/block_sparse_tensor.c:1826-2000) and
// the GPU execution of contractions, transposes and partial traces on packed device tensors;
// the H_eff matvec plan of src/algorithm/chain_ops.c:353-390 in its grouped (three launches,
// transposes folded into operand addressing) and plain (per-block launches) forms.

#include <algorithm>
#include <map>
#include <cstring>
#include <numeric>
#include <unordered_map>

#include "ct_internal.h"

namespace ctg {

static const double* off_ptr(ct_long elems) { return (const double*)(intptr_t)(elems * (ct_long)sizeof(double)); }
static double* off_ptr_mut(ct_long elems) { return (double*)(intptr_t)(elems * (ct_long)sizeof(double)); }

//________________________________________________________________________________________________________________________
//
// grouped GEMM batches

void GemmBatch::add_task(double* C, int M, int N, int nci, int sc_i, int64_t sc_batch, bool accumulate)
{
	GemmTask t;
	t.C = C; t.sc_batch = sc_batch; t.M = M; t.N = N; t.nci = nci > 0 ? nci : 1; t.sc_i = sc_i;
	t.t0 = (int32_t)terms.size(); t.t1 = t.t0; t.accumulate = accumulate ? 1 : 0; t.pad = 0;
	tasks.push_back(t);
}

void GemmBatch::add_term(const double* A, const double* B, int K, int sa_i, int sa_k, int sb_k, int sb_j, int64_t sb_batch)
{
	GemmTerm m;
	m.A = A; m.B = B; m.sb_batch = sb_batch; m.K = K; m.sa_i = sa_i; m.sa_k = sa_k; m.sb_k = sb_k; m.sb_j = sb_j; m.pad = 0;
	terms.push_back(m);
	GemmTask& t = tasks.back();
	t.t1 = (int32_t)terms.size();
	flops += 2.0 * (double)t.M * (double)t.N * (double)K;
}

std::vector<WorkItem> GemmBatch::work_items() const
{
	// tiles in order of decreasing cost, ties in task and tile order: the (cost, tile) keys are
	// sorted, then the items are written once in that order
	struct Key { int cost; int32_t task; int m0, n0; };
	std::vector<Key> keys;
	for (size_t ti = 0; ti < tasks.size(); ti++) {
		const GemmTask& t = tasks[ti];
		int64_t ksum = 0;
		for (int q = t.t0; q < t.t1; q++) ksum += terms[q].K;
		const int cost = (int)std::min<int64_t>(ksum + 4, (int64_t)1 << 30);
		for (int m0 = 0; m0 < t.M; m0 += TILE)
			for (int n0 = 0; n0 < t.N; n0 += TILE) keys.push_back(Key{ cost, (int32_t)ti, m0, n0 });
	}
	std::stable_sort(keys.begin(), keys.end(), [](const Key& a, const Key& b) { return a.cost > b.cost; });
	std::vector<WorkItem> items;
	items.reserve(keys.size());
	for (const Key& k : keys) {
		const GemmTask& t = tasks[(size_t)k.task];
		items.push_back(WorkItem{ t, k.m0, k.n0, k.cost, k.task, t.t1 > t.t0 ? terms[t.t0] : GemmTerm{} });
	}
	return items;
}

// the LDS-tiled form's 64 x 64 tiles of the large tasks, and the other tasks' items (both empty when
// no task is large). Workgroup i runs on XCD i mod 8, and each XCD has its own L2: the tiles of a
// task are cut into groups of up to 4 x 8 tiles (sharing A rows and B columns), the groups are
// dealt to the eight XCDs by load (largest first), and the item order interleaves the XCDs' queues,
// so a group's tiles run together on one XCD and fetch their operand panels into one L2.
void split_lds_items(const GemmBatch& b, const std::vector<WorkItem>& it, std::vector<WorkItem>& lit, std::vector<WorkItem>& rit)
{
	constexpr int NXCD = 8, GM = 4, GN = 8;
	lit.clear();
	rit.clear();
	std::vector<uint8_t> big(b.tasks.size(), 0);
	struct Group { int64_t cost; int32_t task; int mi0, mi1, ni0, ni1; };
	std::vector<Group> groups;
	// cost of a tile: its k-panels times the fraction of its 16 x 16 sub-tiles inside the task (the
	// kernel runs the MFMAs of those only), plus one for the store
	auto tile_cost = [](const GemmTask& t, int64_t panels, int mi, int ni) {
		const int sr = std::min(4, (t.M - mi * LDS_TILE + 15) / 16), sc = std::min(4, (t.N - ni * LDS_TILE + 15) / 16);
		return std::min<int64_t>((panels * sr * sc + 15) / 16 + 1, (int64_t)1 << 30);
	};
	std::vector<int64_t> task_panels(b.tasks.size(), 0);
	for (size_t ti = 0; ti < b.tasks.size(); ti++) {
		const GemmTask& t = b.tasks[ti];
		if (t.M < LDS_TASK_MIN || t.N < LDS_TASK_MIN || t.t1 <= t.t0) continue;
		big[ti] = 1;
		int64_t panels = 0;
		for (int q = t.t0; q < t.t1; q++) panels += (b.terms[q].K + 15) / 16;
		task_panels[ti] = panels;
		const int mt = (t.M + LDS_TILE - 1) / LDS_TILE, nt = (t.N + LDS_TILE - 1) / LDS_TILE;
		for (int mi = 0; mi < mt; mi += GM)
			for (int ni = 0; ni < nt; ni += GN) {
				const int mi1 = std::min(mt, mi + GM), ni1 = std::min(nt, ni + GN);
				int64_t c = 0;
				for (int a = mi; a < mi1; a++)
					for (int bb = ni; bb < ni1; bb++) c += tile_cost(t, panels, a, bb);
				groups.push_back(Group{ c, (int32_t)ti, mi, mi1, ni, ni1 });
			}
	}
	if (groups.empty()) return;
	std::stable_sort(groups.begin(), groups.end(), [](const Group& x, const Group& y) { return x.cost > y.cost; });
	std::vector<std::vector<WorkItem>> q(NXCD);
	std::vector<int64_t> load(NXCD, 0);
	for (const Group& g : groups) {
		const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
		load[(size_t)x] += g.cost;
		const GemmTask& t = b.tasks[(size_t)g.task];
		for (int mi = g.mi0; mi < g.mi1; mi++)
			for (int ni = g.ni0; ni < g.ni1; ni++)
				q[(size_t)x].push_back(WorkItem{ t, mi * LDS_TILE, ni * LDS_TILE, (int)tile_cost(t, task_panels[(size_t)g.task], mi, ni), g.task, b.terms[t.t0] });
	}
	size_t longest = 0;
	for (const auto& v : q) longest = std::max(longest, v.size());
	for (size_t jj = 0; jj < longest; jj++)
		for (int x = 0; x < NXCD; x++)
			if (jj < q[(size_t)x].size()) lit.push_back(q[(size_t)x][jj]);
	for (const WorkItem& w : it) if (!big[(size_t)w.task_index]) rit.push_back(w);
}

void DevGemm::upload(const GemmBatch& b, hipStream_t s)
{
	release();
	const std::vector<WorkItem> it = b.work_items();
	ntasks = (int)b.tasks.size();
	nterms = (int)b.terms.size();
	nitems = (int)it.size();
	flops = b.flops;
	// the three tables in one allocation and one transfer (256-byte aligned parts)
	std::vector<WorkItem> lit, rit;
	split_lds_items(b, it, lit, rit);
	auto al = [](size_t n) { return (n + 255) / 256 * 256; };
	const size_t n_tasks = b.tasks.size() * sizeof(GemmTask), n_terms = b.terms.size() * sizeof(GemmTerm), n_items = it.size() * sizeof(WorkItem);
	const size_t n_lit = lit.size() * sizeof(WorkItem), n_rit = rit.size() * sizeof(WorkItem);
	const size_t o_terms = al(std::max<size_t>(n_tasks, 1)), o_items = o_terms + al(std::max<size_t>(n_terms, 1));
	const size_t o_lit = o_items + al(std::max<size_t>(n_items, 1)), o_rit = o_lit + al(n_lit);
	const size_t total = o_rit + std::max<size_t>(n_rit, 1);
	buf = (char*)dev_alloc(total);
	tasks = (GemmTask*)buf;
	terms = (GemmTerm*)(buf + o_terms);
	items = (WorkItem*)(buf + o_items);
	litems = (WorkItem*)(buf + o_lit);
	ritems = (WorkItem*)(buf + o_rit);
	nlitems = (int)lit.size();
	nritems = (int)rit.size();
	use_lds = false;
	const H2DPart parts[5] = { { b.tasks.data(), n_tasks, 0 }, { b.terms.data(), n_terms, o_terms }, { it.data(), n_items, o_items },
		{ lit.data(), n_lit, o_lit }, { rit.data(), n_rit, o_rit } };
	copy_h2d_parts(buf, parts, 5, o_rit + n_rit, s);
	const int cap = num_compute_units() * 8;
	grid = std::max(1, std::min((nitems + 3) / 4, cap));
	// workgroup-level split of the k-step stream only pays for long streams
	double steps = 0;
	for (const GemmTask& t : b.tasks) {
		double ts = 0;
		for (int q = t.t0; q < t.t1; q++) ts += (b.terms[q].K + 3) / 4;
		steps += ts * (double)(((t.M + TILE - 1) / TILE) * ((t.N + TILE - 1) / TILE));
	}
	const double avg = nitems ? steps / nitems : 0;
	form = (avg > 8.0) ? GEMM_WG_CHUNK : GEMM_WAVE_STREAM;
	avg_steps = avg;
}

void DevGemm::release()
{
	dev_free(buf);
	buf = nullptr;
	tasks = nullptr; terms = nullptr; items = nullptr; litems = nullptr; ritems = nullptr;
	ntasks = nterms = nitems = nlitems = nritems = 0;
	use_lds = false;
}

static void run_batch_plain(const GemmBatch& b, GemmBases bases, hipStream_t s)
{
	for (const GemmTask& t : b.tasks) {
		for (int q = t.t0; q < t.t1; q++) {
			GemmTask one = t;
			one.accumulate = 1;
			launch_plain_gemm_f64(one, b.terms[q], bases, s);
		}
	}
}

//________________________________________________________________________________________________________________________
//
// block_sparse_tensor_dot planning (bst.c:1826-2000)

static Meta dot_result_meta(const Meta& s, int axs, const Meta& t, int axt, int nm)
{
	const int offset_s = (axs == TENSOR_AXIS_RANGE_LEADING ? nm : 0);
	const int offset_t = (axt == TENSOR_AXIS_RANGE_LEADING ? nm : 0);
	const int nds = s.ndim - nm, ndt = t.ndim - nm;
	std::vector<ct_long> dl;
	std::vector<int> ad;
	std::vector<const qnumber*> qn;
	for (int i = 0; i < nds; i++) { dl.push_back(s.dim_logical[offset_s + i]); ad.push_back(s.axis_dir[offset_s + i]); qn.push_back(s.qnums_logical[offset_s + i].data()); }
	for (int i = 0; i < ndt; i++) { dl.push_back(t.dim_logical[offset_t + i]); ad.push_back(t.axis_dir[offset_t + i]); qn.push_back(t.qnums_logical[offset_t + i].data()); }
	return make_meta(nds + ndt, dl.data(), ad.data(), qn.data());
}

/// plan r = s . t: tasks (one per output block) with terms (one per contracted block pair);
/// operand pointers are element offsets into s, t and r (as byte offsets)
static GemmBatch plan_dot(const Meta& s, int axs, const Meta& t, int axt, int nm, const Meta& r, bool accumulate)
{
	const int shift_s = (axs == TENSOR_AXIS_RANGE_LEADING ? 0 : s.ndim - nm);
	const int shift_t = (axt == TENSOR_AXIS_RANGE_LEADING ? 0 : t.ndim - nm);
	const int offset_s = (axs == TENSOR_AXIS_RANGE_LEADING ? nm : 0);
	const int offset_t = (axt == TENSOR_AXIS_RANGE_LEADING ? nm : 0);
	const int nds = s.ndim - nm, ndt = t.ndim - nm;
	const int nlds = (axs == TENSOR_AXIS_RANGE_LEADING ? nm : s.ndim - nm);
	const int nldt = (axt == TENSOR_AXIS_RANGE_LEADING ? nm : t.ndim - nm);
	const bool transa = (axs == TENSOR_AXIS_RANGE_LEADING);
	const bool transb = (axt == TENSOR_AXIS_RANGE_TRAILING);
	ct_long ncontract = 1;
	for (int i = 0; i < nm; i++) ncontract *= t.dim_blocks[shift_t + i];
	GemmBatch b;
	ct_long ir[16], is[16], it[16], ic[16], sd[16], td[16];
	for (ct_long k = 0; k < r.nblocks; k++) {
		if (!r.present(k)) continue;
		if (r.ndim > 0) r.block_index(k, ir);
		ct_long M = 1, N = 1;
		if (r.ndim > 0) {
			ct_long rd[16];
			r.block_dims(k, rd);
			for (int i = 0; i < nds; i++) M *= rd[i];
			for (int i = 0; i < ndt; i++) N *= rd[nds + i];
		}
		b.add_task(off_ptr_mut(r.offset[k]), (int)M, (int)N, (int)N, (int)N, 0, accumulate);
		for (int i = 0; i < nm; i++) ic[i] = 0;
		// the contracted block indices in row-major order; the last one is fixed by the vanishing
		// quantum number sum of s, so only the leading nm - 1 are enumerated and the last is found
		// by binary search (same pairs, same order as scanning all combinations)
		const ct_long nlast = (nm > 0) ? t.dim_blocks[shift_t + nm - 1] : 1;
		const ct_long nprefix = (nm > 0) ? ncontract / std::max<ct_long>(nlast, 1) : 1;
		for (int i = 0; i < nds; i++) is[offset_s + i] = ir[i];
		for (ct_long pp = 0; pp < nprefix && (nm == 0 || nlast > 0); pp++) {
			for (int i = 0; i < nm; i++) is[shift_s + i] = ic[i];
			bool found = true;
			if (nm > 0) {
				const int al = shift_s + nm - 1;
				qnumber rest = 0;
				for (int i = 0; i < s.ndim; i++) if (i != al) rest += s.axis_dir[i] * s.qnums_blocks[i][is[i]];
				const ct_long j = s.find_qnum(al, -s.axis_dir[al] * rest);
				if (j < 0) found = false;
				else { ic[nm - 1] = j; is[al] = j; }
			}
			if (found) {
				for (int i = 0; i < ndt; i++) it[offset_t + i] = ir[nds + i];
				for (int i = 0; i < nm; i++) it[shift_t + i] = ic[i];
				const ct_long sk = s.flat_index(is), tk = t.flat_index(it);
				if (s.present(sk) && t.present(tk)) {
					s.block_dims(sk, sd);
					t.block_dims(tk, td);
					ct_long lds = 1, tds = 1, ldt = 1, tdt = 1;
					for (int i = 0; i < nlds; i++) lds *= sd[i];
					for (int i = nlds; i < s.ndim; i++) tds *= sd[i];
					for (int i = 0; i < nldt; i++) ldt *= td[i];
					for (int i = nldt; i < t.ndim; i++) tdt *= td[i];
					const ct_long K = transa ? lds : tds;
					const int sa_i = transa ? 1 : (int)tds, sa_k = transa ? (int)tds : 1;
					const int sb_k = transb ? 1 : (int)tdt, sb_j = transb ? (int)tdt : 1;
					b.add_term(off_ptr(s.offset[sk]), off_ptr(t.offset[tk]), (int)K, sa_i, sa_k, sb_k, sb_j, 0);
				}
			}
			if (nm > 0) ic[nm - 1] = 0;
			for (int i = nm - 2; i >= 0; i--) { if (++ic[i] < t.dim_blocks[shift_t + i]) break; ic[i] = 0; }
		}
	}
	return b;
}

DevTensor dev_dot(const DevTensor& s, int axs, const DevTensor& t, int axt, int nm, DotForm form, hipStream_t st, double* flops)
{
	DevTensor r(dot_result_meta(s.meta, axs, t.meta, axt, nm));
	const GemmBatch b = plan_dot(s.meta, axs, t.meta, axt, nm, r.meta, form == DotForm::Plain);
	if (flops) *flops = b.flops;
	const GemmBases bases{ s.d, t.d, r.d };
	if (form == DotForm::Plain) {
		r.zero(st);
		run_batch_plain(b, bases, st);
	}
	else {
		// descriptor buffers return to this thread's cache; later uses are ordered behind the launch on st
		DevGemm g;
		g.upload(b, st);
		launch_grouped_gemm_f64(g, bases, st);
		g.release();
	}
	return r;
}

//________________________________________________________________________________________________________________________
//
// transpose (bst.c:785-870) and partial trace (bst.c:1560-1612)

static Meta transpose_meta(const int* perm, const Meta& t)
{
	std::vector<ct_long> dl(t.ndim);
	std::vector<int> ad(t.ndim);
	std::vector<const qnumber*> qn(t.ndim);
	for (int i = 0; i < t.ndim; i++) { dl[i] = t.dim_logical[perm[i]]; ad[i] = t.axis_dir[perm[i]]; qn[i] = t.qnums_logical[perm[i]].data(); }
	return make_meta(t.ndim, dl.data(), ad.data(), qn.data());
}

static std::vector<PermBlock> transpose_blocks(const int* perm, const Meta& t, const Meta& r)
{
	std::vector<PermBlock> pbs;
	ct_long it[16], ir[16], td[16];
	for (ct_long k = 0; k < t.nblocks; k++) {
		if (!t.present(k)) continue;
		t.block_index(k, it);
		for (int i = 0; i < t.ndim; i++) ir[i] = it[perm[i]];
		const ct_long kr = r.flat_index(ir);
		t.block_dims(k, td);
		PermBlock pb;
		memset(&pb, 0, sizeof(pb));
		pb.src = t.offset[k];
		pb.dst = r.offset[kr];
		pb.ndim = t.ndim;
		int64_t st[16], f = 1;
		for (int i = t.ndim - 1; i >= 0; i--) { st[i] = f; f *= td[i]; }
		pb.n = f;
		for (int i = 0; i < t.ndim; i++) { pb.ddim[i] = td[perm[i]]; pb.sstride[i] = st[perm[i]]; }
		pbs.push_back(pb);
	}
	return pbs;
}

DevTensor dev_transpose(const int* perm, const DevTensor& t, hipStream_t st)
{
	if (t.meta.ndim > 6) throw HipError{ hipErrorInvalidValue, "transpose supports at most 6 axes" };
	DevTensor r(transpose_meta(perm, t.meta));
	const std::vector<PermBlock> pbs = transpose_blocks(perm, t.meta, r.meta);
	PermBlock* d = (PermBlock*)dev_alloc(std::max<size_t>(1, pbs.size()) * sizeof(PermBlock));
	copy_h2d_async(d, pbs.data(), pbs.size() * sizeof(PermBlock), st);
	launch_permute_blocks(t.d, r.d, d, (int)pbs.size(), st);
	dev_free(d);
	return r;
}

struct TraceItem { int64_t r_off; int64_t nr; int64_t t_off; int64_t stride; };

__global__ __launch_bounds__(256) void k_partial_trace(const double* __restrict__ t, double* __restrict__ r, const TraceItem* __restrict__ items,
	const int* __restrict__ group_start, int ngroups)
{
	for (int g = blockIdx.x; g < ngroups; g += gridDim.x) {
		const TraceItem first = items[group_start[g]];
		for (int64_t j = threadIdx.x; j < first.nr; j += blockDim.x) {
			double acc = 0;
			for (int q = group_start[g]; q < group_start[g + 1]; q++) {
				const TraceItem it = items[q];
				for (int64_t kk = 0; kk < it.stride; kk++) acc += t[it.t_off + (kk * it.nr + j) * it.stride + kk];
			}
			r[first.r_off + j] = acc;
		}
	}
}

static void plan_trace(const Meta& t, int nt, const Meta& r, std::vector<TraceItem>& items, std::vector<int>& gstart)
{
	const int ndr = t.ndim - 2 * nt;
	ct_long np = 1;
	for (int i = 0; i < nt; i++) np *= t.dim_blocks[i];
	ct_long it[16], td[16];
	for (ct_long k = 0; k < r.nblocks; k++) {
		if (!r.present(k)) continue;
		gstart.push_back((int)items.size());
		for (int i = 0; i < 16; i++) it[i] = 0;
		if (ndr > 0) r.block_index(k, it + nt);
		for (ct_long j = 0; j < np; j++) {
			for (int i = 0; i < nt; i++) it[nt + ndr + i] = it[i];
			const ct_long tk = t.flat_index(it);
			if (t.present(tk)) {
				t.block_dims(tk, td);
				ct_long stride = 1;
				for (int i = 0; i < nt; i++) stride *= td[i];
				items.push_back(TraceItem{ r.offset[k], r.block_size(k), t.offset[tk], stride });
			}
			for (int i = nt - 1; i >= 0; i--) { if (++it[i] < t.dim_blocks[i]) break; it[i] = 0; }
		}
		if ((int)items.size() == gstart.back()) items.push_back(TraceItem{ r.offset[k], r.block_size(k), 0, 0 });
	}
	gstart.push_back((int)items.size());
}

DevTensor dev_cyclic_partial_trace(const DevTensor& t, int nt, hipStream_t st)
{
	const Meta& m = t.meta;
	const int ndr = m.ndim - 2 * nt;
	std::vector<const qnumber*> qn(ndr);
	for (int i = 0; i < ndr; i++) qn[i] = m.qnums_logical[nt + i].data();
	DevTensor r(make_meta(ndr, m.dim_logical.data() + nt, m.axis_dir.data() + nt, qn.data()));
	std::vector<TraceItem> items;
	std::vector<int> gs;
	plan_trace(m, nt, r.meta, items, gs);
	const int ng = (int)gs.size() - 1;
	TraceItem* di = (TraceItem*)dev_alloc(std::max<size_t>(1, items.size()) * sizeof(TraceItem));
	int* dg = (int*)dev_alloc(gs.size() * sizeof(int));
	copy_h2d_async(di, items.data(), items.size() * sizeof(TraceItem), st);
	copy_h2d_async(dg, gs.data(), gs.size() * sizeof(int), st);
	if (ng > 0) {
		hipLaunchKernelGGL(k_partial_trace, dim3(std::min(ng, 8192)), dim3(256), 0, st, t.d, r.d, di, dg, ng);
		CTG_HIP(hipGetLastError());
	}
	dev_free(di);
	dev_free(dg);
	return r;
}

//________________________________________________________________________________________________________________________
//
// reshaping and slicing as batched segment copies (bst.c:950-1290, :1446-1552)

static void run_segments(const double* src, double* dst, const std::vector<CopySeg>& segs, hipStream_t st)
{
	if (segs.empty()) return;
	CopySeg* d = (CopySeg*)dev_alloc(segs.size() * sizeof(CopySeg));
	copy_h2d_async(d, segs.data(), segs.size() * sizeof(CopySeg), st);
	launch_copy_segments(src, dst, d, (int)segs.size(), st);
	dev_free(d);
}

// logical index -> entry index within its block, for one axis (bst.c:882-903)
static std::vector<ct_long> fanin_map(const Meta& m, int ax)
{
	std::vector<ct_long> out((size_t)m.dim_logical[ax]);
	std::vector<ct_long> counter((size_t)m.dim_blocks[ax], 0);
	for (ct_long i = 0; i < m.dim_logical[ax]; i++) out[i] = counter[(size_t)m.find_qnum(ax, m.qnums_logical[ax][i])]++;
	return out;
}

DevTensor dev_flatten_axes(const DevTensor& t, int i_ax, int new_dir, hipStream_t st)
{
	const Meta& m = t.meta;
	const int nd = m.ndim;
	std::vector<ct_long> dl;
	std::vector<int> ad;
	std::vector<std::vector<qnumber>> qs;
	for (int i = 0; i < i_ax; i++) { dl.push_back(m.dim_logical[i]); ad.push_back(m.axis_dir[i]); qs.push_back(m.qnums_logical[i]); }
	const ct_long d0 = m.dim_logical[i_ax], d1 = m.dim_logical[i_ax + 1];
	std::vector<qnumber> qf((size_t)(d0 * d1));
	for (ct_long j = 0; j < d0; j++) for (ct_long k = 0; k < d1; k++)
		qf[j * d1 + k] = new_dir * (m.axis_dir[i_ax] * m.qnums_logical[i_ax][j] + m.axis_dir[i_ax + 1] * m.qnums_logical[i_ax + 1][k]);
	dl.push_back(d0 * d1); ad.push_back(new_dir); qs.push_back(qf);
	for (int i = i_ax + 2; i < nd; i++) { dl.push_back(m.dim_logical[i]); ad.push_back(m.axis_dir[i]); qs.push_back(m.qnums_logical[i]); }
	std::vector<const qnumber*> qp;
	for (auto& q : qs) qp.push_back(q.data());
	DevTensor r(make_meta(nd - 1, dl.data(), ad.data(), qp.data()));
	const Meta& rm = r.meta;
	const std::vector<ct_long> fin = fanin_map(rm, i_ax);
	std::vector<CopySeg> segs;
	ct_long it[16], ir[16], bt[16], br[16], is[16], isr[16];
	for (ct_long k = 0; k < m.nblocks; k++) {
		if (!m.present(k)) continue;
		m.block_index(k, it);
		m.block_dims(k, bt);
		const qnumber q0 = m.qnums_blocks[i_ax][it[i_ax]], q1 = m.qnums_blocks[i_ax + 1][it[i_ax + 1]];
		const qnumber qfl = new_dir * (m.axis_dir[i_ax] * q0 + m.axis_dir[i_ax + 1] * q1);
		for (int i = 0; i < i_ax; i++) ir[i] = it[i];
		ir[i_ax] = rm.find_qnum(i_ax, qfl);
		for (int i = i_ax + 2; i < nd; i++) ir[i - 1] = it[i];
		const ct_long kr = rm.flat_index(ir);
		rm.block_dims(kr, br);
		const std::vector<ct_long> f0 = fanout(m, i_ax, q0), f1 = fanout(m, i_ax + 1, q1);
		ct_long stride = 1;
		for (int i = i_ax + 2; i < nd; i++) stride *= bt[i];
		// one segment per index of the axes before i_ax + 1: for fixed (i_0 .. i_{i_ax}) the entries
		// of consecutive i_{i_ax+1} of this block are consecutive in the fused sector as well (fused
		// logical indices with the same row and the same quantum number of axis i_ax + 1 are
		// exactly those of this block), so the whole run is one contiguous copy
		const ct_long run = bt[i_ax + 1] * stride;
		ct_long nsl = 1;
		for (int i = 0; i < i_ax + 1; i++) { nsl *= bt[i]; is[i] = 0; }
		for (ct_long j = 0; j < nsl; j++) {
			for (int i = 0; i < i_ax; i++) isr[i] = is[i];
			isr[i_ax] = fin[f0[is[i_ax]] * d1 + f1[0]];
			ct_long l = 0;
			for (int i = 0; i <= i_ax; i++) l = l * br[i] + isr[i];
			segs.push_back(CopySeg{ m.offset[k] + j * run, rm.offset[kr] + l * stride, run });
			for (int i = i_ax; i >= 0; i--) { if (++is[i] < bt[i]) break; is[i] = 0; }
		}
	}
	// the segments are disjoint; when they cover every entry no zero fill is needed
	ct_long covered = 0;
	for (const CopySeg& g : segs) covered += g.len;
	if (covered != rm.nelem) r.zero(st);
	run_segments(t.d, r.d, segs, st);
	return r;
}

DevTensor dev_split_axis(const DevTensor& t, int i_ax, const ct_long* nd2, const int* dir2, const qnumber* const* q2, hipStream_t st)
{
	const Meta& m = t.meta;
	const int nd = m.ndim;
	std::vector<ct_long> dl;
	std::vector<int> ad;
	std::vector<const qnumber*> qp;
	for (int i = 0; i < i_ax; i++) { dl.push_back(m.dim_logical[i]); ad.push_back(m.axis_dir[i]); qp.push_back(m.qnums_logical[i].data()); }
	for (int c = 0; c < 2; c++) { dl.push_back(nd2[c]); ad.push_back(dir2[c]); qp.push_back(q2[c]); }
	for (int i = i_ax + 1; i < nd; i++) { dl.push_back(m.dim_logical[i]); ad.push_back(m.axis_dir[i]); qp.push_back(m.qnums_logical[i].data()); }
	DevTensor r(make_meta(nd + 1, dl.data(), ad.data(), qp.data()));
	const Meta& rm = r.meta;
	const std::vector<ct_long> fin = fanin_map(m, i_ax);
	std::vector<CopySeg> segs;
	ct_long ir[16], it[16], br[16], bt[16], is[16], ist[16];
	for (ct_long k = 0; k < rm.nblocks; k++) {
		if (!rm.present(k)) continue;
		rm.block_index(k, ir);
		rm.block_dims(k, br);
		const qnumber q0 = rm.qnums_blocks[i_ax][ir[i_ax]], q1 = rm.qnums_blocks[i_ax + 1][ir[i_ax + 1]];
		const qnumber qfl = m.axis_dir[i_ax] * (rm.axis_dir[i_ax] * q0 + rm.axis_dir[i_ax + 1] * q1);
		for (int i = 0; i < i_ax; i++) it[i] = ir[i];
		it[i_ax] = m.find_qnum(i_ax, qfl);
		for (int i = i_ax + 1; i < nd; i++) it[i] = ir[i + 1];
		if (it[i_ax] < 0) continue;
		const ct_long kt = m.flat_index(it);
		if (!m.present(kt)) continue;
		m.block_dims(kt, bt);
		const std::vector<ct_long> f0 = fanout(rm, i_ax, q0), f1 = fanout(rm, i_ax + 1, q1);
		ct_long stride = 1;
		for (int i = i_ax + 2; i < nd + 1; i++) stride *= br[i];
		// one segment per index of the axes before i_ax + 1 (the run over axis i_ax + 1 is
		// contiguous in the fused sector, as in dev_flatten_axes)
		const ct_long run = br[i_ax + 1] * stride;
		ct_long nsl = 1;
		for (int i = 0; i < i_ax + 1; i++) { nsl *= br[i]; is[i] = 0; }
		for (ct_long j = 0; j < nsl; j++) {
			for (int i = 0; i < i_ax; i++) ist[i] = is[i];
			ist[i_ax] = fin[f0[is[i_ax]] * nd2[1] + f1[0]];
			ct_long l = 0;
			for (int i = 0; i <= i_ax; i++) l = l * bt[i] + ist[i];
			segs.push_back(CopySeg{ m.offset[kt] + l * stride, rm.offset[k] + j * run, run });
			for (int i = i_ax; i >= 0; i--) { if (++is[i] < br[i]) break; is[i] = 0; }
		}
	}
	// result blocks without a source block stay zero: fill only when the segments (disjoint) do
	// not cover every entry
	ct_long covered = 0;
	for (const CopySeg& g : segs) covered += g.len;
	if (covered != rm.nelem) r.zero(st);
	run_segments(t.d, r.d, segs, st);
	return r;
}

//________________________________________________________________________________________________________________________
//
// H_eff matvec plan (chain_ops.c:353-390)

MatvecPlan::~MatvecPlan()
{
	for (int i = 0; i < 3; i++) stage[i].release();
	wapply.release();
	fused.release();
	pair.release();
	for (void* p : owned) dev_free(p);
	for (int i = 0; i < 4; i++) if (ev[i]) (void)hipEventDestroy(ev[i]);
}

static bool same_axis(const Meta& a, int ia, const Meta& b, int ib)
{
	return a.dim_logical[ia] == b.dim_logical[ib] && a.qnums_logical[ia] == b.qnums_logical[ib];
}

// the reference's GEMM count of one matvec (every allocated block pair, zero blocks included)
double reference_flops(const Meta& th, const Meta& w, const Meta& l, const Meta& r)
{
	double f = 0;
	// s1 = theta . R : [a p n c o]
	Meta s1 = dot_result_meta(th, TENSOR_AXIS_RANGE_TRAILING, r, TENSOR_AXIS_RANGE_LEADING, 1);
	f += plan_dot(th, TENSOR_AXIS_RANGE_TRAILING, r, TENSOR_AXIS_RANGE_LEADING, 1, s1, false).flops;
	const int p0[5] = { 1, 2, 0, 3, 4 };
	Meta t = transpose_meta(p0, s1);
	Meta s2 = dot_result_meta(w, TENSOR_AXIS_RANGE_TRAILING, t, TENSOR_AXIS_RANGE_LEADING, 2);
	f += plan_dot(w, TENSOR_AXIS_RANGE_TRAILING, t, TENSOR_AXIS_RANGE_LEADING, 2, s2, false).flops;
	const int p1[5] = { 2, 0, 1, 3, 4 };
	Meta t2 = transpose_meta(p1, s2);
	const int p2[4] = { 0, 3, 1, 2 };
	Meta k = transpose_meta(p2, l);
	Meta s3 = dot_result_meta(k, TENSOR_AXIS_RANGE_TRAILING, t2, TENSOR_AXIS_RANGE_LEADING, 2);
	f += plan_dot(k, TENSOR_AXIS_RANGE_TRAILING, t2, TENSOR_AXIS_RANGE_LEADING, 2, s3, false).flops;
	return f;
}

// Append the three grouped stages of one problem to g. Element offsets: theta / y at x_off, W at
// w_off, L at l_off, R at r_off, its T1 / T2 blocks from the current ends of g's intermediates.
void plan_grouped_tables(GroupedTables& g, const Meta& th, const Meta& W, const Meta& L, const Meta& R,
	int64_t x_off, int64_t w_off, int64_t l_off, int64_t r_off, bool lean)
{
	GemmBatch& s1 = g.s1;
	GemmBatch& s2 = g.s2;
	GemmBatch& s3 = g.s3;
	const int64_t t1_base = g.t1_elems, t2_base = g.t2_elems;
	const qnumber qo = L.qnums_blocks[0][0];
	const ct_long nA = th.dim_blocks[0], nP = th.dim_blocks[1], nB = th.dim_blocks[2];
	const ct_long nN = R.dim_blocks[1], nC = R.dim_blocks[2];
	const ct_long nM = W.dim_blocks[0], nQ = W.dim_blocks[1];
	auto bc = [](const Meta& m, int ax, ct_long i) { return m.block_count[ax][i]; };
	auto qb = [](const Meta& m, int ax, ct_long i) { return m.qnums_blocks[ax][i]; };

	// stage 1: T1(A,P,N,C)[a p n c] = theta(A,P,B) R(B,N,C,O)
	std::vector<ct_long> t1_off((size_t)(nA * nP * nN * nC), -1);
	auto t1_key = [&](ct_long A, ct_long P, ct_long N, ct_long C) { return ((A * nP + P) * nN + N) * nC + C; };
	ct_long t1_elems = 0;
	{
		// the (N, C) of the present R blocks of every B, in (N, C) order
		std::vector<std::vector<std::pair<ct_long, ct_long>>> r_nc((size_t)nB);
		for (ct_long B = 0; B < nB; B++)
			for (ct_long N = 0; N < nN; N++) for (ct_long C = 0; C < nC; C++) {
				const ct_long irr[4] = { B, N, C, 0 };
				if (R.present(R.flat_index(irr))) r_nc[(size_t)B].emplace_back(N, C);
			}
		std::vector<std::tuple<ct_long, ct_long, ct_long, ct_long, ct_long, ct_long>> blocks;  // A P N C Bidx off
		for (ct_long A = 0; A < nA; A++) for (ct_long P = 0; P < nP; P++) {
			const ct_long B = th.find_qnum(2, qb(th, 0, A) + qb(th, 1, P));
			if (B < 0) continue;
			const ct_long ib[3] = { A, P, B };
			if (!th.present(th.flat_index(ib))) continue;
			for (const auto& nc : r_nc[(size_t)B]) {
				const ct_long N = nc.first, C = nc.second;
				t1_off[t1_key(A, P, N, C)] = t1_elems;
				t1_elems += bc(th, 0, A) * bc(th, 1, P) * bc(R, 1, N) * bc(R, 2, C);
				blocks.emplace_back(A, P, N, C, B, t1_off[t1_key(A, P, N, C)]);
			}
		}
		for (auto& bl : blocks) {
			ct_long A, P, N, C, B, off;
			std::tie(A, P, N, C, B, off) = bl;
			const ct_long dA = bc(th, 0, A), dP = bc(th, 1, P), dB = bc(th, 2, B), dN = bc(R, 1, N), dC = bc(R, 2, C);
			const ct_long ith[3] = { A, P, B };
			const ct_long irr[4] = { B, N, C, 0 };
			s1.add_task(off_ptr_mut(t1_base + off), (int)(dA * dP), (int)(dN * dC), (int)(dN * dC), (int)(dN * dC), 0, false);
			s1.add_term(off_ptr(x_off + th.offset[th.flat_index(ith)]), off_ptr(r_off + R.offset[R.flat_index(irr)]), (int)dB, (int)dB, 1, (int)(dN * dC), 1, 0);
		}
	}
	// stage 2: T2(A,M,Q,C)[a m q c] = sum_P W(M,Q,P,N) T1(A,P,N,C)
	std::vector<ct_long> t2_off((size_t)(nA * nM * nQ * nC), -1);
	auto t2_key = [&](ct_long A, ct_long M, ct_long Q, ct_long C) { return ((A * nM + M) * nQ + Q) * nC + C; };
	ct_long t2_elems = 0;
	// (M,Q) pairs by their quantum number sum
	std::map<qnumber, std::vector<std::pair<ct_long, ct_long>>> mq_by_sum;
	for (ct_long M = 0; M < nM; M++) for (ct_long Q = 0; Q < nQ; Q++) mq_by_sum[qb(W, 0, M) + qb(W, 1, Q)].emplace_back(M, Q);
	// the (P, N) of the present W blocks of every (M, Q), in P order
	std::vector<std::vector<std::pair<ct_long, ct_long>>> w_pn((size_t)(nM * nQ));
	for (ct_long M = 0; M < nM; M++) for (ct_long Q = 0; Q < nQ; Q++)
		for (ct_long P = 0; P < nP; P++) {
			const ct_long N = W.find_qnum(3, qb(W, 0, M) + qb(W, 1, Q) - qb(W, 2, P));
			if (N < 0) continue;
			const ct_long iw[4] = { M, Q, P, N };
			if (W.present(W.flat_index(iw))) w_pn[(size_t)(M * nQ + Q)].emplace_back(P, N);
		}
	std::vector<std::pair<ct_long, ct_long>> tl;  // (P, N)
	for (ct_long A = 0; A < nA; A++) for (ct_long C = 0; C < nC; C++) {
		auto grp = mq_by_sum.find(qb(R, 2, C) + qo - qb(th, 0, A));
		if (grp == mq_by_sum.end()) continue;
		for (const auto& mq : grp->second) {
			const ct_long M = mq.first, Q = mq.second;
			tl.clear();
			for (const auto& pn : w_pn[(size_t)(M * nQ + Q)])
				if (t1_off[t1_key(A, pn.first, pn.second, C)] >= 0) tl.push_back(pn);
			if (tl.empty()) continue;
			const ct_long dA = bc(th, 0, A), dM = bc(W, 0, M), dQ = bc(W, 1, Q), dC = bc(R, 2, C);
			const ct_long off = t2_elems;
			t2_off[t2_key(A, M, Q, C)] = off;
			t2_elems += dA * dM * dQ * dC;
			if (lean) {
				// T2 layout and the flop count only: stage 2 runs as the column stream
				for (auto& pn : tl) s2.flops += 2.0 * (double)(dM * dQ) * (double)(dA * dC) * (double)(bc(W, 2, pn.first) * bc(W, 3, pn.second));
				continue;
			}
			s2.add_task(off_ptr_mut(t2_base + off), (int)(dM * dQ), (int)(dA * dC), (int)dC, (int)dC, dM * dQ * dC, false);
			for (auto& pn : tl) {
				const ct_long P = pn.first, N = pn.second;
				const ct_long dP = bc(W, 2, P), dN = bc(W, 3, N);
				const ct_long iw[4] = { M, Q, P, N };
				s2.add_term(off_ptr(w_off + W.offset[W.flat_index(iw)]), off_ptr(t1_base + t1_off[t1_key(A, P, N, C)]), (int)(dP * dN), (int)(dP * dN), 1, (int)dC, 1, dP * dN * dC);
			}
		}
	}
	// column-stream form of stage 2: classes D = qM + qQ = qP + qN, one unit per block pair (A,C)
	{
		WApplyHost& wa = g.wa;
		struct Cls { int64_t wrow = 0, wcol = 0; int K = 0, R = 0; std::vector<std::pair<ct_long, ct_long>> rows, cols; };
		std::map<qnumber, Cls> classes;
		auto get_class = [&](qnumber D) -> const Cls& {
			auto it = classes.find(D);
			if (it != classes.end()) return it->second;
			Cls c;
			for (ct_long M = 0; M < nM; M++) for (ct_long Q = 0; Q < nQ; Q++)
				if (qb(W, 0, M) + qb(W, 1, Q) == D) { c.rows.emplace_back(M, Q); c.R += (int)(bc(W, 0, M) * bc(W, 1, Q)); }
			for (ct_long P = 0; P < nP; P++) for (ct_long N = 0; N < nN; N++)
				if (qb(W, 2, P) + qb(W, 3, N) == D) { c.cols.emplace_back(P, N); c.K += (int)(bc(W, 2, P) * bc(W, 3, N)); }
			c.wrow = (int64_t)(wa.widx.size() / WA_KMAX);
			c.wcol = (int64_t)(wa.widxT.size() / F12_RMAX);
			wa.kmax = std::max(wa.kmax, c.K);
			if (c.K <= WA_KMAX) {
				for (auto& mq : c.rows) {
					const ct_long M = mq.first, Q = mq.second, dM = bc(W, 0, M), dQ = bc(W, 1, Q);
					for (ct_long m = 0; m < dM; m++) for (ct_long q = 0; q < dQ; q++) {
						std::vector<int64_t> row(WA_KMAX, -1);
						int k = 0;
						for (auto& pn : c.cols) {
							const ct_long P = pn.first, N = pn.second, dP = bc(W, 2, P), dN = bc(W, 3, N);
							const ct_long iw[4] = { M, Q, P, N };
							const ct_long kw = W.flat_index(iw);
							const bool pres = W.present(kw);
							for (ct_long pp = 0; pp < dP; pp++) for (ct_long n = 0; n < dN; n++, k++)
								if (pres) row[k] = w_off + W.offset[kw] + ((m * dQ + q) * dP + pp) * dN + n;
						}
						wa.widx.insert(wa.widx.end(), row.begin(), row.end());
					}
				}
				// transposed copy for the fused stages: column k holds rows 0 .. min(R, F12_RMAX)
				std::vector<int64_t> colT((size_t)WA_KMAX * F12_RMAX, -1);
				for (int r = 0; r < std::min(c.R, F12_RMAX); r++)
					for (int k = 0; k < WA_KMAX; k++) colT[(size_t)k * F12_RMAX + r] = wa.widx[(size_t)(c.wrow + r) * WA_KMAX + k];
				wa.widxT.insert(wa.widxT.end(), colT.begin(), colT.end());
			}
			return classes.emplace(D, std::move(c)).first->second;
		};
		for (ct_long A = 0; A < nA; A++) for (ct_long C = 0; C < nC; C++) {
			const qnumber D = qb(R, 2, C) + qo - qb(th, 0, A);
			const Cls& c = get_class(D);
			bool any = false;
			for (auto& mq : c.rows) if (t2_off[t2_key(A, mq.first, mq.second, C)] >= 0) { any = true; break; }
			if (!any) continue;
			const ct_long dA = bc(th, 0, A), dC = bc(R, 2, C);
			WaUnit u;
			u.kb = (int64_t)wa.kbase.size();
			u.rb = (int64_t)wa.rbase.size();
			u.wrow = c.wrow;
			u.wcol = c.wcol;
			u.dA = (int)dA;
			u.dC = (int)dC;
			u.K = c.K;
			u.R = c.R;
			for (auto& pn : c.cols) {
				const ct_long P = pn.first, N = pn.second, dP = bc(W, 2, P), dN = bc(W, 3, N);
				const ct_long o = t1_off[t1_key(A, P, N, C)];
				for (ct_long pp = 0; pp < dP; pp++) for (ct_long n = 0; n < dN; n++) {
					wa.kbase.push_back(o >= 0 ? t1_base + o + (pp * dN + n) * dC : -1);
					wa.kstride.push_back((int)(dP * dN * dC));
				}
			}
			for (auto& mq : c.rows) {
				const ct_long M = mq.first, Q = mq.second, dM = bc(W, 0, M), dQ = bc(W, 1, Q);
				const ct_long o = t2_off[t2_key(A, M, Q, C)];
				for (ct_long m = 0; m < dM; m++) for (ct_long q = 0; q < dQ; q++) {
					wa.rbase.push_back(o >= 0 ? t2_base + o + (m * dQ + q) * dC : -1);
					wa.rstride.push_back((int)(dM * dQ * dC));
				}
			}
			const int unit = (int)wa.units.size();
			wa.units.push_back(u);
			for (int col0 = 0; col0 < (int)(dA * dC); col0 += 64) wa.items.push_back(WaItem{ u, unit, col0 });
			// fused stages 1 + 2: one group per column group (P,N) of the class
			if (lean) { wa.f12_ok = false; continue; }
			if (c.R > F12_RMAX) wa.f12_ok = false;
			const int g0 = (int)wa.groups.size();
			for (auto& pn : c.cols) {
				const ct_long P = pn.first, N = pn.second, dP = bc(W, 2, P), dN = bc(W, 3, N);
				F12Group gr = { -1, -1, (int)dP, (int)dN, 0, 0 };
				const ct_long B = th.find_qnum(2, qb(th, 0, A) + qb(th, 1, P));
				if (B >= 0 && t1_off[t1_key(A, P, N, C)] >= 0) {
					const ct_long ith[3] = { A, P, B };
					const ct_long irr[4] = { B, N, C, 0 };
					gr.th_off = x_off + th.offset[th.flat_index(ith)];
					gr.r_off = r_off + R.offset[R.flat_index(irr)];
					gr.dB = (int)bc(th, 2, B);
				}
				wa.groups.push_back(gr);
			}
			const int ng = (int)wa.groups.size() - g0;
			auto lds_need = [&](int ta, int tc) {
				int n = 0;
				for (int gi = g0; gi < g0 + ng; gi++) {
					const F12Group& gr = wa.groups[gi];
					if (gr.th_off >= 0) n += ta * gr.dP * gr.dB + gr.dB * gr.dN * tc;
				}
				return n;
			};
			int tc = (int)std::min<ct_long>(dC, 64);
			int ta = (int)std::min<ct_long>(dA, std::max(1, 64 / tc));
			while (lds_need(ta, tc) > F12_LDS) {
				if (ta > 1) ta = std::max(1, ta / 2);
				else if (tc > 1) { tc = std::max(1, tc / 2); ta = (int)std::min<ct_long>(dA, std::max(1, 64 / tc)); if (lds_need(ta, tc) > F12_LDS) ta = 1; }
				else break;
			}
			if (lds_need(ta, tc) > F12_LDS) wa.f12_ok = false;
			for (int a0 = 0; a0 < (int)dA; a0 += ta)
				for (int c0 = 0; c0 < (int)dC; c0 += tc) {
					const int ta_i = std::min(ta, (int)dA - a0), tc_i = std::min(tc, (int)dC - c0);
					wa.items12.push_back(F12Item{ u, g0, ng, a0, ta_i, c0, tc_i, lds_need(ta_i, tc_i), 0 });
				}
		}
	}
	// stage 3: y(X,Q,C)[x q c] = sum_(A,M) L(O,A,M,X)^T T2(A,M,Q,C)
	// the present T2 blocks (A, M) of every (Q, C), in (A, M) order; for a given X at most one M
	// of an A matches (qM = qo - qA + qX)
	std::vector<std::vector<std::pair<ct_long, ct_long>>> t2_am((size_t)(nQ * nC));
	for (ct_long A = 0; A < nA; A++) for (ct_long M = 0; M < nM; M++) for (ct_long Q = 0; Q < nQ; Q++) for (ct_long C = 0; C < nC; C++)
		if (t2_off[t2_key(A, M, Q, C)] >= 0) t2_am[(size_t)(Q * nC + C)].emplace_back(A, M);
	ct_long iy[3];
	for (ct_long k = 0; k < th.nblocks; k++) {
		if (!th.present(k)) continue;
		th.block_index(k, iy);
		const ct_long X = iy[0], Q = iy[1], C = iy[2];
		const ct_long dX = bc(th, 0, X), dQ = bc(th, 1, Q), dC = bc(th, 2, C);
		s3.add_task(off_ptr_mut(x_off + th.offset[k]), (int)dX, (int)(dQ * dC), (int)(dQ * dC), (int)(dQ * dC), 0, false);
		for (const auto& am : t2_am[(size_t)(Q * nC + C)]) {
			const ct_long A = am.first, M = am.second;
			if (qb(L, 2, M) != qo - qb(L, 1, A) + qb(L, 3, X)) continue;
			const ct_long il[4] = { 0, A, M, X };
			if (!L.present(L.flat_index(il))) continue;
			const ct_long o2 = t2_off[t2_key(A, M, Q, C)];
			const ct_long dA = bc(L, 1, A), dM = bc(L, 2, M);
			s3.add_term(off_ptr(l_off + L.offset[L.flat_index(il)]), off_ptr(t2_base + o2), (int)(dA * dM), 1, (int)dX, (int)(dQ * dC), 1, 0);
		}
	}
	g.t1_elems += t1_elems;
	g.t2_elems += t2_elems;
	g.stage_bytes[0] += 8.0 * (double)(th.nelem + R.nelem + t1_elems);
	g.stage_bytes[1] += 8.0 * (double)(t1_elems + W.nelem + t2_elems);
	g.stage_bytes[2] += 8.0 * (double)(t2_elems + L.nelem + th.nelem);
}

// Append src (planned with zero intermediate bases) to dst: its T1 / T2 offsets move behind dst's
static const double* shift_ptr(const double* p, int64_t elems) { return (const double*)((intptr_t)p + (intptr_t)(elems * (int64_t)sizeof(double))); }

void merge_grouped_tables(GroupedTables& dst, const GroupedTables& src, int64_t x_off, int64_t w_off, int64_t l_off, int64_t r_off)
{
	const int64_t t1s = dst.t1_elems, t2s = dst.t2_elems;
	auto append = [](GemmBatch& d, const GemmBatch& b, int64_t c_shift, int64_t a_shift, int64_t b_shift) {
		const int32_t q0 = (int32_t)d.terms.size();
		d.tasks.reserve(d.tasks.size() + b.tasks.size());
		d.terms.reserve(d.terms.size() + b.terms.size());
		for (GemmTask t : b.tasks) { t.C = (double*)shift_ptr(t.C, c_shift); t.t0 += q0; t.t1 += q0; d.tasks.push_back(t); }
		for (GemmTerm m : b.terms) { m.A = shift_ptr(m.A, a_shift); m.B = shift_ptr(m.B, b_shift); d.terms.push_back(m); }
		d.flops += b.flops;
	};
	// stage 1: T1 = theta R; stage 2: T2 = W T1; stage 3: y = L^T T2 (y at theta's offset)
	append(dst.s1, src.s1, t1s, x_off, r_off);
	append(dst.s2, src.s2, t2s, w_off, t1s);
	append(dst.s3, src.s3, x_off, l_off, t2s);
	WApplyHost& a = dst.wa;
	const WApplyHost& b = src.wa;
	const int64_t kb0 = (int64_t)a.kbase.size(), rb0 = (int64_t)a.rbase.size(), wr0 = (int64_t)(a.widx.size() / WA_KMAX);
	const int64_t wc0 = (int64_t)(a.widxT.size() / F12_RMAX);
	const int u0 = (int)a.units.size();
	for (WaUnit u : b.units) { u.kb += kb0; u.rb += rb0; u.wrow += wr0; u.wcol += wc0; a.units.push_back(u); }
	for (WaItem it : b.items) { it.unit += u0; it.u.kb += kb0; it.u.rb += rb0; it.u.wrow += wr0; it.u.wcol += wc0; a.items.push_back(it); }
	const int g0 = (int)a.groups.size();
	for (F12Group gr : b.groups) {
		if (gr.th_off >= 0) gr.th_off += x_off;
		if (gr.r_off >= 0) gr.r_off += r_off;
		a.groups.push_back(gr);
	}
	for (F12Item it : b.items12) { it.g0 += g0; it.u.kb += kb0; it.u.rb += rb0; it.u.wrow += wr0; it.u.wcol += wc0; a.items12.push_back(it); }
	a.f12_ok = a.f12_ok && b.f12_ok;
	for (int64_t v : b.kbase) a.kbase.push_back(v >= 0 ? v + t1s : v);
	for (int64_t v : b.rbase) a.rbase.push_back(v >= 0 ? v + t2s : v);
	a.kstride.insert(a.kstride.end(), b.kstride.begin(), b.kstride.end());
	a.rstride.insert(a.rstride.end(), b.rstride.begin(), b.rstride.end());
	for (int64_t v : b.widx) a.widx.push_back(v >= 0 ? v + w_off : v);
	for (int64_t v : b.widxT) a.widxT.push_back(v >= 0 ? v + w_off : v);
	a.kmax = std::max(a.kmax, b.kmax);
	dst.t1_elems += src.t1_elems;
	dst.t2_elems += src.t2_elems;
	for (int i = 0; i < 3; i++) dst.stage_bytes[i] += src.stage_bytes[i];
}

// Host check of the grouped tables of one problem: every operand index inside its buffer and every
// entry of T1, T2 and y written by exactly one output of each form (grouped stages, column-stream
// stage 2, fused stages 1 and 2). Returns 0 or the number of the first failed check.
int check_grouped_tables(const Meta& th, const Meta& W, const Meta& L, const Meta& R)
{
	GroupedTables g;
	plan_grouped_tables(g, th, W, L, R, 0, 0, 0, 0);
	auto elem = [](const void* p) { return (int64_t)((intptr_t)p / (intptr_t)sizeof(double)); };
	auto check_batch = [&](const GemmBatch& b, int64_t na, int64_t nb, int64_t nc, std::vector<int>* writes) -> bool {
		for (const GemmTask& t : b.tasks) {
			const int64_t c0 = elem(t.C);
			for (int i = 0; i < t.M; i++) for (int j = 0; j < t.N; j++) {
				const int jb = j / t.nci, jc = j % t.nci;
				const int64_t ci = c0 + (int64_t)jb * t.sc_batch + jc + (int64_t)i * t.sc_i;
				if (ci < 0 || ci >= nc) return false;
				if (writes) (*writes)[ci]++;
			}
			for (int q = t.t0; q < t.t1; q++) {
				const GemmTerm& m = b.terms[q];
				const int64_t a0 = elem(m.A), b0 = elem(m.B);
				const int jb = (t.N - 1) / t.nci, jc = std::min(t.N - 1, t.nci - 1);
				const int64_t amax = a0 + (int64_t)(t.M - 1) * m.sa_i + (int64_t)(m.K - 1) * m.sa_k;
				const int64_t bmax = b0 + (int64_t)jb * m.sb_batch + (int64_t)jc * m.sb_j + (int64_t)(m.K - 1) * m.sb_k;
				if (a0 < 0 || amax >= na || b0 < 0 || bmax >= nb) return false;
			}
		}
		return true;
	};
	auto once = [](const std::vector<int>& w) { for (int v : w) if (v != 1) return false; return true; };
	std::vector<int> w1((size_t)g.t1_elems, 0), w2((size_t)g.t2_elems, 0), wy((size_t)th.nelem, 0);
	if (!check_batch(g.s1, th.nelem, R.nelem, g.t1_elems, &w1) || !once(w1)) return 1;
	if (!check_batch(g.s2, W.nelem, g.t1_elems, g.t2_elems, &w2) || !once(w2)) return 2;
	if (!check_batch(g.s3, L.nelem, g.t2_elems, th.nelem, &wy) || !once(wy)) return 3;
	// the LDS-tiled split: every output entry of every task in exactly one 64 x 64 tile of a large
	// task or one 32 x 32 item of another task
	for (const GemmBatch* b : { &g.s1, &g.s2, &g.s3 }) {
		const std::vector<WorkItem> it = b->work_items();
		std::vector<WorkItem> lit, rit;
		split_lds_items(*b, it, lit, rit);
		if (lit.empty()) continue;
		std::vector<int64_t> base(b->tasks.size() + 1, 0);
		for (size_t ti = 0; ti < b->tasks.size(); ti++) base[ti + 1] = base[ti] + (int64_t)b->tasks[ti].M * b->tasks[ti].N;
		std::vector<int> cover((size_t)base.back(), 0);
		auto mark = [&](const WorkItem& w, int edge) {
			const GemmTask& t = b->tasks[(size_t)w.task_index];
			for (int i = w.m0; i < std::min(t.M, w.m0 + edge); i++)
				for (int j = w.n0; j < std::min(t.N, w.n0 + edge); j++) cover[(size_t)(base[(size_t)w.task_index] + (int64_t)i * t.N + j)]++;
		};
		for (const WorkItem& w : lit) {
			const GemmTask& t = b->tasks[(size_t)w.task_index];
			if (t.M < LDS_TASK_MIN || t.N < LDS_TASK_MIN || w.first.K != b->terms[t.t0].K) return 8;
			mark(w, LDS_TILE);
		}
		for (const WorkItem& w : rit) mark(w, TILE);
		if (!once(cover)) return 8;
	}
	const WApplyHost& wa = g.wa;
	if (wa.kmax <= WA_KMAX) {
		std::vector<int> ws((size_t)g.t2_elems, 0);
		for (const WaItem& it : wa.items) {
			const WaUnit& u = it.u;
			for (int lane = 0; lane < 64; lane++) {
				const int j = it.col0 + lane;
				if (j >= u.dA * u.dC) continue;
				const int a = j / u.dC, c = j % u.dC;
				for (int k = 0; k < u.K; k++) {
					const int64_t b = wa.kbase[u.kb + k];
					if (b >= 0 && (b + (int64_t)a * wa.kstride[u.kb + k] + c >= g.t1_elems)) return 4;
				}
				for (int r = 0; r < u.R; r++) {
					const int64_t b = wa.rbase[u.rb + r];
					if (b < 0) continue;
					const int64_t i = b + (int64_t)a * wa.rstride[u.rb + r] + c;
					if (i < 0 || i >= g.t2_elems) return 4;
					ws[i]++;
				}
			}
		}
		if (!once(ws)) return 5;
	}
	if (wa.kmax <= WA_KMAX && wa.f12_ok) {
		std::vector<int> wf((size_t)g.t2_elems, 0);
		for (const F12Item& it : wa.items12) {
			const WaUnit& u = it.u;
			if (it.ta * it.tc > 64 || it.lds > F12_LDS || u.R > F12_RMAX) return 6;
			int lds = 0;
			for (int gi = it.g0; gi < it.g0 + it.ng; gi++) {
				const F12Group& gr = wa.groups[gi];
				if (gr.th_off < 0) continue;
				if (gr.th_off + (int64_t)(it.a0 + it.ta) * gr.dP * gr.dB > th.nelem) return 6;
				if (gr.r_off + (int64_t)(gr.dB * gr.dN - 1) * u.dC + it.c0 + it.tc > R.nelem) return 6;
				lds += it.ta * gr.dP * gr.dB + gr.dB * gr.dN * it.tc;
			}
			if (lds != it.lds) return 6;
			for (int ai = 0; ai < it.ta; ai++) for (int ci = 0; ci < it.tc; ci++)
				for (int r = 0; r < u.R; r++) {
					const int64_t b = wa.rbase[u.rb + r];
					if (b < 0) continue;
					const int64_t i = b + (int64_t)(it.a0 + ai) * wa.rstride[u.rb + r] + it.c0 + ci;
					if (i < 0 || i >= g.t2_elems) return 6;
					wf[i]++;
				}
		}
		if (!once(wf)) return 7;
	}
	return 0;
}

static void build_grouped(MatvecPlan& p, bool lean)
{
	GroupedTables g;
	plan_grouped_tables(g, p.theta_meta, p.w.meta, p.l.meta, p.r.meta, 0, 0, 0, 0, lean);
	if (lean && g.wa.kmax > WA_KMAX) {
		g = GroupedTables();
		plan_grouped_tables(g, p.theta_meta, p.w.meta, p.l.meta, p.r.meta, 0, 0, 0, 0, false);
	}
	Meta m1, m2;
	m1.nelem = g.t1_elems;
	m2.nelem = g.t2_elems;
	p.interm.emplace_back(std::move(m1));
	p.interm.emplace_back(std::move(m2));
	p.stage[0].upload(g.s1, p.stream);
	p.stage[1].upload(g.s2, p.stream);
	p.stage[2].upload(g.s3, p.stream);
	for (int i = 0; i < 3; i++) p.stage_bytes[i] = g.stage_bytes[i];
	p.stage_flops[0] = g.s1.flops;
	p.stage_flops[1] = g.s2.flops;
	p.stage_flops[2] = g.s3.flops;
	p.stage_bases[0] = GemmBases{ nullptr, p.r.d, p.interm[0].d };
	p.stage_bases[1] = GemmBases{ p.w.d, p.interm[0].d, p.interm[1].d };
	p.stage_bases[2] = GemmBases{ p.l.d, p.interm[1].d, nullptr };
	// the column-stream stage 2 (FP64 FMAs, no MFMA) serves the FP64 and the split-FP32 forms: in
	// both the MFMA GEMMs are stages 1 and 3
	if (p.mode == CT_MATVEC_GROUPED_FP64 || p.mode == CT_MATVEC_GROUPED_SPLIT) p.wapply.upload(g.wa, p.w.d, p.stream);
	p.nterms = g.s1.terms.size() + g.s2.terms.size() + g.s3.terms.size();
	p.ntasks = g.s1.tasks.size() + g.s2.tasks.size() + g.s3.tasks.size();
	p.nitems = p.stage[0].nitems + p.stage[1].nitems + p.stage[2].nitems;
	p.interm_elems = (double)(g.t1_elems + g.t2_elems);
}

static void build_plain(MatvecPlan& p)
{
	hipStream_t st = p.stream;
	const Meta& th = p.theta_meta;
	// s1 = dot(theta, TRAILING, R, LEADING, 1)
	Meta ms1 = dot_result_meta(th, TENSOR_AXIS_RANGE_TRAILING, p.r.meta, TENSOR_AXIS_RANGE_LEADING, 1);
	GemmBatch b1 = plan_dot(th, TENSOR_AXIS_RANGE_TRAILING, p.r.meta, TENSOR_AXIS_RANGE_LEADING, 1, ms1, true);
	const int p0[5] = { 1, 2, 0, 3, 4 };
	Meta mt = transpose_meta(p0, ms1);
	Meta ms2 = dot_result_meta(p.w.meta, TENSOR_AXIS_RANGE_TRAILING, mt, TENSOR_AXIS_RANGE_LEADING, 2);
	GemmBatch b2 = plan_dot(p.w.meta, TENSOR_AXIS_RANGE_TRAILING, mt, TENSOR_AXIS_RANGE_LEADING, 2, ms2, true);
	const int p1[5] = { 2, 0, 1, 3, 4 };
	Meta mt2 = transpose_meta(p1, ms2);
	const int p2[4] = { 0, 3, 1, 2 };
	Meta mk = transpose_meta(p2, p.l.meta);
	Meta ms3 = dot_result_meta(mk, TENSOR_AXIS_RANGE_TRAILING, mt2, TENSOR_AXIS_RANGE_LEADING, 2);
	GemmBatch b3 = plan_dot(mk, TENSOR_AXIS_RANGE_TRAILING, mt2, TENSOR_AXIS_RANGE_LEADING, 2, ms3, true);
	auto perm_desc = [&](const int* perm, const Meta& from, const Meta& to) {
		std::vector<PermBlock> pbs = transpose_blocks(perm, from, to);
		PermBlock* d = (PermBlock*)dev_alloc(std::max<size_t>(1, pbs.size()) * sizeof(PermBlock));
		copy_h2d_async(d, pbs.data(), pbs.size() * sizeof(PermBlock), st);
		p.owned.push_back(d);
		return std::make_pair(d, (int)pbs.size());
	};
	const auto pt = perm_desc(p0, ms1, mt);
	const auto pt2 = perm_desc(p1, ms2, mt2);
	const auto pk = perm_desc(p2, p.l.meta, mk);
	// trace of s3 over the unit outer bonds: block (0, X, Q, C, 0) of s3 -> block (X, Q, C) of y
	std::vector<CopySeg> segs;
	{
		ct_long iy[3];
		for (ct_long k = 0; k < th.nblocks; k++) {
			if (!th.present(k)) continue;
			th.block_index(k, iy);
			const ct_long is3[5] = { 0, iy[0], iy[1], iy[2], 0 };
			const ct_long k3 = ms3.flat_index(is3);
			if (ms3.present(k3)) segs.push_back(CopySeg{ ms3.offset[k3], th.offset[k], th.block_size(k) });
		}
	}
	CopySeg* dseg = (CopySeg*)dev_alloc(std::max<size_t>(1, segs.size()) * sizeof(CopySeg));
	copy_h2d_async(dseg, segs.data(), segs.size() * sizeof(CopySeg), st);
	p.owned.push_back(dseg);
	const int nseg = (int)segs.size();
	// buffers
	p.interm.emplace_back(std::move(ms1));   // 0 s1
	p.interm.emplace_back(std::move(mt));    // 1 t
	p.interm.emplace_back(std::move(ms2));   // 2 s2
	p.interm.emplace_back(std::move(mt2));   // 3 t2
	p.interm.emplace_back(std::move(mk));    // 4 k
	p.interm.emplace_back(std::move(ms3));   // 5 s3
	double* s1d = p.interm[0].d; double* td = p.interm[1].d; double* s2d = p.interm[2].d;
	double* t2d = p.interm[3].d; double* kd = p.interm[4].d; double* s3d = p.interm[5].d;
	const ct_long n1 = p.interm[0].meta.nelem, n2 = p.interm[2].meta.nelem, n3 = p.interm[5].meta.nelem;
	const double* rd = p.r.d; const double* wd = p.w.d; const double* ld = p.l.d;
	const ct_long ny = th.nelem;
	// stage 1
	p.plain_ops.push_back([=](const double*, double*) { launch_fill(s1d, 0.0, n1, st); });
	for (const GemmTask& t : b1.tasks) for (int q = t.t0; q < t.t1; q++) {
		const GemmTerm term = b1.terms[q];
		p.plain_ops.push_back([=](const double* x, double*) { launch_plain_gemm_f64(t, term, GemmBases{ x, rd, s1d }, st); });
	}
	p.plain_stage_end[0] = (int)p.plain_ops.size();
	// stage 2
	p.plain_ops.push_back([=](const double*, double*) { launch_permute_blocks(s1d, td, pt.first, pt.second, st); });
	p.plain_ops.push_back([=](const double*, double*) { launch_fill(s2d, 0.0, n2, st); });
	for (const GemmTask& t : b2.tasks) for (int q = t.t0; q < t.t1; q++) {
		const GemmTerm term = b2.terms[q];
		p.plain_ops.push_back([=](const double*, double*) { launch_plain_gemm_f64(t, term, GemmBases{ wd, td, s2d }, st); });
	}
	p.plain_stage_end[1] = (int)p.plain_ops.size();
	// stage 3
	p.plain_ops.push_back([=](const double*, double*) { launch_permute_blocks(s2d, t2d, pt2.first, pt2.second, st); });
	p.plain_ops.push_back([=](const double*, double*) { launch_permute_blocks(ld, kd, pk.first, pk.second, st); });
	p.plain_ops.push_back([=](const double*, double*) { launch_fill(s3d, 0.0, n3, st); });
	for (const GemmTask& t : b3.tasks) for (int q = t.t0; q < t.t1; q++) {
		const GemmTerm term = b3.terms[q];
		p.plain_ops.push_back([=](const double*, double*) { launch_plain_gemm_f64(t, term, GemmBases{ kd, t2d, s3d }, st); });
	}
	p.plain_ops.push_back([=](const double*, double* y) { launch_fill(y, 0.0, ny, st); launch_copy_segments(s3d, y, dseg, nseg, st); });
	p.plain_stage_end[2] = (int)p.plain_ops.size();
	p.nterms = b1.terms.size() + b2.terms.size() + b3.terms.size();
	p.ntasks = b1.tasks.size() + b2.tasks.size() + b3.tasks.size();
	p.nitems = p.nterms;
	p.stage_flops[0] = b1.flops;
	p.stage_flops[1] = b2.flops;
	p.stage_flops[2] = b3.flops;
	p.interm_elems = 0;
	for (auto& t : p.interm) p.interm_elems += (double)t.meta.nelem;
}

// Plan-time autotuner: every launch form of every stage runs on the plan's own operands (one
// warm-up and three timed launches, HIP events on the plan stream); the fastest form is kept.
// time every launch form of each grouped stage on the plan's own descriptors (x = 1, y scratch)
// and keep the fastest; stage 2 also tries the column-stream form
void autotune_grouped(DevGemm* stage, const GemmBases* bases, const WApply* wa, bool* use_wapply, bool* use_f12, int64_t nx, hipStream_t s,
	float (*tuned_ms)[GEMM_NUM_FORMS])
{
	double* x = (double*)dev_alloc((size_t)std::max<int64_t>(nx, 1) * sizeof(double));
	double* y = (double*)dev_alloc((size_t)std::max<int64_t>(nx, 1) * sizeof(double));
	launch_fill(x, 1.0, nx, s);
	hipEvent_t e0, e1;
	CTG_HIP(hipEventCreate(&e0));
	CTG_HIP(hipEventCreate(&e1));
	auto time3 = [&](const std::function<void()>& f) {
		f();
		CTG_HIP(hipEventRecord(e0, s));
		for (int rep = 0; rep < 3; rep++) f();
		CTG_HIP(hipEventRecord(e1, s));
		CTG_HIP(hipEventSynchronize(e1));
		float ms = 0;
		CTG_HIP(hipEventElapsedTime(&ms, e0, e1));
		return ms / 3;
	};
	float best1 = 1e30f, best2 = 1e30f;
	for (int st = 0; st < 3; st++) {
		DevGemm& g = stage[st];
		GemmBases b = bases[st];
		if (st == 0) b.A = x;
		if (st == 2) b.C = y;
		float best = 1e30f;
		int best_form = g.form;
		for (int f = 0; f < GEMM_NUM_FORMS; f++) {
			g.form = f;
			const float ms = time3([&] { launch_grouped_gemm_f64(g, b, s); });
			if (tuned_ms) tuned_ms[st][f] = ms;
			if (ms < best) { best = ms; best_form = f; }
		}
		g.form = best_form;
		// the large tasks on the LDS-tiled form, the rest on the best form (CT_GEMM_LDS=0|1 overrides)
		if (g.nlitems > 0) {
			g.use_lds = true;
			const float ms = time3([&] { launch_grouped_gemm_f64(g, b, s); });
			g.use_lds = ms < best;
			if (g.use_lds) best = ms;
			const char* force = getenv("CT_GEMM_LDS");
			if (force && force[0] == '0') g.use_lds = false;
			if (force && force[0] == '1') g.use_lds = true;
		}
		if (st == 0) best1 = best;
		if (st == 1) best2 = best;
	}
	if (use_wapply) {
		*use_wapply = false;
		if (wa && wa->ok) {
			const float ms = time3([&] { launch_wapply(*wa, bases[1].B, bases[1].C, s); });
			*use_wapply = ms < best2;
			if (*use_wapply) best2 = ms;
			// CT_STAGE2=mfma|stream overrides the timing-based choice
			const char* force = getenv("CT_STAGE2");
			if (force && strcmp(force, "mfma") == 0) *use_wapply = false;
			if (force && strcmp(force, "stream") == 0) *use_wapply = true;
		}
	}
	if (use_f12) {
		*use_f12 = false;
		if (wa && wa->ok && wa->f12_ok) {
			const float ms = time3([&] { launch_f12(*wa, x, bases[0].B, bases[1].C, s); });
			*use_f12 = ms < best1 + best2;
			// CT_STAGE12=fused|separate overrides the timing-based choice
			const char* force = getenv("CT_STAGE12");
			if (force && strcmp(force, "separate") == 0) *use_f12 = false;
			if (force && strcmp(force, "fused") == 0) *use_f12 = true;
		}
	}
	CTG_HIP(hipEventDestroy(e0));
	CTG_HIP(hipEventDestroy(e1));
	CTG_HIP(hipStreamSynchronize(s));
	dev_free(x);
	dev_free(y);
}

void launch_grouped_stages(const DevGemm* stage, const GemmBases* bases, const WApply* wa, bool use_wapply, bool use_f12, const double* x, double* y,
	hipStream_t s, hipEvent_t* ev)
{
	const bool f12 = use_f12 && wa && wa->f12_ok;
	for (int st = 0; st < 3; st++) {
		GemmBases b = bases[st];
		if (st == 0) b.A = x;
		if (st == 2) b.C = y;
		if (st == 0 && f12) launch_f12(*wa, x, b.B, bases[1].C, s);
		else if (st == 1 && f12) { /* done by the fused launch */ }
		else if (st == 1 && use_wapply && wa && wa->ok) launch_wapply(*wa, b.B, b.C, s);
		else launch_grouped_gemm_f64(stage[st], b, s);
		if (ev) CTG_HIP(hipEventRecord(ev[st + 1], s));
	}
}

static void autotune_stages(MatvecPlan& p)
{
	autotune_grouped(p.stage, p.stage_bases, &p.wapply, &p.use_wapply, &p.use_f12, p.theta_meta.nelem, p.stream, p.tuned_ms);
	if (p.use_f12 && p.wapply.f12_ok) {
		// one launch for stages 1 and 2: theta, R and W in, T2 out (T1 never written)
		const double t1 = (double)p.interm[0].meta.nelem;
		p.stage_bytes[0] = p.stage_bytes[0] + p.stage_bytes[1] - 16.0 * t1;
		p.stage_bytes[1] = 0;
		p.stage_flops[0] += p.stage_flops[1];
		p.stage_flops[1] = 0;
	}
}

void validate_matvec_operands(const Meta& theta, const Meta& w, const Meta& l, const Meta& r)
{
	if (theta.ndim != 3 || w.ndim != 4 || l.ndim != 4 || r.ndim != 4)
		throw HipError{ hipErrorInvalidValue, "apply_local_hamiltonian expects degrees 3, 4, 4, 4" };
	if (l.dim_logical[0] != 1 || r.dim_logical[3] != 1 || l.qnums_logical[0] != r.qnums_logical[3])
		throw HipError{ hipErrorNotSupported, "outer virtual bonds of the environment blocks must have dimension 1 and equal quantum numbers" };
	if (!same_axis(theta, 2, r, 0) || !same_axis(theta, 1, w, 2) || !same_axis(theta, 0, l, 1) ||
		!same_axis(l, 2, w, 0) || !same_axis(r, 1, w, 3) ||
		!same_axis(theta, 0, l, 3) || !same_axis(theta, 1, w, 1) || !same_axis(theta, 2, r, 2))
		throw HipError{ hipErrorInvalidValue, "incompatible dimensions or quantum numbers of apply_local_hamiltonian operands" };
}

MatvecPlan* build_matvec_plan_dev(const Meta& theta, DevTensor&& w, DevTensor&& l, DevTensor&& r, int mode, hipStream_t s, bool autotune, bool reference_count)
{
	validate_matvec_operands(theta, w.meta, l.meta, r.meta);
	MatvecPlan* p = new MatvecPlan();
	p->mode = mode;
	p->stream = s;
	p->theta_meta = theta;
	p->w = std::move(w);
	p->l = std::move(l);
	p->r = std::move(r);
	try {
		if (mode == CT_MATVEC_PLAIN_FP64) build_plain(*p);
		else if (mode == CT_MATVEC_FUSED_FP64) {
			if (!build_fused_plan(p->fused, theta, p->w.meta, p->l.meta, p->r.meta, s))
				throw HipError{ hipErrorNotSupported, "fused matvec: a right-bond sector column does not fit in LDS" };
			p->stage_flops[0] = p->fused.flops;
			p->stage_bytes[0] = 8.0 * (2.0 * (double)theta.nelem + (double)p->w.meta.nelem + (double)p->l.meta.nelem + (double)p->r.meta.nelem);
			p->nunits = p->fused.nunits;
		}
		else if (mode == CT_MATVEC_PAIR_FP64) {
			// the unit's LDS budget: timed at plan creation over the three budgets (small budgets
			// win for small sectors: more, lighter units), unless CT_PAIR_LDS fixes it
			int lds = 0;
			if (autotune && !getenv("CT_PAIR_LDS")) {
				double* x = (double*)dev_alloc((size_t)std::max<int64_t>(theta.nelem, 1) * sizeof(double));
				double* y = (double*)dev_alloc((size_t)std::max<int64_t>(theta.nelem, 1) * sizeof(double));
				launch_fill(x, 1.0, theta.nelem, s);
				hipEvent_t e0 = nullptr, e1 = nullptr;
				CTG_HIP(hipEventCreate(&e0));
				CTG_HIP(hipEventCreate(&e1));
				float best = 1e30f;
				for (int cand : { 4608, 3072, 2048 }) {
					PairPlan c;
					try {
						if (!build_pair_plan(c, theta, p->w.meta, p->l.meta, p->r.meta, p->w.d, s, cand)) { c.release(); continue; }
						launch_pair_matvec(c, x, p->r.d, p->l.d, y, s);
						CTG_HIP(hipEventRecord(e0, s));
						for (int rep = 0; rep < 3; rep++) launch_pair_matvec(c, x, p->r.d, p->l.d, y, s);
						CTG_HIP(hipEventRecord(e1, s));
						CTG_HIP(hipEventSynchronize(e1));
						float ms = 0;
						CTG_HIP(hipEventElapsedTime(&ms, e0, e1));
						if (ms < best) { best = ms; lds = cand; }
					}
					catch (...) {
						(void)hipStreamSynchronize(s);
						c.release();
						dev_free(x);
						dev_free(y);
						(void)hipEventDestroy(e0);
						(void)hipEventDestroy(e1);
						throw;
					}
					c.release();
				}
				dev_free(x);
				dev_free(y);
				CTG_HIP(hipEventDestroy(e0));
				CTG_HIP(hipEventDestroy(e1));
			}
			if (!build_pair_plan(p->pair, theta, p->w.meta, p->l.meta, p->r.meta, p->w.d, s, lds))
				throw HipError{ hipErrorNotSupported, "sector-pair matvec: a class W_D or a single pair row and column exceeds the unit limits" };
			p->stage_flops[0] = p->pair.flops;
			// units read theta, R, L and W_D and write the partial blocks; the reduction reads them and writes y
			p->stage_bytes[0] = 8.0 * ((double)theta.nelem + (double)p->w.meta.nelem + (double)p->l.meta.nelem + (double)p->r.meta.nelem + (double)p->pair.part_elems);
			p->stage_bytes[1] = 8.0 * ((double)p->pair.part_elems + (double)theta.nelem);
			p->nunits = p->pair.nunits;
			p->ntasks = p->pair.nunits;      // plan statistics: units and reduction items
			p->nitems = p->pair.nred;
			p->interm_elems = (double)p->pair.part_elems;
		}
		else build_grouped(*p, mode == CT_MATVEC_GROUPED_FP64 && !autotune);
		if (mode == CT_MATVEC_GROUPED_FP64 && autotune) autotune_stages(*p);
		// without timing, stage 2 takes the column stream (faster at every size measured)
		if (mode == CT_MATVEC_GROUPED_FP64 && !autotune) p->use_wapply = p->wapply.ok;
		// split-FP32: stages 1 and 3 on FP32 MFMA with sliced operands, stage 2 as the FP64 column
		// stream when it applies (CT_STAGE2=mfma keeps the sliced MFMA stage 2)
		if (mode == CT_MATVEC_GROUPED_SPLIT) {
			const char* f2 = getenv("CT_STAGE2");
			p->use_wapply = p->wapply.ok && !(f2 && strcmp(f2, "mfma") == 0);
		}
		for (int i = 0; i < 4; i++) CTG_HIP(hipEventCreate(&p->ev[i]));
	}
	catch (...) {
		delete p;
		throw;
	}
	// the reference's flop count replans its unfused sequence; plans rebuilt at every bond use the
	// executed count instead
	p->flops = reference_count ? reference_flops(theta, p->w.meta, p->l.meta, p->r.meta)
		: (double)p->stage_flops[0] + p->stage_flops[1] + p->stage_flops[2];
	p->bytes = 8.0 * (2.0 * (double)theta.nelem + (double)p->w.meta.nelem + (double)p->l.meta.nelem + (double)p->r.meta.nelem);
	return p;
}

MatvecPlan* build_matvec_plan(const block_sparse_tensor* a, const block_sparse_tensor* w, const block_sparse_tensor* l, const block_sparse_tensor* r, int mode, bool autotune)
{
	hipStream_t s = default_stream();
	Meta th = meta_of(a);
	DevTensor dw = upload(w, s), dl = upload(l, s), dr = upload(r, s);
	return build_matvec_plan_dev(th, std::move(dw), std::move(dl), std::move(dr), mode, s, autotune);
}

float matvec_stage_time(MatvecPlan& p, int stage, const double* x, double* y, int reps)
{
	hipStream_t s = p.stream;
	auto launch = [&] {
		if (p.mode == CT_MATVEC_FUSED_FP64) {
			if (stage == 0) launch_fused_matvec(p.fused, x, p.w.d, p.l.d, p.r.d, y, s);
			return;
		}
		if (p.mode == CT_MATVEC_PAIR_FP64) {
			if (stage == 0) launch_pair_units(p.pair, x, p.r.d, p.l.d, s);
			else if (stage == 1) launch_pair_reduce(p.pair, y, s);
			return;
		}
		if (p.mode == CT_MATVEC_PLAIN_FP64) {
			const int i0 = stage == 0 ? 0 : p.plain_stage_end[stage - 1];
			for (int i = i0; i < p.plain_stage_end[stage]; i++) p.plain_ops[i](x, y);
			return;
		}
		GemmBases b = p.stage_bases[stage];
		if (stage == 0) b.A = x;
		if (stage == 2) b.C = y;
		if (p.mode == CT_MATVEC_GROUPED_SPLIT) {
			if (stage == 1 && p.use_wapply) launch_wapply(p.wapply, b.B, b.C, s);
			else launch_grouped_gemm_split(p.stage[stage], b, s);
		}
		else if (stage < 2 && p.use_f12 && p.wapply.f12_ok) { if (stage == 0) launch_f12(p.wapply, x, b.B, p.stage_bases[1].C, s); }
		else if (stage == 1 && p.use_wapply && p.wapply.ok) launch_wapply(p.wapply, b.B, b.C, s);
		else launch_grouped_gemm_f64(p.stage[stage], b, s);
	};
	launch();
	CTG_HIP(hipEventRecord(p.ev[0], s));
	for (int r = 0; r < reps; r++) launch();
	CTG_HIP(hipEventRecord(p.ev[1], s));
	CTG_HIP(hipEventSynchronize(p.ev[1]));
	float ms = 0;
	CTG_HIP(hipEventElapsedTime(&ms, p.ev[0], p.ev[1]));
	return ms / reps;
}

void matvec_apply(MatvecPlan& p, const double* x, double* y, float* stage_ms)
{
	hipStream_t s = p.stream;
	if (stage_ms) CTG_HIP(hipEventRecord(p.ev[0], s));
	if (p.mode == CT_MATVEC_FUSED_FP64) {
		launch_fused_matvec(p.fused, x, p.w.d, p.l.d, p.r.d, y, s);
		if (stage_ms) for (int st = 1; st < 4; st++) CTG_HIP(hipEventRecord(p.ev[st], s));
	}
	else if (p.mode == CT_MATVEC_PAIR_FP64) {
		launch_pair_units(p.pair, x, p.r.d, p.l.d, s);
		if (stage_ms) CTG_HIP(hipEventRecord(p.ev[1], s));
		launch_pair_reduce(p.pair, y, s);
		if (stage_ms) for (int st = 2; st < 4; st++) CTG_HIP(hipEventRecord(p.ev[st], s));
	}
	else if (p.mode == CT_MATVEC_PLAIN_FP64) {
		int i = 0;
		for (int st = 0; st < 3; st++) {
			for (; i < p.plain_stage_end[st]; i++) p.plain_ops[i](x, y);
			if (stage_ms) CTG_HIP(hipEventRecord(p.ev[st + 1], s));
		}
	}
	else if (p.mode == CT_MATVEC_GROUPED_SPLIT) {
		for (int st = 0; st < 3; st++) {
			GemmBases b = p.stage_bases[st];
			if (st == 0) b.A = x;
			if (st == 2) b.C = y;
			if (st == 1 && p.use_wapply) launch_wapply(p.wapply, b.B, b.C, s);
			else launch_grouped_gemm_split(p.stage[st], b, s);
			if (stage_ms) CTG_HIP(hipEventRecord(p.ev[st + 1], s));
		}
	}
	else launch_grouped_stages(p.stage, p.stage_bases, &p.wapply, p.use_wapply, p.use_f12, x, y, s, stage_ms ? p.ev : nullptr);
	if (stage_ms) {
		CTG_HIP(hipEventSynchronize(p.ev[3]));
		for (int st = 0; st < 3; st++) CTG_HIP(hipEventElapsedTime(&stage_ms[st], p.ev[st], p.ev[st + 1]));
	}
}

}  // namespace ctg
