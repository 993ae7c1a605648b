// Fused single-launch H_eff matvec in FP64 (apply_local_hamiltonian, src/algorithm/chain_ops.c:353-390,
// with unit outer environment bonds).
//
// The output y(X,Q,C) = sum L(0,A,M,X) W(M,Q,P,N) theta(A,P,B) R(B,N,C,0) keeps the right-bond
// sector C and its column c free in all three contractions, so the work splits into independent
// units of k columns of one sector C. A workgroup per unit:
//   phase 0  stages R(B,N,C)[:, :, c0:c0+k] for every (B,N) in LDS ("panels": per B, rows b,
//            columns (N,n) concatenated, then the k columns innermost);
//   phase 1  T1(A,P)[a p (N n)][kk] = sum_b theta(A,P,B)[a p b] R(B,N,C)[b n c]   (B = qA + qP)
//   phase 2  T2(A,M,Q)[a m q][kk]  = sum_(P,N) sum_(p,n) W(M,Q,P,N)[m q p n] T1(A,P)[a p (N n)]
//   phase 3  y(X,Q,C)[x q c]       = sum_(A,M) sum_(a,m) L(0,A,M,X)[a m x] T2(A,M,Q)[a m q]
// T1 and T2 never leave LDS; each y entry is written once by one thread (or reduced from a fixed
// number of LDS partials in a fixed order), so results are deterministic. MI355X executes FP64
// FMA on the vector ALUs at the same peak as FP64 MFMA, and the per-sector blocks here are small
// and ragged, so the contractions use v_fma_f64 with the k columns as the register tile.
// LDS offsets in the tables are per column and are multiplied by the unit's k at run time.

#include <algorithm>
#include <numeric>
#include <tuple>

#include "ct_internal.h"

namespace ctg {

constexpr int FZ_THREADS = 1024;
constexpr int FZ_KMAX = 4;      // columns per unit
constexpr int FZ_JC = 4;        // T1 columns (N n) per phase-1 item
constexpr int FZ_NXCD = 8;      // MI355X: 8 XCDs; workgroup i runs on XCD i mod 8
constexpr size_t FZ_LDS_BUDGET = 64 * 1024;

struct FzArgs
{
	const FzSector* sectors;
	const FzUnit* units;
	const int* order;
	const int* map;
	const FzCopy* copies;
	const FzTask1* t1;
	const FzTask2* t2;
	const FzTerm2* u2;
	const FzTask3* t3;
	const FzTerm3* u3;
	const double* x;
	const double* w;
	const double* l;
	const double* r;
	double* y;
};

__global__ __launch_bounds__(FZ_THREADS) void k_fused_matvec(FzArgs a)
{
	extern __shared__ double lds[];
	const int uid = a.order[blockIdx.x];
	if (uid < 0) return;
	const FzUnit u = a.units[uid];
	const FzSector sc = a.sectors[u.sector];
	const int k = u.k, c0 = u.c0, tid = threadIdx.x;

	// phase 0: R columns c0 .. c0+k-1 into the LDS panels
	for (int e = tid; e < sc.n0 * k; e += FZ_THREADS) {
		const int epc = e / k, cc = e - epc * k;
		const FzCopy cp = a.copies[a.map[sc.m0 + epc]];
		const int rel = epc - cp.rowstart;
		const int b = rel / cp.dN, n = rel - b * cp.dN;
		lds[(cp.rb + b * cp.nb + cp.j0 + n) * k + cc] = a.r[cp.goff + (int64_t)(b * cp.dN + n) * cp.dC + c0 + cc];
	}
	__syncthreads();

	// phase 1: one item = one row (a p) of T1(A,P) and up to FZ_JC of its (N n) columns
	for (int it = tid; it < sc.rows1; it += FZ_THREADS) {
		const FzTask1 t = a.t1[a.map[sc.m1 + it]];
		const int rel = it - t.rowstart;
		const int ap = rel / t.nJ, jc = (rel - ap * t.nJ) * FZ_JC;
		const int nj = min(FZ_JC, t.nb - jc);
		double acc[FZ_JC][FZ_KMAX];
		#pragma unroll
		for (int j = 0; j < FZ_JC; j++)
			#pragma unroll
			for (int cc = 0; cc < FZ_KMAX; cc++) acc[j][cc] = 0;
		const double* th = a.x + t.th_off + (int64_t)ap * t.dB;
		for (int b = 0; b < t.dB; b++) {
			const double v = th[b];
			const double* rp = lds + (t.rb + b * t.nb + jc) * k;
			#pragma unroll
			for (int j = 0; j < FZ_JC; j++) {
				if (j < nj) {
					#pragma unroll
					for (int cc = 0; cc < FZ_KMAX; cc++)
						if (cc < k) acc[j][cc] = fma(v, rp[j * k + cc], acc[j][cc]);
				}
			}
		}
		double* out = lds + (t.t1 + ap * t.nb + jc) * k;
		#pragma unroll
		for (int j = 0; j < FZ_JC; j++) {
			if (j < nj) {
				#pragma unroll
				for (int cc = 0; cc < FZ_KMAX; cc++)
					if (cc < k) out[j * k + cc] = acc[j][cc];
			}
		}
	}
	__syncthreads();

	// phase 2: one item = one row (a m q) of T2(A,M,Q), all k columns
	for (int it = tid; it < sc.rows2; it += FZ_THREADS) {
		const FzTask2 t = a.t2[a.map[sc.m2 + it]];
		const int rel = it - t.rowstart;
		const int ai = rel / t.dMQ, mq = rel - ai * t.dMQ;
		double acc[FZ_KMAX];
		#pragma unroll
		for (int cc = 0; cc < FZ_KMAX; cc++) acc[cc] = 0;
		for (int q = t.u0; q < t.u1; q++) {
			const FzTerm2 tm = a.u2[q];
			const double* wp = a.w + tm.w_off + (int64_t)mq * tm.dP * tm.dN;
			for (int p = 0; p < tm.dP; p++) {
				const double* tp = lds + (tm.t1 + (ai * tm.dP + p) * tm.nb + tm.j0) * k;
				for (int n = 0; n < tm.dN; n++) {
					const double wv = wp[p * tm.dN + n];
					#pragma unroll
					for (int cc = 0; cc < FZ_KMAX; cc++)
						if (cc < k) acc[cc] = fma(wv, tp[n * k + cc], acc[cc]);
				}
			}
		}
		double* out = lds + (t.t2 + rel) * k;
		#pragma unroll
		for (int cc = 0; cc < FZ_KMAX; cc++)
			if (cc < k) out[cc] = acc[cc];
	}
	__syncthreads();

	// phase 3: one item = one row (x q) of y(X,Q,C) and one of 'split' interleaved parts of its terms
	const int split = sc.split;
	for (int it = tid; it < sc.rows3 * split; it += FZ_THREADS) {
		const int sidx = it / sc.rows3, row = it - sidx * sc.rows3;
		const FzTask3 t = a.t3[a.map[sc.m3 + row]];
		const int rel = row - t.rowstart;
		const int xi = rel / t.dQ, qi = rel - xi * t.dQ;
		double acc[FZ_KMAX];
		#pragma unroll
		for (int cc = 0; cc < FZ_KMAX; cc++) acc[cc] = 0;
		for (int q = t.u0 + sidx; q < t.u1; q += split) {
			const FzTerm3 tm = a.u3[q];
			const double* lp = a.l + tm.l_off + xi;
			const double* tp = lds + (tm.t2 + qi) * k;
			for (int am = 0; am < tm.dAM; am++) {
				const double lv = lp[(int64_t)am * t.dX];
				const double* tq = tp + am * t.dQ * k;
				#pragma unroll
				for (int cc = 0; cc < FZ_KMAX; cc++)
					if (cc < k) acc[cc] = fma(lv, tq[cc], acc[cc]);
			}
		}
		if (split == 1) {
			double* yp = a.y + t.y_off + (int64_t)(xi * t.dQ + qi) * t.dC + c0;
			#pragma unroll
			for (int cc = 0; cc < FZ_KMAX; cc++)
				if (cc < k) yp[cc] = acc[cc];
		}
		else {
			// partials over the R / T1 region, which phase 3 no longer reads
			double* red = lds + (sidx * sc.rows3 + row) * k;
			#pragma unroll
			for (int cc = 0; cc < FZ_KMAX; cc++)
				if (cc < k) red[cc] = acc[cc];
		}
	}
	if (split > 1) {
		__syncthreads();
		for (int it = tid; it < sc.rows3 * k; it += FZ_THREADS) {
			const int row = it / k, cc = it - row * k;
			double v = 0;
			for (int s2 = 0; s2 < split; s2++) v += lds[(s2 * sc.rows3 + row) * k + cc];
			const FzTask3 t = a.t3[a.map[sc.m3 + row]];
			const int rel = row - t.rowstart;
			const int xi = rel / t.dQ, qi = rel - xi * t.dQ;
			a.y[t.y_off + (int64_t)(xi * t.dQ + qi) * t.dC + c0 + cc] = v;
		}
	}
}

//________________________________________________________________________________________________________________________
//
// host planner

struct FusedHost
{
	std::vector<FzSector> sectors;
	std::vector<FzUnit> units;
	std::vector<int> order;
	std::vector<int> map;
	std::vector<FzCopy> copies;
	std::vector<FzTask1> t1;
	std::vector<FzTask2> t2;
	std::vector<FzTerm2> u2;
	std::vector<FzTask3> t3;
	std::vector<FzTerm3> u3;
	size_t lds_doubles = 0;
	double fma = 0;
};

static bool plan_fused_host(FusedHost& h, const Meta& th, const Meta& W, const Meta& L, const Meta& R, size_t lds_budget)
{
	const qnumber qo = L.qnums_blocks[0][0];
	const ct_long nA = th.dim_blocks[0], nP = th.dim_blocks[1], nB = th.dim_blocks[2];
	const ct_long nN = R.dim_blocks[1], nC = R.dim_blocks[2];
	const ct_long nM = W.dim_blocks[0], nQ = W.dim_blocks[1];
	auto bc = [](const Meta& m, int ax, ct_long i) { return m.block_count[ax][i]; };
	auto qb = [](const Meta& m, int ax, ct_long i) { return m.qnums_blocks[ax][i]; };
	const size_t budget = lds_budget / sizeof(double);
	std::vector<double> unit_cost;

	for (ct_long C = 0; C < nC; C++) {
		const int dC = (int)bc(R, 2, C);
		FzSector sc = {};
		sc.dC = dC;
		// R panels: per B, the N with R(B,N,C,0) present, concatenated
		std::vector<int> rb(nB, -1), nbB(nB, 0);
		std::vector<int> j0((size_t)(nB * nN), -1);
		int r_pc = 0;
		sc.m0 = (int)h.map.size();
		for (ct_long B = 0; B < nB; B++) {
			int nb = 0;
			for (ct_long N = 0; N < nN; N++) {
				const ct_long ir[4] = { B, N, C, 0 };
				if (!R.present(R.flat_index(ir))) continue;
				j0[B * nN + N] = nb;
				nb += (int)bc(R, 1, N);
			}
			if (nb == 0) continue;
			rb[B] = r_pc;
			nbB[B] = nb;
			const int dB = (int)bc(R, 0, B);
			for (ct_long N = 0; N < nN; N++) {
				if (j0[B * nN + N] < 0) continue;
				const ct_long ir[4] = { B, N, C, 0 };
				const int dN = (int)bc(R, 1, N);
				FzCopy cp = { R.offset[R.flat_index(ir)], sc.n0, dB, dN, dC, rb[B], nb, j0[B * nN + N] };
				for (int e = 0; e < dB * dN; e++) h.map.push_back((int)h.copies.size());
				h.copies.push_back(cp);
				sc.n0 += dB * dN;
			}
			r_pc += dB * nb;
		}
		// phase 1: T1(A,P) for theta(A,P,B) present with a non-empty panel of B
		std::vector<int> t1off((size_t)(nA * nP), -1);
		std::vector<int> t1B((size_t)(nA * nP), -1);
		int t1_pc = 0;
		double fma_pc = 0;
		sc.m1 = (int)h.map.size();
		for (ct_long A = 0; A < nA; A++) for (ct_long P = 0; P < nP; P++) {
			const ct_long B = th.find_qnum(2, qb(th, 0, A) + qb(th, 1, P));
			if (B < 0 || rb[B] < 0) continue;
			const ct_long ith[3] = { A, P, B };
			if (!th.present(th.flat_index(ith))) continue;
			const int dA = (int)bc(th, 0, A), dP = (int)bc(th, 1, P), dB = (int)bc(th, 2, B);
			const int nb = nbB[B];
			const int nJ = (nb + FZ_JC - 1) / FZ_JC;
			t1off[A * nP + P] = r_pc + t1_pc;
			t1B[A * nP + P] = (int)B;
			FzTask1 t = { th.offset[th.flat_index(ith)], sc.rows1, dB, nb, nJ, rb[B], r_pc + t1_pc };
			for (int e = 0; e < dA * dP * nJ; e++) h.map.push_back((int)h.t1.size());
			h.t1.push_back(t);
			sc.rows1 += dA * dP * nJ;
			t1_pc += dA * dP * nb;
			fma_pc += (double)dA * dP * nb * dB;
		}
		// phase 2: T2(A,M,Q) with qM + qQ + qA - qC - qo = 0 and at least one term
		std::vector<int> t2off((size_t)(nA * nM * nQ), -1);
		int t2_pc = 0;
		const size_t t2_first = h.t2.size();
		sc.m2 = (int)h.map.size();
		for (ct_long A = 0; A < nA; A++) for (ct_long M = 0; M < nM; M++) for (ct_long Q = 0; Q < nQ; Q++) {
			if (qb(W, 0, M) + qb(W, 1, Q) + qb(th, 0, A) - qb(R, 2, C) - qo != 0) continue;
			const int u0 = (int)h.u2.size();
			double kk = 0;
			for (ct_long P = 0; P < nP; P++) {
				const ct_long N = W.find_qnum(3, qb(W, 0, M) + qb(W, 1, Q) - qb(W, 2, P));
				if (N < 0) continue;
				const ct_long iw[4] = { M, Q, P, N };
				if (!W.present(W.flat_index(iw))) continue;
				const int to = t1off[A * nP + P];
				if (to < 0) continue;
				const int B = t1B[A * nP + P];
				if (j0[B * nN + N] < 0) continue;   // R(B,N,C) absent: T1 block is zero
				const int dP = (int)bc(W, 2, P), dN = (int)bc(W, 3, N);
				h.u2.push_back(FzTerm2{ W.offset[W.flat_index(iw)], dP, dN, to, nbB[B], j0[B * nN + N], 0 });
				kk += (double)dP * dN;
			}
			if ((int)h.u2.size() == u0) continue;
			const int dA = (int)bc(th, 0, A), dM = (int)bc(W, 0, M), dQ = (int)bc(W, 1, Q);
			t2off[(A * nM + M) * nQ + Q] = t2_pc;   // relative to the T2 region (rebased below)
			FzTask2 t = { sc.rows2, dM * dQ, t2_pc, u0, (int)h.u2.size(), 0 };
			for (int e = 0; e < dA * dM * dQ; e++) h.map.push_back((int)h.t2.size());
			h.t2.push_back(t);
			sc.rows2 += dA * dM * dQ;
			t2_pc += dA * dM * dQ;
			fma_pc += (double)dA * dM * dQ * kk;
		}
		const int t2_base = r_pc + t1_pc;
		for (size_t i = t2_first; i < h.t2.size(); i++) h.t2[i].t2 += t2_base;
		// phase 3: every present block y(X,Q,C)
		sc.m3 = (int)h.map.size();
		for (ct_long X = 0; X < th.dim_blocks[0]; X++) for (ct_long Q = 0; Q < th.dim_blocks[1]; Q++) {
			const ct_long iy[3] = { X, Q, C };
			const ct_long ky = th.flat_index(iy);
			if (!th.present(ky)) continue;
			const int dX = (int)bc(th, 0, X), dQ = (int)bc(th, 1, Q);
			const int u0 = (int)h.u3.size();
			double kk = 0;
			for (ct_long A = 0; A < nA; A++) {
				const ct_long M = L.find_qnum(2, qo - qb(L, 1, A) + qb(L, 3, X));
				if (M < 0) continue;
				const ct_long il[4] = { 0, A, M, X };
				if (!L.present(L.flat_index(il))) continue;
				const int to = t2off[(A * nM + M) * nQ + Q];
				if (to < 0) continue;
				const int dAM = (int)(bc(L, 1, A) * bc(L, 2, M));
				h.u3.push_back(FzTerm3{ L.offset[L.flat_index(il)], dAM, t2_base + to });
				kk += dAM;
			}
			FzTask3 t = { th.offset[ky], sc.rows3, dX, dQ, dC, u0, (int)h.u3.size() };
			for (int e = 0; e < dX * dQ; e++) h.map.push_back((int)h.t3.size());
			h.t3.push_back(t);
			sc.rows3 += dX * dQ;
			fma_pc += (double)dX * dQ * kk;
		}
		if (sc.rows3 == 0) continue;   // no output block in this sector
		const int lds_pc = r_pc + t1_pc + t2_pc;
		sc.lds_pc = std::max(lds_pc, 1);
		sc.split = 1;
		while (sc.split < 8 && 2 * sc.split * sc.rows3 <= FZ_THREADS && 2 * sc.split * sc.rows3 <= r_pc + t1_pc) sc.split *= 2;
		if ((size_t)sc.lds_pc > budget) return false;
		const int kmax = (int)std::min<size_t>({ (size_t)FZ_KMAX, (size_t)dC, budget / (size_t)sc.lds_pc });
		const int isec = (int)h.sectors.size();
		h.sectors.push_back(sc);
		for (int c0 = 0; c0 < dC; c0 += kmax) {
			const int kk = std::min(kmax, dC - c0);
			h.units.push_back(FzUnit{ isec, c0, kk, 0 });
			unit_cost.push_back(kk * fma_pc + 64.0 * (sc.rows1 + sc.rows2 + sc.rows3));
			h.lds_doubles = std::max(h.lds_doubles, (size_t)sc.lds_pc * kk);
		}
		h.fma += fma_pc * dC;
	}
	// XCD-aware order: longest-processing-time assignment of units to the 8 XCDs, each XCD's
	// units by decreasing cost; workgroup slot*8 + xcd runs the slot-th unit of that XCD
	std::vector<int> idx(h.units.size());
	std::iota(idx.begin(), idx.end(), 0);
	std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return unit_cost[a] > unit_cost[b]; });
	std::vector<std::vector<int>> bins(FZ_NXCD);
	std::vector<double> load(FZ_NXCD, 0.0);
	for (int i : idx) {
		const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
		bins[x].push_back(i);
		load[x] += unit_cost[i];
	}
	size_t slots = 0;
	for (auto& b : bins) slots = std::max(slots, b.size());
	h.order.assign(slots * FZ_NXCD, -1);
	for (int x = 0; x < FZ_NXCD; x++)
		for (size_t s = 0; s < bins[x].size(); s++) h.order[s * FZ_NXCD + x] = bins[x][s];
	return true;
}

template <typename T>
static T* upload_vec(const std::vector<T>& v, hipStream_t s)
{
	T* d = (T*)dev_alloc(std::max<size_t>(1, v.size()) * sizeof(T));
	if (!v.empty()) CTG_HIP(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
	return d;
}

bool build_fused_plan(FusedPlan& f, const Meta& theta, const Meta& w, const Meta& l, const Meta& r, hipStream_t s)
{
	FusedHost h;
	if (!plan_fused_host(h, theta, w, l, r, FZ_LDS_BUDGET)) return false;
	f.sectors = upload_vec(h.sectors, s);
	f.units = upload_vec(h.units, s);
	f.order = upload_vec(h.order, s);
	f.map = upload_vec(h.map, s);
	f.copies = upload_vec(h.copies, s);
	f.t1 = upload_vec(h.t1, s);
	f.t2 = upload_vec(h.t2, s);
	f.u2 = upload_vec(h.u2, s);
	f.t3 = upload_vec(h.t3, s);
	f.u3 = upload_vec(h.u3, s);
	CTG_HIP(hipStreamSynchronize(s));
	f.grid = (int)h.order.size();
	f.nunits = (int)h.units.size();
	f.lds_bytes = std::max<size_t>(h.lds_doubles, 1) * sizeof(double);
	f.flops = 2.0 * h.fma;
	return true;
}

void FusedPlan::release()
{
	dev_free(sectors); dev_free(units); dev_free(order); dev_free(map); dev_free(copies);
	dev_free(t1); dev_free(t2); dev_free(u2); dev_free(t3); dev_free(u3);
	sectors = nullptr; units = nullptr; order = nullptr; map = nullptr; copies = nullptr;
	t1 = nullptr; t2 = nullptr; u2 = nullptr; t3 = nullptr; u3 = nullptr;
}

void launch_fused_matvec(const FusedPlan& f, const double* x, const double* w, const double* l, const double* r, double* y, hipStream_t s)
{
	if (f.grid == 0) return;
	FzArgs a = { f.sectors, f.units, f.order, f.map, f.copies, f.t1, f.t2, f.u2, f.t3, f.u3, x, w, l, r, y };
	hipLaunchKernelGGL(k_fused_matvec, dim3(f.grid), dim3(FZ_THREADS), f.lds_bytes, s, a);
	CTG_HIP(hipGetLastError());
}

}  // namespace ctg
