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
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <tuple>

#include <mutex>
#include "ct_internal.h"

namespace ctg {

constexpr int FZ_THREADS = 1024;
constexpr int FZ_KMAX = 4;      // columns per unit
constexpr int FZ_JC = 4;        // T1 columns (N n) per phase-1 item
constexpr int FZ_LB = 8;        // global loads in flight per thread in the contraction loops
constexpr int FZ_NXCD = 8;      // MI355X: 8 XCDs; workgroup i runs on XCD i mod 8
constexpr size_t FZ_LDS_BUDGET = 64 * 1024;        // per workgroup: two units fit the 160 KB LDS of a CU
constexpr size_t FZ_LDS_MAX = 160 * 1024;          // a sector whose column needs more runs one column per unit up to this

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
		// FZ_LB theta loads in flight per thread, then their FMAs
		for (int b0 = 0; b0 < t.dB; b0 += FZ_LB) {
			double v[FZ_LB];
			#pragma unroll
			for (int i = 0; i < FZ_LB; i++) v[i] = (b0 + i < t.dB) ? th[b0 + i] : 0.0;
			#pragma unroll
			for (int i = 0; i < FZ_LB; i++) {
				if (b0 + i < t.dB) {
					const double* rp = lds + (t.rb + (b0 + i) * t.nb + jc) * k;
					#pragma unroll
					for (int j = 0; j < FZ_JC; j++) {
						if (j < nj) {
							#pragma unroll
							for (int cc = 0; cc < FZ_KMAX; cc++)
								if (cc < k) acc[j][cc] = fma(v[i], rp[j * k + cc], acc[j][cc]);
						}
					}
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
		FzTerm2 nxt;
		if (t.u0 < t.u1) nxt = a.u2[t.u0];
		for (int q = t.u0; q < t.u1; q++) {
			const FzTerm2 tm = nxt;
			if (q + 1 < t.u1) nxt = a.u2[q + 1];
			const double* wp = a.w + tm.w_off + (int64_t)mq * tm.dP * tm.dN;
			const int pn = tm.dP * tm.dN;
			for (int e0 = 0; e0 < pn; e0 += FZ_LB) {
				double wv[FZ_LB];
				#pragma unroll
				for (int i = 0; i < FZ_LB; i++) wv[i] = (e0 + i < pn) ? wp[e0 + i] : 0.0;
				#pragma unroll
				for (int i = 0; i < FZ_LB; i++) {
					if (e0 + i < pn) {
						const int p = (e0 + i) / tm.dN, n = (e0 + i) - p * tm.dN;
						const double* tp = lds + (tm.t1 + (ai * tm.dP + p) * tm.nb + tm.j0 + n) * k;
						#pragma unroll
						for (int cc = 0; cc < FZ_KMAX; cc++)
							if (cc < k) acc[cc] = fma(wv[i], tp[cc], acc[cc]);
					}
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
		FzTerm3 nxt;
		if (t.u0 + sidx < t.u1) nxt = a.u3[t.u0 + sidx];
		for (int q = t.u0 + sidx; q < t.u1; q += split) {
			const FzTerm3 tm = nxt;
			if (q + split < t.u1) nxt = a.u3[q + split];   // next descriptor in flight during this term
			const double* lp = a.l + tm.l_off + xi;
			const double* tp = lds + (tm.t2 + qi) * k;
			for (int am0 = 0; am0 < tm.dAM; am0 += FZ_LB) {
				double lv[FZ_LB];
				#pragma unroll
				for (int i = 0; i < FZ_LB; i++) lv[i] = (am0 + i < tm.dAM) ? lp[(int64_t)(am0 + i) * t.dX] : 0.0;
				#pragma unroll
				for (int i = 0; i < FZ_LB; i++) {
					if (am0 + i < tm.dAM) {
						const double* tq = tp + (am0 + i) * t.dQ * k;
						#pragma unroll
						for (int cc = 0; cc < FZ_KMAX; cc++)
							if (cc < k) acc[cc] = fma(lv[i], tq[cc], acc[cc]);
					}
				}
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
	std::vector<double> cost;   // estimated cost of each unit
	size_t lds_doubles = 0;
	double fma = 0;
};

static bool plan_fused_host(FusedHost& h, const Meta& th, const Meta& W, const Meta& L, const Meta& R, size_t lds_budget, size_t lds_max)
{
	const qnumber qo = L.qnums_blocks[0][0];
	const ct_long nA = th.dim_blocks[0], nP = th.dim_blocks[1], nB = th.dim_blocks[2];
	const ct_long nN = R.dim_blocks[1], nC = R.dim_blocks[2];
	const ct_long nM = W.dim_blocks[0], nQ = W.dim_blocks[1];
	auto bc = [](const Meta& m, int ax, ct_long i) { return m.block_count[ax][i]; };
	auto qb = [](const Meta& m, int ax, ct_long i) { return m.qnums_blocks[ax][i]; };
	const size_t budget = lds_budget / sizeof(double);

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
		if ((size_t)sc.lds_pc * sizeof(double) > lds_max) return false;
		const int kmax = (size_t)sc.lds_pc > budget ? 1 : (int)std::min<size_t>({ (size_t)FZ_KMAX, (size_t)dC, budget / (size_t)sc.lds_pc });
		const int isec = (int)h.sectors.size();
		h.sectors.push_back(sc);
		for (int c0 = 0; c0 < dC; c0 += kmax) {
			const int kk = std::min(kmax, dC - c0);
			h.units.push_back(FzUnit{ isec, c0, kk, 0 });
			h.cost.push_back(kk * fma_pc + 64.0 * (sc.rows1 + sc.rows2 + sc.rows3));
			h.lds_doubles = std::max(h.lds_doubles, (size_t)sc.lds_pc * kk);
		}
		h.fma += fma_pc * dC;
	}
	return true;
}

// XCD-aware order: longest-processing-time assignment of units to the 8 XCDs, each XCD's units by
// decreasing cost; workgroup slot*8 + xcd runs the slot-th unit of that XCD
static void order_units(FusedHost& h)
{
	std::vector<int> idx(h.units.size());
	std::iota(idx.begin(), idx.end(), 0);
	std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return h.cost[a] > h.cost[b]; });
	std::vector<std::vector<int>> bins(FZ_NXCD);
	std::vector<double> load(FZ_NXCD, 0.0);
	for (int i : idx) {
		const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
		bins[x].push_back(i);
		load[x] += h.cost[i];
	}
	size_t slots = 0;
	for (auto& b : bins) slots = std::max(slots, b.size());
	h.order.assign(slots * FZ_NXCD, -1);
	for (int x = 0; x < FZ_NXCD; x++)
		for (size_t s = 0; s < bins[x].size(); s++) h.order[s * FZ_NXCD + x] = bins[x][s];
}

// Append problem b to a, its operands living at element offsets x_off (theta and y), w_off,
// l_off and r_off of concatenated buffers.
static void append_fused(FusedHost& a, const FusedHost& b, int64_t x_off, int64_t w_off, int64_t l_off, int64_t r_off)
{
	const int map0 = (int)a.map.size(), sec0 = (int)a.sectors.size();
	const int cp0 = (int)a.copies.size(), t10 = (int)a.t1.size(), t20 = (int)a.t2.size(), u20 = (int)a.u2.size();
	const int t30 = (int)a.t3.size(), u30 = (int)a.u3.size();
	for (FzSector sc : b.sectors) { sc.m0 += map0; sc.m1 += map0; sc.m2 += map0; sc.m3 += map0; a.sectors.push_back(sc); }
	for (FzUnit u : b.units) { u.sector += sec0; a.units.push_back(u); }
	a.cost.insert(a.cost.end(), b.cost.begin(), b.cost.end());
	// map entries index the table of their phase; the phase of a map range follows from the sectors
	a.map.resize(map0 + b.map.size());
	for (const FzSector& sc : b.sectors) {
		for (int e = 0; e < sc.n0; e++) a.map[map0 + sc.m0 + e] = b.map[sc.m0 + e] + cp0;
		for (int e = 0; e < sc.rows1; e++) a.map[map0 + sc.m1 + e] = b.map[sc.m1 + e] + t10;
		for (int e = 0; e < sc.rows2; e++) a.map[map0 + sc.m2 + e] = b.map[sc.m2 + e] + t20;
		for (int e = 0; e < sc.rows3; e++) a.map[map0 + sc.m3 + e] = b.map[sc.m3 + e] + t30;
	}
	for (FzCopy c : b.copies) { c.goff += r_off; a.copies.push_back(c); }
	for (FzTask1 t : b.t1) { t.th_off += x_off; a.t1.push_back(t); }
	for (FzTask2 t : b.t2) { t.u0 += u20; t.u1 += u20; a.t2.push_back(t); }
	for (FzTerm2 t : b.u2) { t.w_off += w_off; a.u2.push_back(t); }
	for (FzTask3 t : b.t3) { t.y_off += x_off; t.u0 += u30; t.u1 += u30; a.t3.push_back(t); }
	for (FzTerm3 t : b.u3) { t.l_off += l_off; a.u3.push_back(t); }
	a.lds_doubles = std::max(a.lds_doubles, b.lds_doubles);
	a.fma += b.fma;
}

// largest dynamic LDS allocation of one workgroup on the current device (160 KB on MI355X)
static size_t lds_limit()
{
	static const size_t lim = [] {
		int dev = 0, v = 0;
		CTG_HIP(hipGetDevice(&dev));
		CTG_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
		return std::min<size_t>(FZ_LDS_MAX, (size_t)v);
	}();
	return lim;
}

template <typename T>
static T* upload_vec(const std::vector<T>& v, hipStream_t s)
{
	T* d = (T*)dev_alloc(std::max<size_t>(1, v.size()) * sizeof(T));
	copy_h2d_async(d, v.data(), v.size() * sizeof(T), s);
	return d;
}

static void upload_fused(FusedPlan& f, const FusedHost& h, hipStream_t s)
{
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
}

bool build_fused_plan(FusedPlan& f, const Meta& theta, const Meta& w, const Meta& l, const Meta& r, hipStream_t s)
{
	FusedHost h;
	if (!plan_fused_host(h, theta, w, l, r, FZ_LDS_BUDGET, lds_limit())) return false;
	order_units(h);
	upload_fused(f, h, s);
	return true;
}

bool build_fused_batch(FusedPlan& f, const std::vector<FusedProblem>& probs, hipStream_t s)
{
	FusedHost all;
	for (const FusedProblem& pb : probs) {
		FusedHost h;
		if (!plan_fused_host(h, *pb.theta, *pb.w, *pb.l, *pb.r, FZ_LDS_BUDGET, lds_limit())) return false;
		append_fused(all, h, pb.x_off, pb.w_off, pb.l_off, pb.r_off);
	}
	order_units(all);
	upload_fused(f, all, s);
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
	if (f.lds_bytes > 64 * 1024) {
		// the kernel is allowed the device's whole LDS once, before any launch above 64 KB (plans of
		// several host threads launch concurrently; no thread can be left with a smaller grant)
		static std::once_flag granted;
		std::call_once(granted, [] {
			CTG_HIP(hipFuncSetAttribute((const void*)k_fused_matvec, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_limit()));
		});
		if (f.lds_bytes > lds_limit()) throw HipError{ hipErrorInvalidValue, "fused matvec: LDS request above the device limit" };
	}
	hipLaunchKernelGGL(k_fused_matvec, dim3(f.grid), dim3(FZ_THREADS), f.lds_bytes, s, a);
	CTG_HIP(hipGetLastError());
}

//________________________________________________________________________________________________________________________
//
// batch of independent problems

BatchPlan::~BatchPlan()
{
	fused.release();
	for (int i = 0; i < 3; i++) stage[i].release();
	wapply.release();
	dev_free(t1);
	dev_free(t2);
	dev_free(w);
	dev_free(l);
	dev_free(r);
	dev_free(segs);
	dev_free(probs);
}

BatchPlan* build_batch_plan(std::vector<BatchInput>& in, hipStream_t s, bool tune, const ParallelFor* par, bool split)
{
	// the split form has one launch form: nothing to time
	if (split) tune = false;
	using clk = std::chrono::steady_clock;
	const auto secs = [](clk::time_point a, clk::time_point z) { return std::chrono::duration<double>(z - a).count(); };
	const clk::time_point c0 = clk::now();
	std::unique_ptr<BatchPlan> b(new BatchPlan());
	b->stream = s;
	b->nprob = (int)in.size();
	int64_t nx = 0, nw = 0, nl = 0, nr = 0;
	std::vector<FusedProblem> probs;
	b->off.push_back(0);
	for (BatchInput& p : in) {
		const Meta& th = p.theta;
		validate_matvec_operands(th, p.w.meta, p.l.meta, p.r.meta);
		probs.push_back(FusedProblem{ &p.theta, &p.w.meta, &p.l.meta, &p.r.meta, nx, nw, nl, nr });
		nx += th.nelem;
		nw += p.w.meta.nelem;
		nl += p.l.meta.nelem;
		nr += p.r.meta.nelem;
		b->off.push_back(nx);
		if (tune || split) b->flops += reference_flops(th, p.w.meta, p.l.meta, p.r.meta);
		b->bytes += 8.0 * (2.0 * (double)th.nelem + (double)(p.w.meta.nelem + p.l.meta.nelem + p.r.meta.nelem));
	}
	b->w = (double*)dev_alloc((size_t)std::max<int64_t>(nw, 1) * sizeof(double));
	b->l = (double*)dev_alloc((size_t)std::max<int64_t>(nl, 1) * sizeof(double));
	b->r = (double*)dev_alloc((size_t)std::max<int64_t>(nr, 1) * sizeof(double));
	for (size_t i = 0; i < in.size(); i++) {
		if (in[i].w.meta.nelem) launch_copy(in[i].w.d, b->w + probs[i].w_off, in[i].w.meta.nelem, s);
		if (in[i].l.meta.nelem) launch_copy(in[i].l.d, b->l + probs[i].l_off, in[i].l.meta.nelem, s);
		if (in[i].r.meta.nelem) launch_copy(in[i].r.d, b->r + probs[i].r_off, in[i].r.meta.nelem, s);
	}
	const clk::time_point c1 = clk::now();
	// grouped three-stage tables over the concatenated operands; without tuning only the tables of
	// the fixed forms (stage 2 as the column stream) unless the column stream does not apply
	GroupedTables g;
	auto plan_all = [&](bool lean) {
		g = GroupedTables();
		if (par && in.size() > 1) {
			// plan every problem on its own (in parallel), then concatenate
			std::vector<GroupedTables> part(in.size());
			(*par)((int)in.size(), [&](int i) {
				plan_grouped_tables(part[i], in[i].theta, in[i].w.meta, in[i].l.meta, in[i].r.meta, probs[i].x_off, probs[i].w_off,
					probs[i].l_off, probs[i].r_off, lean);
			});
			for (const GroupedTables& t : part) merge_grouped_tables(g, t);
		}
		else {
			for (size_t i = 0; i < in.size(); i++)
				plan_grouped_tables(g, in[i].theta, in[i].w.meta, in[i].l.meta, in[i].r.meta, probs[i].x_off, probs[i].w_off, probs[i].l_off,
					probs[i].r_off, lean);
		}
	};
	// problems whose lean tables were planned ahead (zero operand offsets) are only concatenated
	bool ahead = !tune;
	for (const BatchInput& p : in) ahead = ahead && p.pre && p.pre->wa.kmax <= WA_KMAX;
	if (ahead) {
		g = GroupedTables();
		for (size_t i = 0; i < in.size(); i++)
			merge_grouped_tables(g, *in[i].pre, probs[i].x_off, probs[i].w_off, probs[i].l_off, probs[i].r_off);
	}
	else {
		plan_all(!tune);
		if (!tune && g.wa.kmax > WA_KMAX) plan_all(false);
	}
	const clk::time_point c2 = clk::now();
	b->t1 = (double*)dev_alloc((size_t)std::max<int64_t>(g.t1_elems, 1) * sizeof(double));
	b->t2 = (double*)dev_alloc((size_t)std::max<int64_t>(g.t2_elems, 1) * sizeof(double));
	if (!tune && !split) b->flops = g.s1.flops + g.s2.flops + g.s3.flops;   // executed count (no replanning of the reference sequence)
	b->stage[0].upload(g.s1, s);
	b->stage[1].upload(g.s2, s);
	b->stage[2].upload(g.s3, s);
	b->stage_bases[0] = GemmBases{ nullptr, b->r, b->t1 };
	b->stage_bases[1] = GemmBases{ b->w, b->t1, b->t2 };
	b->stage_bases[2] = GemmBases{ b->l, b->t2, nullptr };
	b->wapply.upload(g.wa, b->w, s);
	if (tune) {
		autotune_grouped(b->stage, b->stage_bases, &b->wapply, &b->use_wapply, &b->use_f12, nx, s, nullptr);
		b->fused_ok = build_fused_batch(b->fused, probs, s);
	}
	else {
		// fixed launch forms per stage (the summation order of a problem does not depend on the
		// batch): stages 1 and 2 one wavefront per tile, stage 3 (long k-chains, K = dA dM) one
		// workgroup per tile
		b->stage[0].form = GEMM_WAVE_STREAM;
		b->stage[1].form = GEMM_WAVE_STREAM;
		b->stage[2].form = GEMM_WG_STREAM;
		b->use_wapply = b->wapply.ok;
		b->split = split;
	}
	const clk::time_point c3 = clk::now();
	// reduction blocks of the batched Lanczos: up to 64 blocks of >= 4096 elements per problem
	std::vector<LzSeg> segs;
	std::vector<LzProb> pr;
	for (int p = 0; p < b->nprob; p++) {
		const int64_t n = b->off[p + 1] - b->off[p];
		const int G = (int)std::max<int64_t>(1, std::min<int64_t>(64, (n + 4095) / 4096));
		LzProb q = { (int)segs.size(), (int)segs.size() + G, n };
		for (int g = 0; g < G; g++) {
			const int64_t lo = b->off[p] + n * g / G, hi = b->off[p] + n * (g + 1) / G;
			segs.push_back(LzSeg{ lo, hi, p, 0 });
		}
		pr.push_back(q);
	}
	b->nblocks = (int)segs.size();
	b->segs = upload_vec(segs, s);
	b->probs = upload_vec(pr, s);
	const clk::time_point c4 = clk::now();
	b->host_s[0] = secs(c0, c1);
	b->host_s[1] = secs(c1, c2);
	b->host_s[2] = secs(c2, c3);
	b->host_s[3] = secs(c3, c4);
	// time both forms on the batch itself and keep the faster (without tuning, the caller's
	// launches follow on the same stream: the staged uploads and copies need no wait)
	if (tune) {
		CTG_HIP(hipStreamSynchronize(s));
		double* x = (double*)dev_alloc((size_t)std::max<int64_t>(nx, 1) * sizeof(double));
		double* y = (double*)dev_alloc((size_t)std::max<int64_t>(nx, 1) * sizeof(double));
		launch_fill(x, 1.0, nx, s);
		hipEvent_t e0, e1;
		CTG_HIP(hipEventCreate(&e0));
		CTG_HIP(hipEventCreate(&e1));
		for (int f = 0; f < (b->fused_ok ? 2 : 1); f++) {
			b->use_fused = (f == 1);
			batch_apply(*b, x, y);
			CTG_HIP(hipEventRecord(e0, s));
			for (int rep = 0; rep < 3; rep++) batch_apply(*b, x, y);
			CTG_HIP(hipEventRecord(e1, s));
			CTG_HIP(hipEventSynchronize(e1));
			CTG_HIP(hipEventElapsedTime(&b->form_ms[f], e0, e1));
			b->form_ms[f] /= 3;
		}
		b->use_fused = b->fused_ok && b->form_ms[1] < b->form_ms[0];
		// CT_BATCH_FORM=grouped|fused overrides the timing-based choice
		const char* force = getenv("CT_BATCH_FORM");
		if (force && strcmp(force, "grouped") == 0) b->use_fused = false;
		if (force && strcmp(force, "fused") == 0 && b->fused_ok) b->use_fused = true;
		CTG_HIP(hipEventDestroy(e0));
		CTG_HIP(hipEventDestroy(e1));
		dev_free(x);
		dev_free(y);
	}
	return b.release();
}

void batch_apply(BatchPlan& b, const double* x, double* y)
{
	if (b.split) {
		for (int st = 0; st < 3; st++) {
			GemmBases g = b.stage_bases[st];
			if (st == 0) g.A = x;
			if (st == 2) g.C = y;
			if (st == 1 && b.use_wapply) launch_wapply(b.wapply, g.B, g.C, b.stream);
			else launch_grouped_gemm_split(b.stage[st], g, b.stream);
		}
		return;
	}
	if (b.use_fused) {
		launch_fused_matvec(b.fused, x, b.w, b.l, b.r, y, b.stream);
		return;
	}
	launch_grouped_stages(b.stage, b.stage_bases, &b.wapply, b.use_wapply, b.use_f12, x, y, b.stream, nullptr);
}

}  // namespace ctg
