// Sector-pair form of the two-site H_eff matvec: the three contractions of apply_local_hamiltonian
// (src/algorithm/chain_ops.c:353-390) in one launch, plus one launch that sums partial results.
//
// With unit outer bonds the reference's sequence is, block by block,
//   T1(A,P,N,C)[a p n c] = sum_b  theta(A,P,B)[a p b] R(B,N,C,o)[b n c 0]           qB = qA + qP
//   T2(A,M,Q,C)[a m q c] = sum_pn W(M,Q,P,N)[m q p n] T1(A,P,N,C)[a p n c]           qM + qQ = qP + qN
//   y(X,Q,C)[x q c]      = sum_am L(o,A,M,X)[0 a m x] T2(A,M,Q,C)[a m q c]           qX = qA + qM - qo
// and every intermediate block belongs to exactly one block pair (A, C) of the left and right bond
// sectors, whose class D = qC - qA + qo fixes the (P,N) "routes" and (M,Q) "rows" involved. A pair,
// or a chunk of its rows a and columns c, is staged in LDS (its theta rows, R columns, L rows and the
// class's dense W_D); a workgroup takes one or several such chunks (small ones packed together up to
// the LDS budget, so that every unit is resident at once), loads all of their data at once, computes
// T1, then T2 = W_D T1, then L^T T2, all on v_mfma_f64_16x16x4f64 with operands from LDS, and writes
// each chunk's contribution to each y(X,Q,C) as a partial block. A second launch sums the partial blocks of every y block in a fixed order (by A,
// then row chunk), so the result does not depend on scheduling. Nothing of T1 / T2 reaches HBM.
//
// The host resolves the block structure into per-unit tables (staging pieces, MFMA tiles with their
// LDS bases and valid extents, the stage-2 row map), so the kernel's per-lane work is address
// arithmetic on uniform descriptors: divisions by small runtime integers use precomputed
// multipliers, the contracted dimension of every staged panel is zero-padded to a multiple of 4,
// lanes outside a tile's valid rows or columns read a clamped in-bounds address (their outputs are
// never stored), and absent T1 rows read an LDS region of zeros.

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include "ct_internal.h"

namespace ctg {

typedef double pr_double4 __attribute__((ext_vector_type(4)));

// Unit record (4-byte words, fixed stride per plan). A unit holds one or more chunks of block pairs
// (small chunks are packed together up to the LDS budget), each in its own LDS region above the
// shared zero / dump area; every LDS offset below is absolute.
//   header (16): n1 n3 npiece n2 | 0 0 0 0 | 0 0 off1 off3 | off2 offp 0 0
//   (offp = 16: the pieces follow the header, so their descriptors load in the header's round trip)
//   stage-1 tiles at off1 (8 words each): abase sth bbase Nr nk vr vc sbase
//   stage-2 tiles at off2 (8 words each): wbase wld|nk4<<16 kb tc | mag(tc) S1 j0 t2out
//     (W_D rows of the tile from LDS wbase at stride wld, nk4 k-steps of four, the chunk's row map at
//     kb, columns j0 .. j0+15 of the chunk's S1 = ta tc columns (a c), output rows at t2out, stride S1)
//   stage-3 tiles at off3 (16 words each): lbase dX t2b dQ*S1 | dM ta|nka<<16 ydelta vr|vc<<16 |
//                                          S1 j0 tc dC | mag(tc) pbase(2) dQ*dC
//     (ydelta: the y index of a partial entry minus its index in the partial blocks)
//   row maps after the tiles (even offsets; 2 words per T1 row k < K4 of a chunk): LDS offset of row k
//                                          at a = 0, c = 0, and its stride in a (0, 0 for absent rows:
//                                          the zero region), staged by the chunk's kind-5 piece
//   pieces at offp (4 words each), 64 consecutive entries e = e0 .. e0+63 of a staged panel whose
//   entry (row, col) = divmod(e, cols) goes to LDS d0 + row dstride + col from source
//   s0 + row sstride + col (or is zero): with (row0, col0) = divmod(e0, cols),
//     w0 = s0 + row0 sstride    w1 = (d0 + row0 dstride) | sstride << 16
//     w2 = dstride | cols << 8 | col0 << 16 | lanes << 24
//     w3 = mag16(cols) | vcols << 17 | kind << 25
//   lane l < lanes: t = col0 + l, row - row0 = t / cols (= t mag16 >> 16), col = t - (row - row0) cols;
//   col < vcols loads, else writes 0. kind: 0 theta (x), 1 R, 2 L, 3 zeros, 4 the W_D table, 5 the
//   unit's own record read as doubles (the row map).
// mag(d) = ceil(2^31 / d): t / d = mulhi(2 t, mag(d)) for 0 <= t < 2^31 / d; mag16(d) = ceil(2^16 / d):
// t / d = (t mag16(d)) >> 16 for t (t + 1) < 2^16 ... exact whenever t cols < 2^16.
constexpr int PR_HDR = 16, PR_T1 = 8, PR_T2 = 8, PR_T3 = 16, PR_PIECE = 4;
// words after the last record: the kernel's piece-slot loads (NB <= 64 slots of 4 words from PR_HDR)
constexpr size_t PR_REC_PAD = PR_HDR + PR_PIECE * 4 * 64;
constexpr int PR_ZERO = 64;   // LDS doubles [0, 64) hold zeros
constexpr int PR_DUMP = 64;   // LDS double taking the stores of lanes outside a tile (never read)
constexpr int PR_WB = 80;     // LDS offset of the W_D copy

__device__ __forceinline__ int pr_q(int t, uint32_t mag) { return (int)__umulhi((uint32_t)t << 1, mag); }
__device__ __forceinline__ int pr_s(uint32_t v) { return __builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ int pr_lane(uint32_t v, int l) { return __builtin_amdgcn_readlane((int)v, l); }
// identity DPP moves give a loaded descriptor registers of their own: the readlanes of it then wait for
// that load only, not for loads issued later in branches the wait counter cannot count
__device__ __forceinline__ void pr_fresh(uint4& d)
{
	d.x = (uint32_t)__builtin_amdgcn_mov_dpp((int)d.x, 0xE4, 0xF, 0xF, false);
	d.y = (uint32_t)__builtin_amdgcn_mov_dpp((int)d.y, 0xE4, 0xF, 0xF, false);
	d.z = (uint32_t)__builtin_amdgcn_mov_dpp((int)d.z, 0xE4, 0xF, 0xF, false);
	d.w = (uint32_t)__builtin_amdgcn_mov_dpp((int)d.w, 0xE4, 0xF, 0xF, false);
}

template <int LDS>
__global__ __launch_bounds__(PR_THREADS) void k_pair_units(const uint32_t* __restrict__ recs, const int rec_words, const double* __restrict__ x,
	const double* __restrict__ rr, const double* __restrict__ ll, const double* __restrict__ wd, double* __restrict__ part, double* __restrict__ dump,
	double* __restrict__ udot)
{
	constexpr int NB = LDS / PR_THREADS;   // staging pieces per wavefront
	static_assert(NB <= 64, "one piece descriptor per lane");
	__shared__ double sm[LDS];
	const int tid = threadIdx.x, lane = tid & 63;
	const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	const int i16 = lane & 15, kq = lane >> 4;
	const uint32_t* __restrict__ g = recs + (size_t)blockIdx.x * (size_t)rec_words;

	// lane j of wavefront w holds the descriptor of piece w + 4 j: loaded at the fixed offset PR_HDR
	// before the header arrives (the record table is padded so that the NB slots of the last record
	// stay in bounds; lanes past NB repeat slot NB - 1, slots past npiece are never used)
	uint4 pd = reinterpret_cast<const uint4*>(g + PR_HDR)[wave + 4 * min(lane, NB - 1)];
	__builtin_amdgcn_sched_barrier(0);
	const int n1 = pr_s(g[0]), n3 = pr_s(g[1]), npiece = pr_s(g[2]), n2 = pr_s(g[3]);
	const int off1 = pr_s(g[10]), off3 = pr_s(g[11]), off2 = pr_s(g[12]);

	// the wavefront's stage-1, stage-2 and stage-3 tiles w + 4 j, one per lane; lanes past a table
	// read the header instead (16 words, in bounds), so every lane issues every load: the wait for
	// pd before the staging then counts exactly the eight loads issued after it, and these
	// descriptors arrive while the staging loads are in flight
	uint4 d1a, d1b, d2a, d2b, d3a, d3b, d3c, d3d;
	{
		const int tj = wave + 4 * lane;
		const uint4* d = reinterpret_cast<const uint4*>(g + (tj < n1 ? off1 + PR_T1 * tj : 0));
		d1a = d[0];
		d1b = d[1];
		d = reinterpret_cast<const uint4*>(g + (tj < n2 ? off2 + PR_T2 * tj : 0));
		d2a = d[0];
		d2b = d[1];
		d = reinterpret_cast<const uint4*>(g + (tj < n3 ? off3 + PR_T3 * tj : 0));
		d3a = d[0];
		d3b = d[1];
		d3c = d[2];
		d3d = d[3];
	}
	__builtin_amdgcn_sched_barrier(0);
	pr_fresh(pd);

	// staging: every load of the wavefront's pieces is issued before the first LDS store
	if (tid < PR_ZERO) sm[tid] = 0.0;
	{
		double v[NB];
		int dst[NB];
		#pragma unroll
		for (int i = 0; i < NB; i++) {
			dst[i] = PR_DUMP;
			v[i] = 0.0;
			if (wave + 4 * i < npiece) {
				const uint32_t w0 = (uint32_t)__builtin_amdgcn_readlane((int)pd.x, i), w1 = (uint32_t)__builtin_amdgcn_readlane((int)pd.y, i);
				const uint32_t w2 = (uint32_t)__builtin_amdgcn_readlane((int)pd.z, i), w3 = (uint32_t)__builtin_amdgcn_readlane((int)pd.w, i);
				if (lane < (int)(w2 >> 24)) {
					const int cols = (int)((w2 >> 8) & 0xffu);
					const int t = (int)((w2 >> 16) & 0xffu) + lane;
					const int q = __mul24(t, (int)(w3 & 0x1ffffu)) >> 16;
					const int c = t - __mul24(q, cols);
					dst[i] = (int)(w1 & 0xffffu) + __mul24(q, (int)(w2 & 0xffu)) + c;
					const int kind = (int)(w3 >> 25);
					if (kind != 3 && c < (int)((w3 >> 17) & 0xffu)) {
						const double* __restrict__ base =
							kind == 0 ? x : (kind == 1 ? rr : (kind == 2 ? ll : (kind == 4 ? wd : reinterpret_cast<const double*>(g))));
						v[i] = base[w0 + (uint32_t)(__mul24(q, (int)(w1 >> 16)) + c)];
					}
				}
			}
		}
		#pragma unroll
		for (int i = 0; i < NB; i++)
			if (wave + 4 * i < npiece) sm[dst[i]] = v[i];
	}
	pr_fresh(d1a);
	pr_fresh(d1b);
	pr_fresh(d2a);
	pr_fresh(d2b);
	pr_fresh(d3a);
	pr_fresh(d3b);
	pr_fresh(d3c);
	pr_fresh(d3d);
	__syncthreads();

	// stage 1: T1 of each route in its own layout [(a p)][(n c)] = theta[(a p), b] R[b, (n c)]
	for (int t = wave, tl = 0; t < n1; t += 4, tl++) {
		const int abase = pr_lane(d1a.x, tl), sth = pr_lane(d1a.y, tl), bbase = pr_lane(d1a.z, tl), Nr = pr_lane(d1a.w, tl);
		const int nk = pr_lane(d1b.x, tl), vr = pr_lane(d1b.y, tl), vc = pr_lane(d1b.z, tl), sbase = pr_lane(d1b.w, tl);
		const int ar = (i16 < vr ? abase + __mul24(i16, sth) : abase) + kq;
		const int bc = (i16 < vc ? bbase + i16 : bbase) + __mul24(kq, Nr);
		// the operands of k-step s + 1 are read before the MFMA of step s is issued
		pr_double4 acc = { 0, 0, 0, 0 };
		if (nk > 0) {
			const int bstep = 4 * Nr;
			int ao = ar, bo = bc;
			double av = sm[ao], bv = sm[bo];
			for (int s = 1; s < nk; s++) {
				ao += 4;
				bo += bstep;
				const double an = sm[ao], bn = sm[bo];
				acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
				av = an;
				bv = bn;
			}
			acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
		}
		#pragma unroll
		for (int r4 = 0; r4 < 4; r4++) {
			const int i = kq + 4 * r4;
			sm[(i16 < vc && i < vr) ? sbase + __mul24(i, Nr) + i16 : PR_DUMP] = acc[r4];
		}
	}
	__syncthreads();

	// stage 2: T2[(M Q m q)][(a c)] = W_D T1 per tile of 16 rows (M Q m q) and 16 columns (a c) of a
	// pair: W_D row i, column k at LDS wbase + i wld + k; T1 row k of column (a, c) at LDS
	// kofs(k) + a kstr(k) + c from the pair's row map (absent rows: 0, 0, the zero region)
	for (int t = wave, tl = 0; t < n2; t += 4, tl++) {
		const int wbase = pr_lane(d2a.x, tl), wq = pr_lane(d2a.y, tl), kbb = pr_lane(d2a.z, tl), tc2 = pr_lane(d2a.w, tl);
		const uint32_t mag2 = (uint32_t)pr_lane(d2b.x, tl);
		const int S1 = pr_lane(d2b.y, tl), j0 = pr_lane(d2b.z, tl), t2o = pr_lane(d2b.w, tl);
		const int wld = wq & 0xffff, nk4 = wq >> 16;
		const int j2 = j0 + i16;
		const bool vj = j2 < S1;
		const int a = vj ? pr_q(j2, mag2) : 0, c = vj ? j2 - a * tc2 : 0;
		const int2* __restrict__ kmap = reinterpret_cast<const int2*>(sm + kbb) + kq;
		const int wrow = wbase + __mul24(i16, wld) + kq;
		// software pipelined: the row-map entry of k-step ks + 2 and the operands of step ks + 1 are
		// read before the MFMA of step ks is issued (the entry index clamped to the last step)
		pr_double4 acc = { 0, 0, 0, 0 };
		if (nk4 > 0) {
			const int2 e0 = kmap[0];
			double wv = sm[wrow], tv = sm[e0.x + __mul24(a, e0.y) + c];
			int2 en = kmap[4 * min(1, nk4 - 1)];
			for (int ks = 1; ks < nk4; ks++) {
				const int2 e2 = kmap[4 * min(ks + 1, nk4 - 1)];
				const double wn = sm[wrow + 4 * ks], tn = sm[en.x + __mul24(a, en.y) + c];
				acc = __builtin_amdgcn_mfma_f64_16x16x4f64(wv, tv, acc, 0, 0, 0);
				wv = wn;
				tv = tn;
				en = e2;
			}
			acc = __builtin_amdgcn_mfma_f64_16x16x4f64(wv, tv, acc, 0, 0, 0);
		}
		#pragma unroll
		for (int r4 = 0; r4 < 4; r4++) sm[vj ? t2o + __mul24(kq + 4 * r4, S1) + j2 : PR_DUMP] = acc[r4];
	}
	__syncthreads();

	// stage 3: partial y(X,Q,C)[x, (q c)] = L(o,A,M,X)[(a m), x]^T T2(A,M,Q,C)[(a m), (q c)]
	double dsum = 0.0;
	for (int t = wave, tl = 0; t < n3; t += 4, tl++) {
		const int lbase = pr_lane(d3a.x, tl), dX = pr_lane(d3a.y, tl), t2b = pr_lane(d3a.z, tl), dQS1 = pr_lane(d3a.w, tl);
		const int dM = pr_lane(d3b.x, tl), tanka = pr_lane(d3b.y, tl), ydelta = pr_lane(d3b.z, tl), vrc = pr_lane(d3b.w, tl);
		const int ta = tanka & 0xffff, nka = tanka >> 16;
		const int S1 = pr_lane(d3c.x, tl), j0 = pr_lane(d3c.y, tl), tc3 = pr_lane(d3c.z, tl), dC = pr_lane(d3c.w, tl);
		const int vr = vrc & 0xffff, vc = vrc >> 16;
		const uint32_t mag3 = (uint32_t)pr_lane(d3d.x, tl);
		const int64_t pbase = (int64_t)((uint64_t)(uint32_t)pr_lane(d3d.y, tl) | ((uint64_t)(uint32_t)pr_lane(d3d.z, tl) << 32));
		const int dQdC = pr_lane(d3d.w, tl);
		const int ax = (i16 < vr ? lbase + i16 : lbase);
		const bool vj = i16 < vc;
		const int j3 = j0 + i16;
		const int q = vj ? pr_q(j3, mag3) : 0, c = vj ? j3 - q * tc3 : 0;
		const int jofs = t2b + __mul24(q, S1) + c;
		// k runs over (m, a): L row (a dM + m) of the staged [a m x] panel, T2 row (m dQ + q) column
		// a tc + c; lanes past the chunk's rows a read the zero region
		// the k-steps (m, s) in one loop, m outer, the operands of the next step read before the MFMA
		// of the current one is issued
		pr_double4 acc = { 0, 0, 0, 0 };
		const int lstep = __mul24(4 * dM, dX), bstep = 4 * tc3;
		const int nsteps = __mul24(dM, nka);
		if (nsteps > 0) {
			int s = 0, lam = ax + __mul24(kq * dM, dX), lbm = jofs + __mul24(kq, tc3), la = lam, lb = lbm;
			double av = sm[kq < ta ? la : 0], bv = sm[kq < ta ? lb : 0];
			for (int i = 1; i < nsteps; i++) {
				s++;
				la += lstep;
				lb += bstep;
				if (s == nka) {
					s = 0;
					lam += dX;
					lbm += dQS1;
					la = lam;
					lb = lbm;
				}
				const bool vk = 4 * s + kq < ta;
				const double an = sm[vk ? la : 0], bn = sm[vk ? lb : 0];
				acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
				av = an;
				bv = bn;
			}
			acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
		}
		if (vj) {
			const int64_t o0 = pbase + (uint32_t)(__mul24(q, dC) + c);
			double* __restrict__ out = part + o0;
			#pragma unroll
			for (int r4 = 0; r4 < 4; r4++)
				if (kq + 4 * r4 < vr) {
					out[(uint32_t)__mul24(kq + 4 * r4, dQdC)] = acc[r4];
					if (udot) dsum += acc[r4] * x[o0 + ydelta + (uint32_t)__mul24(kq + 4 * r4, dQdC)];
				}
		}
	}
	// the unit's share of <x, H x>: its partial entries times x at their y indices, summed over the
	// workgroup in a fixed order
	if (udot) {
		__shared__ double dsh[4];
		#pragma unroll
		for (int off = 32; off > 0; off >>= 1) dsum += __shfl_xor(dsum, off, 64);
		if (lane == 0) dsh[wave] = dsum;
		__syncthreads();
		if (tid == 0) udot[blockIdx.x] = (dsh[0] + dsh[1]) + (dsh[2] + dsh[3]);
	}
}

// y block chunk of up to 64 entries: y[y + e] = sum_s part[p + s * size + e], s = 0 .. nslots-1,
// summed in slot order; the loads of up to PR_RED_BATCH slots are in flight at once (slot indices
// clamped to the last slot, whose repeated values are not added)
constexpr int PR_RED_BATCH = 16;
__device__ __forceinline__ double pr_slot_sum(const double* __restrict__ src, int64_t sz, int nslots)
{
	double acc = 0.0;
	for (int s0 = 0; s0 < nslots; s0 += PR_RED_BATCH) {
		double v[PR_RED_BATCH];
		#pragma unroll
		for (int u = 0; u < PR_RED_BATCH; u++) v[u] = src[(int64_t)min(s0 + u, nslots - 1) * sz];
		#pragma unroll
		for (int u = 0; u < PR_RED_BATCH; u++)
			if (s0 + u < nslots) acc += v[u];
	}
	return acc;
}

__global__ __launch_bounds__(256) void k_pair_reduce(const PrRed* __restrict__ items, const int nitems, const double* __restrict__ part, double* __restrict__ y)
{
	const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
	const int lane = threadIdx.x & 63;
	if (w >= nitems) return;
	const PrRed it = items[w];
	if (lane >= it.n) return;
	y[it.y + lane] = pr_slot_sum(part + it.p + lane, it.size, it.nslots);
}

// The same sums, and dpart[block] = sum over the block's entries of y[e] v[e] (the Lanczos alpha
// partials, v the vector the matvec was applied to): wavefront butterflies, then the four
// wavefront sums in order, so every launch sums in the same order
__global__ __launch_bounds__(256) void k_pair_reduce_dot(const PrRed* __restrict__ items, const int nitems, const double* __restrict__ part,
	double* __restrict__ y, const double* __restrict__ v, double* __restrict__ dpart)
{
	__shared__ double sh[4];
	const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
	const int lane = threadIdx.x & 63;
	double d = 0.0;
	if (w < nitems) {
		const PrRed it = items[w];
		if (lane < it.n) {
			const double acc = pr_slot_sum(part + it.p + lane, it.size, it.nslots);
			y[it.y + lane] = acc;
			d = acc * v[it.y + lane];
		}
	}
	#pragma unroll
	for (int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off, 64);
	if (lane == 0) sh[threadIdx.x >> 6] = d;
	__syncthreads();
	if (threadIdx.x == 0) dpart[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// One step of the deferred-normalisation Lanczos recurrence on a sector-pair plan (the vector part
// of the recurrence's step, as the generic step kernel, with y read from the partial blocks): the
// unit launch wrote the partial blocks of y = H u_j and the per-unit partials of <u_j, y>; here
// beta_{j-1} = sqrt(sum npart) and alpha_j = sum dpart / beta^2 (both sums in a fixed order, the
// same in every workgroup), and per element of one reduction item: y = the sum of its slots in slot
// order, v_j = u_j / beta, u_{j+1} = y / beta - alpha v_j - beta v_{j-1}, and the partials of
// |u_{j+1}|^2 per workgroup. The final step (j = maxiter - 1) writes alpha and v_j only.
__global__ __launch_bounds__(256) void k_pair_lz_step(const PrRed* __restrict__ items, const int nitems, const double* __restrict__ part,
	const double* __restrict__ uj, double* __restrict__ unext, double* __restrict__ vj_out, const double* __restrict__ vjm1, int j, int maxiter,
	const double* __restrict__ dpart, int Gd, const double* __restrict__ npart, int Gn, double thresh, double* __restrict__ npart_out,
	double* __restrict__ scal)
{
	__shared__ double sh[8];
	if (scal[2 * maxiter + 1] != 0) return;
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + wave));
	const bool upd = j < maxiter - 1;
	int64_t i = -1;
	double yv = 0.0, uv = 0.0, bv = 0.0;
	if (w < nitems) {
		const PrRed it = items[w];
		if (lane < it.n) {
			i = it.y + lane;
			uv = uj[i];
			if (upd) {
				yv = pr_slot_sum(part + it.p + lane, it.size, it.nslots);
				if (j > 0) bv = vjm1[i];
			}
		}
	}
	// beta^2 and <u_j, y>: per-thread strided sums, wavefront butterflies, then the four wavefronts
	// (thread tid sums entries tid + 256 k in k order, four loads in flight at a time; entries past
	// the ends add zeros, which leave the sums unchanged)
	double a = 0.0, b = 0.0;
	const int Ga = j > 0 ? Gn : 0;
	for (int k0 = tid; k0 < Ga || k0 < Gd; k0 += 1024) {
		double va[4], vb[4];
		#pragma unroll
		for (int u = 0; u < 4; u++) {
			const int k = k0 + 256 * u;
			va[u] = k < Ga ? npart[k] : 0.0;
			vb[u] = k < Gd ? dpart[k] : 0.0;
		}
		#pragma unroll
		for (int u = 0; u < 4; u++) {
			a += va[u];
			b += vb[u];
		}
	}
	#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		a += __shfl_xor(a, off, 64);
		b += __shfl_xor(b, off, 64);
	}
	if (lane == 0) {
		sh[wave] = a;
		sh[4 + wave] = b;
	}
	__syncthreads();
	const double nsum = (sh[0] + sh[1]) + (sh[2] + sh[3]);
	const double dsum = (sh[4] + sh[5]) + (sh[6] + sh[7]);
	__syncthreads();
	double beta = 1.0;
	if (j > 0) {
		beta = sqrt(nsum);
		// every workgroup has read the stop flag before workgroup 0 may set it (read at entry only)
		if (beta < thresh) {
			if (blockIdx.x == 0 && tid == 0) { scal[maxiter + j - 1] = beta; scal[2 * maxiter] = j; scal[2 * maxiter + 1] = 1; }
			return;
		}
	}
	const double inv = 1 / beta;
	const double alpha = dsum * inv * inv;
	if (blockIdx.x == 0 && tid == 0) {
		scal[j] = alpha;
		if (j > 0) scal[maxiter + j - 1] = beta;
	}
	double acc = 0.0;
	if (i >= 0) {
		const double v = inv * uv;
		if (j > 0) vj_out[i] = v;
		if (upd) {
			const double wi = (j > 0) ? inv * yv - (alpha * v + beta * bv) : yv - alpha * v;
			unext[i] = wi;
			acc = wi * wi;
		}
	}
	if (upd) {
		#pragma unroll
		for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
		if (lane == 0) sh[wave] = acc;
		__syncthreads();
		if (tid == 0) npart_out[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
	}
}

//________________________________________________________________________________________________________________________
//
// host planning

namespace {

struct PrClass
{
	std::vector<std::pair<ct_long, ct_long>> rows, cols;   // (M, Q) and (P, N) blocks
	std::vector<int> rbase, kbase;                          // first W_D row / column of each
	int Rd = 0, Kd = 0, R16 = 0, K4 = 0;
	int64_t wd = 0;                                          // offset of W_D in the table
	bool ok = false;
};

// a stage-3 tile's partial offset: patched with its slot once all slots are known
struct PrSlotRef { int64_t yblock; ct_long A; int ia; size_t unit; int word; };

// one chunk of a block pair, its LDS offsets laid out from PR_WB (moved up when packed into a unit)
struct PrPart
{
	int need = 0;                                  // LDS end offset (doubles)
	int np = 0, n1 = 0, n2 = 0, n3 = 0, mfma = 0;
	size_t kpiece = 0;                             // word of pcs holding the row-map piece's source
	std::vector<uint32_t> pcs, t1s, t2s, t3s, kt;
	std::vector<PrSlotRef> yref;                   // per stage-3 tile: its y block, A and row chunk
	int32_t A = 0;                                 // the pair's left sector
};

inline int cdiv(int a, int b) { return (a + b - 1) / b; }
inline int up4(int a) { return (a + 3) / 4 * 4; }
inline uint32_t magic(int d) { return (uint32_t)((((uint64_t)1 << 31) + (uint64_t)d - 1) / (uint64_t)d); }
inline int qdiv(int t, uint32_t mag) { return (int)(((uint64_t)((uint32_t)t << 1) * mag) >> 32); }

}  // namespace

bool plan_pair_tables(PairHost& h, const Meta& th, const Meta& W, const Meta& L, const Meta& R, int64_t x_off, int64_t w_off, int64_t l_off, int64_t r_off)
{
	const qnumber qo = L.qnums_blocks[0][0];
	const ct_long nA = th.dim_blocks[0], nC = R.dim_blocks[2];
	const ct_long nM = W.dim_blocks[0], nQ = W.dim_blocks[1], nP = W.dim_blocks[2], nN = W.dim_blocks[3];
	auto bc = [](const Meta& m, int ax, ct_long i) { return (int)m.block_count[ax][i]; };
	auto qb = [](const Meta& m, int ax, ct_long i) { return m.qnums_blocks[ax][i]; };

	// classes D = qM + qQ = qP + qN: rows (M,Q), columns (P,N), dense W_D (R16 x K4) as W indices
	std::map<qnumber, PrClass> classes;
	auto get_class = [&](qnumber D) -> PrClass& {
		auto it = classes.find(D);
		if (it != classes.end()) return it->second;
		PrClass& c = classes[D];
		for (ct_long M = 0; M < nM; M++) for (ct_long Q = 0; Q < nQ; Q++)
			if (qb(W, 0, M) + qb(W, 1, Q) == D) { c.rows.emplace_back(M, Q); c.rbase.push_back(c.Rd); c.Rd += bc(W, 0, M) * bc(W, 1, Q); }
		for (ct_long P = 0; P < nP; P++) for (ct_long N = 0; N < nN; N++)
			if (qb(W, 2, P) + qb(W, 3, N) == D) { c.cols.emplace_back(P, N); c.kbase.push_back(c.Kd); c.Kd += bc(W, 2, P) * bc(W, 3, N); }
		c.ok = c.Rd > 0 && c.Kd > 0 && c.Rd <= PR_RMAX && c.Kd <= PR_KMAX;
		if (!c.ok) return c;
		c.R16 = (c.Rd + 15) / 16 * 16;
		c.K4 = up4(c.Kd);
		c.wd = (int64_t)h.wd_idx.size();
		h.wd_idx.resize(h.wd_idx.size() + (size_t)c.R16 * (c.K4 + 1), -1);
		for (size_t ir = 0; ir < c.rows.size(); ir++) {
			const ct_long M = c.rows[ir].first, Q = c.rows[ir].second;
			const int dM = bc(W, 0, M), dQ = bc(W, 1, Q);
			for (size_t ic = 0; ic < c.cols.size(); ic++) {
				const ct_long P = c.cols[ic].first, N = c.cols[ic].second;
				const int dP = bc(W, 2, P), dN = bc(W, 3, N);
				const ct_long iw[4] = { M, Q, P, N };
				const ct_long kw = W.flat_index(iw);
				if (!W.present(kw)) continue;
				for (int m = 0; m < dM; m++) for (int q = 0; q < dQ; q++)
					for (int pp = 0; pp < dP; pp++) for (int n = 0; n < dN; n++) {
						const int row = c.rbase[ir] + m * dQ + q, col = c.kbase[ic] + pp * dN + n;
						h.wd_idx[(size_t)(c.wd + (int64_t)row * (c.K4 + 1) + col)] = w_off + W.offset[kw] + ((ct_long)(m * dQ + q) * dP + pp) * dN + n;
					}
			}
		}
		return c;
	};

	struct VRoute { int64_t th, r; int dP, dB, dN, kbase; };
	struct VRow { int64_t l; int64_t yblock; int dM, dQ, dX, rbase; };
	std::vector<PrSlotRef> refs;
	std::vector<PrPart> parts;
	const int nbp = 4 * (h.lds / PR_THREADS);   // pieces per unit

	std::vector<VRoute> vr;
	std::vector<VRow> vq;
	for (ct_long A = 0; A < nA; A++) {
		const int dA = bc(th, 0, A);
		for (ct_long C = 0; C < nC; C++) {
			const int dC = bc(R, 2, C);
			const qnumber D = qb(R, 2, C) - qb(th, 0, A) + qo;
			PrClass& cl = get_class(D);
			if (cl.rows.empty() || cl.cols.empty()) continue;
			if (!cl.ok) return false;
			vr.clear();
			vq.clear();
			for (size_t ic = 0; ic < cl.cols.size(); ic++) {
				const ct_long P = cl.cols[ic].first, N = cl.cols[ic].second;
				const ct_long B = th.find_qnum(2, qb(th, 0, A) + qb(th, 1, P));
				if (B < 0) continue;
				const ct_long ith[3] = { A, P, B }, irr[4] = { B, N, C, 0 };
				const ct_long kt = th.flat_index(ith), kr = R.flat_index(irr);
				if (!th.present(kt) || !R.present(kr)) continue;
				vr.push_back(VRoute{ x_off + th.offset[kt], r_off + R.offset[kr], bc(th, 1, P), bc(th, 2, B), bc(R, 1, N), cl.kbase[ic] });
			}
			for (size_t ir = 0; ir < cl.rows.size(); ir++) {
				const ct_long M = cl.rows[ir].first, Q = cl.rows[ir].second;
				const ct_long X = L.find_qnum(3, qb(th, 0, A) + qb(W, 0, M) - qo);
				if (X < 0) continue;
				const ct_long il[4] = { 0, A, M, X }, iy[3] = { X, Q, C };
				const ct_long kl = L.flat_index(il), ky = th.flat_index(iy);
				if (!L.present(kl) || !th.present(ky)) continue;
				vq.push_back(VRow{ l_off + L.offset[kl], (int64_t)ky, bc(W, 0, M), bc(W, 1, Q), bc(L, 3, X), cl.rbase[ir] });
			}
			if (vr.empty() || vq.empty()) continue;
			// LDS need of a chunk (ta rows a, tc columns c) and its number of staging pieces
			auto lds_need = [&](int ta, int tc) {
				int l = PR_WB + cl.R16 * (cl.K4 + 1) + cl.K4, t1 = 0, st = 0;
				for (auto& v : vq) l += ta * v.dM * v.dX;
				for (auto& v : vr) { t1 += ta * v.dP * v.dN * tc; st += ta * v.dP * (up4(v.dB) + 1) + up4(v.dB) * v.dN * tc; }
				return l + t1 + std::max(cl.R16 * ta * tc, st);
			};
			auto npieces = [&](int ta, int tc) {
				int n = cdiv(cl.R16 * (cl.K4 + 1), 64) + cdiv(cl.K4, 64);
				for (auto& v : vr) n += cdiv(ta * v.dP * up4(v.dB), 64) + cdiv(v.dB * v.dN * tc, 64) + cdiv((up4(v.dB) - v.dB) * v.dN * tc, 64);
				for (auto& v : vq) n += cdiv(ta * v.dM * v.dX, 64);
				return n;
			};
			// tiles of stages 1, 2 and 3 (one descriptor per lane of a wavefront: at most 4 x 64 each)
			auto tiles = [&](int ta, int tc) {
				int n1 = 0, n3 = 0;
				for (auto& v : vr) n1 += cdiv(ta * v.dP, 16) * cdiv(v.dN * tc, 16);
				for (auto& v : vq) n3 += cdiv(v.dX, 16) * cdiv(v.dQ * tc, 16);
				return std::max(std::max(n1, n3), (cl.R16 / 16) * cdiv(ta * tc, 16));
			};
			auto mfmas = [&](int ta, int tc) {
				int n = 0;
				for (auto& v : vr) n += cdiv(ta * v.dP, 16) * cdiv(v.dN * tc, 16) * (up4(v.dB) / 4);
				n += (cl.R16 / 16) * cdiv(ta * tc, 16) * (cl.K4 / 4);
				for (auto& v : vq) n += cdiv(v.dX, 16) * cdiv(v.dQ * tc, 16) * v.dM * cdiv(ta, 4);
				return n;
			};
			int ta = dA, tc = dC;
			for (auto& v : vr) if (up4(v.dB) > 223 || v.dB > 0xffff || dC > 0xffff) return false;
			while (lds_need(ta, tc) > h.lds || npieces(ta, tc) > nbp || mfmas(ta, tc) > h.mfma_cap || tc > PR_ZERO || tiles(ta, tc) > 4 * 64) {
				if (tc > 4 && (tc >= ta || ta <= 4)) tc = (tc + 1) / 2;
				else if (ta > 1) ta = (ta + 1) / 2;
				else if (tc > 1) tc = (tc + 1) / 2;
				else if (lds_need(ta, tc) > h.lds || npieces(ta, tc) > nbp || tiles(ta, tc) > 4 * 64) return false;
				else break;
			}
			for (int a0 = 0, ia = 0; a0 < dA; a0 += ta, ia++) {
				const int tA = std::min(ta, dA - a0);
				for (int c0 = 0; c0 < dC; c0 += tc) {
					const int tC = std::min(tc, dC - c0);
					const int S1 = tA * tC;
					// LDS layout: zeros, dump | W_D | row map | L rows | T1 of each route | T2 overlaid on
					// the theta and R panels
					const int lwb = PR_WB, lkb = lwb + cl.R16 * (cl.K4 + 1);
					int off = lkb + cl.K4;
					std::vector<int> lds_l(vq.size()), t1b(vr.size()), lth(vr.size()), lr(vr.size());
					for (size_t i = 0; i < vq.size(); i++) { lds_l[i] = off; off += tA * vq[i].dM * vq[i].dX; }
					for (size_t i = 0; i < vr.size(); i++) { t1b[i] = off; off += tA * vr[i].dP * vr[i].dN * tC; }
					const int lt2 = off;
					for (size_t i = 0; i < vr.size(); i++) {
						lth[i] = off; off += tA * vr[i].dP * (up4(vr[i].dB) + 1);
						lr[i] = off; off += up4(vr[i].dB) * vr[i].dN * tC;
					}
					const int need = std::max(off, lt2 + cl.R16 * S1);
					if (need > h.lds || need > 0xffff) return false;
					PrPart pt;
					pt.A = (int32_t)A;
					pt.need = need;
					std::vector<uint32_t>& t1s = pt.t1s;
					std::vector<uint32_t>& t2s = pt.t2s;
					std::vector<uint32_t>& t3s = pt.t3s;
					std::vector<uint32_t>& kt = pt.kt;
					std::vector<uint32_t>& pcs = pt.pcs;
					kt.assign((size_t)cl.K4 * 2, 0u);
					// stage-1 tiles
					for (size_t i = 0; i < vr.size(); i++) {
						const int Mr = tA * vr[i].dP, Nr = vr[i].dN * tC, sth = up4(vr[i].dB) + 1;
						for (int mi = 0; mi < cdiv(Mr, 16); mi++)
							for (int ni = 0; ni < cdiv(Nr, 16); ni++) {
								const uint32_t d[PR_T1] = { (uint32_t)(lth[i] + 16 * mi * sth), (uint32_t)sth, (uint32_t)(lr[i] + 16 * ni), (uint32_t)Nr,
									(uint32_t)(up4(vr[i].dB) / 4), (uint32_t)std::min(16, Mr - 16 * mi), (uint32_t)std::min(16, Nr - 16 * ni),
									(uint32_t)(t1b[i] + 16 * mi * Nr + 16 * ni) };
								t1s.insert(t1s.end(), d, d + PR_T1);
							}
					}
					// stage-2 tiles: 16 rows of W_D by 16 columns (a c) of the chunk
					for (int mi = 0; mi < cl.R16 / 16; mi++)
						for (int ni = 0; ni < cdiv(S1, 16); ni++) {
							const uint32_t d[PR_T2] = { (uint32_t)(lwb + 16 * mi * (cl.K4 + 1)), (uint32_t)(cl.K4 + 1) | ((uint32_t)(cl.K4 / 4) << 16),
								(uint32_t)lkb, (uint32_t)tC, magic(tC), (uint32_t)S1, (uint32_t)(16 * ni), (uint32_t)(lt2 + 16 * mi * S1) };
							t2s.insert(t2s.end(), d, d + PR_T2);
						}
					// stage-2 row map: T1 row k = kbase + p dN + n of a route -> the route's T1 at (a dP + p, n tc + c)
					for (size_t i = 0; i < vr.size(); i++) {
						const int Nr = vr[i].dN * tC;
						for (int p = 0; p < vr[i].dP; p++) for (int n = 0; n < vr[i].dN; n++) {
							const int k = vr[i].kbase + p * vr[i].dN + n;
							kt[(size_t)(2 * k)] = (uint32_t)(t1b[i] + p * Nr + n * tC);
							kt[(size_t)(2 * k + 1)] = (uint32_t)(vr[i].dP * Nr);
						}
					}
					// stage-3 tiles; words 13-14 hold the offset within the partial slot until patched
					for (size_t i = 0; i < vq.size(); i++) {
						const int dX = vq[i].dX, dM = vq[i].dM, dQ = vq[i].dQ, Nq = dQ * tC;
						for (int mi = 0; mi < cdiv(dX, 16); mi++)
							for (int ni = 0; ni < cdiv(Nq, 16); ni++) {
								const int64_t rel = (int64_t)16 * mi * dQ * dC + c0;
								const uint32_t d[PR_T3] = { (uint32_t)(lds_l[i] + 16 * mi), (uint32_t)dX, (uint32_t)(lt2 + vq[i].rbase * S1), (uint32_t)(dQ * S1),
									(uint32_t)dM, (uint32_t)tA | ((uint32_t)cdiv(tA, 4) << 16), 0u, (uint32_t)std::min(16, dX - 16 * mi) | ((uint32_t)std::min(16, Nq - 16 * ni) << 16),
									(uint32_t)S1, (uint32_t)(16 * ni), (uint32_t)tC, (uint32_t)dC, magic(tC),
									(uint32_t)rel, (uint32_t)((uint64_t)rel >> 32), (uint32_t)(dQ * dC) };
								pt.yref.push_back(PrSlotRef{ vq[i].yblock, A, ia, 0, 0 });
								t3s.insert(t3s.end(), d, d + PR_T3);
							}
					}
					// staging pieces: theta rows zero-padded to a multiple of 4 columns, R rows b
					// zero-padded to a multiple of 4 (zero pieces), L rows, W_D and the row map contiguous
					auto add_pieces = [&](int64_t src, int d0, int n, int cols, int dstride, int sstride, int vcols, int kind) {
						for (int e0 = 0; e0 < n; e0 += 64) {
							const int row0 = e0 / cols, col0 = e0 % cols;
							const int64_t s0 = src + (int64_t)row0 * sstride;
							if (s0 < 0 || s0 > 0xffffffffLL) return false;
							const uint32_t d[PR_PIECE] = { (uint32_t)s0, (uint32_t)(d0 + row0 * dstride) | ((uint32_t)sstride << 16),
								(uint32_t)dstride | ((uint32_t)cols << 8) | ((uint32_t)col0 << 16) | ((uint32_t)std::min(64, n - e0) << 24),
								((uint32_t)((65536 + cols - 1) / cols)) | ((uint32_t)vcols << 17) | ((uint32_t)kind << 25) };
							pcs.insert(pcs.end(), d, d + PR_PIECE);
						}
						return true;
					};
					bool ok = true;
					for (size_t i = 0; i < vr.size(); i++) {
						const int dB = vr[i].dB, dB4 = up4(dB), Mr = tA * vr[i].dP, Nr = vr[i].dN * tC;
						ok = ok && add_pieces(vr[i].th + (int64_t)a0 * vr[i].dP * dB, lth[i], Mr * dB4, dB4, dB4 + 1, dB, dB, 0);
						ok = ok && add_pieces(vr[i].r + c0, lr[i], dB * vr[i].dN * tC, tC, tC, dC, tC, 1);
						if (dB4 > dB) ok = ok && add_pieces(0, lr[i] + dB * Nr, (dB4 - dB) * Nr, 64, 64, 64, 64, 3);
					}
					for (size_t i = 0; i < vq.size(); i++) {
						const int n = tA * vq[i].dM * vq[i].dX;
						ok = ok && add_pieces(vq[i].l + (int64_t)a0 * vq[i].dM * vq[i].dX, lds_l[i], n, 64, 64, 64, 64, 2);
					}
					ok = ok && add_pieces(cl.wd, lwb, cl.R16 * (cl.K4 + 1), 64, 64, 64, 64, 4);
					// the row map: the record's own words, copied as doubles (even offset); its source
					// offset is set when the unit's record is assembled
					pt.kpiece = pcs.size();
					ok = ok && add_pieces(0, lkb, cl.K4, 64, 64, 64, 64, 5);
					const int np = (int)(pcs.size() / PR_PIECE);
					if (!ok || np > nbp || cl.K4 > 64) return false;
					pt.np = np;
					pt.n1 = (int)(t1s.size() / PR_T1);
					pt.n2 = (int)(t2s.size() / PR_T2);
					pt.n3 = (int)(t3s.size() / PR_T3);
					pt.mfma = mfmas(tA, tC);
					// flops as executed (algorithmic: no tile padding)
					double f = 0;
					for (auto& v : vr) f += 2.0 * tA * v.dP * v.dB * (double)(v.dN * tC);
					f += 2.0 * (double)cl.Rd * cl.Kd * S1;
					for (auto& v : vq) f += 2.0 * v.dX * (double)(tA * v.dM) * (v.dQ * tC);
					h.flops += f;
					parts.push_back(std::move(pt));
				}
			}
		}
	}
	// XCD groups: workgroup i runs on XCD i mod PR_XCDS, each with its own L2. The left sectors A are
	// dealt to the XCDs by cost (largest first, to the least loaded), and the units of a group are
	// placed on its XCD (PairPlan::upload), so the theta and L rows of a sector are fetched into one
	// L2 rather than into every XCD's
	// estimated unit time (µs, fitted to per-unit stage timings on MI355X): each stage lasts as long
	// as the tiles of its busiest wavefront (tile t runs on wavefront t mod 4), the staging as the
	// pieces per wavefront; stage 3 also grows with its k-steps
	auto steps3 = [](const PrPart& pt) {
		double st = 0.0;
		for (size_t t = 0; t + PR_T3 <= pt.t3s.size(); t += PR_T3) st += (double)pt.t3s[t + 4] * (double)(pt.t3s[t + 5] >> 16);
		return st;
	};
	auto unit_cost = [](double np, double n1, double n2, double n3, double st3) {
		auto w = [](double n) { return std::ceil(n / 4.0); };
		return 3.0 + 0.24 * w(np) + 0.32 * w(n1) + 0.59 * w(n2) + 0.26 * w(n3) + 0.065 * w(st3);
	};
	std::map<int32_t, double> cost_a;
	for (const PrPart& pt : parts) cost_a[pt.A] += 2.0 + 0.06 * pt.np + 0.08 * pt.n1 + 0.1475 * pt.n2 + 0.065 * pt.n3 + 0.016 * steps3(pt);
	std::vector<std::pair<double, int32_t>> by_cost;
	for (const auto& kv : cost_a) by_cost.emplace_back(kv.second, kv.first);
	std::stable_sort(by_cost.begin(), by_cost.end(), [](const std::pair<double, int32_t>& x, const std::pair<double, int32_t>& y) { return x.first > y.first; });
	std::vector<double> xcd_load(PR_XCDS, 0.0);
	std::map<int32_t, int> xcd_of_a;
	for (const auto& ca : by_cost) {
		const int x = (int)(std::min_element(xcd_load.begin(), xcd_load.end()) - xcd_load.begin());
		xcd_load[(size_t)x] += ca.first;
		xcd_of_a[ca.second] = x;
	}
	// units per group: the chunks, largest LDS need first, each into the first unit that still holds
	// it within an LDS budget and the piece, tile and MFMA limits (first fit decreasing), else into a
	// new unit. Packing pays only when it saves a round of workgroups: when a group's chunks exceed
	// its share of the device's resident units (h.slots / PR_XCDS) but fit them once packed, the
	// budget is the smallest that does (small units finish sooner); otherwise every chunk is a unit of
	// its own.
	std::vector<size_t> pord;
	struct PrBin { int used, np, n1, n2, n3, mfma; std::vector<size_t> parts; };
	auto pack = [&](int budget, size_t limit) {
		std::vector<PrBin> bins;
		std::vector<size_t> open;   // units that may take another chunk (each chunk has at least 5 pieces)
		for (size_t pi : pord) {
			const PrPart& pt = parts[pi];
			const int sz = pt.need - PR_WB;
			size_t bi = bins.size();
			for (size_t oi = 0; oi < open.size(); oi++) {
				const PrBin& b = bins[open[oi]];
				if (b.used + sz <= budget && b.np + pt.np <= nbp && b.n1 + pt.n1 <= 4 * 64 && b.n2 + pt.n2 <= 4 * 64 && b.n3 + pt.n3 <= 4 * 64 &&
					b.mfma + pt.mfma <= h.mfma_cap) {
					bi = open[oi];
					break;
				}
			}
			if (bi == bins.size()) {
				if (bins.size() == limit) return std::vector<PrBin>();   // more units than wanted: give up
				bins.push_back(PrBin{ PR_WB, 0, 0, 0, 0, 0, {} });
				open.push_back(bi);
			}
			PrBin& b = bins[bi];
			b.used += sz;
			b.np += pt.np; b.n1 += pt.n1; b.n2 += pt.n2; b.n3 += pt.n3; b.mfma += pt.mfma;
			b.parts.push_back(pi);
			if (b.np + 5 > nbp || budget - b.used < 64) open.erase(std::find(open.begin(), open.end(), bi));
		}
		return bins;
	};
	// the chunks of pord into units for 'limit' resident units: packed with the smallest budget that
	// fits when they exceed it, one per unit when they already fit; empty when they cannot fit
	auto fit = [&](size_t limit) {
		std::vector<PrBin> gb;
		int64_t total = 0;
		int lo = PR_WB;   // a budget below a chunk's need leaves that chunk alone in its unit
		for (size_t pi : pord) total += parts[pi].need - PR_WB;
		if (pord.size() > limit) {
			if (total > (int64_t)limit * (h.lds - PR_WB)) return gb;
			gb = pack(h.lds, limit);
			if (gb.empty()) return gb;
			int hi = h.lds;
			while (hi - lo > 32) {
				const int mid = (lo + hi) / 2;
				std::vector<PrBin> t = pack(mid, limit);
				if (t.empty()) lo = mid;
				else { hi = mid; gb = std::move(t); }
			}
			return gb;
		}
		for (size_t pi : pord) {
			const PrPart& pt = parts[pi];
			gb.push_back(PrBin{ pt.need, pt.np, pt.n1, pt.n2, pt.n3, pt.mfma, { pi } });
		}
		return gb;
	};
	auto by_need = [&]() { std::stable_sort(pord.begin(), pord.end(), [&](size_t i, size_t j) { return parts[i].need > parts[j].need; }); };
	// one round with every unit on its sector's XCD when each XCD's chunks fit its share of the
	// resident units; otherwise the chunks of all sectors together, packed to fit the device when
	// that saves a round, or one per unit (several rounds: placement by sector would leave XCDs idle)
	std::vector<PrBin> bins;
	std::vector<int> bin_xcd;
	const size_t slots_all = (size_t)std::max(1, h.slots);
	const size_t slots_x = (size_t)std::max(1, h.slots / PR_XCDS);
	bool placed = true;
	for (int x = 0; x < PR_XCDS && placed; x++) {
		pord.clear();
		for (size_t i = 0; i < parts.size(); i++)
			if (xcd_of_a[parts[i].A] == x) pord.push_back(i);
		by_need();
		std::vector<PrBin> gb = fit(slots_x);
		if (gb.empty() && !pord.empty()) { placed = false; break; }
		for (PrBin& b : gb) {
			bins.push_back(std::move(b));
			bin_xcd.push_back(x);
		}
	}
	if (!placed) {
		bins.clear();
		bin_xcd.clear();
		pord.resize(parts.size());
		std::iota(pord.begin(), pord.end(), (size_t)0);
		by_need();
		bins = fit(slots_all);
		if (bins.empty())
			for (size_t pi : pord) {
				const PrPart& pt = parts[pi];
				bins.push_back(PrBin{ pt.need, pt.np, pt.n1, pt.n2, pt.n3, pt.mfma, { pi } });
			}
		bin_xcd.assign(bins.size(), -1);
	}
	// records: header | pieces | stage-1 tiles | stage-2 tiles | stage-3 tiles | row maps, every chunk
	// moved up by the LDS of the chunks before it in its unit
	for (size_t bi = 0; bi < bins.size(); bi++) {
		PrBin& b = bins[bi];
		h.xcd.push_back(bin_xcd[bi]);
		std::vector<int> delta(b.parts.size());
		for (size_t q = 0, at = 0; q < b.parts.size(); q++) { delta[q] = (int)at; at += (size_t)(parts[b.parts[q]].need - PR_WB); }
		auto moved = [&](const std::vector<uint32_t>& v, int stride, std::initializer_list<int> words, int dl) {
			std::vector<uint32_t> o = v;
			for (size_t e = 0; e < o.size(); e += (size_t)stride)
				for (int wdx : words) o[e + (size_t)wdx] += (uint32_t)dl;
			return o;
		};
		std::vector<uint32_t> rec(PR_HDR, 0u);
		std::vector<size_t> kword(b.parts.size());
		rec[13] = (uint32_t)rec.size();
		for (size_t q = 0; q < b.parts.size(); q++) {
			const PrPart& pt = parts[b.parts[q]];
			kword[q] = rec.size() + pt.kpiece;
			for (size_t e = 0; e < pt.pcs.size(); e += PR_PIECE) {
				uint32_t w1 = pt.pcs[e + 1];
				const uint32_t d0 = (w1 & 0xffffu) + (uint32_t)delta[q];
				if (d0 > 0xffffu) return false;
				w1 = (w1 & 0xffff0000u) | d0;
				rec.insert(rec.end(), { pt.pcs[e], w1, pt.pcs[e + 2], pt.pcs[e + 3] });
			}
		}
		rec[10] = (uint32_t)rec.size();
		for (size_t q = 0; q < b.parts.size(); q++) {
			const std::vector<uint32_t> v = moved(parts[b.parts[q]].t1s, PR_T1, { 0, 2, 7 }, delta[q]);
			rec.insert(rec.end(), v.begin(), v.end());
		}
		rec[12] = (uint32_t)rec.size();
		for (size_t q = 0; q < b.parts.size(); q++) {
			const std::vector<uint32_t> v = moved(parts[b.parts[q]].t2s, PR_T2, { 0, 2, 7 }, delta[q]);
			rec.insert(rec.end(), v.begin(), v.end());
		}
		rec[11] = (uint32_t)rec.size();
		for (size_t q = 0; q < b.parts.size(); q++) {
			const PrPart& pt = parts[b.parts[q]];
			for (size_t t = 0; t < pt.yref.size(); t++) {
				PrSlotRef rf = pt.yref[t];
				rf.unit = h.recs.size();
				rf.word = (int)(rec.size() + 13);
				refs.push_back(rf);
				std::vector<uint32_t> d(pt.t3s.begin() + (ptrdiff_t)(t * PR_T3), pt.t3s.begin() + (ptrdiff_t)((t + 1) * PR_T3));
				d[0] += (uint32_t)delta[q];
				d[2] += (uint32_t)delta[q];
				rec.insert(rec.end(), d.begin(), d.end());
			}
		}
		for (size_t q = 0; q < b.parts.size(); q++) {
			std::vector<uint32_t> kt = parts[b.parts[q]].kt;
			for (size_t k = 0; k + 1 < kt.size(); k += 2)
				if (kt[k] != 0u || kt[k + 1] != 0u) kt[k] += (uint32_t)delta[q];
			if (rec.size() % 2) rec.push_back(0u);
			rec[kword[q]] = (uint32_t)(rec.size() / 2);
			rec.insert(rec.end(), kt.begin(), kt.end());
		}
		while (rec.size() % 4) rec.push_back(0u);
		rec[0] = (uint32_t)b.n1; rec[1] = (uint32_t)b.n3; rec[2] = (uint32_t)b.np; rec[3] = (uint32_t)b.n2;
		h.max_words = std::max(h.max_words, (int)rec.size());
		h.max_lds = std::max(h.max_lds, b.used);
		double st3 = 0.0;
		for (size_t pi : b.parts) st3 += steps3(parts[pi]);
		h.cost.push_back(unit_cost(b.np, b.n1, b.n2, b.n3, st3));
		h.recs.push_back(std::move(rec));
	}
	// partial slots: one per (y block, A, row chunk), grouped by y block in (A, row chunk) order;
	// the column chunks of a pair write disjoint columns of the same slot
	std::vector<size_t> order(refs.size());
	std::iota(order.begin(), order.end(), (size_t)0);
	std::sort(order.begin(), order.end(), [&](size_t i, size_t j) {
		const PrSlotRef& a = refs[i];
		const PrSlotRef& b = refs[j];
		if (a.yblock != b.yblock) return a.yblock < b.yblock;
		if (a.A != b.A) return a.A < b.A;
		return a.ia < b.ia;
	});
	std::vector<int64_t> slot_base((size_t)th.nblocks, -1);
	std::vector<int> nslots((size_t)th.nblocks, 0);
	// per y block and XCD: the stage-3 tiles that write its partial blocks from units on that XCD
	std::vector<int> votes((size_t)th.nblocks * PR_XCDS, 0);
	for (size_t k = 0; k < order.size();) {
		const int64_t yb = refs[order[k]].yblock;
		const int64_t size = th.block_size(yb);
		slot_base[(size_t)yb] = h.part_elems;
		while (k < order.size() && refs[order[k]].yblock == yb) {
			const PrSlotRef& rs = refs[order[k]];
			const int64_t slot = h.part_elems;
			const ct_long A0 = rs.A;
			const int ia0 = rs.ia;
			for (; k < order.size() && refs[order[k]].yblock == yb && refs[order[k]].A == A0 && refs[order[k]].ia == ia0; k++) {
				const PrSlotRef& ru = refs[order[k]];
				if (ru.unit < h.xcd.size() && h.xcd[ru.unit] >= 0) votes[(size_t)yb * PR_XCDS + (size_t)h.xcd[ru.unit]]++;
				uint32_t* w = &h.recs[ru.unit][(size_t)ru.word];
				const int64_t pb = slot + (int64_t)((uint64_t)w[0] | ((uint64_t)w[1] << 32));
				w[0] = (uint32_t)pb;
				w[1] = (uint32_t)((uint64_t)pb >> 32);
				// y index of partial entry o: o - slot + (x_off + offset of the y block)
				const int64_t yd = x_off + th.offset[yb] - slot;
				if (yd < INT32_MIN || yd > INT32_MAX) return false;
				w[-7] = (uint32_t)(int32_t)yd;
			}
			h.part_elems += size;
			nslots[(size_t)yb]++;
		}
	}
	// reduction items over every y block in serialise order (blocks without slots are zeroed), each
	// queued for the XCD whose units wrote most of its partial blocks (still in that XCD's L2); the
	// four items of workgroup b (it runs on XCD b mod PR_XCDS) come from that XCD's queue while it
	// lasts, else from the longest other queue
	std::vector<std::vector<PrRed>> rq(PR_XCDS);
	for (ct_long k = 0; k < th.nblocks; k++) {
		if (!th.present(k)) continue;
		const int64_t size = th.block_size(k);
		const int* vk = &votes[(size_t)k * PR_XCDS];
		const size_t x = (size_t)(std::max_element(vk, vk + PR_XCDS) - vk);
		for (int64_t e0 = 0; e0 < size; e0 += 64) {
			PrRed it;
			it.y = x_off + th.offset[k] + e0;
			it.p = nslots[(size_t)k] ? slot_base[(size_t)k] + e0 : 0;
			it.n = (int32_t)std::min<int64_t>(64, size - e0);
			it.size = (int32_t)size;
			it.nslots = nslots[(size_t)k];
			it.pad = 0;
			rq[x].push_back(it);
		}
	}
	std::vector<size_t> head(PR_XCDS, 0);
	size_t left = 0;
	for (auto& v : rq) left += v.size();
	for (size_t b = 0; left > 0; b++)
		for (int i = 0; i < 4 && left > 0; i++) {
			size_t x = b % PR_XCDS;
			if (head[x] == rq[x].size()) {
				size_t best = 0;
				for (size_t y = 1; y < PR_XCDS; y++)
					if (rq[y].size() - head[y] > rq[best].size() - head[best]) best = y;
				x = best;
			}
			h.red.push_back(rq[x][head[x]++]);
			left--;
		}
	return true;
}

// Host replay of the unit and reduction launches with the kernels' own index arithmetic (the
// multiplier divisions included), tile order, clamps and dump stores: every operand, LDS and
// partial index in bounds, every LDS operand read after it was written in the current phase (the
// theta / R panels are dead once stage 1 ends), T1 / T2 / partial / y entries written exactly once,
// staging destinations disjoint. y (optional, length nx) receives the product. Returns 0 or a check
// number.
int pair_replay(const PairHost& h, const double* x, int64_t nx, const double* r, int64_t nr, const double* l, int64_t nl, const double* w, int64_t nw, double* y)
{
	const int LDS = h.lds, NB = h.lds / PR_THREADS;
	if (NB > 64) return 1;
	std::vector<double> wd(h.wd_idx.size());
	for (size_t i = 0; i < wd.size(); i++) {
		const int64_t q = h.wd_idx[i];
		if (q >= nw) return 1;
		wd[i] = q >= 0 ? w[q] : 0.0;
	}
	std::vector<double> part((size_t)h.part_elems, 0.0);
	std::vector<uint8_t> pwr((size_t)h.part_elems, 0);
	std::vector<int64_t> pyi((size_t)h.part_elems, -1);   // the y index a stage-3 tile gives each partial entry
	std::vector<double> sm((size_t)LDS);
	std::vector<uint8_t> sw((size_t)LDS);
	const int rec_words = std::max(PR_HDR, (h.max_words + 3) / 4 * 4);
	bool bad = false;
	auto dv = [&](int t, int d, uint32_t mag) { const int q = qdiv(t, mag); if (t < 0 || d < 1 || q != t / d) bad = true; return q; };
	for (size_t u = 0; u < h.recs.size(); u++) {
		const std::vector<uint32_t>& rec = h.recs[u];
		if (rec.size() < (size_t)PR_HDR || (int)rec.size() > rec_words) return 2;
		const uint32_t* g = rec.data();
		const int n1 = (int)g[0], n3 = (int)g[1], npiece = (int)g[2], n2 = (int)g[3];
		const int off1 = (int)g[10], off3 = (int)g[11], off2 = (int)g[12], offp = (int)g[13];
		if ((size_t)off1 + (size_t)PR_T1 * n1 > rec.size() || (size_t)off2 + (size_t)PR_T2 * n2 > rec.size() || (size_t)off3 + (size_t)PR_T3 * n3 > rec.size()) return 3;
		if (offp != PR_HDR || (size_t)PR_HDR + (size_t)PR_PIECE * npiece > (size_t)off1 || off1 % 4 || off2 % 4 || off3 % 4) return 3;
		// the kernel loads NB piece slots from PR_HDR whatever npiece is: within the padded table
		if ((size_t)u * rec_words + PR_HDR + (size_t)PR_PIECE * 4 * NB > h.recs.size() * (size_t)rec_words + PR_REC_PAD) return 3;
		if (npiece > 4 * NB || n1 > 4 * 64 || n2 > 4 * 64 || n3 > 4 * 64) return 5;
		std::vector<double> recd(rec.size() / 2);
		memcpy(recd.data(), rec.data(), recd.size() * sizeof(double));
		const uint32_t* pieces = g + offp;
		std::fill(sw.begin(), sw.end(), 0);
		// LDS entries staged from theta, R or zero pieces: the panels of stage 1, dead afterwards
		std::vector<uint8_t> panel((size_t)LDS, 0);
		auto lw = [&](int a, double v) { if (a < 0 || a >= LDS) { bad = true; return; } sm[(size_t)a] = v; sw[(size_t)a] = 1; };
		auto lr = [&](int a) -> double { if (a < 0 || a >= LDS || a == PR_DUMP || !sw[(size_t)a]) { bad = true; return 0.0; } return sm[(size_t)a]; };
		for (int i = 0; i < PR_ZERO; i++) lw(i, 0.0);
		// staging
		for (int wave = 0; wave < 4; wave++)
			for (int i = 0; i < NB; i++) {
				const int pc = wave + 4 * i;
				if (pc >= npiece) continue;
				const uint32_t* d = pieces + PR_PIECE * pc;
				const int lanes = (int)(d[2] >> 24), cols = (int)((d[2] >> 8) & 0xffu), col0 = (int)((d[2] >> 16) & 0xffu), dstride = (int)(d[2] & 0xffu);
				const int d0 = (int)(d[1] & 0xffffu), sstride = (int)(d[1] >> 16);
				const int mag = (int)(d[3] & 0x1ffffu), vcols = (int)((d[3] >> 17) & 0xffu), kind = (int)(d[3] >> 25);
				if (kind > 5 || cols < 1 || lanes < 1 || lanes > 64 || col0 >= cols) return 6;
				for (int lane = 0; lane < lanes; lane++) {
					const int t = col0 + lane;
					const int q = (t * mag) >> 16, c = t - q * cols;
					if (q != t / cols) return 6;
					const int dd = d0 + q * dstride + c;
					if (dd < PR_WB || dd >= LDS || sw[(size_t)dd]) return 8;
					if (kind == 0 || kind == 1 || kind == 3) panel[(size_t)dd] = 1;
					double v = 0.0;
					if (kind != 3 && c < vcols) {
						const int64_t sidx = (int64_t)d[0] + (int64_t)q * sstride + c;
						const double* base = kind == 0 ? x : (kind == 1 ? r : (kind == 2 ? l : (kind == 4 ? wd.data() : recd.data())));
						const int64_t nb = kind == 0 ? nx : (kind == 1 ? nr : (kind == 2 ? nl : (kind == 4 ? (int64_t)wd.size() : (int64_t)recd.size())));
						if (sidx < 0 || sidx >= nb) return 7;
						v = base[sidx];
					}
					lw(dd, v);
				}
			}
		// stage 1
		std::vector<uint8_t> t1w((size_t)LDS, 0);
		for (int t = 0; t < n1; t++) {
			const uint32_t* td = g + off1 + PR_T1 * t;
			const int abase = (int)td[0], sth = (int)td[1], bbase = (int)td[2], Nr = (int)td[3], nk = (int)td[4];
			const int vr = (int)td[5], vc = (int)td[6], sbase = (int)td[7];
			if (vr < 1 || vr > 16 || vc < 1 || vc > 16) return 9;
			for (int i = 0; i < 16; i++)
				for (int j = 0; j < 16; j++) {
					const int ar = i < vr ? abase + i * sth : abase, bcol = j < vc ? bbase + j : bbase;
					double acc = 0.0;
					for (int k = 0; k < 4 * nk; k++) acc += lr(ar + k) * lr(bcol + k * Nr);
					if (i < vr && j < vc) {
						const int a = sbase + i * Nr + j;
						if (a < PR_WB || a >= LDS || sw[(size_t)a] || t1w[(size_t)a]++) return 10;
						lw(a, acc);
					}
				}
		}
		// the theta / R panels are dead after stage 1; T2 overlays them
		for (int a = 0; a < LDS; a++) if (panel[(size_t)a]) sw[(size_t)a] = 0;
		// stage 2: per tile, W_D rows and the chunk's row map from their LDS copies
		for (int t = 0; t < n2; t++) {
			const uint32_t* td = g + off2 + PR_T2 * t;
			const int wbase = (int)td[0], wld = (int)(td[1] & 0xffffu), nk4 = (int)(td[1] >> 16), kbb = (int)td[2], tc2 = (int)td[3];
			const uint32_t mag2 = td[4];
			const int S1 = (int)td[5], j0 = (int)td[6], t2o = (int)td[7];
			if (tc2 < 1 || tc2 > PR_ZERO || S1 < tc2 || S1 % tc2 || nk4 < 1 || 4 * nk4 >= wld || j0 % 16 || j0 >= S1) return 4;
			std::vector<int> kofs((size_t)(4 * nk4)), kstr((size_t)(4 * nk4));
			for (int k = 0; k < 4 * nk4; k++) {
				const double v = lr(kbb + k);
				int32_t e[2];
				memcpy(e, &v, sizeof(e));
				kofs[(size_t)k] = e[0];
				kstr[(size_t)k] = e[1];
			}
			for (int jl = 0; jl < 16; jl++) {
				const int j2 = j0 + jl;
				const bool vj = j2 < S1;
				const int a = vj ? dv(j2, tc2, mag2) : 0, c = vj ? j2 - a * tc2 : 0;
				std::vector<double> bk((size_t)(4 * nk4));
				for (int k = 0; k < 4 * nk4; k++) bk[(size_t)k] = lr(kofs[(size_t)k] + a * kstr[(size_t)k] + c);
				for (int il = 0; il < 16; il++) {
					double acc = 0.0;
					for (int k = 0; k < 4 * nk4; k++) acc += lr(wbase + il * wld + k) * bk[(size_t)k];
					if (!vj) continue;
					const int addr = t2o + il * S1 + j2;
					if (addr < PR_WB || addr >= LDS || sw[(size_t)addr]) return 11;
					lw(addr, acc);
				}
			}
		}
		// stage 3
		for (int t = 0; t < n3; t++) {
			const uint32_t* td = g + off3 + PR_T3 * t;
			const int lbase = (int)td[0], dX = (int)td[1], t2b = (int)td[2], dQS1 = (int)td[3], dM = (int)td[4];
			const int ta3 = (int)(td[5] & 0xffffu), nka = (int)(td[5] >> 16), ydelta = (int)(int32_t)td[6];
			const int vr = (int)(td[7] & 0xffffu), vc = (int)(td[7] >> 16), S1 = (int)td[8], j0 = (int)td[9];
			const int tc3 = (int)td[10], dC = (int)td[11], dQdC = (int)td[15];
			const int64_t pbase = (int64_t)((uint64_t)td[13] | ((uint64_t)td[14] << 32));
			if (vr < 1 || vr > 16 || vc < 1 || vc > 16 || nka * 4 < ta3 || tc3 < 1 || ta3 * tc3 != S1 || dQS1 % S1) return 12;
			for (int i = 0; i < 16; i++)
				for (int j = 0; j < 16; j++) {
					const int ax = i < vr ? lbase + i : lbase;
					const bool vj = j < vc;
					const int j3 = j0 + j;
					const int q = vj ? dv(j3, tc3, td[12]) : 0, c = vj ? j3 - q * tc3 : 0;
					const int jofs = t2b + q * S1 + c;
					double acc = 0.0;
					for (int m = 0; m < dM; m++)
						for (int a = 0; a < 4 * nka; a++) {
							const bool vk = a < ta3;
							acc += lr(vk ? ax + (a * dM + m) * dX : 0) * lr(vk ? jofs + m * dQS1 + a * tc3 : 0);
						}
					if (i < vr && vj) {
						const int64_t o = pbase + (int64_t)q * dC + c + (int64_t)i * dQdC;
						if (o < 0 || o >= h.part_elems || pwr[(size_t)o]++) return 13;
						part[(size_t)o] = acc;
						pyi[(size_t)o] = o + ydelta;
					}
				}
		}
		if (bad) return 17;
	}
	for (size_t i = 0; i < pwr.size(); i++) if (pwr[i] != 1) return 18;
	std::vector<uint8_t> yw((size_t)nx, 0);
	for (const PrRed& it : h.red) {
		if (it.n < 1 || it.n > 64 || it.y < 0 || it.y + it.n > nx) return 19;
		if (it.nslots > 0 && (it.p < 0 || it.p + (int64_t)(it.nslots - 1) * it.size + it.n > h.part_elems)) return 20;
		for (int e = 0; e < it.n; e++) {
			double acc = 0.0;
			for (int s = 0; s < it.nslots; s++) {
				const int64_t o = it.p + (int64_t)s * it.size + e;
				if (pyi[(size_t)o] != it.y + e) return 23;   // the units' dot indices agree with the reduction
				acc += part[(size_t)o];
			}
			if (yw[(size_t)(it.y + e)]++) return 21;
			if (y) y[it.y + e] = acc;
		}
	}
	for (size_t i = 0; i < yw.size(); i++) if (yw[i] != 1) return 22;
	return 0;
}

template <typename T>
static T* pr_upload(const std::vector<T>& v, hipStream_t s)
{
	T* d = (T*)dev_alloc(std::max<size_t>(1, v.size()) * sizeof(T));
	copy_h2d_async(d, v.data(), v.size() * sizeof(T), s);
	return d;
}

__global__ __launch_bounds__(256) void k_pair_gather_w(const double* __restrict__ w, const int64_t* __restrict__ idx, int64_t n, double* __restrict__ wd)
{
	for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
		const int64_t q = idx[i];
		wd[i] = q >= 0 ? w[q] : 0.0;
	}
}

void PairPlan::upload(const PairHost& h, const double* w, hipStream_t s)
{
	release();
	nunits = (int)h.recs.size();
	rec_words = std::max(PR_HDR, (h.max_words + 3) / 4 * 4);
	// units on their XCD (workgroup i runs on XCD i mod PR_XCDS) when the plan placed them, the XCDs'
	// queues interleaved; an unplaced plan's units in descending cost order (the largest start first)
	bool placed = h.xcd.size() == h.recs.size();
	for (int x : h.xcd) placed = placed && x >= 0;
	std::vector<std::vector<size_t>> q(placed ? PR_XCDS : 1);   // unplaced: one queue in cost order
	for (size_t u = 0; u < h.recs.size(); u++) q[placed ? (size_t)h.xcd[u] : 0].push_back(u);
	size_t longest = 0;
	// placed units start in one round: an XCD dispatches consecutive workgroups of its queue to its
	// CUs in turn, so the queue runs in passes of one unit per CU, descending and ascending in cost
	// by turns, and every CU receives a similar sum of costs
	const size_t cus = (size_t)std::max(1, num_compute_units() / PR_XCDS);
	for (auto& v : q) {
		std::stable_sort(v.begin(), v.end(), [&](size_t a, size_t b) { return h.cost[a] > h.cost[b]; });
		if (placed)
			for (size_t p0 = cus; p0 < v.size(); p0 += 2 * cus) std::reverse(v.begin() + (ptrdiff_t)p0, v.begin() + (ptrdiff_t)std::min(v.size(), p0 + cus));
		longest = std::max(longest, v.size());
	}
	// a shorter queue is padded with empty units (no pieces or tiles; their alpha partial is zero),
	// so that every unit keeps its XCD
	const size_t EMPTY = (size_t)-1;
	std::vector<size_t> ord;
	ord.reserve(q.size() * longest);
	for (size_t j = 0; j < longest; j++)
		for (size_t x = 0; x < q.size(); x++) ord.push_back(j < q[x].size() ? q[x][j] : EMPTY);
	while (!ord.empty() && ord.back() == EMPTY) ord.pop_back();
	nunits = (int)ord.size();
	std::vector<uint32_t> flat((size_t)nunits * (size_t)rec_words + PR_REC_PAD, 0u);
	for (size_t u = 0; u < ord.size(); u++) {
		uint32_t* dst = flat.data() + u * (size_t)rec_words;
		if (ord[u] == EMPTY) {
			// header only: no pieces, tiles or y blocks; every part offset right after the header
			dst[10] = dst[11] = dst[12] = dst[13] = (uint32_t)PR_HDR;
			continue;
		}
		const auto& r = h.recs[ord[u]];
		std::copy(r.begin(), r.end(), dst);
	}
	recs = pr_upload(flat, s);
	red = pr_upload(h.red, s);
	nred = (int)h.red.size();
	part_elems = h.part_elems;
	// 64 extra doubles take the stores of lanes outside a tile
	part = (double*)dev_alloc((size_t)(part_elems + 64) * sizeof(double));
	const int64_t n = (int64_t)h.wd_idx.size();
	wd = (double*)dev_alloc((size_t)std::max<int64_t>(1, n) * sizeof(double));
	if (n > 0) {
		int64_t* idx = pr_upload(h.wd_idx, s);
		hipLaunchKernelGGL(k_pair_gather_w, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 1024)), dim3(256), 0, s, w, (const int64_t*)idx, n, wd);
		CTG_HIP(hipGetLastError());
		dev_free(idx);   // reused only behind the gather kernel on s
	}
	flops = h.flops;
	max_lds = h.max_lds;
	lds = h.lds;
}

void PairPlan::release()
{
	dev_free(recs); dev_free(red); dev_free(part); dev_free(wd);
	recs = nullptr; red = nullptr; part = nullptr; wd = nullptr;
	nunits = 0; nred = 0;
}

// LDS budget (doubles) of a unit: 2048, 3072 or 4608 (CT_PAIR_LDS overrides)
int pair_default_lds()
{
	const char* e = getenv("CT_PAIR_LDS");
	const int v = e ? atoi(e) : 0;
	return (v == 2048 || v == 3072 || v == 4608 || (getenv("CT_PAIR_LDS_ANY") && v > 0 && v <= 65536)) ? v : PR_LDS;
}

// resident units the packing aims for: the device's (or 1024 without a device), CT_PAIR_SLOTS overrides
int pair_default_slots(int device_slots)
{
	const char* e = getenv("CT_PAIR_SLOTS");
	const int v = e ? atoi(e) : 0;
	return v > 0 ? v : device_slots;
}

// MFMAs per unit above which a pair is cut into chunks (CT_PAIR_CAP overrides)
int pair_default_cap()
{
	const char* e = getenv("CT_PAIR_CAP");
	const int v = e ? atoi(e) : 0;
	return v > 0 ? v : 96;
}

// workgroups of the unit launch resident at once on this device (CUs x occupancy of the kernel)
static int pair_resident_slots(int lds)
{
	int per_cu = 0;
	switch (lds) {
	case 2048: CTG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pair_units<2048>, PR_THREADS, 0)); break;
	case 3072: CTG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pair_units<3072>, PR_THREADS, 0)); break;
	default: CTG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pair_units<4608>, PR_THREADS, 0)); break;
	}
	return std::max(1, per_cu) * num_compute_units();
}

bool build_pair_plan(PairPlan& pp, const Meta& th, const Meta& W, const Meta& L, const Meta& R, const double* w, hipStream_t s, int lds)
{
	PairHost h;
	h.mfma_cap = pair_default_cap();
	h.lds = lds > 0 ? lds : pair_default_lds();
	h.slots = pair_default_slots(pair_resident_slots(h.lds));
	if (!plan_pair_tables(h, th, W, L, R, 0, 0, 0, 0)) return false;
	pp.upload(h, w, s);
	return true;
}

bool build_pair_batch(PairPlan& pp, const std::vector<FusedProblem>& probs, const double* w, hipStream_t s)
{
	PairHost h;
	h.mfma_cap = pair_default_cap();
	h.lds = pair_default_lds();
	for (const auto& pb : probs)
		if (!plan_pair_tables(h, *pb.theta, *pb.w, *pb.l, *pb.r, pb.x_off, pb.w_off, pb.l_off, pb.r_off)) return false;
	pp.upload(h, w, s);
	return true;
}

void launch_pair_units(const PairPlan& pp, const double* x, const double* r, const double* l, hipStream_t s, double* udot)
{
	if (pp.nunits <= 0) return;
	const dim3 grid((unsigned)pp.nunits), block(PR_THREADS);
	const uint32_t* recs = pp.recs;
	const double* wd = pp.wd;
	switch (pp.lds) {
	case 2048: hipLaunchKernelGGL(k_pair_units<2048>, grid, block, 0, s, recs, pp.rec_words, x, r, l, wd, pp.part, pp.part + pp.part_elems, udot); break;
	case 3072: hipLaunchKernelGGL(k_pair_units<3072>, grid, block, 0, s, recs, pp.rec_words, x, r, l, wd, pp.part, pp.part + pp.part_elems, udot); break;
	case 4608: hipLaunchKernelGGL(k_pair_units<4608>, grid, block, 0, s, recs, pp.rec_words, x, r, l, wd, pp.part, pp.part + pp.part_elems, udot); break;
	default: throw HipError{ hipErrorInvalidValue, "sector-pair units: unsupported LDS budget" };
	}
	CTG_HIP(hipGetLastError());
}

void launch_pair_reduce(const PairPlan& pp, double* y, hipStream_t s)
{
	if (pp.nred > 0) {
		hipLaunchKernelGGL(k_pair_reduce, dim3((unsigned)((pp.nred + 3) / 4)), dim3(256), 0, s, (const PrRed*)pp.red, pp.nred, (const double*)pp.part, y);
		CTG_HIP(hipGetLastError());
	}
}

void launch_pair_matvec(const PairPlan& pp, const double* x, const double* r, const double* l, double* y, hipStream_t s)
{
	launch_pair_units(pp, x, r, l, s);
	launch_pair_reduce(pp, y, s);
}

int pair_dot_partials(const PairPlan& pp) { return std::max(1, (pp.nred + 3) / 4); }

int pair_step_partials(const PairPlan& pp) { return std::max(std::max(1, pp.nunits), (pp.nred + 3) / 4); }

void launch_pair_lz_step(const PairPlan& pp, const double* uj, double* unext, double* vj_out, const double* vjm1, int j, int maxiter,
	const double* dpart, const double* npart, double thresh, double* npart_out, double* scal, hipStream_t s)
{
	const int G = std::max(1, (pp.nred + 3) / 4);
	hipLaunchKernelGGL(k_pair_lz_step, dim3((unsigned)G), dim3(256), 0, s, (const PrRed*)pp.red, pp.nred, (const double*)pp.part, uj, unext, vj_out, vjm1,
		j, maxiter, dpart, std::max(0, pp.nunits), npart, G, thresh, npart_out, scal);
	CTG_HIP(hipGetLastError());
}

void launch_pair_matvec_dot(const PairPlan& pp, const double* x, const double* r, const double* l, double* y, double* dpart, hipStream_t s)
{
	launch_pair_units(pp, x, r, l, s);
	hipLaunchKernelGGL(k_pair_reduce_dot, dim3((unsigned)pair_dot_partials(pp)), dim3(256), 0, s, (const PrRed*)pp.red, pp.nred, (const double*)pp.part,
		y, x, dpart);
	CTG_HIP(hipGetLastError());
}

//________________________________________________________________________________________________________________________
//
// matvec without allocation: the partial blocks in a file-scope device array instead of the plan's buffer

// Each GPU holds its own copy of this array until the process ends. In host code its name is a
// host-side stand-in, not a device address: host code uses only the per-device address from
// hipGetSymbolAddress below.
static __device__ __align__(256) double scratch[PR_SCRATCH_DOUBLES];

namespace {
std::mutex scratch_mu;
std::vector<double*> scratch_ptr;   // per device: the address of 'scratch' (null until looked up)
}

// the device address of this device's scratch array, looked up once per device with that device current
static double* scratch_address()
{
	int dev = 0;
	CTG_HIP(hipGetDevice(&dev));
	std::lock_guard<std::mutex> lk(scratch_mu);
	if ((int)scratch_ptr.size() <= dev) scratch_ptr.resize((size_t)dev + 1, nullptr);
	if (!scratch_ptr[(size_t)dev]) {
		void* p = nullptr;
		CTG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
		if (!p) throw HipError{ hipErrorInvalidSymbol, "scratch array has no device address" };
		scratch_ptr[(size_t)dev] = (double*)p;
	}
	return scratch_ptr[(size_t)dev];
}

void launch_pair_matvec_noalloc(const PairPlan& pp, const double* x, const double* r, const double* l, double* y, hipStream_t s)
{
	if (pp.part_elems + 64 > PR_SCRATCH_DOUBLES)
		throw HipError{ hipErrorInvalidValue, "matvec without allocation: the plan's partial blocks exceed the scratch array" };
	double* part = scratch_address();
	// every partial entry is written by the unit launch before the reduction reads it: nothing to clear
	if (pp.nunits > 0) {
		const dim3 grid((unsigned)pp.nunits), block(PR_THREADS);
		double* dump = part + pp.part_elems;
		switch (pp.lds) {
		case 2048: hipLaunchKernelGGL(k_pair_units<2048>, grid, block, 0, s, (const uint32_t*)pp.recs, pp.rec_words, x, r, l, (const double*)pp.wd, part, dump, nullptr); break;
		case 3072: hipLaunchKernelGGL(k_pair_units<3072>, grid, block, 0, s, (const uint32_t*)pp.recs, pp.rec_words, x, r, l, (const double*)pp.wd, part, dump, nullptr); break;
		case 4608: hipLaunchKernelGGL(k_pair_units<4608>, grid, block, 0, s, (const uint32_t*)pp.recs, pp.rec_words, x, r, l, (const double*)pp.wd, part, dump, nullptr); break;
		default: throw HipError{ hipErrorInvalidValue, "sector-pair units: unsupported LDS budget" };
		}
		CTG_HIP(hipGetLastError());
	}
	if (pp.nred > 0) {
		hipLaunchKernelGGL(k_pair_reduce, dim3((unsigned)((pp.nred + 3) / 4)), dim3(256), 0, s, (const PrRed*)pp.red, pp.nred, (const double*)part, y);
		CTG_HIP(hipGetLastError());
	}
}

}  // namespace ctg
