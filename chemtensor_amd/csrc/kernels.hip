// HIP kernels of libchemtensor_gpu for gfx950 (CDNA4, 64-wide wavefronts).
//
//  grouped_gemm_f64   persistent grouped FP64 GEMM: every wavefront owns one 32 x 32 output
//                     tile at a time (2 x 2 v_mfma_f64_16x16x4 sub-tiles), sums all terms of
//                     its task in registers and writes the tile once (no atomics, deterministic);
//                     the term descriptors of a task are loaded once, one per lane
//  wapply             stage 2 of the matvec as a column stream (small dense W_D per class)
//  grouped_gemm_split the same schedule with each FP64 operand split into FP32 terms on FP32
//                     MFMA; every MFMA result is lifted into FP64 accumulators
//  plain_gemm_f64     one output block x one contracted block per launch (the per-block form)
//  permute / copy_segments / fill      data movement of transposes, reshapes and slices

#include <hip/hip_runtime.h>
#include <algorithm>
#include <type_traits>
#include "ct_internal.h"

namespace ctg {

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// column offset of task column j in B (term-dependent strides) and in C
__device__ __forceinline__ int64_t b_col_offset(int j, int nci, int64_t sb_batch, int sb_j)
{
	const int jb = j / nci, jc = j - jb * nci;
	return (int64_t)jb * sb_batch + (int64_t)jc * sb_j;
}

// One wavefront computes the 32 x 32 tile (m0, n0) of 'tk' summing terms [tk.t0, tk.t1).
// MFMA 16x16x4 f64 operand layout: lane l supplies A(l % 16, l / 16) and B(l / 16, l % 16);
// the result holds D(l / 16 + 4 r, l % 16) in register r.
__device__ __forceinline__ const double* rebase(const double* base, const double* off)
{
	return (const double*)((const char*)base + (intptr_t)off);
}

// Term descriptors of a task are held one per lane (64 at a time, one vector load) and read with
// v_readlane: walking the terms costs no dependent memory round trip.
struct TermCache
{
	uint32_t w[12];   // the GemmTerm of term base + lane (12 dwords)
	int base;
};

static_assert(sizeof(GemmTerm) == 48, "GemmTerm layout");

__device__ __forceinline__ void term_cache_load(TermCache& c, const GemmTerm* __restrict__ terms, int base, int t1, int lane)
{
	c.base = base;
	const int idx = base + lane;
	if (idx < t1) {
		const uint4* src = (const uint4*)(terms + idx);
		const uint4 a = src[0], b = src[1], d = src[2];
		c.w[0] = a.x; c.w[1] = a.y; c.w[2] = a.z; c.w[3] = a.w;
		c.w[4] = b.x; c.w[5] = b.y; c.w[6] = b.z; c.w[7] = b.w;
		c.w[8] = d.x; c.w[9] = d.y; c.w[10] = d.z; c.w[11] = d.w;
	}
	else {
		#pragma unroll
		for (int i = 0; i < 12; i++) c.w[i] = 0;
	}
}

__device__ __forceinline__ uint32_t term_word(const TermCache& c, int t, int i)
{
	return (uint32_t)__builtin_amdgcn_readlane((int)c.w[i], t - c.base);
}

__device__ __forceinline__ int term_K(const TermCache& c, int t) { return (int)term_word(c, t, 6); }

__device__ __forceinline__ const double* term_ptr(const TermCache& c, int t, int i)
{
	const uint64_t lo = term_word(c, t, i), hi = term_word(c, t, i + 1);
	return (const double*)(intptr_t)(lo | (hi << 32));
}

// The fields of term t: the task's first term comes with the work item (scalar registers), later
// ones from the lane-held cache. Uniform across the wavefront.
struct TermF
{
	const double* A;
	const double* B;
	int64_t sb_batch;
	int K, sa_i, sa_k, sb_k, sb_j;
};

__device__ __forceinline__ TermF term_fields(const TermCache& c, int t, const GemmTerm& first, int t0)
{
	TermF f;
	if (t == t0) {
		f.A = first.A; f.B = first.B; f.sb_batch = first.sb_batch;
		f.K = first.K; f.sa_i = first.sa_i; f.sa_k = first.sa_k; f.sb_k = first.sb_k; f.sb_j = first.sb_j;
	}
	else {
		f.A = term_ptr(c, t, 0); f.B = term_ptr(c, t, 2); f.sb_batch = (int64_t)term_ptr(c, t, 4);
		f.K = term_K(c, t); f.sa_i = (int)term_word(c, t, 7); f.sa_k = (int)term_word(c, t, 8);
		f.sb_k = (int)term_word(c, t, 9); f.sb_j = (int)term_word(c, t, 10);
	}
	return f;
}

__device__ __forceinline__ int term_K_at(const TermCache& c, int t, const GemmTerm& first, int t0)
{
	return (t == t0) ? first.K : term_K(c, t);
}

// the cache of a task's terms: loaded (one vector load) only when the task has more than one term,
// in flight while the first term's operands are loaded
__device__ __forceinline__ void term_cache_start(TermCache& c, const GemmTerm* __restrict__ terms, const GemmTask& tk, int lane)
{
	c.base = tk.t0;
	if (tk.t1 - tk.t0 > 1) term_cache_load(c, terms, tk.t0, tk.t1, lane);
}

// identity DPP moves: the cache words get registers written by a VALU op, so that later readlanes
// wait for nothing (the wait for the cache load happens here, once)
__device__ __forceinline__ void term_cache_settle(TermCache& c)
{
	#pragma unroll
	for (int i = 0; i < 12; i++) c.w[i] = (uint32_t)__builtin_amdgcn_mov_dpp((int)c.w[i], 0xE4, 0xF, 0xF, false);
}

__device__ __forceinline__ void wave_tile_f64(const GemmTask& tk, const GemmTerm* __restrict__ terms, const GemmTerm& first, int m0, int n0, int lane, GemmBases bs)
{
	const int i16 = lane & 15, kq = lane >> 4;
	const bool two_m = (m0 + 16 < tk.M);
	const bool two_n = (n0 + 16 < tk.N);
	const int r0 = m0 + i16, r1 = m0 + 16 + i16;
	const int c0 = n0 + i16, c1 = n0 + 16 + i16;
	const bool vr0 = r0 < tk.M, vr1 = r1 < tk.M, vc0 = c0 < tk.N, vc1 = c1 < tk.N;
	double4_t acc00 = { 0, 0, 0, 0 }, acc01 = { 0, 0, 0, 0 }, acc10 = { 0, 0, 0, 0 }, acc11 = { 0, 0, 0, 0 };
	TermCache tc;
	term_cache_start(tc, terms, tk, lane);
	for (int t = tk.t0; t < tk.t1; t++) {
		if (t - tc.base >= 64) term_cache_load(tc, terms, t, tk.t1, lane);
		const TermF f = term_fields(tc, t, first, tk.t0);
		const double* __restrict__ A = rebase(bs.A, f.A);
		const double* __restrict__ B = rebase(bs.B, f.B);
		const int64_t sb_batch = f.sb_batch;
		const int K = f.K;
		const int sa_i = f.sa_i, sa_k = f.sa_k;
		const int sb_k = f.sb_k, sb_j = f.sb_j;
		const int64_t ao0 = (int64_t)(vr0 ? r0 : 0) * sa_i, ao1 = (int64_t)(vr1 ? r1 : 0) * sa_i;
		const int64_t bo0 = b_col_offset(vc0 ? c0 : 0, tk.nci, sb_batch, sb_j);
		const int64_t bo1 = b_col_offset(vc1 ? c1 : 0, tk.nci, sb_batch, sb_j);
		for (int k0 = 0; k0 < K; k0 += 4) {
			const int k = k0 + kq;
			const bool vk = k < K;
			const int64_t ak = (int64_t)(vk ? k : 0) * sa_k, bk = (int64_t)(vk ? k : 0) * sb_k;
			const double a0 = (vk && vr0) ? A[ao0 + ak] : 0.0;
			const double b0 = (vk && vc0) ? B[bo0 + bk] : 0.0;
			acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc00, 0, 0, 0);
			if (two_n) {
				const double b1 = (vk && vc1) ? B[bo1 + bk] : 0.0;
				acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc01, 0, 0, 0);
				if (two_m) {
					const double a1 = (vk && vr1) ? A[ao1 + ak] : 0.0;
					acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc10, 0, 0, 0);
					acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc11, 0, 0, 0);
				}
			}
			else if (two_m) {
				const double a1 = (vk && vr1) ? A[ao1 + ak] : 0.0;
				acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc10, 0, 0, 0);
			}
		}
	}
	// store: for v_mfma_f64_16x16x4f64 register r of lane l holds D(l / 16 + 4 r, l % 16)
	const int nci = tk.nci;
	#pragma unroll
	for (int sub = 0; sub < 4; sub++) {
		const int mi = sub >> 1, ni = sub & 1;
		const double4_t acc = (sub == 0 ? acc00 : sub == 1 ? acc01 : sub == 2 ? acc10 : acc11);
		const int col = n0 + 16 * ni + i16;
		if (col >= tk.N) continue;
		const int jb = col / nci, jc = col - jb * nci;
		double* __restrict__ C = (double*)rebase(bs.C, tk.C) + (int64_t)jb * tk.sc_batch + jc;
		#pragma unroll
		for (int r = 0; r < 4; r++) {
			const int row = m0 + 16 * mi + kq + 4 * r;
			if (row < tk.M) {
				if (tk.accumulate) C[(int64_t)row * tk.sc_i] += acc[r];
				else C[(int64_t)row * tk.sc_i] = acc[r];
			}
		}
	}
}

__global__ __launch_bounds__(256) void k_grouped_gemm_f64(const GemmTask* __restrict__ tasks, const GemmTerm* __restrict__ terms,
	const WorkItem* __restrict__ items, int nitems, GemmBases bs)
{
	const int lane = threadIdx.x & 63;
	const int nwaves = gridDim.x * 4;
	// the item index is uniform across the wavefront: its descriptors are scalar loads
	for (int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6))); w < nitems; w += nwaves) {
		const WorkItem it = items[w];
		const GemmTask& tk = it.task;
		wave_tile_f64(tk, terms, it.first, it.m0, it.n0, lane, bs);
	}
}

// Workgroup-per-tile form: the four wavefronts of a workgroup share one 32 x 32 output tile.
// The k-steps of all terms of the task are cut into chunks of four; chunk g goes to wavefront
// g % 4, which issues the loads of its four k-steps before their MFMAs (memory-level
// parallelism). The four partial tiles are summed through LDS in a fixed order.
constexpr int CHUNK = 4;

__device__ __forceinline__ void wg_tile_f64(const GemmTask& tk, const GemmTerm* __restrict__ terms, const GemmTerm& first, int m0, int n0, GemmBases bs,
	double* __restrict__ red)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int i16 = lane & 15, kq = lane >> 4;
	const bool two_m = (m0 + 16 < tk.M);
	const bool two_n = (n0 + 16 < tk.N);
	const int r0 = m0 + i16, r1 = m0 + 16 + i16;
	const int c0 = n0 + i16, c1 = n0 + 16 + i16;
	const bool vr0 = r0 < tk.M, vr1 = r1 < tk.M, vc0 = c0 < tk.N, vc1 = c1 < tk.N;
	double4_t acc00 = { 0, 0, 0, 0 }, acc01 = { 0, 0, 0, 0 }, acc10 = { 0, 0, 0, 0 }, acc11 = { 0, 0, 0, 0 };
	int g = 0;
	TermCache tc;
	term_cache_start(tc, terms, tk, lane);
	for (int t = tk.t0; t < tk.t1; t++) {
		if (t - tc.base >= 64) term_cache_load(tc, terms, t, tk.t1, lane);
		const int K = term_K_at(tc, t, first, tk.t0);
		const int nsteps = (K + 3) >> 2;
		const int nchunks = (nsteps + CHUNK - 1) / CHUNK;
		const int c_first = (wave - g % 4 + 4) % 4;
		g += nchunks;
		if (c_first >= nchunks) continue;
		const TermF f = term_fields(tc, t, first, tk.t0);
		const double* __restrict__ A = rebase(bs.A, f.A);
		const double* __restrict__ B = rebase(bs.B, f.B);
		const int64_t sb_batch = f.sb_batch;
		const int sa_i = f.sa_i, sa_k = f.sa_k;
		const int sb_k = f.sb_k, sb_j = f.sb_j;
		const int64_t ao0 = (int64_t)(vr0 ? r0 : 0) * sa_i, ao1 = (int64_t)(vr1 ? r1 : 0) * sa_i;
		const int64_t bo0 = b_col_offset(vc0 ? c0 : 0, tk.nci, sb_batch, sb_j);
		const int64_t bo1 = b_col_offset(vc1 ? c1 : 0, tk.nci, sb_batch, sb_j);
		for (int c = c_first; c < nchunks; c += 4) {
			double a0[CHUNK], a1[CHUNK], b0[CHUNK], b1[CHUNK];
			#pragma unroll
			for (int u = 0; u < CHUNK; u++) {
				const int k = (c * CHUNK + u) * 4 + kq;
				const bool vk = k < K;
				const int64_t ak = (int64_t)(vk ? k : 0) * sa_k, bk = (int64_t)(vk ? k : 0) * sb_k;
				a0[u] = (vk && vr0) ? A[ao0 + ak] : 0.0;
				a1[u] = (vk && vr1) ? A[ao1 + ak] : 0.0;
				b0[u] = (vk && vc0) ? B[bo0 + bk] : 0.0;
				b1[u] = (vk && vc1) ? B[bo1 + bk] : 0.0;
			}
			#pragma unroll
			for (int u = 0; u < CHUNK; u++) {
				acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[u], b0[u], acc00, 0, 0, 0);
				if (two_n) acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[u], b1[u], acc01, 0, 0, 0);
				if (two_m) acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[u], b0[u], acc10, 0, 0, 0);
				if (two_m && two_n) acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[u], b1[u], acc11, 0, 0, 0);
			}
		}
	}
	// reduction over the four wavefronts: red[wave][reg][lane], summed in wave order
	#pragma unroll
	for (int r = 0; r < 4; r++) {
		red[((wave * 16) + r) * 64 + lane] = acc00[r];
		red[((wave * 16) + 4 + r) * 64 + lane] = acc01[r];
		red[((wave * 16) + 8 + r) * 64 + lane] = acc10[r];
		red[((wave * 16) + 12 + r) * 64 + lane] = acc11[r];
	}
	__syncthreads();
	// wavefront w stores sub-tile w: (mi, ni) = (w >> 1, w & 1)
	const int mi = wave >> 1, ni = wave & 1;
	const int col = n0 + 16 * ni + i16;
	if (col < tk.N) {
		const int jb = col / tk.nci, jc = col - jb * tk.nci;
		double* __restrict__ C = (double*)rebase(bs.C, tk.C) + (int64_t)jb * tk.sc_batch + jc;
		#pragma unroll
		for (int r = 0; r < 4; r++) {
			const int row = m0 + 16 * mi + kq + 4 * r;
			if (row < tk.M) {
				const int slot = wave * 4 + r;
				double v = red[(0 * 16 + slot) * 64 + lane];
				v += red[(1 * 16 + slot) * 64 + lane];
				v += red[(2 * 16 + slot) * 64 + lane];
				v += red[(3 * 16 + slot) * 64 + lane];
				if (tk.accumulate) C[(int64_t)row * tk.sc_i] += v;
				else C[(int64_t)row * tk.sc_i] = v;
			}
		}
	}
	__syncthreads();
}

__global__ __launch_bounds__(256) void k_grouped_gemm_f64_wg(const GemmTask* __restrict__ tasks, const GemmTerm* __restrict__ terms,
	const WorkItem* __restrict__ items, int nitems, GemmBases bs)
{
	__shared__ double red[4 * 16 * 64];
	for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
		const WorkItem it = items[w];
		const GemmTask& tk = it.task;
		wg_tile_f64(tk, terms, it.first, it.m0, it.n0, bs, red);
	}
}

// k-step stream: the k-steps (four products each) of all terms of a task form one stream; it is
// consumed in batches of SLOTS k-steps whose loads are all issued before their MFMAs, so a
// batch costs one memory round trip even when it spans several short terms. With nparts > 1,
// batch b of the stream belongs to part b % nparts (the wavefronts of a workgroup).
constexpr int SLOTS = 4;

struct Acc4 { double4_t a00, a01, a10, a11; };

__device__ __forceinline__ void stream_tile_f64(const GemmTask& tk, const GemmTerm* __restrict__ terms, const GemmTerm& first, int m0, int n0, GemmBases bs,
	int part, int nparts, Acc4& acc)
{
	const int lane = threadIdx.x & 63;
	const int i16 = lane & 15, kq = lane >> 4;
	const bool two_m = (m0 + 16 < tk.M);
	const bool two_n = (n0 + 16 < tk.N);
	const int r0 = m0 + i16, r1 = m0 + 16 + i16;
	const int c0 = n0 + i16, c1 = n0 + 16 + i16;
	const bool vr0 = r0 < tk.M, vr1 = r1 < tk.M, vc0 = c0 < tk.N, vc1 = c1 < tk.N;
	TermCache tc;
	term_cache_start(tc, terms, tk, lane);
	auto ensure = [&](int t) {
		if (t - tc.base >= 64) term_cache_load(tc, terms, tk.t0 + ((t - tk.t0) & ~63), tk.t1, lane);
	};
	int t = tk.t0, step = 0;
	// advance the cursor by 'skip' k-steps
	auto advance = [&](int skip) {
		while (t < tk.t1 && skip > 0) {
			ensure(t);
			const int ns = (term_K_at(tc, t, first, tk.t0) + 3) >> 2;
			const int rem = ns - step;
			if (skip >= rem) { skip -= rem; t++; step = 0; }
			else { step += skip; skip = 0; }
		}
	};
	advance(part * SLOTS);
	while (t < tk.t1) {
		double a0[SLOTS], a1[SLOTS], b0[SLOTS], b1[SLOTS];
		int filled = 0;
		#pragma unroll
		for (int u = 0; u < SLOTS; u++) {
			if (t < tk.t1) {
				ensure(t);
				const TermF f = term_fields(tc, t, first, tk.t0);
				const double* __restrict__ A = rebase(bs.A, f.A);
				const double* __restrict__ B = rebase(bs.B, f.B);
				const int64_t sb_batch = f.sb_batch;
				const int K = f.K;
				const int sa_i = f.sa_i, sa_k = f.sa_k;
				const int sb_k = f.sb_k, sb_j = f.sb_j;
				const int k = step * 4 + kq;
				const bool vk = k < K;
				const int64_t ak = (int64_t)(vk ? k : 0) * sa_k, bk = (int64_t)(vk ? k : 0) * sb_k;
				a0[u] = (vk && vr0) ? A[(int64_t)r0 * sa_i + ak] : 0.0;
				a1[u] = (vk && vr1) ? A[(int64_t)r1 * sa_i + ak] : 0.0;
				b0[u] = (vk && vc0) ? B[b_col_offset(c0, tk.nci, sb_batch, sb_j) + bk] : 0.0;
				b1[u] = (vk && vc1) ? B[b_col_offset(c1, tk.nci, sb_batch, sb_j) + bk] : 0.0;
				filled = u + 1;
				step++;
				if (step * 4 >= K) { t++; step = 0; }
			}
			else { a0[u] = 0; a1[u] = 0; b0[u] = 0; b1[u] = 0; }
		}
		#pragma unroll
		for (int u = 0; u < SLOTS; u++) {
			if (u < filled) {
				acc.a00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[u], b0[u], acc.a00, 0, 0, 0);
				if (two_n) acc.a01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[u], b1[u], acc.a01, 0, 0, 0);
				if (two_m) acc.a10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[u], b0[u], acc.a10, 0, 0, 0);
				if (two_m && two_n) acc.a11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[u], b1[u], acc.a11, 0, 0, 0);
			}
		}
		advance((nparts - 1) * SLOTS);
	}
}

__device__ __forceinline__ void store_sub(const GemmTask& tk, int m0, int n0, GemmBases bs, int mi, int ni, const double4_t& v, int lane)
{
	const int i16 = lane & 15, kq = lane >> 4;
	const int col = n0 + 16 * ni + i16;
	if (col >= tk.N) return;
	const int jb = col / tk.nci, jc = col - jb * tk.nci;
	double* __restrict__ C = (double*)rebase(bs.C, tk.C) + (int64_t)jb * tk.sc_batch + jc;
	#pragma unroll
	for (int r = 0; r < 4; r++) {
		const int row = m0 + 16 * mi + kq + 4 * r;
		if (row < tk.M) {
			if (tk.accumulate) C[(int64_t)row * tk.sc_i] += v[r];
			else C[(int64_t)row * tk.sc_i] = v[r];
		}
	}
}

// one wavefront per work item
__global__ __launch_bounds__(256) void k_stream_gemm_f64_wave(const GemmTask* __restrict__ tasks, const GemmTerm* __restrict__ terms,
	const WorkItem* __restrict__ items, int nitems, GemmBases bs)
{
	const int lane = threadIdx.x & 63;
	const int nwaves = gridDim.x * 4;
	for (int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6))); w < nitems; w += nwaves) {
		const WorkItem it = items[w];
		const GemmTask& tk = it.task;
		Acc4 acc;
		acc.a00 = acc.a01 = acc.a10 = acc.a11 = double4_t{ 0, 0, 0, 0 };
		stream_tile_f64(tk, terms, it.first, it.m0, it.n0, bs, 0, 1, acc);
		store_sub(tk, it.m0, it.n0, bs, 0, 0, acc.a00, lane);
		store_sub(tk, it.m0, it.n0, bs, 0, 1, acc.a01, lane);
		store_sub(tk, it.m0, it.n0, bs, 1, 0, acc.a10, lane);
		store_sub(tk, it.m0, it.n0, bs, 1, 1, acc.a11, lane);
	}
}

// one workgroup per work item: the four wavefronts take every fourth batch of the stream and
// their partial tiles are summed through LDS in wavefront order
__global__ __launch_bounds__(256) void k_stream_gemm_f64_wg(const GemmTask* __restrict__ tasks, const GemmTerm* __restrict__ terms,
	const WorkItem* __restrict__ items, int nitems, GemmBases bs)
{
	__shared__ double red[4 * 16 * 64];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
		const WorkItem it = items[w];
		const GemmTask& tk = it.task;
		Acc4 acc;
		acc.a00 = acc.a01 = acc.a10 = acc.a11 = double4_t{ 0, 0, 0, 0 };
		stream_tile_f64(tk, terms, it.first, it.m0, it.n0, bs, wave, 4, acc);
		#pragma unroll
		for (int r = 0; r < 4; r++) {
			red[((wave * 16) + r) * 64 + lane] = acc.a00[r];
			red[((wave * 16) + 4 + r) * 64 + lane] = acc.a01[r];
			red[((wave * 16) + 8 + r) * 64 + lane] = acc.a10[r];
			red[((wave * 16) + 12 + r) * 64 + lane] = acc.a11[r];
		}
		__syncthreads();
		double4_t v;
		#pragma unroll
		for (int r = 0; r < 4; r++) {
			const int slot = wave * 4 + r;
			v[r] = ((red[(0 * 16 + slot) * 64 + lane] + red[(1 * 16 + slot) * 64 + lane]) + red[(2 * 16 + slot) * 64 + lane]) + red[(3 * 16 + slot) * 64 + lane];
		}
		store_sub(tk, it.m0, it.n0, bs, wave >> 1, wave & 1, v, lane);
		__syncthreads();
	}
}

__global__ __launch_bounds__(256) void k_plain_gemm_f64(GemmTask tk, GemmTerm tm, GemmBases bs)
{
	const int lane = threadIdx.x & 63;
	const int tiles_n = (tk.N + TILE - 1) / TILE;
	const int tiles = tiles_n * ((tk.M + TILE - 1) / TILE);
	const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
	if (w >= tiles) return;
	tk.t0 = 0;
	tk.t1 = 1;
	wave_tile_f64(tk, &tm, tm, (w / tiles_n) * TILE, (w % tiles_n) * TILE, lane, bs);
}

//________________________________________________________________________________________________________________________
//
// split-FP32 form (FP64 operands on FP32 MFMA)
//
// For every k-step of four products, the four A values of a row share one power-of-two scale
// 2^EA (the largest |a| of the group is below 2^EA), likewise the four B values of a column
// (2^EB). Each scaled operand is cut into five 10-bit integer slices, x = sum_t x_t 2^(-10(t+1)).
// A slice product is an integer below 2^20, so one v_mfma_f32_16x16x4f32 sums its four products
// exactly, and up to four slice pairs of one level t+u can share an FP32 accumulator (< 2^24).
// Levels t+u <= 4 are kept (15 slice products); each level's exact FP32 sum is scaled by
// 2^(EA+EB-10(t+u+2)) and added into FP64 accumulators. The dropped levels and the truncated
// slices leave an error below 2^-49 |a||b| per product.

constexpr int SPLIT_BITS = 10;
constexpr int SPLIT_TERMS = 5;

// 2^(-SPLIT_BITS (L + 2)) for the levels L = 0 .. 4
__device__ __forceinline__ constexpr double split_level_scale(int L)
{
	return L == 0 ? 0x1p-20 : L == 1 ? 0x1p-30 : L == 2 ? 0x1p-40 : L == 3 ? 0x1p-50 : 0x1p-60;
}

__device__ __forceinline__ int group_exponent(double x)
{
	double m = fabs(x);
	m = fmax(m, __shfl_xor(m, 16));
	m = fmax(m, __shfl_xor(m, 32));
	return (m == 0.0) ? 0 : (ilogb(m) + 1);
}

__device__ __forceinline__ void split_slices(double x, int e, float* sl)
{
	double v = ldexp(x, -e);
	#pragma unroll
	for (int t = 0; t < SPLIT_TERMS; t++) {
		v *= (double)(1 << SPLIT_BITS);
		const double it = trunc(v);
		sl[t] = (float)it;
		v -= it;
	}
}

// The k-steps of the task's terms with (step index mod nparts) == part, summed into acc (the four
// 16 x 16 sub-tiles, register r of lane l holding row 4 (l / 16) + r and column l % 16).
__device__ __forceinline__ void wave_tile_split_part(const GemmTask& tk, const GemmTerm* __restrict__ terms, int m0, int n0, int lane, GemmBases bsp,
	int part, int nparts, double (&acc)[4][4])
{
	const int i16 = lane & 15, kq = lane >> 4;
	#pragma unroll
	for (int s = 0; s < 4; s++)
		#pragma unroll
		for (int r = 0; r < 4; r++) acc[s][r] = 0;
	int step = 0;
	const int rr[2] = { m0 + i16, m0 + 16 + i16 };
	const int cc[2] = { n0 + i16, n0 + 16 + i16 };
	for (int t = tk.t0; t < tk.t1; t++) {
		const GemmTerm tm = terms[t];
		const double* __restrict__ A = rebase(bsp.A, tm.A);
		const double* __restrict__ B = rebase(bsp.B, tm.B);
		int64_t bo[2];
		#pragma unroll
		for (int h = 0; h < 2; h++) bo[h] = b_col_offset(cc[h] < tk.N ? cc[h] : 0, tk.nci, tm.sb_batch, tm.sb_j);
		for (int k0 = 0; k0 < tm.K; k0 += 4) {
			if ((step++ % nparts) != part) continue;
			const int k = k0 + kq;
			const bool vk = k < tm.K;
			float as[2][SPLIT_TERMS], bs[2][SPLIT_TERMS];
			int ea[2], eb[2];
			#pragma unroll
			for (int h = 0; h < 2; h++) {
				const double av = (vk && rr[h] < tk.M) ? A[(int64_t)rr[h] * tm.sa_i + (int64_t)k * tm.sa_k] : 0.0;
				const double bv = (vk && cc[h] < tk.N) ? B[bo[h] + (int64_t)k * tm.sb_k] : 0.0;
				ea[h] = group_exponent(av);
				eb[h] = group_exponent(bv);
				split_slices(av, ea[h], as[h]);
				split_slices(bv, eb[h], bs[h]);
			}
			// the row scales of this k-step's output rows (register r holds row 4 (l / 16) + r)
			int erow[2][4];
			#pragma unroll
			for (int mi = 0; mi < 2; mi++)
				#pragma unroll
				for (int r = 0; r < 4; r++) erow[mi][r] = __shfl(ea[mi], 4 * kq + r);
			#pragma unroll
			for (int s = 0; s < 4; s++) {
				const int mi = s >> 1, ni = s & 1;
				float4_t lv[6];
				#pragma unroll
				for (int q = 0; q < 6; q++) lv[q] = float4_t{ 0, 0, 0, 0 };
				// level L = t + u accumulates in lv[L] (level 4 uses lv[4] and lv[5])
				#pragma unroll
				for (int ta = 0; ta < SPLIT_TERMS; ta++) {
					#pragma unroll
					for (int ub = 0; ub < SPLIT_TERMS - ta; ub++) {
						const int L = ta + ub;
						const int slot = (L == 4 && ta >= 3) ? 5 : L;
						lv[slot] = __builtin_amdgcn_mfma_f32_16x16x4f32(as[mi][ta], bs[ni][ub], lv[slot], 0, 0, 0);
					}
				}
				#pragma unroll
				for (int r = 0; r < 4; r++) {
					// level L's sum scaled by 2^(-SPLIT_BITS (L + 2)): an exact power of two, so one
					// FMA per level rounds exactly as scaling and adding would
					double sum = 0;
					#pragma unroll
					for (int q = 5; q >= 0; q--) {
						const int L = (q == 5 ? 4 : q);
						sum = fma((double)lv[q][r], split_level_scale(L), sum);
					}
					acc[s][r] += ldexp(sum, erow[mi][r] + eb[ni]);
				}
			}
		}
	}
}

// store sub-tile s of a split tile (register r of lane l: row 4 (l / 16) + r, column l % 16)
__device__ __forceinline__ void store_split_sub(const GemmTask& tk, int m0, int n0, GemmBases bsp, int s, const double (&v)[4], int lane)
{
	const int i16 = lane & 15, kq = lane >> 4;
	const int mi = s >> 1, ni = s & 1;
	const int col = n0 + 16 * ni + i16;
	if (col >= tk.N) return;
	const int jb = col / tk.nci, jc = col - jb * tk.nci;
	double* __restrict__ C = (double*)rebase(bsp.C, tk.C) + (int64_t)jb * tk.sc_batch + jc;
	#pragma unroll
	for (int r = 0; r < 4; r++) {
		const int row = m0 + 16 * mi + 4 * kq + r;
		if (row < tk.M) {
			if (tk.accumulate) C[(int64_t)row * tk.sc_i] += v[r];
			else C[(int64_t)row * tk.sc_i] = v[r];
		}
	}
}

// one wavefront per work item
__global__ __launch_bounds__(256) void k_grouped_gemm_split(const GemmTask* __restrict__ tasks, const GemmTerm* __restrict__ terms,
	const WorkItem* __restrict__ items, int nitems, GemmBases bs)
{
	const int lane = threadIdx.x & 63;
	const int nwaves = gridDim.x * 4;
	for (int w = blockIdx.x * 4 + (threadIdx.x >> 6); w < nitems; w += nwaves) {
		const WorkItem it = items[w];
		const GemmTask& tk = it.task;
		double acc[4][4];
		wave_tile_split_part(tk, terms, it.m0, it.n0, lane, bs, 0, 1, acc);
		#pragma unroll
		for (int s = 0; s < 4; s++) store_split_sub(tk, it.m0, it.n0, bs, s, acc[s], lane);
	}
}

// one workgroup per work item (long k-step chains): wavefront w takes the k-steps w, w + 4, ...;
// the four partial tiles are summed through LDS in wavefront order (deterministic), wavefront s
// storing sub-tile s
__global__ __launch_bounds__(256) void k_grouped_gemm_split_wg(const GemmTask* __restrict__ tasks, const GemmTerm* __restrict__ terms,
	const WorkItem* __restrict__ items, int nitems, GemmBases bs)
{
	__shared__ double red[4][4][4][64];   // [wavefront][sub-tile][register][lane], 32 KB
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
		const WorkItem it = items[w];
		const GemmTask& tk = it.task;
		double acc[4][4];
		wave_tile_split_part(tk, terms, it.m0, it.n0, lane, bs, wave, 4, acc);
		#pragma unroll
		for (int s = 0; s < 4; s++)
			#pragma unroll
			for (int r = 0; r < 4; r++) red[wave][s][r][lane] = acc[s][r];
		__syncthreads();
		double v[4];
		#pragma unroll
		for (int r = 0; r < 4; r++) v[r] = ((red[0][wave][r][lane] + red[1][wave][r][lane]) + red[2][wave][r][lane]) + red[3][wave][r][lane];
		store_split_sub(tk, it.m0, it.n0, bs, wave, v, lane);
		__syncthreads();
	}
}

//________________________________________________________________________________________________________________________
//
// data movement

struct PermArgs
{
	int64_t sstride[8];   // source stride of each destination axis
	int64_t ddim[8];      // destination dimensions
	int ndim;
	int64_t n;
};

__global__ __launch_bounds__(256) void k_permute(const double* __restrict__ src, double* __restrict__ dst, PermArgs a)
{
	for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < a.n; e += (int64_t)gridDim.x * blockDim.x) {
		int64_t rem = e, so = 0;
		for (int i = a.ndim - 1; i >= 0; i--) {
			const int64_t q = rem / a.ddim[i];
			so += (rem - q * a.ddim[i]) * a.sstride[i];
			rem = q;
		}
		dst[e] = src[so];
	}
}

__global__ __launch_bounds__(256) void k_permute_blocks(const double* __restrict__ src, double* __restrict__ dst, const PermBlock* __restrict__ blocks, int nblocks)
{
	for (int b = blockIdx.x; b < nblocks; b += gridDim.x) {
		const PermBlock pb = blocks[b];
		for (int64_t e = threadIdx.x; e < pb.n; e += blockDim.x) {
			int64_t rem = e, so = 0;
			// axes from the last to the first; unrolled with constant indices, so the descriptor
			// stays in registers (a loop bounded by ndim indexes it dynamically, in scratch memory)
#pragma unroll
			for (int t = 0; t < 6; t++) {
				const int i = 5 - t;
				if (i < pb.ndim) {
					const int64_t q = rem / pb.ddim[i];
					so += (rem - q * pb.ddim[i]) * pb.sstride[i];
					rem = q;
				}
			}
			dst[pb.dst + e] = src[pb.src + so];
		}
	}
}

// one wavefront per segment (segments are rows of blocks: mostly tens to hundreds of entries)
__global__ __launch_bounds__(256) void k_copy_segments(const double* __restrict__ src, double* __restrict__ dst, const CopySeg* __restrict__ segs, int nsegs)
{
	const int lane = threadIdx.x & 63;
	const int nw = blockDim.x >> 6;
	for (int s = blockIdx.x * nw + (threadIdx.x >> 6); s < nsegs; s += gridDim.x * nw) {
		const CopySeg g = segs[s];
		for (int64_t i = lane; i < g.len; i += 64) dst[g.dst + i] = src[g.src + i];
	}
}

__global__ __launch_bounds__(256) void k_copy_f64(const double* __restrict__ src, double* __restrict__ dst, int64_t n)
{
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_fill(double* __restrict__ p, double v, int64_t n)
{
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

//________________________________________________________________________________________________________________________
//
// LDS-tiled form for large tasks: one workgroup per 64 x 64 output tile, each wavefront a 32 x 32
// quadrant (2 x 2 MFMA 16x16x4 f64 sub-tiles). The k range of every term is cut into panels of 16;
// the A panels (64 rows x 16) and B panels (16 x 64 columns) of the next two panels are in flight
// in two register sets while the current panel's MFMAs run from LDS (LDS double buffer, one barrier
// per panel). Every lane issues every panel load (entries past a term's K or past the task's rows /
// columns read an in-bounds entry and are replaced by zeros; past the last panel the last one is
// loaded again and dropped), so the wait for one register set counts exactly the eight loads of the
// other. The k-steps run in the order of the wavefront forms (terms in order, k ascending in steps
// of four).
constexpr int LK = 16;                 // k-panel
constexpr int LA_LD = LDS_TILE;        // A panel stored k-major: entry (row, k) at k LA_LD + row
constexpr int LB_LD = LDS_TILE;        // B panel: entry (k, col) at k LB_LD + col

struct LdsPanelLoad
{
	double a[4], b[4];
};

__global__ __launch_bounds__(256) void k_grouped_gemm_f64_lds(const GemmTerm* __restrict__ terms, const WorkItem* __restrict__ items, GemmBases bs)
{
	// both panels k-major, unpadded (the MFMA fragments read 16 consecutive rows / columns per k):
	// 32 KB; the two register sets of panels in flight (100 VGPRs) allow four workgroups per CU
	// (five with one set in flight were slower; padding the panels to stride 80 against the read
	// bank conflicts made no difference)
	__shared__ double sa[2][LK * LA_LD];
	__shared__ double sb[2][LK * LB_LD];
	const int tid = threadIdx.x, lane = tid & 63;
	const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	const WorkItem it = items[blockIdx.x];
	const GemmTask& tk = it.task;
	const int m0 = it.m0, n0 = it.n0;
	// loads of this thread: A rows ra (4 consecutive k from ka), B column cb (k rows kb + 4 i)
	const int ra = tid >> 2, ka = (tid & 3) * 4;
	const int cb = tid & 63, kb = tid >> 6;
	const int arow = m0 + ra, bcol = n0 + cb;
	const bool va = arow < tk.M, vb = bcol < tk.N;
	TermCache tc;
	term_cache_start(tc, terms, tk, lane);
	term_cache_settle(tc);
	// the panel (t, k0) into registers: row / column 0 and k = 0 stand in for entries outside the
	// term (every term present here has M, N, K >= 1)
	auto load_panel = [&](int t, int k0, LdsPanelLoad& r) {
		if (t - tc.base >= 64) {
			term_cache_load(tc, terms, t, tk.t1, lane);
			term_cache_settle(tc);
		}
		const TermF f = term_fields(tc, t, it.first, tk.t0);
		const double* __restrict__ A = rebase(bs.A, f.A);
		const double* __restrict__ B = rebase(bs.B, f.B);
		const int64_t ao = (int64_t)(va ? arow : 0) * f.sa_i;
		const int64_t bo = b_col_offset(vb ? bcol : 0, tk.nci, f.sb_batch, f.sb_j);
		#pragma unroll
		for (int i = 0; i < 4; i++) {
			const int k = k0 + ka + i;
			const bool ok = va && k < f.K;
			const double v = A[ao + (int64_t)(ok ? k : 0) * f.sa_k];
			r.a[i] = ok ? v : 0.0;
		}
		#pragma unroll
		for (int i = 0; i < 4; i++) {
			const int k = k0 + kb + 4 * i;
			const bool ok = vb && k < f.K;
			const double v = B[bo + (int64_t)(ok ? k : 0) * f.sb_k];
			r.b[i] = ok ? v : 0.0;
		}
	};
	auto store_panel = [&](int buf, const LdsPanelLoad& r) {
		#pragma unroll
		for (int i = 0; i < 4; i++) sa[buf][(ka + i) * LA_LD + ra] = r.a[i];
		#pragma unroll
		for (int i = 0; i < 4; i++) sb[buf][(kb + 4 * i) * LB_LD + cb] = r.b[i];
	};
	// panels in order: (t, k0) with k0 = 0, 16, ... < K of term t
	int t = tk.t0, k0 = 0;
	// K of term t (the cache moved to cover t first)
	auto term_k = [&](int tt) {
		if (tt - tc.base >= 64) {
			term_cache_load(tc, terms, tt, tk.t1, lane);
			term_cache_settle(tc);
		}
		return term_K_at(tc, tt, it.first, tk.t0);
	};
	auto advance = [&]() {
		k0 += LK;
		while (t < tk.t1 && k0 >= term_k(t)) { t++; k0 = 0; }
	};
	// the next panel into r (or the last one again, dropped, once the panels are exhausted)
	int tp = t, kp = 0;
	auto issue = [&](LdsPanelLoad& r) {
		const bool real = t < tk.t1;
		if (real) { tp = t; kp = k0; }
		load_panel(tp, kp, r);
		if (real) advance();
		return real;
	};
	// a term with K = 0 contributes nothing: skip to the first panel with entries
	while (t < tk.t1 && term_k(t) <= 0) t++;
	double4_t acc00 = { 0, 0, 0, 0 }, acc01 = acc00, acc10 = acc00, acc11 = acc00;
	const int i16 = lane & 15, kq = lane >> 4;
	const int wm = wave >> 1, wn = wave & 1;
	const int ar0 = kq * LA_LD + 32 * wm + i16, ar1 = ar0 + 16;
	const int bc0 = kq * LB_LD + 32 * wn + i16, bc1 = bc0 + 16;
	// 16 x 16 sub-tiles of this wave holding output rows / columns (0, 1 or 2 each way): a wave runs
	// only their MFMAs (the edge tiles of a task whose M or N is not a multiple of 64), and a wave with
	// none still loads and stores its share of the panels
	const int rows = tk.M - m0 - 32 * wm, cols = tk.N - n0 - 32 * wn;
	const int nm = rows <= 0 ? 0 : rows <= 16 ? 1 : 2, nn = cols <= 0 ? 0 : cols <= 16 ? 1 : 2;
	auto mainloop = [&](auto cm, auto cn) {
		constexpr int NM = decltype(cm)::value, NN = decltype(cn)::value;
		auto compute = [&](int buf) {
			const double* __restrict__ pa = sa[buf];
			const double* __restrict__ pb = sb[buf];
			if constexpr (NM > 0 && NN > 0) {
				#pragma unroll
				for (int st = 0; st < LK / 4; st++) {
					const double a0 = pa[ar0 + 4 * st * LA_LD], b0 = pb[bc0 + 4 * st * LB_LD];
					acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc00, 0, 0, 0);
					if constexpr (NN > 1) {
						const double b1 = pb[bc1 + 4 * st * LB_LD];
						acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc01, 0, 0, 0);
						if constexpr (NM > 1) {
							const double a1 = pa[ar1 + 4 * st * LA_LD];
							acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc10, 0, 0, 0);
							acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc11, 0, 0, 0);
						}
					}
					else if constexpr (NM > 1) {
						const double a1 = pa[ar1 + 4 * st * LA_LD];
						acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc10, 0, 0, 0);
					}
				}
			}
		};
		LdsPanelLoad r0, r1;
		issue(r0);
		bool m1 = issue(r1);
		store_panel(0, r0);
		__syncthreads();
		int buf = 0;
		while (true) {
			// LDS buf: panel p; r1: panel p + 1 (when m1); r0 receives panel p + 2
			const bool m2 = issue(r0);
			compute(buf);
			if (!m1) break;
			store_panel(buf ^ 1, r1);
			__syncthreads();
			buf ^= 1;
			// LDS buf: panel p + 1; r0: panel p + 2 (when m2); r1 receives panel p + 3
			m1 = issue(r1);
			compute(buf);
			if (!m2) break;
			store_panel(buf ^ 1, r0);
			__syncthreads();
			buf ^= 1;
		}
	};
	typedef std::integral_constant<int, 0> c0;
	typedef std::integral_constant<int, 1> c1;
	typedef std::integral_constant<int, 2> c2;
	if (t < tk.t1) {
		if (nm == 2 && nn == 2) mainloop(c2(), c2());
		else if (nm == 2 && nn == 1) mainloop(c2(), c1());
		else if (nm == 1 && nn == 2) mainloop(c1(), c2());
		else if (nm == 1 && nn == 1) mainloop(c1(), c1());
		else mainloop(c0(), c0());
	}
	// store: register r of lane l holds D(l / 16 + 4 r, l % 16) of each 16 x 16 sub-tile
	const int nci = tk.nci;
	#pragma unroll
	for (int sub = 0; sub < 4; sub++) {
		const int mi = sub >> 1, ni = sub & 1;
		const double4_t acc = (sub == 0 ? acc00 : sub == 1 ? acc01 : sub == 2 ? acc10 : acc11);
		const int col = n0 + 32 * wn + 16 * ni + i16;
		if (col >= tk.N) continue;
		const int jb = col / nci, jc = col - jb * nci;
		double* __restrict__ C = (double*)rebase(bs.C, tk.C) + (int64_t)jb * tk.sc_batch + jc;
		#pragma unroll
		for (int rr = 0; rr < 4; rr++) {
			const int row = m0 + 32 * wm + 16 * mi + kq + 4 * rr;
			if (row < tk.M) {
				if (tk.accumulate) C[(int64_t)row * tk.sc_i] += acc[rr];
				else C[(int64_t)row * tk.sc_i] = acc[rr];
			}
		}
	}
}

//________________________________________________________________________________________________________________________
//
// launchers

int num_compute_units()
{
	// initialised once (thread-safe): many host threads plan launches concurrently
	static const int cus = [] {
		int dev = 0;
		CTG_HIP(hipGetDevice(&dev));
		hipDeviceProp_t prop;
		CTG_HIP(hipGetDeviceProperties(&prop, dev));
		return prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
	}();
	return cus;
}

// the wavefront / workgroup forms over an item list
static void launch_gemm_form(const DevGemm& g, const WorkItem* items, int nitems, GemmBases b, hipStream_t s)
{
	if (nitems == 0) return;
	const int cap = num_compute_units() * 8;
	const int wave_grid = std::max(1, std::min((nitems + 3) / 4, cap));
	const int wg_grid = std::max(1, std::min(nitems, cap));
	switch (g.form) {
		case GEMM_WAVE_STREAM: hipLaunchKernelGGL(k_stream_gemm_f64_wave, dim3(wave_grid), dim3(256), 0, s, g.tasks, g.terms, items, nitems, b); break;
		case GEMM_WG_CHUNK: hipLaunchKernelGGL(k_grouped_gemm_f64_wg, dim3(wg_grid), dim3(256), 0, s, g.tasks, g.terms, items, nitems, b); break;
		case GEMM_WG_STREAM: hipLaunchKernelGGL(k_stream_gemm_f64_wg, dim3(wg_grid), dim3(256), 0, s, g.tasks, g.terms, items, nitems, b); break;
		default: hipLaunchKernelGGL(k_grouped_gemm_f64, dim3(wave_grid), dim3(256), 0, s, g.tasks, g.terms, items, nitems, b); break;
	}
	CTG_HIP(hipGetLastError());
}

void launch_grouped_gemm_f64(const DevGemm& g, GemmBases b, hipStream_t s)
{
	if (g.nitems == 0) return;
	if (g.use_lds && g.nlitems > 0) {
		// the large tasks on the LDS-tiled form, the others on the chosen form (disjoint outputs)
		hipLaunchKernelGGL(k_grouped_gemm_f64_lds, dim3((unsigned)g.nlitems), dim3(256), 0, s, g.terms, g.litems, b);
		CTG_HIP(hipGetLastError());
		launch_gemm_form(g, g.ritems, g.nritems, b, s);
		return;
	}
	const int wg_grid = std::max(1, std::min(g.nitems, num_compute_units() * 8));
	switch (g.form) {
		case GEMM_WAVE_STREAM: hipLaunchKernelGGL(k_stream_gemm_f64_wave, dim3(g.grid), dim3(256), 0, s, g.tasks, g.terms, g.items, g.nitems, b); break;
		case GEMM_WG_CHUNK: hipLaunchKernelGGL(k_grouped_gemm_f64_wg, dim3(wg_grid), dim3(256), 0, s, g.tasks, g.terms, g.items, g.nitems, b); break;
		case GEMM_WG_STREAM: hipLaunchKernelGGL(k_stream_gemm_f64_wg, dim3(wg_grid), dim3(256), 0, s, g.tasks, g.terms, g.items, g.nitems, b); break;
		default: hipLaunchKernelGGL(k_grouped_gemm_f64, dim3(g.grid), dim3(256), 0, s, g.tasks, g.terms, g.items, g.nitems, b); break;
	}
	CTG_HIP(hipGetLastError());
}

void launch_grouped_gemm_split(const DevGemm& g, GemmBases b, hipStream_t s)
{
	if (g.nitems == 0) return;
	// long k-step chains (the heuristic of the FP64 forms) get a workgroup per tile
	if (g.avg_steps > 8.0) {
		const int wg_grid = std::max(1, std::min(g.nitems, num_compute_units() * 8));
		hipLaunchKernelGGL(k_grouped_gemm_split_wg, dim3(wg_grid), dim3(256), 0, s, g.tasks, g.terms, g.items, g.nitems, b);
	}
	else hipLaunchKernelGGL(k_grouped_gemm_split, dim3(g.grid), dim3(256), 0, s, g.tasks, g.terms, g.items, g.nitems, b);
	CTG_HIP(hipGetLastError());
}

void launch_plain_gemm_f64(const GemmTask& task, const GemmTerm& term, GemmBases b, hipStream_t s)
{
	const int tiles = ((task.M + TILE - 1) / TILE) * ((task.N + TILE - 1) / TILE);
	if (tiles == 0) return;
	hipLaunchKernelGGL(k_plain_gemm_f64, dim3((tiles + 3) / 4), dim3(256), 0, s, task, term, b);
	CTG_HIP(hipGetLastError());
}

void launch_permute(const double* src, double* dst, int ndim, const ct_long* dims, const int* perm, hipStream_t s)
{
	PermArgs a;
	a.ndim = ndim;
	int64_t st[8];
	int64_t f = 1;
	for (int i = ndim - 1; i >= 0; i--) { st[i] = f; f *= dims[i]; }
	a.n = f;
	for (int i = 0; i < ndim; i++) { a.ddim[i] = dims[perm[i]]; a.sstride[i] = st[perm[i]]; }
	if (a.n == 0) return;
	const int64_t nb = (a.n + 255) / 256;
	hipLaunchKernelGGL(k_permute, dim3((unsigned)(nb < 4096 ? nb : 4096)), dim3(256), 0, s, src, dst, a);
	CTG_HIP(hipGetLastError());
}

void launch_permute_blocks(const double* src, double* dst, const PermBlock* d_blocks, int nblocks, hipStream_t s)
{
	if (nblocks == 0) return;
	hipLaunchKernelGGL(k_permute_blocks, dim3(nblocks < 8192 ? nblocks : 8192), dim3(256), 0, s, src, dst, d_blocks, nblocks);
	CTG_HIP(hipGetLastError());
}

void launch_copy_segments(const double* src, double* dst, const CopySeg* d_segs, int nsegs, hipStream_t s)
{
	if (nsegs == 0) return;
	const int nb = (nsegs + 3) / 4;
	hipLaunchKernelGGL(k_copy_segments, dim3(nb < 8192 ? nb : 8192), dim3(256), 0, s, src, dst, d_segs, nsegs);
	CTG_HIP(hipGetLastError());
}

void launch_copy(const double* src, double* dst, ct_long n, hipStream_t s)
{
	if (n <= 0) return;
	const int64_t nb = (n + 255) / 256;
	hipLaunchKernelGGL(k_copy_f64, dim3((unsigned)(nb < 16384 ? nb : 16384)), dim3(256), 0, s, src, dst, (int64_t)n);
	CTG_HIP(hipGetLastError());
}

void launch_fill(double* p, double v, ct_long n, hipStream_t s)
{
	if (n <= 0) return;
	const int64_t nb = (n + 255) / 256;
	hipLaunchKernelGGL(k_fill, dim3((unsigned)(nb < 8192 ? nb : 8192)), dim3(256), 0, s, p, v, (int64_t)n);
	CTG_HIP(hipGetLastError());
}

//________________________________________________________________________________________________________________________
//
// column-stream stage 2: T2 = W_D T1 on the columns (a, c) of a block pair

template <int KM>
__global__ __launch_bounds__(256) void k_wapply(const WaUnit* __restrict__ units, const WaItem* __restrict__ items, int nitems,
	const int64_t* __restrict__ kbase, const int* __restrict__ kstride, const int64_t* __restrict__ rbase, const int* __restrict__ rstride,
	const double* __restrict__ wmat, const double* __restrict__ t1, double* __restrict__ t2)
{
	// the item index is the same for the 64 lanes of a wavefront: descriptor loads go through the scalar cache
	const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
	if (w >= nitems) return;
	const int lane = threadIdx.x & 63;
	const WaItem it = items[w];
	const WaUnit& u = it.u;
	const int j = it.col0 + lane;
	const bool act = j < u.dA * u.dC;
	const int a = act ? j / u.dC : 0, c = act ? j - (j / u.dC) * u.dC : 0;
	double x[KM];
	#pragma unroll
	for (int k = 0; k < KM; k++) {
		double v = 0.0;
		if (k < u.K) {
			const int64_t b = kbase[u.kb + k];
			if (act && b >= 0) v = t1[b + (int64_t)a * kstride[u.kb + k] + c];
		}
		x[k] = v;
	}
	const double* __restrict__ wm = wmat + u.wrow * WA_KMAX;
	for (int r = 0; r < u.R; r++) {
		double y = 0.0;
		#pragma unroll
		for (int k = 0; k < KM; k++) y = fma(wm[(int64_t)r * WA_KMAX + k], x[k], y);
		const int64_t b = rbase[u.rb + r];
		if (act && b >= 0) t2[b + (int64_t)a * rstride[u.rb + r] + c] = y;
	}
}

// stages 1 and 2 fused: one wavefront (workgroup of 64) per tile of a block pair
__global__ __launch_bounds__(64) void k_f12(const F12Item* __restrict__ items, int nitems, const F12Group* __restrict__ groups,
	const int64_t* __restrict__ rbase, const int* __restrict__ rstride, const double* __restrict__ wmatT,
	const double* __restrict__ x, const double* __restrict__ r, double* __restrict__ t2)
{
	__shared__ double sm[F12_LDS];
	const int w = blockIdx.x;
	if (w >= nitems) return;
	const int lane = threadIdx.x;
	const F12Item it = items[w];
	const WaUnit& u = it.u;
	// stage the theta rows a0 .. a0+ta and the R columns c0 .. c0+tc of every present column group
	int off = 0;
	for (int g = 0; g < it.ng; g++) {
		const F12Group gr = groups[it.g0 + g];
		if (gr.th_off < 0) continue;
		const int nth = it.ta * gr.dP * gr.dB;
		const double* __restrict__ ths = x + gr.th_off + (int64_t)it.a0 * gr.dP * gr.dB;
		for (int e = lane; e < nth; e += 64) sm[off + e] = ths[e];
		off += nth;
		const int nr = gr.dB * gr.dN * it.tc;
		const double* __restrict__ rs = r + gr.r_off + it.c0;
		for (int e = lane; e < nr; e += 64) {
			const int row = e / it.tc, ci = e - row * it.tc;
			sm[off + e] = rs[(int64_t)row * u.dC + ci];
		}
		off += nr;
	}
	__syncthreads();
	const bool act = lane < it.ta * it.tc;
	const int ai = act ? lane / it.tc : 0, ci = act ? lane - (lane / it.tc) * it.tc : 0;
	double y[F12_RMAX];
	#pragma unroll
	for (int q = 0; q < F12_RMAX; q++) y[q] = 0.0;
	const double* __restrict__ wm = wmatT + u.wcol * F12_RMAX;   // column k at wm[k * F12_RMAX]
	int k = 0;
	off = 0;
	for (int g = 0; g < it.ng; g++) {
		const F12Group gr = groups[it.g0 + g];
		if (gr.th_off < 0) { k += gr.dP * gr.dN; continue; }
		const double* ths = sm + off + ai * gr.dP * gr.dB;
		const double* rs = sm + off + it.ta * gr.dP * gr.dB + ci;
		for (int p = 0; p < gr.dP; p++) {
			for (int n = 0; n < gr.dN; n++, k++) {
				// T1(A,P,N,C)[a p n c] for this lane's column
				double xv = 0.0;
				for (int b = 0; b < gr.dB; b++) xv = fma(ths[p * gr.dB + b], rs[(b * gr.dN + n) * it.tc], xv);
				const double* __restrict__ wk = wm + (int64_t)k * F12_RMAX;
				#pragma unroll
				for (int q = 0; q < F12_RMAX; q++) y[q] = fma(wk[q], xv, y[q]);   // zero beyond R
			}
		}
		off += it.ta * gr.dP * gr.dB + gr.dB * gr.dN * it.tc;
	}
	if (!act) return;
	const int a = it.a0 + ai, c = it.c0 + ci;
	#pragma unroll
	for (int q = 0; q < F12_RMAX; q++) {
		if (q < u.R) {
			const int64_t b = rbase[u.rb + q];
			if (b >= 0) t2[b + (int64_t)a * rstride[u.rb + q] + c] = y[q];
		}
	}
}

void launch_f12(const WApply& wa, const double* x, const double* r, double* t2, hipStream_t s)
{
	if (wa.nitems12 == 0) return;
	hipLaunchKernelGGL(k_f12, dim3((unsigned)wa.nitems12), dim3(64), 0, s, wa.items12, wa.nitems12, wa.groups, wa.rbase, wa.rstride, wa.wmatT, x, r, t2);
	CTG_HIP(hipGetLastError());
}

// dense W_D copies: wmat[i] = W[widx[i]] (0 where widx < 0)
__global__ __launch_bounds__(256) void k_gather_w(const double* __restrict__ w, const int64_t* __restrict__ widx, int64_t n, double* __restrict__ wmat)
{
	for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
		const int64_t q = widx[i];
		wmat[i] = q >= 0 ? w[q] : 0.0;
	}
}

template <typename T>
static T* upload_array(const std::vector<T>& v, hipStream_t s)
{
	T* d = (T*)dev_alloc(std::max<size_t>(1, v.size()) * sizeof(T));
	copy_h2d_async(d, v.data(), v.size() * sizeof(T), s);
	return d;
}

void WApply::upload(const WApplyHost& h, const double* w, hipStream_t s)
{
	release();
	ok = h.kmax <= WA_KMAX;
	if (!ok) return;
	units = upload_array(h.units, s);
	items = upload_array(h.items, s);
	kbase = upload_array(h.kbase, s);
	rbase = upload_array(h.rbase, s);
	kstride = upload_array(h.kstride, s);
	rstride = upload_array(h.rstride, s);
	int64_t* widx = upload_array(h.widx, s);
	const int64_t n = (int64_t)h.widx.size();
	wmat = (double*)dev_alloc((size_t)std::max<int64_t>(n, 1) * sizeof(double));
	if (n > 0) {
		hipLaunchKernelGGL(k_gather_w, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)), dim3(256), 0, s, w, (const int64_t*)widx, n, wmat);
		CTG_HIP(hipGetLastError());
	}
	dev_free(widx);   // reused only behind the gather kernel on s
	nitems = (int)h.items.size();
	f12_ok = h.f12_ok && !h.items12.empty();
	if (f12_ok) {
		groups = upload_array(h.groups, s);
		items12 = upload_array(h.items12, s);
		nitems12 = (int)h.items12.size();
		int64_t* widxT = upload_array(h.widxT, s);
		const int64_t nt = (int64_t)h.widxT.size();
		wmatT = (double*)dev_alloc((size_t)std::max<int64_t>(nt, 1) * sizeof(double));
		if (nt > 0) {
			hipLaunchKernelGGL(k_gather_w, dim3((unsigned)std::min<int64_t>((nt + 255) / 256, 4096)), dim3(256), 0, s, w, (const int64_t*)widxT, nt, wmatT);
			CTG_HIP(hipGetLastError());
		}
		dev_free(widxT);
	}
	kmax = h.kmax;
}

void WApply::release()
{
	dev_free(units); dev_free(items); dev_free(kbase); dev_free(rbase); dev_free(kstride); dev_free(rstride); dev_free(wmat);
	dev_free(groups); dev_free(items12); dev_free(wmatT);
	wmatT = nullptr;
	units = nullptr; items = nullptr; kbase = nullptr; rbase = nullptr; kstride = nullptr; rstride = nullptr; wmat = nullptr;
	groups = nullptr; items12 = nullptr;
	nitems = 0;
	nitems12 = 0;
	ok = false;
	f12_ok = false;
}

void launch_wapply(const WApply& wa, const double* t1, double* t2, hipStream_t s)
{
	if (wa.nitems == 0) return;
	const dim3 grid((unsigned)((wa.nitems + 3) / 4));
	if (wa.kmax <= 16)
		hipLaunchKernelGGL(k_wapply<16>, grid, dim3(256), 0, s, wa.units, wa.items, wa.nitems, wa.kbase, wa.kstride, wa.rbase, wa.rstride, wa.wmat, t1, t2);
	else
		hipLaunchKernelGGL(k_wapply<32>, grid, dim3(256), 0, s, wa.units, wa.items, wa.nitems, wa.kbase, wa.kstride, wa.rbase, wa.rstride, wa.wmat, t1, t2);
	CTG_HIP(hipGetLastError());
}

}  // namespace ctg
