// Device-resident split of the two-site tensor (mps_split_tensor_svd, src/state/mps.c:1119-1160 ->
// split_block_sparse_matrix_svd, src/algorithm/bond_ops.c:15-138 -> block_sparse_tensor_svd,
// src/tensor/block_sparse_tensor.c:2686-2848).
//
//  k_block_svd    one workgroup per block of the matrix: one-sided (Hestenes) Jacobi with the block
//                 (M x N, M >= N, column-major; a wide block is transposed on load) and the
//                 accumulated rotations (N x N) both held in LDS. The pairs of a sweep follow the
//                 round-robin tournament order, N/2 disjoint pairs per step, one wavefront per
//                 pair; columns whose norm falls below eps ||block||_F are set to zero (they are
//                 numerically zero, and a column in the span of the others would otherwise be
//                 projected out again and again). Sweeps are bounded: a block that has not
//                 converged reports a status, and the caller factorises that matrix on the host.
//  k_slice_gather the retained singular triplets: selected columns (rows) of the block factors,
//                 optionally scaled by their singular values, into the two new site tensors.
//
// The singular values go to the host (a few thousand doubles), where the truncation
// (retained_bond_indices) is decided exactly as for the host path.

#include <algorithm>
#include <mutex>
#include <numeric>

#include "ct_internal.h"

namespace ctg {

constexpr int SVD_THREADS = 256;
constexpr int SVD_MAX_SWEEPS = 60;

__device__ __forceinline__ double wave_sum(double x)
{
	// butterfly: every lane ends with the same (bitwise identical) sum
	#pragma unroll
	for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
	return x;
}

__global__ __launch_bounds__(SVD_THREADS) void k_block_svd(const DevSvdJob* __restrict__ jobs, const double* __restrict__ a, double* __restrict__ u,
	double* __restrict__ vh, double* __restrict__ s, int* __restrict__ status)
{
	extern __shared__ double lds[];
	__shared__ double red[SVD_THREADS / 64];
	__shared__ int rotated;
	const DevSvdJob jb = jobs[blockIdx.x];
	const int r = jb.r, c = jb.c;
	const bool tall = (r >= c);
	const int M = tall ? r : c, N = tall ? c : r;
	double* G = lds;                    // M x N, column-major
	double* V = lds + (size_t)M * N;    // N x N, column-major
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const double* src = a + jb.a;
	// load: G = block (tall) or its transpose (wide); both read the row-major block contiguously
	double part = 0;
	for (int e = tid; e < r * c; e += SVD_THREADS) {
		const int i = e / c, j = e - i * c;
		const double x = src[e];
		if (tall) G[i + (size_t)M * j] = x;
		else G[j + (size_t)M * i] = x;
		part += x * x;
	}
	for (int e = tid; e < N * N; e += SVD_THREADS) {
		const int i = e % N, j = e / N;
		V[e] = (i == j) ? 1.0 : 0.0;
	}
	part = wave_sum(part);
	if (lane == 0) red[wave] = part;
	if (tid == 0) rotated = 0;
	__syncthreads();
	const double fro2 = (red[0] + red[1]) + (red[2] + red[3]);
	constexpr double eps = 2.2204460492503131e-16;
	const double small2 = fro2 * eps * eps;
	const double tol = eps * fmax(sqrt((double)M), 4.0);
	const int np = N + (N & 1);   // players of the tournament (one dummy when N is odd)
	int st = -1;
	for (int sweep = 0; sweep < SVD_MAX_SWEEPS; sweep++) {
		for (int step = 0; step < np - 1; step++) {
			for (int k = wave; k < np / 2; k += SVD_THREADS / 64) {
				int pa, pb;
				if (k == 0) { pa = 0; pb = 1 + step % (np - 1); }
				else { pa = 1 + (step + k) % (np - 1); pb = 1 + (step - k + np - 1) % (np - 1); }
				const int p = min(pa, pb), q = max(pa, pb);
				if (q >= N) continue;
				double* gp = G + (size_t)M * p;
				double* gq = G + (size_t)M * q;
				double al = 0, be = 0, ga = 0;
				for (int i = lane; i < M; i += 64) {
					const double x = gp[i], y = gq[i];
					al += x * x;
					be += y * y;
					ga += x * y;
				}
				al = wave_sum(al);
				be = wave_sum(be);
				ga = wave_sum(ga);
				if (al <= small2 || be <= small2) {
					if (al != 0 && al <= small2) for (int i = lane; i < M; i += 64) gp[i] = 0;
					if (be != 0 && be <= small2) for (int i = lane; i < M; i += 64) gq[i] = 0;
					continue;
				}
				if (fabs(ga) <= tol * (sqrt(al) * sqrt(be))) continue;
				const double zeta = (be - al) / (2 * ga);
				const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1 + zeta * zeta));
				const double cs = 1 / sqrt(1 + t * t), sn = cs * t;
				for (int i = lane; i < M; i += 64) {
					const double x = gp[i], y = gq[i];
					gp[i] = cs * x - sn * y;
					gq[i] = sn * x + cs * y;
				}
				double* vp = V + (size_t)N * p;
				double* vq = V + (size_t)N * q;
				for (int i = lane; i < N; i += 64) {
					const double x = vp[i], y = vq[i];
					vp[i] = cs * x - sn * y;
					vq[i] = sn * x + cs * y;
				}
				if (lane == 0) rotated = 1;
			}
			__syncthreads();
		}
		const int any = rotated;
		__syncthreads();
		if (tid == 0) rotated = 0;
		__syncthreads();
		if (!any) { st = 0; break; }
	}
	if (tid == 0) status[blockIdx.x] = st;
	// singular values (column norms, in column order) and the factors:
	//   tall: block = (G / s) s V^T -> u = G / s (r x N), vh = V^T (N x c)
	//   wide: block^T = (G / s) s V^T -> u = V (r x N), vh = (G / s)^T (N x c)
	double* ub = u + jb.u;
	double* vb = vh + jb.vh;
	for (int l = wave; l < N; l += SVD_THREADS / 64) {
		const double* gl = G + (size_t)M * l;
		double x = 0;
		for (int i = lane; i < M; i += 64) x += gl[i] * gl[i];
		const double sl = sqrt(wave_sum(x));
		if (lane == 0) s[jb.s + l] = sl;
		const double inv = (sl > 0) ? 1.0 / sl : 0.0;
		if (tall) {
			for (int i = lane; i < M; i += 64) ub[(size_t)i * N + l] = gl[i] * inv;
			for (int j = lane; j < N; j += 64) vb[(size_t)l * c + j] = V[j + (size_t)N * l];
		}
		else {
			for (int i = lane; i < N; i += 64) ub[(size_t)i * N + l] = V[i + (size_t)N * l];
			for (int j = lane; j < M; j += 64) vb[(size_t)l * c + j] = gl[j] * inv;
		}
	}
}

// Gram-eigenvector factorisation of one block per workgroup (the default device split): with the
// isometric factor on side n (the rows when 'left', the columns otherwise), G = A A^T (n = r) or
// A^T A (n = c) is formed in LDS and diagonalised by cyclic Jacobi rotations in round-robin
// tournament order (n/2 disjoint pairs per step: row rotations, then column rotations, then the
// eigenvector columns), until a sweep rotates nothing. The eigenpairs are sorted by descending
// eigenvalue (stable) and the leading k = min(r, c) kept: the isometric factor is the eigenvectors,
// the other factor their projection (Z^T A or A Z), which already carries the singular values (so
// nothing is divided by a singular value), and s the norms of the projections, in eigenvalue order
// (the caller sorts them). LDS: G (n x n, row-major) and Z (n x n, column p at Z + p n).
__global__ __launch_bounds__(SVD_THREADS) void k_block_eig(const DevSvdJob* __restrict__ jobs, const double* __restrict__ a, double* __restrict__ u,
	double* __restrict__ vh, double* __restrict__ s, int* __restrict__ status, const int left)
{
	extern __shared__ double lds[];
	__shared__ double red[SVD_THREADS / 64];
	__shared__ double rc[SVD_THREADS / 2], rs[SVD_THREADS / 2];
	__shared__ int rp[SVD_THREADS / 2], rq[SVD_THREADS / 2];
	__shared__ int rotated;
	const DevSvdJob jb = jobs[blockIdx.x];
	const int r = jb.r, c = jb.c;
	const int n = left ? r : c, m = left ? c : r;   // Gram side, contracted side
	const int k = min(r, c);
	double* G = lds;
	double* Z = lds + (size_t)n * n;
	int* sel = reinterpret_cast<int*>(lds + 2 * (size_t)n * n);
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const double* A = a + jb.a;
	// G = A A^T (left) or A^T A: entry (i, j), i <= j, mirrored; Frobenius norm of G
	double part = 0;
	for (int e = tid; e < n * n; e += SVD_THREADS) {
		const int i = e / n, j = e - i * n;
		if (i > j) continue;
		double x = 0;
		if (left) for (int t = 0; t < m; t++) x += A[(size_t)i * c + t] * A[(size_t)j * c + t];
		else for (int t = 0; t < m; t++) x += A[(size_t)t * c + i] * A[(size_t)t * c + j];
		G[i * n + j] = x;
		G[j * n + i] = x;
		part += (i == j ? 1.0 : 2.0) * x * x;
	}
	for (int e = tid; e < n * n; e += SVD_THREADS) Z[e] = (e % (n + 1) == 0) ? 1.0 : 0.0;
	part = wave_sum(part);
	if (lane == 0) red[wave] = part;
	if (tid == 0) rotated = 0;
	__syncthreads();
	const double fro = sqrt((red[0] + red[1]) + (red[2] + red[3]));
	constexpr double eps = 2.2204460492503131e-16;
	// a pair is rotated while its coupling exceeds eps relative to its diagonal and eps ||G||_F, the
	// absolute accuracy of the eigenvalues (below it, rotations among numerically zero eigenvalues
	// only stir the null space and never settle)
	const double floor_abs = eps * fro;
	const int np = n + (n & 1);   // players of the tournament (one dummy when n is odd)
	const int npair = np / 2;
	int st = -1;
	if (n == 1) st = 0;
	for (int sweep = 0; sweep < SVD_MAX_SWEEPS && st != 0; sweep++) {
		for (int step = 0; step < np - 1; step++) {
			// rotation angles of this step's disjoint pairs (one thread per pair)
			for (int kk = tid; kk < npair; kk += SVD_THREADS) {
				int pa, pb;
				if (kk == 0) { pa = 0; pb = 1 + step % (np - 1); }
				else { pa = 1 + (step + kk) % (np - 1); pb = 1 + (step - kk + np - 1) % (np - 1); }
				const int p = min(pa, pb), q = max(pa, pb);
				double cs = 1, sn = 0;
				if (q < n) {
					const double gpp = G[p * n + p], gqq = G[q * n + q], gpq = G[p * n + q];
					if (fabs(gpq) > floor_abs && fabs(gpq) > eps * sqrt(fabs(gpp) * fabs(gqq))) {
						const double zeta = (gqq - gpp) / (2 * gpq);
						const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1 + zeta * zeta));
						cs = 1 / sqrt(1 + t * t);
						sn = cs * t;
						rotated = 1;
					}
				}
				rp[kk] = p;
				rq[kk] = min(q, n - 1);
				rc[kk] = cs;
				rs[kk] = (q < n) ? sn : 0.0;
			}
			__syncthreads();
			// rows p, q of G and columns p, q of Z
			for (int e = tid; e < npair * n; e += SVD_THREADS) {
				const int kk = e / n, j = e - kk * n;
				const double cs = rc[kk], sn = rs[kk];
				if (sn == 0) continue;
				const int p = rp[kk], q = rq[kk];
				const double gp = G[p * n + j], gq = G[q * n + j];
				G[p * n + j] = cs * gp - sn * gq;
				G[q * n + j] = sn * gp + cs * gq;
				const double zp = Z[p * n + j], zq = Z[q * n + j];
				Z[p * n + j] = cs * zp - sn * zq;
				Z[q * n + j] = sn * zp + cs * zq;
			}
			__syncthreads();
			// columns p, q of G
			for (int e = tid; e < npair * n; e += SVD_THREADS) {
				const int kk = e / n, i = e - kk * n;
				const double cs = rc[kk], sn = rs[kk];
				if (sn == 0) continue;
				const int p = rp[kk], q = rq[kk];
				const double gp = G[i * n + p], gq = G[i * n + q];
				G[i * n + p] = cs * gp - sn * gq;
				G[i * n + q] = sn * gp + cs * gq;
			}
			__syncthreads();
		}
		const int any = rotated;
		__syncthreads();
		if (tid == 0) rotated = 0;
		__syncthreads();
		if (!any) st = 0;
	}
	if (tid == 0) status[blockIdx.x] = st;
	// stable descending order of the eigenvalues; the leading k are kept
	for (int i = tid; i < n; i += SVD_THREADS) {
		const double li = G[i * n + i];
		int rank = 0;
		for (int j = 0; j < n; j++) {
			const double lj = G[j * n + j];
			rank += (lj > li || (lj == li && j < i)) ? 1 : 0;
		}
		if (rank < k) sel[rank] = i;
	}
	__syncthreads();
	double* ub = u + jb.u;
	double* vb = vh + jb.vh;
	// one wavefront per kept eigenvector l: the isometric factor is its column of Z, the other factor
	// its projection, and s_l the norm of that projection (accurate to eps ||a||, where sqrt(lambda_l)
	// is only accurate to sqrt(eps) ||a|| for numerically zero eigenvalues)
	for (int l = wave; l < k; l += SVD_THREADS / 64) {
		const double* zl = Z + sel[l] * n;
		double nrm = 0;
		if (left) {
			// u[:, l] = Z[:, sel l] (r entries); vh[l, :] = Z[:, sel l]^T A (c entries)
			for (int i = lane; i < r; i += 64) ub[(size_t)i * k + l] = zl[i];
			for (int j = lane; j < c; j += 64) {
				double x = 0;
				for (int i = 0; i < r; i++) x += zl[i] * A[(size_t)i * c + j];
				vb[(size_t)l * c + j] = x;
				nrm += x * x;
			}
		}
		else {
			// vh[l, :] = Z[:, sel l]^T (c entries); u[:, l] = A Z[:, sel l] (r entries)
			for (int j = lane; j < c; j += 64) vb[(size_t)l * c + j] = zl[j];
			for (int i = lane; i < r; i += 64) {
				double x = 0;
				for (int j = 0; j < c; j++) x += A[(size_t)i * c + j] * zl[j];
				ub[(size_t)i * k + l] = x;
				nrm += x * x;
			}
		}
		nrm = wave_sum(nrm);
		if (lane == 0) s[jb.s + l] = sqrt(nrm);
	}
}

__global__ __launch_bounds__(256) void k_slice_gather(const SliceBlock* __restrict__ blocks, const int* __restrict__ idx, const double* __restrict__ scale,
	const double* __restrict__ src, double* __restrict__ dst)
{
	const SliceBlock b = blocks[blockIdx.x];
	const int n = b.rows * b.cols;
	for (int e = threadIdx.x; e < n; e += blockDim.x) {
		const int i = e / b.cols, j = e - i * b.cols;
		double v;
		if (b.axis == 1) v = src[b.src + (int64_t)i * b.src_ld + idx[b.idx0 + j]] * (scale ? scale[b.idx0 + j] : 1.0);
		else v = src[b.src + (int64_t)idx[b.idx0 + i] * b.src_ld + j] * (scale ? scale[b.idx0 + i] : 1.0);
		dst[b.dst + e] = v;
	}
}

size_t block_svd_lds_bytes(const SvdLayout& lay)
{
	size_t mx = 0;
	for (const SvdLayout::Job& j : lay.jobs) {
		const size_t M = (size_t)std::max(j.r, j.c), N = (size_t)std::min(j.r, j.c);
		mx = std::max(mx, (M * N + N * N) * sizeof(double));
	}
	return mx;
}

size_t block_svd_lds_limit()
{
	static size_t lim = 0;
	static std::once_flag once;
	std::call_once(once, [] {
		int dev = 0, v = 0;
		CTG_HIP(hipGetDevice(&dev));
		CTG_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
		// the kernel's static LDS (reduction slots, flag) comes on top of the dynamic allocation
		lim = (size_t)v > 1024 ? (size_t)v - 1024 : 0;
		CTG_HIP(hipFuncSetAttribute((const void*)k_block_svd, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim));
	});
	return lim;
}

bool launch_block_svd(const SvdLayout& lay, const double* a, double* u, double* vh, double* s, int* status, hipStream_t st)
{
	if (lay.jobs.empty()) return true;
	const size_t lds = block_svd_lds_bytes(lay);
	if (lds > block_svd_lds_limit()) return false;
	std::vector<DevSvdJob> h(lay.jobs.size());
	for (size_t i = 0; i < h.size(); i++) {
		const SvdLayout::Job& j = lay.jobs[i];
		if (j.r > (1 << 20) || j.c > (1 << 20)) return false;
		h[i] = DevSvdJob{ j.a_off, j.u_off, j.vh_off, j.s_off, (int32_t)j.r, (int32_t)j.c };
	}
	// largest blocks first
	std::stable_sort(h.begin(), h.end(), [](const DevSvdJob& x, const DevSvdJob& y) {
		return (int64_t)std::max(x.r, x.c) * std::min(x.r, x.c) > (int64_t)std::max(y.r, y.c) * std::min(y.r, y.c);
	});
	DevSvdJob* d = (DevSvdJob*)dev_alloc(h.size() * sizeof(DevSvdJob));
	copy_h2d_async(d, h.data(), h.size() * sizeof(DevSvdJob), st);
	hipLaunchKernelGGL(k_block_svd, dim3((unsigned)h.size()), dim3(SVD_THREADS), lds, st, d, a, u, vh, s, status);
	CTG_HIP(hipGetLastError());
	dev_free(d);
	return true;
}

size_t block_eig_lds_bytes(const SvdLayout& lay, bool left)
{
	size_t mx = 0;
	for (const SvdLayout::Job& j : lay.jobs) {
		const size_t n = (size_t)(left ? j.r : j.c);
		mx = std::max(mx, 2 * n * n * sizeof(double) + (n + 1) * sizeof(int));
	}
	return (mx + 15) / 16 * 16;
}

size_t block_eig_lds_limit()
{
	static size_t lim = 0;
	static std::once_flag once;
	std::call_once(once, [] {
		int dev = 0, v = 0;
		CTG_HIP(hipGetDevice(&dev));
		CTG_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
		// the kernel's static LDS (reduction slots, rotation tables, flag) comes on top
		lim = (size_t)v > 8192 ? (size_t)v - 8192 : 0;
		CTG_HIP(hipFuncSetAttribute((const void*)k_block_eig, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim));
	});
	return lim;
}

bool launch_block_eig(const SvdLayout& lay, const double* a, double* u, double* vh, double* s, int* status, bool left, hipStream_t st)
{
	if (lay.jobs.empty()) return true;
	const size_t lds = block_eig_lds_bytes(lay, left);
	if (lds > block_eig_lds_limit()) return false;
	std::vector<DevSvdJob> h(lay.jobs.size());
	for (size_t i = 0; i < h.size(); i++) {
		const SvdLayout::Job& j = lay.jobs[i];
		// the kernel's pair tables hold SVD_THREADS / 2 pairs: n <= SVD_THREADS
		if ((left ? j.r : j.c) > SVD_THREADS || j.r > (1 << 20) || j.c > (1 << 20)) return false;
		h[i] = DevSvdJob{ j.a_off, j.u_off, j.vh_off, j.s_off, (int32_t)j.r, (int32_t)j.c };
	}
	// largest Gram matrices first
	std::stable_sort(h.begin(), h.end(), [left](const DevSvdJob& x, const DevSvdJob& y) { return (left ? x.r : x.c) > (left ? y.r : y.c); });
	DevSvdJob* d = (DevSvdJob*)dev_alloc(h.size() * sizeof(DevSvdJob));
	copy_h2d_async(d, h.data(), h.size() * sizeof(DevSvdJob), st);
	hipLaunchKernelGGL(k_block_eig, dim3((unsigned)h.size()), dim3(SVD_THREADS), lds, st, d, a, u, vh, s, status, left ? 1 : 0);
	CTG_HIP(hipGetLastError());
	dev_free(d);
	return true;
}

void plan_slice_2d(const Meta& m, int i_ax, const std::vector<ct_long>& ind, const std::vector<ct_long>& srcpos, const std::vector<double>* scale,
	Meta& out, std::vector<SliceBlock>& blocks, std::vector<int>& idx, std::vector<double>& sc)
{
	std::vector<qnumber> qs(ind.size());
	for (size_t j = 0; j < ind.size(); j++) qs[j] = m.qnums_logical[i_ax][ind[j]];
	ct_long dl[2] = { m.dim_logical[0], m.dim_logical[1] };
	dl[i_ax] = (ct_long)ind.size();
	const qnumber* qn[2] = { i_ax == 0 ? qs.data() : m.qnums_logical[0].data(), i_ax == 1 ? qs.data() : m.qnums_logical[1].data() };
	out = make_meta(2, dl, m.axis_dir.data(), qn);
	blocks.clear();
	idx.clear();
	sc.clear();
	ct_long ir[2], it[2], br[2], bt[2];
	for (ct_long k = 0; k < out.nblocks; k++) {
		if (!out.present(k)) continue;
		out.block_index(k, ir);
		out.block_dims(k, br);
		const qnumber q = out.qnums_blocks[i_ax][ir[i_ax]];
		it[0] = ir[0];
		it[1] = ir[1];
		it[i_ax] = m.find_qnum(i_ax, q);
		const ct_long kt = m.flat_index(it);
		m.block_dims(kt, bt);
		const std::vector<ct_long> fo = fanout(out, i_ax, q);
		SliceBlock b;
		b.src = m.offset[kt];
		b.dst = out.offset[k];
		b.rows = (int32_t)br[0];
		b.cols = (int32_t)br[1];
		b.src_ld = (int32_t)bt[1];
		b.axis = i_ax;
		b.idx0 = (int32_t)idx.size();
		b.pad = 0;
		for (ct_long j = 0; j < br[i_ax]; j++) {
			const ct_long logical = ind[fo[j]];
			idx.push_back((int)srcpos[logical]);
			sc.push_back(scale ? (*scale)[fo[j]] : 1.0);
		}
		blocks.push_back(b);
	}
}

DevTensor device_slice_2d(const DevTensor& src, const Meta& out, const std::vector<SliceBlock>& blocks, const std::vector<int>& idx,
	const std::vector<double>& sc, bool scaled, hipStream_t st)
{
	DevTensor r(out);
	if (blocks.empty()) return r;
	SliceBlock* db = (SliceBlock*)dev_alloc(blocks.size() * sizeof(SliceBlock));
	int* di = (int*)dev_alloc(std::max<size_t>(1, idx.size()) * sizeof(int));
	double* ds = (double*)dev_alloc(std::max<size_t>(1, sc.size()) * sizeof(double));
	copy_h2d_async(db, blocks.data(), blocks.size() * sizeof(SliceBlock), st);
	if (!idx.empty()) copy_h2d_async(di, idx.data(), idx.size() * sizeof(int), st);
	if (!sc.empty()) copy_h2d_async(ds, sc.data(), sc.size() * sizeof(double), st);
	hipLaunchKernelGGL(k_slice_gather, dim3((unsigned)blocks.size()), dim3(256), 0, st, db, di, scaled ? ds : nullptr, src.d, r.d);
	CTG_HIP(hipGetLastError());
	dev_free(db);
	dev_free(di);
	dev_free(ds);
	return r;
}

}  // namespace ctg

// one dense matrix through the device block SVD (a test and utility entry point)
extern "C" int ct_dense_svd_device(const ct_long m, const ct_long n, const double* a, double* u, double* s, double* vh)
{
	using namespace ctg;
	if (m < 1 || n < 1 || !a || !u || !s || !vh) return CT_ERR_ARG;
	try {
		const ct_long k = std::min(m, n);
		SvdLayout lay;
		lay.jobs.push_back(SvdLayout::Job{ 0, 0, 0, 0, m, n });
		hipStream_t st = default_stream();
		double* da = (double*)dev_alloc((size_t)(m * n) * sizeof(double));
		double* du = (double*)dev_alloc((size_t)(m * k) * sizeof(double));
		double* dv = (double*)dev_alloc((size_t)(k * n) * sizeof(double));
		double* ds = (double*)dev_alloc((size_t)k * sizeof(double));
		int* dst = (int*)dev_alloc(sizeof(int));
		copy_h2d_async(da, a, (size_t)(m * n) * sizeof(double), st);
		int status = -1;
		const bool ok = launch_block_svd(lay, da, du, dv, ds, dst, st);
		if (ok) {
			CTG_HIP(hipMemcpyAsync(u, du, (size_t)(m * k) * sizeof(double), hipMemcpyDeviceToHost, st));
			CTG_HIP(hipMemcpyAsync(vh, dv, (size_t)(k * n) * sizeof(double), hipMemcpyDeviceToHost, st));
			CTG_HIP(hipMemcpyAsync(s, ds, (size_t)k * sizeof(double), hipMemcpyDeviceToHost, st));
			CTG_HIP(hipMemcpyAsync(&status, dst, sizeof(int), hipMemcpyDeviceToHost, st));
		}
		CTG_HIP(hipStreamSynchronize(st));
		dev_free(da); dev_free(du); dev_free(dv); dev_free(ds); dev_free(dst);
		if (!ok) return CT_ERR_UNSUPPORTED;
		// descending singular values (stable), columns of u and rows of vh alike
		std::vector<ct_long> ord((size_t)k);
		std::iota(ord.begin(), ord.end(), 0);
		std::stable_sort(ord.begin(), ord.end(), [&](ct_long x, ct_long y) { return s[x] > s[y]; });
		std::vector<double> s2((size_t)k), u2((size_t)(m * k)), v2((size_t)(k * n));
		for (ct_long l = 0; l < k; l++) {
			s2[l] = s[ord[l]];
			for (ct_long i = 0; i < m; i++) u2[(size_t)(i * k + l)] = u[i * k + ord[l]];
			for (ct_long j = 0; j < n; j++) v2[(size_t)(l * n + j)] = vh[ord[l] * n + j];
		}
		std::copy(s2.begin(), s2.end(), s);
		std::copy(u2.begin(), u2.end(), u);
		std::copy(v2.begin(), v2.end(), vh);
		return status;
	}
	catch (const HipError&) {
		return CT_ERR_HIP;
	}
}

// one dense matrix through the device Gram-eigenvector factorisation of the DMRG split (a test and
// utility entry point): left != 0: a = u (isometric, m x k) * vh (k x n, carrying s); left == 0:
// a = u (m x k, carrying s) * vh (isometric); k = min(m, n), s descending
extern "C" int ct_dense_gram_factor_device(const ct_long m, const ct_long n, const double* a, const int left, double* u, double* s, double* vh)
{
	using namespace ctg;
	if (m < 1 || n < 1 || !a || !u || !s || !vh) return CT_ERR_ARG;
	try {
		const ct_long k = std::min(m, n);
		SvdLayout lay;
		lay.jobs.push_back(SvdLayout::Job{ 0, 0, 0, 0, m, n });
		hipStream_t st = default_stream();
		double* da = (double*)dev_alloc((size_t)(m * n) * sizeof(double));
		double* du = (double*)dev_alloc((size_t)(m * k) * sizeof(double));
		double* dv = (double*)dev_alloc((size_t)(k * n) * sizeof(double));
		double* ds = (double*)dev_alloc((size_t)k * sizeof(double));
		int* dst = (int*)dev_alloc(sizeof(int));
		copy_h2d_async(da, a, (size_t)(m * n) * sizeof(double), st);
		int status = -1;
		const bool ok = launch_block_eig(lay, da, du, dv, ds, dst, left != 0, st);
		if (ok) {
			CTG_HIP(hipMemcpyAsync(u, du, (size_t)(m * k) * sizeof(double), hipMemcpyDeviceToHost, st));
			CTG_HIP(hipMemcpyAsync(vh, dv, (size_t)(k * n) * sizeof(double), hipMemcpyDeviceToHost, st));
			CTG_HIP(hipMemcpyAsync(s, ds, (size_t)k * sizeof(double), hipMemcpyDeviceToHost, st));
			CTG_HIP(hipMemcpyAsync(&status, dst, sizeof(int), hipMemcpyDeviceToHost, st));
		}
		CTG_HIP(hipStreamSynchronize(st));
		dev_free(da); dev_free(du); dev_free(dv); dev_free(ds); dev_free(dst);
		if (!ok) return CT_ERR_UNSUPPORTED;
		// descending s (stable), columns of u and rows of vh alike
		std::vector<ct_long> ord((size_t)k);
		std::iota(ord.begin(), ord.end(), 0);
		std::stable_sort(ord.begin(), ord.end(), [&](ct_long x, ct_long y) { return s[x] > s[y]; });
		std::vector<double> s2((size_t)k), u2((size_t)(m * k)), v2((size_t)(k * n));
		for (ct_long l = 0; l < k; l++) {
			s2[l] = s[ord[l]];
			for (ct_long i = 0; i < m; i++) u2[(size_t)(i * k + l)] = u[i * k + ord[l]];
			for (ct_long j = 0; j < n; j++) v2[(size_t)(l * n + j)] = vh[ord[l] * n + j];
		}
		std::copy(s2.begin(), s2.end(), s);
		std::copy(u2.begin(), u2.end(), u);
		std::copy(v2.begin(), v2.end(), vh);
		return status;
	}
	catch (const HipError&) {
		return CT_ERR_HIP;
	}
}
