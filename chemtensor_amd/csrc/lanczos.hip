// Device-resident Lanczos eigensystem (src/util/krylov.c:24-89 and :172-251):
// three-term recurrence without re-orthogonalisation, breakdown test beta < 100 n eps, all
// vectors and the alpha / beta scalars on the GPU in FP64. Reductions are two-pass with a
// fixed summation order (deterministic). The tridiagonal eigenproblem (at most maxiter x
// maxiter) is solved on the host in FP64; the Ritz vectors are formed on the device.

#include <cmath>
#include <cstring>
#include <algorithm>
#include <limits>

#include "ct_internal.h"

namespace ctg {

constexpr int LZ_THREADS = 256;

// sum over the workgroup: butterfly within each wavefront, then the four wavefront sums in order
// (the same value in every thread, the same order in every launch)
__device__ __forceinline__ double block_sum(double v, double* sh)
{
	static_assert(LZ_THREADS == 256, "four wavefronts");
	#pragma unroll
	for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
	if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
	__syncthreads();
	const double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
	__syncthreads();
	return r;
}

// sum of the G partials in a fixed order, identical in every block
__device__ __forceinline__ double sum_partials(const double* __restrict__ part, int G, double* sh)
{
	double v = 0;
	for (int b = threadIdx.x; b < G; b += LZ_THREADS) v += part[b];
	return block_sum(v, sh);
}

// the sums of two partial arrays (each as sum_partials would give it), their loads in one round trip
__device__ __forceinline__ void sum_partials2(const double* __restrict__ pa, int Ga, const double* __restrict__ pb, int Gb, double* sh, double& ra, double& rb)
{
	// each thread sums entries threadIdx.x + k LZ_THREADS in k order, four loads in flight at a time
	// (entries past the ends add zeros, which leave the sums unchanged)
	double a = 0, b = 0;
	for (int i0 = threadIdx.x; i0 < Ga || i0 < Gb; i0 += 4 * LZ_THREADS) {
		double va[4], vb[4];
		#pragma unroll
		for (int u = 0; u < 4; u++) {
			const int i = i0 + u * LZ_THREADS;
			va[u] = i < Ga ? pa[i] : 0.0;
			vb[u] = i < Gb ? pb[i] : 0.0;
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
	if ((threadIdx.x & 63) == 0) {
		sh[threadIdx.x >> 6] = a;
		sh[4 + (threadIdx.x >> 6)] = b;
	}
	__syncthreads();
	ra = (sh[0] + sh[1]) + (sh[2] + sh[3]);
	rb = (sh[4] + sh[5]) + (sh[6] + sh[7]);
	__syncthreads();
}

// scal layout: [0, maxiter) alpha, [maxiter, 2 maxiter) beta, [2 maxiter] numiter, [2 maxiter + 1] stop flag
__global__ __launch_bounds__(LZ_THREADS) void k_lz_dot_partial(const double* __restrict__ x, const double* __restrict__ y, int64_t n,
	double* __restrict__ part, const double* __restrict__ stop)
{
	__shared__ double sh[LZ_THREADS];
	if (stop != nullptr && *stop != 0) return;
	double v = 0;
	for (int64_t i = (int64_t)blockIdx.x * LZ_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * LZ_THREADS) v += x[i] * y[i];
	const double s = block_sum(v, sh);
	if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// v0 = vstart / ||vstart||
__global__ __launch_bounds__(LZ_THREADS) void k_lz_normalize(const double* __restrict__ vs, double* __restrict__ v0, int64_t n, const double* __restrict__ part, int G)
{
	__shared__ double sh[LZ_THREADS];
	const double nrm = sqrt(sum_partials(part, G, sh));
	const double inv = 1 / nrm;
	for (int64_t i = (int64_t)blockIdx.x * LZ_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * LZ_THREADS) v0[i] = inv * vs[i];
}

// The vector-update kernels run on LZ_UPD_GRID(n) blocks; the first LZ_PRE entries of a thread are
// loaded before the partials are summed (one memory round trip instead of two), the rest after.
constexpr int LZ_PRE = 4;
static inline int lz_upd_grid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(1024, (n + LZ_THREADS - 1) / LZ_THREADS)); }

// alpha_j = sum(part); w -= alpha_j v_j + beta_{j-1} v_{j-1}; part2 = partial sums of w^2
__global__ __launch_bounds__(LZ_THREADS) void k_lz_alpha_update(double* __restrict__ w, const double* __restrict__ vj, const double* __restrict__ vjm1,
	int64_t n, int j, int maxiter, const double* __restrict__ part, int G, double* __restrict__ part2, double* __restrict__ scal)
{
	__shared__ double sh[LZ_THREADS];
	if (scal[2 * maxiter + 1] != 0) return;
	const int64_t i0 = (int64_t)blockIdx.x * LZ_THREADS + threadIdx.x, step = (int64_t)gridDim.x * LZ_THREADS;
	const bool upd = j < maxiter - 1;   // final iteration: alpha only
	double wv[LZ_PRE], a[LZ_PRE], b[LZ_PRE];
	#pragma unroll
	for (int k = 0; k < LZ_PRE; k++) {
		const int64_t i = i0 + k * step;
		const bool in = upd && i < n;
		wv[k] = in ? w[i] : 0.0;
		a[k] = in ? vj[i] : 0.0;
		b[k] = (in && j > 0) ? vjm1[i] : 0.0;
	}
	const double alpha = sum_partials(part, G, sh);
	if (blockIdx.x == 0 && threadIdx.x == 0) scal[j] = alpha;
	if (!upd) return;
	const double bprev = (j > 0) ? scal[maxiter + j - 1] : 0.0;
	double v = 0;
	#pragma unroll
	for (int k = 0; k < LZ_PRE; k++) {
		const int64_t i = i0 + k * step;
		if (i < n) {
			const double wi = (j > 0) ? wv[k] - (alpha * a[k] + bprev * b[k]) : wv[k] - alpha * a[k];
			w[i] = wi;
			v += wi * wi;
		}
	}
	for (int64_t i = i0 + LZ_PRE * step; i < n; i += step) {
		double wi = w[i];
		if (j > 0) wi -= alpha * vj[i] + bprev * vjm1[i];
		else wi -= alpha * vj[i];
		w[i] = wi;
		v += wi * wi;
	}
	const double s = block_sum(v, sh);
	if (threadIdx.x == 0) part2[blockIdx.x] = s;
}

// beta_j = ||w||; stop if beta_j < thresh, else v_{j+1} = w / beta_j
__global__ __launch_bounds__(LZ_THREADS) void k_lz_beta_scale(const double* __restrict__ w, double* __restrict__ vnext, int64_t n, int j, int maxiter,
	const double* __restrict__ part2, int G, double thresh, double* __restrict__ scal)
{
	__shared__ double sh[LZ_THREADS];
	if (scal[2 * maxiter + 1] != 0) return;
	const int64_t i0 = (int64_t)blockIdx.x * LZ_THREADS + threadIdx.x, step = (int64_t)gridDim.x * LZ_THREADS;
	double wv[LZ_PRE];
	#pragma unroll
	for (int k = 0; k < LZ_PRE; k++) {
		const int64_t i = i0 + k * step;
		wv[k] = i < n ? w[i] : 0.0;
	}
	const double beta = sqrt(sum_partials(part2, G, sh));
	// every block has read the stop flag before block 0 may set it (it is only read at kernel entry)
	if (beta < thresh) {
		if (blockIdx.x == 0 && threadIdx.x == 0) { scal[maxiter + j] = beta; scal[2 * maxiter] = j + 1; scal[2 * maxiter + 1] = 1; }
		return;
	}
	if (blockIdx.x == 0 && threadIdx.x == 0) scal[maxiter + j] = beta;
	const double inv = 1 / beta;
	#pragma unroll
	for (int k = 0; k < LZ_PRE; k++) {
		const int64_t i = i0 + k * step;
		if (i < n) vnext[i] = inv * wv[k];
	}
	for (int64_t i = i0 + LZ_PRE * step; i < n; i += step) vnext[i] = inv * w[i];
}

// One step of the recurrence with deferred normalisation (device plans): the matvec ran on the
// unnormalised u_j = beta_{j-1} v_j (u_0 = v_0), so y = beta_{j-1} H v_j and dpart holds the partial
// sums of <y, u_j> = beta_{j-1}^2 alpha_j. This kernel finalises beta_{j-1} = ||u_j|| from the previous
// step's partials (breakdown test of krylov.c:58), alpha_j, writes v_j = u_j / beta_{j-1} into the
// Krylov basis and the next unnormalised vector u_{j+1} = H v_j - alpha_j v_j - beta_{j-1} v_{j-1}
// with its partial squared norms. It takes the place of the beta-scale launch: one vector kernel per
// step beyond the matvec. uj and unext may be the same buffer (every entry is read, then written, by
// one thread).
__global__ __launch_bounds__(LZ_THREADS) void k_lz_step(const double* uj, double* unext, double* __restrict__ vj_out, const double* __restrict__ vjm1,
	const double* __restrict__ y, int64_t n, int j, int maxiter, const double* __restrict__ dpart, int Gd, const double* __restrict__ npart, int Gn,
	double thresh, double* __restrict__ npart_out, double* __restrict__ scal)
{
	__shared__ double sh[LZ_THREADS];
	if (scal[2 * maxiter + 1] != 0) return;
	const int64_t i0 = (int64_t)blockIdx.x * LZ_THREADS + threadIdx.x, step = (int64_t)gridDim.x * LZ_THREADS;
	const bool upd = j < maxiter - 1;   // final iteration: alpha and v_j only
	double uv[LZ_PRE], yv[LZ_PRE], bv[LZ_PRE];
	#pragma unroll
	for (int k = 0; k < LZ_PRE; k++) {
		const int64_t i = i0 + k * step;
		const bool in = i < n;
		uv[k] = in ? uj[i] : 0.0;
		yv[k] = (in && upd) ? y[i] : 0.0;
		bv[k] = (in && upd && j > 0) ? vjm1[i] : 0.0;
	}
	// beta^2 and <u_j, y> from their partials, loaded together
	double nsum = 1.0, dsum = 0.0;
	sum_partials2(npart, j > 0 ? Gn : 0, dpart, Gd, sh, nsum, dsum);
	double beta = 1.0;
	if (j > 0) {
		beta = sqrt(nsum);
		// every block has read the stop flag before block 0 may set it (it is only read at kernel entry)
		if (beta < thresh) {
			if (blockIdx.x == 0 && threadIdx.x == 0) { scal[maxiter + j - 1] = beta; scal[2 * maxiter] = j; scal[2 * maxiter + 1] = 1; }
			return;
		}
	}
	const double inv = 1 / beta;
	const double alpha = dsum * inv * inv;
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		scal[j] = alpha;
		if (j > 0) scal[maxiter + j - 1] = beta;
	}
	double acc = 0;
	#pragma unroll
	for (int k = 0; k < LZ_PRE; k++) {
		const int64_t i = i0 + k * step;
		if (i < n) {
			const double v = inv * uv[k];
			if (j > 0) vj_out[i] = v;
			if (upd) {
				const double wi = (j > 0) ? inv * yv[k] - (alpha * v + beta * bv[k]) : yv[k] - alpha * v;
				unext[i] = wi;
				acc += wi * wi;
			}
		}
	}
	for (int64_t i = i0 + LZ_PRE * step; i < n; i += step) {
		const double v = inv * uj[i];
		if (j > 0) vj_out[i] = v;
		if (upd) {
			const double wi = (j > 0) ? inv * y[i] - (alpha * v + beta * vjm1[i]) : y[i] - alpha * v;
			unext[i] = wi;
			acc += wi * wi;
		}
	}
	if (upd) {
		const double s2 = block_sum(acc, sh);
		if (threadIdx.x == 0) npart_out[blockIdx.x] = s2;
	}
}

// u[i, k] = sum_j V[j, i] z[j, k]  (u: n x numeig row-major, z: numiter x numeig row-major)
__global__ __launch_bounds__(LZ_THREADS) void k_lz_ritz(const double* __restrict__ V, const double* __restrict__ z, int numiter, int numeig, int64_t n, double* __restrict__ u)
{
	for (int64_t i = (int64_t)blockIdx.x * LZ_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * LZ_THREADS) {
		for (int k = 0; k < numeig; k++) {
			double acc = 0;
			for (int j = 0; j < numiter; j++) acc += V[(int64_t)j * n + i] * z[j * numeig + k];
			u[i * numeig + k] = acc;
		}
	}
}

void LanczosWork::reserve(int64_t nn, int maxiter, int numeig, int partials)
{
	if (nn != n || cap < maxiter) {
		drop_graph();
		dev_free(v);
		dev_free(w);
		dev_free(scal);
		dev_free(vs);
		dev_free(u);
		v = (double*)dev_alloc((size_t)maxiter * (size_t)nn * sizeof(double));
		w = (double*)dev_alloc((size_t)nn * sizeof(double));
		u = (double*)dev_alloc((size_t)nn * sizeof(double));
		scal = (double*)dev_alloc((size_t)(2 * maxiter + 2) * sizeof(double));
		vs = nullptr;
		n = nn;
		cap = maxiter;
	}
	partials = std::max(partials, 1024);
	if (pcap < partials) {
		drop_graph();
		dev_free(part);
		part = (double*)dev_alloc(3 * (size_t)partials * sizeof(double));
		pcap = partials;
	}
	if (zcap < maxiter * numeig) {
		dev_free(z);
		z = (double*)dev_alloc((size_t)maxiter * numeig * sizeof(double));
		zcap = maxiter * numeig;
	}
}

void LanczosWork::drop_graph()
{
	if (graph) (void)hipGraphExecDestroy(graph);
	graph = nullptr;
	graph_maxiter = 0;
}

LanczosWork::~LanczosWork()
{
	drop_graph();
	dev_free(v);
	dev_free(w);
	dev_free(u);
	dev_free(scal);
	dev_free(part);
	dev_free(z);
	dev_free(vs);
}

// v0 = vstart / |vstart|, then maxiter steps of the recurrence; only kernel launches on s. With
// sync_stop (a matvec that synchronises anyway, such as a host callback) the stop flag is read back
// before every matvec after the first, and the loop ends at a breakdown: the matvec then runs exactly
// numiter times, on normalised Krylov vectors only (krylov.c:34-68 calls afunc numiter times).
static void enqueue_recurrence(LanczosWork& lw, int64_t n, hipStream_t s, const DevMatvec& matvec, const double* vstart, int maxiter,
	bool sync_stop = false, const DevMatvecDot* mvd = nullptr, bool deferred = false)
{
	const int G = (int)std::max<int64_t>(1, std::min<int64_t>(1024, (n + 511) / 512));
	double* V = lw.v;
	double* w = lw.w;
	double* scal = lw.scal;
	double* part = lw.part;
	double* part2 = lw.part + lw.pcap;
	launch_fill(scal, 0.0, 2 * maxiter + 2, s);
	const double thresh = 100.0 * (double)n * 2.2204460492503131e-16;   // krylov.c:58
	hipLaunchKernelGGL(k_lz_dot_partial, dim3(G), dim3(LZ_THREADS), 0, s, vstart, vstart, n, part, (const double*)nullptr);
	hipLaunchKernelGGL(k_lz_normalize, dim3(G), dim3(LZ_THREADS), 0, s, vstart, V, n, part, G);
	CTG_HIP(hipGetLastError());
	if (deferred) {
		// matvec on the unnormalised u_j (u_0 = v_0); k_lz_step finalises beta_{j-1} and alpha_j, writes
		// v_j and u_{j+1}; the squared-norm partials alternate between two areas
		const int Gu = lz_upd_grid(n);
		double* nparts[2] = { lw.part + lw.pcap, lw.part + 2 * (size_t)lw.pcap };
		for (int j = 0; j < maxiter; j++) {
			const double* x = (j == 0) ? V : lw.u;
			if (mvd && mvd->step) {
				// two launches per step: the matvec's units (with the alpha partials), then the step
				// that sums y from the partial blocks as it updates
				mvd->units(x, part);
				mvd->step(x, lw.u, V + (int64_t)j * n, j > 0 ? V + (int64_t)(j - 1) * n : V, j, maxiter, (const double*)part,
					(const double*)nparts[(j + 1) & 1], thresh, nparts[j & 1], scal);
				continue;
			}
			int Ga = G;
			if (mvd) {
				mvd->apply(x, w, part);
				Ga = mvd->partials;
			}
			else {
				matvec(x, w);
				hipLaunchKernelGGL(k_lz_dot_partial, dim3(G), dim3(LZ_THREADS), 0, s, w, x, n, part, scal + 2 * maxiter + 1);
			}
			hipLaunchKernelGGL(k_lz_step, dim3(Gu), dim3(LZ_THREADS), 0, s, x, lw.u, V + (int64_t)j * n, j > 0 ? V + (int64_t)(j - 1) * n : V,
				(const double*)w, n, j, maxiter, (const double*)part, Ga, (const double*)nparts[(j + 1) & 1], Gu, thresh, nparts[j & 1], scal);
			CTG_HIP(hipGetLastError());
		}
		return;
	}
	for (int j = 0; j < maxiter; j++) {
		const double* vj = V + (int64_t)j * n;
		if (sync_stop && j > 0) {
			double stop = 0;
			CTG_HIP(hipMemcpyAsync(&stop, scal + 2 * maxiter + 1, sizeof(double), hipMemcpyDeviceToHost, s));
			CTG_HIP(hipStreamSynchronize(s));
			if (stop != 0) break;
		}
		int Ga = G;
		if (mvd) {
			// the matvec's last launch also writes the alpha partials
			mvd->apply(vj, w, part);
			Ga = mvd->partials;
		}
		else {
			matvec(vj, w);
			hipLaunchKernelGGL(k_lz_dot_partial, dim3(G), dim3(LZ_THREADS), 0, s, w, vj, n, part, scal + 2 * maxiter + 1);
		}
		const int Gu = lz_upd_grid(n);
		hipLaunchKernelGGL(k_lz_alpha_update, dim3(Gu), dim3(LZ_THREADS), 0, s, w, vj, j > 0 ? vj - n : vj, n, j, maxiter, part, Ga, part2, scal);
		if (j < maxiter - 1)
			hipLaunchKernelGGL(k_lz_beta_scale, dim3(Gu), dim3(LZ_THREADS), 0, s, w, V + (int64_t)(j + 1) * n, n, j, maxiter, part2, Gu, thresh, scal);
		CTG_HIP(hipGetLastError());
	}
}

int lanczos_eigensystem(LanczosWork& lw, int64_t n, hipStream_t s, const DevMatvec& matvec, const double* vstart,
	int maxiter, int numeig, double* lambda, double* u, bool use_graph, bool sync_stop, const DevMatvecDot* matvec_dot, bool deferred)
{
	if (sync_stop) deferred = false;   // the host callback sees normalised Krylov vectors
	if (maxiter < 1 || numeig < 1 || numeig > maxiter || n < 1) return CT_ERR_ARG;
	if (matvec_dot && matvec_dot->partials < 1) matvec_dot = nullptr;
	lw.reserve(n, maxiter, numeig, matvec_dot ? std::max(matvec_dot->partials, matvec_dot->step_partials) : 1024);
	if (use_graph) {
		if (!lw.vs) lw.vs = (double*)dev_alloc((size_t)n * sizeof(double));
		launch_copy(vstart, lw.vs, n, s);
		if (!lw.graph || lw.graph_maxiter != maxiter) {
			lw.drop_graph();
			hipGraph_t g = nullptr;
			CaptureScope cs;
			CTG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
			try {
				enqueue_recurrence(lw, n, s, matvec, lw.vs, maxiter, false, matvec_dot, deferred);
			}
			catch (...) {
				(void)hipStreamEndCapture(s, &g);
				if (g) (void)hipGraphDestroy(g);
				throw;
			}
			CTG_HIP(hipStreamEndCapture(s, &g));
			const hipError_t e = hipGraphInstantiate(&lw.graph, g, nullptr, nullptr, 0);
			(void)hipGraphDestroy(g);
			CTG_HIP(e);
			lw.graph_maxiter = maxiter;
		}
		CTG_HIP(hipGraphLaunch(lw.graph, s));
	}
	else enqueue_recurrence(lw, n, s, matvec, vstart, maxiter, sync_stop, matvec_dot, deferred);
	double* V = lw.v;
	std::vector<double> hs((size_t)(2 * maxiter + 2));
	CTG_HIP(hipMemcpyAsync(hs.data(), lw.scal, hs.size() * sizeof(double), hipMemcpyDeviceToHost, s));
	CTG_HIP(hipStreamSynchronize(s));
	const int numiter = (hs[2 * maxiter + 1] != 0) ? (int)hs[2 * maxiter] : maxiter;
	if (numiter < numeig) {
		fprintf(stderr, "Lanczos iteration stopped after %i iterations, cannot compute %i eigenvalues\n", numiter, numeig);
		return -1;
	}
	std::vector<double> d(hs.begin(), hs.begin() + numiter);
	std::vector<double> e((size_t)numiter + 1, 0.0);
	for (int i = 0; i < numiter - 1; i++) e[i] = hs[maxiter + i];
	std::vector<double> z((size_t)numiter * numiter);   // column k (z[k numiter : (k+1) numiter]) = eigenvector k
	if (host_tridiag_eigh(numiter, d.data(), e.data(), z.data()) != 0) {
		fprintf(stderr, "tridiagonal eigensolver failed\n");
		return -2;
	}
	for (int k = 0; k < numeig; k++) lambda[k] = d[k];
	std::vector<double> zr((size_t)numiter * numeig);
	for (int j = 0; j < numiter; j++)
		for (int k = 0; k < numeig; k++) zr[(size_t)j * numeig + k] = z[(size_t)k * numiter + j];
	CTG_HIP(hipMemcpyAsync(lw.z, zr.data(), zr.size() * sizeof(double), hipMemcpyHostToDevice, s));
	const int gr = (int)std::max<int64_t>(1, std::min<int64_t>(4096, (n + LZ_THREADS - 1) / LZ_THREADS));
	hipLaunchKernelGGL(k_lz_ritz, dim3(gr), dim3(LZ_THREADS), 0, s, V, lw.z, numiter, numeig, n, u);
	CTG_HIP(hipGetLastError());
	CTG_HIP(hipStreamSynchronize(s));
	return 0;
}

// device plans normalise each Lanczos vector inside the next step's update (CT_LANCZOS_DEFERRED=0:
// the separate beta-scale launch)
static bool lanczos_deferred()
{
	const char* e = getenv("CT_LANCZOS_DEFERRED");
	return !(e && e[0] == '0');
}

static bool graph_enabled()
{
	const char* e = getenv("CT_LANCZOS_GRAPH");
	return !(e && e[0] == '0');
}

// sector-pair plans fold the partial-block sum into the Lanczos step (CT_PAIR_LZ_FUSED=0: the
// matvec's own partial-sum launch and the generic step)
static bool pair_lz_fused()
{
	const char* e = getenv("CT_PAIR_LZ_FUSED");
	return !(e && e[0] == '0');
}

int lanczos_ground_state(MatvecPlan& p, const double* vstart, int maxiter, double* lambda, double* u)
{
	// the plain form replays thousands of per-block launches; it stays eager
	const bool graph = p.mode != CT_MATVEC_PLAIN_FP64 && graph_enabled();
	// the sector-pair form computes the alpha partials in its partial-sum launch
	DevMatvecDot mvd;
	if (p.mode == CT_MATVEC_PAIR_FP64) {
		mvd.partials = pair_dot_partials(p.pair);
		mvd.apply = [&p](const double* x, double* y, double* part) { launch_pair_matvec_dot(p.pair, x, p.r.d, p.l.d, y, part, p.stream); };
		if (pair_lz_fused()) {
			// deferred recurrence: two launches per step (units with the alpha partials, then the step)
			mvd.step_partials = pair_step_partials(p.pair);
			mvd.units = [&p](const double* x, double* part) { launch_pair_units(p.pair, x, p.r.d, p.l.d, p.stream, part); };
			mvd.step = [&p](const double* uj, double* unext, double* vj_out, const double* vjm1, int j, int maxiter, const double* dpart,
				const double* npart, double thresh, double* npart_out, double* scal) {
				launch_pair_lz_step(p.pair, uj, unext, vj_out, vjm1, j, maxiter, dpart, npart, thresh, npart_out, scal, p.stream);
			};
		}
	}
	return lanczos_eigensystem(p.lz, p.theta_meta.nelem, p.stream,
		[&p](const double* x, double* y) { matvec_apply(p, x, y, nullptr); }, vstart, maxiter, 1, lambda, u, graph, false,
		p.mode == CT_MATVEC_PAIR_FP64 ? &mvd : nullptr, lanczos_deferred());
}

//________________________________________________________________________________________________________________________
//
// batched Lanczos: one recurrence per problem over its segment of the concatenated vectors. Each
// reduction block covers a range of one problem; every block of a problem sums that problem's
// partials in the same order. Per-problem scalars: scal[p * S + ...] with S = 2 maxiter + 2.

__device__ __forceinline__ double sum_range(const double* __restrict__ part, int b0, int b1, double* sh)
{
	double v = 0;
	for (int b = b0 + (int)threadIdx.x; b < b1; b += LZ_THREADS) v += part[b];
	return block_sum(v, sh);
}

__global__ __launch_bounds__(LZ_THREADS) void k_blz_dot_partial(const double* __restrict__ x, const double* __restrict__ y, const LzSeg* __restrict__ segs,
	double* __restrict__ part, const double* __restrict__ scal, int S, int check_stop)
{
	__shared__ double sh[LZ_THREADS];
	const LzSeg g = segs[blockIdx.x];
	if (check_stop && scal[g.prob * S + S - 1] != 0) return;
	double v = 0;
	for (int64_t i = g.begin + threadIdx.x; i < g.end; i += LZ_THREADS) v += x[i] * y[i];
	const double s = block_sum(v, sh);
	if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(LZ_THREADS) void k_blz_normalize(const double* __restrict__ vs, double* __restrict__ v0, const LzSeg* __restrict__ segs,
	const LzProb* __restrict__ probs, const double* __restrict__ part)
{
	__shared__ double sh[LZ_THREADS];
	const LzSeg g = segs[blockIdx.x];
	const LzProb q = probs[g.prob];
	const double inv = 1 / sqrt(sum_range(part, q.b0, q.b1, sh));
	for (int64_t i = g.begin + threadIdx.x; i < g.end; i += LZ_THREADS) v0[i] = inv * vs[i];
}

__global__ __launch_bounds__(LZ_THREADS) void k_blz_alpha_update(double* __restrict__ w, const double* __restrict__ vj, const double* __restrict__ vjm1,
	int j, int maxiter, const LzSeg* __restrict__ segs, const LzProb* __restrict__ probs, const double* __restrict__ part, double* __restrict__ part2,
	double* __restrict__ scal)
{
	__shared__ double sh[LZ_THREADS];
	const int S = 2 * maxiter + 2;
	const LzSeg g = segs[blockIdx.x];
	double* sc = scal + (int64_t)g.prob * S;
	if (sc[S - 1] != 0) return;
	const LzProb q = probs[g.prob];
	const double alpha = sum_range(part, q.b0, q.b1, sh);
	if ((int)blockIdx.x == q.b0 && threadIdx.x == 0) sc[j] = alpha;
	if (j == maxiter - 1) return;
	const double bprev = (j > 0) ? sc[maxiter + j - 1] : 0.0;
	double v = 0;
	for (int64_t i = g.begin + threadIdx.x; i < g.end; i += LZ_THREADS) {
		double wi = w[i];
		if (j > 0) wi -= alpha * vj[i] + bprev * vjm1[i];
		else wi -= alpha * vj[i];
		w[i] = wi;
		v += wi * wi;
	}
	const double s = block_sum(v, sh);
	if (threadIdx.x == 0) part2[blockIdx.x] = s;
}

__global__ __launch_bounds__(LZ_THREADS) void k_blz_beta_scale(const double* __restrict__ w, double* __restrict__ vnext, int j, int maxiter,
	const LzSeg* __restrict__ segs, const LzProb* __restrict__ probs, const double* __restrict__ part2, double* __restrict__ scal)
{
	__shared__ double sh[LZ_THREADS];
	const int S = 2 * maxiter + 2;
	const LzSeg g = segs[blockIdx.x];
	double* sc = scal + (int64_t)g.prob * S;
	if (sc[S - 1] != 0) return;
	const LzProb q = probs[g.prob];
	const double beta = sqrt(sum_range(part2, q.b0, q.b1, sh));
	const double thresh = 100.0 * (double)q.n * 2.2204460492503131e-16;   // krylov.c:58
	if (beta < thresh) {
		// every block of the problem computes the same beta and returns here
		if ((int)blockIdx.x == q.b0 && threadIdx.x == 0) { sc[maxiter + j] = beta; sc[2 * maxiter] = j + 1; sc[S - 1] = 1; }
		return;
	}
	if ((int)blockIdx.x == q.b0 && threadIdx.x == 0) sc[maxiter + j] = beta;
	const double inv = 1 / beta;
	for (int64_t i = g.begin + threadIdx.x; i < g.end; i += LZ_THREADS) vnext[i] = inv * w[i];
}

// u = sum_{j < numiter_p} z_p[j] V_j on each problem's segment
__global__ __launch_bounds__(LZ_THREADS) void k_blz_ritz(const double* __restrict__ V, int64_t ntot, const double* __restrict__ z, const int* __restrict__ nit,
	int maxiter, const LzSeg* __restrict__ segs, double* __restrict__ u)
{
	const LzSeg g = segs[blockIdx.x];
	const int m = nit[g.prob];
	const double* zp = z + (int64_t)g.prob * maxiter;
	for (int64_t i = g.begin + threadIdx.x; i < g.end; i += LZ_THREADS) {
		double acc = 0;
		for (int jj = 0; jj < m; jj++) acc += V[(int64_t)jj * ntot + i] * zp[jj];
		u[i] = acc;
	}
}

void BatchLanczosWork::drop_graph()
{
	if (graph) (void)hipGraphExecDestroy(graph);
	graph = nullptr;
	graph_maxiter = 0;
}

BatchLanczosWork::~BatchLanczosWork()
{
	drop_graph();
	dev_free(v); dev_free(w); dev_free(vs); dev_free(scal); dev_free(part); dev_free(z); dev_free(nit);
}

static void enqueue_batch_recurrence(BatchPlan& b, const double* vstart, int maxiter)
{
	BatchLanczosWork& lw = b.lz;
	hipStream_t s = b.stream;
	const int64_t ntot = b.off.back();
	const int G = b.nblocks;
	const int S = 2 * maxiter + 2;
	double* part = lw.part;
	double* part2 = lw.part + G;
	launch_fill(lw.scal, 0.0, (ct_long)b.nprob * S, s);
	hipLaunchKernelGGL(k_blz_dot_partial, dim3(G), dim3(LZ_THREADS), 0, s, vstart, vstart, b.segs, part, (const double*)lw.scal, S, 0);
	hipLaunchKernelGGL(k_blz_normalize, dim3(G), dim3(LZ_THREADS), 0, s, vstart, lw.v, b.segs, b.probs, (const double*)part);
	CTG_HIP(hipGetLastError());
	for (int j = 0; j < maxiter; j++) {
		const double* vj = lw.v + (int64_t)j * ntot;
		batch_apply(b, vj, lw.w);
		hipLaunchKernelGGL(k_blz_dot_partial, dim3(G), dim3(LZ_THREADS), 0, s, (const double*)lw.w, vj, b.segs, part, (const double*)lw.scal, S, 1);
		hipLaunchKernelGGL(k_blz_alpha_update, dim3(G), dim3(LZ_THREADS), 0, s, lw.w, vj, j > 0 ? vj - ntot : vj, j, maxiter, b.segs, b.probs,
			(const double*)part, part2, lw.scal);
		if (j < maxiter - 1)
			hipLaunchKernelGGL(k_blz_beta_scale, dim3(G), dim3(LZ_THREADS), 0, s, (const double*)lw.w, lw.v + (int64_t)(j + 1) * ntot, j, maxiter,
				b.segs, b.probs, (const double*)part2, lw.scal);
		CTG_HIP(hipGetLastError());
	}
}

int lanczos_ground_state_batch(BatchPlan& b, const double* vstart, int maxiter, double* lambda, int* status, double* u, bool use_graph)
{
	if (maxiter < 1 || b.nprob < 1) return CT_ERR_ARG;
	BatchLanczosWork& lw = b.lz;
	hipStream_t s = b.stream;
	const int64_t ntot = b.off.back();
	const int S = 2 * maxiter + 2;
	if (lw.cap < maxiter) {
		lw.drop_graph();
		dev_free(lw.v); dev_free(lw.scal); dev_free(lw.z);
		lw.v = (double*)dev_alloc((size_t)maxiter * (size_t)ntot * sizeof(double));
		lw.scal = (double*)dev_alloc((size_t)b.nprob * S * sizeof(double));
		lw.z = (double*)dev_alloc((size_t)b.nprob * maxiter * sizeof(double));
		lw.cap = maxiter;
	}
	if (!lw.w) {
		lw.w = (double*)dev_alloc((size_t)ntot * sizeof(double));
		lw.vs = (double*)dev_alloc((size_t)ntot * sizeof(double));
		lw.part = (double*)dev_alloc((size_t)2 * b.nblocks * sizeof(double));
		lw.nit = (int*)dev_alloc((size_t)b.nprob * sizeof(int));
	}
	if (use_graph) {
		launch_copy(vstart, lw.vs, ntot, s);
		if (!lw.graph || lw.graph_maxiter != maxiter) {
			lw.drop_graph();
			hipGraph_t g = nullptr;
			CaptureScope cs;
			CTG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
			try {
				enqueue_batch_recurrence(b, lw.vs, maxiter);
			}
			catch (...) {
				(void)hipStreamEndCapture(s, &g);
				if (g) (void)hipGraphDestroy(g);
				throw;
			}
			CTG_HIP(hipStreamEndCapture(s, &g));
			const hipError_t e = hipGraphInstantiate(&lw.graph, g, nullptr, nullptr, 0);
			(void)hipGraphDestroy(g);
			CTG_HIP(e);
			lw.graph_maxiter = maxiter;
		}
		CTG_HIP(hipGraphLaunch(lw.graph, s));
	}
	else enqueue_batch_recurrence(b, vstart, maxiter);
	std::vector<double> hs((size_t)b.nprob * S);
	CTG_HIP(hipMemcpyAsync(hs.data(), lw.scal, hs.size() * sizeof(double), hipMemcpyDeviceToHost, s));
	CTG_HIP(hipStreamSynchronize(s));
	std::vector<double> zall((size_t)b.nprob * maxiter, 0.0);
	std::vector<int> nit(b.nprob, 0);
	int worst = 0;
	for (int p = 0; p < b.nprob; p++) {
		const double* h = hs.data() + (size_t)p * S;
		const int numiter = (h[S - 1] != 0) ? (int)h[2 * maxiter] : maxiter;
		int st = 0;
		lambda[p] = std::numeric_limits<double>::quiet_NaN();
		if (numiter < 1) st = -1;
		else {
			std::vector<double> d(h, h + numiter), e((size_t)numiter + 1, 0.0), z((size_t)numiter * numiter);
			for (int i = 0; i < numiter - 1; i++) e[i] = h[maxiter + i];
			if (host_tridiag_eigh(numiter, d.data(), e.data(), z.data()) != 0) st = -2;
			else {
				lambda[p] = d[0];
				for (int jj = 0; jj < numiter; jj++) zall[(size_t)p * maxiter + jj] = z[jj];
				nit[p] = numiter;
			}
		}
		if (status) status[p] = st;
		if (st != 0 && worst == 0) worst = st;
	}
	CTG_HIP(hipMemcpyAsync(lw.z, zall.data(), zall.size() * sizeof(double), hipMemcpyHostToDevice, s));
	CTG_HIP(hipMemcpyAsync(lw.nit, nit.data(), nit.size() * sizeof(int), hipMemcpyHostToDevice, s));
	hipLaunchKernelGGL(k_blz_ritz, dim3(b.nblocks), dim3(LZ_THREADS), 0, s, (const double*)lw.v, ntot, (const double*)lw.z, (const int*)lw.nit, maxiter,
		b.segs, u);
	CTG_HIP(hipGetLastError());
	CTG_HIP(hipStreamSynchronize(s));
	return worst;
}

}  // namespace ctg
