/// \file ct_oracle.c
/// \brief CPU FP64 oracle for the two-site DMRG path (TEST INFRASTRUCTURE ONLY).
///
/// A plain-C restatement of the reference algorithms, used only as the checker of the
/// GPU product (tests/, __graft_entry__.smoke(), bench.py cpu_baseline leg).
/// Each function cites the reference source it follows (paths relative to the
/// chemtensor tree). Linear algebra that the reference delegates to BLAS/LAPACK is
/// restated from the published algorithms (LAPACK dgeqr2/dorg2r, dgerq2/dorgr2,
/// Hestenes one-sided Jacobi SVD, implicit-shift QL for tridiagonal matrices).

#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <assert.h>
#include <stdio.h>
#include "ct_oracle.h"

#define ORC_EXPORT __attribute__((visibility("default")))

static void* xmalloc(size_t n) { void* p = malloc(n > 0 ? n : 1); if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); } return p; }
static void* xcalloc(size_t n, size_t s) { void* p = calloc(n > 0 ? n : 1, s > 0 ? s : 1); if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); } return p; }
static ct_long iprod(const ct_long* d, int n) { ct_long p = 1; for (int i = 0; i < n; i++) p *= d[i]; return p; }
static ct_long lmin(ct_long a, ct_long b) { return a < b ? a : b; }
static ct_long lmax(ct_long a, ct_long b) { return a > b ? a : b; }

static void next_index(int ndim, const ct_long* dim, ct_long* index)
{
	for (int i = ndim - 1; i >= 0; i--) {
		index[i]++;
		if (index[i] < dim[i]) return;
		index[i] = 0;
	}
}

static ct_long index_to_offset(int ndim, const ct_long* dim, const ct_long* index)
{
	ct_long off = 0, f = 1;
	for (int i = ndim - 1; i >= 0; i--) { off += f * index[i]; f *= dim[i]; }
	return off;
}

static void offset_to_index(int ndim, const ct_long* dim, ct_long off, ct_long* index)
{
	for (int i = ndim - 1; i >= 0; i--) { index[i] = off % dim[i]; off /= dim[i]; }
}

//________________________________________________________________________________________________________________________
//
// random numbers: PCG32 (published algorithm, src/util/pcg_basic.c:33-151) and
// chemtensor's seeding / Box-Muller / Floyd sampling (src/util/rng.c:15-125)

static uint32_t pcg32_next(struct pcg32_state* rng)
{
	const uint64_t old = rng->state;
	rng->state = old * 6364136223846793005ULL + rng->inc;
	const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
	const uint32_t rot = (uint32_t)(old >> 59u);
	return (xs >> rot) | (xs << ((-rot) & 31));
}

static void pcg32_seed(uint64_t initstate, uint64_t initseq, struct pcg32_state* rng)
{
	rng->state = 0u;
	rng->inc = (initseq << 1u) | 1u;
	pcg32_next(rng);
	rng->state += initstate;
	pcg32_next(rng);
}

// src/util/rng.c:15-27
ORC_EXPORT void orc_seed_rng_state(uint64_t seed, struct orc_rng_state* state)
{
	struct pcg32_state tmp;
	pcg32_seed(seed, 6364136223846793005ULL * seed + 1442695040888963407ULL, &tmp);
	uint64_t w[8];
	for (int i = 0; i < 8; i++) w[i] = pcg32_next(&tmp);
	uint64_t seed1 = (w[0] << 32) | w[1];
	uint64_t seed2 = (w[2] << 32) | w[3];
	uint64_t seq1  = (w[4] << 32) | w[5];
	uint64_t seq2  = (w[6] << 32) | w[7];
	// pcg_basic.c:111-124: the two streams must differ
	const uint64_t mask = ~0ull >> 1;
	if ((seq1 & mask) == (seq2 & mask)) seq2 = ~seq2;
	pcg32_seed(seed1, seq1, &state->gen[0]);
	pcg32_seed(seed2, seq2, &state->gen[1]);
}

// pcg_basic.c:131-135
ORC_EXPORT uint64_t orc_rand_uint64(struct orc_rng_state* state)
{
	const uint64_t hi = pcg32_next(&state->gen[0]);
	const uint64_t lo = pcg32_next(&state->gen[1]);
	return (hi << 32) | lo;
}

// pcg_basic.c:141-150 (rejection sampling for an unbiased bounded draw)
ORC_EXPORT uint64_t orc_rand_interval(uint64_t bound, struct orc_rng_state* state)
{
	const uint64_t threshold = (0 - bound) % bound;
	for (;;) {
		const uint64_t r = orc_rand_uint64(state);
		if (r >= threshold) return r % bound;
	}
}

// rng.c:65-85 (Floyd's algorithm)
ORC_EXPORT void orc_rand_choice(uint64_t bound, uint64_t num_samples, struct orc_rng_state* state, uint64_t* ret)
{
	for (uint64_t i = 0; i < num_samples; i++) {
		const uint64_t j = bound - num_samples + 1 + i;
		const uint64_t t = orc_rand_interval(j, state);
		bool drawn = false;
		for (uint64_t k = 0; k < i; k++) {
			if (ret[k] == t) { drawn = true; break; }
		}
		ret[i] = drawn ? j - 1 : t;
	}
}

// rng.c:93-96
ORC_EXPORT double orc_randu(struct orc_rng_state* state)
{
	return ldexp((double)orc_rand_uint64(state), -64);
}

// rng.c:116-125 (Box-Muller, sine branch)
ORC_EXPORT double orc_randn(struct orc_rng_state* state)
{
	double u1 = orc_randu(state);
	double u2 = orc_randu(state);
	if (u1 == 0) u1 = 1;
	return sqrt(-2 * log(u1)) * sin(6.2831853071795864769 * u2);
}

//________________________________________________________________________________________________________________________
//
// dense tensors (src/tensor/dense_tensor.c)

ORC_EXPORT void orc_allocate_dense_tensor(int ndim, const ct_long* dim, struct dense_tensor* t)
{
	t->dtype = CT_DOUBLE_REAL;
	t->ndim = ndim;
	t->dim = xmalloc((size_t)(ndim > 0 ? ndim : 1) * sizeof(ct_long));
	if (ndim > 0) memcpy(t->dim, dim, ndim * sizeof(ct_long));
	t->data = xcalloc((size_t)iprod(dim, ndim), sizeof(double));
}

ORC_EXPORT void orc_delete_dense_tensor(struct dense_tensor* t)
{
	free(t->data); t->data = NULL;
	free(t->dim); t->dim = NULL;
	t->ndim = 0;
}

ORC_EXPORT ct_long orc_dense_num_elements(const struct dense_tensor* t) { return iprod(t->dim, t->ndim); }

static void copy_dense_tensor(const struct dense_tensor* src, struct dense_tensor* dst)
{
	orc_allocate_dense_tensor(src->ndim, src->dim, dst);
	memcpy(dst->data, src->data, (size_t)orc_dense_num_elements(src) * sizeof(double));
}

// dense_tensor.c:799-935: r[i_0..] = t[...] with r axis i = t axis perm[i]
// (the reference squeezes unit axes and fuses neighbours first; the values are identical)
ORC_EXPORT void orc_dense_tensor_transpose(const int* perm, const struct dense_tensor* t, struct dense_tensor* r)
{
	const int nd = t->ndim;
	if (nd == 0) { orc_allocate_dense_tensor(0, NULL, r); ((double*)r->data)[0] = ((const double*)t->data)[0]; return; }
	ct_long rdim[16];
	for (int i = 0; i < nd; i++) rdim[i] = t->dim[perm[i]];
	orc_allocate_dense_tensor(nd, rdim, r);
	// stride in 'r' of each axis of 't'
	ct_long rstride[16], f = 1;
	for (int i = nd - 1; i >= 0; i--) { rstride[perm[i]] = f; f *= rdim[i]; }
	const ct_long n = iprod(t->dim, nd);
	const double* td = t->data;
	double* rd = r->data;
	ct_long idx[16] = { 0 };
	const ct_long last = t->dim[nd - 1], ls = rstride[nd - 1];
	for (ct_long o = 0; o < n; o += last) {
		ct_long ro = 0;
		for (int i = 0; i < nd - 1; i++) ro += idx[i] * rstride[i];
		for (ct_long j = 0; j < last; j++) rd[ro + j * ls] = td[o + j];
		next_index(nd - 1, t->dim, idx);
	}
}

// C[m x n] += A[m x k] B[k x n], all row-major contiguous
static void gemm_nn_acc(ct_long m, ct_long n, ct_long k, const double* A, const double* B, double* C)
{
	const ct_long JB = 512, PB = 128;
	for (ct_long j0 = 0; j0 < n; j0 += JB) {
		const ct_long j1 = lmin(n, j0 + JB);
		for (ct_long p0 = 0; p0 < k; p0 += PB) {
			const ct_long p1 = lmin(k, p0 + PB);
			for (ct_long i = 0; i < m; i++) {
				double* c = C + i * n;
				for (ct_long p = p0; p < p1; p++) {
					const double a = A[i * k + p];
					const double* b = B + p * n;
					for (ct_long j = j0; j < j1; j++) c[j] += a * b[j];
				}
			}
		}
	}
}

// dense_tensor.c:1761-1838: r <- s @ t + r with the reference's matrix interpretation
// (transa iff the s-axes are leading, transb iff the t-axes are trailing; row-major)
ORC_EXPORT void orc_dense_tensor_dot_update(const struct dense_tensor* s, enum tensor_axis_range axrange_s, const struct dense_tensor* t, enum tensor_axis_range axrange_t, int ndim_mult, struct dense_tensor* r)
{
	const int nlds = (axrange_s == TENSOR_AXIS_RANGE_LEADING ? ndim_mult : s->ndim - ndim_mult);
	const int nldt = (axrange_t == TENSOR_AXIS_RANGE_LEADING ? ndim_mult : t->ndim - ndim_mult);
	const ct_long lds = iprod(s->dim, nlds), tds = iprod(s->dim + nlds, s->ndim - nlds);
	const ct_long ldt = iprod(t->dim, nldt), tdt = iprod(t->dim + nldt, t->ndim - nldt);
	const bool transa = (axrange_s == TENSOR_AXIS_RANGE_LEADING);
	const bool transb = (axrange_t == TENSOR_AXIS_RANGE_TRAILING);
	const ct_long m = transa ? tds : lds;
	const ct_long n = transb ? ldt : tdt;
	const ct_long k = transa ? lds : tds;
	assert(k == (transb ? tdt : ldt));
	const double* A = s->data;
	const double* B = t->data;
	double* Ap = NULL;
	double* Bp = NULL;
	if (transa) {
		Ap = xmalloc((size_t)(m * k) * sizeof(double));
		for (ct_long p = 0; p < k; p++) for (ct_long i = 0; i < m; i++) Ap[i * k + p] = A[p * tds + i];
		A = Ap;
	}
	if (transb) {
		Bp = xmalloc((size_t)(k * n) * sizeof(double));
		for (ct_long j = 0; j < n; j++) for (ct_long p = 0; p < k; p++) Bp[p * n + j] = B[j * tdt + p];
		B = Bp;
	}
	gemm_nn_acc(m, n, k, A, B, r->data);
	free(Ap);
	free(Bp);
}

//________________________________________________________________________________________________________________________
//
// Householder reflectors following LAPACK's conventions (dlarfg, dlarf), so that the
// signs of QR / RQ factors agree with the reference's LAPACK calls.

static double nrm2(ct_long n, const double* x, ct_long incx)
{
	// scaled sum of squares (as in the reference BLAS dnrm2) to avoid overflow
	double scale = 0, ssq = 1;
	for (ct_long i = 0; i < n; i++) {
		const double v = x[i * incx];
		if (v != 0) {
			const double a = fabs(v);
			if (scale < a) { ssq = 1 + ssq * (scale / a) * (scale / a); scale = a; }
			else { ssq += (a / scale) * (a / scale); }
		}
	}
	return scale * sqrt(ssq);
}

// dlarfg: generate H = I - tau v v^T with H (alpha; x) = (beta; 0), v = (1; x_out)
static void larfg(ct_long n, double* alpha, double* x, ct_long incx, double* tau)
{
	if (n <= 1) { *tau = 0; return; }
	const double xnorm = nrm2(n - 1, x, incx);
	if (xnorm == 0) { *tau = 0; return; }
	const double beta = -copysign(hypot(*alpha, xnorm), *alpha);
	*tau = (beta - *alpha) / beta;
	const double sc = 1.0 / (*alpha - beta);
	for (ct_long i = 0; i < n - 1; i++) x[i * incx] *= sc;
	*alpha = beta;
}

// column-major helpers: a(i,j) = a[i + lda*j]
#define CM(a, lda, i, j) ((a)[(i) + (lda) * (j)])

// apply H = I - tau v v^T from the left to C (m x n, column-major, ldc); v has stride incv
static void larf_left(ct_long m, ct_long n, const double* v, ct_long incv, double tau, double* c, ct_long ldc)
{
	if (tau == 0) return;
	for (ct_long j = 0; j < n; j++) {
		double w = 0;
		for (ct_long i = 0; i < m; i++) w += v[i * incv] * CM(c, ldc, i, j);
		w *= tau;
		for (ct_long i = 0; i < m; i++) CM(c, ldc, i, j) -= v[i * incv] * w;
	}
}

// apply H = I - tau v v^T from the right to C (m x n, column-major, ldc); v has stride incv
static void larf_right(ct_long m, ct_long n, const double* v, ct_long incv, double tau, double* c, ct_long ldc)
{
	if (tau == 0) return;
	for (ct_long i = 0; i < m; i++) {
		double w = 0;
		for (ct_long j = 0; j < n; j++) w += CM(c, ldc, i, j) * v[j * incv];
		w *= tau;
		for (ct_long j = 0; j < n; j++) CM(c, ldc, i, j) -= w * v[j * incv];
	}
}

// dense_tensor.c:2253-2420 (reduced mode): a = q r via dgeqr2 + dorg2r
ORC_EXPORT int orc_dense_tensor_qr_fill(const struct dense_tensor* a, struct dense_tensor* q, struct dense_tensor* r)
{
	const ct_long m = a->dim[0], n = a->dim[1], k = lmin(m, n);
	double* t = xmalloc((size_t)(m * n) * sizeof(double));
	double* tau = xmalloc((size_t)k * sizeof(double));
	const double* ad = a->data;
	for (ct_long i = 0; i < m; i++) for (ct_long j = 0; j < n; j++) CM(t, m, i, j) = ad[i * n + j];
	// dgeqr2
	for (ct_long i = 0; i < k; i++) {
		larfg(m - i, &CM(t, m, i, i), &CM(t, m, lmin(i + 1, m - 1), i), 1, &tau[i]);
		if (i < n - 1) {
			const double aii = CM(t, m, i, i);
			CM(t, m, i, i) = 1;
			larf_left(m - i, n - i - 1, &CM(t, m, i, i), 1, tau[i], &CM(t, m, i, i + 1), m);
			CM(t, m, i, i) = aii;
		}
	}
	double* rd = r->data;
	for (ct_long i = 0; i < k; i++) for (ct_long j = 0; j < n; j++) rd[i * n + j] = (j < i ? 0 : CM(t, m, i, j));
	// dorg2r: form the first k columns of Q in place
	for (ct_long i = k - 1; i >= 0; i--) {
		if (i < k - 1) {
			CM(t, m, i, i) = 1;
			larf_left(m - i, k - i - 1, &CM(t, m, i, i), 1, tau[i], &CM(t, m, i, i + 1), m);
		}
		for (ct_long l = i + 1; l < m; l++) CM(t, m, l, i) *= -tau[i];
		CM(t, m, i, i) = 1 - tau[i];
		for (ct_long l = 0; l < i; l++) CM(t, m, l, i) = 0;
	}
	double* qd = q->data;
	for (ct_long i = 0; i < m; i++) for (ct_long j = 0; j < k; j++) qd[i * k + j] = CM(t, m, i, j);
	free(tau);
	free(t);
	return 0;
}

// dense_tensor.c:2680-2900 (reduced mode): a = r q via dgerq2 + dorgr2
ORC_EXPORT int orc_dense_tensor_rq_fill(const struct dense_tensor* a, struct dense_tensor* r, struct dense_tensor* q)
{
	const ct_long m = a->dim[0], n = a->dim[1], k = lmin(m, n);
	double* t = xmalloc((size_t)(m * n) * sizeof(double));
	double* tau = xmalloc((size_t)k * sizeof(double));
	const double* ad = a->data;
	for (ct_long i = 0; i < m; i++) for (ct_long j = 0; j < n; j++) CM(t, m, i, j) = ad[i * n + j];
	// dgerq2: reflector i annihilates row m-k+i left of column n-k+i
	for (ct_long i = k - 1; i >= 0; i--) {
		const ct_long row = m - k + i, col = n - k + i;
		larfg(col + 1, &CM(t, m, row, col), &CM(t, m, row, 0), m, &tau[i]);
		const double aii = CM(t, m, row, col);
		CM(t, m, row, col) = 1;
		larf_right(row, col + 1, &CM(t, m, row, 0), m, tau[i], t, m);
		CM(t, m, row, col) = aii;
	}
	// R (m x k): upper trapezoid ending at the last k columns
	double* rd = r->data;
	const ct_long rshift_row = lmax(m - k, 0);
	const ct_long tshift_col = lmax(n - m, 0);
	for (ct_long i = 0; i < m; i++) {
		const ct_long numzero = lmax(i - rshift_row, 0);
		for (ct_long j = 0; j < k; j++) rd[i * k + j] = (j < numzero ? 0 : CM(t, m, i, tshift_col + j));
	}
	// dorgr2 on the last k rows of t (which hold the reflectors)
	double* qt = xmalloc((size_t)(k * n) * sizeof(double));
	for (ct_long i = 0; i < k; i++) for (ct_long j = 0; j < n; j++) CM(qt, k, i, j) = CM(t, m, rshift_row + i, j);
	for (ct_long i = 0; i < k; i++) {
		const ct_long col = n - k + i;
		CM(qt, k, i, col) = 1;
		larf_right(i, col + 1, &CM(qt, k, i, 0), k, tau[i], qt, k);
		for (ct_long j = 0; j < col; j++) CM(qt, k, i, j) *= -tau[i];
		CM(qt, k, i, col) = 1 - tau[i];
		for (ct_long j = col + 1; j < n; j++) CM(qt, k, i, j) = 0;
	}
	double* qd = q->data;
	for (ct_long i = 0; i < k; i++) for (ct_long j = 0; j < n; j++) qd[i * n + j] = CM(qt, k, i, j);
	free(qt);
	free(tau);
	free(t);
	return 0;
}

// Hestenes one-sided Jacobi SVD of a column-major matrix g (m x n, m >= n): on return
// g holds U*diag(s) column-wise, v (n x n, column-major) the right singular vectors.
static int jacobi_svd_tall(ct_long m, ct_long n, double* g, double* v)
{
	for (ct_long i = 0; i < n * n; i++) v[i] = 0;
	for (ct_long i = 0; i < n; i++) CM(v, n, i, i) = 1;
	// rotation threshold sqrt(m) eps (as in LAPACK dgesvj); exact eps can cycle on rounding
	const double eps = 2.2204460492503131e-16 * fmax(sqrt((double)m), 4.0);
	// columns with norm below eps ||a||_F are numerically zero and set to zero: a column in the span
	// of the others is otherwise projected out again and again, its rounding-noise residual never
	// orthogonal to the rest, until it underflows
	double fro2 = 0;
	for (ct_long i = 0; i < m * n; i++) fro2 += g[i] * g[i];
	const double small2 = fro2 * 2.2204460492503131e-16 * 2.2204460492503131e-16;
	for (int sweep = 0; sweep < 80; sweep++) {
		bool rotated = false;
		for (ct_long p = 0; p < n - 1; p++) {
			for (ct_long q = p + 1; q < n; q++) {
				double alpha = 0, beta = 0, gamma = 0;
				double* gp = &CM(g, m, 0, p);
				double* gq = &CM(g, m, 0, q);
				for (ct_long i = 0; i < m; i++) { alpha += gp[i] * gp[i]; beta += gq[i] * gq[i]; gamma += gp[i] * gq[i]; }
				if (alpha <= small2 || beta <= small2) {
					if (alpha != 0 && alpha <= small2) for (ct_long i = 0; i < m; i++) gp[i] = 0;
					if (beta != 0 && beta <= small2) for (ct_long i = 0; i < m; i++) gq[i] = 0;
					continue;
				}
				if (gamma == 0 || fabs(gamma) <= eps * sqrt(alpha) * sqrt(beta)) continue;
				rotated = true;
				const double zeta = (beta - alpha) / (2 * gamma);
				const double tt = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1 + zeta * zeta));
				const double c = 1 / sqrt(1 + tt * tt), s = c * tt;
				for (ct_long i = 0; i < m; i++) { const double x = gp[i], y = gq[i]; gp[i] = c * x - s * y; gq[i] = s * x + c * y; }
				double* vp = &CM(v, n, 0, p);
				double* vq = &CM(v, n, 0, q);
				for (ct_long i = 0; i < n; i++) { const double x = vp[i], y = vq[i]; vp[i] = c * x - s * y; vq[i] = s * x + c * y; }
			}
		}
		if (!rotated) return 0;
	}
	return -1;
}

// complete the columns of u (m x k, column-major) whose norm is zero to an orthonormal set: each
// takes the unit vector e_i with the largest component outside the span of the columns set so far,
// orthogonalised twice (classical Gram-Schmidt) and normalised
static void complete_orthonormal_columns(ct_long m, ct_long k, double* u, const bool* valid)
{
	double* rowsq = xcalloc((size_t)m, sizeof(double));
	for (ct_long l = 0; l < k; l++) {
		if (!valid[l]) continue;
		for (ct_long i = 0; i < m; i++) rowsq[i] += CM(u, m, i, l) * CM(u, m, i, l);
	}
	for (ct_long j = 0; j < k; j++) {
		if (valid[j]) continue;
		ct_long e = 0;
		for (ct_long i = 1; i < m; i++) if (rowsq[i] < rowsq[e]) e = i;
		double* x = &CM(u, m, 0, j);
		for (ct_long i = 0; i < m; i++) x[i] = (i == e ? 1 : 0);
		for (int pass = 0; pass < 2; pass++) {
			for (ct_long l = 0; l < k; l++) {
				if (l == j || !valid[l]) continue;
				const double* y = &CM(u, m, 0, l);
				double d = 0;
				for (ct_long i = 0; i < m; i++) d += x[i] * y[i];
				for (ct_long i = 0; i < m; i++) x[i] -= d * y[i];
			}
		}
		const double nx = nrm2(m, x, 1);
		for (ct_long i = 0; i < m; i++) { x[i] /= nx; rowsq[i] += x[i] * x[i]; }
		((bool*)valid)[j] = true;
	}
	free(rowsq);
}

// dense_tensor.c:3538-3700: a = u diag(s) vh with s descending (reference: LAPACK dgesvd)
ORC_EXPORT int orc_dense_tensor_svd_fill(const struct dense_tensor* a, struct dense_tensor* u, double* s, struct dense_tensor* vh)
{
	const ct_long m = a->dim[0], n = a->dim[1], k = lmin(m, n);
	const bool tall = (m >= n);
	const ct_long M = tall ? m : n, N = tall ? n : m;
	double* g = xmalloc((size_t)(M * N) * sizeof(double));
	double* v = xmalloc((size_t)(N * N) * sizeof(double));
	const double* ad = a->data;
	// g = a (tall) or a^T (wide), column-major
	for (ct_long i = 0; i < m; i++) for (ct_long j = 0; j < n; j++) {
		if (tall) CM(g, M, i, j) = ad[i * n + j];
		else CM(g, M, j, i) = ad[i * n + j];
	}
	if (jacobi_svd_tall(M, N, g, v) != 0) { fprintf(stderr, "oracle: Jacobi SVD did not converge (%ld x %ld)\n", (long)M, (long)N); free(g); free(v); return -1; }
	double* sv = xmalloc((size_t)N * sizeof(double));
	ct_long* ord = xmalloc((size_t)N * sizeof(ct_long));
	for (ct_long j = 0; j < N; j++) { sv[j] = nrm2(M, &CM(g, M, 0, j), 1); ord[j] = j; }
	// sort descending (stable insertion sort)
	for (ct_long i = 1; i < N; i++) {
		const ct_long o = ord[i]; ct_long j = i;
		while (j > 0 && sv[ord[j - 1]] < sv[o]) { ord[j] = ord[j - 1]; j--; }
		ord[j] = o;
	}
	// left vectors of g: normalised columns (tall: u, wide: v of a)
	double* lu = xmalloc((size_t)(M * N) * sizeof(double));
	bool* valid = xmalloc((size_t)N * sizeof(bool));
	const double smax = (N > 0 ? sv[ord[0]] : 0);
	for (ct_long jj = 0; jj < N; jj++) {
		const ct_long j = ord[jj];
		s[jj] = sv[j];
		valid[jj] = (sv[j] > 0 && sv[j] > 1e-300 && sv[j] >= smax * 1e-300);
		for (ct_long i = 0; i < M; i++) CM(lu, M, i, jj) = valid[jj] ? CM(g, M, i, j) / sv[j] : 0;
	}
	complete_orthonormal_columns(M, N, lu, valid);
	double* ud = u->data;
	double* vd = vh->data;
	if (tall) {
		// a = lu diag(s) v^T: u = lu (m x k), vh = v^T (k x n)
		for (ct_long i = 0; i < m; i++) for (ct_long j = 0; j < k; j++) ud[i * k + j] = CM(lu, M, i, j);
		for (ct_long jj = 0; jj < k; jj++) for (ct_long i = 0; i < n; i++) vd[jj * n + i] = CM(v, N, i, ord[jj]);
	}
	else {
		// a^T = lu diag(s) v^T -> a = v diag(s) lu^T: u = v (m x k), vh = lu^T (k x n)
		for (ct_long i = 0; i < m; i++) for (ct_long jj = 0; jj < k; jj++) ud[i * k + jj] = CM(v, N, i, ord[jj]);
		for (ct_long jj = 0; jj < k; jj++) for (ct_long i = 0; i < n; i++) vd[jj * n + i] = CM(lu, M, i, jj);
	}
	free(valid); free(lu); free(ord); free(sv); free(v); free(g);
	return 0;
}

// Symmetric tridiagonal eigensystem by the implicit-shift QL method (the algorithm of
// EISPACK tql2); replaces the reference's LAPACK_dsteqr("I") call at src/util/krylov.c:230.
// On return diag holds the eigenvalues in ascending order and z (n x n, column-major)
// the eigenvectors in its columns. 'offdiag' has length n-1 and is destroyed.
ORC_EXPORT int orc_symmetric_tridiag_eigh(int n, double* d, double* offdiag, double* z)
{
	double* e = xcalloc((size_t)n + 1, sizeof(double));
	for (int i = 0; i < n - 1; i++) e[i] = offdiag[i];
	for (int i = 0; i < n * n; i++) z[i] = 0;
	for (int i = 0; i < n; i++) z[i + n * i] = 1;
	for (int l = 0; l < n; l++) {
		int iter = 0;
		for (;;) {
			int m;
			for (m = l; m < n - 1; m++) {
				const double dd = fabs(d[m]) + fabs(d[m + 1]);
				if (fabs(e[m]) <= 2.2204460492503131e-16 * dd) break;
			}
			if (m == l) break;
			if (++iter > 60) { free(e); return -1; }
			double g = (d[l + 1] - d[l]) / (2 * e[l]);
			double r = hypot(g, 1.0);
			g = d[m] - d[l] + e[l] / (g + copysign(r, g));
			double s = 1, c = 1, p = 0;
			int i;
			for (i = m - 1; i >= l; i--) {
				double f = s * e[i];
				const double b = c * e[i];
				r = hypot(f, g);
				e[i + 1] = r;
				if (r == 0) { d[i + 1] -= p; e[m] = 0; break; }
				s = f / r; c = g / r;
				g = d[i + 1] - p;
				r = (d[i] - g) * s + 2 * c * b;
				p = s * r;
				d[i + 1] = g + p;
				g = c * r - b;
				for (int k = 0; k < n; k++) {
					f = z[k + n * (i + 1)];
					z[k + n * (i + 1)] = s * z[k + n * i] + c * f;
					z[k + n * i] = c * z[k + n * i] - s * f;
				}
			}
			if (r == 0 && i >= l) continue;
			d[l] -= p; e[l] = g; e[m] = 0;
		}
	}
	// sort ascending
	for (int i = 0; i < n - 1; i++) {
		int k = i;
		for (int j = i + 1; j < n; j++) if (d[j] < d[k]) k = j;
		if (k != i) {
			const double tmp = d[i]; d[i] = d[k]; d[k] = tmp;
			for (int r2 = 0; r2 < n; r2++) { const double t2 = z[r2 + n * i]; z[r2 + n * i] = z[r2 + n * k]; z[r2 + n * k] = t2; }
		}
	}
	free(e);
	return 0;
}

//________________________________________________________________________________________________________________________
//
// block-sparse tensors (src/tensor/block_sparse_tensor.c)

struct qcount { qnumber q; ct_long c; };
static int cmp_qcount(const void* a, const void* b)
{
	const qnumber x = ((const struct qcount*)a)->q, y = ((const struct qcount*)b)->q;
	return (x > y) - (x < y);
}

ORC_EXPORT ct_long orc_bst_num_blocks(const struct block_sparse_tensor* t) { return iprod(t->dim_blocks, t->ndim); }

// bst.c:37-150: sort the distinct quantum numbers of each axis ascending; a block exists
// iff the signed sum of its block quantum numbers is zero; within a block, entries follow
// the first-appearance order of the logical indices carrying that quantum number.
ORC_EXPORT void orc_allocate_block_sparse_tensor(int ndim, const ct_long* dim, const enum tensor_axis_direction* axis_dir, const qnumber** qnums, struct block_sparse_tensor* t)
{
	t->dtype = CT_DOUBLE_REAL;
	t->ndim = ndim;
	if (ndim == 0) {
		t->dim_logical = NULL; t->dim_blocks = NULL; t->axis_dir = NULL;
		t->qnums_logical = NULL; t->qnums_blocks = NULL;
		t->blocks = xcalloc(1, sizeof(struct dense_tensor*));
		t->blocks[0] = xcalloc(1, sizeof(struct dense_tensor));
		orc_allocate_dense_tensor(0, NULL, t->blocks[0]);
		return;
	}
	t->dim_logical = xmalloc(ndim * sizeof(ct_long));
	memcpy(t->dim_logical, dim, ndim * sizeof(ct_long));
	t->dim_blocks = xcalloc(ndim, sizeof(ct_long));
	t->axis_dir = xmalloc(ndim * sizeof(enum tensor_axis_direction));
	memcpy(t->axis_dir, axis_dir, ndim * sizeof(enum tensor_axis_direction));
	t->qnums_logical = xcalloc(ndim, sizeof(qnumber*));
	t->qnums_blocks = xcalloc(ndim, sizeof(qnumber*));
	struct qcount** qc = xmalloc(ndim * sizeof(struct qcount*));
	for (int i = 0; i < ndim; i++) {
		t->qnums_logical[i] = xmalloc((size_t)dim[i] * sizeof(qnumber));
		memcpy(t->qnums_logical[i], qnums[i], dim[i] * sizeof(qnumber));
		qc[i] = xcalloc((size_t)dim[i], sizeof(struct qcount));
		ct_long nq = 0;
		for (ct_long j = 0; j < dim[i]; j++) {
			ct_long k;
			for (k = 0; k < nq; k++) if (qc[i][k].q == qnums[i][j]) { qc[i][k].c++; break; }
			if (k == nq) { qc[i][nq].q = qnums[i][j]; qc[i][nq].c = 1; nq++; }
		}
		qsort(qc[i], nq, sizeof(struct qcount), cmp_qcount);
		t->dim_blocks[i] = nq;
		t->qnums_blocks[i] = xmalloc((size_t)nq * sizeof(qnumber));
		for (ct_long j = 0; j < nq; j++) t->qnums_blocks[i][j] = qc[i][j].q;
	}
	const ct_long nblocks = iprod(t->dim_blocks, ndim);
	t->blocks = xcalloc((size_t)nblocks, sizeof(struct dense_tensor*));
	ct_long* idx = xcalloc(ndim, sizeof(ct_long));
	ct_long* bdim = xmalloc(ndim * sizeof(ct_long));
	for (ct_long k = 0; k < nblocks; k++, next_index(ndim, t->dim_blocks, idx)) {
		qnumber qsum = 0;
		for (int i = 0; i < ndim; i++) qsum += axis_dir[i] * t->qnums_blocks[i][idx[i]];
		if (qsum != 0) continue;
		for (int i = 0; i < ndim; i++) bdim[i] = qc[i][idx[i]].c;
		t->blocks[k] = xcalloc(1, sizeof(struct dense_tensor));
		orc_allocate_dense_tensor(ndim, bdim, t->blocks[k]);
	}
	free(bdim); free(idx);
	for (int i = 0; i < ndim; i++) free(qc[i]);
	free(qc);
}

// bst.c:168-205
ORC_EXPORT void orc_delete_block_sparse_tensor(struct block_sparse_tensor* t)
{
	if (t->blocks == NULL) return;
	const ct_long nblocks = (t->ndim == 0 ? 1 : iprod(t->dim_blocks, t->ndim));
	for (ct_long k = 0; k < nblocks; k++) {
		if (t->blocks[k] != NULL) { orc_delete_dense_tensor(t->blocks[k]); free(t->blocks[k]); }
	}
	free(t->blocks); t->blocks = NULL;
	for (int i = 0; i < t->ndim; i++) { free(t->qnums_blocks[i]); free(t->qnums_logical[i]); }
	free(t->qnums_blocks); free(t->qnums_logical);
	free(t->axis_dir); free(t->dim_blocks); free(t->dim_logical);
	t->qnums_blocks = NULL; t->qnums_logical = NULL; t->axis_dir = NULL; t->dim_blocks = NULL; t->dim_logical = NULL;
}

// bst.c:213-270
ORC_EXPORT void orc_copy_block_sparse_tensor(const struct block_sparse_tensor* src, struct block_sparse_tensor* dst)
{
	orc_allocate_block_sparse_tensor(src->ndim, src->dim_logical, src->axis_dir, (const qnumber**)src->qnums_logical, dst);
	const ct_long nblocks = (src->ndim == 0 ? 1 : iprod(src->dim_blocks, src->ndim));
	for (ct_long k = 0; k < nblocks; k++) {
		if (src->blocks[k] == NULL) continue;
		memcpy(dst->blocks[k]->data, src->blocks[k]->data, (size_t)orc_dense_num_elements(src->blocks[k]) * sizeof(double));
	}
}

// bst.c:278-307
ORC_EXPORT struct dense_tensor* orc_block_sparse_tensor_get_block(const struct block_sparse_tensor* t, const qnumber* qnums)
{
	ct_long idx[16];
	for (int i = 0; i < t->ndim; i++) {
		idx[i] = -1;
		for (ct_long j = 0; j < t->dim_blocks[i]; j++) if (t->qnums_blocks[i][j] == qnums[i]) { idx[i] = j; break; }
		if (idx[i] == -1) return NULL;
	}
	return t->blocks[index_to_offset(t->ndim, t->dim_blocks, idx)];
}

// bst.c:316-334
ORC_EXPORT double orc_block_sparse_tensor_norm2(const struct block_sparse_tensor* t)
{
	double nrm = 0;
	const ct_long nblocks = (t->ndim == 0 ? 1 : iprod(t->dim_blocks, t->ndim));
	for (ct_long k = 0; k < nblocks; k++) {
		const struct dense_tensor* b = t->blocks[k];
		if (b == NULL) continue;
		const double bn = nrm2(orc_dense_num_elements(b), b->data, 1);
		nrm += bn * bn;
	}
	return sqrt(nrm);
}

// bst.c:343-350
ORC_EXPORT void orc_block_sparse_tensor_reverse_axis_directions(struct block_sparse_tensor* t)
{
	for (int i = 0; i < t->ndim; i++) t->axis_dir[i] = -t->axis_dir[i];
}

ORC_EXPORT void orc_scale_block_sparse_tensor(double alpha, struct block_sparse_tensor* t)
{
	const ct_long nblocks = (t->ndim == 0 ? 1 : iprod(t->dim_blocks, t->ndim));
	for (ct_long k = 0; k < nblocks; k++) {
		struct dense_tensor* b = t->blocks[k];
		if (b == NULL) continue;
		double* d = b->data;
		const ct_long n = orc_dense_num_elements(b);
		for (ct_long j = 0; j < n; j++) d[j] *= alpha;
	}
}

// bst.c:497-510 + dense_tensor.c:618-643
ORC_EXPORT void orc_block_sparse_tensor_fill_random_normal(double alpha, double shift, struct orc_rng_state* rng, struct block_sparse_tensor* t)
{
	const ct_long nblocks = (t->ndim == 0 ? 1 : iprod(t->dim_blocks, t->ndim));
	for (ct_long k = 0; k < nblocks; k++) {
		struct dense_tensor* b = t->blocks[k];
		if (b == NULL) continue;
		double* d = b->data;
		const ct_long n = orc_dense_num_elements(b);
		for (ct_long j = 0; j < n; j++) d[j] = alpha * orc_randn(rng) + shift;
	}
}

// logical indices carried by the entries of block 'ib' along one axis (fan-out map)
static void block_fanout(const struct block_sparse_tensor* t, int ax, ct_long ib, ct_long* map, ct_long len)
{
	ct_long c = 0;
	for (ct_long j = 0; j < t->dim_logical[ax]; j++) if (t->qnums_logical[ax][j] == t->qnums_blocks[ax][ib]) map[c++] = j;
	assert(c == len);
	(void)len;
}

// bst.c:518-620 (block-sparse -> dense) and bst.c:642-760 (dense -> block-sparse entries)
static void bst_dense_transfer(struct block_sparse_tensor* s, double* dense, bool to_dense)
{
	const int nd = s->ndim;
	const ct_long nblocks = iprod(s->dim_blocks, nd);
	for (ct_long k = 0; k < nblocks; k++) {
		struct dense_tensor* b = s->blocks[k];
		if (b == NULL) continue;
		ct_long ib[16];
		offset_to_index(nd, s->dim_blocks, k, ib);
		ct_long* maps[16];
		for (int i = 0; i < nd; i++) { maps[i] = xmalloc((size_t)b->dim[i] * sizeof(ct_long)); block_fanout(s, i, ib[i], maps[i], b->dim[i]); }
		ct_long ie[16] = { 0 }, il[16];
		const ct_long n = orc_dense_num_elements(b);
		double* bd = b->data;
		for (ct_long j = 0; j < n; j++, next_index(nd, b->dim, ie)) {
			for (int i = 0; i < nd; i++) il[i] = maps[i][ie[i]];
			const ct_long o = index_to_offset(nd, s->dim_logical, il);
			if (to_dense) dense[o] = bd[j]; else bd[j] = dense[o];
		}
		for (int i = 0; i < nd; i++) free(maps[i]);
	}
}

ORC_EXPORT void orc_block_sparse_to_dense_tensor(const struct block_sparse_tensor* s, struct dense_tensor* t)
{
	orc_allocate_dense_tensor(s->ndim, s->dim_logical, t);
	if (s->ndim == 0) { ((double*)t->data)[0] = ((double*)s->blocks[0]->data)[0]; return; }
	bst_dense_transfer((struct block_sparse_tensor*)s, t->data, true);
}

ORC_EXPORT void orc_dense_to_block_sparse_tensor_entries(const struct dense_tensor* t, struct block_sparse_tensor* s)
{
	bst_dense_transfer(s, t->data, false);
}

// bst.c:785-870
ORC_EXPORT void orc_block_sparse_tensor_transpose(const int* perm, const struct block_sparse_tensor* t, struct block_sparse_tensor* r)
{
	const int nd = t->ndim;
	if (nd == 0) { orc_copy_block_sparse_tensor(t, r); return; }
	ct_long dl[16];
	enum tensor_axis_direction ad[16];
	const qnumber* qn[16];
	for (int i = 0; i < nd; i++) { dl[i] = t->dim_logical[perm[i]]; ad[i] = t->axis_dir[perm[i]]; qn[i] = t->qnums_logical[perm[i]]; }
	// allocating with permuted logical quantum numbers reproduces the permuted block table
	orc_allocate_block_sparse_tensor(nd, dl, ad, qn, r);
	const ct_long nblocks = iprod(t->dim_blocks, nd);
	#pragma omp parallel for schedule(dynamic)
	for (ct_long k = 0; k < nblocks; k++) {
		const struct dense_tensor* bt = t->blocks[k];
		if (bt == NULL) continue;
		ct_long it[16], ir[16];
		offset_to_index(nd, t->dim_blocks, k, it);
		for (int i = 0; i < nd; i++) ir[i] = it[perm[i]];
		const ct_long j = index_to_offset(nd, r->dim_blocks, ir);
		assert(r->blocks[j] != NULL);
		orc_delete_dense_tensor(r->blocks[j]);
		orc_dense_tensor_transpose(perm, bt, r->blocks[j]);
	}
}

// bst.c:882-903: logical index -> (block index, entry index within block)
static void logical_to_block_maps(const qnumber* ql, ct_long dim, const qnumber* qb, ct_long nb, ct_long* map_block, ct_long* map_entry)
{
	ct_long* counter = xcalloc((size_t)nb, sizeof(ct_long));
	for (ct_long i = 0; i < dim; i++) {
		for (ct_long j = 0; j < nb; j++) {
			if (ql[i] == qb[j]) { if (map_block) map_block[i] = j; map_entry[i] = counter[j]++; break; }
		}
	}
	free(counter);
}

static ct_long find_qnum(const qnumber* qb, ct_long nb, qnumber q)
{
	for (ct_long j = 0; j < nb; j++) if (qb[j] == q) return j;
	return -1;
}

// fan-out for an explicit quantum number along a logical axis (bst.c:911-928)
static void fanout_q(const qnumber* ql, ct_long dim, qnumber q, ct_long* map)
{
	ct_long c = 0;
	for (ct_long j = 0; j < dim; j++) if (ql[j] == q) map[c++] = j;
}

// bst.c:950-1115
ORC_EXPORT void orc_block_sparse_tensor_flatten_axes(const struct block_sparse_tensor* t, int i_ax, enum tensor_axis_direction new_axis_dir, struct block_sparse_tensor* r)
{
	const int nd = t->ndim;
	assert(0 <= i_ax && i_ax + 1 < nd);
	{
		ct_long dl[16];
		enum tensor_axis_direction ad[16];
		const qnumber* qn[16];
		for (int i = 0; i < i_ax; i++) { dl[i] = t->dim_logical[i]; ad[i] = t->axis_dir[i]; qn[i] = t->qnums_logical[i]; }
		const ct_long d0 = t->dim_logical[i_ax], d1 = t->dim_logical[i_ax + 1];
		dl[i_ax] = d0 * d1;
		ad[i_ax] = new_axis_dir;
		qnumber* qf = xmalloc((size_t)(d0 * d1) * sizeof(qnumber));
		for (ct_long j = 0; j < d0; j++) for (ct_long k = 0; k < d1; k++)
			qf[j * d1 + k] = new_axis_dir * (t->axis_dir[i_ax] * t->qnums_logical[i_ax][j] + t->axis_dir[i_ax + 1] * t->qnums_logical[i_ax + 1][k]);
		qn[i_ax] = qf;
		for (int i = i_ax + 2; i < nd; i++) { dl[i - 1] = t->dim_logical[i]; ad[i - 1] = t->axis_dir[i]; qn[i - 1] = t->qnums_logical[i]; }
		orc_allocate_block_sparse_tensor(nd - 1, dl, ad, qn, r);
		free(qf);
	}
	ct_long* fanin = xmalloc((size_t)r->dim_logical[i_ax] * sizeof(ct_long));
	logical_to_block_maps(r->qnums_logical[i_ax], r->dim_logical[i_ax], r->qnums_blocks[i_ax], r->dim_blocks[i_ax], NULL, fanin);
	const ct_long nblocks = iprod(t->dim_blocks, nd);
	#pragma omp parallel for schedule(dynamic)
	for (ct_long k = 0; k < nblocks; k++) {
		const struct dense_tensor* bt = t->blocks[k];
		if (bt == NULL) continue;
		ct_long it[16], ir[16];
		offset_to_index(nd, t->dim_blocks, k, it);
		const qnumber q0 = t->qnums_blocks[i_ax][it[i_ax]], q1 = t->qnums_blocks[i_ax + 1][it[i_ax + 1]];
		const qnumber qf = new_axis_dir * (t->axis_dir[i_ax] * q0 + t->axis_dir[i_ax + 1] * q1);
		for (int i = 0; i < i_ax; i++) ir[i] = it[i];
		ir[i_ax] = find_qnum(r->qnums_blocks[i_ax], r->dim_blocks[i_ax], qf);
		for (int i = i_ax + 2; i < nd; i++) ir[i - 1] = it[i];
		struct dense_tensor* br = r->blocks[index_to_offset(nd - 1, r->dim_blocks, ir)];
		assert(br != NULL);
		const ct_long b0 = bt->dim[i_ax], b1 = bt->dim[i_ax + 1];
		ct_long* f0 = xmalloc((size_t)b0 * sizeof(ct_long));
		ct_long* f1 = xmalloc((size_t)b1 * sizeof(ct_long));
		fanout_q(t->qnums_logical[i_ax], t->dim_logical[i_ax], q0, f0);
		fanout_q(t->qnums_logical[i_ax + 1], t->dim_logical[i_ax + 1], q1, f1);
		ct_long* mb = xmalloc((size_t)(b0 * b1) * sizeof(ct_long));
		for (ct_long j0 = 0; j0 < b0; j0++) for (ct_long j1 = 0; j1 < b1; j1++)
			mb[j0 * b1 + j1] = fanin[f0[j0] * t->dim_logical[i_ax + 1] + f1[j1]];
		const ct_long nslices = iprod(bt->dim, i_ax + 2);
		const ct_long stride = iprod(bt->dim + i_ax + 2, nd - i_ax - 2);
		ct_long is[16] = { 0 }, isr[16];
		for (ct_long j = 0; j < nslices; j++, next_index(i_ax + 2, bt->dim, is)) {
			for (int i = 0; i < i_ax; i++) isr[i] = is[i];
			isr[i_ax] = mb[is[i_ax] * b1 + is[i_ax + 1]];
			const ct_long l = index_to_offset(i_ax + 1, br->dim, isr);
			memcpy((double*)br->data + l * stride, (const double*)bt->data + j * stride, (size_t)stride * sizeof(double));
		}
		free(mb); free(f1); free(f0);
	}
	free(fanin);
}

// bst.c:1123-1290
ORC_EXPORT void orc_block_sparse_tensor_split_axis(const struct block_sparse_tensor* t, int i_ax, const ct_long new_dim_logical[2], const enum tensor_axis_direction new_axis_dir[2], const qnumber* new_qnums_logical[2], struct block_sparse_tensor* r)
{
	const int nd = t->ndim;
	assert(new_dim_logical[0] * new_dim_logical[1] == t->dim_logical[i_ax]);
	{
		ct_long dl[16];
		enum tensor_axis_direction ad[16];
		const qnumber* qn[16];
		for (int i = 0; i < i_ax; i++) { dl[i] = t->dim_logical[i]; ad[i] = t->axis_dir[i]; qn[i] = t->qnums_logical[i]; }
		for (int c = 0; c < 2; c++) { dl[i_ax + c] = new_dim_logical[c]; ad[i_ax + c] = new_axis_dir[c]; qn[i_ax + c] = new_qnums_logical[c]; }
		for (int i = i_ax + 1; i < nd; i++) { dl[i + 1] = t->dim_logical[i]; ad[i + 1] = t->axis_dir[i]; qn[i + 1] = t->qnums_logical[i]; }
		orc_allocate_block_sparse_tensor(nd + 1, dl, ad, qn, r);
	}
	ct_long* fanin = xmalloc((size_t)t->dim_logical[i_ax] * sizeof(ct_long));
	logical_to_block_maps(t->qnums_logical[i_ax], t->dim_logical[i_ax], t->qnums_blocks[i_ax], t->dim_blocks[i_ax], NULL, fanin);
	const ct_long nblocks = iprod(r->dim_blocks, nd + 1);
	#pragma omp parallel for schedule(dynamic)
	for (ct_long k = 0; k < nblocks; k++) {
		struct dense_tensor* br = r->blocks[k];
		if (br == NULL) continue;
		ct_long ir[16], it[16];
		offset_to_index(nd + 1, r->dim_blocks, k, ir);
		const qnumber q0 = r->qnums_blocks[i_ax][ir[i_ax]], q1 = r->qnums_blocks[i_ax + 1][ir[i_ax + 1]];
		const qnumber qf = t->axis_dir[i_ax] * (r->axis_dir[i_ax] * q0 + r->axis_dir[i_ax + 1] * q1);
		for (int i = 0; i < i_ax; i++) it[i] = ir[i];
		it[i_ax] = find_qnum(t->qnums_blocks[i_ax], t->dim_blocks[i_ax], qf);
		assert(it[i_ax] >= 0);
		for (int i = i_ax + 1; i < nd; i++) it[i] = ir[i + 1];
		const struct dense_tensor* bt = t->blocks[index_to_offset(nd, t->dim_blocks, it)];
		assert(bt != NULL);
		const ct_long b0 = br->dim[i_ax], b1 = br->dim[i_ax + 1];
		ct_long* f0 = xmalloc((size_t)b0 * sizeof(ct_long));
		ct_long* f1 = xmalloc((size_t)b1 * sizeof(ct_long));
		fanout_q(r->qnums_logical[i_ax], r->dim_logical[i_ax], q0, f0);
		fanout_q(r->qnums_logical[i_ax + 1], r->dim_logical[i_ax + 1], q1, f1);
		ct_long* mb = xmalloc((size_t)(b0 * b1) * sizeof(ct_long));
		for (ct_long j0 = 0; j0 < b0; j0++) for (ct_long j1 = 0; j1 < b1; j1++)
			mb[j0 * b1 + j1] = fanin[f0[j0] * r->dim_logical[i_ax + 1] + f1[j1]];
		const ct_long nslices = iprod(br->dim, i_ax + 2);
		const ct_long stride = iprod(br->dim + i_ax + 2, nd + 1 - i_ax - 2);
		ct_long is[16] = { 0 }, ist[16];
		for (ct_long j = 0; j < nslices; j++, next_index(i_ax + 2, br->dim, is)) {
			for (int i = 0; i < i_ax; i++) ist[i] = is[i];
			ist[i_ax] = mb[is[i_ax] * b1 + is[i_ax + 1]];
			const ct_long l = index_to_offset(i_ax + 1, bt->dim, ist);
			memcpy((double*)br->data + j * stride, (const double*)bt->data + l * stride, (size_t)stride * sizeof(double));
		}
		free(mb); free(f1); free(f0);
	}
	free(fanin);
}

// bst.c:1446-1552
ORC_EXPORT void orc_block_sparse_tensor_slice(const struct block_sparse_tensor* t, int i_ax, const ct_long* ind, ct_long nind, struct block_sparse_tensor* r)
{
	const int nd = t->ndim;
	{
		ct_long dl[16];
		const qnumber* qn[16];
		qnumber* qs = xmalloc((size_t)nind * sizeof(qnumber));
		for (ct_long j = 0; j < nind; j++) qs[j] = t->qnums_logical[i_ax][ind[j]];
		for (int i = 0; i < nd; i++) { dl[i] = t->dim_logical[i]; qn[i] = t->qnums_logical[i]; }
		dl[i_ax] = nind; qn[i_ax] = qs;
		orc_allocate_block_sparse_tensor(nd, dl, t->axis_dir, qn, r);
		free(qs);
	}
	const ct_long nblocks = iprod(r->dim_blocks, nd);
	for (ct_long k = 0; k < nblocks; k++) {
		struct dense_tensor* br = r->blocks[k];
		if (br == NULL) continue;
		ct_long ir[16], it[16];
		offset_to_index(nd, r->dim_blocks, k, ir);
		const qnumber q = r->qnums_blocks[i_ax][ir[i_ax]];
		memcpy(it, ir, nd * sizeof(ct_long));
		it[i_ax] = find_qnum(t->qnums_blocks[i_ax], t->dim_blocks[i_ax], q);
		const struct dense_tensor* bt = t->blocks[index_to_offset(nd, t->dim_blocks, it)];
		assert(bt != NULL);
		ct_long* fo = xmalloc((size_t)br->dim[i_ax] * sizeof(ct_long));
		fanout_q(r->qnums_logical[i_ax], r->dim_logical[i_ax], q, fo);
		ct_long* fin = xcalloc((size_t)t->dim_logical[i_ax], sizeof(ct_long));
		ct_long c = 0;
		for (ct_long j = 0; j < t->dim_logical[i_ax]; j++) if (t->qnums_logical[i_ax][j] == q) fin[j] = c++;
		// dense_tensor_slice_fill along axis i_ax
		const ct_long outer = iprod(bt->dim, i_ax), inner = iprod(bt->dim + i_ax + 1, nd - i_ax - 1);
		const ct_long nt = bt->dim[i_ax], nr = br->dim[i_ax];
		for (ct_long o = 0; o < outer; o++) for (ct_long j = 0; j < nr; j++) {
			const ct_long js = fin[ind[fo[j]]];
			memcpy((double*)br->data + (o * nr + j) * inner, (const double*)bt->data + (o * nt + js) * inner, (size_t)inner * sizeof(double));
		}
		free(fin); free(fo);
	}
}

// bst.c:1560-1612 with dense_tensor.c:269-345
ORC_EXPORT void orc_block_sparse_tensor_cyclic_partial_trace(const struct block_sparse_tensor* t, int ndim_trace, struct block_sparse_tensor* r)
{
	if (ndim_trace == 0) { orc_copy_block_sparse_tensor(t, r); return; }
	const int nd = t->ndim, ndr = nd - 2 * ndim_trace;
	orc_allocate_block_sparse_tensor(ndr, t->dim_logical + ndim_trace, t->axis_dir + ndim_trace, (const qnumber**)(t->qnums_logical + ndim_trace), r);
	const ct_long nblocks_r = (ndr == 0 ? 1 : iprod(r->dim_blocks, ndr));
	const ct_long nblocks_p = iprod(t->dim_blocks, ndim_trace);
	for (ct_long k = 0; k < nblocks_r; k++) {
		struct dense_tensor* br = r->blocks[k];
		if (br == NULL) continue;
		ct_long it[16] = { 0 };
		if (ndr > 0) offset_to_index(ndr, r->dim_blocks, k, it + ndim_trace);
		for (ct_long j = 0; j < nblocks_p; j++, next_index(ndim_trace, t->dim_blocks, it)) {
			for (int i = 0; i < ndim_trace; i++) it[ndim_trace + ndr + i] = it[i];
			const struct dense_tensor* bt = t->blocks[index_to_offset(nd, t->dim_blocks, it)];
			assert(bt != NULL);
			const ct_long stride = iprod(bt->dim, ndim_trace);
			const ct_long nr = orc_dense_num_elements(br);
			const double* td = bt->data;
			double* rd = br->data;
			for (ct_long kk = 0; kk < stride; kk++) for (ct_long jj = 0; jj < nr; jj++) rd[jj] += td[(kk * nr + jj) * stride + kk];
		}
	}
}

// bst.c:1654-1705 with dense_tensor_multiply_pointwise_fill (dense_tensor.c:1300-1420)
ORC_EXPORT void orc_block_sparse_tensor_multiply_pointwise_vector(const struct block_sparse_tensor* s, const double* vec, ct_long nvec, enum tensor_axis_range axrange, struct block_sparse_tensor* r)
{
	const int nd = s->ndim;
	const int i_ax = (axrange == TENSOR_AXIS_RANGE_LEADING ? 0 : nd - 1);
	assert(nvec == s->dim_logical[i_ax]);
	(void)nvec;
	orc_allocate_block_sparse_tensor(nd, s->dim_logical, s->axis_dir, (const qnumber**)s->qnums_logical, r);
	const ct_long nblocks = iprod(r->dim_blocks, nd);
	for (ct_long k = 0; k < nblocks; k++) {
		struct dense_tensor* br = r->blocks[k];
		if (br == NULL) continue;
		const struct dense_tensor* bs = s->blocks[k];
		ct_long ir[16];
		offset_to_index(nd, r->dim_blocks, k, ir);
		ct_long* map = xmalloc((size_t)br->dim[i_ax] * sizeof(ct_long));
		block_fanout(r, i_ax, ir[i_ax], map, br->dim[i_ax]);
		const ct_long n = orc_dense_num_elements(br), nt = br->dim[i_ax];
		const double* sd = bs->data;
		double* rd = br->data;
		if (axrange == TENSOR_AXIS_RANGE_LEADING) {
			const ct_long inner = n / nt;
			for (ct_long j = 0; j < nt; j++) for (ct_long i = 0; i < inner; i++) rd[j * inner + i] = sd[j * inner + i] * vec[map[j]];
		}
		else {
			const ct_long outer = n / nt;
			for (ct_long i = 0; i < outer; i++) for (ct_long j = 0; j < nt; j++) rd[i * nt + j] = sd[i * nt + j] * vec[map[j]];
		}
		free(map);
	}
}

// bst.c:1826-2000
ORC_EXPORT void orc_block_sparse_tensor_dot(const struct block_sparse_tensor* s, enum tensor_axis_range axrange_s, const struct block_sparse_tensor* t, enum tensor_axis_range axrange_t, int ndim_mult, struct block_sparse_tensor* r)
{
	const int shift_s = (axrange_s == TENSOR_AXIS_RANGE_LEADING ? 0 : s->ndim - ndim_mult);
	const int shift_t = (axrange_t == TENSOR_AXIS_RANGE_LEADING ? 0 : t->ndim - ndim_mult);
	for (int i = 0; i < ndim_mult; i++) {
		assert(s->dim_logical[shift_s + i] == t->dim_logical[shift_t + i]);
		assert(s->axis_dir[shift_s + i] == -t->axis_dir[shift_t + i]);
	}
	if (s->ndim + t->ndim == 2 * ndim_mult) {
		orc_allocate_block_sparse_tensor(0, NULL, NULL, NULL, r);
		const ct_long nc = iprod(s->dim_blocks, ndim_mult);
		ct_long ic[16] = { 0 };
		for (ct_long m = 0; m < nc; m++, next_index(ndim_mult, s->dim_blocks, ic)) {
			qnumber qsum = 0;
			for (int i = 0; i < s->ndim; i++) qsum += s->axis_dir[i] * s->qnums_blocks[i][ic[i]];
			if (qsum != 0) continue;
			const ct_long ib = index_to_offset(s->ndim, s->dim_blocks, ic);
			orc_dense_tensor_dot_update(s->blocks[ib], axrange_s, t->blocks[ib], axrange_t, ndim_mult, r->blocks[0]);
		}
		return;
	}
	const int offset_s = (axrange_s == TENSOR_AXIS_RANGE_LEADING ? ndim_mult : 0);
	const int offset_t = (axrange_t == TENSOR_AXIS_RANGE_LEADING ? ndim_mult : 0);
	const int nds = s->ndim - ndim_mult, ndt = t->ndim - ndim_mult, ndr = nds + ndt;
	{
		ct_long dl[16];
		enum tensor_axis_direction ad[16];
		const qnumber* qn[16];
		for (int i = 0; i < nds; i++) { dl[i] = s->dim_logical[offset_s + i]; ad[i] = s->axis_dir[offset_s + i]; qn[i] = s->qnums_logical[offset_s + i]; }
		for (int i = 0; i < ndt; i++) { dl[nds + i] = t->dim_logical[offset_t + i]; ad[nds + i] = t->axis_dir[offset_t + i]; qn[nds + i] = t->qnums_logical[offset_t + i]; }
		orc_allocate_block_sparse_tensor(ndr, dl, ad, qn, r);
	}
	const ct_long nblocks = iprod(r->dim_blocks, ndr);
	const ct_long ncontract = iprod(t->dim_blocks + shift_t, ndim_mult);
	#pragma omp parallel for schedule(dynamic)
	for (ct_long k = 0; k < nblocks; k++) {
		struct dense_tensor* br = r->blocks[k];
		if (br == NULL) continue;
		ct_long ir[16], is[16], itt[16], ic[16] = { 0 };
		offset_to_index(ndr, r->dim_blocks, k, ir);
		for (ct_long m = 0; m < ncontract; m++, next_index(ndim_mult, t->dim_blocks + shift_t, ic)) {
			for (int i = 0; i < nds; i++) is[offset_s + i] = ir[i];
			for (int i = 0; i < ndim_mult; i++) is[shift_s + i] = ic[i];
			qnumber qsum = 0;
			for (int i = 0; i < s->ndim; i++) qsum += s->axis_dir[i] * s->qnums_blocks[i][is[i]];
			if (qsum != 0) continue;
			for (int i = 0; i < ndt; i++) itt[offset_t + i] = ir[nds + i];
			for (int i = 0; i < ndim_mult; i++) itt[shift_t + i] = ic[i];
			const struct dense_tensor* bs = s->blocks[index_to_offset(s->ndim, s->dim_blocks, is)];
			const struct dense_tensor* bt = t->blocks[index_to_offset(t->ndim, t->dim_blocks, itt)];
			assert(bs != NULL && bt != NULL);
			orc_dense_tensor_dot_update(bs, axrange_s, bt, axrange_t, ndim_mult, br);
		}
	}
}

// bst.c:3111-3180
ORC_EXPORT ct_long orc_block_sparse_tensor_num_elements_blocks(const struct block_sparse_tensor* t)
{
	const ct_long nblocks = (t->ndim == 0 ? 1 : iprod(t->dim_blocks, t->ndim));
	ct_long n = 0;
	for (ct_long k = 0; k < nblocks; k++) if (t->blocks[k] != NULL) n += orc_dense_num_elements(t->blocks[k]);
	return n;
}

ORC_EXPORT void orc_block_sparse_tensor_serialize_entries(const struct block_sparse_tensor* t, double* entries)
{
	const ct_long nblocks = (t->ndim == 0 ? 1 : iprod(t->dim_blocks, t->ndim));
	ct_long c = 0;
	for (ct_long k = 0; k < nblocks; k++) {
		const struct dense_tensor* b = t->blocks[k];
		if (b == NULL) continue;
		const ct_long n = orc_dense_num_elements(b);
		memcpy(entries + c, b->data, (size_t)n * sizeof(double));
		c += n;
	}
}

ORC_EXPORT void orc_block_sparse_tensor_deserialize_entries(struct block_sparse_tensor* t, const double* entries)
{
	const ct_long nblocks = (t->ndim == 0 ? 1 : iprod(t->dim_blocks, t->ndim));
	ct_long c = 0;
	for (ct_long k = 0; k < nblocks; k++) {
		struct dense_tensor* b = t->blocks[k];
		if (b == NULL) continue;
		const ct_long n = orc_dense_num_elements(b);
		memcpy(b->data, entries + c, (size_t)n * sizeof(double));
		c += n;
	}
}

// bst.c:2402-2536 (reduced mode)
ORC_EXPORT int orc_block_sparse_tensor_qr(const struct block_sparse_tensor* a, struct block_sparse_tensor* q, struct block_sparse_tensor* r)
{
	ct_long dim_interm = 0;
	qnumber* qi = xcalloc((size_t)a->dim_logical[0] + 1, sizeof(qnumber));
	for (ct_long j = 0; j < a->dim_blocks[1]; j++) for (ct_long i = 0; i < a->dim_blocks[0]; i++) {
		if (a->axis_dir[0] * a->qnums_blocks[0][i] + a->axis_dir[1] * a->qnums_blocks[1][j] != 0) continue;
		const struct dense_tensor* b = a->blocks[i * a->dim_blocks[1] + j];
		const ct_long k = lmin(b->dim[0], b->dim[1]);
		for (ct_long l = 0; l < k; l++) qi[dim_interm + l] = a->qnums_blocks[1][j];
		dim_interm += k;
	}
	const bool dummy = (dim_interm == 0);
	if (dummy) { dim_interm = 1; qi[0] = -a->axis_dir[0] * a->axis_dir[1] * a->qnums_blocks[0][0]; }
	const ct_long dq[2] = { a->dim_logical[0], dim_interm };
	const qnumber* qq[2] = { a->qnums_logical[0], qi };
	orc_allocate_block_sparse_tensor(2, dq, a->axis_dir, qq, q);
	const ct_long dr[2] = { dim_interm, a->dim_logical[1] };
	const enum tensor_axis_direction adr[2] = { -a->axis_dir[1], a->axis_dir[1] };
	const qnumber* qr[2] = { qi, a->qnums_logical[1] };
	orc_allocate_block_sparse_tensor(2, dr, adr, qr, r);
	free(qi);
	if (dummy) {
		const qnumber qb[2] = { a->qnums_blocks[0][0], -a->axis_dir[0] * a->axis_dir[1] * a->qnums_blocks[0][0] };
		struct dense_tensor* bq = orc_block_sparse_tensor_get_block(q, qb);
		((double*)bq->data)[0] = 1;
		return 0;
	}
	int ret = 0;
	for (ct_long i = 0; i < a->dim_blocks[0]; i++) for (ct_long j = 0; j < a->dim_blocks[1]; j++) {
		const struct dense_tensor* ba = a->blocks[i * a->dim_blocks[1] + j];
		if (ba == NULL) continue;
		const ct_long k = find_qnum(q->qnums_blocks[1], q->dim_blocks[1], a->qnums_blocks[1][j]);
		struct dense_tensor* bq = q->blocks[i * q->dim_blocks[1] + k];
		struct dense_tensor* br = r->blocks[k * r->dim_blocks[1] + j];
		if (orc_dense_tensor_qr_fill(ba, bq, br) != 0) ret = -1;
	}
	return ret;
}

// bst.c:2544-2678 (reduced mode)
ORC_EXPORT int orc_block_sparse_tensor_rq(const struct block_sparse_tensor* a, struct block_sparse_tensor* r, struct block_sparse_tensor* q)
{
	ct_long dim_interm = 0;
	qnumber* qi = xcalloc((size_t)a->dim_logical[1] + 1, sizeof(qnumber));
	for (ct_long i = 0; i < a->dim_blocks[0]; i++) for (ct_long j = 0; j < a->dim_blocks[1]; j++) {
		if (a->axis_dir[0] * a->qnums_blocks[0][i] + a->axis_dir[1] * a->qnums_blocks[1][j] != 0) continue;
		const struct dense_tensor* b = a->blocks[i * a->dim_blocks[1] + j];
		const ct_long k = lmin(b->dim[0], b->dim[1]);
		for (ct_long l = 0; l < k; l++) qi[dim_interm + l] = a->qnums_blocks[0][i];
		dim_interm += k;
	}
	const bool dummy = (dim_interm == 0);
	if (dummy) { dim_interm = 1; qi[0] = -a->axis_dir[0] * a->axis_dir[1] * a->qnums_blocks[1][0]; }
	const ct_long dr[2] = { a->dim_logical[0], dim_interm };
	const enum tensor_axis_direction adr[2] = { a->axis_dir[0], -a->axis_dir[0] };
	const qnumber* qr[2] = { a->qnums_logical[0], qi };
	orc_allocate_block_sparse_tensor(2, dr, adr, qr, r);
	const ct_long dq[2] = { dim_interm, a->dim_logical[1] };
	const qnumber* qq[2] = { qi, a->qnums_logical[1] };
	orc_allocate_block_sparse_tensor(2, dq, a->axis_dir, qq, q);
	free(qi);
	if (dummy) {
		const qnumber qb[2] = { -a->axis_dir[0] * a->axis_dir[1] * a->qnums_blocks[1][0], a->qnums_blocks[1][0] };
		struct dense_tensor* bq = orc_block_sparse_tensor_get_block(q, qb);
		((double*)bq->data)[0] = 1;
		return 0;
	}
	int ret = 0;
	for (ct_long i = 0; i < a->dim_blocks[0]; i++) for (ct_long j = 0; j < a->dim_blocks[1]; j++) {
		const struct dense_tensor* ba = a->blocks[i * a->dim_blocks[1] + j];
		if (ba == NULL) continue;
		const ct_long k = find_qnum(q->qnums_blocks[0], q->dim_blocks[0], a->qnums_blocks[0][i]);
		struct dense_tensor* br = r->blocks[i * r->dim_blocks[1] + k];
		struct dense_tensor* bq = q->blocks[k * q->dim_blocks[1] + j];
		if (orc_dense_tensor_rq_fill(ba, br, bq) != 0) ret = -1;
	}
	return ret;
}

// bst.c:2686-2848; *s is allocated (length *ns) and must be freed by the caller
ORC_EXPORT int orc_block_sparse_tensor_svd(const struct block_sparse_tensor* a, struct block_sparse_tensor* u, double** s, ct_long* ns, struct block_sparse_tensor* vh)
{
	ct_long dim_interm = 0;
	qnumber* qi = xcalloc((size_t)a->dim_logical[1] + 1, sizeof(qnumber));
	for (ct_long j = 0; j < a->dim_blocks[1]; j++) for (ct_long i = 0; i < a->dim_blocks[0]; i++) {
		if (a->axis_dir[0] * a->qnums_blocks[0][i] + a->axis_dir[1] * a->qnums_blocks[1][j] != 0) continue;
		const struct dense_tensor* b = a->blocks[i * a->dim_blocks[1] + j];
		const ct_long k = lmin(b->dim[0], b->dim[1]);
		for (ct_long l = 0; l < k; l++) qi[dim_interm + l] = a->qnums_blocks[1][j];
		dim_interm += k;
	}
	const bool dummy = (dim_interm == 0);
	if (dummy) { dim_interm = 1; qi[0] = -a->axis_dir[0] * a->axis_dir[1] * a->qnums_blocks[0][0]; }
	const ct_long du[2] = { a->dim_logical[0], dim_interm };
	const qnumber* qu[2] = { a->qnums_logical[0], qi };
	orc_allocate_block_sparse_tensor(2, du, a->axis_dir, qu, u);
	*ns = dim_interm;
	*s = xcalloc((size_t)dim_interm, sizeof(double));
	const ct_long dv[2] = { dim_interm, a->dim_logical[1] };
	const enum tensor_axis_direction adv[2] = { -a->axis_dir[1], a->axis_dir[1] };
	const qnumber* qv[2] = { qi, a->qnums_logical[1] };
	orc_allocate_block_sparse_tensor(2, dv, adv, qv, vh);
	free(qi);
	if (dummy) {
		const qnumber qb[2] = { a->qnums_blocks[0][0], -a->axis_dir[0] * a->axis_dir[1] * a->qnums_blocks[0][0] };
		struct dense_tensor* bu = orc_block_sparse_tensor_get_block(u, qb);
		((double*)bu->data)[0] = 1;
		return 0;
	}
	int ret = 0;
	for (ct_long i = 0; i < a->dim_blocks[0]; i++) for (ct_long j = 0; j < a->dim_blocks[1]; j++) {
		const struct dense_tensor* ba = a->blocks[i * a->dim_blocks[1] + j];
		if (ba == NULL) continue;
		const ct_long k = find_qnum(u->qnums_blocks[1], u->dim_blocks[1], a->qnums_blocks[1][j]);
		struct dense_tensor* bu = u->blocks[i * u->dim_blocks[1] + k];
		struct dense_tensor* bv = vh->blocks[k * vh->dim_blocks[1] + j];
		double* bs = xmalloc((size_t)bu->dim[1] * sizeof(double));
		if (orc_dense_tensor_svd_fill(ba, bu, bs, bv) != 0) ret = -1;
		ct_long l0;
		for (l0 = 0; l0 < u->dim_logical[1]; l0++) if (u->qnums_logical[1][l0] == a->qnums_blocks[1][j]) break;
		memcpy(*s + l0, bs, (size_t)bu->dim[1] * sizeof(double));
		free(bs);
	}
	return ret;
}

//________________________________________________________________________________________________________________________
//
// truncation (src/algorithm/truncation.c) and bond operations (src/algorithm/bond_ops.c)

// truncation.c:13-27
ORC_EXPORT double orc_von_neumann_entropy(const double* sigma, ct_long n)
{
	double s = 0;
	for (ct_long i = 0; i < n; i++) if (sigma[i] > 0) { const double sq = sigma[i] * sigma[i]; s -= sq * log(sq); }
	return s;
}

struct val_idx { double v; ct_long i; };
static int cmp_val_idx(const void* p1, const void* p2)
{
	const struct val_idx* x = p1;
	const struct val_idx* y = p2;
	if (x->v < y->v) return -1;
	if (y->v < x->v) return 1;
	return 0;
}

// truncation.c:110-222; 'ind' must have room for n entries; returns the number retained
ORC_EXPORT ct_long orc_retained_bond_indices(const double* sigma, ct_long n, double tol, bool relative_thresh, ct_long max_vdim, ct_long* ind, struct trunc_info* info)
{
	info->tol_eff = tol;
	struct val_idx* ss = xmalloc((size_t)n * sizeof(struct val_idx));
	for (ct_long i = 0; i < n; i++) { ss[i].v = sigma[i]; ss[i].i = i; }
	qsort(ss, n, sizeof(struct val_idx), cmp_val_idx);
	double sqsum = 0;
	for (ct_long i = 0; i < n; i++) { ss[i].v = ss[i].v * ss[i].v; sqsum += ss[i].v; }
	if (sqsum == 0) { free(ss); info->norm_sigma = 0; info->entropy = 0; return 0; }
	if (relative_thresh) for (ct_long i = 0; i < n; i++) ss[i].v /= sqsum;
	for (ct_long i = 1; i < n; i++) ss[i].v += ss[i - 1].v;
	if (max_vdim < n) {
		info->tol_eff = fmax(tol, ss[n - max_vdim - 1].v);
		for (ct_long i = 0; i < n - max_vdim; i++) ss[i].v = 0;
	}
	double* accum = xmalloc((size_t)n * sizeof(double));
	for (ct_long i = 0; i < n; i++) accum[ss[i].i] = ss[i].v;
	free(ss);
	ct_long num = 0;
	for (ct_long i = 0; i < n; i++) if (accum[i] > tol) ind[num++] = i;
	free(accum);
	if (num == 0) { info->norm_sigma = 0; info->entropy = 0; return 0; }
	double* ret = xmalloc((size_t)num * sizeof(double));
	for (ct_long i = 0; i < num; i++) ret[i] = sigma[ind[i]];
	info->norm_sigma = nrm2(num, ret, 1);
	for (ct_long i = 0; i < num; i++) ret[i] /= info->norm_sigma;
	info->entropy = orc_von_neumann_entropy(ret, num);
	free(ret);
	return num;
}

// bond_ops.c:15-138
ORC_EXPORT int orc_split_block_sparse_matrix_svd(const struct block_sparse_tensor* a, double tol, bool relative_thresh, ct_long max_vdim, bool renormalize, enum singular_value_distr svd_distr,
	struct block_sparse_tensor* a0, struct block_sparse_tensor* a1, struct trunc_info* info)
{
	struct block_sparse_tensor u, vh;
	double* s;
	ct_long ns;
	int ret = orc_block_sparse_tensor_svd(a, &u, &s, &ns, &vh);
	if (ret < 0) return ret;
	ct_long* ind = xmalloc((size_t)ns * sizeof(ct_long));
	const ct_long num = orc_retained_bond_indices(s, ns, tol, relative_thresh, max_vdim, ind, info);
	if (num == 0) {
		const ct_long i0[1] = { 0 };
		orc_block_sparse_tensor_slice(&u, 1, i0, 1, a0);
		orc_block_sparse_tensor_slice(&vh, 0, i0, 1, a1);
		orc_delete_block_sparse_tensor(&u); orc_delete_block_sparse_tensor(&vh);
		const double zero[1] = { 0 };
		struct block_sparse_tensor tmp;
		if (svd_distr == SVD_DISTR_LEFT) {
			orc_block_sparse_tensor_multiply_pointwise_vector(a0, zero, 1, TENSOR_AXIS_RANGE_TRAILING, &tmp);
			orc_delete_block_sparse_tensor(a0); *a0 = tmp;
		}
		else {
			orc_block_sparse_tensor_multiply_pointwise_vector(a1, zero, 1, TENSOR_AXIS_RANGE_LEADING, &tmp);
			orc_delete_block_sparse_tensor(a1); *a1 = tmp;
		}
		free(ind); free(s);
		return 0;
	}
	orc_block_sparse_tensor_slice(&u, 1, ind, num, a0);
	orc_block_sparse_tensor_slice(&vh, 0, ind, num, a1);
	orc_delete_block_sparse_tensor(&u); orc_delete_block_sparse_tensor(&vh);
	double* sr = xmalloc((size_t)num * sizeof(double));
	for (ct_long i = 0; i < num; i++) sr[i] = s[ind[i]];
	if (renormalize) {
		const double scale = nrm2(ns, s, 1) / info->norm_sigma;
		for (ct_long i = 0; i < num; i++) sr[i] *= scale;
	}
	struct block_sparse_tensor tmp;
	if (svd_distr == SVD_DISTR_LEFT) {
		orc_block_sparse_tensor_multiply_pointwise_vector(a0, sr, num, TENSOR_AXIS_RANGE_TRAILING, &tmp);
		orc_delete_block_sparse_tensor(a0); *a0 = tmp;
	}
	else {
		orc_block_sparse_tensor_multiply_pointwise_vector(a1, sr, num, TENSOR_AXIS_RANGE_LEADING, &tmp);
		orc_delete_block_sparse_tensor(a1); *a1 = tmp;
	}
	free(sr); free(ind); free(s);
	return 0;
}

//________________________________________________________________________________________________________________________
//
// Krylov subspace methods (src/util/krylov.c)

// krylov.c:24-89 (three-term recurrence without re-orthogonalisation)
ORC_EXPORT void orc_lanczos_iteration_d(ct_long n, orc_linear_func_d afunc, const void* adata, const double* vstart, int maxiter, double* alpha, double* beta, double* v, int* numiter)
{
	double* w = xmalloc((size_t)n * sizeof(double));
	memcpy(v, vstart, (size_t)n * sizeof(double));
	const double nrm = nrm2(n, v, 1);
	for (ct_long i = 0; i < n; i++) v[i] *= 1 / nrm;
	for (int j = 0; j < maxiter - 1; j++) {
		afunc(n, adata, &v[j * n], w);
		double a = 0;
		for (ct_long i = 0; i < n; i++) a += w[i] * v[j * n + i];
		alpha[j] = a;
		if (j > 0) for (ct_long i = 0; i < n; i++) w[i] -= alpha[j] * v[j * n + i] + beta[j - 1] * v[(j - 1) * n + i];
		else for (ct_long i = 0; i < n; i++) w[i] -= alpha[j] * v[j * n + i];
		beta[j] = nrm2(n, w, 1);
		if (beta[j] < 100 * n * 2.2204460492503131e-16) { *numiter = j + 1; free(w); return; }
		const double ib = 1 / beta[j];
		for (ct_long i = 0; i < n; i++) v[(j + 1) * n + i] = ib * w[i];
	}
	{
		const int j = maxiter - 1;
		afunc(n, adata, &v[j * n], w);
		double a = 0;
		for (ct_long i = 0; i < n; i++) a += w[i] * v[j * n + i];
		alpha[j] = a;
	}
	free(w);
	*numiter = maxiter;
}

// krylov.c:172-251 (release build: the debug-only symmetry probe at :178-212 is compiled out)
ORC_EXPORT int orc_eigensystem_krylov_symmetric(ct_long n, orc_linear_func_d afunc, const void* adata, const double* vstart, int maxiter, int numeig, double* lambda, double* u_ritz)
{
	double* alpha = xmalloc((size_t)maxiter * sizeof(double));
	double* beta = xmalloc((size_t)(maxiter > 1 ? maxiter - 1 : 1) * sizeof(double));
	double* v = xmalloc((size_t)(maxiter * n) * sizeof(double));
	int numiter;
	orc_lanczos_iteration_d(n, afunc, adata, vstart, maxiter, alpha, beta, v, &numiter);
	if (numiter < numeig) {
		fprintf(stderr, "Lanczos iteration stopped after %i iterations, cannot compute %i eigenvalues\n", numiter, numeig);
		free(v); free(beta); free(alpha);
		return -1;
	}
	double* z = xmalloc((size_t)(numiter * numiter) * sizeof(double));
	if (orc_symmetric_tridiag_eigh(numiter, alpha, beta, z) != 0) { free(z); free(v); free(beta); free(alpha); return -2; }
	memcpy(lambda, alpha, (size_t)numeig * sizeof(double));
	// u_ritz[i, e] = sum_j v[j, i] z[j, e]
	for (ct_long i = 0; i < n; i++) for (int e = 0; e < numeig; e++) {
		double acc = 0;
		for (int j = 0; j < numiter; j++) acc += v[j * n + i] * z[j + numiter * e];
		u_ritz[i * numeig + e] = acc;
	}
	free(z); free(v); free(beta); free(alpha);
	return 0;
}

// dense row-major matrix-vector product used by the Krylov fixtures: data = (ct_long n, double a[n*n])
ORC_EXPORT void orc_dense_matvec(ct_long n, const void* data, const double* v, double* ret)
{
	const double* a = (const double*)((const char*)data + sizeof(ct_long));
	for (ct_long i = 0; i < n; i++) { double acc = 0; for (ct_long j = 0; j < n; j++) acc += a[i * n + j] * v[j]; ret[i] = acc; }
}

//________________________________________________________________________________________________________________________
//
// MPS (src/state/mps.c)

// mps.c:36-53
ORC_EXPORT void orc_allocate_mps(int nsites, ct_long d, const qnumber* qsite, const ct_long* dim_bonds, const qnumber** qbonds, struct mps* mps)
{
	mps->nsites = nsites; mps->d = d;
	mps->qsite = xmalloc((size_t)d * sizeof(qnumber));
	memcpy(mps->qsite, qsite, (size_t)d * sizeof(qnumber));
	mps->a = xcalloc((size_t)nsites, sizeof(struct block_sparse_tensor));
	for (int i = 0; i < nsites; i++) {
		const ct_long dim[3] = { dim_bonds[i], d, dim_bonds[i + 1] };
		const enum tensor_axis_direction ad[3] = { TENSOR_AXIS_OUT, TENSOR_AXIS_OUT, TENSOR_AXIS_IN };
		const qnumber* qn[3] = { qbonds[i], qsite, qbonds[i + 1] };
		orc_allocate_block_sparse_tensor(3, dim, ad, qn, &mps->a[i]);
	}
}

ORC_EXPORT void orc_delete_mps(struct mps* mps)
{
	for (int i = 0; i < mps->nsites; i++) orc_delete_block_sparse_tensor(&mps->a[i]);
	free(mps->a); mps->a = NULL;
	free(mps->qsite); mps->qsite = NULL;
	mps->nsites = 0; mps->d = 0;
}

// mps.c:93-178
ORC_EXPORT void orc_construct_random_mps(int nsites, ct_long d, const qnumber* qsite, qnumber qnum_sector, ct_long max_vdim, struct orc_rng_state* rng, struct mps* mps)
{
	ct_long* db = xmalloc((size_t)(nsites + 1) * sizeof(ct_long));
	qnumber** qb = xmalloc((size_t)(nsites + 1) * sizeof(qnumber*));
	db[0] = 1; qb[0] = xmalloc(sizeof(qnumber)); qb[0][0] = 0;
	db[nsites] = 1; qb[nsites] = xmalloc(sizeof(qnumber)); qb[nsites][0] = qnum_sector;
	for (int l = 1; l < (nsites + 1) / 2; l++) {
		const ct_long df = db[l - 1] * d;
		qnumber* qf = xmalloc((size_t)df * sizeof(qnumber));
		for (ct_long i = 0; i < db[l - 1]; i++) for (ct_long j = 0; j < d; j++) qf[i * d + j] = qb[l - 1][i] + qsite[j];
		db[l] = lmin(df, max_vdim);
		qb[l] = xmalloc((size_t)db[l] * sizeof(qnumber));
		if (df <= max_vdim) memcpy(qb[l], qf, (size_t)db[l] * sizeof(qnumber));
		else {
			uint64_t* idx = xmalloc((size_t)max_vdim * sizeof(uint64_t));
			orc_rand_choice(df, max_vdim, rng, idx);
			for (ct_long i = 0; i < max_vdim; i++) qb[l][i] = qf[idx[i]];
			free(idx);
		}
		free(qf);
	}
	for (int l = nsites - 1; l >= (nsites + 1) / 2; l--) {
		const ct_long df = db[l + 1] * d;
		qnumber* qf = xmalloc((size_t)df * sizeof(qnumber));
		for (ct_long i = 0; i < db[l + 1]; i++) for (ct_long j = 0; j < d; j++) qf[i * d + j] = qb[l + 1][i] - qsite[j];
		db[l] = lmin(df, max_vdim);
		qb[l] = xmalloc((size_t)db[l] * sizeof(qnumber));
		if (df <= max_vdim) memcpy(qb[l], qf, (size_t)db[l] * sizeof(qnumber));
		else {
			uint64_t* idx = xmalloc((size_t)max_vdim * sizeof(uint64_t));
			orc_rand_choice(df, max_vdim, rng, idx);
			for (ct_long i = 0; i < max_vdim; i++) qb[l][i] = qf[idx[i]];
			free(idx);
		}
		free(qf);
	}
	orc_allocate_mps(nsites, d, qsite, db, (const qnumber**)qb, mps);
	for (int l = 0; l <= nsites; l++) free(qb[l]);
	free(qb); free(db);
	for (int l = 0; l < nsites; l++) {
		const ct_long nelem = iprod(mps->a[l].dim_logical, 3);
		orc_block_sparse_tensor_fill_random_normal(1.0 / sqrt((double)nelem), 0, rng, &mps->a[l]);
	}
}

// mps.c:513-550
ORC_EXPORT void orc_mps_local_orthonormalize_qr(struct block_sparse_tensor* a, struct block_sparse_tensor* a_next)
{
	const ct_long dl[2] = { a->dim_logical[0], a->dim_logical[1] };
	qnumber* ql[2];
	for (int i = 0; i < 2; i++) { ql[i] = xmalloc((size_t)dl[i] * sizeof(qnumber)); memcpy(ql[i], a->qnums_logical[i], (size_t)dl[i] * sizeof(qnumber)); }
	const enum tensor_axis_direction ad[2] = { TENSOR_AXIS_OUT, TENSOR_AXIS_OUT };
	struct block_sparse_tensor am, q, r;
	orc_block_sparse_tensor_flatten_axes(a, 0, TENSOR_AXIS_OUT, &am);
	orc_delete_block_sparse_tensor(a);
	orc_block_sparse_tensor_qr(&am, &q, &r);
	orc_delete_block_sparse_tensor(&am);
	orc_block_sparse_tensor_split_axis(&q, 0, dl, ad, (const qnumber**)ql, a);
	orc_delete_block_sparse_tensor(&q);
	for (int i = 0; i < 2; i++) free(ql[i]);
	struct block_sparse_tensor upd;
	orc_block_sparse_tensor_dot(&r, TENSOR_AXIS_RANGE_TRAILING, a_next, TENSOR_AXIS_RANGE_LEADING, 1, &upd);
	orc_delete_block_sparse_tensor(a_next);
	*a_next = upd;
	orc_delete_block_sparse_tensor(&r);
}

// mps.c:560-601
ORC_EXPORT void orc_mps_local_orthonormalize_rq(struct block_sparse_tensor* a, struct block_sparse_tensor* a_prev)
{
	const ct_long dr[2] = { a->dim_logical[1], a->dim_logical[2] };
	qnumber* qr_[2];
	for (int i = 0; i < 2; i++) { qr_[i] = xmalloc((size_t)dr[i] * sizeof(qnumber)); memcpy(qr_[i], a->qnums_logical[1 + i], (size_t)dr[i] * sizeof(qnumber)); }
	const enum tensor_axis_direction ad[2] = { TENSOR_AXIS_OUT, TENSOR_AXIS_IN };
	struct block_sparse_tensor am, r, q;
	orc_block_sparse_tensor_flatten_axes(a, 1, TENSOR_AXIS_IN, &am);
	orc_delete_block_sparse_tensor(a);
	orc_block_sparse_tensor_rq(&am, &r, &q);
	orc_delete_block_sparse_tensor(&am);
	orc_block_sparse_tensor_split_axis(&q, 1, dr, ad, (const qnumber**)qr_, a);
	orc_delete_block_sparse_tensor(&q);
	for (int i = 0; i < 2; i++) free(qr_[i]);
	struct block_sparse_tensor upd;
	orc_block_sparse_tensor_dot(a_prev, TENSOR_AXIS_RANGE_TRAILING, &r, TENSOR_AXIS_RANGE_LEADING, 1, &upd);
	orc_delete_block_sparse_tensor(a_prev);
	*a_prev = upd;
	orc_delete_block_sparse_tensor(&r);
}

// mps.c:609-760
ORC_EXPORT double orc_mps_orthonormalize_qr(struct mps* mps, enum mps_orthonormalization_mode mode)
{
	const int L = mps->nsites;
	if (mode == MPS_ORTHONORMAL_LEFT) {
		for (int i = 0; i < L - 1; i++) orc_mps_local_orthonormalize_qr(&mps->a[i], &mps->a[i + 1]);
		const int i = L - 1;
		const ct_long dt[3] = { mps->a[i].dim_logical[2], 1, mps->a[i].dim_logical[2] };
		const enum tensor_axis_direction ad[3] = { TENSOR_AXIS_OUT, TENSOR_AXIS_OUT, TENSOR_AXIS_IN };
		const qnumber qz[1] = { 0 };
		const qnumber* qn[3] = { mps->a[i].qnums_logical[2], qz, mps->a[i].qnums_logical[2] };
		struct block_sparse_tensor tail;
		orc_allocate_block_sparse_tensor(3, dt, ad, qn, &tail);
		((double*)tail.blocks[0]->data)[0] = 1;
		orc_mps_local_orthonormalize_qr(&mps->a[i], &tail);
		double norm = 0;
		if (tail.blocks[0] != NULL) {
			norm = ((double*)tail.blocks[0]->data)[0];
			if (norm < 0) { orc_scale_block_sparse_tensor(-1, &mps->a[i]); norm = -norm; }
		}
		orc_delete_block_sparse_tensor(&tail);
		return norm;
	}
	for (int i = L - 1; i > 0; i--) orc_mps_local_orthonormalize_rq(&mps->a[i], &mps->a[i - 1]);
	const ct_long dh[3] = { mps->a[0].dim_logical[0], 1, mps->a[0].dim_logical[0] };
	const enum tensor_axis_direction ad[3] = { TENSOR_AXIS_OUT, TENSOR_AXIS_OUT, TENSOR_AXIS_IN };
	const qnumber qz[1] = { 0 };
	const qnumber* qn[3] = { mps->a[0].qnums_logical[0], qz, mps->a[0].qnums_logical[0] };
	struct block_sparse_tensor head;
	orc_allocate_block_sparse_tensor(3, dh, ad, qn, &head);
	((double*)head.blocks[0]->data)[0] = 1;
	orc_mps_local_orthonormalize_rq(&mps->a[0], &head);
	double norm = 0;
	if (head.blocks[0] != NULL) {
		norm = ((double*)head.blocks[0]->data)[0];
		if (norm < 0) { orc_scale_block_sparse_tensor(-1, &mps->a[0]); norm = -norm; }
	}
	orc_delete_block_sparse_tensor(&head);
	return norm;
}

// mps.c:1166-1180
ORC_EXPORT void orc_mps_merge_tensor_pair(const struct block_sparse_tensor* a0, const struct block_sparse_tensor* a1, struct block_sparse_tensor* a)
{
	struct block_sparse_tensor t;
	orc_block_sparse_tensor_dot(a0, TENSOR_AXIS_RANGE_TRAILING, a1, TENSOR_AXIS_RANGE_LEADING, 1, &t);
	orc_block_sparse_tensor_flatten_axes(&t, 1, TENSOR_AXIS_OUT, a);
	orc_delete_block_sparse_tensor(&t);
}

// mps.c:1119-1160
ORC_EXPORT int orc_mps_split_tensor_svd(const struct block_sparse_tensor* a, const ct_long d[2], const qnumber* new_qsite[2], double tol, ct_long max_vdim, bool renormalize,
	enum singular_value_distr svd_distr, struct block_sparse_tensor* a0, struct block_sparse_tensor* a1, struct trunc_info* info)
{
	struct block_sparse_tensor a2, tmp, am;
	const enum tensor_axis_direction ad[2] = { TENSOR_AXIS_OUT, TENSOR_AXIS_OUT };
	orc_block_sparse_tensor_split_axis(a, 1, d, ad, new_qsite, &a2);
	orc_block_sparse_tensor_flatten_axes(&a2, 0, TENSOR_AXIS_OUT, &tmp);
	orc_block_sparse_tensor_flatten_axes(&tmp, 1, TENSOR_AXIS_IN, &am);
	orc_delete_block_sparse_tensor(&tmp);
	struct block_sparse_tensor m0, m1;
	int ret = orc_split_block_sparse_matrix_svd(&am, tol, true, max_vdim, renormalize, svd_distr, &m0, &m1, info);
	orc_delete_block_sparse_tensor(&am);
	if (ret < 0) { orc_delete_block_sparse_tensor(&a2); return ret; }
	orc_block_sparse_tensor_split_axis(&m0, 0, a2.dim_logical, a2.axis_dir, (const qnumber**)a2.qnums_logical, a0);
	orc_block_sparse_tensor_split_axis(&m1, 1, a2.dim_logical + 2, a2.axis_dir + 2, (const qnumber**)(a2.qnums_logical + 2), a1);
	orc_delete_block_sparse_tensor(&m1); orc_delete_block_sparse_tensor(&m0);
	orc_delete_block_sparse_tensor(&a2);
	return 0;
}

// mps.c:1188-1210: contract all site tensors; the state vector is returned as a new array
ORC_EXPORT void orc_mps_to_statevector(const struct mps* mps, double** vec, ct_long* n)
{
	struct block_sparse_tensor t;
	orc_copy_block_sparse_tensor(&mps->a[0], &t);
	for (int i = 1; i < mps->nsites; i++) {
		struct block_sparse_tensor t2;
		orc_mps_merge_tensor_pair(&t, &mps->a[i], &t2);
		orc_delete_block_sparse_tensor(&t);
		t = t2;
	}
	struct dense_tensor dt;
	orc_block_sparse_to_dense_tensor(&t, &dt);
	*n = orc_dense_num_elements(&dt);
	*vec = dt.data;
	free(dt.dim);
	orc_delete_block_sparse_tensor(&t);
}

//________________________________________________________________________________________________________________________
//
// MPO (src/operator/mpo.c, src/operator/hamiltonian.c)

ORC_EXPORT void orc_delete_mpo(struct mpo* mpo)
{
	for (int i = 0; i < mpo->nsites; i++) orc_delete_block_sparse_tensor(&mpo->a[i]);
	free(mpo->a); mpo->a = NULL;
	free(mpo->qsite); mpo->qsite = NULL;
	mpo->nsites = 0; mpo->d = 0;
}

// mpo.c:255-282
ORC_EXPORT void orc_mpo_merge_tensor_pair(const struct block_sparse_tensor* a0, const struct block_sparse_tensor* a1, struct block_sparse_tensor* a)
{
	struct block_sparse_tensor t, tp, tmp;
	orc_block_sparse_tensor_dot(a0, TENSOR_AXIS_RANGE_TRAILING, a1, TENSOR_AXIS_RANGE_LEADING, 1, &t);
	const int perm[6] = { 0, 1, 3, 2, 4, 5 };
	orc_block_sparse_tensor_transpose(perm, &t, &tp);
	orc_delete_block_sparse_tensor(&t);
	orc_block_sparse_tensor_flatten_axes(&tp, 1, TENSOR_AXIS_OUT, &tmp);
	orc_delete_block_sparse_tensor(&tp);
	orc_block_sparse_tensor_flatten_axes(&tmp, 2, TENSOR_AXIS_IN, a);
	orc_delete_block_sparse_tensor(&tmp);
}

// mpo.c:287-320: dense matrix (n x n) of the whole operator, returned as a new array
ORC_EXPORT void orc_mpo_to_matrix(const struct mpo* mpo, double** mat, ct_long* n)
{
	struct block_sparse_tensor t;
	orc_copy_block_sparse_tensor(&mpo->a[0], &t);
	for (int i = 1; i < mpo->nsites; i++) {
		struct block_sparse_tensor t2;
		orc_mpo_merge_tensor_pair(&t, &mpo->a[i], &t2);
		orc_delete_block_sparse_tensor(&t);
		t = t2;
	}
	struct dense_tensor dt;
	orc_block_sparse_to_dense_tensor(&t, &dt);
	*n = t.dim_logical[1];
	*mat = dt.data;
	free(dt.dim);
	orc_delete_block_sparse_tensor(&t);
}

// assemble an MPO from dense bulk tensors W[ml][p'][p][mr] of a finite-state machine with
// start state 0 and final state nb-1 (first site keeps row 0, last site keeps column nb-1)
static void mpo_from_fsm(int nsites, ct_long d, const qnumber* qsite, int nb, const qnumber* qbond, const double* w, struct mpo* mpo)
{
	mpo->nsites = nsites; mpo->d = d;
	mpo->qsite = xmalloc((size_t)d * sizeof(qnumber));
	memcpy(mpo->qsite, qsite, (size_t)d * sizeof(qnumber));
	mpo->a = xcalloc((size_t)nsites, sizeof(struct block_sparse_tensor));
	const qnumber qzero[1] = { 0 };
	for (int l = 0; l < nsites; l++) {
		const int ml0 = (l == 0 ? 0 : 0), nml = (l == 0 ? 1 : nb);
		const int mr0 = (l == nsites - 1 ? nb - 1 : 0), nmr = (l == nsites - 1 ? 1 : nb);
		const ct_long dim[4] = { nml, d, d, nmr };
		const enum tensor_axis_direction ad[4] = { TENSOR_AXIS_OUT, TENSOR_AXIS_OUT, TENSOR_AXIS_IN, TENSOR_AXIS_IN };
		const qnumber* qn[4] = { l == 0 ? qzero : qbond, qsite, qsite, l == nsites - 1 ? qzero : qbond };
		struct dense_tensor dt;
		orc_allocate_dense_tensor(4, dim, &dt);
		double* dd = dt.data;
		for (int a = 0; a < nml; a++) for (ct_long p = 0; p < d; p++) for (ct_long q = 0; q < d; q++) for (int b = 0; b < nmr; b++)
			dd[((a * d + p) * d + q) * nmr + b] = w[(((ml0 + a) * d + p) * d + q) * nb + (mr0 + b)];
		orc_allocate_block_sparse_tensor(4, dim, ad, qn, &mpo->a[l]);
		orc_dense_to_block_sparse_tensor_entries(&dt, &mpo->a[l]);
		orc_delete_dense_tensor(&dt);
	}
}

static void fsm_put(double* w, int nb, ct_long d, int a, int b, double coeff, const double* op)
{
	for (ct_long p = 0; p < d; p++) for (ct_long q = 0; q < d; q++) w[(((ct_long)a * d + p) * d + q) * nb + b] += coeff * op[p * d + q];
}

// the operator of hamiltonian.c:102-155: sum_j J (Sx Sx + Sy Sy + D Sz Sz) - h Sz,
// qsite = {+1, -1} (2 Sz); bond basis: 0 start, 1 after S+, 2 after S-, 3 after Sz, 4 final
ORC_EXPORT void orc_construct_heisenberg_xxz_1d_mpo(int nsites, double J, double D, double h, struct mpo* mpo)
{
	const ct_long d = 2;
	const qnumber qsite[2] = { 1, -1 };
	const double id[4] = { 1, 0, 0, 1 }, sup[4] = { 0, 1, 0, 0 }, sdn[4] = { 0, 0, 1, 0 }, sz[4] = { 0.5, 0, 0, -0.5 };
	const int nb = 5;
	const qnumber qbond[5] = { 0, 2, -2, 0, 0 };
	double* w = xcalloc((size_t)(nb * nb * d * d), sizeof(double));
	fsm_put(w, nb, d, 0, 0, 1, id);
	fsm_put(w, nb, d, 0, 1, 1, sup);
	fsm_put(w, nb, d, 1, 4, 0.5 * J, sdn);
	fsm_put(w, nb, d, 0, 2, 1, sdn);
	fsm_put(w, nb, d, 2, 4, 0.5 * J, sup);
	fsm_put(w, nb, d, 0, 3, 1, sz);
	fsm_put(w, nb, d, 3, 4, J * D, sz);
	fsm_put(w, nb, d, 0, 4, -h, sz);
	fsm_put(w, nb, d, 4, 4, 1, id);
	mpo_from_fsm(nsites, d, qsite, nb, qbond, w, mpo);
	free(w);
}

static qnumber encode_pair(qnumber qa, qnumber qb) { return qa * 65536 + qb; }

static void kron2(const double* a, const double* b, double* r)
{
	for (int i1 = 0; i1 < 2; i1++) for (int i2 = 0; i2 < 2; i2++) for (int j1 = 0; j1 < 2; j1++) for (int j2 = 0; j2 < 2; j2++)
		r[(i1 * 2 + i2) * 4 + (j1 * 2 + j2)] = a[i1 * 2 + j1] * b[i2 * 2 + j2];
}

// the operator of hamiltonian.c:240-366: -t sum (hopping with Jordan-Wigner strings)
// - mu sum (n_up + n_dn) + u sum (n_up - 1/2)(n_dn - 1/2); site basis |n_up n_dn> with
// index 2 n_up + n_dn; qsite = pack(N, 2Sz); bond basis: 0 start, 1..4 open hops, 5 final
ORC_EXPORT void orc_construct_fermi_hubbard_1d_mpo(int nsites, double t, double u, double mu, struct mpo* mpo)
{
	const ct_long d = 4;
	const qnumber qsite[4] = { encode_pair(0, 0), encode_pair(1, -1), encode_pair(1, 1), encode_pair(2, 0) };
	const double id[4] = { 1, 0, 0, 1 }, cr[4] = { 0, 0, 1, 0 }, an[4] = { 0, 1, 0, 0 }, z[4] = { 1, 0, 0, -1 };
	double Id[16], CZ[16], AZ[16], IC[16], IA[16], CI[16], AI[16], ZA[16], ZC[16], loc[16];
	kron2(id, id, Id); kron2(cr, z, CZ); kron2(an, z, AZ); kron2(id, cr, IC); kron2(id, an, IA);
	kron2(cr, id, CI); kron2(an, id, AI); kron2(z, an, ZA); kron2(z, cr, ZC);
	const double ntot[4] = { 0, 1, 1, 2 }, nint[4] = { 0.25, -0.25, -0.25, 0.25 };
	for (int i = 0; i < 16; i++) loc[i] = 0;
	for (int i = 0; i < 4; i++) loc[i * 4 + i] = -mu * ntot[i] + u * nint[i];
	const int nb = 6;
	const qnumber qbond[6] = { 0, encode_pair(1, 1), encode_pair(-1, -1), encode_pair(1, -1), encode_pair(-1, 1), 0 };
	double* w = xcalloc((size_t)(nb * nb * d * d), sizeof(double));
	fsm_put(w, nb, d, 0, 0, 1, Id);
	fsm_put(w, nb, d, 0, 1, 1, CZ); fsm_put(w, nb, d, 1, 5, -t, AI);
	fsm_put(w, nb, d, 0, 2, 1, AZ); fsm_put(w, nb, d, 2, 5, -t, CI);
	fsm_put(w, nb, d, 0, 3, 1, IC); fsm_put(w, nb, d, 3, 5, -t, ZA);
	fsm_put(w, nb, d, 0, 4, 1, IA); fsm_put(w, nb, d, 4, 5, -t, ZC);
	fsm_put(w, nb, d, 0, 5, 1, loc);
	fsm_put(w, nb, d, 5, 5, 1, Id);
	mpo_from_fsm(nsites, d, qsite, nb, qbond, w, mpo);
	free(w);
}

//________________________________________________________________________________________________________________________
//
// chain operations (src/algorithm/chain_ops.c)

// chain_ops.c:14-46
ORC_EXPORT void orc_create_dummy_operator_block_right(const struct block_sparse_tensor* a, const struct block_sparse_tensor* b, const struct block_sparse_tensor* w, struct block_sparse_tensor* r)
{
	const ct_long dim[6] = { 1, 1, 1, 1, 1, 1 };
	const enum tensor_axis_direction ad[6] = { TENSOR_AXIS_OUT, TENSOR_AXIS_OUT, TENSOR_AXIS_IN, TENSOR_AXIS_IN, TENSOR_AXIS_IN, TENSOR_AXIS_OUT };
	const qnumber* qn[6] = { a->qnums_logical[2], w->qnums_logical[3], b->qnums_logical[2], a->qnums_logical[2], w->qnums_logical[3], b->qnums_logical[2] };
	struct block_sparse_tensor s, t;
	orc_allocate_block_sparse_tensor(6, dim, ad, qn, &s);
	((double*)s.blocks[0]->data)[0] = 1;
	orc_block_sparse_tensor_flatten_axes(&s, 4, TENSOR_AXIS_IN, &t);
	orc_delete_block_sparse_tensor(&s);
	orc_block_sparse_tensor_flatten_axes(&t, 3, TENSOR_AXIS_IN, r);
	orc_delete_block_sparse_tensor(&t);
}

// chain_ops.c:53-85
ORC_EXPORT void orc_create_dummy_operator_block_left(const struct block_sparse_tensor* a, const struct block_sparse_tensor* b, const struct block_sparse_tensor* w, struct block_sparse_tensor* l)
{
	const ct_long dim[6] = { 1, 1, 1, 1, 1, 1 };
	const enum tensor_axis_direction ad[6] = { TENSOR_AXIS_OUT, TENSOR_AXIS_OUT, TENSOR_AXIS_IN, TENSOR_AXIS_IN, TENSOR_AXIS_IN, TENSOR_AXIS_OUT };
	const qnumber* qn[6] = { a->qnums_logical[0], w->qnums_logical[0], b->qnums_logical[0], a->qnums_logical[0], w->qnums_logical[0], b->qnums_logical[0] };
	struct block_sparse_tensor s, t;
	orc_allocate_block_sparse_tensor(6, dim, ad, qn, &s);
	((double*)s.blocks[0]->data)[0] = 1;
	orc_block_sparse_tensor_flatten_axes(&s, 0, TENSOR_AXIS_OUT, &t);
	orc_delete_block_sparse_tensor(&s);
	orc_block_sparse_tensor_flatten_axes(&t, 0, TENSOR_AXIS_OUT, l);
	orc_delete_block_sparse_tensor(&t);
}

// shallow copy of 'b' with reversed axis directions (real entries: no conjugation needed)
static struct block_sparse_tensor reversed_view(const struct block_sparse_tensor* b)
{
	struct block_sparse_tensor br = *b;
	br.axis_dir = xmalloc((size_t)b->ndim * sizeof(enum tensor_axis_direction));
	for (int i = 0; i < b->ndim; i++) br.axis_dir[i] = -b->axis_dir[i];
	return br;
}

// chain_ops.c:116-165
ORC_EXPORT void orc_contraction_operator_step_right(const struct block_sparse_tensor* a, const struct block_sparse_tensor* b, const struct block_sparse_tensor* w, const struct block_sparse_tensor* r, struct block_sparse_tensor* r_next)
{
	struct block_sparse_tensor s, t;
	orc_block_sparse_tensor_dot(a, TENSOR_AXIS_RANGE_TRAILING, r, TENSOR_AXIS_RANGE_LEADING, 1, &s);
	const int perm0[5] = { 1, 2, 0, 3, 4 };
	orc_block_sparse_tensor_transpose(perm0, &s, &t);
	orc_delete_block_sparse_tensor(&s);
	orc_block_sparse_tensor_dot(w, TENSOR_AXIS_RANGE_TRAILING, &t, TENSOR_AXIS_RANGE_LEADING, 2, &s);
	orc_delete_block_sparse_tensor(&t);
	const int perm1[5] = { 4, 2, 0, 1, 3 };
	orc_block_sparse_tensor_transpose(perm1, &s, &t);
	orc_delete_block_sparse_tensor(&s);
	struct block_sparse_tensor br = reversed_view(b);
	orc_block_sparse_tensor_dot(&t, TENSOR_AXIS_RANGE_TRAILING, &br, TENSOR_AXIS_RANGE_TRAILING, 2, &s);
	free(br.axis_dir);
	orc_delete_block_sparse_tensor(&t);
	const int perm2[4] = { 1, 2, 3, 0 };
	orc_block_sparse_tensor_transpose(perm2, &s, r_next);
	orc_delete_block_sparse_tensor(&s);
}

// chain_ops.c:196-245
ORC_EXPORT void orc_contraction_operator_step_left(const struct block_sparse_tensor* a, const struct block_sparse_tensor* b, const struct block_sparse_tensor* w, const struct block_sparse_tensor* l, struct block_sparse_tensor* l_next)
{
	struct block_sparse_tensor s, t;
	struct block_sparse_tensor br = reversed_view(b);
	orc_block_sparse_tensor_dot(l, TENSOR_AXIS_RANGE_TRAILING, &br, TENSOR_AXIS_RANGE_LEADING, 1, &s);
	free(br.axis_dir);
	const int perm0[5] = { 0, 1, 4, 2, 3 };
	orc_block_sparse_tensor_transpose(perm0, &s, &t);
	orc_delete_block_sparse_tensor(&s);
	orc_block_sparse_tensor_dot(&t, TENSOR_AXIS_RANGE_TRAILING, w, TENSOR_AXIS_RANGE_LEADING, 2, &s);
	orc_delete_block_sparse_tensor(&t);
	const int perm1[5] = { 1, 3, 4, 2, 0 };
	orc_block_sparse_tensor_transpose(perm1, &s, &t);
	orc_delete_block_sparse_tensor(&s);
	orc_block_sparse_tensor_dot(a, TENSOR_AXIS_RANGE_LEADING, &t, TENSOR_AXIS_RANGE_LEADING, 2, &s);
	orc_delete_block_sparse_tensor(&t);
	const int perm2[4] = { 3, 0, 1, 2 };
	orc_block_sparse_tensor_transpose(perm2, &s, l_next);
	orc_delete_block_sparse_tensor(&s);
}

// chain_ops.c:253-267
ORC_EXPORT void orc_compute_right_operator_blocks(const struct mps* psi, const struct mps* chi, const struct mpo* op, struct block_sparse_tensor* r_list)
{
	const int L = op->nsites;
	orc_create_dummy_operator_block_right(&psi->a[L - 1], &chi->a[L - 1], &op->a[L - 1], &r_list[L - 1]);
	for (int i = L - 1; i > 0; i--) orc_contraction_operator_step_right(&psi->a[i], &chi->a[i], &op->a[i], &r_list[i], &r_list[i - 1]);
}

// chain_ops.c:353-390
ORC_EXPORT void orc_apply_local_hamiltonian(const struct block_sparse_tensor* a, const struct block_sparse_tensor* w, const struct block_sparse_tensor* l, const struct block_sparse_tensor* r, struct block_sparse_tensor* b)
{
	struct block_sparse_tensor s, t, k;
	orc_block_sparse_tensor_dot(a, TENSOR_AXIS_RANGE_TRAILING, r, TENSOR_AXIS_RANGE_LEADING, 1, &s);
	const int perm0[5] = { 1, 2, 0, 3, 4 };
	orc_block_sparse_tensor_transpose(perm0, &s, &t);
	orc_delete_block_sparse_tensor(&s);
	orc_block_sparse_tensor_dot(w, TENSOR_AXIS_RANGE_TRAILING, &t, TENSOR_AXIS_RANGE_LEADING, 2, &s);
	orc_delete_block_sparse_tensor(&t);
	const int perm1[5] = { 2, 0, 1, 3, 4 };
	orc_block_sparse_tensor_transpose(perm1, &s, &t);
	orc_delete_block_sparse_tensor(&s);
	const int perm2[4] = { 0, 3, 1, 2 };
	orc_block_sparse_tensor_transpose(perm2, l, &k);
	orc_block_sparse_tensor_dot(&k, TENSOR_AXIS_RANGE_TRAILING, &t, TENSOR_AXIS_RANGE_LEADING, 2, &s);
	orc_delete_block_sparse_tensor(&k);
	orc_delete_block_sparse_tensor(&t);
	orc_block_sparse_tensor_cyclic_partial_trace(&s, 1, b);
	orc_delete_block_sparse_tensor(&s);
}

//________________________________________________________________________________________________________________________
//
// DMRG (src/algorithm/dmrg.c)

struct local_ham { const struct block_sparse_tensor* w; const struct block_sparse_tensor* l; const struct block_sparse_tensor* r; struct block_sparse_tensor* a; };

// dmrg.c:28-50
static void apply_local_hamiltonian_wrapper(ct_long n, const void* data, const double* v, double* ret)
{
	const struct local_ham* h = data;
	(void)n;
	orc_block_sparse_tensor_deserialize_entries(h->a, v);
	struct block_sparse_tensor ha;
	orc_apply_local_hamiltonian(h->a, h->w, h->l, h->r, &ha);
	orc_block_sparse_tensor_serialize_entries(&ha, ret);
	orc_delete_block_sparse_tensor(&ha);
}

// dmrg.c:86-147
static int minimize_local_energy(const struct block_sparse_tensor* w, const struct block_sparse_tensor* l, const struct block_sparse_tensor* r,
	const struct block_sparse_tensor* a_start, int maxiter, double* en_min, struct block_sparse_tensor* a_opt)
{
	const ct_long n = orc_block_sparse_tensor_num_elements_blocks(a_start);
	double* vstart = xmalloc((size_t)n * sizeof(double));
	orc_block_sparse_tensor_serialize_entries(a_start, vstart);
	orc_allocate_block_sparse_tensor(a_start->ndim, a_start->dim_logical, a_start->axis_dir, (const qnumber**)a_start->qnums_logical, a_opt);
	struct local_ham h = { .w = w, .l = l, .r = r, .a = a_opt };
	double* u = xmalloc((size_t)n * sizeof(double));
	int ret = orc_eigensystem_krylov_symmetric(n, apply_local_hamiltonian_wrapper, &h, vstart, maxiter, 1, en_min, u);
	if (ret == 0) orc_block_sparse_tensor_deserialize_entries(a_opt, u);
	free(u); free(vstart);
	return ret;
}

// dmrg.c:262-399
ORC_EXPORT int orc_dmrg_twosite(const struct mpo* H, int num_sweeps, int maxiter_lanczos, double tol_split, ct_long max_vdim, struct mps* psi, double* en_sweeps, double* entropy)
{
	const int L = H->nsites;
	const double nrm = orc_mps_orthonormalize_qr(psi, MPS_ORTHONORMAL_RIGHT);
	if (nrm == 0) printf("Warning: in 'dmrg_twosite': initial MPS has norm zero (possibly due to mismatching quantum numbers)\n");
	struct block_sparse_tensor* lb = xmalloc((size_t)L * sizeof(struct block_sparse_tensor));
	struct block_sparse_tensor* rb = xmalloc((size_t)L * sizeof(struct block_sparse_tensor));
	orc_compute_right_operator_blocks(psi, psi, H, rb);
	orc_create_dummy_operator_block_left(&psi->a[0], &psi->a[0], &H->a[0], &lb[0]);
	for (int i = 1; i < L; i++) orc_copy_block_sparse_tensor(&lb[0], &lb[i]);
	struct block_sparse_tensor* h2 = xmalloc((size_t)(L - 1) * sizeof(struct block_sparse_tensor));
	for (int i = 0; i < L - 1; i++) orc_mpo_merge_tensor_pair(&H->a[i], &H->a[i + 1], &h2[i]);
	int status = 0;
	for (int n = 0; n < num_sweeps && status == 0; n++) {
		double en = 0;
		for (int i = 0; i < L - 2 && status == 0; i++) {
			struct block_sparse_tensor ac, ao;
			orc_mps_merge_tensor_pair(&psi->a[i], &psi->a[i + 1], &ac);
			orc_delete_block_sparse_tensor(&psi->a[i]);
			orc_delete_block_sparse_tensor(&psi->a[i + 1]);
			int ret = minimize_local_energy(&h2[i], &lb[i], &rb[i + 1], &ac, maxiter_lanczos, &en, &ao);
			orc_delete_block_sparse_tensor(&ac);
			if (ret < 0) { status = ret; orc_delete_block_sparse_tensor(&ao); break; }
			const ct_long dp[2] = { psi->d, psi->d };
			const qnumber* qp[2] = { psi->qsite, psi->qsite };
			struct trunc_info info;
			ret = orc_mps_split_tensor_svd(&ao, dp, qp, tol_split, max_vdim, false, SVD_DISTR_RIGHT, &psi->a[i], &psi->a[i + 1], &info);
			orc_delete_block_sparse_tensor(&ao);
			if (ret < 0) { status = ret; break; }
			orc_delete_block_sparse_tensor(&lb[i + 1]);
			orc_contraction_operator_step_left(&psi->a[i], &psi->a[i], &H->a[i], &lb[i], &lb[i + 1]);
		}
		for (int i = L - 2; i >= 0 && status == 0; i--) {
			struct block_sparse_tensor ac, ao;
			orc_mps_merge_tensor_pair(&psi->a[i], &psi->a[i + 1], &ac);
			orc_delete_block_sparse_tensor(&psi->a[i]);
			orc_delete_block_sparse_tensor(&psi->a[i + 1]);
			int ret = minimize_local_energy(&h2[i], &lb[i], &rb[i + 1], &ac, maxiter_lanczos, &en, &ao);
			orc_delete_block_sparse_tensor(&ac);
			if (ret < 0) { status = ret; orc_delete_block_sparse_tensor(&ao); break; }
			const ct_long dp[2] = { psi->d, psi->d };
			const qnumber* qp[2] = { psi->qsite, psi->qsite };
			struct trunc_info info;
			ret = orc_mps_split_tensor_svd(&ao, dp, qp, tol_split, max_vdim, false, SVD_DISTR_LEFT, &psi->a[i], &psi->a[i + 1], &info);
			orc_delete_block_sparse_tensor(&ao);
			if (ret < 0) { status = ret; break; }
			entropy[i] = info.entropy;
			orc_delete_block_sparse_tensor(&rb[i]);
			orc_contraction_operator_step_right(&psi->a[i + 1], &psi->a[i + 1], &H->a[i + 1], &rb[i + 1], &rb[i]);
		}
		if (status != 0) break;
		{
			struct block_sparse_tensor tt;
			const ct_long dim[3] = { psi->a[0].dim_logical[0], 1, psi->a[0].dim_logical[0] };
			const enum tensor_axis_direction ad[3] = { TENSOR_AXIS_OUT, TENSOR_AXIS_OUT, TENSOR_AXIS_IN };
			const qnumber qz[1] = { 0 };
			const qnumber* qn[3] = { psi->a[0].qnums_logical[0], qz, psi->a[0].qnums_logical[0] };
			orc_allocate_block_sparse_tensor(3, dim, ad, qn, &tt);
			orc_mps_local_orthonormalize_rq(&psi->a[0], &tt);
			orc_delete_block_sparse_tensor(&tt);
		}
		en_sweeps[n] = en;
	}
	for (int i = 0; i < L - 1; i++) orc_delete_block_sparse_tensor(&h2[i]);
	free(h2);
	for (int i = 0; i < L; i++) { orc_delete_block_sparse_tensor(&rb[i]); orc_delete_block_sparse_tensor(&lb[i]); }
	free(rb); free(lb);
	return status;
}

ORC_EXPORT int orc_bst_get_meta(const struct block_sparse_tensor* t, ct_long* dim_logical, ct_long* dim_blocks, int* axis_dir)
{
	for (int i = 0; i < t->ndim; i++) { dim_logical[i] = t->dim_logical[i]; dim_blocks[i] = t->dim_blocks[i]; axis_dir[i] = t->axis_dir[i]; }
	return t->ndim;
}
