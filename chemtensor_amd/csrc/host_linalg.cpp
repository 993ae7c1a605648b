// Host FP64 linear algebra of the DMRG sweep: the steps whose stability needs FP64 and whose
// sizes are small (tridiagonal Lanczos eigenproblem, truncating block SVD, QR/RQ of the
// orthonormalisation). Replaces the reference's LAPACK calls (dsteqr at src/util/krylov.c:230,
// dgesvd at src/tensor/dense_tensor.c:3644, dgeqrf/dorgqr and dgerqf/dorgrq at
// dense_tensor.c:2391-2879). QR and RQ use LAPACK's Householder sign convention (dlarfg), so
// the gauge of the factors matches the reference; the SVD is Hestenes one-sided Jacobi.

#include <immintrin.h>

#include <algorithm>
#include <cmath>
#include <cfloat>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "ct_internal.h"

namespace ctg {

#define CM(a, lda, i, j) ((a)[(size_t)(i) + (size_t)(lda) * (size_t)(j)])

static double nrm2s(ct_long n, const double* x, ct_long inc)
{
	double scale = 0, ssq = 1;
	for (ct_long i = 0; i < n; i++) {
		const double v = x[i * inc];
		if (v == 0) continue;
		const double a = std::fabs(v);
		if (scale < a) { ssq = 1 + ssq * (scale / a) * (scale / a); scale = a; }
		else ssq += (a / scale) * (a / scale);
	}
	return scale * std::sqrt(ssq);
}

// sqrt(a^2 + b^2) without overflow or underflow: the larger magnitude scaled out (about one ulp;
// std::hypot rounds correctly but costs several times as much, once per QL rotation)
static inline double hypot_fast(double a, double b)
{
	a = std::fabs(a);
	b = std::fabs(b);
	// NaN and infinite input propagate (NaN stays NaN, an infinity gives infinity) instead of being
	// dropped by max / min or turned into inf / inf
	if (!(a <= DBL_MAX) || !(b <= DBL_MAX)) return a + b;
	const double m = std::max(a, b), n = std::min(a, b);
	if (m == 0) return 0;
	const double r = n / m;
	return m * std::sqrt(1.0 + r * r);
}

// Implicit-shift QL iteration on the symmetric tridiagonal (d, offd) (the algorithm of EISPACK
// tql2), applying the rotations to the columns of z (n x n, column-major) as it goes: with z = I on
// entry the columns become the eigenvectors of the tridiagonal, with z = Q those of Q T Q^T.
// Eigenvalues (and columns) are returned in ascending order. An off-diagonal element deflates when
// it is below eps times its two diagonal neighbours or, with anorm > 0, below eps * anorm (absolute
// accuracy relative to the norm: clusters of eigenvalues at the rounding level then converge).
static int ql_vectors(int n, double* d, const double* offd, double* z, double anorm = 0)
{
	std::vector<double> e((size_t)n + 1, 0.0);
	for (int i = 0; i < n - 1; i++) e[i] = offd[i];
	const double eps = 2.2204460492503131e-16;
	for (int l = 0; l < n; l++) {
		for (int iter = 0;; iter++) {
			int m = l;
			for (; m < n - 1; m++) if (std::fabs(e[m]) <= eps * (std::fabs(d[m]) + std::fabs(d[m + 1])) || std::fabs(e[m]) <= eps * anorm) break;
			if (m == l) break;
			if (iter >= 60) return -1;
			double g = (d[l + 1] - d[l]) / (2 * e[l]);
			double r = hypot_fast(g, 1.0);
			g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
			double s = 1, c = 1, p = 0;
			bool deflated = false;
			for (int i = m - 1; i >= l; i--) {
				double f = s * e[i];
				const double b = c * e[i];
				r = hypot_fast(f, g);
				e[i + 1] = r;
				if (r == 0) { d[i + 1] -= p; e[m] = 0; deflated = true; break; }
				s = f / r;
				c = g / r;
				g = d[i + 1] - p;
				r = (d[i] - g) * s + 2 * c * b;
				p = s * r;
				d[i + 1] = g + p;
				g = c * r - b;
				double* zi = z + (size_t)n * i;
				double* zj = z + (size_t)n * (i + 1);
				for (int k = 0; k < n; k++) {
					const double x = zi[k], y = zj[k];
					zj[k] = s * x + c * y;
					zi[k] = c * x - s * y;
				}
			}
			if (deflated) continue;
			d[l] -= p;
			e[l] = g;
			e[m] = 0;
		}
	}
	std::vector<int> ord(n);
	std::iota(ord.begin(), ord.end(), 0);
	std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return d[a] < d[b]; });
	std::vector<double> d2(n), z2((size_t)n * n);
	for (int i = 0; i < n; i++) {
		d2[i] = d[ord[i]];
		for (int k = 0; k < n; k++) z2[k + (size_t)n * i] = z[k + (size_t)n * ord[i]];
	}
	std::copy(d2.begin(), d2.end(), d);
	std::copy(z2.begin(), z2.end(), z);
	return 0;
}

int host_tridiag_eigh(int n, double* d, double* offd, double* z)
{
	for (int i = 0; i < n * n; i++) z[i] = 0;
	for (int i = 0; i < n; i++) z[i + n * i] = 1;
	return ql_vectors(n, d, offd, z);
}

// Householder reflector (LAPACK dlarfg convention): H (alpha; x) = (beta; 0)
static void larfg(ct_long n, double& alpha, double* x, ct_long inc, double& tau)
{
	if (n <= 1) { tau = 0; return; }
	const double xn = nrm2s(n - 1, x, inc);
	if (xn == 0) { tau = 0; return; }
	const double beta = -std::copysign(std::hypot(alpha, xn), alpha);
	tau = (beta - alpha) / beta;
	const double sc = 1.0 / (alpha - beta);
	for (ct_long i = 0; i < n - 1; i++) x[i * inc] *= sc;
	alpha = beta;
}

// horizontal sum of the four lanes, (l0 + l1) + (l2 + l3)
static inline double hsum4(__m256d v)
{
	alignas(32) double l[4];
	_mm256_store_pd(l, v);
	return (l[0] + l[1]) + (l[2] + l[3]);
}

// c (m x n, column-major, leading dimension ldc) <- (I - tau v v^T) c. Four columns at a time:
// their dot products with v in four-lane partial sums, then the rank-1 update.
static void apply_left(ct_long m, ct_long n, const double* v, ct_long incv, double tau, double* c, ct_long ldc)
{
	if (tau == 0) return;
	if (incv != 1) {
		for (ct_long j = 0; j < n; j++) {
			double w = 0;
			for (ct_long i = 0; i < m; i++) w += v[i * incv] * CM(c, ldc, i, j);
			w *= tau;
			for (ct_long i = 0; i < m; i++) CM(c, ldc, i, j) -= v[i * incv] * w;
		}
		return;
	}
	const ct_long m4 = m & ~(ct_long)3;
	ct_long j = 0;
	for (; j + 4 <= n; j += 4) {
		double* c0 = &CM(c, ldc, 0, j);
		double* c1 = c0 + ldc;
		double* c2 = c1 + ldc;
		double* c3 = c2 + ldc;
		__m256d a0 = _mm256_setzero_pd(), a1 = a0, a2 = a0, a3 = a0;
		for (ct_long i = 0; i < m4; i += 4) {
			const __m256d vv = _mm256_loadu_pd(v + i);
			a0 = _mm256_fmadd_pd(vv, _mm256_loadu_pd(c0 + i), a0);
			a1 = _mm256_fmadd_pd(vv, _mm256_loadu_pd(c1 + i), a1);
			a2 = _mm256_fmadd_pd(vv, _mm256_loadu_pd(c2 + i), a2);
			a3 = _mm256_fmadd_pd(vv, _mm256_loadu_pd(c3 + i), a3);
		}
		double w0 = hsum4(a0), w1 = hsum4(a1), w2 = hsum4(a2), w3 = hsum4(a3);
		for (ct_long i = m4; i < m; i++) { w0 += v[i] * c0[i]; w1 += v[i] * c1[i]; w2 += v[i] * c2[i]; w3 += v[i] * c3[i]; }
		w0 *= tau; w1 *= tau; w2 *= tau; w3 *= tau;
		const __m256d b0 = _mm256_set1_pd(w0), b1 = _mm256_set1_pd(w1), b2 = _mm256_set1_pd(w2), b3 = _mm256_set1_pd(w3);
		for (ct_long i = 0; i < m4; i += 4) {
			const __m256d vv = _mm256_loadu_pd(v + i);
			_mm256_storeu_pd(c0 + i, _mm256_fnmadd_pd(vv, b0, _mm256_loadu_pd(c0 + i)));
			_mm256_storeu_pd(c1 + i, _mm256_fnmadd_pd(vv, b1, _mm256_loadu_pd(c1 + i)));
			_mm256_storeu_pd(c2 + i, _mm256_fnmadd_pd(vv, b2, _mm256_loadu_pd(c2 + i)));
			_mm256_storeu_pd(c3 + i, _mm256_fnmadd_pd(vv, b3, _mm256_loadu_pd(c3 + i)));
		}
		for (ct_long i = m4; i < m; i++) { c0[i] -= v[i] * w0; c1[i] -= v[i] * w1; c2[i] -= v[i] * w2; c3[i] -= v[i] * w3; }
	}
	for (; j < n; j++) {
		double* cj = &CM(c, ldc, 0, j);
		__m256d a0 = _mm256_setzero_pd();
		for (ct_long i = 0; i < m4; i += 4) a0 = _mm256_fmadd_pd(_mm256_loadu_pd(v + i), _mm256_loadu_pd(cj + i), a0);
		double w = hsum4(a0);
		for (ct_long i = m4; i < m; i++) w += v[i] * cj[i];
		w *= tau;
		for (ct_long i = 0; i < m; i++) cj[i] -= v[i] * w;
	}
}

// c (m x n, column-major) <- c (I - tau v v^T), v with stride incv: w = c v accumulated column by
// column (each w_i sums j = 0 .. n-1 in order), then c -= tau w v^T column by column.
static void apply_right(ct_long m, ct_long n, const double* v, ct_long incv, double tau, double* c, ct_long ldc)
{
	if (tau == 0 || m == 0) return;
	std::vector<double> w((size_t)m, 0.0);
	for (ct_long j = 0; j < n; j++) {
		const double vj = v[j * incv];
		const double* cj = &CM(c, ldc, 0, j);
		for (ct_long i = 0; i < m; i++) w[i] += cj[i] * vj;
	}
	for (ct_long i = 0; i < m; i++) w[i] *= tau;
	for (ct_long j = 0; j < n; j++) {
		const double vj = v[j * incv];
		double* cj = &CM(c, ldc, 0, j);
		for (ct_long i = 0; i < m; i++) cj[i] -= w[i] * vj;
	}
}

int host_qr(ct_long m, ct_long n, const double* a, double* q, double* r)
{
	const ct_long k = std::min(m, n);
	std::vector<double> t((size_t)(m * n)), tau((size_t)k);
	for (ct_long i = 0; i < m; i++) for (ct_long j = 0; j < n; j++) CM(t.data(), m, i, j) = a[i * n + j];
	double* T = t.data();
	for (ct_long i = 0; i < k; i++) {
		larfg(m - i, CM(T, m, i, i), &CM(T, m, std::min(i + 1, m - 1), i), 1, tau[i]);
		if (i < n - 1) {
			const double aii = CM(T, m, i, i);
			CM(T, m, i, i) = 1;
			apply_left(m - i, n - i - 1, &CM(T, m, i, i), 1, tau[i], &CM(T, m, i, i + 1), m);
			CM(T, m, i, i) = aii;
		}
	}
	for (ct_long i = 0; i < k; i++) for (ct_long j = 0; j < n; j++) r[i * n + j] = (j < i ? 0 : CM(T, m, i, j));
	for (ct_long i = k - 1; i >= 0; i--) {
		if (i < k - 1) {
			CM(T, m, i, i) = 1;
			apply_left(m - i, k - i - 1, &CM(T, m, i, i), 1, tau[i], &CM(T, m, i, i + 1), m);
		}
		for (ct_long l = i + 1; l < m; l++) CM(T, m, l, i) *= -tau[i];
		CM(T, m, i, i) = 1 - tau[i];
		for (ct_long l = 0; l < i; l++) CM(T, m, l, i) = 0;
	}
	for (ct_long i = 0; i < m; i++) for (ct_long j = 0; j < k; j++) q[i * k + j] = CM(T, m, i, j);
	return 0;
}

int host_rq(ct_long m, ct_long n, const double* a, double* r, double* q)
{
	const ct_long k = std::min(m, n);
	std::vector<double> t((size_t)(m * n)), tau((size_t)k);
	for (ct_long i = 0; i < m; i++) for (ct_long j = 0; j < n; j++) CM(t.data(), m, i, j) = a[i * n + j];
	double* T = t.data();
	for (ct_long i = k - 1; i >= 0; i--) {
		const ct_long row = m - k + i, col = n - k + i;
		larfg(col + 1, CM(T, m, row, col), &CM(T, m, row, 0), m, tau[i]);
		const double aii = CM(T, m, row, col);
		CM(T, m, row, col) = 1;
		apply_right(row, col + 1, &CM(T, m, row, 0), m, tau[i], T, m);
		CM(T, m, row, col) = aii;
	}
	const ct_long rshift = std::max<ct_long>(m - k, 0), cshift = std::max<ct_long>(n - m, 0);
	for (ct_long i = 0; i < m; i++) {
		const ct_long nz = std::max<ct_long>(i - rshift, 0);
		for (ct_long j = 0; j < k; j++) r[i * k + j] = (j < nz ? 0 : CM(T, m, i, cshift + j));
	}
	std::vector<double> qt((size_t)(k * n));
	double* Q = qt.data();
	for (ct_long i = 0; i < k; i++) for (ct_long j = 0; j < n; j++) CM(Q, k, i, j) = CM(T, m, rshift + i, j);
	for (ct_long i = 0; i < k; i++) {
		const ct_long col = n - k + i;
		CM(Q, k, i, col) = 1;
		apply_right(i, col + 1, &CM(Q, k, i, 0), k, tau[i], Q, k);
		for (ct_long j = 0; j < col; j++) CM(Q, k, i, j) *= -tau[i];
		CM(Q, k, i, col) = 1 - tau[i];
		for (ct_long j = col + 1; j < n; j++) CM(Q, k, i, j) = 0;
	}
	for (ct_long i = 0; i < k; i++) for (ct_long j = 0; j < n; j++) q[i * n + j] = CM(Q, k, i, j);
	return 0;
}

// One-sided (Hestenes) Jacobi on the columns of g (M x N, column-major), accumulating the rotations
// in v (N x N). The two column norms and the inner product of a pair are formed exactly in one
// pass (four independent partial sums each, vectorisable, fixed order): norms carried over from
// earlier rotations would cancel to tiny or negative values for rank-deficient columns.
// A column whose norm falls below eps * ||A||_F is numerically zero and is set to zero: a column in
// the span of the others would otherwise be projected out again and again, each rotation leaving a
// residual of rounding noise eps times smaller that is not orthogonal to the rest, until it
// underflows (the limit is within the backward error of the decomposition; zero columns are
// completed to an orthonormal set by the caller).
static int jacobi(ct_long M, ct_long N, double* g, double* v)
{
	constexpr double eps = 2.2204460492503131e-16;
	std::fill(v, v + (size_t)(N * N), 0.0);
	for (ct_long i = 0; i < N; i++) CM(v, N, i, i) = 1;
	const double tol = eps * std::max(std::sqrt((double)M), 4.0);
	double fro2 = 0;
	for (ct_long i = 0; i < M * N; i++) fro2 += g[i] * g[i];
	const double small2 = fro2 * eps * eps;
	double worst = 0;   // largest |gamma| / sqrt(alpha beta) of the last sweep
	for (int sweep = 0; sweep < 80; sweep++) {
		bool rotated = false;
		worst = 0;
		for (ct_long p = 0; p < N - 1; p++) {
			for (ct_long q = p + 1; q < N; q++) {
				double* gp = &CM(g, M, 0, p);
				double* gq = &CM(g, M, 0, q);
				double a4[4] = { 0, 0, 0, 0 }, b4[4] = { 0, 0, 0, 0 }, g4[4] = { 0, 0, 0, 0 };
				ct_long i = 0;
				for (; i + 4 <= M; i += 4) {
					for (int l = 0; l < 4; l++) {
						const double x = gp[i + l], y = gq[i + l];
						a4[l] += x * x;
						b4[l] += y * y;
						g4[l] += x * y;
					}
				}
				for (; i < M; i++) { a4[0] += gp[i] * gp[i]; b4[0] += gq[i] * gq[i]; g4[0] += gp[i] * gq[i]; }
				const double al = (a4[0] + a4[1]) + (a4[2] + a4[3]);
				const double be = (b4[0] + b4[1]) + (b4[2] + b4[3]);
				const double ga = (g4[0] + g4[1]) + (g4[2] + g4[3]);
				if (al <= small2 || be <= small2) {
					if (al != 0 && al <= small2) std::fill(gp, gp + M, 0.0);
					if (be != 0 && be <= small2) std::fill(gq, gq + M, 0.0);
					continue;
				}
				const double rel = std::fabs(ga) / (std::sqrt(al) * std::sqrt(be));
				if (rel <= tol) continue;
				worst = std::max(worst, rel);
				rotated = true;
				const double zeta = (be - al) / (2 * ga);
				const double t = std::copysign(1.0, zeta) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
				const double c = 1 / std::sqrt(1 + t * t), s = c * t;
				for (ct_long k = 0; k < M; k++) { const double x = gp[k], y = gq[k]; gp[k] = c * x - s * y; gq[k] = s * x + c * y; }
				double* vp = &CM(v, N, 0, p);
				double* vq = &CM(v, N, 0, q);
				for (ct_long k = 0; k < N; k++) { const double x = vp[k], y = vq[k]; vp[k] = c * x - s * y; vq[k] = s * x + c * y; }
			}
		}
		if (!rotated) return 0;
	}
	fprintf(stderr, "chemtensor_gpu: Jacobi SVD (%lld x %lld) did not converge (off-orthogonality %.3e)\n", (long long)M, (long long)N, worst);
	return -1;
}

int host_svd(ct_long m, ct_long n, const double* a, double* u, double* s, double* vh)
{
	const ct_long k = std::min(m, n);
	const bool tall = (m >= n);
	const ct_long M = tall ? m : n, N = tall ? n : m;
	std::vector<double> g((size_t)(M * N)), v((size_t)(N * N));
	for (ct_long i = 0; i < m; i++) for (ct_long j = 0; j < n; j++) {
		if (tall) CM(g.data(), M, i, j) = a[i * n + j];
		else CM(g.data(), M, j, i) = a[i * n + j];
	}
	if (jacobi(M, N, g.data(), v.data()) != 0) return -1;
	std::vector<double> sv((size_t)N);
	for (ct_long j = 0; j < N; j++) sv[j] = nrm2s(M, &CM(g.data(), M, 0, j), 1);
	std::vector<ct_long> ord((size_t)N);
	std::iota(ord.begin(), ord.end(), 0);
	std::stable_sort(ord.begin(), ord.end(), [&](ct_long x, ct_long y) { return sv[x] > sv[y]; });
	std::vector<double> lu((size_t)(M * N), 0.0);
	std::vector<char> valid((size_t)N, 0);
	for (ct_long jj = 0; jj < N; jj++) {
		const ct_long j = ord[jj];
		s[jj] = sv[j];
		valid[jj] = sv[j] > 0;
		if (valid[jj]) for (ct_long i = 0; i < M; i++) CM(lu.data(), M, i, jj) = CM(g.data(), M, i, j) / sv[j];
	}
	// complete columns of zero singular values to an orthonormal set: each takes the unit vector e_i
	// with the largest component outside the current span (1 - sum of squares of row i over the set
	// columns, at least (M - set columns) / M on average), orthogonalised twice and normalised
	std::vector<double> rowsq((size_t)M, 0.0);
	for (ct_long l = 0; l < N; l++) if (valid[l]) {
		const double* y = &CM(lu.data(), M, 0, l);
		for (ct_long i = 0; i < M; i++) rowsq[i] += y[i] * y[i];
	}
	for (ct_long j = 0; j < N; j++) {
		if (valid[j]) continue;
		const ct_long e = std::min_element(rowsq.begin(), rowsq.end()) - rowsq.begin();
		double* x = &CM(lu.data(), M, 0, j);
		for (ct_long i = 0; i < M; i++) x[i] = (i == e ? 1 : 0);
		for (int pass = 0; pass < 2; pass++)
			for (ct_long l = 0; l < N; l++) {
				if (l == j || !valid[l]) continue;
				const double* y = &CM(lu.data(), M, 0, l);
				double d = 0;
				for (ct_long i = 0; i < M; i++) d += x[i] * y[i];
				for (ct_long i = 0; i < M; i++) x[i] -= d * y[i];
			}
		const double nx = nrm2s(M, x, 1);
		for (ct_long i = 0; i < M; i++) { x[i] /= nx; rowsq[i] += x[i] * x[i]; }
		valid[j] = 1;
	}
	if (tall) {
		for (ct_long i = 0; i < m; i++) for (ct_long j = 0; j < k; j++) u[i * k + j] = CM(lu.data(), M, i, j);
		for (ct_long jj = 0; jj < k; jj++) for (ct_long i = 0; i < n; i++) vh[jj * n + i] = CM(v.data(), N, i, ord[jj]);
	}
	else {
		for (ct_long i = 0; i < m; i++) for (ct_long jj = 0; jj < k; jj++) u[i * k + jj] = CM(v.data(), N, i, ord[jj]);
		for (ct_long jj = 0; jj < k; jj++) for (ct_long i = 0; i < n; i++) vh[jj * n + i] = CM(lu.data(), M, i, jj);
	}
	return 0;
}

// Gram matrix of the rows of b (k x n, row-major): c (k x k, column-major, lower triangle and
// diagonal) = b b^T. Tiles of 2 x 4 rows, four-lane partial sums per entry (fixed order).
static void gram_rows(ct_long k, ct_long n, const double* b, double* c)
{
	const ct_long n4 = n & ~(ct_long)3;
	for (ct_long i0 = 0; i0 < k; i0 += 2) {
		const ct_long ni = std::min<ct_long>(2, k - i0);
		for (ct_long j0 = 0; j0 < i0 + ni; j0 += 4) {
			const ct_long nj = std::min<ct_long>(4, k - j0);
			const double* bi[2];
			const double* bj[4];
			for (int a = 0; a < 2; a++) bi[a] = b + (size_t)(i0 + std::min<ct_long>(a, ni - 1)) * n;
			for (int a = 0; a < 4; a++) bj[a] = b + (size_t)(j0 + std::min<ct_long>(a, nj - 1)) * n;
			__m256d acc[2][4];
			for (int a = 0; a < 2; a++) for (int q = 0; q < 4; q++) acc[a][q] = _mm256_setzero_pd();
			for (ct_long t = 0; t < n4; t += 4) {
				const __m256d x0 = _mm256_loadu_pd(bi[0] + t), x1 = _mm256_loadu_pd(bi[1] + t);
				for (int q = 0; q < 4; q++) {
					const __m256d y = _mm256_loadu_pd(bj[q] + t);
					acc[0][q] = _mm256_fmadd_pd(x0, y, acc[0][q]);
					acc[1][q] = _mm256_fmadd_pd(x1, y, acc[1][q]);
				}
			}
			for (int a = 0; a < ni; a++)
				for (int q = 0; q < nj; q++) {
					if (j0 + q > i0 + a) continue;
					double x = hsum4(acc[a][q]);
					for (ct_long t = n4; t < n; t++) x += bi[a][t] * bj[q][t];
					CM(c, k, i0 + a, j0 + q) = x;
				}
		}
	}
}

// right (k x n, row-major) = z^T b with z (k x k, column-major) and b (k x n, row-major): tiles of
// 6 rows x 8 columns of right accumulated in registers over the k rows of b (each entry sums
// i = 0 .. k-1 in order)
static void project_rows(ct_long k, ct_long n, const double* z, const double* b, double* right)
{
	const ct_long n8 = n & ~(ct_long)7;
	for (ct_long l0 = 0; l0 < k; l0 += 6) {
		const ct_long nl = std::min<ct_long>(6, k - l0);
		const double* zc[6];
		for (int l = 0; l < 6; l++) zc[l] = z + (size_t)k * (size_t)(l0 + std::min<ct_long>(l, nl - 1));
		for (ct_long t0 = 0; t0 < n8; t0 += 8) {
			__m256d acc[6][2];
			for (int l = 0; l < 6; l++) acc[l][0] = acc[l][1] = _mm256_setzero_pd();
			for (ct_long i = 0; i < k; i++) {
				const double* bi = b + (size_t)i * n + t0;
				const __m256d b0 = _mm256_loadu_pd(bi), b1 = _mm256_loadu_pd(bi + 4);
				for (int l = 0; l < 6; l++) {
					const __m256d zz = _mm256_broadcast_sd(zc[l] + i);
					acc[l][0] = _mm256_fmadd_pd(zz, b0, acc[l][0]);
					acc[l][1] = _mm256_fmadd_pd(zz, b1, acc[l][1]);
				}
			}
			for (int l = 0; l < nl; l++) {
				_mm256_storeu_pd(right + (size_t)(l0 + l) * n + t0, acc[l][0]);
				_mm256_storeu_pd(right + (size_t)(l0 + l) * n + t0 + 4, acc[l][1]);
			}
		}
		for (ct_long l = 0; l < nl; l++) {
			double* r = right + (size_t)(l0 + l) * n;
			for (ct_long t = n8; t < n; t++) {
				double x = 0;
				for (ct_long i = 0; i < k; i++) x += zc[l][i] * b[(size_t)i * n + t];
				r[t] = x;
			}
		}
	}
}

// Householder reduction of the symmetric c (k x k, column-major, lower triangle used) to
// tridiagonal form T = Q^T c Q (d: diagonal, e: subdiagonal); q receives Q explicitly.
static void tridiagonalize(ct_long k, double* c, double* d, double* e, double* q)
{
	std::vector<double> tau((size_t)std::max<ct_long>(k, 1), 0.0), p((size_t)k), w((size_t)k);
	for (ct_long j = 0; j + 2 < k; j++) {
		const ct_long r = k - j - 1;          // length of the reflector
		double* v = &CM(c, k, j + 1, j);
		larfg(r, v[0], v + 1, 1, tau[j]);
		e[j] = v[0];
		d[j] = CM(c, k, j, j);
		const double t = tau[j];
		if (t == 0) continue;
		v[0] = 1;
		// p = t * A22 v with A22 = c[j+1:, j+1:] (lower triangle)
		double* A = &CM(c, k, j + 1, j + 1);
		std::fill(p.begin(), p.begin() + r, 0.0);
		for (ct_long cj = 0; cj < r; cj++) {
			const double* col = A + (size_t)cj * k;
			const double vj = v[cj];
			double s4[4] = { 0, 0, 0, 0 };
			ct_long i = cj + 1;
			for (; i + 4 <= r; i += 4)
				for (int l = 0; l < 4; l++) { p[i + l] += col[i + l] * vj; s4[l] += col[i + l] * v[i + l]; }
			double sum = (s4[0] + s4[1]) + (s4[2] + s4[3]);
			for (; i < r; i++) { p[i] += col[i] * vj; sum += col[i] * v[i]; }
			p[cj] += col[cj] * vj + sum;
		}
		double pv = 0;
		for (ct_long i = 0; i < r; i++) { p[i] *= t; pv += p[i] * v[i]; }
		const double alpha = -0.5 * t * pv;
		for (ct_long i = 0; i < r; i++) w[i] = p[i] + alpha * v[i];
		// A22 -= v w^T + w v^T (lower triangle)
		for (ct_long cj = 0; cj < r; cj++) {
			double* col = A + (size_t)cj * k;
			const double vj = v[cj], wj = w[cj];
			for (ct_long i = cj; i < r; i++) col[i] -= v[i] * wj + w[i] * vj;
		}
	}
	if (k >= 2) { d[k - 2] = CM(c, k, k - 2, k - 2); e[k - 2] = CM(c, k, k - 1, k - 2); }
	if (k >= 1) d[k - 1] = CM(c, k, k - 1, k - 1);
	// Q = H_0 H_1 ... H_{k-3}, accumulated backwards; H_j acts on rows and columns j+1 .. k-1
	std::fill(q, q + (size_t)(k * k), 0.0);
	for (ct_long i = 0; i < k; i++) CM(q, k, i, i) = 1;
	for (ct_long j = k - 3; j >= 0; j--) {
		if (tau[j] == 0) continue;
		double* v = &CM(c, k, j + 1, j);
		v[0] = 1;
		apply_left(k - j - 1, k - j - 1, v, 1, tau[j], &CM(q, k, j + 1, j + 1), k);
	}
}

// Factorisation of b (k x n, row-major, k <= n) as b = z * right with z (k x k, column-major)
// orthogonal: the eigenvectors of b b^T in descending order of their eigenvalue, right = z^T b.
static int gram_factor(ct_long k, ct_long n, const double* b, std::vector<double>& z, std::vector<double>& right)
{
	std::vector<double> c((size_t)(k * k), 0.0), d((size_t)k), e((size_t)std::max<ct_long>(k, 1), 0.0), q((size_t)(k * k));
	gram_rows(k, n, b, c.data());
	tridiagonalize(k, c.data(), d.data(), e.data(), q.data());
	double anorm = 0;
	for (ct_long i = 0; i < k; i++) anorm = std::max(anorm, std::fabs(d[i]) + (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < k ? std::fabs(e[i]) : 0.0));
	if (ql_vectors((int)k, d.data(), e.data(), q.data(), anorm) != 0) return -1;
	// descending order
	z.assign((size_t)(k * k), 0.0);
	for (ct_long l = 0; l < k; l++) std::copy(&CM(q.data(), k, 0, k - 1 - l), &CM(q.data(), k, 0, k - 1 - l) + k, &CM(z.data(), k, 0, l));
	right.assign((size_t)(k * n), 0.0);
	project_rows(k, n, z.data(), b, right.data());
	return 0;
}

int host_svd_factor(ct_long m, ct_long n, const double* a, int iso, double* left, double* s, double* right)
{
	const ct_long k = std::min(m, n);
	if (k == 0) return 0;
	if (iso == 1) {
		// a^T = left' right' with left' isometric: a = right'^T left'^T
		std::vector<double> at((size_t)(m * n)), l2((size_t)(n * k)), r2((size_t)(k * m));
		for (ct_long i = 0; i < m; i++) for (ct_long j = 0; j < n; j++) at[(size_t)j * m + i] = a[(size_t)i * n + j];
		const int ret = host_svd_factor(n, m, at.data(), 0, l2.data(), s, r2.data());
		if (ret != 0) return ret;
		for (ct_long i = 0; i < m; i++) for (ct_long l = 0; l < k; l++) left[(size_t)i * k + l] = r2[(size_t)l * m + i];
		for (ct_long l = 0; l < k; l++) for (ct_long j = 0; j < n; j++) right[(size_t)l * n + j] = l2[(size_t)j * k + l];
		return 0;
	}
	std::vector<double> z, rt, qm, rm;
	const double* b = a;
	if (m > n) {
		// a = Q R: factor R (n x n) and left = Q z
		qm.resize((size_t)(m * n));
		rm.resize((size_t)(n * n));
		if (host_qr(m, n, a, qm.data(), rm.data()) != 0) return -1;
		b = rm.data();
	}
	if (gram_factor(k, n, b, z, rt) != 0) return -1;
	// singular values as the norms of the rows of right (sorted descending; ties keep their order)
	std::vector<double> sv((size_t)k);
	for (ct_long l = 0; l < k; l++) sv[l] = nrm2s(n, rt.data() + (size_t)l * n, 1);
	std::vector<ct_long> ord((size_t)k);
	std::iota(ord.begin(), ord.end(), 0);
	std::stable_sort(ord.begin(), ord.end(), [&](ct_long x, ct_long y) { return sv[x] > sv[y]; });
	for (ct_long l = 0; l < k; l++) {
		s[l] = sv[ord[l]];
		std::copy(rt.data() + (size_t)ord[l] * n, rt.data() + (size_t)(ord[l] + 1) * n, right + (size_t)l * n);
	}
	if (m <= n) {
		for (ct_long i = 0; i < m; i++) for (ct_long l = 0; l < k; l++) left[(size_t)i * k + l] = CM(z.data(), k, i, ord[l]);
	}
	else {
		// left = Q z (m x n times n x n): rows of Q against the reordered rows of z^T
		std::vector<double> zr((size_t)(k * k));
		for (ct_long i = 0; i < k; i++) for (ct_long l = 0; l < k; l++) zr[(size_t)i * k + l] = CM(z.data(), k, i, ord[l]);
		for (ct_long r = 0; r < m; r++) {
			double* out = left + (size_t)r * k;
			std::fill(out, out + k, 0.0);
			const double* qr = qm.data() + (size_t)r * n;
			for (ct_long i = 0; i < k; i++) {
				const double x = qr[i];
				const double* zi = zr.data() + (size_t)i * k;
				for (ct_long l = 0; l < k; l++) out[l] += x * zi[l];
			}
		}
	}
	return 0;
}

}  // namespace ctg
