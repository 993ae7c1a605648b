// Test driver for the product's host SVDs (chemtensor_amd/csrc/host_linalg.cpp): reads matrices
// from a binary file (int64 count, then per matrix int64 m, int64 n, m*n doubles row-major) and
// writes, per matrix, int32 status followed by u (m x k), s (k), vh (k x n) row-major.
// Mode (third argument): "jacobi" (host_svd, default), "factor0" / "factor1" (host_svd_factor with
// the isometry on the left / right), "tridiag" (host_tridiag_eigh of 2 x n inputs [d; e]: status,
// eigenvalues, eigenvectors).
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "ct_internal.h"

int main(int argc, char** argv)
{
	if (argc != 3 && argc != 4) return 2;
	const std::string mode = (argc == 4) ? argv[3] : "jacobi";
	FILE* in = fopen(argv[1], "rb");
	FILE* out = fopen(argv[2], "wb");
	if (!in || !out) return 2;
	int64_t count = 0;
	if (fread(&count, 8, 1, in) != 1) return 2;
	for (int64_t c = 0; c < count; c++) {
		int64_t mn[2];
		if (fread(mn, 8, 2, in) != 2) return 2;
		const int64_t m = mn[0], n = mn[1], k = m < n ? m : n;
		std::vector<double> a((size_t)(m * n)), u((size_t)(m * k)), s((size_t)k), vh((size_t)(k * n));
		if (fread(a.data(), 8, a.size(), in) != a.size()) return 2;
		if (mode == "tridiag") {
			// a = [d; e] (2 x n, e[n-1] unused): int32 status, then the n eigenvalues and the n x n
			// eigenvector matrix of host_tridiag_eigh
			if (m != 2) return 2;
			std::vector<double> d(a.begin(), a.begin() + n), e(a.begin() + n, a.end()), z((size_t)(n * n));
			const int32_t rc = ctg::host_tridiag_eigh((int)n, d.data(), e.data(), z.data());
			fwrite(&rc, 4, 1, out);
			fwrite(d.data(), 8, d.size(), out);
			fwrite(z.data(), 8, z.size(), out);
			continue;
		}
		const int32_t rc = (mode == "jacobi") ? ctg::host_svd(m, n, a.data(), u.data(), s.data(), vh.data())
			: ctg::host_svd_factor(m, n, a.data(), mode == "factor1" ? 1 : 0, u.data(), s.data(), vh.data());
		fwrite(&rc, 4, 1, out);
		fwrite(u.data(), 8, u.size(), out);
		fwrite(s.data(), 8, s.size(), out);
		fwrite(vh.data(), 8, vh.size(), out);
	}
	fclose(out);
	fclose(in);
	return 0;
}
