"""Host-side parts of the product library (no GPU needed): the C ABI is complete, and the
seeded inputs (random MPS, MPOs, bond quantum numbers) are bit-identical to the oracle's."""

import ctypes as C
import re

import numpy as np

import chemtensor_amd as ct
from oracle import oracle as orc


def declared_functions():
    text = open(ct.HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = []
    for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(([^;{]*)\)\s*;", text, flags=re.M):
        if m.group(1) not in ("if", "while", "for", "return", "sizeof") and not m.group(0).startswith("typedef"):
            names.append(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = ct.lib()
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layout_matches_reference_abi():
    # include/tensor/block_sparse_tensor.h:18-28 and dense_tensor.h:17-23 on LP64
    assert C.sizeof(ct.BlockSparseTensor) == 56
    assert C.sizeof(ct.DenseTensor) == 24
    assert C.sizeof(ct.MPSStruct) == 32


def _site_equal(pt, ot):
    """product site tensor (struct pointer) vs oracle BST: same metadata and entries"""
    h = orc.BST(pt, owned=False)
    d1, a1, q1 = h.meta()
    d2, a2, q2 = ot.meta()
    assert list(d1) == list(d2) and list(a1) == list(a2)
    assert all(np.array_equal(x, y) for x, y in zip(q1, q2))
    assert np.array_equal(h.serialize(), ot.serialize())


def test_random_mps_bit_identical_to_oracle():
    for (L, d, qsite, sector, chi, seed) in [(6, 4, None, (7, 1), 32, 42), (16, 2, [1, -1], 0, 32, 1), (9, 2, [1, -1], 1, 7, 5)]:
        if qsite is None:
            qsite = [orc.encode_pair(0, 0), orc.encode_pair(1, -1), orc.encode_pair(1, 1), orc.encode_pair(2, 0)]
            sector = orc.encode_pair(*sector)
        p = ct.MPS.random(L, d, qsite, sector, chi, seed)
        o = orc.MPS.random(L, d, qsite, sector, chi, seed)
        for i in range(L):
            _site_equal(p.site_ptr(i), o.site(i))


def test_mpo_constructors_bit_identical_to_oracle():
    for pm, om in [(ct.MPO.xxz(7, 0.3, 1.7, -0.2), orc.MPO.xxz(7, 0.3, 1.7, -0.2)),
                   (ct.MPO.hubbard(5, 1.1, 3.3, 0.4), orc.MPO.hubbard(5, 1.1, 3.3, 0.4))]:
        for i in range(pm.nsites):
            _site_equal(C.cast(C.pointer(pm.s.a[i]), C.c_void_p).value, om.site(i))


def test_product_mpo_matches_reference_fixture(golden):
    """The product's Hubbard MPO (through the oracle's contraction to a dense matrix) equals the
    reference's fixture test/operator/data/test_fermi_hubbard_1d_mpo.hdf5."""
    pm = ct.MPO.hubbard(4, 11. / 9, 13. / 4, 3. / 7)
    mat = np.zeros(256 * 256)
    n = orc.lib().orc_h_mpo_matrix(pm.ptr, orc.dptr(mat), mat.size)
    assert n == 256
    ref = golden("test_fermi_hubbard_1d_mpo")["D_fermi_hubbard_1d_mat"]
    assert np.max(np.abs(mat.reshape(256, 256) - ref)) < 1e-13


def _svd_stress_matrices(rng):
    """matrices that are hard for one-sided Jacobi: columns in the span of the others (random
    low rank, exact triplicates), graded columns down to 1e-15, zero and one-element cases"""
    mats = []
    for (m, n, r) in [(60, 40, 7), (40, 60, 5), (97, 95, 30), (33, 33, 1)]:
        mats.append(rng.standard_normal((m, r)) @ rng.standard_normal((r, n)))
    # triplicated rows: one-sided Jacobi without a zero-column cutoff never converges on these
    for (m, r, n) in [(40, 30, 60), (30, 29, 30), (60, 60, 62)]:
        x = np.repeat(rng.standard_normal((m, r)), 3, axis=1)[:, :n]
        mats.append(x.T.copy())
        mats.append(x)
    mats.append(rng.standard_normal((80, 75)) * 10.0 ** (-15.0 * np.arange(75) / 75))
    mats.append(np.zeros((5, 3)))
    mats.append(np.array([[2.5]]))
    mats.append(rng.standard_normal((50, 30)))
    return mats


def _check_svd(a, u, s, vh):
    k = min(a.shape)
    scale = max(1.0, np.linalg.norm(a))
    assert np.max(np.abs((u * s) @ vh - a)) < 1e-13 * scale
    assert np.max(np.abs(u.T @ u - np.eye(k))) < 1e-12
    assert np.max(np.abs(vh @ vh.T - np.eye(k))) < 1e-12
    assert np.all(np.diff(s) <= 0) and np.all(s >= 0)
    assert np.max(np.abs(s - np.linalg.svd(a, compute_uv=False))) < 1e-13 * scale


def _run_host_svd(tmp_path, mats, mode):
    """run the product's host SVD routines (compiled here from their source with a small driver)"""
    import os
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        import pytest
        pytest.skip("g++ not available")
    here = os.path.dirname(os.path.abspath(__file__))
    csrc = os.path.join(os.path.dirname(here), "chemtensor_amd", "csrc")
    exe = tmp_path / "svd_check"
    if not exe.exists():
        subprocess.run([gxx, "-O2", "-std=c++17", "-march=x86-64-v3", "-I", csrc, "-I", "/opt/rocm/include",
                        "-D__HIP_PLATFORM_AMD__", "-x", "c++", os.path.join(here, "host_svd_check.cpp"),
                        "-x", "c++", os.path.join(csrc, "host_linalg.cpp"), "-o", str(exe)], check=True)
    with open(tmp_path / "in.bin", "wb") as f:
        np.array([len(mats)], dtype=np.int64).tofile(f)
        for a in mats:
            np.array(a.shape, dtype=np.int64).tofile(f)
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
    subprocess.run([str(exe), str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), mode], check=True, timeout=120)
    raw = open(tmp_path / "out.bin", "rb").read()
    pos = 0
    out = []
    for a in mats:
        m, n = a.shape
        k = min(m, n)
        rc = int(np.frombuffer(raw, np.int32, 1, pos)[0]); pos += 4
        u = np.frombuffer(raw, np.float64, m * k, pos).reshape(m, k); pos += 8 * m * k
        s = np.frombuffer(raw, np.float64, k, pos); pos += 8 * k
        vh = np.frombuffer(raw, np.float64, k * n, pos).reshape(k, n); pos += 8 * k * n
        out.append((rc, u, s, vh))
    assert pos == len(raw)
    return out


def _svd_inputs(golden):
    mats = _svd_stress_matrices(np.random.default_rng(7))
    z = golden("test_dense_tensor_svd")
    return mats + [z[k] for k in sorted(z.files) if z[k].dtype == np.float64 and z[k].ndim == 2]


def test_host_svd_rank_deficient_and_fixtures(golden, tmp_path):
    """the product's host Jacobi SVD against numpy's LAPACK SVD on rank-deficient inputs and the
    dense SVD fixture"""
    mats = _svd_inputs(golden)
    for a, (rc, u, s, vh) in zip(mats, _run_host_svd(tmp_path, mats, "jacobi")):
        assert rc == 0, a.shape
        _check_svd(a, u, s, vh)


def test_host_split_factor_rank_deficient_and_fixtures(golden, tmp_path):
    """the split factorisation (Gram eigenvectors of the shorter side, the other factor by
    projection): left @ right reproduces a, the isometric factor is orthonormal, and the squared
    singular values agree with numpy's LAPACK SVD to eps * ||a||^2 (the eigenvalue accuracy of the
    Gram matrix)"""
    mats = _svd_inputs(golden)
    for iso in (0, 1):
        for a, (rc, left, s, right) in zip(mats, _run_host_svd(tmp_path, mats, "factor%d" % iso)):
            assert rc == 0, a.shape
            k = min(a.shape)
            scale = max(1.0, np.linalg.norm(a))
            assert np.max(np.abs(left @ right - a)) < 1e-13 * scale
            iso_f = left.T @ left if iso == 0 else right @ right.T
            assert np.max(np.abs(iso_f - np.eye(k))) < 1e-12
            # the other factor carries the singular values: its columns / rows have norms s
            other = np.linalg.norm(right, axis=1) if iso == 0 else np.linalg.norm(left, axis=0)
            assert np.max(np.abs(other - s)) < 1e-13 * scale
            assert np.all(np.diff(s) <= 0) and np.all(s >= 0)
            s_np = np.linalg.svd(a, compute_uv=False)
            assert np.max(np.abs(s ** 2 - s_np ** 2)) < 1e-13 * scale ** 2


def test_host_tridiag_eigh_nonfinite_and_numpy(tmp_path):
    """The product's tridiagonal eigensolver (host_tridiag_eigh, the dsteqr replacement of
    krylov.c:230): eigenvalues equal numpy's on finite input; a NaN or an infinity in the input is not
    rotated away as if finite (the result is an error status or carries a non-finite value)."""
    import os
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        import pytest
        pytest.skip("g++ not available")
    here = os.path.dirname(os.path.abspath(__file__))
    csrc = os.path.join(os.path.dirname(here), "chemtensor_amd", "csrc")
    exe = tmp_path / "tridiag_check"
    subprocess.run([gxx, "-O2", "-std=c++17", "-march=x86-64-v3", "-I", csrc, "-I", "/opt/rocm/include",
                    "-D__HIP_PLATFORM_AMD__", "-x", "c++", os.path.join(here, "host_svd_check.cpp"),
                    "-x", "c++", os.path.join(csrc, "host_linalg.cpp"), "-o", str(exe)], check=True)
    rng = np.random.default_rng(11)
    n = 7
    cases = []
    d, e = rng.standard_normal(n), rng.standard_normal(n)
    cases.append(("finite", d, e))
    for what, val in (("nan_off", np.nan), ("inf_off", np.inf), ("nan_diag", np.nan), ("inf_diag", -np.inf)):
        dd, ee = d.copy(), e.copy()
        if what.endswith("off"):
            ee[2] = val
        else:
            dd[3] = val
        cases.append((what, dd, ee))
    with open(tmp_path / "in.bin", "wb") as f:
        np.array([len(cases)], dtype=np.int64).tofile(f)
        for _, dd, ee in cases:
            np.array([2, n], dtype=np.int64).tofile(f)
            np.concatenate([dd, ee]).astype(np.float64).tofile(f)
    subprocess.run([str(exe), str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), "tridiag"], check=True, timeout=60)
    raw = open(tmp_path / "out.bin", "rb").read()
    pos = 0
    for what, dd, ee in cases:
        rc = int(np.frombuffer(raw, np.int32, 1, pos)[0]); pos += 4
        lam = np.frombuffer(raw, np.float64, n, pos); pos += 8 * n
        z = np.frombuffer(raw, np.float64, n * n, pos); pos += 8 * n * n
        if what == "finite":
            t = np.diag(dd) + np.diag(ee[:-1], 1) + np.diag(ee[:-1], -1)
            assert rc == 0
            assert np.max(np.abs(np.sort(lam) - np.linalg.eigvalsh(t))) < 1e-12
        else:
            assert rc != 0 or not (np.all(np.isfinite(lam)) and np.all(np.isfinite(z))), what
    assert pos == len(raw)
