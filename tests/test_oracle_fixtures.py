"""Pin the CPU oracle against the reference's own golden fixtures (test/**/data/*.hdf5,
converted to tests/golden/*.npz) and, for the RNG, against the reference's own RNG sources
compiled into oracle/_ref. These mirror the reference's C tests (test/**/test_*.c)."""

import ctypes as C

import numpy as np
import pytest

from oracle import oracle as orc


def _qnums(z, n, key="A_qnums"):
    return [z[f"{key}{i}"].astype(np.int32) for i in range(n)]


def _real_imag(a):
    return np.ascontiguousarray(a.real), np.ascontiguousarray(a.imag)


# ---------------------------------------------------------------------------------------
# random numbers: bit-exact against the reference's rng.c / pcg_basic.c

def test_rng_matches_reference_build():
    ref = orc.ref_rng_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    ref.seed_rng_state.argtypes = [C.c_uint64, C.c_void_p]
    ref.rand_uint64.restype = C.c_uint64
    ref.rand_uint64.argtypes = [C.c_void_p]
    ref.randn.restype = C.c_double
    ref.randn.argtypes = [C.c_void_p]
    ref.rand_choice.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p, C.POINTER(C.c_uint64)]
    ref.rand_interval.restype = C.c_uint64
    ref.rand_interval.argtypes = [C.c_uint64, C.c_void_p]
    L = orc.lib()
    for seed in [0, 1, 42, 287, 2**40 + 17]:
        st = C.create_string_buffer(64)
        ref.seed_rng_state(seed, st)
        want_u = np.array([ref.rand_uint64(st) for _ in range(200)], dtype=np.uint64)
        got_u = np.zeros(200, dtype=np.uint64)
        L.orc_h_rng_draws(seed, 0, 200, 0, None, got_u.ctypes.data_as(orc.UP))
        assert np.array_equal(want_u, got_u)

        st = C.create_string_buffer(64)
        ref.seed_rng_state(seed, st)
        want_n = np.array([ref.randn(st) for _ in range(500)])
        got_n = np.zeros(500)
        L.orc_h_rng_draws(seed, 1, 500, 0, got_n.ctypes.data_as(orc.DP), None)
        assert np.array_equal(want_n, got_n)  # bit-exact (same libm on the same host)

        st = C.create_string_buffer(64)
        ref.seed_rng_state(seed, st)
        want_c = (C.c_uint64 * 37)()
        ref.rand_choice(1000, 37, st, want_c)
        got_c = np.zeros(37, dtype=np.uint64)
        L.orc_h_rng_draws(seed, 2, 37, 1000, None, got_c.ctypes.data_as(orc.UP))
        assert np.array_equal(np.array(list(want_c), dtype=np.uint64), got_c)

        st = C.create_string_buffer(64)
        ref.seed_rng_state(seed, st)
        want_i = np.array([ref.rand_interval(12345677, st) for _ in range(100)], dtype=np.uint64)
        got_i = np.zeros(100, dtype=np.uint64)
        L.orc_h_rng_draws(seed, 3, 100, 12345677, None, got_i.ctypes.data_as(orc.UP))
        assert np.array_equal(want_i, got_i)


def test_randn_statistics():
    """Box-Muller output is standard normal (independent of the reference build)."""
    got = np.zeros(200000)
    orc.lib().orc_h_rng_draws(7, 1, got.size, 0, got.ctypes.data_as(orc.DP), None)
    assert abs(got.mean()) < 0.01 and abs(got.std() - 1) < 0.01


def test_rand_choice_distinct():
    got = np.zeros(300, dtype=np.uint64)
    orc.lib().orc_h_rng_draws(3, 2, 300, 310, None, got.ctypes.data_as(orc.UP))
    assert len(set(got.tolist())) == 300 and got.max() < 310


# ---------------------------------------------------------------------------------------
# block-sparse bookkeeping

def test_block_sparse_transpose_fixture(golden):
    # test/tensor/test_block_sparse_tensor.c:385-450
    z = golden("test_block_sparse_tensor_transpose")
    t = orc.BST.from_dense(z["D_t"], z["A_axis_dir"], _qnums(z, 4))
    tp = orc.transpose([1, 3, 2, 0], t)
    assert np.array_equal(tp.to_dense(), z["D_t_tp"])


def test_block_sparse_dense_roundtrip(golden):
    z = golden("test_block_sparse_tensor_transpose")
    qn = _qnums(z, 4)
    ad = z["A_axis_dir"]
    t = orc.BST.from_dense(z["D_t"], ad, qn)
    d = t.to_dense()
    # entries outside the quantum-number sectors are dropped, all others kept exactly
    mask = np.zeros(d.shape, dtype=bool)
    idx = np.indices(d.shape)
    qsum = sum(ad[i] * qn[i][idx[i]] for i in range(4))
    mask = (qsum == 0)
    assert np.array_equal(d, np.where(mask, z["D_t"], 0))
    # serialisation is a bijection onto the stored entries
    v = t.serialize()
    assert v.size == mask.sum()
    t2 = orc.BST.zeros(d.shape, ad, qn)
    t2.deserialize(v)
    assert np.array_equal(t2.to_dense(), d)


def test_block_sparse_reshape_fixture(golden):
    # test/tensor/test_block_sparse_tensor.c:456-520 (complex data, real and imaginary parts separately)
    z = golden("test_block_sparse_tensor_reshape")
    qn = _qnums(z, 5)
    for part in _real_imag(z["D_t"]):
        t = orc.BST.from_dense(part, z["A_axis_dir"], qn)
        dense = t.to_dense()
        tf = orc.flatten(t, 1, orc.AXIS_OUT)
        assert np.array_equal(tf.to_dense(), dense.reshape(5, 28, 11, 3))
        dims, ad, q = t.meta()
        ts = orc.split(tf, 1, dims[1:3], ad[1:3], q[1:3])
        assert np.array_equal(ts.to_dense(), dense)
        assert np.array_equal(ts.serialize(), t.serialize())


def test_block_sparse_dot_fixture(golden):
    # test/tensor/test_block_sparse_tensor.c:1117-1230 (complex via four real products)
    z = golden("test_block_sparse_tensor_dot")
    ad = z["A_axis_dir"].astype(np.int32)
    qn = _qnums(z, 8)
    nm = 3
    ad_t = np.array([(-1 if i < nm else 1) * ad[5 - nm + i] for i in range(6)], dtype=np.int32)
    sr, si = _real_imag(z["D_s"])
    tr, ti = _real_imag(z["D_t"])
    S = [orc.BST.from_dense(x, ad[:5], qn[:5]) for x in (sr, si)]
    T = [orc.BST.from_dense(x, ad_t, qn[2:]) for x in (tr, ti)]
    ref = z["D_r"]
    for axs in (orc.LEADING, orc.TRAILING):
        Sp = S if axs == orc.TRAILING else [orc.transpose([2, 3, 4, 0, 1], x) for x in S]
        for axt in (orc.LEADING, orc.TRAILING):
            Tp = T if axt == orc.LEADING else [orc.transpose([3, 4, 5, 0, 1, 2], x) for x in T]
            rr = orc.dot(Sp[0], axs, Tp[0], axt, nm).to_dense() - orc.dot(Sp[1], axs, Tp[1], axt, nm).to_dense()
            ri = orc.dot(Sp[0], axs, Tp[1], axt, nm).to_dense() + orc.dot(Sp[1], axs, Tp[0], axt, nm).to_dense()
            assert np.max(np.abs(rr + 1j * ri - ref)) < 1e-13


def test_block_sparse_dot_matches_tensordot():
    """Closed form: the block-sparse product equals numpy.tensordot of the densified operands."""
    rng = np.random.default_rng(5)
    dims = [6, 4, 5, 7]
    qn = [rng.integers(-2, 3, size=d).astype(np.int32) for d in dims]
    a = orc.BST.zeros(dims[:3], [1, 1, -1], qn[:3]).fill_random(11)
    b = orc.BST.zeros([5, 3, 7], [1, -1, 1], [qn[2], rng.integers(-1, 2, size=3).astype(np.int32), qn[3]]).fill_random(12)
    r = orc.dot(a, orc.TRAILING, b, orc.LEADING, 1)
    want = np.tensordot(a.to_dense(), b.to_dense(), axes=(2, 0))
    assert np.max(np.abs(r.to_dense() - want)) < 1e-13


def test_block_sparse_cyclic_partial_trace_fixture(golden):
    # test/tensor/test_block_sparse_tensor.c:163-230 (complex64 data; tolerance of the reference test)
    z = golden("test_block_sparse_tensor_cyclic_partial_trace")
    qn = _qnums(z, 7)
    ad = z["A_axis_dir"]
    ref = z["D_t_tr"]
    got = []
    for part in _real_imag(z["D_t"]):
        t = orc.BST.from_dense(part, ad, qn)
        got.append(orc.partial_trace(t, 2).to_dense())
    # the reference compares only within the quantum-number sectors of the result
    tr_ref = [orc.BST.from_dense(x, ad[2:5], qn[2:5]).to_dense() for x in _real_imag(ref)]
    assert np.max(np.abs(got[0] - tr_ref[0])) < 5e-6
    assert np.max(np.abs(got[1] - tr_ref[1])) < 5e-6


# ---------------------------------------------------------------------------------------
# dense factorisations (test/tensor/test_dense_tensor.c: qr, rq, svd)

def _dense_mats(z):
    return [z[k] for k in sorted(z.files) if z[k].dtype == np.float64]


def test_dense_qr_rq_fixtures(golden):
    L = orc.lib()
    for name, rq in (("test_dense_tensor_qr", 0), ("test_dense_tensor_rq", 1)):
        for a in _dense_mats(golden(name)):
            m, n = a.shape
            k = min(m, n)
            a = np.ascontiguousarray(a)
            if not rq:
                q = np.zeros((m, k)); r = np.zeros((k, n))
            else:
                q = np.zeros((k, n)); r = np.zeros((m, k))
            assert L.orc_h_qr_dense(m, n, orc.dptr(a), orc.dptr(q), orc.dptr(r), rq) == 0
            if not rq:
                assert np.max(np.abs(q @ r - a)) < 1e-13
                assert np.max(np.abs(q.T @ q - np.eye(k))) < 1e-13
                assert np.allclose(np.tril(r, -1), 0)
                # same reflector convention as LAPACK dgeqrf (numpy's qr calls it)
                r_np = np.linalg.qr(a, mode="r")
                assert np.max(np.abs(np.abs(r) - np.abs(r_np))) < 1e-12
                assert np.array_equal(np.sign(np.diag(r)), np.sign(np.diag(r_np)))
            else:
                assert np.max(np.abs(r @ q - a)) < 1e-13
                assert np.max(np.abs(q @ q.T - np.eye(k))) < 1e-13


def test_dense_svd_fixture(golden):
    L = orc.lib()
    for a in _dense_mats(golden("test_dense_tensor_svd")):
        m, n = a.shape
        k = min(m, n)
        a = np.ascontiguousarray(a)
        u = np.zeros((m, k)); s = np.zeros(k); vh = np.zeros((k, n))
        assert L.orc_h_svd_dense(m, n, orc.dptr(a), orc.dptr(u), orc.dptr(s), orc.dptr(vh)) == 0
        assert np.max(np.abs((u * s) @ vh - a)) < 1e-13
        assert np.max(np.abs(u.T @ u - np.eye(k))) < 1e-13
        assert np.max(np.abs(vh @ vh.T - np.eye(k))) < 1e-13
        assert np.max(np.abs(s - np.linalg.svd(a, compute_uv=False))) < 1e-13
        assert np.all(np.diff(s) <= 0)


def test_svd_rank_deficient_completes_isometry():
    L = orc.lib()
    rng = np.random.default_rng(1)
    a = np.ascontiguousarray(rng.standard_normal((9, 2)) @ rng.standard_normal((2, 6)))
    u = np.zeros((9, 6)); s = np.zeros(6); vh = np.zeros((6, 6))
    assert L.orc_h_svd_dense(9, 6, orc.dptr(a), orc.dptr(u), orc.dptr(s), orc.dptr(vh)) == 0
    assert np.max(np.abs(u.T @ u - np.eye(6))) < 1e-12
    assert np.max(np.abs((u * s) @ vh - a)) < 1e-13


def test_tridiagonal_eigh_matches_numpy():
    L = orc.lib()
    rng = np.random.default_rng(2)
    for n in (1, 2, 5, 25):
        d = rng.standard_normal(n)
        e = rng.standard_normal(max(n - 1, 0))
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        dd, ee, z = d.copy(), np.ascontiguousarray(np.append(e, 0.0)), np.zeros(n * n)
        assert L.orc_h_tridiag_eigh(n, orc.dptr(dd), orc.dptr(ee), orc.dptr(z)) == 0
        Z = z.reshape(n, n).T  # columns are eigenvectors
        w = np.linalg.eigvalsh(T)
        assert np.max(np.abs(dd - w)) < 1e-13
        assert np.max(np.abs(T @ Z - Z * dd)) < 1e-12


# ---------------------------------------------------------------------------------------
# Krylov (test/util/test_krylov.c)

def test_lanczos_iteration_fixture(golden):
    z = golden("test_lanczos_iteration_d")
    a = np.ascontiguousarray(z["D_a"])
    n = a.shape[0]
    maxiter = z["D_alpha"].size
    alpha = np.zeros(maxiter); beta = np.zeros(maxiter - 1); v = np.zeros((maxiter, n))
    numiter = C.c_int(0)
    orc.lib().orc_h_lanczos_dense(n, orc.dptr(a), orc.dptr(np.ascontiguousarray(z["D_vstart"])), maxiter,
                                  orc.dptr(alpha), orc.dptr(beta), orc.dptr(v), C.byref(numiter))
    assert numiter.value == maxiter
    assert np.max(np.abs(alpha - z["D_alpha"])) < 1e-13
    assert np.max(np.abs(beta - z["D_beta"])) < 1e-13
    assert np.max(np.abs(v - z["D_v"])) < 1e-13


def test_eigensystem_krylov_fixture(golden):
    z = golden("test_eigensystem_krylov_symmetric")
    a = np.ascontiguousarray(z["D_a"])
    n = a.shape[0]
    numeig = z["D_lambda"].size
    lam = np.zeros(numeig); u = np.zeros((n, numeig))
    ret = orc.lib().orc_h_eigensystem_dense(n, orc.dptr(a), orc.dptr(np.ascontiguousarray(z["D_vstart"])), 35, numeig,
                                            orc.dptr(lam), orc.dptr(u))
    assert ret == 0
    assert np.max(np.abs(lam - z["D_lambda"])) < 1e-13
    uref = z["D_u_ritz"]
    for e in range(numeig):  # Ritz vectors are fixed up to sign
        s = np.sign(np.dot(u[:, e], uref[:, e]))
        assert np.max(np.abs(s * u[:, e] - uref[:, e])) < 1e-12


# ---------------------------------------------------------------------------------------
# truncation and splitting

def test_retained_bond_indices_fixture(golden):
    # test/algorithm/test_truncation.c:8-80
    z = golden("test_retained_bond_indices")
    sigma = np.ascontiguousarray(z["D_sigma"])
    n = sigma.size
    ind_ref = z["D_ind"]
    for i in range(2):
        ind = np.zeros(n, dtype=np.int64)
        info = np.zeros(3)
        num = orc.lib().orc_h_retained_bond_indices(orc.dptr(sigma), n, float(z["A_tol"]) if i == 0 else 1e-5, 1,
                                                    n if i == 0 else ind_ref.size, orc.lptr(ind), orc.dptr(info))
        assert num == ind_ref.size
        assert np.array_equal(ind[:num], ind_ref)
        assert abs(info[0] - float(z["D_norm_sigma"])) < 1e-13
        assert abs(info[1] - float(z["D_entropy"])) < 1e-13


def test_mps_split_tensor_svd_fixture(golden):
    # test/state/test_mps.c:518-630
    z = golden("test_mps_split_tensor_svd")
    q0 = z["A_qsite0"].astype(np.int32)
    q1 = z["A_qsite1"].astype(np.int32)
    qf = (q0[:, None] + q1[None, :]).reshape(-1).astype(np.int32)
    qb0 = z["A_qbonds0"].astype(np.int32)
    qb1 = z["A_qbonds1"].astype(np.int32)
    a = orc.BST.from_dense(z["D_a_pair"], [1, 1, -1], [qb0, qf, qb1])
    L = orc.lib()
    for i in range(2):
        for distr in (0, 1):
            p0, p1 = orc.P(), orc.P()
            info = np.zeros(3)
            tol = 0.0 if i == 0 else float(z["A_tol"])
            ret = L.orc_h_mps_split_svd(a.ptr, 4, 5, orc.iptr(q0), orc.iptr(q1), tol, 100, 0, distr,
                                        C.byref(p0), C.byref(p1), orc.dptr(info))
            assert ret == 0
            a0, a1 = orc.BST(p0.value), orc.BST(p1.value)
            m = orc.BST(L.orc_h_mps_merge(a0.ptr, a1.ptr))
            ref = a.to_dense() if i == 0 else z["D_a_mrg"]
            assert np.max(np.abs(m.to_dense() - ref)) < 1e-13


# ---------------------------------------------------------------------------------------
# MPO construction (test/operator/test_hamiltonian.c)

def test_heisenberg_xxz_mpo_fixture(golden):
    mpo = orc.MPO.xxz(7, 14. / 25, 13. / 8, 2. / 7)
    got = mpo.to_matrix()
    assert np.max(np.abs(got - golden("test_heisenberg_xxz_1d_mpo")["D_heisenberg_xxz_1d_mat"])) < 1e-13


def test_fermi_hubbard_mpo_fixture(golden):
    mpo = orc.MPO.hubbard(4, 11. / 9, 13. / 4, 3. / 7)
    got = mpo.to_matrix()
    assert np.max(np.abs(got - golden("test_fermi_hubbard_1d_mpo")["D_fermi_hubbard_1d_mat"])) < 1e-13


def test_svd_stress_matrices_oracle():
    """the oracle's Jacobi SVD on inputs with columns in the span of the others, triplicated
    columns and graded columns, against numpy's LAPACK SVD"""
    from test_product_host import _check_svd, _svd_stress_matrices
    L = orc.lib()
    for a in _svd_stress_matrices(np.random.default_rng(7)):
        m, n = a.shape
        k = min(m, n)
        a = np.ascontiguousarray(a)
        u = np.zeros((m, k)); s = np.zeros(k); vh = np.zeros((k, n))
        assert L.orc_h_svd_dense(m, n, orc.dptr(a), orc.dptr(u), orc.dptr(s), orc.dptr(vh)) == 0, a.shape
        _check_svd(a, u, s, vh)
