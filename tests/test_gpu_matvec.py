"""GPU parity of the H_eff matvec path (HIP product through the C ABI) against the CPU oracle.

Tolerances: relative Frobenius error <= 1e-12 for every matvec form (the bar of the FP64
reference path); bit-exact block bookkeeping (identical packed layouts and lengths)."""

import ctypes as C
import sys

import numpy as np
import pytest

import chemtensor_amd as ct
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_mfma_tile_layout_single_block():
    """One dense block (all quantum numbers zero): the grouped FP64 MFMA dot equals numpy for
    shapes that exercise partial tiles, K tails and both transposition flags."""
    rng = np.random.default_rng(0)
    for (m, k, n) in [(1, 1, 1), (16, 4, 16), (17, 5, 33), (40, 37, 29), (64, 3, 70)]:
        say("case", m, k, n)
        a = rng.standard_normal((m, k))
        b = rng.standard_normal((k, n))
        A = orc.BST.from_dense(a, [1, -1], [np.zeros(m, np.int32), np.zeros(k, np.int32)])
        B = orc.BST.from_dense(b, [1, -1], [np.zeros(k, np.int32), np.zeros(n, np.int32)])
        r = ct.dot(A, ct.TRAILING, B, ct.LEADING, 1)
        assert rel(r.serialize(), (a @ b).reshape(-1)) < 1e-14
        # A^T as leading axes, B^T as trailing axes
        At = orc.BST.from_dense(a.T.copy(), [-1, 1], [np.zeros(k, np.int32), np.zeros(m, np.int32)])
        Bt = orc.BST.from_dense(b.T.copy(), [1, -1], [np.zeros(n, np.int32), np.zeros(k, np.int32)])
        r2 = ct.dot(At, ct.LEADING, Bt, ct.TRAILING, 1)
        assert rel(r2.serialize(), (a @ b).reshape(-1)) < 1e-14


def test_dot_fixture(golden):
    """test/tensor/test_block_sparse_tensor.c:1117-1230 on the GPU (complex via four real products)."""
    z = golden("test_block_sparse_tensor_dot")
    ad = z["A_axis_dir"].astype(np.int32)
    qn = [z[f"A_qnums{i}"].astype(np.int32) for i in range(8)]
    ad_t = np.array([(-1 if i < 3 else 1) * ad[2 + i] for i in range(6)], dtype=np.int32)
    S = [orc.BST.from_dense(np.ascontiguousarray(x), ad[:5], qn[:5]) for x in (z["D_s"].real, z["D_s"].imag)]
    T = [orc.BST.from_dense(np.ascontiguousarray(x), ad_t, qn[2:]) for x in (z["D_t"].real, z["D_t"].imag)]
    ref = z["D_r"]
    for axs in (ct.LEADING, ct.TRAILING):
        Sp = S if axs == ct.TRAILING else [orc.transpose([2, 3, 4, 0, 1], x) for x in S]
        for axt in (ct.LEADING, ct.TRAILING):
            say("axis ranges", axs, axt)
            Tp = T if axt == ct.LEADING else [orc.transpose([3, 4, 5, 0, 1, 2], x) for x in T]
            outs = {}
            for i in range(2):
                for j in range(2):
                    g = ct.dot(Sp[i], axs, Tp[j], axt, 3)
                    o = orc.dot(Sp[i], axs, Tp[j], axt, 3)
                    v = g.serialize()
                    assert v.size == o.nelem()
                    o.deserialize(v)
                    outs[i, j] = o.to_dense()
            got = (outs[0, 0] - outs[1, 1]) + 1j * (outs[0, 1] + outs[1, 0])
            assert np.max(np.abs(got - ref)) < 1e-13


def test_transpose_fixture(golden):
    z = golden("test_block_sparse_tensor_transpose")
    t = orc.BST.from_dense(z["D_t"], z["A_axis_dir"], [z[f"A_qnums{i}"].astype(np.int32) for i in range(4)])
    g = ct.transpose([1, 3, 2, 0], t)
    o = orc.transpose([1, 3, 2, 0], t)
    assert np.array_equal(g.serialize(), o.serialize())


def test_cyclic_partial_trace_fixture(golden):
    """test/tensor/test_block_sparse_tensor.c:163-230 on the GPU: the reference's complex64 tensor
    (7 axes, trace over 2) split into real and imaginary parts (the trace is linear), each traced by
    block_sparse_tensor_cyclic_partial_trace against the reference result (tolerance of the reference
    test, 5e-6 for single-precision data) and against the oracle's trace (1e-12)."""
    z = golden("test_block_sparse_tensor_cyclic_partial_trace")
    qn = [z[f"A_qnums{i}"].astype(np.int32) for i in range(7)]
    ad = z["A_axis_dir"].astype(np.int32)
    for part_t, part_r in zip((z["D_t"].real, z["D_t"].imag), (z["D_t_tr"].real, z["D_t_tr"].imag)):
        t = orc.BST.from_dense(np.ascontiguousarray(part_t, dtype=np.float64), ad, qn)
        g = ct.cyclic_partial_trace(t, 2)
        o = orc.partial_trace(t, 2)
        want = o.to_dense()
        v = g.serialize()
        assert v.size == o.nelem()
        o.deserialize(v)
        got = o.to_dense()
        assert np.max(np.abs(got - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))
        ref = orc.BST.from_dense(np.ascontiguousarray(part_r, dtype=np.float64), ad[2:5], qn[2:5]).to_dense()
        assert np.max(np.abs(got - ref)) < 5e-6


@pytest.mark.parametrize("nt,seed", [(1, 3), (2, 4), (0, 5)])
def test_cyclic_partial_trace_random(nt, seed):
    """Random block-sparse tensors with traced outer axes of dimension > 1 (several blocks per traced
    axis) against the oracle's trace at 1e-13, and a dense einsum closed form."""
    rng = np.random.default_rng(seed)
    dims_t = [int(d) for d in rng.integers(3, 6, size=nt)]
    dims_r = [7, 6]
    qn_t = [rng.integers(-1, 2, size=d).astype(np.int32) for d in dims_t]
    qn_r = [rng.integers(-1, 2, size=d).astype(np.int32) for d in dims_r]
    dims = dims_t + dims_r + dims_t
    ad = np.array([1] * nt + [1, -1] + [-1] * nt, dtype=np.int32)
    t = orc.BST.zeros(dims, ad, qn_t + qn_r + qn_t).fill_random(seed)
    g = ct.cyclic_partial_trace(t, nt)
    o = orc.partial_trace(t, nt)
    want = o.to_dense()
    v = g.serialize()
    assert v.size == o.nelem()
    o.deserialize(v)
    assert np.max(np.abs(o.to_dense() - want)) < 1e-13
    d = t.to_dense()
    if nt == 1:
        dense = np.einsum("iabi->ab", d)
    elif nt == 2:
        dense = np.einsum("ijabij->ab", d)
    else:
        dense = d
    assert np.max(np.abs(o.to_dense() - dense)) < 1e-13


def test_mpo_merge_matches_oracle():
    mpo = orc.MPO.hubbard(6, 1.0, 4.0, 0.3)
    for i in range(5):
        g = ct.mpo_merge_tensor_pair(mpo.site(i), mpo.site(i + 1))
        o = orc.BST(orc.lib().orc_h_mpo_merge(mpo.site(i).ptr, mpo.site(i + 1).ptr))
        assert g.nelem() == o.nelem()
        assert np.array_equal(g.serialize(), o.serialize())


@pytest.mark.parametrize("model,L,chi", [("xxz", 12, 16), ("hubbard", 8, 24), ("hubbard", 32, 64), ("xxz", 32, 128), ("hubbard", 32, 256),
                                         ("hubbard", 32, 512), ("xxz", 32, 512)])
def test_matvec_all_forms_match_oracle(model, L, chi):
    say("matvec", model, L, chi)
    theta, h2, left, right, _ = orc.matvec_problem(model, L, chi, 1234)
    want = orc.apply_local_hamiltonian(theta, h2, left, right).serialize()
    x = ct.DeviceArray.from_numpy(theta.serialize())
    for mode in (ct.PLAIN_FP64, ct.GROUPED_FP64, ct.GROUPED_SPLIT, ct.FUSED_FP64):
        say("  mode", mode)
        plan = ct.MatvecPlan(theta, h2, left, right, mode)
        assert plan.dim == theta.nelem()
        y = ct.DeviceArray(plan.dim)
        plan.apply(x, y)
        plan.synchronize()
        got = y.numpy()
        err = rel(got, want)
        say("  rel err", err)
        assert err < TOL
        # repeated application is deterministic (no atomics, fixed summation order)
        plan.apply(x, y)
        plan.synchronize()
        assert np.array_equal(y.numpy(), got)


@pytest.mark.parametrize("stage2", ["mfma", "stream", "fused12"])
def test_grouped_stage2_forms_match_oracle(stage2, monkeypatch):
    """The stage-2 forms of the grouped matvec (MFMA tiles, column stream, stages 1 and 2 fused)
    against the oracle."""
    monkeypatch.setenv("CT_STAGE2", "mfma" if stage2 == "mfma" else "stream")
    monkeypatch.setenv("CT_STAGE12", "fused" if stage2 == "fused12" else "separate")
    for model, L, chi in [("hubbard", 32, 512), ("xxz", 32, 256), ("hubbard", 8, 24)]:
        theta, h2, left, right, _ = orc.matvec_problem(model, L, chi, 4242)
        want = orc.apply_local_hamiltonian(theta, h2, left, right).serialize()
        plan = ct.MatvecPlan(theta, h2, left, right, ct.GROUPED_FP64)
        expect = {"mfma": 0, "stream": 4, "fused12": 5}[stage2]
        if stage2 == "mfma":
            assert plan.forms()[0][1] not in (4, 5)
        else:
            assert plan.forms()[0][1] == expect
        x = ct.DeviceArray.from_numpy(theta.serialize())
        y = ct.DeviceArray(plan.dim)
        plan.apply(x, y)
        plan.synchronize()
        err = rel(y.numpy(), want)
        say("  stage2", stage2, model, L, chi, "rel err", err)
        assert err < TOL


@pytest.mark.parametrize("stage2", ["mfma", "stream"])
def test_split_form_stage2_options_match_oracle(stage2, monkeypatch):
    """The split-FP32 form with stage 2 as sliced FP32 MFMA tiles and as the FP64 column stream
    (its default), against the oracle."""
    monkeypatch.setenv("CT_STAGE2", stage2)
    for model, L, chi in [("hubbard", 32, 512), ("xxz", 32, 256), ("hubbard", 8, 24)]:
        theta, h2, left, right, _ = orc.matvec_problem(model, L, chi, 4243)
        want = orc.apply_local_hamiltonian(theta, h2, left, right).serialize()
        plan = ct.MatvecPlan(theta, h2, left, right, ct.GROUPED_SPLIT)
        x = ct.DeviceArray.from_numpy(theta.serialize())
        y = ct.DeviceArray(plan.dim)
        plan.apply(x, y)
        plan.synchronize()
        err = rel(y.numpy(), want)
        say("  split stage2", stage2, model, L, chi, "rel err", err)
        assert err < TOL


@pytest.mark.parametrize("model,L,chi", [("hubbard", 16, 48), ("hubbard", 32, 256), ("xxz", 32, 512), ("hubbard", 32, 512)])
def test_apply_local_hamiltonian_dropin_with_reference_structs(model, L, chi):
    """The oracle's structs (reference layout) go straight into the product's drop-in entry point
    (operands copied to the device on every call), up to the headline size."""
    say("drop-in", model, L, chi)
    theta, h2, left, right, _ = orc.matvec_problem(model, L, chi, 77)
    b = ct.apply_local_hamiltonian(theta, h2, left, right)
    want = orc.apply_local_hamiltonian(theta, h2, left, right)
    dims_b, ad_b, qn_b = b.meta()
    dims_w, ad_w, qn_w = want.meta()
    assert list(dims_b) == list(dims_w) and list(ad_b) == list(ad_w)
    assert all(np.array_equal(x, y) for x, y in zip(qn_b, qn_w))
    assert rel(b.serialize(), want.serialize()) < TOL


def test_generated_problem_matches_oracle_generator():
    """ct_generate_matvec_problem (product) and the oracle's generator give identical inputs."""
    for model, L, chi in [("xxz", 16, 32), ("hubbard", 32, 128)]:
        g = ct.generate_matvec_problem(model, L, chi, 4321)
        o = orc.matvec_problem(model, L, chi, 4321)[:4]
        for gt, ot in zip(g, o):
            assert gt.nelem() == ot.nelem()
            assert np.array_equal(gt.serialize(), ot.serialize())


def test_lanczos_ground_state_matches_oracle():
    """Device Lanczos (krylov.c:172-251) on physical environments equals the oracle's."""
    mpo = orc.MPO.hubbard(8, 1.0, 4.0, 0.5)
    psi = orc.MPS.random(8, 4, mpo.qsite, orc.encode_pair(8, 0), 12, 5)
    theta, h2, left, right = orc.local_problem(mpo, psi, 3)
    n = theta.nelem()
    vstart = theta.serialize()
    for maxiter, mode in ((3, ct.GROUPED_FP64), (10, ct.GROUPED_FP64), (25, ct.GROUPED_FP64), (25, ct.FUSED_FP64)):
        say("maxiter", maxiter, "mode", mode)
        plan = ct.MatvecPlan(theta, h2, left, right, mode)
        lam, u = plan.lanczos(ct.DeviceArray.from_numpy(vstart), maxiter)
        lam_o = np.zeros(1)
        u_o = np.zeros(n)
        h = orc.lib()
        # oracle: same matrix-free operator through the oracle's own Krylov code
        ham = theta.copy()

        def afunc(v):
            ham.deserialize(v)
            return orc.apply_local_hamiltonian(ham, h2, left, right).serialize()
        # dense H_eff from the oracle matvec (small n) -> oracle Lanczos on the dense matrix
        H = np.stack([afunc(e) for e in np.eye(n)], axis=1)
        ret = h.orc_h_eigensystem_dense(n, orc.dptr(np.ascontiguousarray(H)), orc.dptr(vstart), maxiter, 1, orc.dptr(lam_o), orc.dptr(u_o))
        assert ret == 0
        assert abs(lam - lam_o[0]) < 1e-10 * max(1, abs(lam_o[0]))
        ug = u.numpy()
        s = np.sign(np.dot(ug, u_o))
        assert rel(s * ug, u_o) < 1e-9


@pytest.mark.parametrize("deferred", ["1", "0"])
@pytest.mark.parametrize("mode", [ct.GROUPED_FP64, ct.PAIR_FP64])
def test_lanczos_breakdown_device_plan(mode, deferred, monkeypatch):
    """Device plans at a Krylov breakdown (maxiter above the dimension of the local problem): the
    recurrence stops where the oracle's does (krylov.c:58) and the ground energy and Ritz vector agree
    with it, with each vector's normalisation deferred into the next step's update (the default) and
    in a launch of its own (CT_LANCZOS_DEFERRED=0), eagerly and replayed from a graph."""
    monkeypatch.setenv("CT_LANCZOS_DEFERRED", deferred)
    mpo = orc.MPO.xxz(6, 1.0, 0.5, 0.0)
    psi = orc.MPS.random(6, 2, mpo.qsite, 0, 4, 3)
    theta, h2, left, right = orc.local_problem(mpo, psi, 1)
    n = theta.nelem()
    vstart = theta.serialize()
    ham = theta.copy()

    def afunc(v):
        ham.deserialize(v)
        return orc.apply_local_hamiltonian(ham, h2, left, right).serialize()
    H = np.stack([afunc(e) for e in np.eye(n)], axis=1)
    for maxiter in (n + 3, n + 12):
        lam_o = np.zeros(1)
        u_o = np.zeros(n)
        assert orc.lib().orc_h_eigensystem_dense(n, orc.dptr(np.ascontiguousarray(H)), orc.dptr(vstart), maxiter, 1, orc.dptr(lam_o),
                                                 orc.dptr(u_o)) == 0
        for graph in ("1", "0"):
            monkeypatch.setenv("CT_LANCZOS_GRAPH", graph)
            plan = ct.MatvecPlan(theta, h2, left, right, mode)
            lam, u = plan.lanczos(ct.DeviceArray.from_numpy(vstart), maxiter)
            say("n", n, "maxiter", maxiter, "graph", graph, "lambda", lam, "oracle", lam_o[0])
            assert abs(lam - lam_o[0]) < 1e-10 * max(1, abs(lam_o[0]))
            ug = u.numpy()
            # without re-orthogonalisation the Krylov vectors drift from orthogonality near a
            # breakdown, and the last of them carry rounding noise amplified by 1 / beta (beta near
            # the breakdown threshold): the Ritz vector (norm included) agrees with the oracle's to
            # 1e-8, not to rounding
            say("  |u| - |u_oracle|", np.linalg.norm(ug) - np.linalg.norm(u_o), "rel", rel(np.sign(np.dot(ug, u_o)) * ug, u_o))
            assert abs(np.linalg.norm(ug) - np.linalg.norm(u_o)) < 1e-8
            assert rel(np.sign(np.dot(ug, u_o)) * ug, u_o) < 1e-8


def test_eigensystem_krylov_symmetric_fixture(golden):
    """The exported eigensystem_krylov_symmetric (host callback, several eigenpairs) reproduces the
    reference fixture test/util/data/test_eigensystem_krylov_symmetric.hdf5 (maxiter 35)."""
    z = golden("test_eigensystem_krylov_symmetric")
    a = np.ascontiguousarray(z["D_a"])
    numeig = z["D_lambda"].size
    say("eigensystem n", a.shape[0], "numeig", numeig)
    lam, u = ct.eigensystem_krylov_symmetric(lambda v: a @ v, z["D_vstart"], 35, numeig)
    assert np.max(np.abs(lam - z["D_lambda"])) < 1e-13
    uref = z["D_u_ritz"]
    for e in range(numeig):  # Ritz vectors are fixed up to sign
        s = np.sign(np.dot(u[:, e], uref[:, e]))
        assert np.max(np.abs(s * u[:, e] - uref[:, e])) < 1e-12
    with pytest.raises(RuntimeError):
        ct.eigensystem_krylov_symmetric(lambda v: a @ v, z["D_vstart"], 2, 3)   # numeig > maxiter


@pytest.mark.parametrize("mode", [ct.FUSED_FP64, ct.GROUPED_FP64])
def test_lanczos_graph_replays_match_eager(mode, monkeypatch):
    """The Lanczos recurrence captured in a HIP graph gives bit-identical results to the eager
    launches, on every replay, with other eager work and synchronisations between the replays and
    a different start vector on the last replay."""
    theta, h2, left, right, _ = orc.matvec_problem("hubbard", 16, 64, 99)
    plan = ct.MatvecPlan(theta, h2, left, right, mode)
    rng = np.random.default_rng(3)
    starts = [rng.standard_normal(plan.dim) for _ in range(2)]
    eager = []
    monkeypatch.setenv("CT_LANCZOS_GRAPH", "0")
    for v in starts:
        lam, u = plan.lanczos(ct.DeviceArray.from_numpy(v), 12)
        eager.append((lam, u.numpy()))
    monkeypatch.setenv("CT_LANCZOS_GRAPH", "1")
    x = ct.DeviceArray.from_numpy(theta.serialize())
    y = ct.DeviceArray(plan.dim)
    for rep, which in enumerate((0, 0, 0, 1)):
        say("graph replay", rep, "start", which)
        lam, u = plan.lanczos(ct.DeviceArray.from_numpy(starts[which]), 12)
        assert lam == eager[which][0]
        assert np.array_equal(u.numpy(), eager[which][1])
        plan.apply(x, y)          # eager work between replays
        plan.synchronize()
