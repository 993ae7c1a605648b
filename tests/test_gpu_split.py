"""GPU parity of the device split (block_sparse_tensor_svd's per-block factorisation,
block_sparse_tensor.c:2686-2848, with the truncation of bond_ops.c:15-138) in its two device forms:
the Gram-eigenvector factorisation (the default) and the one-sided Jacobi SVD. The kernels against
numpy's LAPACK SVD on hard inputs and the reference's SVD fixture, and whole DMRG sweeps with each
device split against the host split and the oracle."""

import os
import sys

import numpy as np
import pytest

import chemtensor_amd as ct
from oracle import oracle as orc
from test_product_host import _svd_stress_matrices

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_dense_svd_device_matches_lapack(golden):
    """reconstruction 1e-13, orthonormal singular vectors of the nonzero singular values 1e-12,
    singular values within 1e-13 ||a|| of numpy's (zero singular values come with zero vectors)"""
    mats = _svd_stress_matrices(np.random.default_rng(7))
    z = golden("test_dense_tensor_svd")
    mats += [z[k] for k in sorted(z.files) if z[k].dtype == np.float64 and z[k].ndim == 2]
    for a in mats:
        say("matrix", a.shape)
        u, s, vh = ct.dense_svd_device(a)
        scale = max(1.0, np.linalg.norm(a))
        assert np.max(np.abs((u * s) @ vh - a)) < 1e-13 * scale
        assert np.all(np.diff(s) <= 0) and np.all(s >= 0)
        assert np.max(np.abs(s - np.linalg.svd(a, compute_uv=False))) < 1e-13 * scale
        nz = s > 1e-12 * scale
        k = int(nz.sum())
        assert np.max(np.abs(u[:, nz].T @ u[:, nz] - np.eye(k)), initial=0) < 1e-12
        assert np.max(np.abs(vh[nz] @ vh[nz].T - np.eye(k)), initial=0) < 1e-12


def test_dense_gram_factor_device_matches_lapack(golden):
    """The default device factorisation on the same inputs, both isometric sides: the isometric factor
    orthonormal (1e-12), the other factor's row / column norms equal to s (1e-13 ||a||), s^2 within
    1e-13 ||a||^2 of numpy's (the eigenvalue accuracy of the Gram matrix), descending. a = u vh to
    1e-13 ||a|| when the isometric side is the shorter one; when it is the longer one (its Gram matrix
    is singular) directions with s^2 below the Gram matrix's accuracy eps ||a||^2 are not resolved,
    so the reconstruction holds to sqrt(n eps) ||a||."""
    mats = _svd_stress_matrices(np.random.default_rng(7))
    z = golden("test_dense_tensor_svd")
    mats += [z[k] for k in sorted(z.files) if z[k].dtype == np.float64 and z[k].ndim == 2]
    for a in mats:
        for left in (True, False):
            m, n = a.shape
            u, s, vh = ct.dense_gram_factor_device(a, left)
            scale = max(1.0, np.linalg.norm(a))
            k = min(m, n)
            rec = np.max(np.abs(u @ vh - a))
            iso = u.T @ u if left else vh @ vh.T
            other = np.linalg.norm(vh, axis=1) if left else np.linalg.norm(u, axis=0)
            s_np = np.linalg.svd(a, compute_uv=False)
            say("matrix", a.shape, "left" if left else "right", "rec %.1e" % (rec / scale))
            assert np.max(np.abs(iso - np.eye(k))) < 1e-12
            assert np.max(np.abs(other - s)) < 1e-13 * scale
            assert np.all(np.diff(s) <= 0) and np.all(s >= 0)
            assert np.max(np.abs(s ** 2 - s_np ** 2)) < 1e-13 * scale ** 2
            longer = (m if left else n) > k
            assert rec < (np.sqrt(max(m, n) * 2.2e-16) if longer else 1e-13) * scale


@pytest.mark.parametrize("split", ["device", "svd"])
@pytest.mark.parametrize("model,L,chi,sweeps,maxiter,tol,seed", [
    ("xxz", 16, 32, 3, 10, 1e-12, 1),
    ("hubbard", 10, 48, 2, 8, 1e-10, 7),
])
def test_dmrg_device_split_matches_host_split_and_oracle(model, L, chi, sweeps, maxiter, tol, seed, split):
    def run(split):
        os.environ["CT_SPLIT"] = split
        try:
            if model == "xxz":
                mpo = ct.MPO.xxz(L, 1.0, 0.5, 0.0)
                psi = ct.MPS.random(L, 2, mpo.qsite, 0, chi, seed)
            else:
                mpo = ct.MPO.hubbard(L, 1.0, 4.0, 0.0)
                psi = ct.MPS.random(L, 4, mpo.qsite, ct.encode_pair(L, 0), chi, seed)
            en, ent = ct.dmrg(mpo, psi, num_sweeps=sweeps, maxiter_lanczos=maxiter, tol_split=tol, max_vdim=chi)
            return en, ent, psi.bond_dims()
        finally:
            os.environ.pop("CT_SPLIT", None)
    en_d, ent_d, bd_d = run(split)
    en_h, ent_h, bd_h = run("host")
    say("device", en_d.tolist(), "host", en_h.tolist())
    assert np.max(np.abs(en_d - en_h)) < 1e-10
    assert np.max(np.abs(ent_d - ent_h)) < 1e-8
    assert bd_d == bd_h
    if model == "xxz":
        om = orc.MPO.xxz(L, 1.0, 0.5, 0.0)
        op = orc.MPS.random(L, 2, om.qsite, 0, chi, seed)
    else:
        om = orc.MPO.hubbard(L, 1.0, 4.0, 0.0)
        op = orc.MPS.random(L, 4, om.qsite, orc.encode_pair(L, 0), chi, seed)
    ret, en_o, ent_o = orc.dmrg(om, op, sweeps, maxiter, tol, chi)
    assert ret == 0
    assert np.max(np.abs(en_d - en_o)) < 1e-10
    assert np.max(np.abs(ent_d - ent_o)) < 1e-8
