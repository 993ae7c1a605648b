"""Pin the oracle's H_eff matvec and two-site DMRG: closed forms (numpy einsum on densified
operands), invariants (symmetry of H_eff with physical environments), exact diagonalisation,
and the reference's published notebook energies (examples/dmrg/basic_dmrg_fermi_hubbard.ipynb)."""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as orc


def heff_dense(theta, h2, left, right):
    """Closed form of apply_local_hamiltonian (src/algorithm/chain_ops.c:353-390) on dense arrays:
    y[a', p', b'] = sum L[o, a, m, a'] W[m, p', p, n] theta[a, p, b] R[b, n, b', o]."""
    return np.einsum("oama,mqpn,apb,bnco->aqc".replace("oama", "oamx").replace("aqc", "xqc"),
                     left.to_dense(), h2.to_dense(), theta.to_dense(), right.to_dense())


@pytest.mark.parametrize("model,L,chi", [("xxz", 12, 16), ("hubbard", 8, 24)])
def test_apply_local_hamiltonian_closed_form(model, L, chi):
    theta, h2, left, right, _ = orc.matvec_problem(model, L, chi, 1234)
    y = orc.apply_local_hamiltonian(theta, h2, left, right)
    want = heff_dense(theta, h2, left, right)
    # entries outside the sectors of y must vanish in the closed form as well
    assert np.max(np.abs(y.to_dense() - want)) < 1e-12 * max(1.0, np.max(np.abs(want)))
    # the result has the block structure of theta (it is the next Lanczos vector)
    assert y.nelem() == theta.nelem()


def test_apply_local_hamiltonian_symmetric_with_physical_environments():
    mpo = orc.MPO.hubbard(8, 1.0, 4.0, 0.5)
    psi = orc.MPS.random(8, 4, mpo.qsite, orc.encode_pair(8, 0), 20, 99)
    theta, h2, left, right = orc.local_problem(mpo, psi, 3)
    n = theta.nelem()
    rng = np.random.default_rng(0)
    x, z = rng.standard_normal(n), rng.standard_normal(n)
    tx, tz = theta.copy(), theta.copy()
    tx.deserialize(x)
    tz.deserialize(z)
    hx = orc.apply_local_hamiltonian(tx, h2, left, right).serialize()
    hz = orc.apply_local_hamiltonian(tz, h2, left, right).serialize()
    assert abs(np.dot(z, hx) - np.dot(x, hz)) < 1e-11 * np.linalg.norm(hx) * np.linalg.norm(z)


def test_dmrg_notebook_fermi_hubbard():
    """examples/dmrg/basic_dmrg_fermi_hubbard.ipynb:197, :215, :236, :304-307 (seed 42)."""
    mpo = orc.MPO.hubbard(6, 1.0, 4.0, 1.5)
    psi = orc.MPS.random(6, 4, mpo.qsite, orc.encode_pair(7, 1), 32, 42)
    ret, en, ent = orc.dmrg(mpo, psi, 4, 5, 1e-8, 32)
    assert ret == 0
    en_ref = [-18.48431136702258, -18.484358903148298, -18.484358904136563, -18.484358904136524]
    assert np.max(np.abs(en - en_ref)) < 1e-10
    assert psi.bond_dims() == [1, 4, 16, 30, 16, 4, 1]
    ent_ref = [1.057297431102482, 1.0833060473946288, 1.1618575108707525, 1.0833062031796814, 1.0572974453589872]
    assert np.max(np.abs(ent - ent_ref)) < 1e-8
    v = psi.to_vector(4)
    assert abs(np.linalg.norm(v) - 1) < 1e-12
    # <psi|H|psi> equals the last sweep energy; ED ground energy from the notebook (:372)
    h = mpo.to_matrix()
    assert abs(v @ h @ v - en[-1]) < 1e-10
    assert abs(np.linalg.eigvalsh(h)[0] - (-18.484358962762183)) < 1e-10


def xxz_sparse_sector(L, J, D, h, sz2):
    """Sparse XXZ Hamiltonian restricted to total 2Sz = sz2 (basis bit 1 = spin down)."""
    states = [s for s in range(1 << L) if L - 2 * bin(s).count("1") == sz2]
    index = {s: i for i, s in enumerate(states)}
    rows, cols, vals = [], [], []
    for i, s in enumerate(states):
        diag = 0.0
        for j in range(L):
            zj = 0.5 - ((s >> (L - 1 - j)) & 1)
            diag -= h * zj
            if j < L - 1:
                zk = 0.5 - ((s >> (L - 2 - j)) & 1)
                diag += J * D * zj * zk
                if zj != zk:
                    t = s ^ (1 << (L - 1 - j)) ^ (1 << (L - 2 - j))
                    rows.append(index[t]); cols.append(i); vals.append(0.5 * J)
        rows.append(i); cols.append(i); vals.append(diag)
    n = len(states)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def test_xxz_mpo_matches_sparse_ed_operator():
    L = 8
    mpo = orc.MPO.xxz(L, 1.0, 0.5, 0.3)
    h = mpo.to_matrix()
    # restrict the MPO matrix to the 2Sz = 0 sector (qsite {+1, -1}, site 0 most significant)
    states = [s for s in range(1 << L) if L - 2 * bin(s).count("1") == 0]
    hs = h[np.ix_(states, states)]
    want = xxz_sparse_sector(L, 1.0, 0.5, 0.3, 0).toarray()
    assert np.max(np.abs(hs - want)) < 1e-13


def test_dmrg_xxz_config1_against_ed():
    """Config: XXZ L=16, J=1, D=0.5, h=0, 2Sz=0, chi=32, 4 sweeps, 10 Lanczos, tol 1e-12."""
    mpo = orc.MPO.xxz(16, 1.0, 0.5, 0.0)
    psi = orc.MPS.random(16, 2, mpo.qsite, 0, 32, 1)
    ret, en, _ = orc.dmrg(mpo, psi, 4, 10, 1e-12, 32)
    assert ret == 0
    e0 = spla.eigsh(xxz_sparse_sector(16, 1.0, 0.5, 0.0, 0), k=1, which="SA", tol=1e-14)[0][0]
    # variational, monotone over sweeps, and close to the exact energy at chi = 32
    assert en[-1] >= e0 - 1e-10
    assert abs(en[-1] - e0) < 1e-6
    assert np.all(np.diff(en) < 1e-9)
