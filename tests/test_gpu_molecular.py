"""Two-site DMRG on the GPU with the spin molecular Hamiltonian MPO (configuration 3's operator at
small orbital counts) against exact diagonalisation of the same Hamiltonian in the target sector."""

import sys

import numpy as np
import pytest

import chemtensor_amd as ct
import chemtensor_amd.compat as chemtensor

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sector_ground_energy(h, n, N, twosz):
    # local basis index 2 n_up + n_dn, site 0 most significant
    occ = np.array([[0, 0], [0, 1], [1, 0], [1, 1]])   # (n_up, n_dn) per local index
    idx = np.indices([4] * n).reshape(n, -1).T
    nup = occ[idx][:, :, 0].sum(axis=1)
    ndn = occ[idx][:, :, 1].sum(axis=1)
    sel = np.nonzero((nup + ndn == N) & (nup - ndn == twosz))[0]
    return np.linalg.eigvalsh(h[np.ix_(sel, sel)])[0]


@pytest.mark.parametrize("n,N,chi,seed", [(4, 4, 32, 1), (6, 6, 64, 2)])
def test_dmrg_spin_molecular_matches_exact_diagonalisation(n, N, chi, seed):
    t, v = ct.molecular_coefficients(n, seed)
    mpo = chemtensor.construct_spin_molecular_hamiltonian_mpo(t, v)
    h = mpo.to_matrix()
    assert np.max(np.abs(h - h.T)) < 1e-12
    e0 = sector_ground_energy(h, n, N, 0)
    psi, en, ent = chemtensor.dmrg(mpo, num_sweeps=4, maxiter_lanczos=20, tol_split=1e-12, max_vdim=chi,
                                   qnum_sector=chemtensor.encode_quantum_number_pair(N, 0), rng_seed=seed)
    say("molecular n", n, "energies", en, "exact", e0)
    assert abs(en[-1] - e0) < 1e-8
    vec = psi.to_statevector()
    assert abs(np.linalg.norm(vec) - 1) < 1e-10
    assert abs(vec @ h @ vec - en[-1]) < 1e-9


def test_dmrg_spin_molecular_split_fp32_matches_fp64(monkeypatch):
    # the precision option on the molecular operator: the same seeded sweeps with split-FP32
    # matvecs (CT_DMRG_MATVEC=split) agree with the FP64 sweeps within 1e-10 per sweep and reach
    # the exact ground energy
    n, N, chi, seed = 6, 6, 64, 3
    t, v = ct.molecular_coefficients(n, seed)
    mpo = chemtensor.construct_spin_molecular_hamiltonian_mpo(t, v)
    e0 = sector_ground_energy(mpo.to_matrix(), n, N, 0)
    kw = dict(num_sweeps=4, maxiter_lanczos=20, tol_split=1e-12, max_vdim=chi,
              qnum_sector=chemtensor.encode_quantum_number_pair(N, 0), rng_seed=seed)
    _, en64, _ = chemtensor.dmrg(mpo, **kw)
    monkeypatch.setenv("CT_DMRG_MATVEC", "split")
    _, en32, _ = chemtensor.dmrg(mpo, **kw)
    say("molecular split", en32, "fp64", en64, "exact", e0)
    assert np.max(np.abs(np.array(en32) - np.array(en64))) < 1e-10 * max(1.0, abs(e0))
    assert abs(en32[-1] - e0) < 1e-8
