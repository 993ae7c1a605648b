"""The reference's notebook flow (examples/dmrg/basic_dmrg_fermi_hubbard.ipynb) through the
reference-style Python surface chemtensor_amd.compat, with dmrg_twosite on the GPU."""

import sys

import numpy as np
import pytest

import chemtensor_amd.compat as chemtensor

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_notebook_flow_fermi_hubbard():
    nsites, t, u, mu = 6, 1.0, 4.0, 1.5
    hamiltonian = chemtensor.construct_fermi_hubbard_1d_mpo(nsites, t, u, mu)
    assert hamiltonian.d == 4
    assert [chemtensor.decode_quantum_number_pair(q) for q in hamiltonian.qsite] == [(0, 0), (1, -1), (1, 1), (2, 0)]
    qnum_sector = chemtensor.encode_quantum_number_pair(7, 1)
    psi, en_sweeps, entropy = chemtensor.dmrg(hamiltonian, num_sweeps=4, maxiter_lanczos=5, tol_split=1e-8, max_vdim=32,
                                              qnum_sector=qnum_sector)
    say("energies", en_sweeps)
    # the notebook's printed results (basic_dmrg_fermi_hubbard.ipynb:304-307)
    en_ref = [-18.48431136702258, -18.484358903148298, -18.484358904136563, -18.484358904136524]
    assert isinstance(en_sweeps, list) and isinstance(entropy, list)
    assert np.max(np.abs(np.array(en_sweeps) - en_ref)) < 1e-10
    assert psi.bond_dims == [1, 4, 16, 30, 16, 4, 1]
    psi_vec = psi.to_statevector()
    assert psi_vec.shape == (4 ** nsites,)
    assert abs(np.linalg.norm(psi_vec) - 1) < 1e-12
    h_mat = hamiltonian.to_matrix()
    assert abs(np.vdot(psi_vec, h_mat @ psi_vec) - en_sweeps[-1]) < 1e-10
    w_ref = np.linalg.eigvalsh(h_mat)
    say("exact ground state", w_ref[0], "difference", en_sweeps[-1] - w_ref[0])
    assert en_sweeps[-1] - w_ref[0] > -1e-10 and en_sweeps[-1] - w_ref[0] < 1e-6
