"""The reference's Python surface (chemtensor_amd.compat, after python/pymodule.c) on the CPU:
quantum-number pair helpers, dense MPO / MPS conversions against the oracle's, argument checks of
dmrg."""

import numpy as np
import pytest

import chemtensor_amd as ct
import chemtensor_amd.compat as chemtensor
from oracle import oracle as orc


def test_quantum_number_pairs_roundtrip():
    # qnumber.h:47-70; (7, 1) is the notebook's sector
    assert chemtensor.encode_quantum_number_pair(7, 1) == 7 * 65536 + 1
    for qa in (-3, -1, 0, 1, 2, 20):
        for qb in (-5, -1, 0, 1, 4):
            q = chemtensor.encode_quantum_number_pair(qa, qb)
            assert chemtensor.decode_quantum_number_pair(q) == (qa, qb)
            assert q == ct.encode_pair(qa, qb)


@pytest.mark.parametrize("model,L", [("xxz", 6), ("hubbard", 4)])
def test_mpo_to_matrix_matches_oracle(model, L):
    if model == "xxz":
        m, o = chemtensor.construct_heisenberg_xxz_1d_mpo(L, 1.0, 0.5, 0.2), orc.MPO.xxz(L, 1.0, 0.5, 0.2)
    else:
        m, o = chemtensor.construct_fermi_hubbard_1d_mpo(L, 1.0, 4.0, 1.5), orc.MPO.hubbard(L, 1.0, 4.0, 1.5)
    h = m.to_matrix()
    assert h.shape == (m.d ** L, m.d ** L)
    assert np.max(np.abs(h - o.to_matrix())) < 1e-14
    assert np.allclose(h, h.T)
    assert m.bond_dims[0] == 1 and m.bond_dims[-1] == 1 and len(m.bond_dims) == L + 1


def test_mps_to_statevector_matches_oracle():
    L, d = 5, 4
    qsite = [orc.encode_pair(0, 0), orc.encode_pair(1, -1), orc.encode_pair(1, 1), orc.encode_pair(2, 0)]
    sector = orc.encode_pair(5, 1)
    p = chemtensor.MPS(ct.MPS.random(L, d, qsite, sector, 16, 3))
    o = orc.MPS.random(L, d, qsite, sector, 16, 3)
    v = p.to_statevector()
    assert np.max(np.abs(v - o.to_vector(d))) < 1e-15
    assert abs(p.norm() - np.linalg.norm(v)) < 1e-12


def test_dmrg_argument_checks():
    # pymodule.c:3272-3295: the same exceptions and messages, raised before any GPU work
    h = chemtensor.construct_heisenberg_xxz_1d_mpo(6, 1.0, 0.5, 0.0)
    with pytest.raises(ValueError, match="'num_sweeps' must be a positive integer, received 0"):
        chemtensor.dmrg(h, num_sweeps=0)
    with pytest.raises(ValueError, match="'maxiter_lanczos' must be a positive integer"):
        chemtensor.dmrg(h, maxiter_lanczos=-2)
    with pytest.raises(ValueError, match=r"'tol_split' must be in the interval \[0, 1\]"):
        chemtensor.dmrg(h, tol_split=2.0)
    with pytest.raises(ValueError, match="'max_vdim' must be a positive integer"):
        chemtensor.dmrg(h, max_vdim=0)
    with pytest.raises(SyntaxError):
        chemtensor.dmrg("not an MPO")
