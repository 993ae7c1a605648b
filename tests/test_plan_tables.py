"""Host-side validation of the grouped matvec tables (no GPU needed): for several models, chain
lengths and bond dimensions, every operand index of the grouped stages, the column-stream stage 2
and the fused stages 1 and 2 lies inside its buffer, and every entry of T1, T2 and the result is
written by exactly one output."""

import pytest

import chemtensor_amd as ct
from oracle import oracle as orc


@pytest.mark.parametrize("model,L,chi,seed", [("xxz", 12, 16, 1), ("hubbard", 8, 24, 2), ("hubbard", 32, 64, 3),
                                              ("xxz", 32, 128, 4), ("hubbard", 24, 128, 5), ("xxz", 40, 256, 6)])
def test_plan_tables_valid(model, L, chi, seed):
    theta, h2, left, right, _ = orc.matvec_problem(model, L, chi, seed)
    assert ct.matvec_plan_check(theta, h2, left, right) == 0


def test_plan_tables_physical_environments():
    """Environments built by contraction from a DMRG state (oracle), not random blocks."""
    mpo = orc.MPO.hubbard(8, 1.0, 4.0, 0.5)
    psi = orc.MPS.random(8, 4, mpo.qsite, orc.encode_pair(8, 0), 12, 5)
    for i in range(7):
        theta, h2, left, right = orc.local_problem(mpo, psi, i)
        assert ct.matvec_plan_check(theta, h2, left, right) == 0


def test_plan_check_rejects_incompatible_operands():
    theta, h2, left, right, _ = orc.matvec_problem("xxz", 12, 16, 1)
    other = orc.matvec_problem("xxz", 12, 24, 9)
    assert ct.matvec_plan_check(theta, h2, left, other[3]) == ct.CT_ERR_ARG


@pytest.mark.parametrize("n,chi", [(12, 128), (20, 512)])
def test_plan_tables_molecular(n, chi):
    """The spin molecular operator's centre-bond problem (large W blocks): the grouped tables and the
    LDS-tiled split of the large tasks (64 x 64 tiles, the other tasks' 32 x 32 items) cover every
    output entry once."""
    t, v = ct.molecular_coefficients(n, 5)
    m = ct.MPO.spin_molecular(t, v)
    om = orc.MPOView(m.ptr, m.nsites, m.d, m.qsite)
    theta, h2, left, right = orc.matvec_problem_mpo(om, orc.encode_pair(n, 0), chi, 2024)
    assert ct.matvec_plan_check(theta, h2, left, right) == 0
