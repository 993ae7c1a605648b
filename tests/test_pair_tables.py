"""Host replay of the sector-pair matvec tables (ct_matvec_pair_check, no GPU): the unit records
and reduction items are walked with the kernels' own index arithmetic, tile order and guards.
Every operand, LDS and partial index must be in bounds, every LDS operand read after it was
written, every T1 / T2 / partial / result entry written exactly once, and the replayed product
must equal the oracle's apply_local_hamiltonian (chain_ops.c:353-390) to 1e-13."""

import numpy as np
import pytest

import chemtensor_amd as ct
from oracle import oracle as orc


def check(theta, h2, left, right, cap=0):
    st, y = ct.matvec_pair_check(theta, h2, left, right, cap)
    assert st == 0
    want = orc.apply_local_hamiltonian(theta, h2, left, right).serialize()
    assert y.size == want.size
    assert np.linalg.norm(y - want) <= 1e-13 * max(np.linalg.norm(want), 1e-300)


@pytest.mark.parametrize("model,L,chi,seed", [("hubbard", 16, 32, 11), ("xxz", 12, 16, 2), ("hubbard", 32, 128, 3), ("xxz", 32, 256, 4),
                                              ("hubbard", 32, 512, 7), ("hubbard", 8, 24, 5), ("xxz", 20, 64, 8)])
def test_pair_tables_random_blocks(model, L, chi, seed):
    check(*orc.matvec_problem(model, L, chi, seed)[:4])


@pytest.mark.parametrize("cap", [1, 8, 24, 100000])
def test_pair_tables_chunking(cap):
    """Pairs cut into row and column chunks (small MFMA caps) and left whole (large cap)."""
    check(*orc.matvec_problem("hubbard", 32, 256, 21)[:4], cap=cap)
    check(*orc.matvec_problem("xxz", 24, 128, 22)[:4], cap=cap)


@pytest.mark.parametrize("model,L,chi,seed,cap", [("hubbard", 32, 128, 31, 0), ("xxz", 20, 64, 32, 0), ("hubbard", 32, 256, 33, 0),
                                                  ("hubbard", 8, 24, 34, 0), ("hubbard", 32, 128, 35, 24), ("xxz", 24, 128, 36, 16)])
def test_pair_tables_packed_units(model, L, chi, seed, cap, monkeypatch):
    """Chunks packed into units (several chunks per workgroup, their LDS regions moved up) when the
    chunks exceed the resident units but fit them packed: the target is set to 2/3 and 1/2 of the
    chunk count, the packing is checked to have happened, and the replay (liveness of the moved
    regions included) must reproduce the oracle's product."""
    theta, h2, left, right = orc.matvec_problem(model, L, chi, seed)[:4]
    monkeypatch.setenv("CT_PAIR_SLOTS", "100000000")
    chunks = int(ct.matvec_pair_stats(theta, h2, left, right, cap)["units"])
    packed = False
    for slots in (chunks * 2 // 3, chunks // 2, chunks - max(1, chunks // 10)):
        monkeypatch.setenv("CT_PAIR_SLOTS", str(max(1, slots)))
        units = int(ct.matvec_pair_stats(theta, h2, left, right, cap)["units"])
        # packed into the target, or (when the chunks cannot fit it) left one per unit
        assert units <= max(1, slots) or units == chunks
        packed = packed or units < chunks
        check(theta, h2, left, right, cap=cap)
    if model == "hubbard":   # small sectors: chunks well below the LDS budget always pack
        assert packed


def test_pair_tables_physical_environments():
    """Environments built by contraction from DMRG states (oracle), every bond of the chain,
    including the edges where bond dimensions are 1."""
    mpo = orc.MPO.hubbard(8, 1.0, 4.0, 0.5)
    psi = orc.MPS.random(8, 4, mpo.qsite, orc.encode_pair(8, 0), 12, 5)
    for i in range(7):
        check(*orc.local_problem(mpo, psi, i))
    mpo = orc.MPO.xxz(10, 1.0, 0.7, 0.2)
    psi = orc.MPS.random(10, 2, mpo.qsite, 2, 16, 6)
    for i in range(9):
        check(*orc.local_problem(mpo, psi, i))


def test_pair_tables_xxz_large_sectors():
    """XXZ at chi = 512: sectors of about 100 states, every pair chunked to fit the LDS."""
    check(*orc.matvec_problem("xxz", 32, 512, 5)[:4])
