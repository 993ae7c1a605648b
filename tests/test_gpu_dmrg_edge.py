"""dmrg_twosite on the GPU against the oracle at the edges of the sweep: two and three sites (one or
two bonds), bond dimension 1 and 3, zero truncation tolerance, odd sectors."""

import sys

import numpy as np
import pytest

import chemtensor_amd as ct
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("model,L,chi,sweeps,maxiter,tol,seed", [
    ("xxz", 2, 4, 2, 5, 1e-10, 1),
    ("xxz", 3, 8, 2, 5, 1e-10, 2),
    ("hubbard", 2, 16, 2, 4, 1e-10, 3),
    ("hubbard", 3, 16, 2, 6, 1e-10, 4),
    ("xxz", 8, 1, 2, 4, 1e-10, 5),
    ("xxz", 9, 3, 2, 4, 0.0, 6),
    ("hubbard", 5, 64, 3, 8, 0.0, 7),
])
def test_dmrg_edges_match_oracle(model, L, chi, sweeps, maxiter, tol, seed):
    say("dmrg edge", model, L, chi, sweeps, maxiter, tol, seed)
    if model == "xxz":
        pm, om, sector = ct.MPO.xxz(L, 1.0, 0.5, 0.0), orc.MPO.xxz(L, 1.0, 0.5, 0.0), L % 2
    else:
        pm, om = ct.MPO.hubbard(L, 1.0, 4.0, 0.0), orc.MPO.hubbard(L, 1.0, 4.0, 0.0)
        sector = ct.encode_pair(L, L % 2)
    pp = ct.MPS.random(L, pm.d, pm.qsite, sector, chi, seed)
    op = orc.MPS.random(L, om.d, om.qsite, sector, chi, seed)
    ret, en_o, ent_o = orc.dmrg(om, op, sweeps, maxiter, tol, chi)
    assert ret == 0
    en_g, ent_g = ct.dmrg(pm, pp, sweeps, maxiter, tol, chi)
    say("gpu", en_g.tolist(), "oracle", en_o.tolist())
    assert np.max(np.abs(en_g - en_o)) < 1e-10
    assert pp.bond_dims() == op.bond_dims()
    assert np.max(np.abs(ent_g - ent_o)) < 1e-8
