"""GPU parity of the product's dmrg_twosite (device-resident sweep) against the oracle and the
reference's published notebook energies. Energies within 1e-10; states normalised."""

import sys

import numpy as np
import pytest

import chemtensor_amd as ct
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def statevector(mps, d):
    n = d ** mps.nsites
    out = np.zeros(n)
    assert orc.lib().orc_h_mps_vector(mps.ptr, orc.dptr(out), n) == n
    return out


def test_dmrg_notebook_fermi_hubbard_on_gpu():
    """examples/dmrg/basic_dmrg_fermi_hubbard.ipynb:197/:304-307 (seed 42) through the GPU drop-in."""
    mpo = ct.MPO.hubbard(6, 1.0, 4.0, 1.5)
    psi = ct.MPS.random(6, 4, mpo.qsite, ct.encode_pair(7, 1), 32, 42)
    en, ent = ct.dmrg(mpo, psi, num_sweeps=4, maxiter_lanczos=5, tol_split=1e-8, max_vdim=32)
    say("energies", en.tolist())
    en_ref = [-18.48431136702258, -18.484358903148298, -18.484358904136563, -18.484358904136524]
    assert np.max(np.abs(en - en_ref)) < 1e-10
    assert psi.bond_dims() == [1, 4, 16, 30, 16, 4, 1]
    ent_ref = [1.057297431102482, 1.0833060473946288, 1.1618575108707525, 1.0833062031796814, 1.0572974453589872]
    assert np.max(np.abs(ent - ent_ref)) < 1e-8
    v = statevector(psi, 4)
    assert abs(np.linalg.norm(v) - 1) < 1e-12


@pytest.mark.parametrize("model,L,chi,sweeps,maxiter,tol,seed", [
    ("xxz", 16, 32, 4, 10, 1e-12, 1),        # configuration: XXZ L=16, J=1, D=0.5, h=0, chi=32, 4 sweeps, 10 Lanczos
    ("hubbard", 10, 48, 2, 8, 1e-10, 7),
])
def test_dmrg_matches_oracle(model, L, chi, sweeps, maxiter, tol, seed):
    say("dmrg", model, L, chi)
    if model == "xxz":
        pm, om, sector = ct.MPO.xxz(L, 1.0, 0.5, 0.0), orc.MPO.xxz(L, 1.0, 0.5, 0.0), 0
    else:
        pm, om, sector = ct.MPO.hubbard(L, 1.0, 4.0, 0.0), orc.MPO.hubbard(L, 1.0, 4.0, 0.0), ct.encode_pair(L, 0)
    pp = ct.MPS.random(L, pm.d, pm.qsite, sector, chi, seed)
    op = orc.MPS.random(L, om.d, om.qsite, sector, chi, seed)
    en_g, ent_g = ct.dmrg(pm, pp, sweeps, maxiter, tol, chi)
    ret, en_o, ent_o = orc.dmrg(om, op, sweeps, maxiter, tol, chi)
    assert ret == 0
    say("gpu", en_g.tolist(), "oracle", en_o.tolist())
    assert np.max(np.abs(en_g - en_o)) < 1e-10
    assert pp.bond_dims() == op.bond_dims()
    assert np.max(np.abs(ent_g - ent_o)) < 1e-8


def test_dmrg_split_fp32_matvec_energies_within_1e10():
    """the precision option over whole sweeps (CT_DMRG_MATVEC=split: split-FP32 MFMA matvecs, FP64
    Lanczos scalars, eigensolver and split): energies within 1e-10 of the FP64 sweep"""
    import os
    def run():
        mpo = ct.MPO.hubbard(12, 1.0, 4.0, 0.0)
        psi = ct.MPS.random(12, 4, mpo.qsite, ct.encode_pair(12, 0), 48, 11)
        en, _ = ct.dmrg(mpo, psi, num_sweeps=3, maxiter_lanczos=10, tol_split=1e-10, max_vdim=48)
        return en
    en64 = run()
    os.environ["CT_DMRG_MATVEC"] = "split"
    try:
        en32 = run()
    finally:
        os.environ.pop("CT_DMRG_MATVEC", None)
    say("fp64", en64.tolist(), "split", en32.tolist())
    assert np.max(np.abs(en32 - en64)) < 1e-10
