"""Two-site DMRG on the GPU with the spin molecular Hamiltonian MPO (the seeded-integral operator at
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


def test_dmrg_spin_molecular_matches_oracle():
    """The seeded-integral molecular operator through both sweeps: the product's spin molecular MPO (reference
    struct layout) drives the product's dmrg_twosite and the oracle's (a restatement of dmrg.c:262-399)
    from the same seeded start state. n = 8 orbitals, N = 8, 2Sz = 0, chi = 64 (reached on the
    inner bonds): energies within 1e-10, bond dimensions equal, entropies within 1e-8."""
    from oracle import oracle as orc
    n, N, chi, seed, sweeps, maxiter, tol = 8, 8, 64, 4, 3, 10, 1e-12
    t, v = ct.molecular_coefficients(n, seed)
    pm = ct.MPO.spin_molecular(t, v)
    sector = ct.encode_pair(N, 0)
    pp = ct.MPS.random(n, 4, pm.qsite, sector, chi, seed)
    op = orc.MPS.random(n, 4, pm.qsite, sector, chi, seed)
    say("molecular oracle parity n", n, "chi", chi)
    en_g, ent_g = ct.dmrg(pm, pp, sweeps, maxiter, tol, chi)
    ret, en_o, ent_o = orc.dmrg(pm, op, sweeps, maxiter, tol, chi)
    assert ret == 0
    say("gpu", en_g.tolist(), "oracle", en_o.tolist(), "bonds", pp.bond_dims())
    assert np.max(np.abs(en_g - en_o)) < 1e-10 * max(1.0, np.max(np.abs(en_o)))
    assert pp.bond_dims() == op.bond_dims()
    assert max(pp.bond_dims()) == chi
    assert np.max(np.abs(ent_g - ent_o)) < 1e-8


@pytest.mark.parametrize("n,chi", [(8, 64), (12, 128), (20, 512)])
def test_molecular_matvec_forms_match_oracle(n, chi):
    """The H_eff matvec with the spin molecular MPO (seeded integrals; at n = 20 the centre pair h2 is
    [712, 16, 16, 712] and chi = 512, the molecular workload's size) in every form that plans it, against
    the oracle's apply_local_hamiltonian on the same seeded inputs: 1e-12 relative; the product's
    generator gives the oracle generator's inputs bit for bit."""
    from oracle import oracle as orc
    t, v = ct.molecular_coefficients(n, 5)
    pm = ct.MPO.spin_molecular(t, v)
    om = orc.MPOView(pm.ptr, pm.nsites, pm.d, pm.qsite)
    theta, h2, left, right = orc.matvec_problem_mpo(om, orc.encode_pair(n, 0), chi, 2024)
    say("molecular matvec n", n, "chi", chi, "h2", list(h2.meta()[0]))
    gen = ct.generate_matvec_problem_mpo(pm, ct.encode_pair(n, 0), chi, 2024)
    for gt, ot in zip(gen, (theta, h2, left, right)):
        assert gt.nelem() == ot.nelem()
        assert np.array_equal(gt.serialize(), ot.serialize())
    want = orc.apply_local_hamiltonian(theta, h2, left, right).serialize()
    x = ct.DeviceArray.from_numpy(theta.serialize())
    modes = (ct.GROUPED_FP64, ct.GROUPED_SPLIT, ct.FUSED_FP64, ct.PAIR_FP64) + ((ct.PLAIN_FP64,) if n < 20 else ())
    for mode in modes:
        try:
            plan = ct.MatvecPlan(theta, h2, left, right, mode)
        except RuntimeError as e:
            # the single-launch and sector-pair forms bound the W_D class size (MPO bond pairs of a
            # class): plans beyond it are refused, never run
            assert mode in (ct.FUSED_FP64, ct.PAIR_FP64) and "unsupported" in str(e), (mode, str(e))
            say("  mode", mode, "refused:", e)
            continue
        y = ct.DeviceArray(plan.dim)
        plan.apply(x, y)
        plan.synchronize()
        got = y.numpy()
        err = np.max(np.abs(got - want)) / np.max(np.abs(want))
        say("  mode", mode, "rel err", err)
        assert err < 1e-12


@pytest.mark.parametrize("n,chi", [(12, 128), (20, 512)])
def test_molecular_matvec_lds_form(n, chi, monkeypatch):
    """The LDS-tiled GEMM form (64 x 64 workgroup tiles for the large tasks of the molecular
    operator) forced on and off, against the oracle at 1e-12 and against each other (the forms
    chosen without it split the k-steps of a tile differently, so the sums agree to rounding)."""
    from oracle import oracle as orc
    t, v = ct.molecular_coefficients(n, 5)
    pm = ct.MPO.spin_molecular(t, v)
    om = orc.MPOView(pm.ptr, pm.nsites, pm.d, pm.qsite)
    theta, h2, left, right = orc.matvec_problem_mpo(om, orc.encode_pair(n, 0), chi, 2024)
    want = orc.apply_local_hamiltonian(theta, h2, left, right).serialize()
    x = ct.DeviceArray.from_numpy(theta.serialize())
    ys = {}
    for lds in ("1", "0"):
        monkeypatch.setenv("CT_GEMM_LDS", lds)
        plan = ct.MatvecPlan(theta, h2, left, right, ct.GROUPED_FP64)
        forms = plan.forms()[0]
        y = ct.DeviceArray(plan.dim)
        plan.apply(x, y)
        plan.synchronize()
        ys[lds] = y.numpy()
        err = np.max(np.abs(ys[lds] - want)) / np.max(np.abs(want))
        say("molecular n", n, "CT_GEMM_LDS", lds, "forms", list(forms), "rel err", err)
        assert err < 1e-12
        if lds == "1":
            assert 6 in list(forms)
    assert np.max(np.abs(ys["1"] - ys["0"])) < 1e-13 * np.max(np.abs(ys["0"]))
