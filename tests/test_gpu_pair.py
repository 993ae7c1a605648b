"""GPU parity of the sector-pair matvec form (CT_MATVEC_PAIR_FP64: the three contractions of every
left/right sector pair in one launch, partial blocks summed by a second launch) against the CPU
oracle's apply_local_hamiltonian (chain_ops.c:353-390).

Tolerance: relative Frobenius error <= 1e-12 (the bar of the FP64 reference path); repeated
application bit-identical (fixed summation order, no atomics)."""

import sys

import numpy as np
import pytest

import chemtensor_amd as ct
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def run(theta, h2, left, right):
    want = orc.apply_local_hamiltonian(theta, h2, left, right).serialize()
    x = ct.DeviceArray.from_numpy(theta.serialize())
    plan = ct.MatvecPlan(theta, h2, left, right, ct.PAIR_FP64)
    assert plan.dim == theta.nelem()
    y = ct.DeviceArray(plan.dim)
    plan.apply(x, y)
    plan.synchronize()
    got = y.numpy()
    err = rel(got, want)
    say("  rel err", err)
    assert err < TOL
    plan.apply(x, y)
    plan.synchronize()
    assert np.array_equal(y.numpy(), got)
    return plan, got


@pytest.mark.parametrize("model,L,chi", [("xxz", 12, 16), ("hubbard", 8, 24), ("hubbard", 32, 64), ("xxz", 32, 128), ("hubbard", 32, 256),
                                         ("hubbard", 32, 512), ("xxz", 32, 512), ("hubbard", 48, 1024)])
def test_pair_matches_oracle(model, L, chi):
    say("pair matvec", model, L, chi)
    run(*orc.matvec_problem(model, L, chi, 1234)[:4])


@pytest.mark.parametrize("cap", ["1", "8", "100000"])
def test_pair_chunking_matches_oracle(cap, monkeypatch):
    """Pairs cut into many row / column chunks (small caps) or left whole."""
    monkeypatch.setenv("CT_PAIR_CAP", cap)
    for model, L, chi in [("hubbard", 32, 256), ("xxz", 24, 128)]:
        say("pair matvec cap", cap, model, L, chi)
        run(*orc.matvec_problem(model, L, chi, 77)[:4])


def test_pair_physical_environments():
    """Environments built by contraction from DMRG states: every bond, including the chain edges
    where bond dimensions are 1."""
    mpo = orc.MPO.hubbard(8, 1.0, 4.0, 0.5)
    psi = orc.MPS.random(8, 4, mpo.qsite, orc.encode_pair(8, 0), 12, 5)
    for i in range(7):
        say("bond", i)
        run(*orc.local_problem(mpo, psi, i))


def test_pair_lanczos_matches_grouped():
    """Device Lanczos on the pair form: graph replays equal eager launches, and the ground energy
    equals the grouped form's to 1e-12."""
    theta, h2, left, right, _ = orc.matvec_problem("hubbard", 32, 256, 99)
    v0 = ct.DeviceArray.from_numpy(np.random.default_rng(3).standard_normal(theta.nelem()))
    res = {}
    for mode in (ct.GROUPED_FP64, ct.PAIR_FP64):
        plan = ct.MatvecPlan(theta, h2, left, right, mode)
        lam, u = plan.lanczos(v0, 20)
        res[mode] = (lam, u.numpy())
    assert abs(res[ct.PAIR_FP64][0] - res[ct.GROUPED_FP64][0]) < 1e-12 * abs(res[ct.GROUPED_FP64][0])
    plan = ct.MatvecPlan(theta, h2, left, right, ct.PAIR_FP64)
    import os
    vals = []
    for flag in ("1", "1", "0"):
        os.environ["CT_LANCZOS_GRAPH"] = flag
        lam, u = plan.lanczos(v0, 20)
        vals.append((lam, u.numpy()))
    os.environ.pop("CT_LANCZOS_GRAPH", None)
    for lam, u in vals[1:]:
        assert lam == vals[0][0]
        assert np.array_equal(u, vals[0][1])


@pytest.mark.parametrize("model,L,chi", [("hubbard", 32, 512), ("xxz", 24, 128), ("hubbard", 16, 64)])
def test_pair_lanczos_step_from_partials(model, L, chi, monkeypatch):
    """The Lanczos recurrence on a pair plan with the partial-block sum folded into the step kernel
    (two launches per step: units writing the alpha partials per unit, then the step) against the
    same recurrence on the matvec's own partial-sum launch (CT_PAIR_LZ_FUSED=0): ground energies
    within 1e-12 relative (only the alpha sums run in another order), Ritz vectors within 1e-7, and
    graph replays bit-identical to eager launches."""
    theta, h2, left, right = orc.matvec_problem(model, L, chi, 31)[:4]
    v0 = ct.DeviceArray.from_numpy(np.random.default_rng(5).standard_normal(theta.nelem()))
    res = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("CT_PAIR_LZ_FUSED", flag)
        plan = ct.MatvecPlan(theta, h2, left, right, ct.PAIR_FP64)
        lam, u = plan.lanczos(v0, 16)
        res[flag] = (lam, u.numpy())
        say("pair lanczos", model, L, chi, "fused step", flag, lam)
    assert abs(res["1"][0] - res["0"][0]) <= 1e-12 * abs(res["0"][0])
    u1, u0 = res["1"][1], res["0"][1]
    assert min(np.linalg.norm(u1 - u0), np.linalg.norm(u1 + u0)) < 1e-7 * np.linalg.norm(u0)
    monkeypatch.setenv("CT_PAIR_LZ_FUSED", "1")
    plan = ct.MatvecPlan(theta, h2, left, right, ct.PAIR_FP64)
    vals = []
    for flag in ("1", "1", "0"):
        monkeypatch.setenv("CT_LANCZOS_GRAPH", flag)
        lam, u = plan.lanczos(v0, 16)
        vals.append((lam, u.numpy()))
    for lam, u in vals[1:]:
        assert lam == vals[0][0]
        assert np.array_equal(u, vals[0][1])


@pytest.mark.parametrize("model,L,chi,seed", [("hubbard", 32, 128, 41), ("hubbard", 32, 256, 42), ("xxz", 20, 64, 43)])
def test_pair_packed_units_match_oracle(model, L, chi, seed, monkeypatch):
    """Units holding several packed chunks (the packing target set below the chunk count) against
    the oracle at 1e-12, bit-reproducible, for every LDS budget."""
    theta, h2, left, right = orc.matvec_problem(model, L, chi, seed)[:4]
    for lds in ("2048", "3072", "4608"):
        monkeypatch.setenv("CT_PAIR_LDS", lds)
        monkeypatch.setenv("CT_PAIR_SLOTS", "100000000")
        chunks = int(ct.matvec_pair_stats(theta, h2, left, right)["units"])
        monkeypatch.setenv("CT_PAIR_SLOTS", str(max(1, chunks * 2 // 3)))
        units = int(ct.matvec_pair_stats(theta, h2, left, right)["units"])
        say("packed pair", model, L, chi, "lds", lds, "chunks", chunks, "units", units)
        run(theta, h2, left, right)
