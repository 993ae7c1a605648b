"""GPU parity of the batched fused form: several independent H_eff problems (different models,
chain lengths and bond dimensions) in one launch, and the batched Lanczos (one recurrence per
problem), against the oracle and against the single-problem plans."""

import sys

import numpy as np
import pytest

import chemtensor_amd as ct
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
TOL = 1e-12

CASES = [("xxz", 16, 64, 1), ("hubbard", 24, 128, 2), ("xxz", 40, 256, 3), ("hubbard", 16, 32, 4),
         ("hubbard", 48, 512, 5), ("xxz", 64, 512, 6), ("hubbard", 20, 96, 7)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def problems():
    return [orc.matvec_problem(m, L, chi, seed) for (m, L, chi, seed) in CASES]


@pytest.mark.parametrize("form", ["grouped", "fused"])
def test_batched_matvec_matches_oracle(problems, form, monkeypatch):
    say("batch of", len(problems), "form", form)
    monkeypatch.setenv("CT_BATCH_FORM", form)
    batch = ct.MatvecBatch([p[:4] for p in problems])
    assert batch.fused == (form == "fused")
    off = batch.offsets
    assert list(np.diff(off)) == [p[0].nelem() for p in problems]
    x = np.concatenate([p[0].serialize() for p in problems])
    dx = ct.DeviceArray.from_numpy(x)
    dy = ct.DeviceArray(batch.dim)
    batch.apply(dx, dy)
    batch.synchronize()
    y = dy.numpy()
    for i, p in enumerate(problems):
        want = orc.apply_local_hamiltonian(*p[:4]).serialize()
        err = rel(y[off[i]:off[i + 1]], want)
        say("  problem", CASES[i], "rel err", err)
        assert err < TOL
    batch.apply(dx, dy)
    batch.synchronize()
    assert np.array_equal(dy.numpy(), y)


@pytest.mark.parametrize("form", ["grouped", "fused"])
def test_batched_lanczos_matches_single_plans(problems, form, monkeypatch):
    monkeypatch.setenv("CT_BATCH_FORM", form)
    batch = ct.MatvecBatch([p[:4] for p in problems])
    off = batch.offsets
    rng = np.random.default_rng(11)
    v0 = rng.standard_normal(batch.dim)
    maxiter = 15
    lam_g, st_g, u_g = batch.lanczos(ct.DeviceArray.from_numpy(v0), maxiter)
    monkeypatch.setenv("CT_LANCZOS_GRAPH", "0")
    lam_e, st_e, u_e = batch.lanczos(ct.DeviceArray.from_numpy(v0), maxiter)
    assert np.array_equal(st_g, np.zeros(len(problems))) and np.array_equal(st_e, st_g)
    assert np.array_equal(lam_g, lam_e)                      # graph replay == eager launches
    assert np.array_equal(u_g.numpy(), u_e.numpy())
    ug = u_g.numpy()
    for i, p in enumerate(problems):
        plan = ct.MatvecPlan(*p[:4], ct.FUSED_FP64)
        lam1, u1 = plan.lanczos(ct.DeviceArray.from_numpy(v0[off[i]:off[i + 1]]), maxiter)
        say("  problem", CASES[i], "batched", lam_g[i], "single", lam1)
        assert abs(lam_g[i] - lam1) <= 1e-12 * abs(lam1)
        a, b = ug[off[i]:off[i + 1]], u1.numpy()
        assert rel(np.sign(np.dot(a, b)) * a, b) < 1e-8


def test_batched_split_fp32_matches_oracle(problems):
    """the precision option in batches (ct_matvec_batch_create_mode, CT_MATVEC_GROUPED_SPLIT): every
    problem of the batch against the oracle at 1e-12, bit-reproducible, and the batched Lanczos on the
    split batch within 1e-10 of the FP64 batch's ground energies"""
    say("split batch of", len(problems))
    batch = ct.MatvecBatch([p[:4] for p in problems], mode=ct.GROUPED_SPLIT)
    assert not batch.fused
    off = batch.offsets
    x = np.concatenate([p[0].serialize() for p in problems])
    dx = ct.DeviceArray.from_numpy(x)
    dy = ct.DeviceArray(batch.dim)
    batch.apply(dx, dy)
    batch.synchronize()
    y = dy.numpy()
    for i, p in enumerate(problems):
        want = orc.apply_local_hamiltonian(*p[:4]).serialize()
        err = rel(y[off[i]:off[i + 1]], want)
        say("  problem", CASES[i], "rel err", err)
        assert err < TOL
    batch.apply(dx, dy)
    batch.synchronize()
    assert np.array_equal(dy.numpy(), y)
    lam32, st32, _ = batch.lanczos(dx, 20)
    lam64, st64, _ = ct.MatvecBatch([p[:4] for p in problems]).lanczos(dx, 20)
    assert np.all(st32 == 0) and np.all(st64 == 0)
    say("  lanczos split", lam32, "fp64", lam64)
    assert np.max(np.abs(np.array(lam32) - np.array(lam64)) / np.maximum(1.0, np.abs(lam64))) < 1e-10
