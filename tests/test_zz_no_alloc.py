"""ct_matvec_apply_noalloc (the sector-pair matvec with its partial blocks in a file-scope device
array instead of a library allocation) against ct_matvec_apply on the same plan: bit-identical
results on two calls in a row, the second with a smaller problem, both within 1e-12 of the oracle's
apply_local_hamiltonian (chain_ops.c:353-390); non-pair plans are refused."""

import numpy as np
import pytest

import chemtensor_amd as ct
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def test_noalloc_matches_apply_twice_shrinking():
    rng = np.random.default_rng(5)
    for model, L, chi, seed in [("hubbard", 32, 512, 71), ("hubbard", 16, 64, 72)]:
        print("noalloc", model, L, chi, flush=True)
        theta, h2, left, right, _ = orc.matvec_problem(model, L, chi, seed)
        want = orc.apply_local_hamiltonian(theta, h2, left, right).serialize()
        plan = ct.MatvecPlan(theta, h2, left, right, ct.PAIR_FP64)
        x = ct.DeviceArray.from_numpy(theta.serialize())
        y_ref = ct.DeviceArray.from_numpy(rng.standard_normal(plan.dim))
        y_na = ct.DeviceArray.from_numpy(rng.standard_normal(plan.dim))
        plan.apply(x, y_ref)
        assert ct.lib().ct_matvec_apply_noalloc(plan.h, x.ptr, y_na.ptr) == 0
        plan.synchronize()
        a, b = y_ref.numpy(), y_na.numpy()
        assert np.array_equal(a, b)
        assert np.linalg.norm(b - want) <= 1e-12 * np.linalg.norm(want)


def test_noalloc_refuses_other_forms():
    theta, h2, left, right, _ = orc.matvec_problem("xxz", 12, 16, 3)
    plan = ct.MatvecPlan(theta, h2, left, right, ct.GROUPED_FP64)
    x = ct.DeviceArray.from_numpy(theta.serialize())
    y = ct.DeviceArray(plan.dim)
    assert ct.lib().ct_matvec_apply_noalloc(plan.h, x.ptr, y.ptr) == ct.CT_ERR_UNSUPPORTED
