"""GPU checks of the Krylov drop-in eigensystem_krylov_symmetric (src/util/krylov.c:172-251) at a
Lanczos breakdown: the caller's afunc runs exactly numiter times, on normalised Krylov vectors only,
as the reference's lanczos_iteration_d (krylov.c:34-68) calls it, and the eigenvalues equal the
oracle's."""

import ctypes as C

import numpy as np
import pytest

import chemtensor_amd as ct
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,maxiter", [(5, 12), (3, 3), (8, 30)])
def test_krylov_breakdown_calls_afunc_numiter_times(n, maxiter):
    rng = np.random.default_rng(n + maxiter)
    a = rng.standard_normal((n, n))
    a = np.ascontiguousarray(a + a.T)
    vstart = rng.standard_normal(n)
    alpha = np.zeros(maxiter)
    beta = np.zeros(maxiter)
    v = np.zeros((maxiter, n))
    numiter = C.c_int(0)
    orc.lib().orc_h_lanczos_dense(n, orc.dptr(a), orc.dptr(vstart), maxiter, orc.dptr(alpha), orc.dptr(beta), orc.dptr(v), C.byref(numiter))
    calls = []

    def afunc(x):
        calls.append(x.copy())
        return a @ x

    numeig = 1
    lam, u = ct.eigensystem_krylov_symmetric(afunc, vstart, maxiter, numeig)
    assert len(calls) == numiter.value
    assert numiter.value <= max(n, 1) + 1
    for x in calls:
        assert np.all(np.isfinite(x))
        assert abs(np.linalg.norm(x) - 1.0) < 1e-12
    lo = np.zeros(numeig)
    uo = np.zeros((n, numeig))
    assert orc.lib().orc_h_eigensystem_dense(n, orc.dptr(a), orc.dptr(vstart), maxiter, numeig, orc.dptr(lo), orc.dptr(uo)) == 0
    assert abs(lam[0] - lo[0]) < 1e-12 * max(1.0, abs(lo[0]))
    assert abs(lam[0] - np.linalg.eigvalsh(a)[0]) < 1e-10 * max(1.0, abs(lo[0]))
