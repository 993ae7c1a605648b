"""GPU tests of the serving loop: a solve gives the same energies alone and in a mix of concurrent
solves (continuous batching), requests that arrive while others run are admitted and finish, the
HBM budget limits admission, and the served energies agree with the standalone dmrg_twosite."""

import sys
import time

import numpy as np
import pytest

import chemtensor_amd as ct

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def req_xxz(L, chi, delta, seed, **kw):
    return ct.serve_request("xxz", L, (1.0, delta, 0.0), 0, chi, seed=seed, **kw)


def req_hub(L, chi, u, seed, **kw):
    return ct.serve_request("hubbard", L, (1.0, u, 0.0), ct.encode_pair(L, 0), chi, seed=seed, **kw)


MIX = [req_xxz(12, 24, 0.5, 1), req_hub(8, 32, 4.0, 2), req_xxz(16, 32, 1.3, 3), req_hub(10, 24, 2.0, 4),
       req_xxz(20, 16, 0.2, 5), req_hub(6, 16, 8.0, 6)]


def run_alone(req):
    srv = ct.Server(50e9, 8)
    t = srv.submit(req)
    r = srv.poll(t, wait=True)
    en = srv.energies(t)
    srv.close()
    return r, en


def test_serve_same_result_alone_and_in_mix():
    alone = []
    for i, q in enumerate(MIX):
        r, en = run_alone(q)
        say("alone", i, r.status, r.sweeps, en.tolist())
        assert r.status == 0
        alone.append(en)
    srv = ct.Server(50e9, 16)
    tickets = []
    for i, q in enumerate(MIX):          # staggered arrivals: later requests join running ones
        tickets.append(srv.submit(q))
        time.sleep(0.005)
    res = []
    for i, t in enumerate(tickets):
        r = srv.poll(t, wait=True)
        en = srv.energies(t)
        say("mix", i, r.status, r.sweeps, en.tolist(), "admit", r.t_admit, "done", r.t_done, "latency", r.t_done - r.t_submit)
        assert r.status == 0
        assert len(en) == len(alone[i])
        assert np.max(np.abs(en - alone[i])) <= 1e-10 * max(1.0, np.max(np.abs(alone[i])))
        res.append((r.t_admit, r.t_done))
    st = srv.stats()
    say("stats", st)
    # every later request was admitted while an earlier one was still running (the solves overlapped)
    for i in range(1, len(res)):
        assert any(res[j][0] <= res[i][0] < res[j][1] for j in range(i)), i
    assert st["max_batch"] >= 2      # work of different requests shared batched launches
    srv.close()


def test_serve_matches_standalone_dmrg():
    q = req_xxz(16, 32, 0.5, 1, num_sweeps=4, maxiter_lanczos=10, tol_split=1e-12, tol_energy=0.0)
    r, en = run_alone(q)
    mpo = ct.MPO.xxz(16, 1.0, 0.5, 0.0)
    psi = ct.MPS.random(16, 2, mpo.qsite, 0, 32, 1)
    en_ref, _ = ct.dmrg(mpo, psi, num_sweeps=4, maxiter_lanczos=10, tol_split=1e-12, max_vdim=32)
    say("served", en.tolist(), "dmrg_twosite", en_ref.tolist())
    assert r.status == 0 and len(en) == 4
    assert np.max(np.abs(en - en_ref)) < 1e-10


def test_serve_budget_limits_admission():
    # budget below two requests: the second is admitted only after the first finishes
    q = req_xxz(12, 16, 0.5, 7, num_sweeps=2, tol_energy=0.0)
    srv = ct.Server(1.0, 8)              # tiny budget: one request at a time (the first is always admitted)
    t1, t2 = srv.submit(q), srv.submit(q)
    r1, r2 = srv.poll(t1, wait=True), srv.poll(t2, wait=True)
    assert r1.status == 0 and r2.status == 0
    assert r2.t_admit >= r1.t_done - 1e-6
    assert r1.energy == r2.energy
    srv.close()


def test_serve_many_concurrent_requests():
    # 24 requests in flight at once: their splits run on all pool threads, device blocks freed on
    # one thread are reused by others (after that thread's stream is synchronised), batches hold
    # many requests; every request must finish and sampled ones must equal standalone DMRG
    rng = np.random.default_rng(11)
    reqs, spec = [], []
    for i in range(24):
        if i % 2 == 0:
            L, chi, delta = int(rng.integers(8, 17)) // 2 * 2, int(rng.integers(12, 41)), float(rng.uniform(0, 2))
            reqs.append(req_xxz(L, chi, delta, 100 + i, num_sweeps=3, maxiter_lanczos=8, tol_split=1e-12, tol_energy=0.0))
            spec.append(("xxz", L, chi, delta, 100 + i))
        else:
            L, chi, u = int(rng.integers(6, 13)) // 2 * 2, int(rng.integers(12, 41)), float(rng.uniform(0, 8))
            reqs.append(req_hub(L, chi, u, 100 + i, num_sweeps=3, maxiter_lanczos=8, tol_split=1e-12, tol_energy=0.0))
            spec.append(("hubbard", L, chi, u, 100 + i))
    srv = ct.Server(50e9, 32)
    tickets = [srv.submit(q) for q in reqs]
    res = [srv.poll(t, wait=True) for t in tickets]
    ens = [srv.energies(t) for t in tickets]
    st = srv.stats()
    srv.close()
    say("many: statuses", [r.status for r in res], "max_batch", st["max_batch"])
    assert all(r.status == 0 for r in res)
    assert all(len(e) == 3 for e in ens)
    for i in (0, 1, 6, 13):
        model, L, chi, p, seed = spec[i]
        if model == "xxz":
            mpo = ct.MPO.xxz(L, 1.0, p, 0.0)
            psi = ct.MPS.random(L, 2, mpo.qsite, 0, chi, seed)
        else:
            mpo = ct.MPO.hubbard(L, 1.0, p, 0.0)
            psi = ct.MPS.random(L, 4, mpo.qsite, ct.encode_pair(L, 0), chi, seed)
        en_ref, _ = ct.dmrg(mpo, psi, num_sweeps=3, maxiter_lanczos=8, tol_split=1e-12, max_vdim=chi)
        say("many: request", i, spec[i], ens[i].tolist(), en_ref.tolist())
        assert np.max(np.abs(ens[i] - en_ref)) < 1e-10 * max(1.0, np.max(np.abs(en_ref)))
