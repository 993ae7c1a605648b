"""Benchmark of the two-site DMRG H_eff matvec on one MI355X.

Workload: one H_eff matvec (apply_local_hamiltonian) at the centre bond pair of a Fermi-Hubbard
chain, L = 32, t = 1, U = 4, half filling, U(1) x U(1) quantum numbers (N, 2Sz), MPO bond dim 6,
chi = 512 on both bonds; bond quantum numbers drawn as construct_random_mps draws them, theta and
the environment blocks N(0,1) from seeds (synthetic data). A step is one matvec on device-resident
inputs. The value is algorithmic TFLOP/s (2 m n k summed over the reference's GEMM calls) of the
fastest FP64 batched form (fused single launch, three grouped launches, or the sector-pair unit
launch plus its partial-sum launch); the plain per-block FP64 form and the split-FP32 form are timed
beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--chi 512] [--no-cpu]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK = 78.6   # TFLOP/s, MI355X dense FP64 matrix
FP32_MFMA_PEAK = 157.3  # TFLOP/s, MI355X dense FP32 matrix


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_form(ct, plan, x, y, steps, warmup):
    """Average ms per matvec (events around K steps) and per-stage ms (events between stages)."""
    for _ in range(warmup):
        plan.apply(x, y)
    plan.synchronize()
    e0, e1 = ct.Event(), ct.Event()
    e0.record(plan.stream)
    for _ in range(steps):
        plan.apply(x, y)
    e1.record(plan.stream)
    total = e0.elapsed_ms(e1) / steps
    stages = np.zeros(3)
    for _ in range(steps):
        stages += plan.apply_timed(x, y)
    plan.synchronize()
    return total, stages / steps


def cpu_baseline(model, nsites, chi, seed, budget_s=10.0):
    """The oracle (C restatement of the reference's FP64 path, OpenMP over output blocks) on the
    host cores: matvecs of the same workload for about budget_s seconds."""
    from oracle import oracle as orc
    theta, h2, left, right, _ = orc.matvec_problem(model, nsites, chi, seed)
    orc.apply_local_hamiltonian(theta, h2, left, right)
    n, t0 = 0, time.perf_counter()
    while True:
        y = orc.apply_local_hamiltonian(theta, h2, left, right)
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 2000:
            break
    dt = (time.perf_counter() - t0) / n
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dt, n, threads, y.serialize()


def bench_batch(ct, nprob, seed, steps, warmup, lanczos_iter):
    """Batched launch: nprob independent H_eff problems (XXZ or Hubbard, L in 16..64, chi drawn from
    {64, 128, 256, 512, 1024}, per-instance seeds) applied in one fused launch, and their Lanczos
    ground states with one recurrence per problem (one launch of every kind per Lanczos step)."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(nprob):
        model = ["xxz", "hubbard"][int(rng.integers(2))]
        L = int(rng.integers(16, 65)) // 2 * 2
        chi = int(rng.choice([64, 128, 256, 512, 1024]))
        specs.append((model, L, chi, seed + 1000 * i))
    t0 = time.perf_counter()
    probs = [ct.generate_matvec_problem(*sp) for sp in specs]
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    batch = ct.MatvecBatch(probs)
    t_plan = time.perf_counter() - t0
    x = ct.DeviceArray.from_numpy(np.concatenate([p[0].serialize() for p in probs]))
    y = ct.DeviceArray(batch.dim)
    for _ in range(warmup):
        batch.apply(x, y)
    batch.synchronize()
    e0, e1 = ct.Event(), ct.Event()
    e0.record(batch.stream)
    for _ in range(steps):
        batch.apply(x, y)
    e1.record(batch.stream)
    ms = e0.elapsed_ms(e1) / steps
    # the precision option on the same batch: split-FP32 stages 1 and 3
    t0 = time.perf_counter()
    sbatch = ct.MatvecBatch(probs, mode=ct.GROUPED_SPLIT)
    t_splan = time.perf_counter() - t0
    ys = ct.DeviceArray(batch.dim)
    for _ in range(warmup):
        sbatch.apply(x, ys)
    sbatch.synchronize()
    e0.record(sbatch.stream)
    for _ in range(steps):
        sbatch.apply(x, ys)
    e1.record(sbatch.stream)
    ms_split = e0.elapsed_ms(e1) / steps
    batch.apply(x, y)
    batch.synchronize()
    y64, y32 = y.numpy(), ys.numpy()
    split_err = float(np.linalg.norm(y32 - y64) / np.linalg.norm(y64))
    log(f"batch of {nprob} split-FP32: {ms_split:.4f} ms per batched matvec (rel err vs FP64 batch {split_err:.1e}), plan {t_splan:.2f}s")
    batch.lanczos(x, lanczos_iter)
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        lam, st, _u = batch.lanczos(x, lanczos_iter)
    lz_ms = (time.perf_counter() - t0) / reps * 1e3
    log(f"batch of {nprob}: {ms:.4f} ms per batched matvec ({batch.flops / (ms * 1e-3) / 1e12:.3f} TFLOP/s), "
        f"Lanczos {lanczos_iter} steps {lz_ms:.3f} ms; generate {t_gen:.1f}s plan {t_plan:.2f}s")
    return {
        "workload": f"{nprob} independent two-site H_eff matvecs (XXZ / Fermi-Hubbard, L 16-64, chi from {{64,128,256,512,1024}}) in one batched launch (fused, or one grouped launch per stage)",
        "problems": [f"{m} L={L} chi={c}" for (m, L, c, _) in specs],
        "elements": int(batch.dim),
        "flops_per_batched_matvec": batch.flops,
        "ms_per_batched_matvec": round(ms, 5),
        "tflops": round(batch.flops / (ms * 1e-3) / 1e12, 4),
        "hbm_gbs_algorithmic": round(batch.bytes / (ms * 1e-3) / 1e9, 2),
        "lanczos_maxiter": lanczos_iter,
        "lanczos_ms_per_batched_solve": round(lz_ms, 4),
        "lanczos_ms_per_step": round(lz_ms / lanczos_iter, 5),
        "lanczos_status_nonzero": int(np.count_nonzero(st)),
        "plan_seconds": round(t_plan, 3),
        "form": "fused (one launch)" if batch.fused else "grouped (three launches)",
        "form_ms_grouped_fused": [round(v, 5) for v in batch.form_ms],
        "split_fp32": {"ms_per_batched_matvec": round(ms_split, 5), "tflops": round(batch.flops / (ms_split * 1e-3) / 1e12, 4),
                       "rel_err_vs_fp64_batch": split_err, "speedup_vs_fp64": round(ms / ms_split, 3)},
    }


def bench_dmrg_precision(ct, seed):
    """the precision option over whole sweeps: dmrg_twosite with FP64 matvecs and with split-FP32
    matvecs (CT_DMRG_MATVEC=split; Lanczos scalars, eigensolver and split stay FP64) from the same
    seeded start state; the energy difference per sweep"""
    out = []
    for model, L, chi, sweeps in (("hubbard", 16, 64, 4), ("xxz", 32, 64, 4)):
        runs = {}
        for form in ("fp64", "split"):
            if form == "split":
                os.environ["CT_DMRG_MATVEC"] = "split"
            try:
                if model == "hubbard":
                    mpo = ct.MPO.hubbard(L, 1.0, 4.0, 0.0)
                    psi = ct.MPS.random(L, 4, mpo.qsite, ct.encode_pair(L, 0), chi, seed)
                else:
                    mpo = ct.MPO.xxz(L, 1.0, 0.5, 0.0)
                    psi = ct.MPS.random(L, 2, mpo.qsite, 0, chi, seed)
                t0 = time.perf_counter()
                en, _ent = ct.dmrg(mpo, psi, num_sweeps=sweeps, maxiter_lanczos=10, tol_split=1e-10, max_vdim=chi)
                runs[form] = (en, time.perf_counter() - t0)
            finally:
                os.environ.pop("CT_DMRG_MATVEC", None)
        err = [abs(a - b) for a, b in zip(runs["split"][0], runs["fp64"][0])]
        out.append({"model": model, "nsites": L, "max_vdim": chi, "sweeps": sweeps,
                    "energies_fp64": [float(x) for x in runs["fp64"][0]], "energy_error_split_fp32_per_sweep": err,
                    "seconds_fp64": round(runs["fp64"][1], 3), "seconds_split_fp32": round(runs["split"][1], 3)})
        log(f"dmrg {model} L={L} chi={chi}: split-FP32 energy error per sweep {['%.1e' % e for e in err]}")
    return out


def bench_molecular(ct, n, chi, sweeps, seed, split):
    """Two-site DMRG on a synthetic spin molecular Hamiltonian (n spatial orbitals,
    sector N = n, 2 Sz = 0), FP64 matvecs and split-FP32 matvecs from the same seeded start state;
    energy per sweep and the split form's energy error per sweep"""
    t, v = ct.molecular_coefficients(n, seed)
    t0 = time.perf_counter()
    mpo = ct.MPO.spin_molecular(t, v)
    t_mpo = time.perf_counter() - t0
    s = mpo.s
    mdims = [int(s.a[i].dim_logical[3]) for i in range(n - 1)]
    runs = {}
    for form in (("fp64", "split") if split else ("fp64",)):
        if form == "split":
            os.environ["CT_DMRG_MATVEC"] = "split"
        try:
            psi = ct.MPS.random(n, 4, mpo.qsite, ct.encode_pair(n, 0), chi, seed)
            t0 = time.perf_counter()
            en, _ent = ct.dmrg(mpo, psi, num_sweeps=sweeps, maxiter_lanczos=10, tol_split=1e-10, max_vdim=chi)
            runs[form] = (en, time.perf_counter() - t0, max(psi.bond_dims()))
        finally:
            os.environ.pop("CT_DMRG_MATVEC", None)
    out = {"workload": f"two-site DMRG, spin molecular Hamiltonian n={n} orbitals (seeded integrals: t N(0,1)/n symmetric, "
                       f"(ij|kl) 8-fold symmetric N(0,0.1^2) made PSD as a pair matrix), sector N={n} 2Sz=0, chi={chi}, "
                       f"{sweeps} sweeps, 10 Lanczos steps",
           "mpo_max_bond": max(mdims), "mpo_seconds": round(t_mpo, 3),
           "energies_fp64": [float(x) for x in runs["fp64"][0]], "seconds_fp64": round(runs["fp64"][1], 3),
           "mps_max_bond": runs["fp64"][2]}
    log(f"molecular n={n} chi={chi}: MPO bond <= {max(mdims)} ({t_mpo:.2f} s), FP64 {sweeps} sweeps {runs['fp64'][1]:.2f} s, "
        f"E = {runs['fp64'][0][-1]:.12f}")
    if split:
        err = [abs(a - b) for a, b in zip(runs["split"][0], runs["fp64"][0])]
        out["energy_error_split_fp32_per_sweep"] = err
        out["seconds_split_fp32"] = round(runs["split"][1], 3)
        log(f"molecular split-FP32 energy error per sweep {['%.1e' % e for e in err]} ({runs['split'][1]:.2f} s)")
    return out


def split_mfma_use(plan, stage_ms):
    """The split-FP32 form's matrix-core rate: every FP64 product of the sliced stages (1 and 3) is 15
    exact FP32 slice products on v_mfma_f32_16x16x4f32; against the FP32 MFMA peak."""
    sf = plan.stage_flops()
    t = (stage_ms[0] + stage_ms[2]) * 1e-3
    tf = 15.0 * (sf[0] + sf[2]) / t / 1e12 if t > 0 else 0.0
    return {"slice_products_per_fp64_product": 15, "stages": [1, 3], "fp32_mfma_tflops_executed": round(tf, 3),
            "fp32_mfma_peak_tflops": FP32_MFMA_PEAK, "fp32_mfma_peak_frac": round(tf / FP32_MFMA_PEAK, 4)}


def bench_molecular_matvec(ct, n, chi, seed, steps, warmup):
    """The molecular workload's compute-heavy kernel: one H_eff matvec at the centre bond pair of the spin
    molecular Hamiltonian (n orbitals, h2 = [712, 16, 16, 712] at n = 20, the MPO bond between them 862),
    chi on both outer bonds, bond quantum numbers as construct_random_mps draws them in the sector
    N = n, 2Sz = 0, theta / L / R N(0,1); every FP64 form that plans it and the split-FP32 form"""
    t, v = ct.molecular_coefficients(n, seed)
    mpo = ct.MPO.spin_molecular(t, v)
    theta, h2, left, right = ct.generate_matvec_problem_mpo(mpo, ct.encode_pair(n, 0), chi, seed)
    x = ct.DeviceArray.from_numpy(theta.serialize())
    out = {"workload": f"single two-site H_eff matvec, spin molecular Hamiltonian n={n} (seeded integrals), centre bond pair, "
                       f"chi={chi}, h2 {h2.meta()[0]}",
           "theta_elements": int(theta.nelem()), "forms": {}}
    ys = {}
    for name, mode in (("grouped_fp64", ct.GROUPED_FP64), ("pair_fp64", ct.PAIR_FP64), ("fused_fp64", ct.FUSED_FP64),
                       ("split_fp32", ct.GROUPED_SPLIT)):
        try:
            plan = ct.MatvecPlan(theta, h2, left, right, mode)
        except RuntimeError as e:
            out["forms"][name] = {"refused": str(e)}
            continue
        y = ct.DeviceArray(plan.dim)
        ms, stage_ms = time_form(ct, plan, x, y, steps, warmup)
        f = {"ms_per_matvec": round(ms, 5), "tflops": round(plan.flops / (ms * 1e-3) / 1e12, 4),
             "fp64_mfma_peak_frac": round(plan.flops / (ms * 1e-3) / 1e12 / FP64_MFMA_PEAK, 4),
             "stage_ms": [round(float(v), 5) for v in stage_ms]}
        if name == "split_fp32":
            f["matrix_core_use"] = split_mfma_use(plan, stage_ms)
        out["forms"][name] = f
        out["flops_per_matvec"] = plan.flops
        ys[name] = y.numpy()
        log(f"molecular matvec n={n} chi={chi} {name}: {ms:.4f} ms ({f['tflops']} TFLOP/s)")
    for k in ys:
        if k != "grouped_fp64":
            out["forms"][k]["rel_err_vs_grouped"] = float(np.linalg.norm(ys[k] - ys["grouped_fp64"]) / np.linalg.norm(ys["grouped_fp64"]))
    return out


def bench_serve(ct, nreq, rate, L_range, chi_range, seed, budget, max_active, timeout_s=None):
    """Serving loop: nreq ground-state solve requests with Poisson arrivals at `rate` per second,
    each an XXZ chain (Delta ~ U[0,2]) or a Fermi-Hubbard chain (U ~ U[0,8]) with L and chi drawn
    uniformly from the given ranges, Lanczos 10 steps per bond, stop when a sweep changes the energy
    by less than 1e-10 (at most 12 sweeps). Latency = finish - submit."""
    import threading
    if timeout_s is None:
        timeout_s = max(240.0, 1.0 * nreq)
    rng = np.random.default_rng(seed)
    reqs, desc = [], []
    for i in range(nreq):
        L = int(rng.integers(L_range[0], L_range[1] + 1)) // 2 * 2
        chi = int(rng.integers(chi_range[0], chi_range[1] + 1))
        if rng.integers(2) == 0:
            delta = float(rng.uniform(0, 2))
            reqs.append(ct.serve_request("xxz", L, (1.0, delta, 0.0), 0, chi, num_sweeps=12, maxiter_lanczos=10,
                                         tol_split=1e-10, tol_energy=1e-10, seed=seed + i))
            desc.append(f"xxz L={L} chi={chi} D={delta:.3f}")
        else:
            u = float(rng.uniform(0, 8))
            reqs.append(ct.serve_request("hubbard", L, (1.0, u, 0.0), ct.encode_pair(L, 0), chi, num_sweeps=12,
                                         maxiter_lanczos=10, tol_split=1e-10, tol_energy=1e-10, seed=seed + i))
            desc.append(f"hubbard L={L} chi={chi} U={u:.3f}")
    gaps = rng.exponential(1.0 / rate, nreq)
    srv = ct.Server(budget, max_active)
    tickets = [None] * nreq

    def submitter():
        for i in range(nreq):
            time.sleep(gaps[i])
            tickets[i] = srv.submit(reqs[i])

    t0 = time.perf_counter()
    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    res = []
    t_note = time.perf_counter()
    for i in range(nreq):
        while True:
            r = srv.poll(tickets[i], wait=False)
            if r.status != 1 or time.perf_counter() - t0 > timeout_s:
                break
            time.sleep(0.01)
            if time.perf_counter() - t_note > 30:      # progress of long runs
                t_note = time.perf_counter()
                log(f"serve: {i}/{nreq} finished in order after {t_note - t0:.0f}s")
        res.append(r)
    makespan = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    done = [r for r in res if r.status == 0]
    lat = np.array([r.t_done - r.t_submit for r in done]) if done else np.zeros(1)
    log(f"serve: {len(done)}/{nreq} solved in {makespan:.1f}s, p50 {np.percentile(lat, 50):.2f}s p99 {np.percentile(lat, 99):.2f}s")
    return {
        "sample": (f"bounded sample: {nreq} of the 2000 requests of the serving workload (run --serve 2000 for all of them)"
                   if nreq < 2000 else "the serving workload in full (2000 requests)"),
        "workload": f"{nreq} DMRG ground-state solves, Poisson arrivals at {rate}/s, XXZ (Delta~U[0,2]) or Hubbard (U~U[0,8]), "
                    f"L~U{{{L_range[0]}..{L_range[1]}}}, chi~U{{{chi_range[0]}..{chi_range[1]}}}, stop at |dE_sweep| < 1e-10",
        "requests": nreq, "solved": len(done), "failed": int(sum(1 for r in res if r.status < 0)),
        "failures": [{"request": desc[i], "status": int(r.status)} for i, r in enumerate(res) if r.status != 0],
        "arrival_rate_per_s": rate,
        "solves_per_s": round(len(done) / makespan, 4),
        "latency_p50_s": round(float(np.percentile(lat, 50)), 4),
        "latency_p99_s": round(float(np.percentile(lat, 99)), 4),
        "latency_mean_s": round(float(lat.mean()), 4),
        "queue_wait_p50_s": round(float(np.percentile([r.t_admit - r.t_submit for r in done] or [0], 50)), 4),
        "sweeps_mean": round(float(np.mean([r.sweeps for r in done])) if done else 0, 2),
        "makespan_s": round(makespan, 3),
        "slowest": ({"request": desc[int(np.argmax([r.t_done - r.t_submit if r.status == 0 else -1 for r in res]))],
                     "latency_s": round(float(lat.max()), 4),
                     "sweeps": int(res[int(np.argmax([r.t_done - r.t_submit if r.status == 0 else -1 for r in res]))].sweeps)}
                    if done else None),
        "hbm_budget_bytes": budget, "max_active": max_active,
        "server_stats": st,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--chi", type=int, default=512)
    ap.add_argument("--nsites", type=int, default=32)
    ap.add_argument("--model", default="hubbard")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch", type=int, default=64, help="problems of the batched-launch measurement (0: skip)")
    ap.add_argument("--serve", type=int, default=16, help="requests of the serving measurement (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-L", default="16,64", help="chain length range")
    ap.add_argument("--serve-chi", default="32,512", help="bond dimension range")
    ap.add_argument("--serve-budget", type=float, default=200e9, help="HBM admission budget (bytes)")
    ap.add_argument("--lanczos-iter", type=int, default=20, help="Lanczos steps of the timed ground-state solve")
    ap.add_argument("--dmrg-precision", type=int, default=1, help="energy error of whole sweeps with split-FP32 matvecs (0: skip)")
    ap.add_argument("--molecular", type=int, default=20, help="orbitals of the molecular DMRG measurement (0: skip)")
    ap.add_argument("--molecular-chi", type=int, default=512)
    ap.add_argument("--molecular-sweeps", type=int, default=6)
    ap.add_argument("--molecular-split", type=int, default=1, help="also run the sweeps with split-FP32 matvecs")
    ap.add_argument("--molecular-matvec", type=int, default=1, help="time the molecular centre-bond matvec (n, chi of --molecular)")
    ap.add_argument("--hbm-traffic-kib", default="",
                    help="FETCH_SIZE,WRITE_SIZE (KiB per launch) of the dominant kernel from rocprofv3 --pmc runs")
    ap.add_argument("--fetch-scale", type=float, default=2.0,
                    help="FETCH_SIZE scale from the plain-copy calibration (bench.py --calibrate-copy)")
    ap.add_argument("--forms", default="fused_fp64,grouped_fp64,pair_fp64,split_fp32,plain_fp64",
                    help="comma list of matvec forms to time (grouped_fp64 is always included)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB ten times (scale for HBM counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: the library has no multi-GPU path (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    import chemtensor_amd as ct
    if not ct.gpu_available():
        print("no GPU visible: the HIP product path cannot run here", file=sys.stderr)
        sys.exit(3)

    if args.calibrate_copy:
        n = args.calibrate_copy * 1024 * 1024 // 8
        src, dst = ct.DeviceArray(n), ct.DeviceArray(n)
        for _ in range(10):
            ct.check(ct.lib().ct_device_copy(dst.ptr, src.ptr, n), "ct_device_copy")
        print(json.dumps({"calibrate_copy_bytes_read": n * 8, "calibrate_copy_bytes_written": n * 8, "launches": 10}))
        return

    t0 = time.perf_counter()
    theta, h2, left, right = ct.generate_matvec_problem(args.model, args.nsites, args.chi, args.seed)
    log(f"generated inputs in {time.perf_counter() - t0:.2f}s: |theta| = {theta.nelem()}")
    x = ct.DeviceArray.from_numpy(theta.serialize())

    forms = {}
    wanted = set(args.forms.split(",")) | {"grouped_fp64"}
    for name, mode in (("fused_fp64", ct.FUSED_FP64), ("grouped_fp64", ct.GROUPED_FP64), ("pair_fp64", ct.PAIR_FP64), ("split_fp32", ct.GROUPED_SPLIT),
                       ("plain_fp64", ct.PLAIN_FP64)):
        if name not in wanted:
            continue
        t0 = time.perf_counter()
        plan = ct.MatvecPlan(theta, h2, left, right, mode)
        y = ct.DeviceArray(plan.dim)
        steps = args.steps if mode != ct.PLAIN_FP64 else max(3, args.steps // 20)
        warm = args.warmup if mode != ct.PLAIN_FP64 else 2
        ms, stage_ms = time_form(ct, plan, x, y, steps, warm)
        forms[name] = dict(plan=plan, y=y, ms=ms, stage_ms=stage_ms, steps=steps, plan_s=time.perf_counter() - t0)
        log(f"{name}: {ms:.4f} ms/matvec, stages {np.round(stage_ms, 4).tolist()} ms, plan {forms[name]['plan_s']:.2f}s")

    head = min((k for k in ("fused_fp64", "grouped_fp64", "pair_fp64") if k in forms), key=lambda k: forms[k]["ms"])
    g = forms[head]
    plan = g["plan"]
    flops = plan.flops
    sflops = plan.stage_flops()
    ms = g["ms"]
    value = flops / (ms * 1e-3) / 1e12
    sbytes = plan.stage_bytes()
    # per-launch kernel durations: each stage launched back to back on the plan's stream
    nstage = {"fused_fp64": 1, "pair_fp64": 2}.get(head, 3)
    kernel_ms = [plan.stage_time(i, x, g["y"], 100) for i in range(nstage)]
    dom = int(np.argmax(kernel_ms))
    # every stage is below the FP64 ridge point (about 10 flop/byte), so HBM is its roofline
    intensity = sflops[dom] / sbytes[dom]
    achieved = sbytes[dom] / (kernel_ms[dom] * 1e-3) / 1e9
    traffic, traffic_source = None, None
    if args.hbm_traffic_kib:
        fetch_kib, write_kib = (float(v) for v in args.hbm_traffic_kib.split(","))
        traffic = round((fetch_kib * args.fetch_scale + write_kib) * 1024.0)
        traffic_source = "--hbm-traffic-kib"
    y_g = forms["grouped_fp64"]["y"].numpy()
    names = ["k_grouped_gemm_f64", "k_stream_gemm_f64_wave", "k_grouped_gemm_f64_wg", "k_stream_gemm_f64_wg", "k_wapply", "k_f12",
             "k_grouped_gemm_f64_lds"]
    if head == "fused_fp64":
        kernel_name = "k_fused_matvec"
    elif head == "pair_fp64":
        kernel_name = ["k_pair_units", "k_pair_reduce"][dom]
    else:
        kernel_name = names[plan.forms()[0][dom]]
    workload = (f"single two-site H_eff matvec, Fermi-Hubbard L={args.nsites} t=1 U=4 half filling, centre bond, chi={args.chi}"
                if args.model == "hubbard" else f"single two-site H_eff matvec, XXZ L={args.nsites}, centre bond, chi={args.chi}")
    if traffic is None:
        # HBM bytes per launch of this kernel on this workload from the committed rocprofv3 --pmc
        # FETCH_SIZE / WRITE_SIZE passes (scaled by the copy calibration), when they cover it
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic.json")
        if os.path.exists(path):
            with open(path) as f:
                tab = json.load(f)
            kib = tab["kib_per_launch"].get(f"{kernel_name}:{dom + 1}")
            if tab.get("workload") == workload and kib is not None:
                traffic = round((kib[0] * tab["fetch_scale"] + kib[1] * tab["write_scale"]) * 1024.0)
                traffic_source = "profiles/hbm_traffic.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; copy-calibrated)"
    if head == "fused_fp64":
        kernel_desc = "k_fused_matvec (one launch per matvec: theta.R, W.T1, L.T2 with T1/T2 in LDS)"
        form_desc = "fused FP64 (one launch per matvec, intermediates in LDS)"
    elif head == "pair_fp64":
        kernel_desc = [f"k_pair_units (theta.R, W.T1, L^T.T2 of every sector-pair chunk in LDS; {plan.stats()['tasks']} units, partial blocks out)",
                       "k_pair_reduce (sums the partial blocks of every y block in a fixed order)"][dom]
        form_desc = "sector-pair FP64 MFMA (one unit launch with T1/T2 in LDS + one partial-sum launch per matvec)"
    else:
        kernel_desc = f"{kernel_name}: grouped stage {dom + 1} ({['theta.R' if kernel_name != 'k_f12' else 'theta.R and W.T1 fused', 'W.T1', 'L.T2'][dom]})"
        form_desc = "batched grouped FP64 MFMA (3 launches per matvec)"

    result = {
        "metric": "H_eff matvec TFLOP/s @chi=512 (split-FP32 vs FP64); solves/s + p99 latency",
        "value": round(value, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded construct_random_mps bond quantum numbers, N(0,1) blocks)",
        "config": {
            "workload": workload,
            "chi": args.chi, "nsites": args.nsites, "model": args.model, "theta_elements": int(plan.dim),
            "form": form_desc,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "traffic_source": traffic_source,
            "kernel": kernel_desc,
            "kernel_algorithmic_bytes_per_launch": sbytes[dom],
            "kernel_flops_per_launch": sflops[dom],
            "kernel_flop_per_byte": round(intensity, 3),
            "kernel_ms": round(float(kernel_ms[dom]), 5),
            "kernel_tflops": round(sflops[dom] / (kernel_ms[dom] * 1e-3) / 1e12, 4),
            "kernel_ms_all_stages": [round(float(v), 5) for v in kernel_ms],
            "fp64_mfma_peak_tflops": FP64_MFMA_PEAK,
        },
        "stage_flops": sflops.tolist(),
        "stage_bytes": sbytes.tolist(),
        "flops_per_matvec_reference_count": flops,
        "flops_per_matvec_executed": float(np.sum(sflops)),
        "hbm_gbs_algorithmic": round(plan.bytes / (ms * 1e-3) / 1e9, 2),
        "forms": {k: {"ms_per_matvec": round(v["ms"], 5), "tflops": round(flops / (v["ms"] * 1e-3) / 1e12, 4),
                      "stage_ms": [round(float(s), 5) for s in v["stage_ms"]]} for k, v in forms.items()},
        "headline_form": head,
        "speedup_headline_vs_plain": round(forms["plain_fp64"]["ms"] / ms, 2) if "plain_fp64" in forms else None,
        "speedup_split_vs_plain": round(forms["plain_fp64"]["ms"] / forms["split_fp32"]["ms"], 2) if {"plain_fp64", "split_fp32"} <= set(forms) else None,
        "plan_stats": plan.stats(),
        "autotune_grouped": {"forms": forms["grouped_fp64"]["plan"].forms()[0],
                             "ms_per_form": [[round(x, 5) for x in r] for r in forms["grouped_fp64"]["plan"].forms()[1]],
                     "form_names": ["wave_simple", "wave_stream", "wg_chunk", "wg_stream", "column_stream", "fused_stages_1_2", "lds64_large_tasks"]},
    }
    if "split_fp32" in forms:
        result["forms"]["split_fp32"]["matrix_core_use"] = split_mfma_use(forms["split_fp32"]["plan"], forms["split_fp32"]["stage_ms"])
    ys = {k: v["y"].numpy() for k, v in forms.items()}
    for k in ys:
        if k != "grouped_fp64":
            result[f"rel_err_{k}_vs_grouped"] = float(np.linalg.norm(ys[k] - y_g) / np.linalg.norm(y_g))

    # Lanczos ground-state solve on the headline plan: recurrence captured in a HIP graph vs eager launches
    lz = {}
    v0 = ct.DeviceArray.from_numpy(np.random.default_rng(args.seed).standard_normal(plan.dim))
    for label, flag in (("graph", "1"), ("eager", "0")):
        os.environ["CT_LANCZOS_GRAPH"] = flag
        plan.lanczos(v0, args.lanczos_iter)
        t0 = time.perf_counter()
        reps = 20
        for _ in range(reps):
            lam, _u = plan.lanczos(v0, args.lanczos_iter)
        lz[label] = (time.perf_counter() - t0) / reps * 1e3
    os.environ.pop("CT_LANCZOS_GRAPH", None)
    result["lanczos"] = {"maxiter": args.lanczos_iter, "ms_per_solve_graph": round(lz["graph"], 4),
                         "ms_per_solve_eager": round(lz["eager"], 4), "ms_per_step_graph": round(lz["graph"] / args.lanczos_iter, 5),
                         "lambda0": lam, "form": head}
    log(f"lanczos {args.lanczos_iter} steps: graph {lz['graph']:.3f} ms, eager {lz['eager']:.3f} ms")

    if args.dmrg_precision:
        result["dmrg_precision"] = bench_dmrg_precision(ct, args.seed)

    if args.molecular > 0:
        result["molecular"] = bench_molecular(ct, args.molecular, args.molecular_chi, args.molecular_sweeps, args.seed,
                                              args.molecular_split)
        if args.molecular_matvec:
            result["molecular"]["centre_matvec"] = bench_molecular_matvec(ct, args.molecular, args.molecular_chi, args.seed,
                                                                          max(10, args.steps // 5), 3)

    if args.batch > 0:
        result["batched"] = bench_batch(ct, args.batch, args.seed, max(10, args.steps // 5), 3, args.lanczos_iter)

    if args.serve > 0:
        lr = tuple(int(v) for v in args.serve_L.split(","))
        cr = tuple(int(v) for v in args.serve_chi.split(","))
        result["serving"] = bench_serve(ct, args.serve, args.serve_rate, lr, cr, args.seed, args.serve_budget, 64)

    if not args.no_cpu:
        dt, n, threads, y_cpu = cpu_baseline(args.model, args.nsites, args.chi, args.seed)
        result["cpu_baseline"] = {
            "value": round(flops / dt / 1e12, 6), "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "sample": f"{n} oracle matvecs of the same workload ({dt * 1e3:.2f} ms each), OpenMP over output blocks",
        }
        for k, yy in ys.items():
            result["forms"][k]["rel_err_vs_cpu_fp64"] = float(np.linalg.norm(yy - y_cpu) / np.linalg.norm(y_cpu))
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
