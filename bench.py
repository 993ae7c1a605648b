"""Benchmark of the two-site DMRG H_eff matvec on one MI355X.

Workload: one H_eff matvec (apply_local_hamiltonian) at the centre bond pair of a Fermi-Hubbard
chain, L = 32, t = 1, U = 4, half filling, U(1) x U(1) quantum numbers (N, 2Sz), MPO bond dim 6,
chi = 512 on both bonds; bond quantum numbers drawn as construct_random_mps draws them, theta and
the environment blocks N(0,1) from seeds (synthetic data). A step is one matvec on device-resident
inputs. The value is algorithmic TFLOP/s (2 m n k summed over the reference's GEMM calls) of the
fastest FP64 batched form (fused single launch, or three grouped launches); the plain per-block FP64
form and the split-FP32 form are timed beside it.

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
    ap.add_argument("--hbm-traffic-kib", default="",
                    help="FETCH_SIZE,WRITE_SIZE (KiB per launch) of the dominant kernel from rocprofv3 --pmc runs")
    ap.add_argument("--fetch-scale", type=float, default=2.0,
                    help="FETCH_SIZE scale from the plain-copy calibration (bench.py --calibrate-copy)")
    ap.add_argument("--forms", default="fused_fp64,grouped_fp64,split_fp32,plain_fp64",
                    help="comma list of matvec forms to time (grouped_fp64 is always included)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB ten times (scale for HBM counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus})", file=sys.stderr)
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
    for name, mode in (("fused_fp64", ct.FUSED_FP64), ("grouped_fp64", ct.GROUPED_FP64), ("split_fp32", ct.GROUPED_SPLIT), ("plain_fp64", ct.PLAIN_FP64)):
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

    head = min((k for k in ("fused_fp64", "grouped_fp64") if k in forms), key=lambda k: forms[k]["ms"])
    g = forms[head]
    plan = g["plan"]
    flops = plan.flops
    sflops = plan.stage_flops()
    ms = g["ms"]
    value = flops / (ms * 1e-3) / 1e12
    sbytes = plan.stage_bytes()
    dom = int(np.argmax(g["stage_ms"]))
    # every stage is below the FP64 ridge point (about 10 flop/byte), so HBM is its roofline
    intensity = sflops[dom] / sbytes[dom]
    achieved = sbytes[dom] / (g["stage_ms"][dom] * 1e-3) / 1e9
    traffic = None
    if args.hbm_traffic_kib:
        fetch_kib, write_kib = (float(v) for v in args.hbm_traffic_kib.split(","))
        traffic = round((fetch_kib * args.fetch_scale + write_kib) * 1024.0)
    y_g = forms["grouped_fp64"]["y"].numpy()
    if head == "fused_fp64":
        kernel_desc = "k_fused_matvec (one launch per matvec: theta.R, W.T1, L.T2 with T1/T2 in LDS)"
        form_desc = "fused FP64 (one launch per matvec, intermediates in LDS)"
    else:
        kernel_desc = f"grouped FP64 GEMM, stage {dom + 1} ({['theta.R', 'W.T1', 'L.T2'][dom]}), form {plan.forms()[0][dom]}"
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
            "workload": f"single two-site H_eff matvec, Fermi-Hubbard L={args.nsites} t=1 U=4 half filling, centre bond, chi={args.chi}"
                        if args.model == "hubbard" else f"single two-site H_eff matvec, XXZ L={args.nsites}, centre bond, chi={args.chi}",
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
            "kernel": kernel_desc,
            "kernel_algorithmic_bytes_per_launch": sbytes[dom],
            "kernel_flops_per_launch": sflops[dom],
            "kernel_flop_per_byte": round(intensity, 3),
            "kernel_ms": round(float(g["stage_ms"][dom]), 5),
            "kernel_tflops": round(sflops[dom] / (g["stage_ms"][dom] * 1e-3) / 1e12, 4),
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
                     "form_names": ["wave_simple", "wave_stream", "wg_chunk", "wg_stream"]},
    }
    ys = {k: v["y"].numpy() for k, v in forms.items()}
    for k in ys:
        if k != "grouped_fp64":
            result[f"rel_err_{k}_vs_grouped"] = float(np.linalg.norm(ys[k] - y_g) / np.linalg.norm(y_g))

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
