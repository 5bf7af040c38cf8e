#!/usr/bin/env python3
"""Throughput benchmark of the OpenSBLI explicit-solver hot path on one MI355X.

A "step" is one RK3 time step (three RK stages) over the whole grid with the
state resident in HBM.  value = grid-point RK-stage updates per second
(GPts*stage/s) = npoints * 3 * K / t.  Default workload: the 3-D viscous
Taylor-Green vortex at 128^3 (Ma=0.1, Re=1600, Pr=0.71), 4th-order central
Feiereisen split form + constant-viscosity terms, RK3-LS, FP64.  dt is the
reference TGV app's 0.003385 (apps/taylor_green_vortex/taylor_green_vortex.py:12):
5e-3 exceeds the RK3/central-4 acoustic stability limit at Ma=0.1 on 128^3.

The roofline entry uses the library's own HIP-event timing of its dominant
kernel family over the timed region (events recorded on the context's
stream); the CPU baseline runs the oracle on a bounded sample of the same
workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GPts·stage/s FP64 vs mixed-prec at stated L2 error; ensemble p50/p99 latency"
HBM_PEAK_GBS = 8000.0
# compulsory bytes per grid-point RK stage: read q, read du, write du, write q (FP64 state)
BYTES_PER_PT_STAGE = {1: 96.0, 2: 128.0, 3: 160.0}


# HBM bytes per launch of the dominant kernel measured with rocprofv3 PMC
# FETCH_SIZE/WRITE_SIZE (calibrated against a copy of known size); the
# summaries hold a line "traffic_bytes_per_launch: N"
PMC_PROFILES = {("central_fused", "tgv_central_128", "fp64"): "central_fused_tgv128_pmc.md"}


def pmc_traffic(family, workload, precision):
    name = PMC_PROFILES.get((family, workload, precision))
    if not name:
        return None
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def make_problem(name, precision):
    from opensbli_amd import native as nat, problems as P
    prec = nat.PREC_MIXED if precision == "mixed" else nat.PREC_FP64
    if name == "tgv_central_128":
        return P.tgv(n=128, precision=prec)
    if name == "tgv_teno5_256":
        return P.tgv(n=256, steps=100, scheme="teno", minf=1.25, cfl=0.5, dt=0.0, teno_ct=1e-5, precision=prec)
    if name == "sod_1024":
        return P.sod(precision=prec)
    if name == "tgv_teno6_128":
        return P.tgv_teno6_validation()
    if name == "tgv_central4weno5_128":
        return P.tgv_central4_weno5()
    if name.startswith("kh_"):
        return P.kelvin_helmholtz(n=int(name.split("_")[1]), precision=prec)
    raise SystemExit(f"unknown workload {name}")


def ensure_built():
    from opensbli_amd import native as nat
    if not os.path.exists(nat.HIP_LIBRARY):
        import subprocess
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "opensbli_amd", "csrc")])
    oracle = os.path.join(ROOT, "oracle", "libosb_oracle.so")
    if not os.path.exists(oracle):
        import subprocess
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle")])
    return oracle


def time_gpu(problem, steps, warmup):
    """Wall time of `steps` steps (captured time steps replayed as hipGraphs),
    then the same number of steps again, launched eagerly with HIP events
    around every kernel, for the per-kernel durations."""
    import torch
    from opensbli_amd import native as nat
    hip = nat.load_hip()
    with hip.context(problem.desc) as c:
        c.set_state(problem.q0)
        c.step(warmup)
        c.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        c.step(steps)
        c.sync()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        state = c.get_state()
        c.profile_enable(True)
        c.step(steps)
        c.sync()
        prof = c.profile_read()
        c.profile_enable(False)
    return t1 - t0, prof, state


def cpu_baseline(oracle_path, problem, budget_s=15.0):
    from opensbli_amd import native as nat
    lib = nat.Library(oracle_path)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    with lib.context(problem.desc) as c:
        c.set_state(problem.q0)
        t0 = time.perf_counter()
        c.step(1)
        t1 = time.perf_counter()
        per = max(t1 - t0, 1e-6)
        n = max(1, min(problem.steps, int(budget_s / per)))
        t0 = time.perf_counter()
        c.step(n)
        t1 = time.perf_counter()
    value = problem.desc.npoints * 3.0 * n / (t1 - t0) / 1e9
    return {"value": value, "unit": "GPts·stage/s", "cores": cores, "kind": "port",
            "sample": f"{problem.name}: {n} of the workload's steps after 1 warm-up step, oracle OpenMP build, {cores} threads"}


def _rel_l2(a, b):
    import numpy as np
    per_var = [float(np.linalg.norm(a[v] - b[v]) / max(np.linalg.norm(b[v]), 1e-300)) for v in range(b.shape[0])]
    return float(np.linalg.norm(a - b) / np.linalg.norm(b)), per_var


def ensemble_bench(steps, count=64, seed=20261203):
    """KH ensemble: plain FP64 (each instance alone, one after another), then
    the batched FP64 and batched mixed forms; mixed error against the FP64
    results (the FP64 path is bitwise equal to the oracle)."""
    import numpy as np
    import torch
    from opensbli_amd import native as nat, problems as P
    hip = nat.load_hip()
    probs = P.kh_ensemble(count=count, seed=seed, steps=steps, precision=nat.PREC_FP64)
    work = sum(p.desc.npoints * 3.0 * steps for p in probs)
    sizes = sorted({p.desc.n[0] for p in probs})
    # plain FP64 path: one context per instance, stepped alone
    t_plain = 0.0
    for p in probs:
        with hip.context(p.desc) as c:
            c.set_state(p.q0)
            c.step(1)
            c.sync()
            t0 = time.perf_counter()
            c.step(steps)
            c.sync()
            t_plain += time.perf_counter() - t0
    res = {}
    finals = {}
    for label, prec in (("batched_fp64", nat.PREC_FP64), ("batched_mixed", nat.PREC_MIXED)):
        descs = []
        for p in probs:
            d = nat.ProblemDesc(**{**p.desc.__dict__, "precision": prec})
            descs.append(d)
        with hip.batch(descs) as b:
            for i, p in enumerate(probs):
                b.set_state(i, p.q0)
            b.step(1)                      # autotune + graph capture + one step
            b.sync()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            b.step(steps)
            b.sync()
            t1 = time.perf_counter()
            res[label] = work / (t1 - t0) / 1e9
            res[label + "_describe"] = b.describe()
            finals[label] = [b.get_state(i) for i in range(len(probs))]
    worst_state, worst_var = 0.0, 0.0
    for a, bref in zip(finals["batched_mixed"], finals["batched_fp64"]):
        st, pv = _rel_l2(a, bref)
        worst_state = max(worst_state, st)
        worst_var = max(worst_var, max(pv))
    tuning = res["batched_mixed_describe"]["groups"][0]["tuning"]
    return {
        "workload": f"kh_ensemble_{count}", "instances": count, "sizes": sizes, "seed": seed,
        "steps_timed": steps, "work_gpts_stage": work / 1e9,
        "plain_fp64": work / t_plain / 1e9,
        "batched_fp64": res["batched_fp64"],
        "batched_mixed": res["batched_mixed"],
        "speedup_batched_mixed_vs_plain_fp64": res["batched_mixed"] * t_plain / work * 1e9,
        "mixed_rel_l2_vs_fp64_worst_instance": {"state": worst_state, "per_variable_max": worst_var},
        "autotuned_face_flux_grid": {k: v["grid"] for k, v in tuning.items() if k.startswith("face_flux")},
    }


def serving_bench(count=64, rate=50.0, seed=20261203, budget_gib=8.0, max_solves=32, chunk=25):
    from opensbli_amd import native as nat, serving as S
    hip = nat.load_hip()
    reqs = S.make_stream(count=count, rate=rate, seed=seed)
    rep = S.serve(hip, reqs, int(budget_gib * 2**30), max_solves=max_solves, nstreams=4, chunk=chunk)
    out = {"workload": f"serving_{count}", "requests": count, "rate_per_s": rate, "seed": seed,
           "hbm_budget_gib": budget_gib, "max_solves": max_solves, "chunk_steps": chunk, "policy": "fifo",
           "p50_time_to_solution_s": rep.p50, "p99_time_to_solution_s": rep.p99,
           "aggregate_gpts_stage_per_s": rep.throughput, "makespan_s": rep.makespan,
           "peak_running": rep.peak_running, "peak_device_gib": rep.peak_bytes / 2**30}
    # the same stream admitted least-work-first
    reqs = S.make_stream(count=count, rate=rate, seed=seed)
    sj = S.serve(hip, reqs, int(budget_gib * 2**30), max_solves=max_solves, nstreams=4, chunk=chunk, policy="sjf")
    out["sjf"] = {"p50_time_to_solution_s": sj.p50, "p99_time_to_solution_s": sj.p99,
                  "aggregate_gpts_stage_per_s": sj.throughput, "makespan_s": sj.makespan}
    return out


def reference_benchmarks(steps=20, warmup=2):
    """The reference's own performance-benchmark methods at their validation
    size (apps/performance_benchmark/{TENO6,Central4-WENO5}, 128^3, FP64):
    grid-point RK-stage updates per second over `steps` captured steps."""
    out = {}
    for name in ("tgv_teno6_128", "tgv_central4weno5_128"):
        pr = make_problem(name, "fp64")
        dt, prof, _ = time_gpu(pr, steps, warmup)
        out[name] = {"value": pr.desc.npoints * 3.0 * steps / dt / 1e9, "unit": "GPts*stage/s", "steps_timed": steps,
                     "kernel_ms_per_step": {k: v[0] / steps for k, v in prof.items() if v[1]}}
    return out


def mixed_headline(problem_name, steps, warmup):
    """The headline workload in mixed precision, with its error against FP64."""
    from opensbli_amd import native as nat
    pm = make_problem(problem_name, "mixed")
    dt_m, _, qm = time_gpu(pm, steps, warmup)
    pf = make_problem(problem_name, "fp64")
    hip = nat.load_hip()
    with hip.context(pf.desc) as c:
        c.set_state(pf.q0)
        c.step(warmup + steps)
        qf = c.get_state()
    st, pv = _rel_l2(qm, qf)
    return {"value": pm.desc.npoints * 3.0 * steps / dt_m / 1e9, "rel_l2_vs_fp64": {"state": st, "per_variable": pv}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="tgv_central_128")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the mixed-precision, KH-ensemble and serving lines added to the default run")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--serving-requests", type=int, default=64)
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the timed run, copy a 256 MiB FP64 buffer 5 times (PMC byte-counter calibration)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "the multi-GPU decomposition is design-only in this build.", file=sys.stderr)
        sys.exit(2)

    oracle_path = ensure_built()
    if args.workload in ("kh_ensemble", "serving"):
        extra = ensemble_bench(args.ensemble_steps) if args.workload == "kh_ensemble" else \
            serving_bench(args.serving_requests)
        value = extra["batched_mixed"] if args.workload == "kh_ensemble" else extra["aggregate_gpts_stage_per_s"]
        print(json.dumps({"metric": METRIC, "value": value, "unit": "GPts·stage/s", "n_gpus": 1,
                          "steps": args.ensemble_steps, "warmup": 1, "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": None, "dtype": "f64-state/f32-flux" if args.workload == "kh_ensemble" else "f64",
                          "data": "synthetic (seeded)", "config": {"workload": extra["workload"]}, "detail": extra}),
              flush=True)
        return
    problem = make_problem(args.workload, args.precision)
    dt_wall, prof, state = time_gpu(problem, args.steps, args.warmup)
    import numpy as np
    if not np.all(np.isfinite(state)):
        raise SystemExit("non-finite state after the timed steps")
    npts = problem.desc.npoints
    value = npts * 3.0 * args.steps / dt_wall / 1e9

    # roofline of the dominant kernel family: algorithmic bytes of one RK stage over
    # the grid (read q, read du, write du, write q) per launch of a per-stage kernel
    fam, (ms, nlaunch) = max(prof.items(), key=lambda kv: kv[1][0])
    avg_s = ms / 1e3 / max(nlaunch, 1)
    ndim = problem.desc.ndim
    bytes_alg = BYTES_PER_PT_STAGE[ndim] * npts
    achieved = bytes_alg / avg_s / 1e9
    stage_s = dt_wall / (3.0 * args.steps)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": pmc_traffic(fam, args.workload, args.precision),
                "kernel": fam, "kernel_avg_us": avg_s * 1e6, "kernel_launches": nlaunch,
                "bytes_alg_per_launch": bytes_alg,
                "stage_achieved_GBs": bytes_alg / stage_s / 1e9}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "GPts·stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt_wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64-state/f32-flux",
        "data": "synthetic (analytic initial field)",
        "config": {"workload": problem.name, "grid": list(problem.desc.n), "precision": args.precision,
                   "scheme": args.workload.split("_")[1], "steps_timed": args.steps},
        "roofline": roofline,
        "kernel_ms": {k: v[0] for k, v in prof.items() if v[1]},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(oracle_path, problem, args.cpu_budget)
    if not args.no_extras and args.workload == "tgv_central_128" and args.precision == "fp64":
        out["mixed_precision"] = mixed_headline(args.workload, args.steps, args.warmup)
        out["ensemble"] = ensemble_bench(args.ensemble_steps)
        out["serving"] = serving_bench(args.serving_requests)
        out["reference_benchmarks"] = reference_benchmarks()
    if args.copy_probe:
        import torch
        a = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        out["copy_probe_bytes"] = a.numel() * 8
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
