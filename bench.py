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


WORKLOADS = ("tgv_central_128", "tgv_teno5_256", "sod_1024", "tgv_teno6_128", "tgv_central4weno5_128")


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
    """Wall time of `steps` steps (captured time steps replayed as hipGraphs) on
    a stream the context adopts, with HIP events recorded on that stream around
    them; then the same number of steps again, launched eagerly with HIP events
    around every kernel, for each kernel family's share of the GPU time."""
    import torch
    from opensbli_amd import native as nat
    hip = nat.load_hip()
    stream = torch.cuda.Stream()
    with hip.context(problem.desc) as c:
        c.set_stream(stream.cuda_stream)
        c.set_state(problem.q0)
        c.step(warmup)
        c.sync()
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(stream)
        c.step(steps)
        ev1.record(stream)
        c.sync()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        state = c.get_state()
        path = c.describe()
        c.profile_enable(True)
        c.step(steps)
        c.sync()
        prof = c.profile_read()
        c.profile_enable(False)
    return t1 - t0, prof, state, ev0.elapsed_time(ev1) / 1e3, path


def cpu_info():
    """Host CPU model and the CPUs this process may run on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, avail


def _cpu_leg(oracle_path, workload, threads, budget_s):
    """Time the oracle on a bounded sample of the workload in a child process
    with `threads` OpenMP threads bound to cores; returns its JSON record."""
    import subprocess
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg", workload, "--cpu-budget", str(budget_s)],
                         env=env, capture_output=True, text=True, check=True)
    return json.loads(out.stdout.strip().splitlines()[-1])


def cpu_leg_main(workload, budget_s):
    """Child side of _cpu_leg: run the oracle (no GPU use in this process)."""
    from opensbli_amd import native as nat
    problem = make_problem(workload, "fp64")
    lib = nat.Library(os.path.join(ROOT, "oracle", "libosb_oracle.so"))
    with lib.context(problem.desc) as c:
        c.set_state(problem.q0)
        t0 = time.perf_counter()
        c.step(1)
        per = max(time.perf_counter() - t0, 1e-6)
        n = max(1, min(problem.steps, int(budget_s / per)))
        t0 = time.perf_counter()
        c.step(n)
        t1 = time.perf_counter()
    print(json.dumps({"steps": n, "seconds": t1 - t0, "gpts_stage_per_s": problem.desc.npoints * 3.0 * n / (t1 - t0) / 1e9}))


def cpu_baseline(oracle_path, workload, problem, budget_s=15.0):
    """The oracle on the host cores: all the cores this process may use (at
    most OMP_NUM_THREADS when set), OpenMP threads bound one per core, and one
    thread, each on a bounded sample of the workload's steps."""
    model, avail = cpu_info()
    threads = min(avail, int(os.environ.get("OMP_NUM_THREADS", avail)))
    many = _cpu_leg(oracle_path, workload, threads, budget_s)
    one = _cpu_leg(oracle_path, workload, 1, budget_s)
    return {"value": many["gpts_stage_per_s"], "unit": "GPts·stage/s", "cores": threads, "kind": "port",
            "sample": f"{problem.name}: {many['steps']} of the workload's steps after 1 warm-up step, oracle OpenMP "
                      f"build, {threads} threads (OMP_PROC_BIND=close, OMP_PLACES=cores)",
            "cpu_model": model, "cpus_available": avail,
            "single_thread": {"value": one["gpts_stage_per_s"], "cores": 1,
                              "sample": f"{problem.name}: {one['steps']} steps after 1 warm-up step, 1 thread"}}


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
        "autotuned_face_resid_width": {k: tuning[k] for k in ("resid_w", "resid_w_us") if k in tuning},
    }


def tgv_ensemble_bench(steps=200, count=48, seed=20261204):
    """3-D TGV ensemble (FP64 central): each instance stepped alone, one after
    another, against the batched form (one fused stage launch per stage for every
    instance); the batched results are bitwise those of the instances alone."""
    import numpy as np
    import torch
    from opensbli_amd import native as nat, problems as P
    hip = nat.load_hip()
    probs = P.tgv_ensemble(count=count, seed=seed, steps=steps)
    work = sum(p.desc.npoints * 3.0 * steps for p in probs)
    t_plain, alone = 0.0, []
    for p in probs:
        with hip.context(p.desc) as c:
            c.set_state(p.q0)
            c.step(1)
            c.sync()
            t0 = time.perf_counter()
            c.step(steps)
            c.sync()
            t_plain += time.perf_counter() - t0
            alone.append(c.get_state())
    with hip.batch([p.desc for p in probs]) as b:
        for i, p in enumerate(probs):
            b.set_state(i, p.q0)
        b.step(1)
        b.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        t_batch = time.perf_counter() - t0
        info = b.describe()
        same = all(np.array_equal(b.get_state(i), q) for i, q in enumerate(alone))
    return {
        "workload": f"tgv_ensemble_{count}", "instances": count, "seed": seed,
        "sizes": sorted({p.desc.n[0] for p in probs}), "steps_timed": steps, "work_gpts_stage": work / 1e9,
        "plain_fp64": work / t_plain / 1e9, "batched_fp64": work / t_batch / 1e9,
        "speedup_batched_vs_plain": t_plain / t_batch, "batched_bitwise_equal_alone": same,
        "groups": [dict({k: g[k] for k in ("kind", "members")},
                        **{k: g["tuning"][k] for k in ("tile", "tile_us") if k in g.get("tuning", {})})
                   for g in info["groups"]], "solo": info["solo"],
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
    # a load one GPU sustains: the same request mix at a lower arrival rate, so
    # the time-to-solution is service time (overlapped with the other solves on
    # the device) rather than queueing; offered load = work per second of
    # arrivals, its fraction of the saturated throughput measured above
    srate, scount = 1.5, 24
    reqs = S.make_stream(count=scount, rate=srate, seed=seed + 1)
    su = S.serve(hip, reqs, int(budget_gib * 2**30), max_solves=max_solves, nstreams=4, chunk=chunk)
    offered = sum(r.work for r in reqs) / max(r.arrival for r in reqs) / 1e9
    out["sustained"] = {"requests": scount, "rate_per_s": srate, "seed": seed + 1,
                        "p50_time_to_solution_s": su.p50, "p99_time_to_solution_s": su.p99,
                        "aggregate_gpts_stage_per_s": su.throughput, "makespan_s": su.makespan,
                        "offered_gpts_stage_per_s": offered, "offered_fraction_of_saturated": offered / rep.throughput,
                        "peak_running": su.peak_running}
    return out


def reference_benchmarks(steps=20, warmup=2):
    """The reference's own performance-benchmark methods at their validation
    size (apps/performance_benchmark/{TENO6,Central4-WENO5}, 128^3, FP64):
    grid-point RK-stage updates per second over `steps` captured steps."""
    out = {}
    for name in ("tgv_teno6_128", "tgv_central4weno5_128"):
        pr = make_problem(name, "fp64")
        dt, prof, _, _, _ = time_gpu(pr, steps, warmup)
        out[name] = {"value": pr.desc.npoints * 3.0 * steps / dt / 1e9, "unit": "GPts*stage/s", "steps_timed": steps,
                     "kernel_ms_per_step": {k: v[0] / steps for k, v in prof.items() if v[1]}}
    return out


def kernel_roofline(problem, workload, precision, prof, t_events, dt_wall, steps):
    """Roofline entry of the dominant kernel family: algorithmic bytes of one RK
    stage over the grid (read q, read du, write du, write q) per launch, over
    that family's average launch duration in the timed region -- the region's
    HIP-event time on the context's stream times the family's share of the
    GPU time in the per-kernel profile pass, over its launches."""
    fam, (ms, nlaunch) = max(prof.items(), key=lambda kv: kv[1][0])
    total_ms = sum(v[0] for v in prof.values())
    share = ms / total_ms if total_ms > 0 else 1.0
    avg_s = t_events * share / max(nlaunch, 1)
    npts = problem.desc.npoints
    bytes_alg = BYTES_PER_PT_STAGE[problem.desc.ndim] * npts
    achieved = bytes_alg / avg_s / 1e9
    stage_s = dt_wall / (3.0 * steps)
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": pmc_traffic(fam, workload, precision),
            "kernel": fam, "kernel_avg_us": avg_s * 1e6, "kernel_launches": nlaunch,
            "kernel_share_of_gpu_time": share, "kernel_avg_us_profile_pass": ms / max(nlaunch, 1) * 1e3,
            "timed_region_events_s": t_events, "bytes_alg_per_launch": bytes_alg,
            "stage_achieved_GBs": bytes_alg / stage_s / 1e9}


def profile_value(name, key):
    """A `key: value` line of a committed profile summary (None if absent)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":")[1])
    return None


FP64_VALU_PEAK_TFLOPS = 78.6
FP32_VALU_PEAK_TFLOPS = 157.3      # published, with packed FP32 (off in this build)


def mixed_valu(key, stage_s):
    """FP32 and FP64 VALU rates of a mixed-precision stage from the rocprofv3
    summary: FP32 against its published peak, FP64 against its own, and their
    sum against the FP64 rate, which bounds it without packed FP32."""
    f32 = profile_value("mixed_precision_valu_pmc.md", key + "_mixed_fp32_flops_per_stage")
    f64 = profile_value("mixed_precision_valu_pmc.md", key + "_mixed_fp64_flops_per_stage")
    if not f32 or not f64:
        return None
    t32, t64 = f32 / stage_s / 1e12, f64 / stage_s / 1e12
    return {"fp32_flops_per_stage": f32, "fp64_flops_per_stage": f64,
            "fp32_TFLOPs": t32, "fp32_frac_of_157.3": t32 / FP32_VALU_PEAK_TFLOPS,
            "fp64_TFLOPs": t64, "fp64_frac_of_78.6": t64 / FP64_VALU_PEAK_TFLOPS,
            "sum_frac_of_78.6": (t32 + t64) / FP64_VALU_PEAK_TFLOPS}


def teno256_line(steps=10, warmup=2):
    """The supersonic TGV at 256^3 (TENO5, C_T = 1e-5, Roe, LLF, Sutherland
    viscous terms, CFL 0.5 from the device reduction), FP64 and mixed: GPts*stage/s,
    the HBM roofline of the stage, the FP64 VALU fraction of the FP64 stage (FP64
    flops per stage from the rocprofv3 SQ_INSTS_VALU_FLOPS_FP64 summary), and the
    mixed state's error against the FP64 run after the same steps."""
    out = {"workload": "tgv_teno5_256", "grid": [256, 256, 256], "steps_timed": steps}
    states = {}
    for prec in ("fp64", "mixed"):
        pr = make_problem("tgv_teno5_256", prec)
        dt, prof, st, t_ev, _ = time_gpu(pr, steps, warmup)
        npts = pr.desc.npoints
        stage_s = t_ev / (3.0 * steps)
        rec = {"value": npts * 3.0 * steps / dt / 1e9, "ms_per_step": dt / steps * 1e3,
               "kernel_ms_per_step": {k: v[0] / steps for k, v in prof.items() if v[1]},
               "hbm": {"bytes_alg_per_stage": BYTES_PER_PT_STAGE[3] * npts,
                       "achieved_GBs": BYTES_PER_PT_STAGE[3] * npts / stage_s / 1e9, "peak_GBs": HBM_PEAK_GBS}}
        rec["hbm"]["frac"] = rec["hbm"]["achieved_GBs"] / HBM_PEAK_GBS
        if prec == "fp64":
            flops = profile_value("tgv_teno5_256_pmc.md", "fp64_flops_per_stage")
            if flops:
                rec["valu_fp64"] = {"flops_per_stage": flops, "achieved_TFLOPs": flops / stage_s / 1e12,
                                    "peak_TFLOPs": FP64_VALU_PEAK_TFLOPS,
                                    "frac": flops / stage_s / 1e12 / FP64_VALU_PEAK_TFLOPS}
        else:
            valu = mixed_valu("teno256", stage_s)
            if valu:
                rec["valu"] = valu
        out[prec] = rec
        states[prec] = st
    st_err, pv = _rel_l2(states["mixed"], states["fp64"])
    out["mixed"]["rel_l2_vs_fp64"] = {"state": st_err, "per_variable": pv}
    return out


def mixed_headline(problem_name, steps, warmup):
    """The headline workload in mixed precision, with its error against FP64."""
    from opensbli_amd import native as nat
    pm = make_problem(problem_name, "mixed")
    dt_m, _, qm, t_ev, _ = time_gpu(pm, steps, warmup)
    pf = make_problem(problem_name, "fp64")
    hip = nat.load_hip()
    with hip.context(pf.desc) as c:
        c.set_state(pf.q0)
        c.step(warmup + steps)
        qf = c.get_state()
    st, pv = _rel_l2(qm, qf)
    out = {"value": pm.desc.npoints * 3.0 * steps / dt_m / 1e9, "rel_l2_vs_fp64": {"state": st, "per_variable": pv}}
    valu = mixed_valu({"tgv_central_128": "tgv128", "kh_1024": "kh1024"}.get(problem_name, ""), t_ev / (3.0 * steps))
    if valu:
        out["valu"] = valu
    return out


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
    ap.add_argument("--tgv-ensemble-steps", type=int, default=200)
    ap.add_argument("--cpu-leg", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the timed run, copy a 2 GiB FP64 buffer of random values 5 times (PMC byte-counter "
                         "calibration; larger than the 256 MB last-level cache)")
    args = ap.parse_args()
    if args.cpu_leg:
        cpu_leg_main(args.cpu_leg, args.cpu_budget)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "the multi-GPU decomposition is design-only in this build.", file=sys.stderr)
        sys.exit(2)

    oracle_path = ensure_built()
    if args.workload in ("kh_ensemble", "tgv_ensemble", "serving"):
        extra = {"kh_ensemble": lambda: ensemble_bench(args.ensemble_steps), "tgv_ensemble": lambda: tgv_ensemble_bench(args.tgv_ensemble_steps),
                 "serving": lambda: serving_bench(args.serving_requests)}[args.workload]()
        value = {"kh_ensemble": "batched_mixed", "tgv_ensemble": "batched_fp64",
                 "serving": "aggregate_gpts_stage_per_s"}[args.workload]
        value = extra[value]
        print(json.dumps({"metric": METRIC, "value": value, "unit": "GPts·stage/s", "n_gpus": 1,
                          "steps": extra.get("steps_timed", args.ensemble_steps), "warmup": 1, "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": None, "dtype": "f64-state/f32-flux" if args.workload == "kh_ensemble" else "f64",
                          "data": "synthetic (seeded)", "config": {"workload": extra["workload"]}, "detail": extra}),
              flush=True)
        return
    problem = make_problem(args.workload, args.precision)
    dt_wall, prof, state, t_events, path = time_gpu(problem, args.steps, args.warmup)
    import numpy as np
    if not np.all(np.isfinite(state)):
        raise SystemExit("non-finite state after the timed steps")
    npts = problem.desc.npoints
    value = npts * 3.0 * args.steps / dt_wall / 1e9

    # roofline of the dominant kernel family: algorithmic bytes of one RK stage over
    # the grid (read q, read du, write du, write q) per launch of a per-stage kernel
    roofline = kernel_roofline(problem, args.workload, args.precision, prof, t_events, dt_wall, args.steps)
    # the same kernel against the FP64 vector peak: FP64 flops per launch from
    # the rocprofv3 SQ_INSTS_VALU_FLOPS_FP64 summary of this workload
    flops = profile_value("central_fused_tgv128_pmc.md", "fp64_flops_per_launch") \
        if (args.workload, args.precision, roofline.get("kernel")) == ("tgv_central_128", "fp64", "central_fused") else None
    if flops:
        tf = flops / (roofline["kernel_avg_us"] * 1e-6) / 1e12
        roofline["valu_fp64"] = {"flops_per_launch": flops, "achieved_TFLOPs": tf, "peak_TFLOPs": FP64_VALU_PEAK_TFLOPS,
                                 "frac": tf / FP64_VALU_PEAK_TFLOPS}
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
        "kernel_path": path,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(oracle_path, args.workload, problem, args.cpu_budget)
    if not args.no_extras and args.workload == "tgv_central_128" and args.precision == "fp64":
        out["mixed_precision"] = mixed_headline(args.workload, args.steps, args.warmup)
        out["ensemble"] = ensemble_bench(args.ensemble_steps)
        out["tgv_ensemble"] = tgv_ensemble_bench(args.tgv_ensemble_steps)
        out["serving"] = serving_bench(args.serving_requests)
        out["reference_benchmarks"] = reference_benchmarks()
        out["tgv_teno5_256"] = teno256_line()
    if args.copy_probe:
        import torch
        a = torch.rand(256 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        out["copy_probe_bytes"] = a.numel() * 8
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
