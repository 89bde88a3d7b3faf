#!/usr/bin/env python3
"""Benchmark of the CATMIP inner loop on one MI355X.

Workload: CDM inversion, 5,000 LOS points on a 100 x 50 lattice over 30 x 30
km, 16,384 chains, 10 parameters, full exponential data covariance (sill
64 mm^2, range 5 km), FP32 forward with FP64 near-singular sides, FP32 MFMA
whitening.  A step is one Metropolis step of every chain at a fixed
temperature: Philox proposal, support and free-surface verify, prior, CDM
forward + residual, whitening GEMM + reduction, accept/reject.  The sample
state is resident on the GPU when the timed region starts.

Prints one JSON line: likelihood evaluations per second (value), accepted
samples per second, the roofline of the dominant kernel (HIP-event durations
on its launch stream inside the timed region) and the FP64 CPU oracle timed
on a bounded sample of chains on one host core.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

# MI355X published peaks
PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0

# Algorithmic FP32 operations per (chain, observation) of the CDM forward +
# residual, counted from the source expressions of altar_amd/csrc/cdm_math.hpp
# and sources.hip: add, subtract, multiply = 1 each (a fused multiply-add = 2),
# reciprocal, reciprocal square root, logarithm and atan2 = 1 each.  12 vertices
# x 10 + 12 sides x 177 + 10 for the LOS projection and residual; 72 of them
# are reciprocals, reciprocal square roots, logarithms or atan2.
CDM_FLOPS_PER_PAIR = 2254

# HBM traffic per launch measured by rocprofv3 (FETCH_SIZE, WRITE_SIZE) on the
# default workload; lines "kernel <name> fetch_bytes <f> write_bytes <w>"
TRAFFIC_FILE = os.path.join(HERE, "profiles", "hbm_traffic_cdm5k_16384.txt")


def measured_traffic(kernel):
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                w = line.split()
                if len(w) >= 6 and w[0] == "kernel" and w[1] == kernel:
                    return int(w[3]) + int(w[5])
    except OSError:
        pass
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--chains", type=int, default=16384)
    ap.add_argument("--obs", type=int, default=5000)
    ap.add_argument("--workload", default="cdm", choices=["cdm", "linear", "mogi", "joint"],
                    help="cdm: the headline case; linear: 4,000 x 400 static slip, full Cd; "
                         "mogi: 1,000-point point source, diagonal Cd, annealed to beta = 1; "
                         "joint: Mogi + CDM, 2 x 20,000 InSAR + 600 GNSS, block-diagonal Cd, 65,536 chains")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16r", "bf16"])
    ap.add_argument("--fp64-sides", type=float, default=0.2,
                    help="CDM sides with |sin(beta)| below this are evaluated in FP64")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-da", action="store_true", help="skip the surrogate delayed-acceptance leg")
    ap.add_argument("--no-other", action="store_true",
                    help="skip the precision variants of the headline step and the other workloads")
    ap.add_argument("--joint-da", action="store_true",
                    help="joint workload: also train the surrogate and time delayed acceptance against plain steps")
    ap.add_argument("--da-train", type=int, default=200_000, help="training fields for the surrogate")
    ap.add_argument("--da-stage-steps", type=int, default=13,
                    help="Metropolis steps per annealing stage before the timed DA stage (training fields are "
                         "collected while beta < 0.1: 13 steps of 16,384 chains fill the 200,000-field store)")
    ap.add_argument("--da-epochs", type=int, default=1000, help="training epochs of the surrogate network")
    ap.add_argument("--da-substeps", default="128",
                    help="surrogate sub-steps per exact step in the trajectory leg, or a comma-separated "
                         "list (the first is the reported one; 1 disables the leg)")
    ap.add_argument("--da-substep-scale", type=float, default=1.6,
                    help="sub-step proposal scale relative to the Metropolis proposal")
    ap.add_argument("--da-reference-steps", type=int, default=6000,
                    help="steps of the surrogate-free plain Metropolis reference run (moments, autocorrelation); "
                         "0 skips the mixing and moment study")
    ap.add_argument("--da-reference-burn", type=int, default=1000,
                    help="burn-in steps of the plain reference run")
    ap.add_argument("--da-mixing-steps", default="3000,800,800",
                    help="steps of the two-stage, the trajectory and the biased-surrogate trajectory runs of the "
                         "mixing study")
    ap.add_argument("--da-substep-target", type=float, default=0.3,
                    help="sub-step acceptance rate the pilot trajectories tune the sub-step scale to (0: fixed)")
    ap.add_argument("--trajectory-kernel", type=int, default=0, metavar="SUBSTEPS",
                    help="only time altar_da_trajectory_fused on the CDM chains (SUBSTEPS sub-steps, network "
                         "with its initial weights), for counter runs")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration)")
    return ap.parse_args()


def copy_calibration(mib):
    """A plain copy of known size: FETCH_SIZE / WRITE_SIZE of this dispatch
    calibrate the HBM traffic counters of the kernels."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_mib": mib, "bytes_read": n * 4, "bytes_written": n * 4}))


def trajectory_kernel(args, subs, reps=3):
    """The fused trajectory kernel alone: CDM chains at the first annealing
    temperature, the 10 -> 256 x 3 -> 64 network with its initial weights
    and a fixed random PCA scaling (the kernel's cost does not depend on the
    weights).  HIP-event time per launch on the launch stream."""
    import torch
    from altar_amd import Job, _lib
    from altar_amd.bayesian import Annealer, COV, Metropolis
    from altar_amd.cases import cdm_case
    from altar_amd.cdm import CDM
    from altar_amd.surrogate import Surrogate
    torch.cuda.init()
    C = args.chains
    case = cdm_case(N=args.obs)
    job = Job(chains=C, steps=1, seed=5)
    model = CDM(case["psets"], case["x"], case["y"], case["los"], case["d"], case["oid"], case["cd"], nu=0.25,
                job=job).initialize()
    ann = Annealer(Metropolis(), COV(), seed=5).initialize(job)
    ann.model = model
    ann.worker.start(ann)
    ann.worker.cool(ann)
    step = ann.worker.step
    ann.sampler.prepareSamplingPDF(ann, step)
    sur = Surrogate(model, K=64, capacity=1024, max_batch=C, seed=2)
    dev = step.theta.device
    rng = np.random.default_rng(3)
    e = torch.from_numpy(rng.normal(0.0, 1.0, 64)).to(dev)
    s_ = torch.from_numpy(rng.uniform(0.05, 0.2, 64)).to(dev)
    P = step.parameters
    f64 = lambda *sh: torch.empty(sh, dtype=torch.float64, device=dev)
    post_s, phi, pphi, postphi = f64(C), f64(C, P), f64(C), f64(C)
    index = torch.empty(2 * C, dtype=torch.int32, device=dev)
    cnt = torch.zeros(3, dtype=torch.int64, device=dev)
    L = (ann.sampler.sigma_chol * 1.6).contiguous()
    p = _lib.ptr

    def launch(k):
        _lib.call("altar_da_trajectory_fused", sur.mlp.handle, model._prior.handle, model.source.handle,
                  p(step.theta), p(step.prior), p(post_s), p(phi), p(pphi), p(postphi), p(L), p(e), p(s_), 64, 3.0,
                  float(model.normalization), float(step.beta), C, P, 5, 1, k, subs, p(index), p(cnt),
                  _lib.stream())
    launch(0)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for k in range(reps):
        launch(k + 1)
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / reps
    flops = 2.0 * (P * 256 + 2 * 256 * 256 + 256 * 64) * C * (subs + 1)
    print(json.dumps({"trajectory_kernel": {"chains": C, "substeps": subs, "ms_per_launch": ms,
                                            "us_per_substep": 1e3 * ms / subs,
                                            "network_tflops": flops / (ms * 1e-3) / 1e12,
                                            "substep_counts": cnt.tolist()}}))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


_BASELINE_DIR = None


def cpu_baseline(case, theta, seconds, omp=False):
    """FP64 oracle (oracle/oracle.c): CDM residuals + whitened L2 on a bounded
    sample of the chains; one thread, or OpenMP over chains.  Both builds are
    compiled here, on the host that runs them, with gcc -O3 -march=native."""
    global _BASELINE_DIR
    import tempfile
    from oracle import oracle as o
    from altar_amd import norms
    if _BASELINE_DIR is None:
        _BASELINE_DIR = tempfile.mkdtemp(prefix="altar_oracle_")
        o.baseline_build(_BASELINE_DIR)
    variant = "native_omp" if omp else "native"
    layout = (0, 2, 3, 4, 7, -1)
    W = o.exact_whitener(case["cd"])
    norm = norms.normalization(case["cd"])
    args = (case["x"], case["y"], case["los"], case["d"], case["oid"], 0.25)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1) if omp else 1

    def run(n):
        t0 = time.perf_counter()
        r = o.cdm_residuals(theta[:n], layout, *args, omp=variant)
        o.l2_llk(r, W, norm, omp=variant)
        return time.perf_counter() - t0

    run(threads)                       # warm-up (thread pool, caches)
    n1 = 8 * threads if omp else 2     # whole chunks of 4 chains for every thread
    t1 = run(n1)
    n = int(max(1, min(len(theta), seconds / max(t1, 1e-6) * n1)))
    t = run(n)
    return {"value": n / t, "unit": "likelihood evals/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "build": "gcc -O3 -march=native -fno-fast-math",
            "sample": f"{n} chains x {len(case['x'])} obs, FP64 CDM + full-Cd whitening, "
                      f"{threads} thread{'s' if threads > 1 else ''}, {t:.1f} s"}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return 0
    if args.trajectory_kernel:
        trajectory_kernel(args, args.trajectory_kernel)
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(chain-parallel sharding is a design, not built)"}))
        return 2
    if args.workload != "cdm":
        print(json.dumps(side_workload(args)), flush=True)
        return 0
    import torch
    from altar_amd import Job, _lib
    from altar_amd.bayesian import Annealer, COV, Metropolis
    from altar_amd.cases import cdm_case
    from altar_amd.cdm import CDM

    torch.cuda.init()
    dev = torch.device("cuda:0")
    C, N = args.chains, args.obs
    t_setup = time.perf_counter()
    case = cdm_case(N=N)
    job = Job(chains=C, steps=1, seed=777 + 2)
    model = CDM(case["psets"], case["x"], case["y"], case["los"], case["d"], case["oid"], case["cd"],
                nu=0.25, precision=args.precision, job=job, fp64_sides=args.fp64_sides)
    model.initialize()
    annealer = Annealer(Metropolis(), COV(), seed=job.seed).initialize(job)
    annealer.model = model
    worker = annealer.worker
    worker.start(annealer)
    worker.cool(annealer)           # first beta stage: beta, Sigma, resampling
    step = worker.step
    if not step.beta > 0.0:
        raise RuntimeError(f"the first annealing stage did not advance (beta = {step.beta})")
    sampler = annealer.sampler
    sampler.prepareSamplingPDF(annealer, step)
    sampler.steps = args.warmup
    sampler.walkChains(annealer, step)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t_setup

    src = model.source
    _lib.call("altar_source_timing", src.handle, 1)
    torch.cuda.synchronize()
    sampler.steps = args.steps
    t0 = time.perf_counter()
    acc, rej, unl = sampler.walkChains(annealer, step)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    _lib.call("altar_source_timing", src.handle, 0)
    import ctypes
    ms = (ctypes.c_double * 4)()
    cnt = (ctypes.c_longlong * 4)()
    _lib.call("altar_source_timers", src.handle, ms, cnt, 4)

    f32s, f64s, skipped, invalid = src.side_census(C)
    valid = max(C - invalid, 1)
    sides = {"fp64_threshold": model.fp64_sides, "fp32_per_chain": f32s / valid,
             "fp64_per_chain": f64s / valid, "invalid_chains": invalid}

    evals = C * args.steps
    value = evals / elapsed
    fwd_ms = ms[0] / max(cnt[0], 1)
    gemm_ms = ms[1] / max(cnt[1], 1)
    Np = ((N + 127) // 128) * 128
    whiten_flops = C * N * (N + 1)          # lower-triangular |W r|^2: N(N+1) per chain
    fwd_flops = C * N * CDM_FLOPS_PER_PAIR
    kernels = {
        "cdm_forward": {"avg_ms": fwd_ms, "tflops": fwd_flops / (fwd_ms * 1e-3) / 1e12 if fwd_ms else None,
                        "bound": "valu", "peak": PEAK_FP32_TFLOPS},
        "whitening_gemm": {"avg_ms": gemm_ms,
                           "tflops": whiten_flops / (gemm_ms * 1e-3) / 1e12 if gemm_ms else None,
                           "bound": "mfma",
                           "peak": PEAK_FP32_TFLOPS if args.precision == "fp32" else PEAK_BF16_TFLOPS},
    }
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"])
    kd = kernels[dom]
    # HBM bytes per launch of the dominant kernel, from the committed rocprofv3
    # counter summary of this workload (FP32 default precision only)
    traffic = measured_traffic(dom) if args.precision == "fp32" and C == 16384 and N == 5000 else None
    roofline = {"bound": kd["bound"], "achieved": kd["tflops"], "peak": kd["peak"], "unit": "TFLOP/s",
                "frac": kd["tflops"] / kd["peak"] if kd["tflops"] else None, "traffic": traffic,
                "kernel": dom}
    out = {
        "metric": "likelihood evals/s and accepted samples/s, CDM 5k obs: plain vs surrogate-DA",
        "value": value,
        "unit": "likelihood evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision if args.precision == "fp32" else args.precision + "+fp32acc",
        "data": "synthetic (CDM lattice case, exponential Cd, seeded)",
        "config": {"workload": "CDM plain Metropolis step, 16,384 chains x 5,000 LOS, full Cd",
                   "chains": C, "observations": N, "parameters": model.parameters,
                   "beta": step.beta, "precision": args.precision, "parallelism": "single GPU"},
        "accepted_samples_per_s": acc / elapsed,
        "acceptance": {"accepted": acc, "rejected": rej, "unlikely": unl},
        "kernels": kernels,
        "cdm_sides": sides,
        "roofline": roofline,
        "setup_s": setup_s,
    }
    if not args.no_cpu:
        theta = step.theta.cpu().numpy()
        out["cpu_baseline"] = cpu_baseline(case, theta, args.cpu_seconds)
        out["cpu_baseline_all_cores"] = cpu_baseline(case, theta, args.cpu_seconds / 2, omp=True)
    # the headline CDM step at the other precisions, on the same chains
    variants = {}
    if not args.no_other:
        for label, prec, thr in (("bf16r", "bf16r", args.fp64_sides),
                                 ("fp32_every_side_fp64", "fp32", 2.0),
                                 ("fp64_reference_precision", "fp64", 2.0)):
            variants[label] = cdm_variant_timing(annealer, model, step, prec, thr, args.steps)
        model.precision = args.precision
        model.source.fp64_sides(args.fp64_sides)
        model.fp64_sides = args.fp64_sides
    if variants:
        out["cdm_variants"] = variants
    if not args.no_da:
        out["surrogate_da"] = da_leg(args, model, C, stage_steps=args.da_stage_steps)
    if not args.no_other:
        out["other_workloads"] = other_workloads(args)
    print(json.dumps(out), flush=True)
    return 0


def cdm_variant_timing(annealer, model, step, precision, fp64_sides, steps):
    """The headline step with another whitening precision or FP64-side
    threshold: Metropolis steps on a copy of the stage's chains, with the
    forward and whitening kernels timed by HIP events on their stream."""
    import ctypes
    import torch
    from altar_amd import _lib
    model.precision = precision
    model.source.fp64_sides(fp64_sides)
    model.fp64_sides = fp64_sides
    st = step.clone()
    sampler = annealer.sampler
    keep = sampler.steps
    sampler.steps = 2
    sampler.walkChains(annealer, st)           # warm-up (workspaces of this precision)
    src = model.source
    _lib.call("altar_source_timing", src.handle, 1)
    sampler.steps = steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    acc, rej, unl = sampler.walkChains(annealer, st)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    _lib.call("altar_source_timing", src.handle, 0)
    sampler.steps = keep
    ms = (ctypes.c_double * 4)()
    cnt = (ctypes.c_longlong * 4)()
    _lib.call("altar_source_timers", src.handle, ms, cnt, 4)
    C, N = st.theta.shape[0], model.observations
    fwd, gemm = ms[0] / max(cnt[0], 1), ms[1] / max(cnt[1], 1)
    f32s, f64s, skipped, invalid = src.side_census(C)
    return {"precision": precision, "fp64_sides_threshold": fp64_sides, "steps": steps,
            "ms_per_step": elapsed / steps * 1e3, "likelihood_evals_per_s": C * steps / elapsed,
            "accepted_samples_per_s": acc / elapsed, "cdm_forward_ms": fwd,
            "cdm_forward_tflops": C * N * CDM_FLOPS_PER_PAIR / (fwd * 1e-3) / 1e12,
            "whitening_ms": gemm, "whitening_tflops_algorithmic": C * N * (N + 1) / (gemm * 1e-3) / 1e12,
            "whitening": {"fp32": "gemm_nt_f32 (MFMA)", "bf16r": "gemm_nt_bf16x3 (MFMA)",
                          "fp64": "gemm_nt_f64_sumsq (FP64 MFMA)"}.get(precision),
            "fp64_sides_per_chain": f64s / max(C - invalid, 1)}


def other_workloads(args):
    """The other workloads of the build, each on its own handles with a few
    timed steps (side_workload): linear FP32 and BF16 + refinement, Mogi
    (with its whole anneal), and the joint Mogi + CDM case at full size."""
    import copy
    res = {}
    for label, wl, prec, steps in (("linear_fp32", "linear", "fp32", 10), ("linear_bf16r", "linear", "bf16r", 10),
                                   ("mogi_fp64", "mogi", "fp32", 10), ("joint_fp32", "joint", "fp32", 3)):
        a = copy.copy(args)
        a.workload, a.precision, a.steps, a.warmup, a.joint_da = wl, prec, steps, 1, False
        a.chains = 16384
        t0 = time.perf_counter()
        r = side_workload(a)
        r["wall_s"] = time.perf_counter() - t0
        for k in ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "warmup"):
            r.pop(k, None)
        res[label] = r
    return res


def side_workload(args):
    """The other two cases, same timing rules (sample state resident, K timed
    Metropolis steps at the first annealing temperature):
    * linear: G 4,000 x 400, full exponential Cd, 16,384 chains; the
      likelihood is one pre-whitened GEMM |W G theta - W d|^2 (FP32 or split
      BF16 + refinement);
    * mogi: 1,000 points, diagonal Cd, 1,024 chains, FP64; also the wall time
      of the whole anneal to beta = 1 (5 steps per stage)."""
    import torch
    from altar_amd import Job, _lib
    from altar_amd.bayesian import Annealer, COV, Metropolis
    from altar_amd import cases

    torch.cuda.init()
    if args.workload == "linear":
        from altar_amd.linear import Linear
        C = args.chains
        case = cases.linear_case()
        job = Job(chains=C, steps=1, seed=777 + 1)
        model = Linear(case["psets"], case["G"], case["d"], case["cd"], precision=args.precision, job=job)
        N, P = case["G"].shape
        flops_per_eval = 2 * N * P
        name = f"linear static slip, {C:,} chains x {N:,} obs x {P} params, full Cd (pre-whitened GEMM)"
    elif args.workload == "joint":
        from altar_amd.joint import Joint
        C = 65536 if args.chains == 16384 else args.chains
        case = cases.joint_case()
        job = Job(chains=C, steps=1, seed=777 + 4)
        model = Joint(case["psets"], case["datasets"], precision=args.precision, job=job,
                      fp64_sides=args.fp64_sides)
        N, P = sum(len(d.x) for d in case["datasets"]), 17
        # lower-triangular whitening of each block: n_b (n_b + 1) flops per chain
        flops_per_eval = sum(len(d.x) * (len(d.x) + 1) for d in case["datasets"])
        name = (f"joint Mogi + CDM, {C:,} chains x {N:,} obs (2 InSAR tracks + GNSS), 17 params, "
                f"block-diagonal Cd")
    else:
        from altar_amd.mogi import Mogi
        C = 1024
        case = cases.mogi_case()
        job = Job(chains=C, steps=5, seed=777 + 0)
        model = Mogi(case["psets"], case["x"], case["y"], case["los"], case["d"], case["oid"], case["cd"],
                     precision="fp64", job=job)
        N, P = len(case["x"]), 5
        flops_per_eval = None
        name = f"mogi point source, {C:,} chains x {N:,} obs, diagonal Cd, FP64"
    model.initialize()
    ann = Annealer(Metropolis(), COV(), seed=job.seed).initialize(job)
    ann.model = model
    w = ann.worker
    t_setup = time.perf_counter()
    w.start(ann)
    w.cool(ann)
    step = w.step
    if not step.beta > 0.0:
        raise RuntimeError(f"the first annealing stage did not advance (beta = {step.beta})")
    sampler = ann.sampler
    sampler.prepareSamplingPDF(ann, step)
    sampler.steps = args.warmup
    sampler.walkChains(ann, step)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t_setup
    src = model.source
    _lib.call("altar_source_timing", src.handle, 1)
    sampler.steps = args.steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    acc, rej, unl = sampler.walkChains(ann, step)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    _lib.call("altar_source_timing", src.handle, 0)
    import ctypes
    ms = (ctypes.c_double * 4)()
    cnt = (ctypes.c_longlong * 4)()
    _lib.call("altar_source_timers", src.handle, ms, cnt, 4)
    out = {"metric": "likelihood evals/s", "value": C * args.steps / elapsed, "unit": "likelihood evals/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp64" if args.workload == "mogi" else args.precision,
           "data": "synthetic (seeded)", "config": {"workload": name, "beta": step.beta},
           "accepted_samples_per_s": acc / elapsed, "acceptance": {"accepted": acc, "rejected": rej,
                                                                  "unlikely": unl},
           "setup_s": setup}
    if cnt[1]:
        g = ms[1] / cnt[1]
        out["kernels"] = {"gemm": {"avg_ms": g, "tflops": C * flops_per_eval / (g * 1e-3) / 1e12
                                   if flops_per_eval else None}}
    if cnt[0]:
        f = ms[0] / cnt[0]
        out.setdefault("kernels", {})["forward"] = {"avg_ms": f}
        if args.workload == "joint":
            pairs = C * model.observations
            out["kernels"]["forward"]["cdm_tflops"] = pairs * CDM_FLOPS_PER_PAIR / (f * 1e-3) / 1e12
    if cnt[2]:
        out.setdefault("kernels", {})["mogi_forward"] = {"avg_ms": ms[2] / cnt[2]}
    if args.workload == "joint":
        f32s, f64s, skipped, invalid = src.side_census(C)
        valid = max(C - invalid, 1)
        out["cdm_sides"] = {"fp64_threshold": model.fp64_sides, "fp32_per_chain": f32s / valid,
                            "fp64_per_chain": f64s / valid, "invalid_chains": invalid}
    if args.workload == "joint" and args.joint_da:
        # the surrogate covers the whole whitened residual field (both sources);
        # 3 steps per stage, 196,608 training fields (3 exact steps of 65,536 chains)
        out["surrogate_da"] = da_leg(args, model, C, seed=777 + 5, capacity=3 * C, stage_steps=3,
                                     fit_kwargs=dict(epochs=20, batch=4096, lr=2e-3, pca_sample=16384),
                                     validate_chains=512, moments=False)
    if args.workload == "mogi":
        # the whole anneal from a fresh sample
        ann2 = Annealer(Metropolis(steps=5), COV(), seed=job.seed).initialize(job)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ann2.posterior(model)
        torch.cuda.synchronize()
        out["anneal"] = {"wall_s": time.perf_counter() - t1, "stages": len(ann2.history),
                         "final_beta": ann2.worker.beta}
    return out


class JumpMonitor:
    """Squared jump distance of every chain per step, in units of the
    starting sample's standard deviation per parameter, summed over the
    parameters and averaged over the chains (expected squared jump distance,
    Pasarica and Gelman 2010)."""

    def __init__(self, step, sd):
        self.step, self.inv_sd, self.prev, self.total, self.steps = step, 1.0 / sd, None, 0.0, 0

    def on(self, event, controller):
        if event == "chainAdvanceStart":
            self.prev = self.step.theta.clone()
        elif event == "chainAdvanceFinish" and self.prev is not None:
            dx = (self.step.theta - self.prev) * self.inv_sd
            self.total += float((dx * dx).sum(1).mean())
            self.steps += 1


def fused_kernel_timing(ann, da, start, subs, reps=3):
    """HIP-event time of one altar_da_trajectory_fused launch (all chains,
    ``subs`` sub-steps) on the launch stream, and the network's algorithmic
    rate: 2 (P 256 + 2 * 256 * 256 + 256 * 64) flops per chain and
    evaluation, subs + 1 evaluations per launch."""
    import torch
    from altar_amd import _lib
    model, sur = ann.model, da.surrogate
    C, P = start.theta.shape
    if not da.fused_supported(model, P):
        return None
    ws = da._workspace(C, P, start.theta.device)
    cnt = torch.zeros(3, dtype=torch.int64, device=start.theta.device)
    L = (da.sigma_chol * da.substep_scale).contiguous()
    p = _lib.ptr

    def launch(k):
        _lib.call("altar_da_trajectory_fused", sur.mlp.handle, model._prior.handle, model.source.handle,
                  p(start.theta), p(start.prior), p(ws["post_s"]), p(ws["phi"]), p(ws["prior_phi"]),
                  p(ws["post_phi"]), p(L), p(sur.e_d), p(sur.s_d), sur.K, float(sur.cperp),
                  float(model.normalization), float(start.beta), C, P, 99, 0, k, subs, p(ws["index"]), p(cnt),
                  _lib.stream())
    launch(0)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for k in range(reps):
        launch(k + 1)
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / reps
    flops = 2.0 * (P * 256 + 2 * 256 * 256 + 256 * 64) * C * (subs + 1)
    return {"ms_per_launch": ms, "us_per_substep": 1e3 * ms / subs, "network_tflops": flops / (ms * 1e-3) / 1e12,
            "network_flops_per_launch": flops, "substeps": subs}


def surrogate_llk_error(ann, sur, step, n):
    """Surrogate against exact data llk on the first n chains of the stage:
    the spread of the difference, times beta, is what the exact stage of
    delayed acceptance sees (its acceptance falls as beta sd grows)."""
    import torch
    from altar_amd import _lib
    theta = step.theta[:n].contiguous()
    exact = torch.empty(n, dtype=torch.float64, device=theta.device)
    surr = torch.empty(n, dtype=torch.float64, device=theta.device)
    model = ann.model
    model.source.data_llk(model.restrict(theta), exact, model.precision)
    sur.data_llk(theta, surr)
    d = (exact - surr).cpu().numpy()
    return {"chains": n, "beta": float(step.beta), "mean_nats": float(d.mean()), "sd_nats": float(d.std()),
            "beta_times_sd": float(step.beta * d.std())}


class Trace:
    """Records a walk: the ensemble mean of theta (P, FP64) after every exact
    step, and theta itself (FP32) for the steps in [first, first + keep)."""

    def __init__(self, step, steps, first, keep):
        import torch
        C, P = step.theta.shape
        dev = step.theta.device
        self.step, self.t, self.first, self.keep = step, 0, first, keep
        self.means = torch.empty((steps, P), dtype=torch.float64, device=dev)
        self.window = torch.empty((keep, C, P), dtype=torch.float32, device=dev)

    def on(self, event, controller):
        if event != "chainAdvanceFinish":
            return
        th = self.step.theta
        if self.t < self.means.shape[0]:
            self.means[self.t] = th.mean(0)
        j = self.t - self.first
        if 0 <= j < self.keep:
            self.window[j].copy_(th)
        self.t += 1


def autocorrelation_times(window, max_lag=None):
    """Integrated autocorrelation time of every parameter from a (L, C, P)
    record of C chains: each step is centred on its ensemble mean (so a
    common drift of the population, e.g. while it forgets its start, does
    not enter), rho_k is the chain-averaged lag-k autocovariance over the
    pooled variance, tau = -1 + 2 sum_m Gamma_m with Gamma_m = rho_2m +
    rho_2m+1 summed while positive and kept non-increasing (Geyer 1992, the
    initial monotone sequence).  Returns (tau per parameter, whether the
    sequence ran past max_lag, so that tau is only a lower bound, and
    rho at lags 1, 10, 100)."""
    import torch
    L, C, P = window.shape
    max_lag = min(max_lag or L // 2, L - 1)
    n = 1 << (2 * L - 1).bit_length()
    taus, cut, rhos = [], [], []
    for j in range(P):
        x = window[:, :, j].double()
        x = x - x.mean(1, keepdim=True)
        var = (x * x).mean()
        f = torch.fft.rfft(x, n=n, dim=0)
        ac = torch.fft.irfft(f * f.conj(), n=n, dim=0)[: max_lag + 1].sum(1)
        ac = ac / (C * (L - torch.arange(max_lag + 1, device=x.device, dtype=torch.float64)))
        rho = (ac / var).cpu().numpy()
        tau, prev, m, truncated = -1.0, np.inf, 0, True
        while 2 * m + 1 <= max_lag:
            g = rho[2 * m] + rho[2 * m + 1]
            if g <= 0:
                truncated = False
                break
            g = min(g, prev)
            tau += 2.0 * g
            prev = g
            m += 1
        taus.append(max(tau, 1.0))
        cut.append(truncated)
        rhos.append([float(rho[k]) if k <= max_lag else None for k in (1, 10, 100)])
    return np.array(taus), np.array(cut), rhos


def mixing_and_moments(ann, da, plain, start, subs, times, ref_steps, ref_burn, da_steps, traj_steps,
                       biased_steps, bias_sd=1.0, seed=41):
    """Effective samples per second from measured autocorrelation, and
    posterior moments at the stage's temperature judged against a
    surrogate-free reference.
    * Every run starts from the stage's resampled population.
    * Plain Metropolis runs ref_steps steps: its ensemble means after ref_burn
      steps are the reference moments, and its autocorrelation comes from
      those steps.
    * Two-stage delayed acceptance runs da_steps steps, surrogate trajectories
      traj_steps steps with the trained surrogate and biased_steps with a
      deliberately biased one (its PCA-coefficient offsets shifted by
      N(0, bias_sd^2) draws, a smooth error of the surrogate llk, which slows
      its mixing); each discards its first fifth.
    * tau_k per parameter (Geyer's initial monotone sequence over the chain
      average); effective samples per second = C / (tau_k * seconds per step)
      with the timed seconds per step of ``times``; the reported figure is
      the worst parameter's.
    * Standard error of a run's mean of parameter k: sd_k sqrt(tau_k / (C L))
      over its L kept steps.  z_k = (mean - reference mean) / sqrt(se^2 +
      se_ref^2); chi2 = sum_k z_k^2 (10 parameters: 18.3 is the 95% point).
    * The same against plain Metropolis's means over the run's own kept
      steps (vs_plain_same_window): every run starts from the same population,
      so a slow drift of the population away from it enters both alike."""
    import torch
    sur = da.surrogate
    C, P = start.theta.shape
    seed0, keep = ann.seed, (da.steps, da.substeps)
    e_true = sur.e_d
    shift = torch.from_numpy(np.random.default_rng(seed).normal(0.0, bias_sd, sur.K)).to(e_true.device)
    runs = {}

    def walk(name, sampler, steps, burn, window, fn):
        st = start.clone()
        tr = Trace(st, steps, burn, window)
        ann.dispatcher.subscribe(tr)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        try:
            fn(st, steps)
        finally:
            ann.dispatcher.monitors.remove(tr)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        tau, cut, rhos = autocorrelation_times(tr.window)
        w = tr.window.double()
        runs[name] = {"steps": steps, "burn": burn, "wall_s": wall, "tau": tau, "tau_lower_bound": cut, "rho": rhos,
                      "mean": tr.means[burn:].mean(0).cpu().numpy(), "sd": w.reshape(-1, P).std(0).cpu().numpy(),
                      "kept": steps - burn, "means": tr.means.cpu().numpy()}

    def run_plain(st, steps):
        ann.sampler = plain
        keep_p = plain.steps
        plain.steps = steps
        try:
            plain.walkChains(ann, st)
        finally:
            plain.steps = keep_p

    def run_da(st, steps, substeps):
        ann.sampler = da
        da.steps, da.substeps = steps, substeps
        if substeps > 1:
            da._walk_trajectory(ann, st)
        else:
            da._walk_da(ann, st)

    try:
        ann.seed = seed0 + 101
        walk("plain", plain, ref_steps, ref_burn, ref_steps - ref_burn, run_plain)
        ann.seed = seed0 + 202
        b = da_steps // 5
        walk("da", da, da_steps, b, da_steps - b, lambda st, n: run_da(st, n, 1))
        ann.seed = seed0 + 303
        b = traj_steps // 5
        walk("trajectory", da, traj_steps, b, traj_steps - b, lambda st, n: run_da(st, n, subs))
        sur.e_d = (e_true + shift).contiguous()
        ann.seed = seed0 + 404
        b = biased_steps // 5
        walk("trajectory_biased_surrogate", da, biased_steps, b, biased_steps - b,
             lambda st, n: run_da(st, n, subs))
    finally:
        sur.e_d = e_true
        ann.seed = seed0
        da.steps, da.substeps = keep
        ann.sampler = da
    ref = runs["plain"]
    ref_se = ref["sd"] * np.sqrt(ref["tau"] / (C * ref["kept"]))
    out = {"start": "the stage's resampled population", "bias_sd": bias_sd}
    for name, r in runs.items():
        ms = times.get("trajectory" if name.startswith("trajectory") else name)
        ess_s = C / (r["tau"] * ms * 1e-3) if ms else None
        se = r["sd"] * np.sqrt(r["tau"] / (C * r["kept"]))
        o = {"steps": r["steps"], "burn_in": r["burn"], "wall_s": r["wall_s"], "ms_per_step_timed": ms,
             "tau": r["tau"].tolist(), "tau_is_lower_bound": r["tau_lower_bound"].tolist(),
             "rho_lag_1_10_100": r["rho"],
             "tau_max": float(r["tau"].max()),
             "effective_samples_per_s_min": float(ess_s.min()) if ess_s is not None else None,
             "mean": r["mean"].tolist(), "sd": r["sd"].tolist(), "mean_se": se.tolist()}
        if name != "plain":
            z = (r["mean"] - ref["mean"]) / np.sqrt(se ** 2 + ref_se ** 2)
            o["vs_plain_reference"] = {"z": z.tolist(), "max_abs_z": float(np.abs(z).max()),
                                       "chi2": float((z ** 2).sum()),
                                       "sd_ratio": (r["sd"] / ref["sd"]).tolist()}
            w0, w1 = r["burn"], r["steps"]
            if w1 <= ref["steps"]:
                ref_w = ref["means"][w0:w1].mean(0)
                ref_se_w = ref["sd"] * np.sqrt(ref["tau"] / (C * (w1 - w0)))
                zw = (r["mean"] - ref_w) / np.sqrt(se ** 2 + ref_se_w ** 2)
                o["vs_plain_same_window"] = {"steps": [w0, w1], "z": zw.tolist(),
                                             "max_abs_z": float(np.abs(zw).max()),
                                             "chi2": float((zw ** 2).sum())}
        out[name] = o
    base = out["plain"]["effective_samples_per_s_min"]
    for name in ("da", "trajectory", "trajectory_biased_surrogate"):
        e = out[name]["effective_samples_per_s_min"]
        out[name]["gain_effective_samples_per_s"] = e / base if (e and base) else None
    return out


def da_leg(args, model, C, seed=777 + 3, capacity=None, stage_steps=5, fit_kwargs=None, validate_chains=1024,
           moments=True):
    """Anneal with field collection until beta >= 0.1, fit the surrogate
    (PCA of whitened residuals + BF16 MLP), then time three samplers from the
    same state at the next temperature: plain Metropolis, two-stage delayed
    acceptance, and delayed acceptance along surrogate trajectories of
    --da-substeps sub-steps.  Each is timed on its own; a second, identical
    run of each (same streams, same result) measures its squared jump
    distance per step."""
    import torch
    from altar_amd import Job
    from altar_amd.bayesian import Annealer, COV, Metropolis
    from altar_amd.bayesian.da import DelayedAcceptance
    from altar_amd.surrogate import Surrogate

    job = Job(chains=C, steps=stage_steps, seed=seed)
    model.job = job
    t0 = time.perf_counter()
    sur = Surrogate(model, K=64, capacity=capacity or args.da_train, max_batch=C, seed=3)
    da = DelayedAcceptance(sur, train_beta=0.1, max_rmse=2e-3, steps=stage_steps, validate_chains=validate_chains,
                           fit_kwargs=fit_kwargs or dict(epochs=args.da_epochs, batch=4096, lr=2e-3))
    ann = Annealer(da, COV(), seed=job.seed).initialize(job)
    ann.model = model
    w = ann.worker
    w.start(ann)
    t_fit = None
    extra = 0
    while w.beta + job.tolerance < 1:
        w.cool(ann)
        stats = w.walk(ann)
        w.resample(ann, stats)
        if sur.trained and t_fit is None:
            t_fit = time.perf_counter() - t0
        if sur.trained and w.beta >= 0.1:
            # past training: stop once a stage ran delayed acceptance (the
            # surrogate validated), or after a few refits on the posterior
            extra += 1
            if da.mode_history[-1]["mode"] == "da" or extra >= 4:
                break
    setup = time.perf_counter() - t0
    if not sur.trained:
        return {"error": "surrogate not trained", "beta": w.beta}
    # one more stage: temperature, covariance, resampling; then time the samplers
    w.cool(ann)
    step = w.step
    da.prepareSamplingPDF(ann, step)
    ok = da._validate(ann, step)
    llk_err = surrogate_llk_error(ann, sur, step, n=min(2048, C)) if sur.trained else None
    start = step.clone()
    sd = start.theta.std(0)
    P = start.theta.shape[1]
    K = max(2, args.steps)
    da.steps = K
    plain = Metropolis(steps=K)
    plain.scaling = da.scaling
    plain.sigma_chol = da.sigma_chol
    sweep = [max(1, int(v)) for v in str(args.da_substeps).split(",")]
    subs = sweep[0]
    da.substep_scale = args.da_substep_scale
    tuning = None
    if args.da_substep_target and ok and da.fused_supported(model, start.theta.shape[1]):
        # the stage's pilot tuning of the sub-step scale, once, outside the
        # timed runs (which then share the scale, so the replays repeat them)
        da.substep_target = args.da_substep_target
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        da.tune_substep_scale(ann, start, da._workspace(C, start.theta.shape[1], start.theta.device))
        torch.cuda.synchronize()
        tuning = {"target": da.substep_target, "scale": da.substep_scale, "ms": 1e3 * (time.perf_counter() - t3)}
    da.substep_target = None

    def walk(kind):
        if kind == "plain":
            ann.sampler = plain
            return lambda st: plain.walkChains(ann, st)
        ann.sampler = da
        da.substeps = subs if kind == "trajectory" else 1
        return (lambda st: da._walk_trajectory(ann, st)) if kind == "trajectory" else (lambda st: da._walk_da(ann, st))

    def run(kind):
        st = start.clone()
        fn = walk(kind)
        ev0 = da.exact_evals
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        acc, rej, unl = fn(st)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t1
        exact = C * K if kind == "plain" else da.exact_evals - ev0
        # replay with the jump monitor (same streams: the same moves)
        st2 = start.clone()
        mon = JumpMonitor(st2, sd)
        ann.dispatcher.subscribe(mon)
        try:
            walk(kind)(st2)
        finally:
            ann.dispatcher.monitors.remove(mon)
        same = bool(torch.equal(st2.theta, st.theta))
        esjd = mon.total / max(mon.steps, 1)
        return {"ms_per_step": dt / K * 1e3, "accepted": acc, "exact_evals": exact,
                "accepted_samples_per_s": acc / dt, "exact_evals_per_s": exact / dt,
                "esjd_per_step": esjd, "esjd_per_s": mon.total / dt, "replay_identical": same}

    res = {"plain": run("plain")}
    if ok:
        res["da"] = run("da")
        runs = []
        for n in sweep:
            if n < 2:
                continue
            subs = n
            sub0 = list(da.substep_counts)
            r = run("trajectory")
            sc = [b - a for a, b in zip(sub0, da.substep_counts)]
            r.update(substeps=n, substep_scale=da.substep_scale, substep_acceptance=sc[0] / max(sum(sc), 1),
                     substep_tuning=tuning)
            runs.append(r)
        if runs:
            res["trajectory"] = runs[0]
            subs = sweep[0]
            res["trajectory"]["fused_kernel"] = fused_kernel_timing(ann, da, start, subs)
            if moments and args.da_reference_steps > 0:
                times = {k: res[k]["ms_per_step"] for k in ("plain", "da", "trajectory") if k in res}
                ds, ts, bs = (int(v) for v in str(args.da_mixing_steps).split(","))
                res["mixing"] = mixing_and_moments(ann, da, plain, start, subs, times, args.da_reference_steps,
                                                   args.da_reference_burn, ds, ts, bs)
        if len(runs) > 1:
            res["trajectory_sweep"] = runs
    ann.sampler = da
    pl = res["plain"]
    out = {
        "beta": step.beta, "steps": K, "validation_rmse_m": sur.validation_rmse, "surrogate_used": bool(ok),
        "refits": da.refits, "modes": [m["mode"] for m in da.mode_history],
        "training_fields": sur.count, "fit_s": t_fit, "setup_s": setup, "surrogate_llk_error": llk_err,
        "mlp_flops_per_chain": 2 * (model.parameters * 256 + 256 * 256 * 2 + 256 * 64),
    }
    out.update(res)
    for kind, r in [("da", res.get("da")), ("trajectory", res.get("trajectory"))] + \
            [(f"sweep{i}", r) for i, r in enumerate(res.get("trajectory_sweep", []))]:
        if r is not None and pl["accepted"]:
            r["gain_accepted_per_s"] = r["accepted_samples_per_s"] / pl["accepted_samples_per_s"]
            r["gain_esjd_per_s"] = r["esjd_per_s"] / pl["esjd_per_s"] if pl["esjd_per_s"] else None
    # the headline gains: accepted samples per second (the exact stage's
    # count A of Metropolis.py:217), then effective samples per second from
    # the measured autocorrelation of the worst parameter
    mix = res.get("mixing", {})
    for kind in ("da", "trajectory"):
        if kind in res and "gain_accepted_per_s" in res[kind]:
            out[f"{kind}_gain_accepted_per_s"] = res[kind]["gain_accepted_per_s"]
            out[f"{kind}_gain_effective_samples_per_s"] = mix.get(kind, {}).get("gain_effective_samples_per_s")
            out[f"{kind}_gain_esjd_per_s"] = res[kind]["gain_esjd_per_s"]
    out["gain_accepted_per_s"] = out.get("da_gain_accepted_per_s")
    return out

if __name__ == "__main__":
    sys.exit(main())
