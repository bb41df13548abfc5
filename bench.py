#!/usr/bin/env python
"""Headline benchmark: batched Tsit5 forward solves per second on one MI355X.

Workload (the reference path's uneven batched configuration): 65,536 Van der Pol
oscillators y'' = mu (1 - y^2) y' - y with mu ~ LogUniform[0.1, 5], y0 ~ U[-2, 2]^2,
t1 ~ U[5, 50] per instance, Tsit5 + PIDController(rtol=1e-6, atol=1e-9), dt0 = 0.1,
SaveAt(ts = 64 points scaled to each [0, t1]), max_steps = 16384; inputs from
numpy.random.default_rng(0) in that draw order, resident on the GPU before timing.

One "step" is one complete batched solve (one persistent-kernel launch, replayed from a
HIP graph).  Two device forms are timed: the float64 plain form (bit-identical to the
FP64 oracle) and the float32-state / float64-time form; `value` is the faster form whose
max relative error against the FP64 reference is <= 1e-6.  The CPU baseline is the FP64
C oracle (OpenMP over the host cores) on the same inputs.

Prints one JSON line.  `--gpus N` with N > 1 exits: this build measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6     # MI355X published FP64 (vector and matrix) peak
FP32_PEAK_TFLOPS = 157.3    # MI355X published FP32 vector peak
# Algorithmic flops (Tsit5 on a d=2 system with a 5-flop vector field):
#   per attempted step 2*(42 + 7 + 13 + 6) + 6*5 = 166; per save 35 + 14*d = 63.
FLOP_PER_STEP = 166
FLOP_PER_SAVE = 63


def make_inputs(seed=0, B=65536, n_ts=64):
    import numpy as np
    rng = np.random.default_rng(seed)
    mu = np.exp(rng.uniform(np.log(0.1), np.log(5.0), size=B))
    y0 = rng.uniform(-2.0, 2.0, size=(B, 2))
    t1 = rng.uniform(5.0, 50.0, size=B)
    ts = np.linspace(0.0, 1.0, n_ts)[None, :] * t1[:, None]
    return mu, y0, t1, ts


def prepare(dtype, mu, y0, t1, ts):
    import torch
    import diffrax_amd as dfx
    from diffrax_amd import _integrate
    dev = torch.device("cuda", 0)
    return _integrate._prepare_erk(
        dfx.ODETerm(dfx.VanDerPol()), dfx.Tsit5(), 0.0,
        torch.tensor(t1, dtype=torch.float64, device=dev), 0.1,
        torch.tensor(y0, dtype=dtype, device=dev), torch.tensor(mu, dtype=dtype, device=dev),
        dfx.SaveAt(ts=torch.tensor(ts, dtype=torch.float64)),
        dfx.PIDController(rtol=1e-6, atol=1e-9), 16384, None)


def time_form(plan, steps, warmup):
    import torch
    from diffrax_amd import _autotune
    cfg, tune_ms = _autotune.tune(plan)
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plan.launch()      # warm the code object before capture
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        plan.launch()
    for _ in range(warmup):
        g.replay()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        a.record()
        g.replay()
        b.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [a.elapsed_time(b) for a, b in evs]
    return dict(ms_per_step=wall * 1e3 / steps, kernel_ms_mean=statistics.fmean(per),
                kernel_ms_median=statistics.median(per), launch=list(cfg), tune_ms=tune_ms)


def cpu_baseline(mu, y0, t1, ts, reps=3):
    import numpy as np
    import oracle
    cores = min(len(os.sched_getaffinity(0)), 16)
    times = []
    out = None
    for _ in range(reps):
        t = time.perf_counter()
        out = oracle.solve(y0, t1, mu[:, None], ts=ts, solver="tsit5", vf="vdp", dt0=0.1,
                           rtol=1e-6, atol=1e-9, max_steps=16384, nthreads=cores)
        times.append(time.perf_counter() - t)
    med = statistics.median(times)
    cpu_model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    _ = np
    return dict(value=y0.shape[0] / med, unit="solves/s", cores=cores, kind="port",
                sample=f"full workload ({y0.shape[0]} trajectories), FP64 C oracle, "
                       f"OpenMP {cores} threads, median of {reps} runs ({med:.3f} s each); "
                       f"host CPU: {cpu_model}; JAX-CPU diffrax: unavailable"), out


def read_traffic(kernel_tag):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if present."""
    path = os.path.join(ROOT, "profiles", f"{kernel_tag}_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--forms", default="f64,f32", help="device forms to time")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (PMC byte-counter scale)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch
    import oracle

    mu, y0, t1, ts = make_inputs()
    B = y0.shape[0]
    forms = {}
    results = {}
    plans = {}
    for form in args.forms.split(","):
        dtype = torch.float64 if form == "f64" else torch.float32
        plan = prepare(dtype, mu, y0, t1, ts)
        plans[form] = plan
        forms[form] = time_form(plan, args.steps, args.warmup)
        torch.cuda.synchronize()
        results[form] = dict(ys=plan.t["ys"].double().cpu().numpy(),
                             stats=plan.t["stats"].cpu().numpy(),
                             result=plan.t["result"].cpu().numpy())

    cpu = None
    ref = None
    if not args.no_cpu:
        cpu, ref = cpu_baseline(mu, y0, t1, ts)
    ref_ys = ref["ys"] if ref is not None else results["f64"]["ys"]
    for form in forms:
        st = results[form]["stats"]
        attempted = int(st[:, 0].sum())
        saves = int(np.isfinite(results[form]["ys"][:, :, 0]).sum())
        flops = attempted * FLOP_PER_STEP + saves * FLOP_PER_SAVE
        forms[form].update(
            solves_per_s=B / (forms[form]["ms_per_step"] / 1e3),
            attempted_steps=attempted, accepted_steps=int(st[:, 1].sum()),
            max_steps_one_trajectory=int(st[:, 0].max()),
            algorithmic_flop=flops,
            achieved_tflops=flops / (forms[form]["kernel_ms_mean"] * 1e-3) / 1e12,
            max_rel_err_vs_fp64=oracle.max_rel_err(results[form]["ys"], ref_ys, 1e-9),
            all_successful=bool((results[form]["result"] == 0).all()),
        )
        if ref is not None:
            forms[form]["step_counts_equal_fp64_oracle"] = bool(
                np.array_equal(results[form]["stats"], ref["stats"]))

    eligible = [f for f in forms if forms[f]["max_rel_err_vs_fp64"] <= 1e-6]
    head = max(eligible, key=lambda f: forms[f]["solves_per_s"]) if eligible else "f64"
    h = forms[head]
    peak = FP64_PEAK_TFLOPS
    roof = {"bound": "valu", "achieved": round(h["achieved_tflops"], 4), "peak": peak,
            "unit": "TFLOP/s", "frac": round(h["achieved_tflops"] / peak, 5),
            "traffic": read_traffic("erk_small_tsit5_vdp_" + head),
            "kernel": f"erk_small_kernel<Tsit5, VanDerPol, {'double' if head == 'f64' else 'float'}>",
            "note": "FP64 VALU peak; algorithmic flop = 166/attempted step + 63/save from "
                    "the kernel's own step counters"}
    line = {
        "metric": "batched Tsit5 forward solves/s at err<=1e-6 vs plain FP64",
        "value": round(B / (h["ms_per_step"] / 1e3), 1),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(h["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "f64" else "f32 state / f64 time",
        "data": "synthetic (numpy default_rng(0))",
        "config": {"workload": "65,536 uneven Van der Pol, Tsit5 + PID(1e-6, 1e-9), "
                               "t1~U[5,50], 64 saves per instance, forward only",
                   "batch": B, "n_ts": 64, "form": head, "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "forms": {f: {k: v for k, v in forms[f].items() if k != "tune_ms"} for f in forms},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
