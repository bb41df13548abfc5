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

The same line carries the other workloads of the path under "workloads" (not the
headline; each with its own timing, error and CPU-baseline fields):
  neural_ode_forward: MLP 64 -> 256 -> 256 -> 64 (softplus, tanh output), Xavier-normal
      weights, zero biases, batch 8,192, y0 ~ N(0, 1), t in [0, 1], Tsit5 + PID(1e-5, 1e-7),
      dt0 = 0.1, 32 saves; float32 and bfloat16-MFMA forms.
  neural_ode_grad: the same solve plus the checkpointed discrete adjoint of
      loss = mean((ys - targets)^2), targets ~ N(0, 0.5^2), gradients for all 98,880 MLP
      parameters and y0.
  vdp_forward_grad: the headline solve plus the discrete-adjoint gradient (the default
      RecursiveCheckpointAdjoint) of loss = mean(ys^2) for y0 and mu; float64 (bit-identical
      to the FP64 oracle's adjoint on a 512-instance subset) and float32-state forms.
  ou_sde_heun: 262,144 paths of a 16-dim Ornstein-Uhlenbeck SDE, Heun, 1,000 steps on
      [0, 1], float32 Brownian increments [1000, 262144, 16] (16.8 GB) resident in HBM;
      forward, and forward + adjoint of loss = mean y(1)^2 for theta, mu, sigma.
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


def prepare(dtype, mu, y0, t1, ts, rtol=1e-6, atol=1e-9):
    import torch
    import diffrax_amd as dfx
    from diffrax_amd import _integrate
    dev = torch.device("cuda", 0)
    return _integrate._prepare_erk(
        dfx.ODETerm(dfx.VanDerPol()), dfx.Tsit5(), 0.0,
        torch.tensor(t1, dtype=torch.float64, device=dev), 0.1,
        torch.tensor(y0, dtype=dtype, device=dev), torch.tensor(mu, dtype=dtype, device=dev),
        dfx.SaveAt(ts=torch.tensor(ts, dtype=torch.float64)),
        dfx.PIDController(rtol=rtol, atol=atol), 16384, None)


def matched_error(mu, y0, t1, ts, tight, target, steps, warmup, sub):
    """The float32-state form at the loosest tolerance (rtol r, atol r / 1000; r = 1e-6
    halved until it matches) whose normwise error against the tight FP64 solve of the
    checked subset is at most `target`, the FP64 form's at rtol 1e-6; timed there."""
    import numpy as np
    import torch
    scale = float(np.max(np.abs(tight)))
    tried = []
    r = 1e-6
    for _ in range(6):
        plan = prepare(torch.float32, mu, y0, t1, ts, rtol=r, atol=r * 1e-3)
        plan.launch()
        torch.cuda.synchronize()
        ys = plan.t["ys"][:sub].double().cpu().numpy()
        ok = bool((plan.t["result"] == 0).all())
        err = float(np.max(np.abs(ys - tight)) / scale)
        tried.append({"rtol": r, "normwise_err": err, "all_successful": ok})
        if err <= target and ok:
            t = time_form(plan, steps, warmup)
            B = y0.shape[0]
            res = {"matched": True, "rtol": r, "atol": r * 1e-3, "normwise_err": err,
                   "ms_per_step": round(t["ms_per_step"], 4),
                   "kernel_ms_mean": round(t["kernel_ms_mean"], 4), "launch": t["launch"],
                   "solves_per_s": round(B / (t["ms_per_step"] / 1e3), 1), "tried": tried}
            plan.t.clear()
            return res
        plan.t.clear()
        r *= 0.5
    return {"matched": False, "tried": tried,
            "note": "no tolerance down to rtol 3.1e-8 brings the float32-state form to the "
                    "FP64 form's error: its float32 rounding, amplified by the neutrally "
                    "stable limit-cycle phase, sets the floor"}


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


def cpu_baseline(mu, y0, t1, ts, min_seconds=10.0, max_reps=200):
    """The FP64 C oracle on the full workload, repeated until ~10 s of CPU time."""
    import numpy as np
    import oracle
    cores = min(len(os.sched_getaffinity(0)), 16)
    times = []
    out = None
    while (sum(times) < min_seconds and len(times) < max_reps) or len(times) < 3:
        t = time.perf_counter()
        out = oracle.solve(y0, t1, mu[:, None], ts=ts, solver="tsit5", vf="vdp", dt0=0.1,
                           rtol=1e-6, atol=1e-9, max_steps=16384, nthreads=cores)
        times.append(time.perf_counter() - t)
    reps = len(times)
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
                       f"OpenMP {cores} threads, median of {reps} runs ({med:.3f} s each, "
                       f"{sum(times):.1f} s in total); "
                       f"host CPU: {cpu_model}; JAX-CPU diffrax: unavailable"), out


MLP_FLOP_PER_EVAL = 2 * (64 * 256 + 256 * 256 + 256 * 64)   # 196,608 MFMA flop per trajectory
BF16_PEAK_TFLOPS = 2500.0   # MI355X dense bfloat16 MFMA
F32_MFMA_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _event_time(fn, steps, warmup):
    """Mean and median event time (ms) of fn() on the current stream, and wall ms/step."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [a.elapsed_time(b) for a, b in evs]
    return statistics.fmean(per), statistics.median(per), wall * 1e3 / steps


def _graphed(launch):
    import torch
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        launch()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        launch()
    return g


def mlp_inputs(seed=0, B=8192, n_ts=32):
    import numpy as np
    rng = np.random.default_rng(seed)
    xav = lambda o, i: rng.normal(0.0, np.sqrt(2.0 / (i + o)), size=(o, i))
    W1, W2, W3 = xav(256, 64), xav(256, 256), xav(64, 256)
    y0 = rng.normal(size=(B, 64))
    targets = rng.normal(0.0, 0.5, size=(B, n_ts, 64))
    ts = np.linspace(0.0, 1.0, n_ts)
    return [W1, np.zeros(256), W2, np.zeros(256), W3, np.zeros(64)], y0, ts, targets


def mlp_grad_error(gplan, W, y0, ts, check):
    """Gradient check on `check` rows: the device adjoint of sum(w * ys) with w nonzero on
    those rows only (weights gradients then come from them alone), against torch float64
    autograd through a replay of the device's accepted steps of those rows."""
    import numpy as np
    import torch
    from diffrax_amd import _mlp
    from oracle import pyoracle
    rng = np.random.default_rng(17)
    B = y0.shape[0]
    rows = np.sort(rng.choice(B, size=check, replace=False))
    w = torch.zeros(gplan.ys.shape, dtype=torch.float32, device=gplan.ys.device)
    wr = rng.normal(size=(check,) + tuple(gplan.ys.shape[1:])).astype(np.float32)
    w[torch.as_tensor(rows, device=w.device)] = torch.tensor(wr, device=w.device)
    gplan.launch()
    g = _mlp.backward_mlp(gplan, w)     # (y0_bar, W1, b1, W2, b2, W3, b3)
    torch.cuda.synchronize()
    grid = gplan.grid.cpu().numpy()[rows]
    nacc = gplan.stats[:, 1].cpu().numpy()[rows]
    Pd = [torch.tensor(np.asarray(x, dtype=np.float32).astype(np.float64), requires_grad=True)
          for x in W]
    W1, b1, W2, b2, W3, b3 = Pd
    sp = torch.nn.functional.softplus
    f = lambda t, y: torch.tanh(sp(sp(y @ W1.T + b1) @ W2.T + b2) @ W3.T + b3)
    Y0 = torch.tensor(y0[rows].astype(np.float32).astype(np.float64), requires_grad=True)
    loss = 0.0
    for i in range(check):
        steps = [tuple(grid[i, n]) for n in range(nacc[i])]
        out = pyoracle.replay(f, "tsit5", steps, Y0[i], list(ts))
        loss = loss + (torch.stack(out) * torch.tensor(wr[i].astype(np.float64))).sum()
    loss.backward()
    got = [g[0][torch.as_tensor(rows, device=g[0].device)]] + list(g[1:])
    errs = [float((a.double().cpu() - r).abs().max() / r.abs().max())
            for a, r in zip(got, [Y0.grad] + [p.grad for p in Pd])]
    worst = max(errs) if all(e == e for e in errs) else float("nan")   # a NaN stays visible
    return dict(grad_max_rel_err_vs_fp64_replay=worst, grad_check_rows=check)


def bench_mlp(steps, warmup, with_cpu, check=256):
    """Neural-ODE forward and forward+gradient, float32 and bfloat16-MFMA forms."""
    import numpy as np
    import torch
    import diffrax_amd as dfx
    from diffrax_amd import _mlp
    import oracle
    dev = torch.device("cuda", 0)
    W, y0, ts, targets = mlp_inputs()
    B = y0.shape[0]
    ctrl = dfx.PIDController(rtol=1e-5, atol=1e-7)
    y0d = torch.tensor(y0, dtype=torch.float32, device=dev)
    tgt = torch.tensor(targets, dtype=torch.float32, device=dev)
    q = oracle.mlp_pack(*W)
    y0c = y0[:check].astype(np.float32).astype(np.float64)
    cores = min(len(os.sched_getaffinity(0)), 16)
    ref = oracle.solve(y0c, 1.0, q, ts=ts, solver="tsit5", vf="mlp", hidden=256, dt0=0.1,
                       rtol=1e-5, atol=1e-7, nthreads=cores)
    true = oracle.solve(y0c, 1.0, q, ts=ts, solver="tsit5", vf="mlp", hidden=256, dt0=0.1,
                        rtol=1e-12, atol=1e-12, max_steps=100000, nthreads=cores)["ys"]
    cpu_times = []
    if with_cpu:   # the whole batch, repeated until ~10 s of CPU time
        y0f = y0.astype(np.float32).astype(np.float64)
        while (sum(cpu_times) < 10.0 and len(cpu_times) < 100) or len(cpu_times) < 2:
            t = time.perf_counter()
            oracle.solve(y0f, 1.0, q, ts=ts, solver="tsit5", vf="mlp", hidden=256, dt0=0.1,
                         rtol=1e-5, atol=1e-7, nthreads=cores)
            cpu_times.append(time.perf_counter() - t)
    fwd, grad = {}, {}
    for matmul in ("f32", "bf16"):
        vf = dfx.MLPField(*(torch.tensor(w, dtype=torch.float32, device=dev) for w in W),
                          matmul=matmul)
        term = dfx.ODETerm(vf)
        plan = _mlp.prepare_mlp(term, dfx.Tsit5(), 0.0, 1.0, 0.1, y0d, dfx.SaveAt(ts=ts), ctrl,
                                4096, dev)
        g = _graphed(plan.launch)
        k_mean, k_med, wall = _event_time(g.replay, steps, warmup)
        st = plan.stats.cpu().numpy()
        ys = plan.ys[:check].double().cpu().numpy()
        flop = float(((6 * st[:, 0] + 1) * MLP_FLOP_PER_EVAL).sum())
        peak = F32_MFMA_PEAK_TFLOPS if matmul == "f32" else BF16_PEAK_TFLOPS
        tf = flop / (k_mean * 1e-3) / 1e12
        fwd[matmul] = dict(
            solves_per_s=round(B / (wall * 1e-3), 1), ms_per_solve=round(wall, 4),
            kernel_ms_mean=round(k_mean, 4), kernel_ms_median=round(k_med, 4),
            attempted_steps_mean=float(st[:, 0].mean()), attempted_steps_max=int(st[:, 0].max()),
            algorithmic_tflops=round(tf, 3), mfma_peak_tflops=peak, frac=round(tf / peak, 5),
            max_rel_err_vs_fp64=oracle.max_rel_err(ys, ref["ys"], 1e-7),
            max_abs_err_vs_tight_fp64=float(np.max(np.abs(ys - true))),
            fp64_oracle_abs_err_vs_tight=float(np.max(np.abs(ref["ys"] - true))),
            all_successful=bool((plan.result == 0).all().item()))
        # forward + gradient (checkpointed discrete adjoint), loss = mean((ys - targets)^2)
        gplan = _mlp.prepare_mlp(term, dfx.Tsit5(), 0.0, 1.0, 0.1, y0d, dfx.SaveAt(ts=ts), ctrl,
                                 4096, dev, ckpt_every=_mlp.checkpoint_every(
                                     dfx.RecursiveCheckpointAdjoint(), 4096))
        out = {}

        def fwd_grad():
            gplan.launch()
            ys_bar = (gplan.ys - tgt) * (2.0 / gplan.ys.numel())
            out["g"] = _mlp.backward_mlp(gplan, ys_bar)

        k_mean, k_med, wall = _event_time(fwd_grad, steps, warmup)
        gst = gplan.stats.cpu().numpy()
        acc = gst[:, 1].astype(np.float64)
        # MFMA work of one gradient: the forward's evaluations, then per accepted step the
        # segment recompute (6 per step + the first stage), the stage VJPs (7, each the
        # three transposed products) and the weight-gradient GEMM rows (7)
        gflop = float((((6 * gst[:, 0] + 1) + (6 * acc + 1) + 7 * acc + 7 * acc)
                       * MLP_FLOP_PER_EVAL).sum())
        gtf = gflop / (k_mean * 1e-3) / 1e12
        grad[matmul] = dict(solves_per_s=round(B / (wall * 1e-3), 1), ms_per_solve=round(wall, 4),
                            event_ms_mean=round(k_mean, 4),
                            checkpoint_every=int(gplan.prob.ckpt_every),
                            grad_finite=bool(all(torch.isfinite(x).all().item() for x in out["g"])),
                            algorithmic_tflops=round(gtf, 3), mfma_peak_tflops=peak,
                            frac=round(gtf / peak, 5),
                            **mlp_grad_error(gplan, W, y0, ts, check))
        if matmul == "f32":
            grad["f32_backsolve"] = bench_mlp_backsolve(plan, tgt, out["g"], steps)
        del plan, gplan, g
    cpu = None
    if cpu_times:
        med = statistics.median(cpu_times)
        cpu = dict(value=round(B / med, 2), unit="solves/s", cores=cores, kind="port",
                   sample=f"the full batch ({B} trajectories), FP64 C oracle forward, OpenMP "
                          f"{cores} threads, median of {len(cpu_times)} runs ({med:.3f} s each, "
                          f"{sum(cpu_times):.1f} s in total); host CPU: {_cpu_model()}")
    return dict(neural_ode_forward=dict(batch=B, forms=fwd, cpu_baseline=cpu,
                                        error_subset=check),
                neural_ode_grad=dict(batch=B, forms=grad,
                                     loss="mean((ys - targets)^2), targets ~ N(0, 0.25)"))


def bench_mlp_backsolve(plan, tgt, g_discrete, steps):
    """Forward + BacksolveAdjoint (adjoint seminorm, the forward's tolerances) of the same
    loss; wall clock, since the host reads one counter per backsolve launch."""
    import torch
    import diffrax_amd as dfx
    from diffrax_amd import _mlp
    adj = dfx.BacksolveAdjoint(stepsize_controller=dfx.PIDController(
        rtol=1e-5, atol=1e-7, norm=dfx.adjoint_rms_seminorm))
    back = _mlp.backsolve_problem(plan, adj)

    def fwd_back():
        plan.launch()
        return _mlp.backsolve_mlp(plan, back, (plan.ys - tgt) * (2.0 / plan.ys.numel()))
    g, res = fwd_back()
    torch.cuda.synchronize()
    n = max(2, steps // 4)
    t = time.perf_counter()
    for _ in range(n):
        g, res = fwd_back()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) * 1e3 / n
    B = plan.ys.shape[0]
    diffs = [float((a - b).abs().max() / b.abs().max()) for a, b in zip(g, g_discrete)]
    diff = max(diffs) if all(e == e for e in diffs) else float("nan")
    return dict(ms_per_solve=round(ms, 4), solves_per_s=round(B / (ms * 1e-3), 1),
                adjoint="BacksolveAdjoint(stepsize_controller=PIDController(1e-5, 1e-7, "
                        "norm=adjoint_rms_seminorm))",
                iterations_per_launch=_mlp.BACKSOLVE_CHUNK,
                all_successful=bool((res == 0).all().item()),
                grad_normwise_diff_vs_discrete_adjoint=diff)


def bench_sde(steps, warmup, with_cpu, P=262144, d=16, N=1000, cpu_paths=16384):
    """OU SDE, Heun, 1000 steps: forward and forward + adjoint (float32)."""
    import numpy as np
    import torch
    from diffrax_amd import _sde
    import oracle
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(0)
    theta = rng.uniform(0.5, 2.0, size=d)
    mu = rng.normal(0.0, 1.0, size=d)
    sigma = rng.uniform(0.1, 0.5, size=d)
    dt = 1.0 / N
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    dW = torch.randn((N, P, d), generator=gen, dtype=torch.float32, device=dev)
    dW.mul_(float(np.sqrt(dt)))
    y0 = torch.zeros((P, d), dtype=torch.float32, device=dev)
    prm = [torch.tensor(a, dtype=torch.float32, device=dev) for a in (theta, mu, sigma)]
    grid = torch.tensor(_sde.constant_grid(0.0, 1.0, dt), dtype=torch.float64, device=dev)
    fwd = _sde._Sde(y0, *prm, dW, grid, False, True, False, None)
    k_mean, k_med, wall = _event_time(fwd.forward, steps, warmup)
    alg_bytes = P * d * N * 4 + 2 * P * d * 4
    out = dict(paths=P, d=d, steps=N,
               forward=dict(paths_per_s=round(P / (k_mean * 1e-3), 1),
                            kernel_ms_mean=round(k_mean, 4), kernel_ms_median=round(k_med, 4),
                            ms_per_solve=round(wall, 4),
                            algorithmic_bytes=alg_bytes,
                            achieved_gbs=round(alg_bytes / (k_mean * 1e-3) / 1e9, 1),
                            frac=round(alg_bytes / (k_mean * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)))
    gs = _sde._Sde(y0, *prm, dW, grid, False, True, True, None)

    def fwd_grad():
        gs.forward()
        ys_bar = gs.ys * (2.0 / gs.ys.numel())
        out["g"] = gs.adjoint(ys_bar, False)

    k_mean, k_med, wall = _event_time(fwd_grad, steps, warmup)
    out["forward_grad"] = dict(paths_per_s=round(P / (wall * 1e-3), 1),
                               ms_per_solve=round(wall, 4), event_ms_mean=round(k_mean, 4),
                               checkpoint_every=int(gs.prob.checkpoint_every),
                               algorithmic_bytes=2 * alg_bytes,
                               achieved_gbs=round(2 * alg_bytes / (wall * 1e-3) / 1e9, 1))
    out.pop("g", None)
    # Errors on the last 512 paths (dW offsets past 2^31): the forward against the FP64
    # oracle, and the adjoint of sum(w * y(1)) with w nonzero on those paths only against
    # torch float64 autograd through the same fixed-step replay.
    last = slice(P - 512, P)
    dW_last = dW[:, last].double().cpu().numpy()
    q = np.concatenate([theta, mu, sigma])
    y0n = np.zeros((512, d))
    ref = oracle.solve(y0n, 1.0, q, dW=dW_last, solver="heun", vf="ou", controller="constant",
                       dt0=dt, save_t1=True, max_steps=100000)
    gs.forward()
    ys_last = gs.ys[last].double().cpu().numpy()
    out["forward"]["normwise_err_vs_fp64_oracle_last512"] = float(
        np.max(np.abs(ys_last - ref["ys"])) / np.max(np.abs(ref["ys"])))
    w = torch.zeros_like(gs.ys)
    wr = np.random.default_rng(9).normal(size=(512, 1, d))
    w[last] = torch.tensor(wr, dtype=torch.float32, device=dev)
    _, thb, mub, sgb = gs.adjoint(w, False)
    from oracle import pyoracle
    TH, MU, SG = (torch.tensor(a, requires_grad=True) for a in (theta, mu, sigma))
    dWt = torch.tensor(dW_last)
    g64 = grid.cpu().numpy()
    rep = pyoracle.replay(lambda t, y: TH * (MU - y), "heun", list(zip(g64[:-1], g64[1:])),
                          torch.tensor(y0n), [], g_dw=lambda n, t, y: SG * dWt[n], save_t1=True)
    (rep[-1] * torch.tensor(wr[:, 0])).sum().backward()
    out["forward_grad"]["grad_normwise_err_vs_fp64_replay_last512"] = max(
        float((a.cpu() - r).abs().max() / r.abs().max())
        for a, r in ((thb, TH.grad), (mub, MU.grad), (sgb, SG.grad)))
    if with_cpu:
        cores = min(len(os.sched_getaffinity(0)), 16)
        dWc = dW[:, :cpu_paths].double().cpu().numpy()
        times = []
        while (sum(times) < 10.0 and len(times) < 100) or len(times) < 2:
            t = time.perf_counter()
            oracle.solve(np.zeros((cpu_paths, d)), 1.0, np.concatenate([theta, mu, sigma]),
                         dW=dWc, solver="heun", vf="ou", controller="constant", dt0=dt,
                         save_t1=True, max_steps=100000, nthreads=cores)
            times.append(time.perf_counter() - t)
        med = statistics.median(times)
        out["cpu_baseline"] = dict(value=round(cpu_paths / med, 1), unit="paths/s (forward)",
                                   cores=cores, kind="port",
                                   sample=f"{cpu_paths} of the {P} paths (all 1000 steps), FP64 C "
                                          f"oracle, OpenMP {cores} threads, median of "
                                          f"{len(times)} runs ({med:.3f} s each, "
                                          f"{sum(times):.1f} s in total)")
    del dW, fwd, gs
    torch.cuda.empty_cache()
    return {"ou_sde_heun": out}


# Discrete adjoint of one accepted Tsit5 step on Van der Pol (d = 2), algorithmic flop:
# stage recompute 2*(42 + 7) + 7*5 = 133; y1 transpose 2*(12 + 1) = 26; per stage
# w = control*k_bar (2), vjp (14), accumulations (2 + 1) over 7 stages = 133; a-transpose
# 2*2*21 = 84: 376 per accepted step.  Per save: Tsit5 weights (35) + 2*(1 + 14) = 65.
# The checkpoint replay repeats the forward's attempted steps (FLOP_PER_STEP each, no
# saves) before the sweep.
ADJ_FLOP_PER_STEP = 376
ADJ_FLOP_PER_SAVE = 65


def bench_vdp_grad(mu, y0, t1, ts, steps, warmup, with_cpu, forms=("f64", "f32"), check=512):
    """The headline solve plus its gradient: the discrete adjoint (RecursiveCheckpointAdjoint,
    the default adjoint) of loss = mean(ys^2) with respect to y0 and mu, for all 65,536
    instances.  One step = forward with sqrt(max_steps)-spaced checkpoints + the cotangent
    + the adjoint (segment replays from the checkpoints and reverse sweeps), replayed from
    one HIP graph."""
    import ctypes
    import numpy as np
    import torch
    from diffrax_amd import _lib
    import oracle
    lib = _lib.load()
    dev = torch.device("cuda", 0)
    B = y0.shape[0]
    cores = min(len(os.sched_getaffinity(0)), 16)
    out = {}
    ref = None
    sub = slice(0, check)
    kw = dict(solver="tsit5", vf="vdp", dt0=0.1, rtol=1e-6, atol=1e-9, max_steps=16384)
    # FP64 oracle on a subset: forward with recorded steps, then its discrete adjoint
    rf = oracle.solve(y0[sub], t1[sub], mu[sub, None], ts=ts[sub], record_grid=True,
                      nthreads=cores, **kw)
    n_el = float(B * ts.shape[1] * 2)
    ryb, rqb = oracle.adjoint(t1[sub], mu[sub, None], rf, rf["ys"] * (2.0 / n_el), ts=ts[sub],
                              nthreads=cores, **kw)
    for form in forms:
        dtype = torch.float64 if form == "f64" else torch.float32
        plan = prepare(dtype, mu, y0, t1, ts)
        plan.record_checkpoints()           # sqrt(max_steps) checkpoints, the default
        T = plan.t
        ys = T["ys"]
        gys = torch.empty_like(ys)
        y0_bar = torch.empty((B, 2), dtype=dtype, device=dev)
        mu_bar = torch.empty((B, 1), dtype=dtype, device=dev)
        ws = torch.empty(plan.adjoint_workspace_bytes(), dtype=torch.uint8, device=dev)
        extra_bytes = T["ckpt"].numel() * 8 + ws.numel()

        def adjoint():
            plan.adjoint(gys, y0_bar, mu_bar, workspace=ws)

        def fwd_grad():
            plan.launch()
            torch.mul(ys, 2.0 / n_el, out=gys)      # d mean(ys^2) / d ys
            adjoint()

        g_all = _graphed(fwd_grad)
        k_mean, k_med, wall = _event_time(g_all.replay, steps, warmup)
        g_fwd = _graphed(plan.launch)
        f_mean, _, _ = _event_time(g_fwd.replay, steps, warmup)
        g_adj = _graphed(adjoint)
        a_mean, a_med, _ = _event_time(g_adj.replay, steps, warmup)
        g_all.replay()
        torch.cuda.synchronize()
        st = T["stats"].cpu().numpy()
        acc = int(st[:, 1].sum())
        saves = int(B * ts.shape[1])
        attempted = int(st[:, 0].sum())
        flop = attempted * FLOP_PER_STEP + acc * ADJ_FLOP_PER_STEP + saves * ADJ_FLOP_PER_SAVE
        yb = y0_bar[sub].double().cpu().numpy()
        qb = mu_bar[sub].double().cpu().numpy()
        o = dict(solves_per_s=round(B / (wall * 1e-3), 1), ms_per_step=round(wall, 4),
                 event_ms_mean=round(k_mean, 4), event_ms_median=round(k_med, 4),
                 forward_recording_ms=round(f_mean, 4), adjoint_kernel_ms=round(a_mean, 4),
                 checkpoint_every=int(plan.problem.ckpt_every),
                 gradient_memory_bytes=int(extra_bytes),
                 accepted_steps=acc, adjoint_algorithmic_flop=flop,
                 adjoint_achieved_tflops=round(flop / (a_mean * 1e-3) / 1e12, 4),
                 adjoint_frac_of_peak=round(flop / (a_mean * 1e-3) / 1e12 /
                                            (FP64_PEAK_TFLOPS if form == "f64"
                                             else FP32_PEAK_TFLOPS), 5),
                 grad_normwise_err_vs_fp64_oracle_subset=max(
                     float(np.max(np.abs(yb - ryb)) / np.max(np.abs(ryb))),
                     float(np.max(np.abs(qb - rqb)) / np.max(np.abs(rqb)))))
        if form == "f64":
            o["grad_bit_identical_to_fp64_oracle_subset"] = bool(
                np.array_equal(yb, ryb) and np.array_equal(qb, rqb))
        out[form] = o
        T.clear()
        del g_all, g_fwd, g_adj, gys, y0_bar, mu_bar, ws
        torch.cuda.empty_cache()
    res = dict(workload="65,536 uneven Van der Pol, Tsit5 + PID(1e-6, 1e-9), 64 saves, "
                        "forward + discrete-adjoint gradient of mean(ys^2) for y0 and mu",
               check_subset=check, forms=out)
    if with_cpu:
        n_cpu = 4096
        csub = slice(0, n_cpu)
        ckw = dict(kw, max_steps=1024)   # no instance takes more than 869 steps
        times = []
        while (sum(times) < 10.0 and len(times) < 500) or len(times) < 3:
            t = time.perf_counter()
            f = oracle.solve(y0[csub], t1[csub], mu[csub, None], ts=ts[csub], record_grid=True,
                             nthreads=cores, **ckw)
            oracle.adjoint(t1[csub], mu[csub, None], f, f["ys"] * (2.0 / n_el), ts=ts[csub],
                           nthreads=cores, **ckw)
            times.append(time.perf_counter() - t)
        med = statistics.median(times)
        res["cpu_baseline"] = dict(
            value=round(n_cpu / med, 1), unit="solves/s (forward + gradient)", cores=cores,
            kind="port",
            sample=f"first {n_cpu} of the {B} instances, FP64 C oracle forward with recorded "
                   f"steps + its discrete adjoint, OpenMP {cores} threads, max_steps 1024 (no "
                   f"instance needs more than 869), median of {len(times)} runs "
                   f"({med:.3f} s each, {sum(times):.1f} s in total)")
    return {"vdp_forward_grad": res}


def read_traffic(kernel_tag):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if present."""
    path = os.path.join(ROOT, "profiles", f"{kernel_tag}_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def sweep_batch(steps, warmup):
    """The headline solve at 1, 2 and 4 waves per SIMD (1,024 SIMDs x 64 lanes = 65,536).
    With one wave per SIMD nothing hides the latency of a step's dependent FP64 chain; the
    time for 2x and 4x the instances shows how much issue capacity one wave leaves idle."""
    import torch
    from diffrax_amd import _autotune
    base = None
    for B in (65536, 131072, 262144):
        mu, y0, t1, ts = make_inputs(B=B)
        for block in (64, 256):
            plan = prepare(torch.float64, mu, y0, t1, ts)
            tune = _autotune.tune
            _autotune.tune = lambda p, verbose=False, c=(block, B): (setattr(p, "launch_cfg", c) or (c, {}))
            try:
                r = time_form(plan, steps, warmup)
            finally:
                _autotune.tune = tune
            st = plan.t["stats"].cpu().numpy()
            if base is None:
                base = r["kernel_ms_mean"]
            print(json.dumps({"batch": B, "waves_per_simd": B // 65536, "block": block,
                              "kernel_ms_mean": round(r["kernel_ms_mean"], 4),
                              "solves_per_s": round(B / (r["kernel_ms_mean"] * 1e-3), 1),
                              "time_vs_one_wave": round(r["kernel_ms_mean"] / base, 3),
                              "max_steps_one_trajectory": int(st[:, 0].max()),
                              "mean_steps": round(float(st[:, 0].mean()), 2)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--forms", default="f64,f32", help="device forms to time")
    ap.add_argument("--workloads", default="grad,mlp,sde",
                    help="extra workloads to time besides the headline (grad, mlp, sde; '' for "
                         "none)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (PMC byte-counter scale)")
    ap.add_argument("--sweep-batch", action="store_true",
                    help="only time the FP64 Van der Pol solve at 1, 2 and 4 waves per SIMD "
                         "(65,536 / 131,072 / 262,144 instances), one JSON line per batch")
    args = ap.parse_args()
    if args.sweep_batch:
        sweep_batch(args.steps, args.warmup)
        return
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

    # Matched-error view on a subset: each form's error against a tight FP64 reference
    # (rtol = atol = 1e-12), next to the FP64 oracle's own error at the workload tolerance.
    sub = 512
    tight = oracle.solve(y0[:sub], t1[:sub], mu[:sub, None], ts=ts[:sub], solver="tsit5",
                         vf="vdp", dt0=0.1, rtol=1e-12, atol=1e-12, max_steps=200000,
                         nthreads=min(len(os.sched_getaffinity(0)), 16))["ys"]
    scale = np.max(np.abs(tight))
    for form in forms:
        forms[form]["normwise_err_vs_tight_fp64_subset"] = float(
            np.max(np.abs(results[form]["ys"][:sub] - tight)) / scale)
    if ref is not None:
        ref_err = float(np.max(np.abs(ref["ys"][:sub] - tight)) / scale)
        for form in forms:
            forms[form]["fp64_oracle_normwise_err_vs_tight_subset"] = ref_err

    # Matched error: the FP64 form at the workload tolerance sets the error (normwise, vs
    # the tight FP64 solve of the subset); the float32-state form is run at the tolerance
    # that matches it.  `value` is the faster of the forms at matched error; the stricter
    # gate (max relative error <= 1e-6 against the plain FP64 solve) is reported beside it.
    matched = None
    if "f64" in forms and "f32" in forms:
        target = forms["f64"]["normwise_err_vs_tight_fp64_subset"]
        matched = {"target_normwise_err": target,
                   "f64": {"rtol": 1e-6, "normwise_err": target,
                           "solves_per_s": round(forms["f64"]["solves_per_s"], 1),
                           "ms_per_step": round(forms["f64"]["ms_per_step"], 4)},
                   "f32": matched_error(mu, y0, t1, ts, tight, target, args.steps,
                                        args.warmup, sub)}
    gate = {f: forms[f]["max_rel_err_vs_fp64"] <= 1e-6 for f in forms}
    eligible = [f for f in forms if gate[f]]
    head = max(eligible, key=lambda f: forms[f]["solves_per_s"]) if eligible else "f64"
    if matched is not None and matched["f32"].get("matched") and \
            matched["f32"]["solves_per_s"] > forms[head]["solves_per_s"]:
        head = "f32"
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
        "matched_error": matched,
        "gate_max_rel_err_le_1e-6_vs_fp64": gate,
    }
    if head == "f32" and matched is not None:
        # the float32-state form at its matched tolerance carries the headline
        line["value"] = matched["f32"]["solves_per_s"]
        line["ms_per_step"] = matched["f32"]["ms_per_step"]
        line["config"]["workload"] = line["config"]["workload"].replace(
            "PID(1e-6, 1e-9)", f"PID({matched['f32']['rtol']:.3g}, {matched['f32']['atol']:.3g})")
    for p in plans.values():
        p.t.clear()
    plans.clear()
    torch.cuda.empty_cache()
    extra = {}
    wl = [w for w in args.workloads.split(",") if w]
    if "grad" in wl:
        extra.update(bench_vdp_grad(mu, y0, t1, ts, args.steps, args.warmup, not args.no_cpu))
    if "mlp" in wl:
        extra.update(bench_mlp(args.steps, args.warmup, not args.no_cpu))
    if "sde" in wl:
        extra.update(bench_sde(max(3, args.steps // 4), 1, not args.no_cpu))
    if extra:
        line["workloads"] = extra
    print(json.dumps(line))


if __name__ == "__main__":
    main()
