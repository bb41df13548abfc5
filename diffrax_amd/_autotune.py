"""Launch-shape autotuner for the persistent ERK kernels.

For a prepared plan it times a few candidate launch shapes with HIP events and keeps the
fastest: for the small-system kernel the workgroup size and the number of persistent
trajectory slots (all instances, or half with the refill queue), for the neural-ODE
kernel the wavefronts per 32-trajectory tile (4 or 8; the tile's LDS layout is the same,
the waves split its output columns and state components).  Results are cached in memory
and, if DIFFRAX_AMD_TUNE_CACHE names a file, as JSON keyed by (kernel signature, batch,
device name)."""
from __future__ import annotations

import json
import os

import torch

_CACHE: dict = {}


def _key(plan):
    """Everything that selects a different kernel or changes its work or memory traffic:
    the (solver, field, precision, d) kernel, the controller kind (PID coefficients) and
    tolerances (shared or per instance), the save layout (ts count and sharing, t0 / t1 /
    every-k-th-step rows), max_steps, the dense / clipping / event variant, step and
    checkpoint recording, and the batch and device."""
    p = plan.problem
    io = getattr(plan, "io", None)
    has = lambda f: int(bool(io is not None and getattr(io, f)))
    name = torch.cuda.get_device_name(torch.cuda.current_device())
    pid = f"{p.pcoeff}/{p.icoeff}/{p.dcoeff}/{p.error_order}"
    tol = f"{p.rtol!r}/{p.atol!r}/{has('rtol')}/{has('atol')}"
    save = (f"{p.n_ts}/{int(p.ts_stride != 0)}/{int(p.t1_stride != 0)}/{p.save_t0}/"
            f"{p.save_t1}/{p.save_steps}/{p.max_steps}")
    extras = (f"{p.n_step_ts}/{p.n_jump_ts}/{p.n_events}/{has('dense_ts')}/{has('grid')}/"
              f"{p.ckpt_every}")
    return (f"erk:{p.solver}:{p.vf}:{p.precision}:{p.d}:{p.controller}:{pid}:{tol}:{save}:"
            f"{extras}:{p.batch}:{name}")


def _load_file():
    path = os.environ.get("DIFFRAX_AMD_TUNE_CACHE")
    if path and os.path.exists(path):
        with open(path) as f:
            _CACHE.update(json.load(f))


def _save_file():
    path = os.environ.get("DIFFRAX_AMD_TUNE_CACHE")
    if path:
        with open(path, "w") as f:
            json.dump(_CACHE, f, indent=1)


def candidates(batch):
    out = []
    for block in (64, 128, 256):
        for frac in (1, 2):
            lanes = max(block, (batch // frac) // block * block)
            if frac > 1 and lanes >= batch:
                continue
            out.append((block, lanes))
    return out


def time_launch(launch, reps=3):
    """Mean event time (ms) of reps calls of launch() after one untimed call."""
    launch()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        launch()
    end.record()
    end.synchronize()
    return start.elapsed_time(end) / reps


def time_plan(plan, launch_cfg, reps=3):
    plan.launch_cfg = launch_cfg
    return time_launch(plan.launch, reps)


def tune(plan, verbose=False):
    """Pick the fastest launch shape for this plan; returns (block, lanes) and timings."""
    if not _CACHE:
        _load_file()
    key = _key(plan)
    if key in _CACHE:
        plan.launch_cfg = tuple(_CACHE[key]["best"])
        return plan.launch_cfg, _CACHE[key]["ms"]
    # Shapes often differ by a few percent, about the spread of one timing: each candidate
    # is timed in two interleaved rounds of 5 launches and keeps its faster round.
    timings = {}
    cands = candidates(int(plan.problem.batch))
    for _ in range(2):
        for cfg in cands:
            name = f"{cfg[0]}x{cfg[1]}"
            ms = time_plan(plan, cfg, reps=5)
            timings[name] = min(ms, timings.get(name, ms))
    if verbose:
        for name, ms in timings.items():
            print(f"  tune {name}: {ms:.3f} ms", flush=True)
    best = min(timings, key=timings.get)
    b, l = (int(x) for x in best.split("x"))
    plan.launch_cfg = (b, l)
    _CACHE[key] = {"best": [b, l], "ms": timings}
    _save_file()
    return plan.launch_cfg, timings


def _key_mlp(plan):
    """The neural-ODE kernel variant (precision, controller kind, gradient recording, dt0
    given or selected), its save layout and tolerances, the batch and the device."""
    p = plan.prob
    name = torch.cuda.get_device_name(torch.cuda.current_device())
    pid = f"{p.pcoeff}/{p.icoeff}/{p.dcoeff}/{p.error_order}/{p.rtol!r}/{p.atol!r}"
    return (f"mlp:{p.precision}:{pid}:{int(p.dt0 == p.dt0)}:{p.n_ts}/{int(p.ts_stride != 0)}/"
            f"{p.save_t0}/{p.save_t1}/{p.save_steps}:"
            f"{p.max_steps}:{p.ckpt_every}:{p.t0!r}/{p.t1!r}:{p.batch}:{name}")


def tune_mlp(plan, verbose=False):
    """Pick the faster wave count (4 or 8) for a neural-ODE plan; sets plan.prob.waves and
    returns it with the timings."""
    if not _CACHE:
        _load_file()
    key = _key_mlp(plan)
    if key in _CACHE:
        plan.prob.waves = int(_CACHE[key]["best"])
        return plan.prob.waves, _CACHE[key]["ms"]
    timings = {}
    for _ in range(2):
        for waves in (4, 8):
            plan.prob.waves = waves
            ms = time_launch(plan.launch, reps=5)
            timings[str(waves)] = min(ms, timings.get(str(waves), ms))
    if verbose:
        for name, ms in timings.items():
            print(f"  tune waves {name}: {ms:.3f} ms", flush=True)
    best = min(timings, key=timings.get)
    plan.prob.waves = int(best)
    _CACHE[key] = {"best": int(best), "ms": timings}
    _save_file()
    return plan.prob.waves, timings
