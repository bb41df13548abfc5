"""Solver / controller state export and resume on the small-system kernel
(SaveAt(solver_state=True, controller_state=True, made_jump=True), _integrate.py:1489-1500;
diffeqsolve(solver_state=, controller_state=, made_jump=), 1252-1270, 1309).

Bars: the exported states and the resumed solves equal the Python oracle's (whose resume
semantics tests/test_oracle_resume.py pins) bit for bit, except with non-default PID
exponents: there the two sides' pow differ in the last bit, and over tens of steps with
rejections on Van der Pol the states drift apart by up to 1.5e-10 (measured), so those
cases are held to 1e-8 relative (floor 1e-8) with equal step counts; a solve
cut short by max_steps and resumed on the GPU from its exported state continues the uncut
GPU solve bit for bit."""
import numpy as np
import pytest
import torch

import diffrax_amd as dfx
from oracle import pyoracle

pytestmark = pytest.mark.gpu

SOLVERS = {"tsit5": dfx.Tsit5(), "dopri5": dfx.Dopri5(), "heun": dfx.Heun()}
CASES = [
    ("tsit5", dict(rtol=1e-6, atol=1e-9)),
    ("tsit5", dict(rtol=1e-6, atol=1e-9, pcoeff=0.3, icoeff=0.5, dcoeff=0.1)),
    ("dopri5", dict(rtol=1e-7, atol=1e-9, pcoeff=0.2, icoeff=0.4)),
    ("heun", dict(rtol=1e-4, atol=1e-6)),
    ("tsit5", None),    # ConstantStepSize
    ("heun", None),
]
IDS = ["tsit5-I", "tsit5-PID", "dopri5-PI", "heun-I", "tsit5-const", "heun-const"]
T1, DT0 = 6.0, 0.05


def _inputs(B=5, seed=4):
    rng = np.random.default_rng(seed)
    mu = np.exp(rng.uniform(np.log(0.3), np.log(3.0), size=B))
    y0 = rng.uniform(-2.0, 2.0, size=(B, 2))
    return mu, y0


def _f(mu):
    return lambda t, y: np.array([y[1], mu * (1 - y[0] * y[0]) * y[1] - y[0]])


def _ctrl(kw):
    return dfx.ConstantStepSize() if kw is None else dfx.PIDController(**kw)


def _okw(kw):
    return dict(controller="constant") if kw is None else dict(controller="pid", **kw)


def _generic(kw):
    return kw is not None and (kw.get("pcoeff", 0) or kw.get("dcoeff", 0))


def _close(a, b, exact, rtol=1e-8):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    if exact:
        return np.array_equal(a, b)
    return bool(np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), 1e-8)))


ALL_STATE = dict(solver_state=True, controller_state=True, made_jump=True)


def _gpu(solver, kw, t0, t1, dt0, y0, mu, max_steps, **resume):
    return dfx.diffeqsolve(dfx.ODETerm(dfx.VanDerPol()), SOLVERS[solver], t0, t1, dt0,
                           torch.tensor(y0), mu, saveat=dfx.SaveAt(t1=True, **ALL_STATE),
                           stepsize_controller=_ctrl(kw), max_steps=max_steps, throw=False,
                           **resume)


def _np(x):
    return x.cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)


def _check_states(sol, b, ref, exact, batched=True):
    pick = (lambda v: _np(v)[b]) if batched else _np
    if ref["solver_state"] is None:
        assert sol.solver_state is None
    else:
        assert bool(pick(sol.solver_state[0])) == ref["solver_state"][0]
        assert _close(pick(sol.solver_state[1]), ref["solver_state"][1], exact)
    cs = tuple(pick(v) for v in sol.controller_state)
    if len(cs) == 3:
        # inverse scaled errors: 1 / norm of a stage difference, whose cancellation
        # magnifies the drift above (1.1e-8 measured; 1e-6 relative there)
        assert _close(cs[:2], ref["controller_state"][:2], exact, rtol=1e-6)
        assert bool(cs[2]) == ref["controller_state"][2]
    else:
        assert tuple(int(v) for v in cs[:2]) == ref["controller_state"][:2]
        assert tuple(float(v) for v in cs[2:]) == ref["controller_state"][2:]
    assert bool(pick(sol.made_jump)) == ref["made_jump"]


@pytest.mark.parametrize("solver,kw", CASES, ids=IDS)
@pytest.mark.parametrize("m", [7, 40])
def test_exported_state_matches_oracle(solver, kw, m):
    mu, y0 = _inputs()
    sol = _gpu(solver, kw, 0.0, T1, DT0, y0, torch.tensor(mu), m)
    torch.cuda.synchronize()
    exact = not _generic(kw)
    for b in range(len(mu)):
        ref = pyoracle.solve(_f(mu[b]), solver, 0.0, T1, DT0, y0[b], [], save_t1=True,
                             max_steps=m, **_okw(kw))
        got = _np(sol.ys)[b, -1]
        assert _close(got, ref["ys"][-1], exact), (b, got - ref["ys"][-1])
        assert tuple(int(_np(sol.stats[k])[b]) for k in
                     ("num_steps", "num_accepted_steps", "num_rejected_steps")) == ref["stats"]
        assert _close(_np(sol.ts)[b, -1], ref["ts"][-1], exact)
        _check_states(sol, b, ref, exact)


@pytest.mark.parametrize("solver,kw", [c for c in CASES if not _generic(c[1])],
                         ids=[i for c, i in zip(CASES, IDS) if not _generic(c[1])])
def test_interrupted_and_resumed_continues_the_uncut_solve(solver, kw):
    mu, y0 = _inputs()
    m = 23
    full = _gpu(solver, kw, 0.0, T1, DT0, y0, torch.tensor(mu), 100000)
    part = _gpu(solver, kw, 0.0, T1, DT0, y0, torch.tensor(mu), m)
    torch.cuda.synchronize()
    for b in range(len(mu)):
        # the next step's end comes from the oracle, whose interrupted solve equals the
        # kernel's bit for bit (test_exported_state_matches_oracle)
        ref = pyoracle.solve(_f(mu[b]), solver, 0.0, T1, DT0, y0[b], [], save_t1=True,
                             max_steps=m, **_okw(kw))
        t0 = float(_np(part.ts)[b, -1])
        assert t0 == ref["tprev"]
        dt0 = ref["tnext"] - t0
        assert t0 + dt0 == ref["tnext"]
        ss = None if part.solver_state is None else (part.solver_state[0][b],
                                                      part.solver_state[1][b])
        rest = dfx.diffeqsolve(
            dfx.ODETerm(dfx.VanDerPol()), SOLVERS[solver], t0, T1, dt0,
            part.ys[b, -1].clone(), float(mu[b]), saveat=dfx.SaveAt(t1=True, **ALL_STATE),
            stepsize_controller=_ctrl(kw), max_steps=100000, throw=False,
            solver_state=ss, controller_state=tuple(v[b] for v in part.controller_state),
            made_jump=part.made_jump[b])
        torch.cuda.synchronize()
        assert np.array_equal(_np(rest.ys)[-1], _np(full.ys)[b, -1])
        for k in ("num_steps", "num_accepted_steps", "num_rejected_steps"):
            assert int(_np(part.stats[k])[b]) + int(_np(rest.stats[k])) == int(_np(full.stats[k])[b])
        assert _close(_np(rest.controller_state[0]), _np(full.controller_state[0])[b], True)
        if full.solver_state is not None:
            assert np.array_equal(_np(rest.solver_state[1]), _np(full.solver_state[1])[b])


@pytest.mark.parametrize("solver,kw", CASES, ids=IDS)
@pytest.mark.parametrize("dt0", [0.05, None])
def test_batched_resume_matches_oracle(solver, kw, dt0):
    """Solve [0, 2], then resume the whole batch over [2, 6] from the exported states: with
    the controller state and dt0 given the first step is [2, 2 + dt0] unclamped; with dt0
    None the initial step is selected afresh (PID only)."""
    if kw is None and dt0 is None:
        pytest.skip("ConstantStepSize needs dt0")
    mu, y0 = _inputs(seed=9)
    kw2 = None if kw is None else dict(kw, dtmax=0.04)   # a clamp the resume must skip
    first = _gpu(solver, kw2, 0.0, 2.0, DT0, y0, torch.tensor(mu), 100000)
    rest = _gpu(solver, kw2, 2.0, T1, dt0, _np(first.ys)[:, -1], torch.tensor(mu), 100000,
                solver_state=first.solver_state, controller_state=first.controller_state,
                made_jump=first.made_jump)
    torch.cuda.synchronize()
    exact = not _generic(kw)
    for b in range(len(mu)):
        ss = None if first.solver_state is None else (bool(_np(first.solver_state[0])[b]),
                                                       _np(first.solver_state[1])[b])
        cs = tuple(_np(v)[b].item() for v in first.controller_state)
        ref = pyoracle.solve(_f(mu[b]), solver, 2.0, T1, dt0, _np(first.ys)[b, -1], [],
                             save_t1=True, max_steps=100000, solver_state=ss,
                             controller_state=cs, made_jump=bool(_np(first.made_jump)[b]),
                             **_okw(kw2))
        assert _close(_np(rest.ys)[b, -1], ref["ys"][-1], exact)
        got = tuple(int(_np(rest.stats[k])[b]) for k in
                    ("num_steps", "num_accepted_steps", "num_rejected_steps"))
        if exact:
            assert got == tuple(ref["stats"])
        _check_states(rest, b, ref, exact)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_fresh_state_values_change_nothing(dtype):
    """Passing the initial states the reference builds reproduces the fresh solve, in both
    precisions (the float32-state form stores f0 exactly in the float64 record)."""
    mu, y0 = _inputs(B=64, seed=2)
    B = len(mu)
    kw = dict(rtol=1e-6, atol=1e-9)
    run = lambda **r: dfx.diffeqsolve(
        dfx.ODETerm(dfx.VanDerPol()), dfx.Tsit5(), 0.0, 4.0, 0.1, torch.tensor(y0, dtype=dtype),
        torch.tensor(mu, dtype=dtype), saveat=dfx.SaveAt(ts=[1.0, 2.0, 3.0], t1=True),
        stepsize_controller=dfx.PIDController(**kw), max_steps=4096, **r)
    a = run()
    b = run(solver_state=(torch.ones(B, dtype=torch.bool), torch.zeros(B, 2, dtype=dtype)),
            controller_state=(1.0, 1.0, False), made_jump=False)
    torch.cuda.synchronize()
    assert torch.equal(a.ys, b.ys) and torch.equal(a.ts, b.ts)
    assert torch.equal(a.stats["num_steps"], b.stats["num_steps"])
