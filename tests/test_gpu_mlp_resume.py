"""Solver / controller state export and resume on the neural-ODE kernel
(SaveAt(solver_state=True, controller_state=True, made_jump=True), _integrate.py:1489-1500;
diffeqsolve(solver_state=, controller_state=, made_jump=), 1252-1270, 1309), forward only.

The kernel's own arithmetic is the reference here (its float32 steps are not the FP64
oracle's; the oracle's resume semantics are pinned in tests/test_oracle_resume.py and the
small-system kernel's against them in tests/test_gpu_resume.py).  Bars:
- a solve cut short by max_steps and resumed from its exported state, with the uncut
  solve's next step as dt0, continues the uncut solve bit for bit (final state, step
  counts, final controller state);
- the initial states the reference builds, passed explicitly, reproduce the fresh solve
  bit for bit;
- the exported FSAL stage is the vector field at the final state, bit for bit
  (dfx_mlp_eval runs the same tile arithmetic)."""
import numpy as np
import pytest
import torch

import diffrax_amd as dfx
from diffrax_amd._mlp import mlp_eval
from tests.test_gpu_mlp import mlp_weights

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-7, 1e-9
ALL_STATE = dict(solver_state=True, controller_state=True, made_jump=True)


def _setup(matmul, B=16, seed=60):
    W = mlp_weights(seed)
    y0 = np.random.default_rng(seed + 1).normal(size=(B, 64))
    term = dfx.ODETerm(dfx.MLPField(*(torch.tensor(w, dtype=torch.float32) for w in W),
                                    matmul=matmul))
    return term, torch.tensor(y0, dtype=torch.float32)


def _solve(term, y0, t0, dt0, max_steps, saveat, **resume):
    return dfx.diffeqsolve(term, dfx.Tsit5(), t0, 1.0, dt0, y0, saveat=saveat,
                           stepsize_controller=dfx.PIDController(rtol=RTOL, atol=ATOL),
                           max_steps=max_steps, throw=False, **resume)


@pytest.mark.parametrize("matmul", ["f32", "bf16"])
def test_interrupted_and_resumed_continues_the_uncut_solve(cuda, matmul):
    term, y0 = _setup(matmul)
    m = 5
    full = _solve(term, y0, 0.0, 0.05, 4096, dfx.SaveAt(t1=True, **ALL_STATE))
    knots = _solve(term, y0, 0.0, 0.05, 4096, dfx.SaveAt(steps=True))
    part = _solve(term, y0, 0.0, 0.05, m, dfx.SaveAt(t1=True, **ALL_STATE))
    probe = _solve(term, y0, 0.0, 0.05, m + 1, dfx.SaveAt(t1=True))
    torch.cuda.synchronize()
    # trajectories whose first m + 1 attempted steps were all accepted: the uncut solve's
    # next step then ends at its (m + 1)-th knot
    ok = [b for b in range(len(y0)) if int(probe.stats["num_rejected_steps"][b]) == 0]
    assert len(ok) >= 4
    for b in ok[:6]:
        t0 = float(part.ts[b, -1])
        assert t0 == float(knots.ts[b, m - 1])
        tnext = float(knots.ts[b, m])
        dt0 = tnext - t0
        assert t0 + dt0 == tnext
        rest = _solve(term, part.ys[b:b + 1, -1].clone(), t0, dt0, 4096,
                      dfx.SaveAt(t1=True, **ALL_STATE),
                      solver_state=(part.solver_state[0][b:b + 1], part.solver_state[1][b:b + 1]),
                      controller_state=tuple(v[b:b + 1] for v in part.controller_state),
                      made_jump=part.made_jump[b:b + 1])
        torch.cuda.synchronize()
        assert torch.equal(rest.ys[0, -1], full.ys[b, -1]), b
        for k in ("num_steps", "num_accepted_steps", "num_rejected_steps"):
            assert int(part.stats[k][b]) + int(rest.stats[k][0]) == int(full.stats[k][b])
        for x, y in zip(rest.controller_state, full.controller_state):
            assert torch.equal(x[0], y[b])
        assert torch.equal(rest.solver_state[1][0], full.solver_state[1][b])


@pytest.mark.parametrize("matmul", ["f32", "bf16"])
def test_fresh_state_values_and_exported_state(cuda, matmul):
    term, y0 = _setup(matmul, B=40, seed=62)
    B = len(y0)
    ts = [0.25, 0.5, 1.0]
    a = _solve(term, y0, 0.0, 0.1, 4096, dfx.SaveAt(ts=ts))
    b = _solve(term, y0, 0.0, 0.1, 4096, dfx.SaveAt(ts=ts, **ALL_STATE),
               solver_state=(torch.ones(B, dtype=torch.bool), torch.zeros(B, 64)),
               controller_state=(1.0, 1.0, False), made_jump=False)
    torch.cuda.synchronize()
    assert torch.equal(a.ys, b.ys) and torch.equal(a.stats["num_steps"], b.stats["num_steps"])
    first, f0 = b.solver_state
    assert not bool(first.any()) and not bool(b.made_jump.any())
    f_end = mlp_eval(term.vector_field, b.ys[:, -1].contiguous())
    assert torch.equal(f0, f_end)
    inv, pinv, at = b.controller_state
    assert bool((inv > 0).all()) and not bool(at.any())
    with pytest.raises(NotImplementedError, match="state"):
        dfx.diffeqsolve(term, dfx.Tsit5(), 0.0, 1.0, 0.1, y0.clone().requires_grad_(True),
                        saveat=dfx.SaveAt(t1=True, controller_state=True),
                        stepsize_controller=dfx.PIDController(rtol=RTOL, atol=ATOL))
