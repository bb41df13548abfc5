"""GPU tests of ClipStepSizeController(PIDController, step_ts, jump_ts) on the small-system
kernel: the float64 form must equal the FP64 oracle (whose clip logic the CPU tests pin)
bit for bit -- saved states, step statistics and the accepted step grid -- forwards and
backwards, for Tsit5, Dopri5 and Heun, with per-instance end times."""
import numpy as np
import pytest
import torch

import diffrax_amd as dfx
import oracle
from diffrax_amd import _integrate

pytestmark = pytest.mark.gpu

# 2.5 and the float just after it: after stepping to whichever of the two the solve meets
# first, the controller's bump skips the other (clip.py:323-333) -- 2.5 forwards, the
# larger one backwards.
STEP_TS = [1.0, 2.5, np.nextafter(2.5, np.inf), 7.0]
JUMP_TS = [3.3, 8.8]
SOLVERS = {"tsit5": dfx.Tsit5, "dopri5": dfx.Dopri5, "heun": dfx.Heun}


@pytest.mark.parametrize("solver", sorted(SOLVERS))
@pytest.mark.parametrize("backwards", [False, True])
def test_clip_matches_oracle(cuda, solver, backwards):
    B = 48
    rng = np.random.default_rng(21)
    y0 = rng.uniform(0.5, 2.0, size=(B, 2))
    t0 = 10.0 if backwards else 0.0
    t1 = rng.uniform(0.0, 4.0, size=B) if backwards else rng.uniform(6.0, 10.0, size=B)
    q = np.array([1.5, 1.0, 3.0, 1.0])
    rtol = 1e-4 if solver == "heun" else 1e-6   # second order: keep within max_steps
    ctrl = dfx.ClipStepSizeController(dfx.PIDController(rtol=rtol, atol=1e-8),
                                      step_ts=STEP_TS, jump_ts=JUMP_TS)
    plan = _integrate._prepare_erk(
        dfx.ODETerm(dfx.LotkaVolterra()), SOLVERS[solver](), t0, torch.tensor(t1), None,
        torch.tensor(y0), tuple(q), dfx.SaveAt(steps=True), ctrl, 4096, None, record_grid=True)
    plan.launch()
    torch.cuda.synchronize()
    ref = oracle.solve(y0, t1, q, solver=solver, vf="lv", t0=t0, rtol=rtol, atol=1e-8,
                       save_steps=True, record_grid=True, step_ts=STEP_TS, jump_ts=JUMP_TS)
    st = plan.t["stats"].cpu().numpy()
    assert np.array_equal(st, ref["stats"])
    assert (plan.t["result"].cpu().numpy() == 0).all()
    assert np.array_equal(plan.t["ys"].cpu().numpy(), ref["ys"])
    assert np.array_equal(plan.t["ts_out"].cpu().numpy(), ref["ts"])
    g = plan.t["grid"].cpu().numpy()
    for b in range(B):
        n = st[b, 1]
        assert np.array_equal(g[b, :n], ref["grid"][b, :n])
        ends = ref["ts"][b, :n]
        lo, hi = min(t0, t1[b]), max(t0, t1[b])
        for s in (1.0, 7.0):                  # every ordinary step time in the span is hit
            if lo < s < hi:
                assert s in ends
        if lo < 2.5 < hi:
            first, other = STEP_TS[2], STEP_TS[1]
            if not backwards:
                first, other = other, first
            assert first in ends and other not in ends


def test_pid_meta_builds_clip_and_solves(cuda):
    ctrl = dfx.PIDController(rtol=1e-6, atol=1e-8, step_ts=[3.0, 4.0])
    assert isinstance(ctrl, dfx.ClipStepSizeController)
    sol = dfx.diffeqsolve(dfx.ODETerm(dfx.LinearField()), dfx.Dopri5(), 0.0, 5.0, None,
                          torch.tensor([1.0], dtype=torch.float64), torch.tensor([[-0.2]]),
                          saveat=dfx.SaveAt(steps=True), stepsize_controller=ctrl)
    ts = sol.ts.cpu().numpy()
    assert 3.0 in ts and 4.0 in ts
