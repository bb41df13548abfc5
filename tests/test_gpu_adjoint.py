"""GPU tests of the discrete adjoint of the small-system solve (the default
RecursiveCheckpointAdjoint, _adjoint.py:264-312): torch.autograd through diffeqsolve,
computed by dfx_erk_adjoint, which replays each segment of accepted steps from the
forward's checkpoints and sweeps it in reverse.

- float64: y0 and args gradients bit-identical to the oracle's discrete adjoint
  (oracle.adjoint) on the oracle's own recorded steps, for every field the kernel covers
  and each solver, with SaveAt t0 + ts + t1, steps, per-instance save times, reverse time,
  step clipping and a zero-length solve (the oracle is pinned against torch autograd of an
  independent replay and against closed forms on the CPU), for checkpoint spacings from
  one accepted step to the whole solve;
- the checkpoints hold the oracle's recorded steps bit for bit (start state, times);
- shared args: the per-instance gradients summed over the batch (1e-13, torch's order);
- float32-state form within 2e-4 relative of the float64 oracle;
- the reference's test_against recipe (test_adjoint.py:34-214) against the closed form at
  its tolerance (1e-5), and within 1e-5 of BacksolveAdjoint, as the reference requires of
  the two adjoints.
"""
import numpy as np
import pytest
import torch
from scipy.linalg import expm

import diffrax_amd as dfx
import oracle

pytestmark = pytest.mark.gpu

SOLVERS = {"tsit5": dfx.Tsit5, "dopri5": dfx.Dopri5, "heun": dfx.Heun}
FIELDS = {"lv": (dfx.LotkaVolterra, 2), "vdp": (dfx.VanDerPol, 2),
          "linear2": (dfx.LinearField, 2), "linear4": (dfx.LinearField, 4)}


def _case(field, B, rng):
    cls, d = FIELDS[field]
    y0 = rng.uniform(0.5, 2.0, size=(B, d))
    if field == "lv":
        q = np.array([1.5, 1.0, 3.0, 1.0]) * rng.uniform(0.8, 1.2, size=(B, 4))
    elif field == "vdp":
        q = rng.uniform(0.2, 2.0, size=(B, 1))
    else:
        q = (rng.normal(size=(B, d, d)) * 0.5 - np.eye(d)).reshape(B, d * d)
    return cls, d, y0, q


def _args(field, qt, d):
    if field == "lv":
        return tuple(qt[:, i] for i in range(4))
    if field == "vdp":
        return qt[:, 0]
    return qt.reshape(-1, d, d)


def _run(field, solver, y0, q, w_fn, *, t0=0.0, t1=3.0, dt0=0.05, saveat, ctrl,
         dtype=torch.float64, cuda=None, adjoint=dfx.RecursiveCheckpointAdjoint()):
    cls, d = FIELDS[field]
    y0t = torch.tensor(y0, dtype=dtype, device=cuda, requires_grad=True)
    qt = torch.tensor(q, dtype=dtype, device=cuda, requires_grad=True)
    sol = dfx.diffeqsolve(dfx.ODETerm(cls()), SOLVERS[solver](), t0, t1, dt0, y0t,
                          _args(field, qt, d), saveat=saveat, stepsize_controller=ctrl,
                          max_steps=2000, adjoint=adjoint)
    w = w_fn(tuple(sol.ys.shape))
    wt = torch.tensor(w, dtype=dtype, device=cuda)
    gy, gq = torch.autograd.grad((sol.ys * torch.nan_to_num(wt)).sum(), (y0t, qt))
    return sol, w, gy.cpu().numpy(), gq.cpu().numpy()


@pytest.mark.parametrize("solver", sorted(SOLVERS))
@pytest.mark.parametrize("field", sorted(FIELDS))
@pytest.mark.parametrize("save", ["t0+ts+t1", "steps", "steps3+ts+t1"])
def test_bitwise_against_oracle(cuda, field, solver, save):
    rng = np.random.default_rng([sorted(FIELDS).index(field), len(solver), len(save)])
    B = 96
    _, d, y0, q = _case(field, B, rng)
    ts = np.array([0.0, 0.5, 1.0, 2.25, 3.0]) if save != "steps" else None
    flags = dict(save_t0=save == "t0+ts+t1", save_t1=save != "steps",
                 save_steps={"steps": 1, "steps3+ts+t1": 3}.get(save, 0))
    saveat = dfx.SaveAt(t0=flags["save_t0"], t1=flags["save_t1"], ts=ts,
                        steps=flags["save_steps"])
    rtol = 1e-4 if solver == "heun" else 1e-6
    ctrl = dfx.PIDController(rtol=rtol, atol=1e-8)
    sol, w, gy, gq = _run(field, solver, y0, q, lambda s: rng.normal(size=s), saveat=saveat,
                          ctrl=ctrl, cuda=cuda)
    kw = dict(solver=solver, vf="linear" if field.startswith("linear") else field, dt0=0.05,
              rtol=rtol, atol=1e-8, max_steps=2000, **flags)
    fwd = oracle.solve(y0, 3.0, q, ts=ts, record_grid=True, **kw)
    assert (fwd["result"] == 0).all()
    ys = sol.ys.detach().cpu().numpy()
    assert np.array_equal(ys, fwd["ys"])
    w = np.where(np.isfinite(fwd["ys"]), w, 0.0)   # steps rows past the last step
    yb, qb = oracle.adjoint(3.0, q, fwd, w, ts=ts, **kw)
    assert np.array_equal(gy, yb)
    assert np.array_equal(gq, qb)


def test_checkpoints_match_oracle_steps(cuda):
    rng = np.random.default_rng(1)
    _, d, y0, q = _case("vdp", 64, rng)
    ts = np.linspace(0.0, 3.0, 7)
    plan = dfx._integrate._prepare_erk(
        dfx.ODETerm(dfx.VanDerPol()), dfx.Tsit5(), 0.0, 3.0, 0.05,
        torch.tensor(y0), torch.tensor(q[:, 0]), dfx.SaveAt(ts=ts),
        dfx.PIDController(rtol=1e-6, atol=1e-8), 2000, None)
    plan.record_grid()
    plan.record_checkpoints(checkpoints=400)     # every 5 accepted steps
    K = int(plan.problem.ckpt_every)
    assert K == 5 and tuple(plan.t["ckpt"].shape) == (64, 400, d + 8)
    plan.launch()
    torch.cuda.synchronize()
    fwd = oracle.solve(y0, 3.0, q, ts=ts, record_grid=True, solver="tsit5", vf="vdp",
                       dt0=0.05, rtol=1e-6, atol=1e-8, max_steps=2000)
    n = plan.t["stats"][:, 1].cpu().numpy()
    assert np.array_equal(n, fwd["stats"][:, 1])
    g = plan.t["grid"].cpu().numpy()
    gy = plan.t["grid_y"].cpu().numpy()
    ck = plan.t["ckpt"].cpu().numpy()
    for b in range(64):
        assert np.array_equal(g[b, :n[b]], fwd["grid"][b, :n[b]])
        assert np.array_equal(gy[b, :n[b]], fwd["grid_y"][b, :n[b]])
        for c in range((n[b] - 1) // K + 1):
            s = c * K
            assert np.array_equal(ck[b, c, :d], fwd["grid_y"][b, s])      # start state
            assert ck[b, c, d] == fwd["grid"][b, s, 0]                   # t_prev
            tl = fwd["grid"][b, s - 1, 1] if c > 0 else -np.inf
            assert ck[b, c, d + 4] == tl                                 # previous step's end
        ints = ck[b, :, d + 5:].copy().view(np.int32)
        assert (ints[0, :] == 0).all()    # no steps, no rejections, indices 0 at c = 0


@pytest.mark.parametrize("checkpoints", [1, 3, 45, 700, 2000])
def test_checkpoint_spacing_bitwise(cuda, checkpoints):
    """Segments of 2,000 (one segment), 667, 45, 3 and 1 accepted steps: the same
    gradients bit for bit."""
    rng = np.random.default_rng(21)
    B = 64
    _, d, y0, q = _case("vdp", B, rng)
    ts = np.array([0.0, 0.7, 1.9, 3.0])
    y0t = torch.tensor(y0, device=cuda, requires_grad=True)
    qt = torch.tensor(q, device=cuda, requires_grad=True)
    sol = dfx.diffeqsolve(dfx.ODETerm(dfx.VanDerPol()), dfx.Dopri5(), 0.0, 3.0, 0.05, y0t,
                          qt[:, 0], saveat=dfx.SaveAt(t0=True, ts=ts, t1=True),
                          stepsize_controller=dfx.PIDController(rtol=1e-7, atol=1e-9),
                          adjoint=dfx.RecursiveCheckpointAdjoint(checkpoints=checkpoints),
                          max_steps=2000)
    w = rng.normal(size=tuple(sol.ys.shape))
    gy, gq = torch.autograd.grad((sol.ys * torch.tensor(w, device=cuda)).sum(), (y0t, qt))
    kw = dict(solver="dopri5", vf="vdp", dt0=0.05, rtol=1e-7, atol=1e-9, max_steps=2000,
              save_t0=True, save_t1=True)
    fwd = oracle.solve(y0, 3.0, q, ts=ts, record_grid=True, **kw)
    yb, qb = oracle.adjoint(3.0, q, fwd, w, ts=ts, **kw)
    assert np.array_equal(gy.cpu().numpy(), yb)
    assert np.array_equal(gq.cpu().numpy(), qb)


@pytest.mark.parametrize("checkpoints", [None, 2000])
def test_reverse_time_per_instance_ts_clip(cuda, checkpoints):
    """Step clipping and jumps; with 2,000 checkpoints every accepted step is a restart
    point, so restarts right after a jump (made_jump) are covered."""
    rng = np.random.default_rng(12)
    B = 64
    _, d, y0, q = _case("lv", B, rng)
    ts = np.sort(rng.uniform(0.0, 5.0, size=(B, 6)), axis=1)[:, ::-1].copy()
    ctrl = dfx.ClipStepSizeController(dfx.PIDController(rtol=1e-6, atol=1e-8),
                                      step_ts=[1.0, 2.0, 3.0], jump_ts=[2.5])
    sol, w, gy, gq = _run("lv", "dopri5", y0, q, lambda s: rng.normal(size=s), t0=5.0, t1=0.0,
                          dt0=-0.1, saveat=dfx.SaveAt(ts=ts, t1=True), ctrl=ctrl, cuda=cuda,
                          adjoint=dfx.RecursiveCheckpointAdjoint(checkpoints=checkpoints))
    kw = dict(solver="dopri5", vf="lv", t0=5.0, dt0=-0.1, rtol=1e-6, atol=1e-8, save_t1=True,
              step_ts=[1.0, 2.0, 3.0], jump_ts=[2.5], max_steps=2000)
    fwd = oracle.solve(y0, 0.0, q, ts=ts, record_grid=True, **kw)
    assert np.array_equal(sol.ys.detach().cpu().numpy(), fwd["ys"])
    yb, qb = oracle.adjoint(0.0, q, fwd, w, ts=ts, **kw)
    assert np.array_equal(gy, yb) and np.array_equal(gq, qb)


def test_zero_length_and_shared_args(cuda):
    rng = np.random.default_rng(3)
    B = 50
    y0 = rng.uniform(0.5, 2.0, size=(B, 2))
    q = np.array([1.5, 1.0, 3.0, 1.0])
    # zero-length: rows are y0
    y0t = torch.tensor(y0, device=cuda, requires_grad=True)
    sol = dfx.diffeqsolve(dfx.ODETerm(dfx.LotkaVolterra()), dfx.Tsit5(), 1.0, 1.0, 0.1, y0t,
                          tuple(q), saveat=dfx.SaveAt(t0=True, ts=[1.0, 1.0], t1=True))
    gy, = torch.autograd.grad((sol.ys * torch.arange(8.0, device=cuda).reshape(4, 2)).sum(),
                              (y0t,))
    assert torch.equal(gy.cpu(), torch.tensor([[12.0, 16.0]], dtype=torch.float64).expand(B, 2))
    # shared args: the gradient is the sum of the per-instance ones
    qt = torch.tensor(q, device=cuda, requires_grad=True)
    ts = np.array([1.0, 2.0, 4.0])
    sol = dfx.diffeqsolve(dfx.ODETerm(dfx.LotkaVolterra()), dfx.Tsit5(), 0.0, 4.0, 0.1,
                          torch.tensor(y0, device=cuda), tuple(qt[i] for i in range(4)),
                          saveat=dfx.SaveAt(ts=ts),
                          stepsize_controller=dfx.PIDController(rtol=1e-6, atol=1e-8))
    gq, = torch.autograd.grad(sol.ys.sum(), (qt,))
    kw = dict(solver="tsit5", vf="lv", dt0=0.1, rtol=1e-6, atol=1e-8)
    fwd = oracle.solve(y0, 4.0, q, ts=ts, record_grid=True, **kw)
    _, qb = oracle.adjoint(4.0, q, fwd, np.ones_like(fwd["ys"]), ts=ts, **kw)
    np.testing.assert_allclose(gq.cpu().numpy(), qb.sum(0), rtol=1e-13)


def test_float32_state_form(cuda):
    rng = np.random.default_rng(8)
    B = 128
    _, d, y0, q = _case("lv", B, rng)
    y0 = y0.astype(np.float32).astype(np.float64)
    q = q.astype(np.float32).astype(np.float64)
    ts = np.array([1.0, 2.0, 3.0])
    w = rng.normal(size=(B, 3, 2)).astype(np.float32).astype(np.float64)
    _, _, gy, gq = _run("lv", "tsit5", y0, q, lambda s: w, saveat=dfx.SaveAt(ts=ts),
                        ctrl=dfx.PIDController(rtol=1e-6, atol=1e-8), dtype=torch.float32,
                        cuda=cuda)
    kw = dict(solver="tsit5", vf="lv", dt0=0.05, rtol=1e-6, atol=1e-8, max_steps=2000)
    fwd = oracle.solve(y0, 3.0, q, ts=ts, record_grid=True, **kw)
    yb, qb = oracle.adjoint(3.0, q, fwd, w, ts=ts, **kw)
    # float32 state (unit roundoff 6e-8) over ~100 Lotka-Volterra steps, whose float32
    # accepted grid also differs from the FP64 one: measured 1.07e-4 (2e-4 bound)
    assert oracle.max_rel_err(gy, yb, 1.0) <= 2e-4
    assert oracle.max_rel_err(gq, qb, 1.0) <= 2e-4


def test_against_closed_form_and_backsolve(cuda):
    A = np.array([[0.1, -1.0], [1.0, -0.1]])
    y0 = np.array([0.9, 5.4])
    for st0 in (True, False):
        for st1 in (True, False):
            for ts in (None, [0.3], [2.0], [9.5], [1.0, 7.0], [0.3, 7.0, 9.5]):
                if not st0 and not st1 and ts is None:
                    continue
                grads = []
                for adj in (dfx.RecursiveCheckpointAdjoint(), dfx.BacksolveAdjoint()):
                    y0t = torch.tensor(y0, device=cuda, requires_grad=True)
                    At = torch.tensor(A, device=cuda, requires_grad=True)
                    sol = dfx.diffeqsolve(
                        dfx.ODETerm(dfx.LinearField()), dfx.Tsit5(), 0.3, 9.5, None, y0t, At,
                        saveat=dfx.SaveAt(t0=st0, t1=st1, ts=ts), adjoint=adj,
                        stepsize_controller=dfx.PIDController(rtol=1e-8, atol=1e-8))
                    grads.append([g.cpu().numpy() for g in
                                  torch.autograd.grad(sol.ys.sum(), (y0t, At))])
                ey, eA = np.zeros(2), np.zeros((2, 2))
                for tk in sol.ts.cpu().numpy():
                    dt = tk - 0.3
                    ey += expm(A.T * dt) @ np.ones(2)
                    M = np.zeros((4, 4))
                    M[:2, :2] = M[2:, 2:] = A.T
                    M[:2, 2:] = np.outer(np.ones(2), y0)
                    eA += expm(M * dt)[:2, 2:]
                (gy, gA), (by, bA) = grads
                np.testing.assert_allclose(gy, ey, rtol=1e-5, atol=1e-5)
                np.testing.assert_allclose(gA, eA, rtol=1e-5, atol=1e-5)
                np.testing.assert_allclose(gy, by, rtol=1e-5, atol=1e-5)
                np.testing.assert_allclose(gA, bA, rtol=1e-5, atol=1e-5)
