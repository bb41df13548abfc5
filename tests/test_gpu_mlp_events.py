"""Terminating events on the neural-ODE kernel: Event with affine conditions
(_event.py:13-56), checked at every step end and, with a root finder, located on the
step's interpolant (_integrate.py:542-633, 1333-1426).

The kernel integrates in float32, so its step sequence is not the FP64 oracle's (module
docstring of test_gpu_mlp.py).  The bars therefore split in two:
- against the same kernel's own solve without the event (the event check does not change
  the steps): without a root finder the solve ends at the first accepted step whose end
  changes the condition's sign, with that knot's state bit for bit; with one it ends
  inside that step, at the dense output's value there bit for bit (same interpolant
  arithmetic), where the condition is within 1e-5 of zero; SaveAt(ts) rows before the
  event are the eventless rows bit for bit and the rest are inf;
- against the FP64 oracle: the root-found event time within 1e-6 of a tight FP64 solve's
  (rtol = atol = 1e-12)."""
import numpy as np
import pytest
import torch

import diffrax_amd as dfx
import oracle
from tests.test_gpu_mlp import ATOL, RTOL, mlp_weights

pytestmark = pytest.mark.gpu

B = 70     # two tiles, the second ragged
E0 = (1.0, [0.1] + [0.0] * 63, -0.5)           # t - 0.5 + 0.1 y0: rises through 0 near t = 0.5
E1 = (1.0, [0.0, -0.1] + [0.0] * 62, -0.4)     # t - 0.4 - 0.1 y1


def _cond(e, boolean=False):
    return dfx.AffineCondition(t=e[0], y=e[1], offset=e[2], boolean=boolean)


def _g(e, t, y):
    part = e[0] * t
    for c, v in zip(e[1], y):
        part += c * float(v)
    return part + e[2]


def _setup(matmul, seed=20):
    W = mlp_weights(seed)
    y0 = np.random.default_rng(seed + 1).normal(size=(B, 64))
    term = dfx.ODETerm(dfx.MLPField(*(torch.tensor(w, dtype=torch.float32) for w in W),
                                    matmul=matmul))
    return W, y0, term


def _solve(term, y0, saveat, event=None, launch=None, max_steps=64):
    return dfx.diffeqsolve(term, dfx.Tsit5(), 0.0, 1.0, 0.1,
                           torch.as_tensor(y0, dtype=torch.float32), saveat=saveat,
                           stepsize_controller=dfx.PIDController(rtol=RTOL, atol=ATOL),
                           event=event, max_steps=max_steps, launch=launch)


def _first_crossing(e, t0y0, knots_t, knots_y):
    """Index of the first knot whose condition sign differs from the one at (t0, y0)."""
    s0 = _g(e, 0.0, t0y0) > 0
    for k in range(len(knots_t)):
        if (_g(e, float(knots_t[k]), knots_y[k]) > 0) != s0:
            return k
    return None


@pytest.mark.parametrize("matmul", ["f32", "bf16"])
def test_event_ends_at_the_crossing_step(cuda, matmul):
    W, y0, term = _setup(matmul)
    y0f = y0.astype(np.float32)
    steps = _solve(term, y0, dfx.SaveAt(steps=True, t1=True))
    dense = _solve(term, y0, dfx.SaveAt(t1=True, dense=True))
    ts = np.linspace(0.0, 1.0, 11)
    plain = _solve(term, y0, dfx.SaveAt(ts=ts))
    at_step = _solve(term, y0, dfx.SaveAt(t1=True), dfx.Event(_cond(E0)))
    rooted = _solve(term, y0, dfx.SaveAt(ts=ts, t1=True), dfx.Event(_cond(E0), dfx.Newton()))
    torch.cuda.synchronize()
    acc = steps.stats["num_accepted_steps"].cpu().numpy()
    kt, ky = steps.ts.cpu().numpy(), steps.ys.cpu().numpy()
    for sol in (at_step, rooted):
        assert (sol.result.cpu().numpy() == int(dfx.RESULTS.event_occurred)).all()
        assert sol.event_mask.cpu().numpy().all()
    # the event row is the last finite one (SaveAt(t1) writes at the save index)
    r_ts = rooted.ts.cpu().numpy()
    last = np.isfinite(r_ts).sum(1) - 1
    t_root = r_ts[np.arange(B), last]
    for b in range(B):
        k = _first_crossing(E0, y0f[b], kt[b, :acc[b]], ky[b, :acc[b]])
        assert k is not None
        # no root finder: the crossing step's end, its state and its accepted-step count
        assert float(at_step.ts[b, -1]) == kt[b, k]
        assert np.array_equal(at_step.ys[b, -1].cpu().numpy(), ky[b, k])
        assert int(at_step.stats["num_accepted_steps"][b]) == k + 1
        # root finder: inside that step, on its interpolant, at a zero of the condition
        lo = 0.0 if k == 0 else kt[b, k - 1]
        assert lo < t_root[b] <= kt[b, k]
        want = dense.evaluate(torch.tensor([[t_root[b]]] * B, dtype=torch.float64))[b, 0]
        n = last[b]
        assert torch.equal(rooted.ys[b, n], want)
        assert abs(_g(E0, t_root[b], rooted.ys[b, n].cpu().numpy())) < 1e-5
        # SaveAt(ts): the eventless rows up to the event, inf after
        assert n == int(np.searchsorted(ts, t_root[b], side="right"))
        assert torch.equal(rooted.ys[b, :n], plain.ys[b, :n])
        assert torch.isinf(rooted.ys[b, n + 1:]).all()
    # the oracle's event times at tight tolerance
    q = oracle.mlp_pack(*W)
    y0m = y0f.astype(np.float64)
    ev = [(E0[0], E0[1], E0[2], False, None)]
    kw = dict(solver="tsit5", vf="mlp", hidden=256, dt0=0.1, save_t1=True, events=ev,
              event_root_find=True)
    true = oracle.solve(y0m, 1.0, q, rtol=1e-12, atol=1e-12, max_steps=100000, **kw)
    ora = oracle.solve(y0m, 1.0, q, rtol=RTOL, atol=ATOL, **kw)
    assert true["event_mask"].all()
    e_dev = np.max(np.abs(t_root - true["ts"][:, -1]))
    e_ora = np.max(np.abs(ora["ts"][:, -1] - true["ts"][:, -1]))
    print(f"{matmul}: event time error {e_dev:.2e} (FP64 oracle at the same tolerance "
          f"{e_ora:.2e})")
    assert e_dev < 1e-6      # measured 2.7e-8 (f32), 3.1e-8 (bf16)


@pytest.mark.parametrize("matmul", ["f32", "bf16"])
def test_event_list_masks_and_direction(cuda, matmul):
    """Two conditions: the mask names, of the conditions that change sign in the first step
    where any does, the first in list order (_integrate.py:623-628), and the root find
    follows that one; a direction the condition never crosses in leaves the solve as it
    was, bit for bit; four and eight wavefronts give the same bits."""
    W, y0, term = _setup(matmul, seed=30)
    y0f = y0.astype(np.float32)
    event = dfx.Event([_cond(E0), _cond(E1)], dfx.Newton())
    sol8 = _solve(term, y0, dfx.SaveAt(t1=True), event, launch=8)
    sol4 = _solve(term, y0, dfx.SaveAt(t1=True), event, launch=4)
    one = [_solve(term, y0, dfx.SaveAt(t1=True), dfx.Event(_cond(e), dfx.Newton()))
           for e in (E0, E1)]
    steps = _solve(term, y0, dfx.SaveAt(steps=True, t1=True))
    torch.cuda.synchronize()
    assert torch.equal(sol4.ys, sol8.ys) and torch.equal(sol4.ts, sol8.ts)
    mask = np.stack([m.cpu().numpy() for m in sol4.event_mask], axis=-1)
    acc = steps.stats["num_accepted_steps"].cpu().numpy()
    kt, ky = steps.ts.cpu().numpy(), steps.ys.cpu().numpy()
    for b in range(B):
        ks = [_first_crossing(e, y0f[b], kt[b, :acc[b]], ky[b, :acc[b]]) for e in (E0, E1)]
        k = min(x for x in ks if x is not None)
        first = 0 if ks[0] == k else 1
        assert mask[b].tolist() == [first == 0, first == 1]
        assert torch.equal(sol4.ts[b], one[first].ts[b]) and torch.equal(sol4.ys[b], one[first].ys[b])
    assert mask[:, 0].any() and mask[:, 1].any()
    # direction: E0 only rises, so a falling-only event never triggers
    down = _solve(term, y0, dfx.SaveAt(t1=True), dfx.Event(_cond(E0), direction=False))
    plain = _solve(term, y0, dfx.SaveAt(t1=True))
    torch.cuda.synchronize()
    assert (down.result.cpu().numpy() == 0).all() and not down.event_mask.cpu().numpy().any()
    assert torch.equal(down.ys, plain.ys)
    assert torch.equal(down.stats["num_steps"], plain.stats["num_steps"])


def test_boolean_condition(cuda):
    """A boolean condition ends the solve at the first accepted step end where it holds, or
    at t0 if it holds there.  Without a root finder the state is that knot's; with one it
    is the step's interpolant at its end (_integrate.py:730-736, 774-778).  Gradients
    through an evented neural-ODE solve raise."""
    W, y0, term = _setup("f32", seed=40)
    steps = _solve(term, y0, dfx.SaveAt(steps=True, t1=True))
    dense = _solve(term, y0, dfx.SaveAt(t1=True, dense=True))
    late = (1.0, [0.0] * 64, -0.45)                 # t > 0.45
    sols = [_solve(term, y0, dfx.SaveAt(t1=True), dfx.Event(_cond(late, True), rf))
            for rf in (None, dfx.Newton())]
    torch.cuda.synchronize()
    acc = steps.stats["num_accepted_steps"].cpu().numpy()
    for b in range(B):
        kt = steps.ts[b, :acc[b]].cpu().numpy()
        k = int(np.argmax(kt > 0.45))
        for sol in sols:
            assert float(sol.ts[b, -1]) == kt[k]
        assert torch.equal(sols[0].ys[b, -1], steps.ys[b, k])
        end = dense.evaluate(torch.tensor([[kt[k]]] * B, dtype=torch.float64), left=True)
        assert torch.equal(sols[1].ys[b, -1], end[b, 0])
    at_t0 = _solve(term, y0, dfx.SaveAt(t0=True, ts=[0.5], t1=True),
                   dfx.Event(_cond((0.0, [0.0] * 64, 1.0), True)))
    torch.cuda.synchronize()
    assert (at_t0.stats["num_steps"].cpu().numpy() == 0).all()
    assert (at_t0.result.cpu().numpy() == int(dfx.RESULTS.event_occurred)).all()
    y0t = torch.tensor(y0, dtype=torch.float32, device=cuda)
    # rows: t0, then the t1 row at the save index (the event at t0), then the unsaved ts
    assert torch.equal(at_t0.ys[:, 0], y0t) and torch.equal(at_t0.ys[:, 1], y0t)
    assert (at_t0.ts[:, 1].cpu().numpy() == 0.0).all()
    assert torch.isinf(at_t0.ys[:, 2]).all()
    with pytest.raises(NotImplementedError, match="events"):
        _solve(term, torch.tensor(y0, dtype=torch.float32).requires_grad_(True),
               dfx.SaveAt(t1=True), dfx.Event(_cond(E0)))


@pytest.mark.parametrize("matmul", ["f32", "bf16"])
def test_event_in_reverse_time(cuda, matmul):
    """t0 = 1 > t1 = 0: conditions see direction-normalised time, as the small-system kernel
    and the oracle; the root-found event time within 1e-6 of a tight FP64 oracle solve's."""
    W, y0, term = _setup(matmul, seed=70)
    cond = (-1.0, [0.1] + [0.0] * 63, -0.5)   # -t_norm - 0.5 + 0.1 y0 = t - 0.5 + 0.1 y0
    sol = dfx.diffeqsolve(term, dfx.Tsit5(), 1.0, 0.0, -0.1,
                          torch.tensor(y0, dtype=torch.float32), saveat=dfx.SaveAt(t1=True),
                          stepsize_controller=dfx.PIDController(rtol=RTOL, atol=ATOL),
                          event=dfx.Event(_cond(cond), dfx.Newton()), max_steps=64)
    torch.cuda.synchronize()
    assert (sol.result.cpu().numpy() == int(dfx.RESULTS.event_occurred)).all()
    q = oracle.mlp_pack(*W)
    y0m = y0.astype(np.float32).astype(np.float64)
    true = oracle.solve(y0m, 0.0, q, t0=1.0, solver="tsit5", vf="mlp", hidden=256, dt0=-0.1,
                        save_t1=True, rtol=1e-12, atol=1e-12, max_steps=100000,
                        events=[(*cond, False, None)], event_root_find=True)
    assert true["event_mask"].all()
    t_dev = sol.ts[:, -1].cpu().numpy()
    err = np.max(np.abs(t_dev - true["ts"][:, -1]))
    print(f"{matmul}: reverse-time event time error {err:.2e}")
    assert np.all((t_dev > 0.0) & (t_dev < 1.0)) and err < 1e-6


def test_event_with_step_saves(cuda):
    """SaveAt(steps=True, t1=True) with an event: without a root finder the rows are the
    knots up to and including the crossing step's, with no separate t1 row (the last step
    was saved, _integrate.py:847-877); with one the crossing step's knot, after the event
    time, is dropped and the t1 row records the event state (_integrate.py:781-806)."""
    W, y0, term = _setup("f32", seed=80)
    y0f = y0.astype(np.float32)
    steps = _solve(term, y0, dfx.SaveAt(steps=True, t1=True))
    plain = _solve(term, y0, dfx.SaveAt(steps=True, t1=True), dfx.Event(_cond(E0)))
    root = _solve(term, y0, dfx.SaveAt(steps=True, t1=True), dfx.Event(_cond(E0), dfx.Newton()))
    one = _solve(term, y0, dfx.SaveAt(t1=True), dfx.Event(_cond(E0), dfx.Newton()))
    torch.cuda.synchronize()
    acc = steps.stats["num_accepted_steps"].cpu().numpy()
    kt, ky = steps.ts.cpu().numpy(), steps.ys.cpu().numpy()
    for b in range(B):
        k = _first_crossing(E0, y0f[b], kt[b, :acc[b]], ky[b, :acc[b]])
        assert torch.equal(plain.ts[b, :k + 1], steps.ts[b, :k + 1])
        assert torch.equal(plain.ys[b, :k + 1], steps.ys[b, :k + 1])
        assert torch.isinf(plain.ts[b, k + 1:]).all()
        assert torch.equal(root.ts[b, :k], steps.ts[b, :k])
        assert torch.equal(root.ts[b, k], one.ts[b, 0]) and torch.equal(root.ys[b, k], one.ys[b, 0])
        assert torch.isinf(root.ts[b, k + 1:]).all()
