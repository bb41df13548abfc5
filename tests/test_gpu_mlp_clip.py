"""ClipStepSizeController(PIDController, step_ts, jump_ts) on the neural-ODE kernel
(clip.py:120-428; PIDController(step_ts=, jump_ts=) builds the same).

The kernel integrates in float32, so its steps are not the FP64 oracle's (module docstring
of test_gpu_mlp.py).  Bars:
- the clipping itself, exactly: every step_ts time inside (t0, t1) is an accepted step's
  end, and after every jump_ts time j the next step starts at nextup(j) (the saved step
  time, clip.py:82-101); times outside (t0, t1) change nothing, bit for bit;
- accuracy at matched tolerance: the rtol-scaled error against a tight FP64 solve at most
  the FP64 oracle's with the same clipping, plus the kernel's float32 floor (the largest
  of its errors at rtol 1e-7, 1e-8, 1e-9), plus one unit;
- four and eight wavefronts, and both ways of building the controller, give the same bits;
- gradients (RecursiveCheckpointAdjoint) against torch float64 autograd through a replay of
  the same clipped accepted-step grid, 1e-4 relative as for the unclipped solve
  (test_gpu_mlp_grad.py); BacksolveAdjoint with clipping raises."""
import numpy as np
import pytest
import torch

import diffrax_amd as dfx
import oracle
from tests.test_gpu_mlp import ATOL, RTOL, _scaled, mlp_weights

pytestmark = pytest.mark.gpu

B = 70
STEP_TS = [0.25, 0.5, 0.75]
JUMP_TS = [0.4]


def _setup(matmul, seed=50):
    W = mlp_weights(seed)
    y0 = np.random.default_rng(seed + 1).normal(size=(B, 64))
    term = dfx.ODETerm(dfx.MLPField(*(torch.tensor(w, dtype=torch.float32) for w in W),
                                    matmul=matmul))
    return W, y0, term


def _ctrl(rtol=RTOL, atol=ATOL, step_ts=STEP_TS, jump_ts=JUMP_TS):
    return dfx.ClipStepSizeController(dfx.PIDController(rtol=rtol, atol=atol), step_ts=step_ts,
                                      jump_ts=jump_ts)


def _solve(term, y0, saveat, ctrl, launch=None, dt0=0.1):
    return dfx.diffeqsolve(term, dfx.Tsit5(), 0.0, 1.0, dt0,
                           torch.as_tensor(y0, dtype=torch.float32), saveat=saveat,
                           stepsize_controller=ctrl, max_steps=4096, launch=launch)


@pytest.mark.parametrize("matmul", ["f32", "bf16"])
def test_clipped_steps_land_on_the_times(cuda, matmul):
    W, y0, term = _setup(matmul)
    for dt0 in (0.1, None):
        st = _solve(term, y0, dfx.SaveAt(steps=True), _ctrl(), dt0=dt0)
        torch.cuda.synchronize()
        acc = st.stats["num_accepted_steps"].cpu().numpy()
        kt = st.ts.cpu().numpy()
        for b in range(B):
            knots = set(kt[b, :acc[b]].tolist())
            for s in STEP_TS:
                assert s in knots, (b, s)
            for j in JUMP_TS:
                assert np.nextafter(j, np.inf) in knots and j not in knots, (b, j)
            assert kt[b, acc[b] - 1] == 1.0
    # times outside (t0, t1) leave the solve as it was, bit for bit
    plain = _solve(term, y0, dfx.SaveAt(ts=np.linspace(0, 1, 5)),
                   dfx.PIDController(rtol=RTOL, atol=ATOL))
    out = _solve(term, y0, dfx.SaveAt(ts=np.linspace(0, 1, 5)),
                 _ctrl(step_ts=[-1.0, 2.0], jump_ts=[3.0]))
    torch.cuda.synchronize()
    assert torch.equal(plain.ys, out.ys)
    assert torch.equal(plain.stats["num_steps"], out.stats["num_steps"])


@pytest.mark.parametrize("matmul", ["f32", "bf16"])
def test_clipped_accuracy_matched_tolerance(cuda, matmul):
    W, y0, term = _setup(matmul, seed=52)
    ts = np.linspace(0.0, 1.0, 9)
    sol = _solve(term, y0, dfx.SaveAt(ts=ts), _ctrl())
    q = oracle.mlp_pack(*W)
    y0m = y0.astype(np.float32).astype(np.float64)
    kw = dict(ts=ts, solver="tsit5", vf="mlp", hidden=256, dt0=0.1)
    true = oracle.solve(y0m, 1.0, q, rtol=1e-12, atol=1e-12, max_steps=100000, **kw)["ys"]
    ref = oracle.solve(y0m, 1.0, q, rtol=RTOL, atol=ATOL, step_ts=STEP_TS, jump_ts=JUMP_TS, **kw)
    e_floor = 0.0
    for rt in (1e-7, 1e-8, 1e-9):
        fl = _solve(term, y0, dfx.SaveAt(ts=ts), _ctrl(rtol=rt, atol=rt / 100))
        e_floor = max(e_floor, _scaled(fl.ys.double().cpu().numpy(), true))
    e_dev = _scaled(sol.ys.double().cpu().numpy(), true)
    e_ora = _scaled(ref["ys"], true)
    st = sol.stats["num_steps"].cpu().numpy()
    print(f"{matmul}: rtol-scaled err {e_dev:.3f} (oracle {e_ora:.3f}, floor {e_floor:.3f}); "
          f"steps {st.min()}-{st.max()} (oracle {ref['stats'][:, 0].min()}-"
          f"{ref['stats'][:, 0].max()})")
    assert int(sol.result.abs().max()) == 0
    assert e_dev <= e_ora + e_floor + 1.0


@pytest.mark.parametrize("matmul", ["f32", "bf16"])
def test_clip_waves_forms_and_gradients(cuda, matmul):
    W, y0, term = _setup(matmul, seed=54)
    ts = np.linspace(0.0, 1.0, 6)
    a = _solve(term, y0, dfx.SaveAt(ts=ts, t1=True), _ctrl(), launch=4)
    b = _solve(term, y0, dfx.SaveAt(ts=ts, t1=True), _ctrl(), launch=8)
    c = _solve(term, y0, dfx.SaveAt(ts=ts, t1=True),
               dfx.PIDController(rtol=RTOL, atol=ATOL, step_ts=STEP_TS, jump_ts=JUMP_TS))
    torch.cuda.synchronize()
    assert torch.equal(a.ys, b.ys) and torch.equal(a.ys, c.ys)
    assert torch.equal(a.stats["num_steps"], b.stats["num_steps"])
    y = torch.tensor(y0, dtype=torch.float32).requires_grad_(True)
    with pytest.raises(NotImplementedError, match="ClipStepSizeController"):
        dfx.diffeqsolve(term, dfx.Tsit5(), 0.0, 1.0, 0.1, y, saveat=dfx.SaveAt(t1=True),
                        stepsize_controller=_ctrl(), adjoint=dfx.BacksolveAdjoint())


@pytest.mark.parametrize("matmul", ["f32", "bf16"])
def test_clipped_gradient_against_fp64_replay(cuda, matmul):
    from diffrax_amd import _mlp
    from tests.test_gpu_mlp_grad import _replay_grads, _weights
    C, max_steps = 2, 32
    Wn = _weights(5)
    rng = np.random.default_rng(56)
    y0 = rng.normal(size=(24, 64)).astype(np.float32)
    ts = np.linspace(0.0, 1.0, 8)
    wloss = rng.normal(size=(24, 8, 64)).astype(np.float32)
    dev = torch.device("cuda", 0)
    P = [torch.tensor(w, dtype=torch.float32, device=dev, requires_grad=True) for w in Wn]
    y0t = torch.tensor(y0, device=dev, requires_grad=True)
    sol = dfx.diffeqsolve(dfx.ODETerm(dfx.MLPField(*P, matmul=matmul)), dfx.Tsit5(), 0.0, 1.0,
                          0.1, y0t, saveat=dfx.SaveAt(ts=ts), max_steps=max_steps,
                          stepsize_controller=_ctrl(),
                          adjoint=dfx.RecursiveCheckpointAdjoint(checkpoints=max_steps // C))
    (sol.ys * torch.tensor(wloss, device=dev)).sum().backward()
    # the clipped accepted grid of the same solve
    plan = _mlp.prepare_mlp(dfx.ODETerm(dfx.MLPField(*[p.detach() for p in P], matmul=matmul)),
                            dfx.Tsit5(), 0.0, 1.0, 0.1, torch.tensor(y0), dfx.SaveAt(ts=ts),
                            _ctrl(), max_steps, dev, ckpt_every=C)
    plan.launch()
    torch.cuda.synchronize()
    assert torch.equal(plan.ys, sol.ys.detach())
    grid = plan.grid.cpu().numpy()
    nacc = plan.stats[:, 1].cpu().numpy()
    assert nacc.max() > C
    for b in range(len(y0)):   # the jump: a step ends just before 0.4, the next starts after
        ends = set(grid[b, :nacc[b], 1].tolist())
        starts = set(grid[b, :nacc[b], 0].tolist())
        assert np.nextafter(0.4, -np.inf) in ends and np.nextafter(0.4, np.inf) in starts
    ref, _ = _replay_grads(Wn, y0, ts, wloss, grid, nacc)
    grads = [y0t.grad.double().cpu()] + [p.grad.double().cpu() for p in P]
    worst = max(float((g - r).abs().max() / r.abs().max()) for g, r in zip(grads, ref))
    print(f"{matmul}: clipped gradient rel err {worst:.2e}; accepted steps "
          f"{nacc.min()}-{nacc.max()}")
    assert worst < 1e-4


def test_clipped_steps_in_reverse_time(cuda):
    """t0 = 1 > t1 = 0: the clipping works on the direction-scaled times (clip.py:231-241),
    so step_ts are still step ends and after a jump the next step starts just past it in
    the direction of integration, at nextafter(j, -inf)."""
    W, y0, term = _setup("f32", seed=58)
    st = dfx.diffeqsolve(term, dfx.Tsit5(), 1.0, 0.0, -0.1,
                         torch.as_tensor(y0, dtype=torch.float32), saveat=dfx.SaveAt(steps=True),
                         stepsize_controller=_ctrl(), max_steps=4096)
    torch.cuda.synchronize()
    acc = st.stats["num_accepted_steps"].cpu().numpy()
    kt = st.ts.cpu().numpy()
    for b in range(B):
        knots = set(kt[b, :acc[b]].tolist())
        for s in STEP_TS:
            assert s in knots, (b, s)
        for j in JUMP_TS:
            assert np.nextafter(j, -np.inf) in knots and j not in knots, (b, j)
        assert kt[b, acc[b] - 1] == 0.0
