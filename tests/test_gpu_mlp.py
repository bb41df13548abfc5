"""GPU tests of the neural-ODE path (MLP 64 -> 256 -> 256 -> 64, Tsit5 + PID).

Vector field: the MFMA tile (dfx_mlp_eval) against a float64 torch evaluation of the same
float32-weight MLP, both forms to 2e-6 absolute (outputs of size ~1).  The bfloat16 form
carries every weight and every layer input as three bfloat16 pieces (v = p0 + p1 + p2
exactly) and sums the six products p_i q_j with i + j <= 2 on bf16 MFMA with float32
accumulation, so it evaluates the float32-weight model to float32 MFMA accuracy.

Solutions: the device forms integrate in float32, so their step sequences need not equal
the FP64 oracle's (float32 rounding moves the embedded error estimate when the true local
error is near 1e-9).  They are therefore judged by accuracy at matched tolerance: the
rtol-scaled error max |y - y_true| / (atol + rtol |y_true|) against a tight FP64 solve
(rtol = atol = 1e-12) of the float32-weight model must be at most that of the FP64 oracle
at the same rtol/atol, + 1 for float32 rounding, for both forms."""
import numpy as np
import pytest
import torch

import diffrax_amd as dfx
from diffrax_amd import _mlp
import oracle

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-5, 1e-7


def mlp_weights(seed=0):
    rng = np.random.default_rng(seed)
    xav = lambda o, i: rng.normal(0.0, np.sqrt(2.0 / (i + o)), size=(o, i))
    W1, W2, W3 = xav(256, 64), xav(256, 256), xav(64, 256)
    b1, b2, b3 = np.zeros(256), np.zeros(256), np.zeros(64)
    return W1, b1, W2, b2, W3, b3


def _vf(W, matmul):
    return dfx.MLPField(*(torch.tensor(w, dtype=torch.float32) for w in W), matmul=matmul)


def _scaled(y, y_true):
    return float(np.max(np.abs(y - y_true) / (ATOL + RTOL * np.abs(y_true))))


def _f64_vf(W, y):
    return _vf(W, "f32")(0.0, torch.tensor(y.astype(np.float32), dtype=torch.float64), None).numpy()


@pytest.mark.parametrize("matmul", ["f32", "bf16"])
def test_mlp_vector_field_on_mfma(cuda, matmul):
    """The MFMA tile against a float64 torch evaluation of the same MLP (dfx_mlp_eval)."""
    from diffrax_amd._mlp import mlp_eval
    W = mlp_weights(3)
    W = [W[0], np.random.default_rng(5).normal(0, 0.1, 256), W[2],
         np.random.default_rng(6).normal(0, 0.1, 256), W[4], np.random.default_rng(7).normal(0, 0.1, 64)]
    y0 = np.random.default_rng(4).normal(size=(100, 64))  # 3 full tiles + a ragged one
    f = mlp_eval(_vf(W, matmul), torch.tensor(y0, dtype=torch.float32, device="cuda"))
    f = f.double().cpu().numpy()
    err = np.max(np.abs(f - _f64_vf(W, y0)))
    print(f"{matmul}: max |f_gpu - f_fp64| = {err:.3e}")
    assert err < 2e-6


@pytest.mark.parametrize("matmul", ["f32", "bf16"])
def test_mlp_solution_accuracy_matched_tolerance(cuda, matmul):
    W = mlp_weights(0)
    y0 = np.random.default_rng(1).normal(size=(64, 64))
    ts = np.linspace(0.0, 1.0, 32)
    sol = dfx.diffeqsolve(dfx.ODETerm(_vf(W, matmul)), dfx.Tsit5(), 0.0, 1.0, 0.1,
                          torch.tensor(y0, dtype=torch.float32), saveat=dfx.SaveAt(ts=ts),
                          stepsize_controller=dfx.PIDController(rtol=RTOL, atol=ATOL))
    q = oracle.mlp_pack(*W)
    y0m = y0.astype(np.float32).astype(np.float64)
    ref = oracle.solve(y0m, 1.0, q, ts=ts, solver="tsit5", vf="mlp", hidden=256, dt0=0.1,
                       rtol=RTOL, atol=ATOL)
    true = oracle.solve(y0m, 1.0, q, ts=ts, solver="tsit5", vf="mlp", hidden=256, dt0=0.1,
                        rtol=1e-12, atol=1e-12, max_steps=100000)["ys"]
    ys = sol.ys.double().cpu().numpy()
    e_dev, e_ora = _scaled(ys, true), _scaled(ref["ys"], true)
    st = sol.stats["num_steps"].cpu().numpy()
    a_dev, a_ora = np.max(np.abs(ys - true)), np.max(np.abs(ref["ys"] - true))
    print(f"{matmul}: rtol-scaled err vs tight ref {e_dev:.3f} (FP64 oracle at same tol "
          f"{e_ora:.3f}); abs err {a_dev:.3e} (oracle {a_ora:.3e}); steps "
          f"{st.min()}-{st.max()} (oracle {ref['stats'][:, 0].min()}-{ref['stats'][:, 0].max()})")
    assert np.isfinite(ys).all()
    assert e_dev <= e_ora + 1.0


def test_device_resident_ts(cuda):
    """SaveAt(ts) as a device tensor: the same rows as host ts, bit for bit, through the
    solve and its gradient (the backward pass reads its segment count from a pinned copy
    queued behind the forward); an unordered device ts still raises."""
    W = mlp_weights(1)
    y0 = torch.tensor(np.random.default_rng(2).normal(size=(40, 64)), dtype=torch.float32)
    ts = torch.linspace(0.0, 1.0, 5, dtype=torch.float64)
    ctrl = dfx.PIDController(rtol=RTOL, atol=ATOL)

    def solve(ts_, y0_):
        return dfx.diffeqsolve(dfx.ODETerm(_vf(W, "f32")), dfx.Tsit5(), 0.0, 1.0, 0.1, y0_,
                               saveat=dfx.SaveAt(ts=ts_), stepsize_controller=ctrl,
                               max_steps=64)
    a, b = solve(ts, y0), solve(ts.to(cuda), y0)
    assert torch.equal(a.ys, b.ys)
    ga, gb = (y0.clone().requires_grad_(True) for _ in range(2))
    solve(ts, ga).ys.square().sum().backward()
    solve(ts.to(cuda), gb).ys.square().sum().backward()
    assert torch.equal(ga.grad, gb.grad)
    with pytest.raises(RuntimeError, match="increasing or decreasing"):
        solve(torch.tensor([0.0, 0.5, 0.25, 0.75, 1.0], dtype=torch.float64, device=cuda), y0)


@pytest.mark.parametrize("matmul", ["f32", "bf16"])
@pytest.mark.parametrize("ctrl_kw", [dict(), dict(pcoeff=0.3, icoeff=0.4),
                                     dict(pcoeff=0.1, icoeff=0.3, dcoeff=0.05)],
                         ids=["I", "PI", "PID"])
def test_mlp_initial_step_and_pid_terms(cuda, matmul, ctrl_kw):
    """dt0=None (_select_initial_step, pid.py:33-81) and PID p / d terms (pid.py:394-567)
    on the neural-ODE kernel, against the FP64 oracle with the same controller: the first
    accepted step within 1e-4 of the oracle's, and the solution's rtol-scaled error against
    a tight solve at most the oracle's plus the kernel's own floor plus one unit (the two
    errors add).  The float32 form evaluates f to ~8e-7 relative (hardware exp2 / log2),
    ~1e-6 absolute over [0, 1], which the atol of 1e-7 scales to several units on small
    components, where a PI or PID controller takes the FP64 oracle well below one.  The
    floor is that rounding error: the kernel's error, same controller, at rtol 1e-7, 1e-8
    and 1e-9, where truncation is below 0.01 units; it varies from run to run (3 to 16
    units measured over these cases), so the largest of the three samples stands for it."""
    W = mlp_weights(4)
    y0 = np.random.default_rng(5).normal(size=(48, 64))
    ts = np.linspace(0.0, 1.0, 16)
    ctrl = dfx.PIDController(rtol=RTOL, atol=ATOL, **ctrl_kw)
    sol = dfx.diffeqsolve(dfx.ODETerm(_vf(W, matmul)), dfx.Tsit5(), 0.0, 1.0, None,
                          torch.tensor(y0, dtype=torch.float32), saveat=dfx.SaveAt(ts=ts),
                          stepsize_controller=ctrl)
    q = oracle.mlp_pack(*W)
    y0m = y0.astype(np.float32).astype(np.float64)
    kw = dict(ts=ts, solver="tsit5", vf="mlp", hidden=256, dt0=None, rtol=RTOL, atol=ATOL,
              pcoeff=ctrl_kw.get("pcoeff", 0.0), icoeff=ctrl_kw.get("icoeff", 1.0),
              dcoeff=ctrl_kw.get("dcoeff", 0.0))
    ref = oracle.solve(y0m, 1.0, q, record_grid=True, **kw)
    true = oracle.solve(y0m, 1.0, q, ts=ts, solver="tsit5", vf="mlp", hidden=256, dt0=0.1,
                        rtol=1e-12, atol=1e-12, max_steps=100000)["ys"]
    ys = sol.ys.double().cpu().numpy()
    assert np.isfinite(ys).all() and int(sol.result.abs().max()) == 0
    e_dev, e_ora = _scaled(ys, true), _scaled(ref["ys"], true)
    # the first step size of every trajectory (the oracle's grid row 0) against the kernel's
    plan = _mlp.prepare_mlp(
        dfx.ODETerm(_vf(W, matmul)), dfx.Tsit5(), 0.0, 1.0, None,
        torch.tensor(y0, dtype=torch.float32), dfx.SaveAt(ts=ts), ctrl, 4096,
        torch.device("cuda", 0), ckpt_every=64)
    plan.launch()
    g = plan.grid[:, 0].cpu().numpy()
    dt_dev = g[:, 1] - g[:, 0]
    dt_ora = ref["grid"][:, 0, 1] - ref["grid"][:, 0, 0]
    rel = np.max(np.abs(dt_dev - dt_ora) / dt_ora)
    st = sol.stats["num_steps"].cpu().numpy()
    print(f"{matmul} {ctrl_kw}: first step rel diff {rel:.2e}; rtol-scaled err {e_dev:.3f} "
          f"(oracle {e_ora:.3f}); steps {st.min()}-{st.max()} (oracle "
          f"{ref['stats'][:, 0].min()}-{ref['stats'][:, 0].max()})")
    e_floor = 0.0
    for rt in (1e-7, 1e-8, 1e-9):
        floor = dfx.diffeqsolve(
            dfx.ODETerm(_vf(W, matmul)), dfx.Tsit5(), 0.0, 1.0, None,
            torch.tensor(y0, dtype=torch.float32), saveat=dfx.SaveAt(ts=ts), max_steps=100000,
            stepsize_controller=dfx.PIDController(rtol=rt, atol=rt / 100, **ctrl_kw))
        e_floor = max(e_floor, _scaled(floor.ys.double().cpu().numpy(), true))
    print(f"  kernel floor (rtol 1e-7 .. 1e-9): rtol-scaled err {e_floor:.3f}")
    assert rel < 1e-4
    assert e_dev <= e_ora + e_floor + 1.0


@pytest.mark.parametrize("matmul", ["f32", "bf16"])
def test_wave_count_gives_the_same_bits(cuda, matmul):
    """Four and eight wavefronts per tile split the same products and the controller norms
    sum in the same order, so solves (integral and PID control, given and automatic dt0,
    with gradient recording) and gradients are bit-identical; launch='auto' picks one."""
    W = mlp_weights(6)
    y0 = torch.tensor(np.random.default_rng(8).normal(size=(100, 64)), dtype=torch.float32)
    ts = np.linspace(0.0, 1.0, 9)
    P = [torch.tensor(w, dtype=torch.float32, device=cuda) for w in W]

    def solve(launch, ctrl, dt0, grad=False):
        params = [p.clone().requires_grad_(grad) for p in P]
        y = y0.to(cuda).clone().requires_grad_(grad)
        sol = dfx.diffeqsolve(dfx.ODETerm(dfx.MLPField(*params, matmul=matmul)), dfx.Tsit5(),
                              0.0, 1.0, dt0, y, saveat=dfx.SaveAt(ts=ts),
                              stepsize_controller=ctrl, launch=launch)
        if not grad:
            return [sol.ys.detach(), sol.stats["num_steps"]]
        sol.ys.square().sum().backward()
        return [sol.ys.detach(), y.grad] + [p.grad for p in params]

    for ctrl, dt0 in ((dfx.PIDController(rtol=RTOL, atol=ATOL), 0.1),
                      (dfx.PIDController(rtol=RTOL, atol=ATOL, pcoeff=0.2, icoeff=0.5), None)):
        for grad in (False, True):
            a, b = solve(4, ctrl, dt0, grad), solve(8, ctrl, dt0, grad)
            assert all(torch.equal(x, y) for x, y in zip(a, b))
    c = solve("auto", dfx.PIDController(rtol=RTOL, atol=ATOL), 0.1)
    assert torch.equal(c[0], solve(4, dfx.PIDController(rtol=RTOL, atol=ATOL), 0.1)[0])
    with pytest.raises(ValueError, match="wavefronts"):
        solve(6, dfx.PIDController(rtol=RTOL, atol=ATOL), 0.1)


@pytest.mark.parametrize("matmul", ["f32", "bf16"])
def test_mlp_dense_output_and_step_saves(cuda, matmul):
    """SaveAt(dense=True) and SaveAt(steps=k) on the neural-ODE kernel.  Dense output
    evaluated at the SaveAt(ts) times returns the saved rows bit for bit (same interpolant
    arithmetic), at t0 returns y0, outside [t0, t1] NaN; its derivative at the knots equals
    the vector field at the knot states to float32 MFMA accuracy (Tsit5's interpolant has
    the FSAL stage as its end slopes), and in between to the interpolant's accuracy.
    SaveAt(steps=k) rows are the accepted-step knots of the same solve (every k-th), in
    time order with the ts rows, with the t1 row dropped when the last step was saved."""
    W = mlp_weights(9)
    B = 70
    y0 = torch.tensor(np.random.default_rng(10).normal(size=(B, 64)), dtype=torch.float32,
                      device=cuda)
    ts = np.linspace(0.0, 1.0, 7)
    term = dfx.ODETerm(_vf(W, matmul))
    ctrl = dfx.PIDController(rtol=RTOL, atol=ATOL)
    sol = dfx.diffeqsolve(term, dfx.Tsit5(), 0.0, 1.0, 0.1, y0,
                          saveat=dfx.SaveAt(ts=ts, dense=True), stepsize_controller=ctrl)
    assert torch.equal(sol.evaluate(torch.tensor(ts)), sol.ys)
    assert torch.equal(sol.evaluate(0.0), y0)
    assert torch.isnan(sol.evaluate(1.5)).all() and torch.isnan(sol.evaluate(-0.1)).all()
    per = torch.tensor(np.tile(ts[None, 2:5], (B, 1)))
    assert torch.equal(sol.evaluate(per), sol.ys[:, 2:5])
    plan = sol.interpolation.plan
    n = int(sol.stats["num_accepted_steps"].max())
    knots_t = plan.dense["dense_ts"][:, 1:n + 1]
    knots_y = plan.dense["dense_ys"][:, 1:n + 1]
    # derivative at the step ends (left: the step that ends there) = f(knot state)
    from diffrax_amd._mlp import mlp_eval
    acc = sol.stats["num_accepted_steps"]
    dev = sol.derivative(plan.dense["dense_ts"][torch.arange(B), acc].reshape(B, 1),
                         left=True)[:, 0]
    f_end = mlp_eval(term.vector_field, plan.dense["dense_ys"][torch.arange(B), acc])
    assert float((dev - f_end).abs().max()) < 2e-5
    # in between: against the vector field at the interpolated state
    mid = torch.tensor(np.tile([0.33, 0.61], (B, 1)))
    dmid = sol.derivative(mid)
    fmid = mlp_eval(term.vector_field, sol.evaluate(mid).reshape(-1, 64)).reshape(B, 2, 64)
    assert float((dmid - fmid).abs().max()) < 1e-3
    # SaveAt(steps=1): the knots; steps=2: every other one; combined with ts: time order
    st = dfx.diffeqsolve(term, dfx.Tsit5(), 0.0, 1.0, 0.1, y0,
                         saveat=dfx.SaveAt(steps=True, t1=True), stepsize_controller=ctrl,
                         max_steps=64)
    assert torch.equal(st.stats["num_accepted_steps"], acc)
    for b in range(B):
        k = int(acc[b])
        assert torch.equal(st.ts[b, :k], knots_t[b, :k])
        assert torch.equal(st.ys[b, :k], knots_y[b, :k])
        assert torch.isinf(st.ts[b, k:]).all()   # t1 was the last step's row
    s2 = dfx.diffeqsolve(term, dfx.Tsit5(), 0.0, 1.0, 0.1, y0,
                         saveat=dfx.SaveAt(steps=2, ts=ts), stepsize_controller=ctrl,
                         max_steps=64)
    for b in range(B):
        row = s2.ts[b][torch.isfinite(s2.ts[b])]
        assert bool((row[1:] >= row[:-1]).all())
        k = int(acc[b])
        want = torch.cat([knots_t[b, 1:k:2], torch.tensor(ts, device=cuda)]).sort().values
        assert torch.equal(row.sort().values, want)
    with pytest.raises(NotImplementedError, match="gradients"):
        y = y0.clone().requires_grad_(True)
        dfx.diffeqsolve(term, dfx.Tsit5(), 0.0, 1.0, 0.1, y,
                        saveat=dfx.SaveAt(ts=ts, dense=True), stepsize_controller=ctrl)


@pytest.mark.parametrize("matmul", ["f32", "bf16"])
def test_per_trajectory_save_times(cuda, matmul):
    """SaveAt(ts=[B, n]) on the neural-ODE kernel (a vmapped SaveAt(ts)): every trajectory of
    a tile is integrated independently, so each row equals the solve of that trajectory
    alone with its own ts, bit for bit, forward and y0 gradient (discrete adjoint and
    BacksolveAdjoint); the weight gradients, summed over trajectories, to float32
    summation order."""
    W = mlp_weights(12)
    B = 40
    rng = np.random.default_rng(13)
    y0 = torch.tensor(rng.normal(size=(B, 64)), dtype=torch.float32, device=cuda)
    ts = np.sort(rng.uniform(0.0, 1.0, size=(B, 6)), axis=1)
    ts[:, -1] = 1.0
    ctrl = dfx.PIDController(rtol=RTOL, atol=ATOL)
    P = [torch.tensor(w, dtype=torch.float32, device=cuda) for w in W]
    w_out = torch.tensor(rng.normal(size=(B, 6, 64)), dtype=torch.float32, device=cuda)

    def run(y, t, adjoint, rows):
        Pg = [p.clone().requires_grad_(True) for p in P]
        yg = y.clone().requires_grad_(True)
        sol = dfx.diffeqsolve(dfx.ODETerm(dfx.MLPField(*Pg, matmul=matmul)), dfx.Tsit5(), 0.0,
                              1.0, 0.1, yg, saveat=dfx.SaveAt(ts=t), stepsize_controller=ctrl,
                              adjoint=adjoint, max_steps=256)
        (sol.ys * w_out[rows]).sum().backward()
        return sol.ys.detach(), yg.grad, [p.grad for p in Pg]

    back = dfx.BacksolveAdjoint(stepsize_controller=dfx.PIDController(
        rtol=1e-6, atol=1e-8, norm=dfx.adjoint_rms_seminorm))
    for adjoint in (dfx.RecursiveCheckpointAdjoint(), back):
        ys, gy, gw = run(y0, torch.tensor(ts, device=cuda), adjoint, slice(None))
        tot = [torch.zeros_like(p) for p in P]
        for b in range(B):
            ysb, gyb, gwb = run(y0[b:b + 1], torch.tensor(ts[b:b + 1], device=cuda), adjoint,
                                slice(b, b + 1))
            assert torch.equal(ysb, ys[b:b + 1]) and torch.equal(gyb, gy[b:b + 1])
            tot = [t + g for t, g in zip(tot, gwb)]
        for t, g in zip(tot, gw):
            assert float((t - g).abs().max()) <= 1e-5 * float(g.abs().max())
    # every row the same as the shared ts: the shared solve's bits
    shared = dfx.diffeqsolve(dfx.ODETerm(dfx.MLPField(*P, matmul=matmul)), dfx.Tsit5(), 0.0,
                             1.0, 0.1, y0, saveat=dfx.SaveAt(ts=ts[0]),
                             stepsize_controller=ctrl)
    rows = dfx.diffeqsolve(dfx.ODETerm(dfx.MLPField(*P, matmul=matmul)), dfx.Tsit5(), 0.0,
                           1.0, 0.1, y0, saveat=dfx.SaveAt(ts=np.tile(ts[:1], (B, 1))),
                           stepsize_controller=ctrl)
    assert torch.equal(shared.ys, rows.ys) and torch.equal(shared.ts, rows.ts)


def test_mlp_save_fn_and_subs(cuda):
    """SaveAt(fn=...) and SaveAt(subs=[...]) on the neural-ODE path (the host applies fn to
    the saved rows and runs one solve per SubSaveAt, as for the small-system kernel)."""
    W = mlp_weights(14)
    y0 = torch.tensor(np.random.default_rng(15).normal(size=(33, 64)), dtype=torch.float32,
                      device=cuda)
    term = dfx.ODETerm(_vf(W, "f32"))
    ctrl = dfx.PIDController(rtol=RTOL, atol=ATOL)
    ts = [0.25, 0.5, 1.0]
    plain = dfx.diffeqsolve(term, dfx.Tsit5(), 0.0, 1.0, 0.1, y0, saveat=dfx.SaveAt(ts=ts),
                            stepsize_controller=ctrl)
    sq = dfx.diffeqsolve(term, dfx.Tsit5(), 0.0, 1.0, 0.1, y0,
                         saveat=dfx.SaveAt(ts=ts, fn=lambda t, y, args: y.square().sum()),
                         stepsize_controller=ctrl)
    assert torch.equal(sq.ys, plain.ys.square().sum(-1))
    subs = dfx.diffeqsolve(term, dfx.Tsit5(), 0.0, 1.0, 0.1, y0,
                           saveat=dfx.SaveAt(subs=[dfx.SubSaveAt(ts=ts), dfx.SubSaveAt(t1=True)]),
                           stepsize_controller=ctrl)
    t1 = dfx.diffeqsolve(term, dfx.Tsit5(), 0.0, 1.0, 0.1, y0, saveat=dfx.SaveAt(t1=True),
                         stepsize_controller=ctrl)
    assert torch.equal(subs.ys[0], plain.ys) and torch.equal(subs.ys[1], t1.ys)
