"""GPU tests of the Heun SDE path (diagonal-noise OU) and its discrete adjoint.

Forward: the float64 form must equal the FP64 oracle bit for bit; the float32 form is
held to a normwise relative error max|y - y_ref| / max|y_ref| < 1e-5 after 1000 steps
(components cross zero, so a componentwise ratio is not meaningful).
Gradient: the adjoint kernel against (a) torch float64 autograd through a replay of the
same fixed step grid (oracle/pyoracle.py) to 1e-10 relative, and (b) central finite
differences of the FP64 oracle along random directions (h = 1e-6) to 1e-6 relative."""
import numpy as np
import pytest
import torch

import diffrax_amd as dfx
import oracle
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def _inputs(P, d, N, seed=0, dt=1e-3):
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0.5, 2.0, size=d)
    mu = rng.normal(0.0, 1.0, size=d)
    sigma = rng.uniform(0.1, 0.5, size=d)
    dW = (rng.standard_normal(size=(N, P, d)) * np.sqrt(dt)).astype(np.float32)
    y0 = rng.normal(0.0, 0.3, size=(P, d))
    return theta, mu, sigma, dW, y0


def _gpu(y0, theta, mu, sigma, dW, dtype, t1=1.0, dt0=1e-3, grad=False):
    dev = torch.device("cuda", 0)
    y0t = torch.tensor(y0, dtype=dtype, device=dev, requires_grad=grad)
    th = torch.tensor(theta, dtype=dtype, device=dev, requires_grad=grad)
    m = torch.tensor(mu, dtype=dtype, device=dev, requires_grad=grad)
    sg = torch.tensor(sigma, dtype=dtype, device=dev, requires_grad=grad)
    terms = dfx.MultiTerm(dfx.ODETerm(dfx.OUDrift()),
                          dfx.ControlTerm(dfx.DiagonalNoise(), dfx.IncrementPath(torch.tensor(dW))))
    sol = dfx.diffeqsolve(terms, dfx.Heun(), 0.0, t1, dt0, y0t, (th, m, sg),
                          saveat=dfx.SaveAt(t1=True), stepsize_controller=dfx.ConstantStepSize())
    return sol, (y0t, th, m, sg)


def _oracle(y0, theta, mu, sigma, dW, t1=1.0, dt0=1e-3):
    q = np.concatenate([theta, mu, sigma])
    return oracle.solve(y0, t1, q, dW=dW.astype(np.float64), solver="heun", vf="ou",
                        controller="constant", dt0=dt0, save_t1=True, max_steps=100000)


def test_sde_forward_f64_bit_identical(cuda):
    theta, mu, sigma, dW, y0 = _inputs(P=512, d=16, N=1000)
    sol, _ = _gpu(y0, theta, mu, sigma, dW, torch.float64)
    ref = _oracle(y0, theta, mu, sigma, dW)
    ys = sol.ys.detach().cpu().numpy()
    print("sde f64 max abs diff", np.max(np.abs(ys - ref["ys"])))
    assert np.array_equal(ys, ref["ys"])
    assert int(sol.stats["num_steps"][0]) == 1000 == int(ref["stats"][0, 0])


def test_sde_forward_f32_within_tolerance(cuda):
    theta, mu, sigma, dW, y0 = _inputs(P=4096, d=16, N=1000, seed=1)
    sol, _ = _gpu(y0, theta, mu, sigma, dW, torch.float32)
    ref = _oracle(y0, theta, mu, sigma, dW)
    ys = sol.ys.detach().cpu().numpy()
    err = np.max(np.abs(ys - ref["ys"])) / np.max(np.abs(ref["ys"]))
    print("sde f32 normwise rel err", err)
    assert err < 1e-5


def test_sde_adjoint_vs_torch_replay(cuda):
    P, d, N = 8, 16, 100
    theta, mu, sigma, dW, y0 = _inputs(P=P, d=d, N=N, seed=2)
    dt0 = 1.0 / N
    sol, (y0t, th, m, sg) = _gpu(y0, theta, mu, sigma, dW, torch.float64, t1=1.0, dt0=dt0,
                                 grad=True)
    w = np.random.default_rng(3).normal(size=(P, 1, d))
    loss = (sol.ys * torch.tensor(w, device=sol.ys.device)).sum()
    loss.backward()
    # reference: torch float64 autograd through the fixed grid
    from diffrax_amd._sde import constant_grid
    grid = constant_grid(0.0, 1.0, dt0)
    steps = list(zip(grid[:-1], grid[1:]))
    Y0 = torch.tensor(y0, dtype=torch.float64, requires_grad=True)
    TH = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    MU = torch.tensor(mu, dtype=torch.float64, requires_grad=True)
    SG = torch.tensor(sigma, dtype=torch.float64, requires_grad=True)
    dWt = torch.tensor(dW.astype(np.float64))
    f = lambda t, y: TH * (MU - y)
    g_dw = lambda n, t, y: SG * dWt[n]
    out = pyoracle.replay(f, "heun", steps, Y0, [], g_dw=g_dw, save_t1=True)
    ref_loss = (out[-1] * torch.tensor(w[:, 0, :])).sum()
    ref_loss.backward()
    for name, got, want in (("y0", y0t.grad, Y0.grad), ("theta", th.grad, TH.grad),
                            ("mu", m.grad, MU.grad), ("sigma", sg.grad, SG.grad)):
        got = got.detach().cpu().numpy()
        want = want.numpy()
        rel = np.max(np.abs(got - want)) / np.max(np.abs(want))
        print(f"adjoint {name}: rel err {rel:.3e}")
        assert rel < 1e-10, name


def test_sde_adjoint_vs_central_differences(cuda):
    P, d, N = 16, 16, 200
    theta, mu, sigma, dW, y0 = _inputs(P=P, d=d, N=N, seed=4)
    dt0 = 1.0 / N
    sol, (y0t, th, m, sg) = _gpu(y0, theta, mu, sigma, dW, torch.float64, dt0=dt0, grad=True)
    loss = (sol.ys ** 2).mean()
    loss.backward()
    grads = np.concatenate([th.grad.cpu().numpy(), m.grad.cpu().numpy(), sg.grad.cpu().numpy()])

    def L(params):
        r = oracle.solve(y0, 1.0, params, dW=dW.astype(np.float64), solver="heun", vf="ou",
                         controller="constant", dt0=dt0, save_t1=True, max_steps=100000)
        return float(np.mean(r["ys"] ** 2))

    p0 = np.concatenate([theta, mu, sigma])
    rng = np.random.default_rng(5)
    h = 1e-6
    worst = 0.0
    for _ in range(8):
        v = rng.normal(size=p0.shape)
        v /= np.linalg.norm(v)
        fd = (L(p0 + h * v) - L(p0 - h * v)) / (2 * h)
        an = float(grads @ v)
        worst = max(worst, abs(fd - an) / max(abs(an), 1e-8))
    print("sde adjoint vs central FD worst rel", worst)
    assert worst < 1e-6


def test_sde_adjoint_f32_vs_f64(cuda):
    """The float32 adjoint (FMA-contracted, per-segment float32 cotangent sums) against the
    float64 adjoint, which the two tests above pin to the replay and to finite differences.
    1000 steps, 32-step checkpoint segments, 256 paths; bar 1e-4 relative per parameter."""
    P, d, N = 256, 16, 1000
    theta, mu, sigma, dW, y0 = _inputs(P=P, d=d, N=N, seed=6)
    w = np.random.default_rng(7).normal(size=(P, 1, d))
    grads = {}
    for dt in (torch.float64, torch.float32):
        sol, leaves = _gpu(y0, theta, mu, sigma, dW, dt, grad=True)
        (sol.ys * torch.tensor(w, dtype=dt, device=sol.ys.device)).sum().backward()
        grads[dt] = [x.grad.double().cpu().numpy() for x in leaves]
    for name, g32, g64 in zip(("y0", "theta", "mu", "sigma"), grads[torch.float32],
                              grads[torch.float64]):
        rel = np.max(np.abs(g32 - g64)) / np.max(np.abs(g64))
        print(f"f32 adjoint {name}: rel err {rel:.3e}")
        assert rel < 1e-4, name


SAVE_TS = np.array([0.0, 0.0132, 0.25, 0.2571, 0.5, 0.66, 0.999, 1.0])


@pytest.mark.parametrize("k,with_ts", [(1, False), (7, False), (10, False), (0, True),
                                       (7, True)])
def test_sde_step_saves_forward_and_adjoint(cuda, k, with_ts):
    """SaveAt(t0=True, ts=..., steps=k, t1=True) on the SDE path: the rows (t0; per step the
    ts inside it, interpolated by Heun's Hermite polynomial, then the state when k divides
    the step count; t1 unless k divides the step count) equal the FP64 oracle's bit for
    bit, and the adjoint with cotangents on every row matches torch float64 autograd
    through the fixed-grid replay to 1e-10."""
    P, d, N = 24, 16, 100
    ts = SAVE_TS if with_ts else None
    n_ts = 0 if ts is None else ts.size
    theta, mu, sigma, dW, y0 = _inputs(P=P, d=d, N=N, seed=21)
    dt0 = 1.0 / N
    dev = torch.device("cuda", 0)
    y0t = torch.tensor(y0, dtype=torch.float64, device=dev, requires_grad=True)
    th = torch.tensor(theta, dtype=torch.float64, device=dev, requires_grad=True)
    m = torch.tensor(mu, dtype=torch.float64, device=dev, requires_grad=True)
    sg = torch.tensor(sigma, dtype=torch.float64, device=dev, requires_grad=True)
    terms = dfx.MultiTerm(dfx.ODETerm(dfx.OUDrift()),
                          dfx.ControlTerm(dfx.DiagonalNoise(), dfx.IncrementPath(torch.tensor(dW))))
    sol = dfx.diffeqsolve(terms, dfx.Heun(), 0.0, 1.0, dt0, y0t, (th, m, sg),
                          saveat=dfx.SaveAt(t0=True, ts=ts, steps=k, t1=True),
                          stepsize_controller=dfx.ConstantStepSize())
    n_rows = 1 + n_ts + (N // k if k else 0) + (0 if k and N % k == 0 else 1)
    assert tuple(sol.ys.shape) == (P, n_rows, d) and tuple(sol.ts.shape) == (P, n_rows)
    q = np.concatenate([theta, mu, sigma])
    ref = oracle.solve(y0, 1.0, q, ts=ts, dW=dW.astype(np.float64), solver="heun", vf="ou",
                       controller="constant", dt0=dt0, save_t0=True, save_t1=True,
                       save_steps=k, max_steps=N)
    assert np.array_equal(sol.ys.detach().cpu().numpy(), ref["ys"][:, :n_rows])
    assert np.array_equal(sol.ts.cpu().numpy(), ref["ts"][:, :n_rows])
    w = np.random.default_rng(22).normal(size=(P, n_rows, d))
    (sol.ys * torch.tensor(w, device=dev)).sum().backward()
    from diffrax_amd._sde import constant_grid
    grid = constant_grid(0.0, 1.0, dt0)
    steps = list(zip(grid[:-1], grid[1:]))
    Y0 = torch.tensor(y0, dtype=torch.float64, requires_grad=True)
    TH, MU, SG = (torch.tensor(v, dtype=torch.float64, requires_grad=True)
                  for v in (theta, mu, sigma))
    dWt = torch.tensor(dW.astype(np.float64))
    out = pyoracle.replay(lambda t, y: TH * (MU - y), "heun", steps, Y0,
                          [] if ts is None else list(ts), g_dw=lambda n, t, y: SG * dWt[n],
                          save_t1=not (k and N % k == 0), save_steps=k)
    rows = torch.stack([Y0] + out, dim=1)
    assert rows.shape[1] == n_rows
    (rows * torch.tensor(w)).sum().backward()
    for name, got, want in (("y0", y0t.grad, Y0.grad), ("theta", th.grad, TH.grad),
                            ("mu", m.grad, MU.grad), ("sigma", sg.grad, SG.grad)):
        got, want = got.detach().cpu().numpy(), want.numpy()
        rel = np.max(np.abs(got - want)) / np.max(np.abs(want))
        assert rel < 1e-10, (name, rel)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_sde_dense_output(cuda, dtype):
    """SaveAt(dense=True) on the SDE path (dfx_sde_dense_eval): evaluated at the SaveAt(ts)
    times it returns the saved rows bit for bit (the same Hermite arithmetic, which the
    FP64 oracle pins in test_sde_step_saves_forward_and_adjoint), shared or per path; at
    t0 exactly y0; at the step ends the SaveAt(steps=True) states from either side, to the
    rounding of the polynomial's sum at theta = 0 or 1 (as the reference's polyval); outside
    [t0, t1] NaN.  Its derivative matches central differences of the evaluation inside a
    step (1e-6 relative, float64).  SaveAt(dense=True) alone saves no rows; gradients
    through a dense solve raise."""
    P, d, N = 24, 16, 100
    theta, mu, sigma, dW, y0 = _inputs(P=P, d=d, N=N, seed=31)
    dev = torch.device("cuda", 0)
    args = tuple(torch.tensor(v, dtype=dtype, device=dev) for v in (theta, mu, sigma))
    terms = dfx.MultiTerm(dfx.ODETerm(dfx.OUDrift()),
                          dfx.ControlTerm(dfx.DiagonalNoise(), dfx.IncrementPath(torch.tensor(dW))))
    y0t = torch.tensor(y0, dtype=dtype, device=dev)
    solve = lambda saveat: dfx.diffeqsolve(terms, dfx.Heun(), 0.0, 1.0, 1.0 / N, y0t, args,
                                           saveat=saveat,
                                           stepsize_controller=dfx.ConstantStepSize())
    sol = solve(dfx.SaveAt(ts=SAVE_TS, dense=True))
    steps = solve(dfx.SaveAt(steps=True))
    assert torch.equal(sol.evaluate(torch.tensor(SAVE_TS)), sol.ys)
    grid = steps.ts[0].cpu().numpy()
    tol = 1e-12 if dtype == torch.float64 else 2e-5
    for left, rows in ((True, slice(None)), (False, slice(0, -1))):
        got = sol.evaluate(torch.tensor(grid[rows]), left=left)
        assert float((got - steps.ys[:, rows]).abs().max()) <= tol
    assert torch.equal(sol.evaluate(0.0), y0t)
    assert torch.isnan(sol.evaluate(1.5)).all() and torch.isnan(sol.evaluate(-0.1)).all()
    per = torch.tensor(np.tile(SAVE_TS[None, 1:3], (P, 1)))
    assert torch.equal(sol.evaluate(per), sol.ys[:, 1:3])
    only = solve(dfx.SaveAt(dense=True))
    assert only.ys is None and torch.equal(only.evaluate(torch.tensor(SAVE_TS)), sol.ys)
    if dtype == torch.float64:
        t, h = 0.4537, 1e-6
        fd = (sol.evaluate(t + h) - sol.evaluate(t - h)) / (2 * h)
        der = sol.derivative(t)
        assert float((der - fd).abs().max()) <= 1e-6 * float(der.abs().max())
    with pytest.raises(NotImplementedError, match="dense"):
        g = y0t.clone().requires_grad_(True)
        dfx.diffeqsolve(terms, dfx.Heun(), 0.0, 1.0, 1.0 / N, g, args,
                        saveat=dfx.SaveAt(t1=True, dense=True),
                        stepsize_controller=dfx.ConstantStepSize())
