"""GPU correctness at the full benchmark sizes of the neural-ODE gradient and the OU SDE.

Diagonal-noise OU, Heun, 262,144 paths x 16 components x 1,000 steps: the increments
dW [1000, 262144, 16] float32 (16.8 GB) are drawn on the device from a fixed seed; the
FP64 oracle reads the increments of the paths it checks.  Offsets into dW pass 2^31
from step 512 on, so every checked path exercises the kernels' 64-bit indexing, and the
last 512 paths are the furthest from the base.
- float64 forward: bit-identical to the oracle on the first 16 and the last 512 paths;
- float32 forward: within 1e-5 normwise of the oracle on the same paths;
- gradients of sum(w * y(1)) with w nonzero on the last 512 paths only, so that theta,
  mu and sigma cotangents come from those paths alone: float64 within 1e-10 and float32
  within 1e-5 normwise of torch float64 autograd through the same fixed-step replay
  (oracle/pyoracle.py, pinned by central differences in test_gpu_sde.py).

Neural ODE, 8,192 trajectories, MLP 64 -> 256 -> 256 -> 64, Tsit5 + PID(1e-5, 1e-7),
32 saves, sqrt(N) checkpoints: the cotangent of ys is nonzero on 256 rows spread over the
batch, so the weight gradients come from those rows alone; y0 and every weight gradient
against torch float64 autograd through a replay of the device's accepted steps of those
rows (pinned by central differences in test_gpu_mlp_grad.py): 1e-4 for both forms (the bar
of the small-batch test)."""
import math

import numpy as np
import pytest
import torch

import diffrax_amd as dfx
import oracle
from diffrax_amd import _mlp
from diffrax_amd._sde import constant_grid
from oracle import pyoracle

pytestmark = pytest.mark.gpu

P_SDE, D_SDE, N_SDE = 262144, 16, 1000


@pytest.fixture(scope="module")
def sde_case(cuda):
    rng = np.random.default_rng(2024)
    theta = rng.uniform(0.5, 2.0, size=D_SDE)
    mu = rng.normal(0.0, 1.0, size=D_SDE)
    sigma = rng.uniform(0.1, 0.5, size=D_SDE)
    y0 = rng.normal(0.0, 0.3, size=(P_SDE, D_SDE))
    gen = torch.Generator(device=cuda).manual_seed(2024)
    dW = torch.empty((N_SDE, P_SDE, D_SDE), dtype=torch.float32, device=cuda)
    dW.normal_(0.0, math.sqrt(1.0 / N_SDE), generator=gen)
    rows = np.r_[0:16, P_SDE - 512:P_SDE]
    dW_rows = dW[:, torch.as_tensor(rows, device=cuda)].double().cpu().numpy()
    case = dict(theta=theta, mu=mu, sigma=sigma, y0=y0, dW=dW, rows=rows, dW_rows=dW_rows)
    del dW
    yield case
    case.clear()          # the 16.8 GB of increments go back before the next module
    torch.cuda.empty_cache()


def _sde_solve(c, dtype, grad):
    dev = c["dW"].device
    leaves = [torch.tensor(v, dtype=dtype, device=dev, requires_grad=grad)
              for v in (c["y0"], c["theta"], c["mu"], c["sigma"])]
    terms = dfx.MultiTerm(dfx.ODETerm(dfx.OUDrift()),
                          dfx.ControlTerm(dfx.DiagonalNoise(), dfx.IncrementPath(c["dW"])))
    sol = dfx.diffeqsolve(terms, dfx.Heun(), 0.0, 1.0, 1.0 / N_SDE, leaves[0],
                          tuple(leaves[1:]), saveat=dfx.SaveAt(t1=True),
                          stepsize_controller=dfx.ConstantStepSize(), max_steps=N_SDE)
    return sol, leaves


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_sde_full_size_forward(sde_case, dtype):
    c = sde_case
    with torch.no_grad():
        sol, _ = _sde_solve(c, dtype, grad=False)
    ys = sol.ys[torch.as_tensor(c["rows"], device=sol.ys.device)].double().cpu().numpy()
    q = np.concatenate([c["theta"], c["mu"], c["sigma"]])
    ref = oracle.solve(c["y0"][c["rows"]], 1.0, q, dW=c["dW_rows"], solver="heun", vf="ou",
                       controller="constant", dt0=1.0 / N_SDE, save_t1=True, max_steps=100000)
    if dtype == torch.float64:
        assert np.array_equal(ys, ref["ys"])
    else:
        err = np.max(np.abs(ys - ref["ys"])) / np.max(np.abs(ref["ys"]))
        print(f"full-size SDE f32 forward normwise rel err {err:.3e}")
        assert err < 1e-5


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_sde_full_size_gradient(sde_case, dtype, tol):
    c = sde_case
    last = slice(P_SDE - 512, P_SDE)
    w = np.zeros((P_SDE, 1, D_SDE))
    w[last] = np.random.default_rng(9).normal(size=(512, 1, D_SDE))
    sol, leaves = _sde_solve(c, dtype, grad=True)
    (sol.ys * torch.tensor(w, dtype=dtype, device=sol.ys.device)).sum().backward()
    got = [leaves[0].grad[last].double().cpu().numpy()] + \
          [x.grad.double().cpu().numpy() for x in leaves[1:]]
    # torch float64 autograd through the same fixed-step Heun replay of the last 512 paths
    grid = constant_grid(0.0, 1.0, 1.0 / N_SDE)
    steps = list(zip(grid[:-1], grid[1:]))
    Y0 = torch.tensor(c["y0"][last], requires_grad=True)
    TH, MU, SG = (torch.tensor(c[k], requires_grad=True) for k in ("theta", "mu", "sigma"))
    dWt = torch.tensor(c["dW_rows"][:, 16:])
    out = pyoracle.replay(lambda t, y: TH * (MU - y), "heun", steps, Y0, [],
                          g_dw=lambda n, t, y: SG * dWt[n], save_t1=True)
    (out[-1] * torch.tensor(w[last, 0])).sum().backward()
    for name, g, r in zip(("y0", "theta", "mu", "sigma"), got,
                          (Y0.grad, TH.grad, MU.grad, SG.grad)):
        r = r.numpy()
        rel = np.max(np.abs(g - r)) / np.max(np.abs(r))
        print(f"full-size SDE {dtype} gradient {name}: normwise rel err {rel:.3e}")
        assert rel < tol, name


def _mlp_weights():
    rng = np.random.default_rng(5)
    xav = lambda o, i: rng.normal(0.0, np.sqrt(2.0 / (i + o)), size=(o, i))
    return [xav(256, 64), rng.normal(0, 0.1, 256), xav(256, 256), rng.normal(0, 0.1, 256),
            xav(64, 256), rng.normal(0, 0.1, 64)]


@pytest.mark.parametrize("matmul,tol", [("f32", 1e-4), ("bf16", 1e-4)])
def test_neural_ode_full_batch_gradient(cuda, matmul, tol):
    B, n_ts = 8192, 32
    Wn = _mlp_weights()
    rng = np.random.default_rng(6)
    y0 = rng.normal(size=(B, 64)).astype(np.float32)
    ts = np.linspace(0.0, 1.0, n_ts)
    rows = np.sort(rng.choice(B, size=256, replace=False))
    w = np.zeros((B, n_ts, 64), dtype=np.float32)
    w[rows] = rng.normal(size=(256, n_ts, 64)).astype(np.float32)
    P = [torch.tensor(x, dtype=torch.float32, device=cuda, requires_grad=True) for x in Wn]
    y0t = torch.tensor(y0, device=cuda, requires_grad=True)
    ctrl = dfx.PIDController(rtol=1e-5, atol=1e-7)
    sol = dfx.diffeqsolve(dfx.ODETerm(dfx.MLPField(*P, matmul=matmul)), dfx.Tsit5(), 0.0, 1.0,
                          0.1, y0t, saveat=dfx.SaveAt(ts=ts), stepsize_controller=ctrl,
                          max_steps=64)
    (sol.ys * torch.tensor(w, device=cuda)).sum().backward()
    # the accepted steps of the same solve (the kernels are deterministic)
    plan = _mlp.prepare_mlp(dfx.ODETerm(dfx.MLPField(*[p.detach() for p in P], matmul=matmul)),
                            dfx.Tsit5(), 0.0, 1.0, 0.1, torch.tensor(y0), dfx.SaveAt(ts=ts),
                            ctrl, 64, cuda, ckpt_every=8)
    plan.launch()
    torch.cuda.synchronize()
    assert torch.equal(plan.ys, sol.ys.detach())
    grid = plan.grid.cpu().numpy()[rows]
    nacc = plan.stats[:, 1].cpu().numpy()[rows]
    Pd = [torch.tensor(np.asarray(x, dtype=np.float32).astype(np.float64), requires_grad=True)
          for x in Wn]
    W1, b1, W2, b2, W3, b3 = Pd
    sp = torch.nn.functional.softplus
    f = lambda t, y: torch.tanh(sp(sp(y @ W1.T + b1) @ W2.T + b2) @ W3.T + b3)
    Y0 = torch.tensor(y0[rows].astype(np.float64), requires_grad=True)
    loss = 0.0
    for i in range(len(rows)):
        steps = [tuple(grid[i, n]) for n in range(nacc[i])]
        out = pyoracle.replay(f, "tsit5", steps, Y0[i], list(ts))
        loss = loss + (torch.stack(out) * torch.tensor(w[rows[i]].astype(np.float64))).sum()
    loss.backward()
    got = [y0t.grad[torch.as_tensor(rows, device=cuda)].double().cpu()] + \
          [p.grad.double().cpu() for p in P]
    worst = 0.0
    for name, g, r in zip(("y0", "W1", "b1", "W2", "b2", "W3", "b3"), got, [Y0.grad] + [
            p.grad for p in Pd]):
        rel = float((g - r).abs().max() / r.abs().max())
        print(f"full-batch neural ODE {matmul} gradient {name}: rel err {rel:.3e}")
        worst = max(worst, rel) if rel == rel else float("inf")   # NaN fails
    assert worst < tol
    # rows without a cotangent get none
    others = np.setdiff1d(np.arange(B), rows)
    assert float(y0t.grad[torch.as_tensor(others, device=cuda)].abs().max()) == 0.0
