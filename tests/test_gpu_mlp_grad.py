"""GPU tests of the neural-ODE discrete adjoint (RecursiveCheckpointAdjoint semantics).

Reference gradient: torch float64 autograd through a replay of the same accepted step grid
(the reference's gradient is the derivative of exactly that replay: step-size factors are
stop_gradient, pid.py:498-530).  The replay itself is pinned by central differences along
random parameter directions.  Bar: 1e-4 relative (max |g - g_ref| / max |g_ref| per
parameter) for both forms; the bfloat16 form's adjoint runs its products on three-piece
bf16 operands (the float32-weight model, as its forward)."""
import numpy as np
import pytest
import torch

import diffrax_amd as dfx
from diffrax_amd import _mlp
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def _weights(seed=0):
    rng = np.random.default_rng(seed)
    xav = lambda o, i: rng.normal(0.0, np.sqrt(2.0 / (i + o)), size=(o, i))
    return [xav(256, 64), rng.normal(0, 0.1, 256), xav(256, 256), rng.normal(0, 0.1, 256),
            xav(64, 256), rng.normal(0, 0.1, 64)]


def _torch_mlp(P):
    W1, b1, W2, b2, W3, b3 = P
    sp = torch.nn.functional.softplus
    return lambda t, y: torch.tanh(sp(sp(y @ W1.T + b1) @ W2.T + b2) @ W3.T + b3)


def _setup(matmul, B=40, n_ts=8, seed=0, C=2):
    Wn = _weights(seed)
    rng = np.random.default_rng(seed + 7)
    y0 = rng.normal(size=(B, 64)).astype(np.float32)
    ts = np.linspace(0.0, 1.0, n_ts)
    wloss = rng.normal(size=(B, n_ts, 64)).astype(np.float32)
    return Wn, y0, ts, wloss


def _device_grads(Wn, y0, ts, wloss, matmul, C, max_steps=None):
    max_steps = max_steps or C * C
    dev = torch.device("cuda", 0)
    P = [torch.tensor(w, dtype=torch.float32, device=dev, requires_grad=True) for w in Wn]
    y0t = torch.tensor(y0, device=dev, requires_grad=True)
    vf = dfx.MLPField(*P, matmul=matmul)
    sol = dfx.diffeqsolve(dfx.ODETerm(vf), dfx.Tsit5(), 0.0, 1.0, 0.1, y0t,
                          saveat=dfx.SaveAt(ts=ts), max_steps=max_steps,
                          stepsize_controller=dfx.PIDController(rtol=1e-5, atol=1e-7),
                          adjoint=dfx.RecursiveCheckpointAdjoint(checkpoints=max_steps // C))
    loss = (sol.ys * torch.tensor(wloss, device=dev)).sum()
    loss.backward()
    # the accepted grid of the same solve (deterministic kernels: same steps)
    plan = _mlp.prepare_mlp(dfx.ODETerm(dfx.MLPField(*[p.detach() for p in P], matmul=matmul)),
                            dfx.Tsit5(), 0.0, 1.0, 0.1, torch.tensor(y0), dfx.SaveAt(ts=ts),
                            dfx.PIDController(rtol=1e-5, atol=1e-7), max_steps, dev, ckpt_every=C)
    plan.launch()
    torch.cuda.synchronize()
    assert torch.equal(plan.ys, sol.ys.detach())
    grid = plan.grid.cpu().numpy()
    nacc = plan.stats[:, 1].cpu().numpy()
    grads = [y0t.grad.double().cpu()] + [p.grad.double().cpu() for p in P]
    return grads, grid, nacc


def _replay_grads(Wn, y0, ts, wloss, grid, nacc):
    P = [torch.tensor(np.asarray(w, dtype=np.float32).astype(np.float64), requires_grad=True) for w in Wn]
    f = _torch_mlp(P)
    Y0 = torch.tensor(y0.astype(np.float64), requires_grad=True)
    loss = 0.0
    for b in range(y0.shape[0]):
        steps = [tuple(grid[b, n]) for n in range(nacc[b])]
        out = pyoracle.replay(f, "tsit5", steps, Y0[b], list(ts))
        loss = loss + (torch.stack(out) * torch.tensor(wloss[b].astype(np.float64))).sum()
    loss.backward()
    return [Y0.grad] + [p.grad for p in P], loss.item()


@pytest.mark.parametrize("matmul,tol", [("f32", 1e-4), ("bf16", 1e-4)])
def test_mlp_adjoint_against_fp64_replay(cuda, matmul, tol):
    C = 2   # checkpoint every 2 accepted steps: several segments per trajectory
    Wn, y0, ts, wloss = _setup(matmul, C=C)
    grads, grid, nacc = _device_grads(Wn, y0, ts, wloss, matmul, C, max_steps=32)
    print(f"{matmul}: accepted steps {nacc.min()}-{nacc.max()}")
    assert nacc.max() > C  # more than one segment
    ref, _ = _replay_grads(Wn, y0, ts, wloss, grid, nacc)
    names = ["y0", "W1", "b1", "W2", "b2", "W3", "b3"]
    worst = 0.0
    for name, g, r in zip(names, grads, ref):
        rel = float((g - r).abs().max() / r.abs().max())
        print(f"  {name}: rel err {rel:.3e}")
        worst = max(worst, rel) if rel == rel else float("inf")   # NaN fails
    assert worst < tol


def test_replay_gradient_matches_central_differences(cuda):
    """Pins the replay reference: d loss / d theta along random directions vs central FD."""
    Wn, y0, ts, wloss = _setup("f32", B=4)
    grads, grid, nacc = _device_grads(Wn, y0, ts, wloss, "f32", 2)
    ref, _ = _replay_grads(Wn, y0, ts, wloss, grid, nacc)
    rng = np.random.default_rng(11)
    base = [np.asarray(w, dtype=np.float32).astype(np.float64) for w in Wn]

    def L(params):
        P = [torch.tensor(p) for p in params]
        f = _torch_mlp(P)
        tot = 0.0
        for b in range(y0.shape[0]):
            steps = [tuple(grid[b, n]) for n in range(nacc[b])]
            out = pyoracle.replay(f, "tsit5", steps, torch.tensor(y0[b].astype(np.float64)), list(ts))
            tot += float((torch.stack(out) * torch.tensor(wloss[b].astype(np.float64))).sum())
        return tot

    h = 1e-6
    for _ in range(4):
        v = [rng.normal(size=p.shape) for p in base]
        nrm = np.sqrt(sum(float((x * x).sum()) for x in v))
        v = [x / nrm for x in v]
        fd = (L([p + h * x for p, x in zip(base, v)]) - L([p - h * x for p, x in zip(base, v)])) / (2 * h)
        an = sum(float((r.numpy() * x).sum()) for r, x in zip(ref[1:], v))
        print("FD", fd, "replay", an)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))
