"""GPU tests of BacksolveAdjoint on the neural-ODE kernel (dfx_mlp_backsolve): the
continuous adjoint (_adjoint.py:646-735) of the MLP 64 -> 256 -> 256 -> 64 field with the
adjoint seminorm (_misc.py:85-116) as the backward controller's norm.

- Against the FP64 oracle's continuous adjoint (pyoracle.backsolve, pinned on the CPU
  against the C oracle, the closed form and the reference's seminorm recipe), driven by
  torch float64 vjps of the same float32-weight MLP from the same forward saves, on rows
  that carry a cotangent (the weight gradients come from those rows alone).  The kernel
  computes in float32, so step-size decisions can differ from the float64 oracle's near
  the acceptance threshold; the bar is 1e-4 normwise.
- Against the discrete adjoint (RecursiveCheckpointAdjoint) on the same smooth solve: the
  two adjoints are different discretisations of one gradient.  The reference holds them
  within atol 1e-5 and rtol 1e-5 elementwise in float64 (test_adjoint.py:156-159); the
  kernels sum in float32 over 64 trajectories, so the bar here is 1e-5 normwise
  (max |diff| / max |g|), the elementwise figure printed beside it.
- Guards: the default norm, and a non-Tsit5 backward solver, raise NotImplementedError.
"""
import numpy as np
import pytest
import torch

import diffrax_amd as dfx
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def _weights(seed=11):
    rng = np.random.default_rng(seed)
    xav = lambda o, i: rng.normal(0.0, np.sqrt(2.0 / (i + o)), size=(o, i))
    return [xav(256, 64), rng.normal(0, 0.1, 256), xav(256, 256), rng.normal(0, 0.1, 256),
            xav(64, 256), rng.normal(0, 0.1, 64)]


def _aug_field(W):
    """(f, vjp_y(-a), vjp_W(-a)) of the float32-weight MLP in torch float64; parameters in
    the order W1, b1, W2, b2, W3, b3, flattened."""
    P = [torch.tensor(np.asarray(w, dtype=np.float32).astype(np.float64)) for w in W]
    sp = torch.nn.functional.softplus

    def f(t, z):
        y = torch.tensor(z[:64], requires_grad=True)
        Pr = [p.clone().requires_grad_(True) for p in P]
        W1, b1, W2, b2, W3, b3 = Pr
        fy = torch.tanh(W3 @ sp(W2 @ sp(W1 @ y + b1) + b2) + b3)
        g = torch.autograd.grad(fy, [y] + Pr, -torch.tensor(z[64:128]))
        return np.concatenate([fy.detach().numpy(), g[0].numpy()] +
                              [x.numpy().ravel() for x in g[1:]])
    return f


def _solve(W, y0, ts, adjoint, matmul="f32", fwd_tol=(1e-5, 1e-7), dev="cuda"):
    P = [torch.tensor(w, dtype=torch.float32, device=dev, requires_grad=True) for w in W]
    y0t = torch.tensor(y0, device=dev, requires_grad=True)
    sol = dfx.diffeqsolve(dfx.ODETerm(dfx.MLPField(*P, matmul=matmul)), dfx.Tsit5(), 0.0, 1.0,
                          0.1, y0t, saveat=dfx.SaveAt(ts=ts),
                          stepsize_controller=dfx.PIDController(*fwd_tol),
                          adjoint=adjoint, max_steps=256)
    return sol, y0t, P


def _seminorm_adjoint(rtol=1e-6, atol=1e-8):
    return dfx.BacksolveAdjoint(stepsize_controller=dfx.PIDController(
        rtol, atol, norm=dfx.adjoint_rms_seminorm))


@pytest.mark.parametrize("matmul", ["f32", "bf16"])
def test_backsolve_against_fp64_oracle(cuda, matmul):
    W = _weights()
    rng = np.random.default_rng(12)
    B, rows = 40, [0, 13, 39]
    y0 = rng.normal(size=(B, 64)).astype(np.float32)
    ts = np.linspace(0.0, 1.0, 6)
    w = np.zeros((B, ts.size, 64), dtype=np.float32)
    w[rows] = rng.normal(size=(len(rows), ts.size, 64)).astype(np.float32)
    sol, y0t, P = _solve(W, y0, ts, _seminorm_adjoint(), matmul)
    (sol.ys * torch.tensor(w, device=cuda)).sum().backward()
    ys = sol.ys.detach().double().cpu().numpy()
    tso = sol.ts.detach().cpu().numpy()
    f_aug = _aug_field(W)
    ref_y, ref_q = [], np.zeros(98880)
    for r in rows:
        ay, aq, worst, _ = pyoracle.backsolve(f_aug, 64, 98880, "tsit5", 0.0, -0.1, tso[r],
                                              ys[r], w[r].astype(np.float64), save_t0=False,
                                              rtol=1e-6, atol=1e-8, seminorm=True, fused=False)
        assert worst == 0
        ref_y.append(ay)
        ref_q += aq
    got_y = y0t.grad.double().cpu().numpy()[rows]
    rel = np.max(np.abs(got_y - np.array(ref_y))) / np.max(np.abs(ref_y))
    print(f"{matmul}: y0 gradient vs FP64 oracle backsolve: {rel:.3e}")
    assert rel < 1e-4
    o = 0
    for name, p in zip(("W1", "b1", "W2", "b2", "W3", "b3"), P):
        n = p.numel()
        r = ref_q[o:o + n].reshape(p.shape)
        o += n
        g = p.grad.double().cpu().numpy()
        e = np.max(np.abs(g - r)) / np.max(np.abs(r))
        print(f"{matmul}: {name} gradient vs FP64 oracle backsolve: {e:.3e}")
        assert e < 1e-4, name
    others = np.setdiff1d(np.arange(B), rows)
    assert float(y0t.grad[torch.as_tensor(others, device=cuda)].abs().max()) == 0.0


def test_backsolve_against_discrete_adjoint(cuda):
    W = _weights(21)
    rng = np.random.default_rng(22)
    B = 64
    y0 = rng.normal(size=(B, 64)).astype(np.float32)
    ts = np.linspace(0.0, 1.0, 5)
    w = rng.normal(size=(B, ts.size, 64)).astype(np.float32)
    grads = {}
    for name, adj in (("discrete", dfx.RecursiveCheckpointAdjoint()),
                      ("backsolve", _seminorm_adjoint(1e-7, 1e-9))):
        sol, y0t, P = _solve(W, y0, ts, adj, fwd_tol=(1e-7, 1e-9))
        (sol.ys * torch.tensor(w, device=cuda)).sum().backward()
        grads[name] = [y0t.grad.double().cpu().numpy()] + [p.grad.double().cpu().numpy()
                                                            for p in P]
    worst = 0.0
    for nm, a, b in zip(("y0", "W1", "b1", "W2", "b2", "W3", "b3"), grads["backsolve"],
                        grads["discrete"]):
        e = np.max(np.abs(a - b)) / np.max(np.abs(b))
        ew = np.max(np.abs(a - b) / (1.0 + np.abs(b)))
        print(f"backsolve vs discrete adjoint, {nm}: normwise {e:.3e}, "
              f"elementwise |diff| / (1 + |g|) {ew:.3e}")
        worst = max(worst, e) if e == e else float("inf")   # NaN fails
    assert worst < 1e-5


def test_backsolve_guards(cuda):
    W = _weights()
    y0 = np.zeros((4, 64), dtype=np.float32)
    with pytest.raises(NotImplementedError, match="adjoint seminorm"):
        sol, _, _ = _solve(W, y0, [0.5, 1.0], dfx.BacksolveAdjoint())
        sol.ys.sum().backward()
    bad = dfx.BacksolveAdjoint(solver=dfx.Dopri5(), stepsize_controller=dfx.PIDController(
        1e-6, 1e-8, norm=dfx.adjoint_rms_seminorm))
    with pytest.raises(NotImplementedError, match="Tsit5"):
        sol, _, _ = _solve(W, y0, [0.5, 1.0], bad)
        sol.ys.sum().backward()
    # the backward solves restart from the returned ys: editing them in place raises
    sol, _, _ = _solve(W, y0, [0.5, 1.0], _seminorm_adjoint())
    loss = sol.ys.sum()
    sol.ys.detach().mul_(2.0)   # shares storage and version counter with sol.ys
    with pytest.raises(RuntimeError, match="modified by an inplace operation"):
        loss.backward()


def test_backsolve_launch_split(cuda):
    """Rows finish in different launches: the result does not depend on how many lock-step
    iterations a launch runs (y0 cotangents bit for bit; the weight sums, added launch by
    launch, to float32 rounding)."""
    from diffrax_amd import _mlp
    W = _weights(31)
    rng = np.random.default_rng(32)
    B = 64
    y0 = rng.normal(size=(B, 64)).astype(np.float32)
    ts = np.linspace(0.0, 1.0, 5)
    ctrl = dfx.PIDController(1e-7, 1e-9)
    vf = dfx.MLPField(*[torch.tensor(w, dtype=torch.float32) for w in W], matmul="f32")
    plan = _mlp.prepare_mlp(dfx.ODETerm(vf), dfx.Tsit5(), 0.0, 1.0, 0.1, torch.tensor(y0),
                            dfx.SaveAt(ts=ts), ctrl, 256, torch.device("cuda", 0))
    plan.launch()
    back = _mlp.backsolve_problem(plan, _seminorm_adjoint(1e-7, 1e-9))
    w = torch.tensor(rng.normal(size=(B, ts.size, 64)).astype(np.float32), device=cuda)
    runs = {k: _mlp.backsolve_mlp(plan, back, w, chunk=k) for k in (1, 3, 64)}
    ref, res = runs[64]
    assert int(res.abs().max()) == 0
    for k in (1, 3):
        g, r = runs[k]
        assert torch.equal(g[0], ref[0]) and torch.equal(r, res)
        for a, b in zip(g[1:], ref[1:]):
            e = float((a - b).abs().max() / b.abs().max())
            assert e < 1e-5, (k, e)
