"""Writes the golden fixtures in this folder from the FP64 oracle.

Each fixture is a small seeded subset of one benchmark configuration: inputs plus the
oracle's saved states and step statistics.  The CPU test suite regenerates them and
requires bit equality (drift of the oracle fails; the gradient fixture, which goes through
torch float64 autograd, to 1e-12 relative), the GPU suite compares the kernels with them.
Run from the repository root:  python tests/golden/make_golden.py

  c1_dopri5_lv_b16   Dopri5 + PID Lotka-Volterra, 16 instances
  c2_tsit5_vdp_b16   Tsit5 + PID uneven Van der Pol, 16 instances
  c3_mlp_tsit5_b8    neural ODE (64 -> 256 -> 256 -> 64) forward, 8 instances, with a
                     tight (rtol = atol = 1e-12) solve and the accepted step grid
  c4_mlp_grad_b4     gradient of sum(ys * w) through the accepted grid (discrete adjoint
                     semantics): dL/dy0 and dL/dtheta projected on 32 seeded unit directions
  c5_heun_ou_p8      Heun OU SDE, 8 paths x 16 dims x 100 steps
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402


def c1(seed=100, B=16):
    rng = np.random.default_rng(seed)
    y0 = rng.uniform(0.5, 2.0, size=(B, 2))
    q = np.array([1.5, 1.0, 3.0, 1.0])
    ts = np.linspace(0.0, 10.0, 100)
    r = oracle.solve(y0, 10.0, q, ts=ts, solver="dopri5", vf="lv", dt0=0.1, rtol=1e-6, atol=1e-8)
    return dict(y0=y0, q=q, ts=ts, ys=r["ys"], stats=r["stats"], result=r["result"])


def c2(seed=101, B=16):
    rng = np.random.default_rng(seed)
    mu = np.exp(rng.uniform(np.log(0.1), np.log(5.0), size=B))
    y0 = rng.uniform(-2.0, 2.0, size=(B, 2))
    t1 = rng.uniform(5.0, 50.0, size=B)
    ts = np.linspace(0.0, 1.0, 64)[None, :] * t1[:, None]
    r = oracle.solve(y0, t1, mu[:, None], ts=ts, solver="tsit5", vf="vdp", dt0=0.1,
                     rtol=1e-6, atol=1e-9, max_steps=16384)
    return dict(mu=mu, y0=y0, t1=t1, ts=ts, ys=r["ys"], stats=r["stats"], result=r["result"])


def c5(seed=102, P=8, d=16, N=100):
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0.5, 2.0, size=d)
    mu = rng.normal(0.0, 1.0, size=d)
    sigma = rng.uniform(0.1, 0.5, size=d)
    dW = (rng.standard_normal(size=(N, P, d)) * np.sqrt(1.0 / N)).astype(np.float32)
    y0 = np.zeros((P, d))
    q = np.concatenate([theta, mu, sigma])
    r = oracle.solve(y0, 1.0, q, dW=dW.astype(np.float64), solver="heun", vf="ou",
                     controller="constant", dt0=1.0 / N, save_t1=True, max_steps=10 * N)
    return dict(theta=theta, mu=mu, sigma=sigma, dW=dW, y0=y0, ys=r["ys"], stats=r["stats"])


def mlp_weights(rng):
    xav = lambda o, i: rng.normal(0.0, np.sqrt(2.0 / (i + o)), size=(o, i)).astype(np.float32)
    W1, W2, W3 = xav(256, 64), xav(256, 256), xav(64, 256)
    b1, b2, b3 = (rng.normal(0.0, 0.1, size=n).astype(np.float32) for n in (256, 256, 64))
    return W1, b1, W2, b2, W3, b3


MLP_KEYS = ("W1", "b1", "W2", "b2", "W3", "b3")


def _mlp_solve(W, y0, ts, **kw):
    q = oracle.mlp_pack(*[np.asarray(w, dtype=np.float64) for w in W])
    return oracle.solve(np.asarray(y0, dtype=np.float64), 1.0, q, ts=ts, solver="tsit5",
                        vf="mlp", hidden=256, dt0=0.1, **kw)


def c3(seed=103, B=8, n_ts=8):
    rng = np.random.default_rng(seed)
    W = mlp_weights(rng)
    y0 = rng.normal(size=(B, 64)).astype(np.float32)
    ts = np.linspace(0.0, 1.0, n_ts)
    r = _mlp_solve(W, y0, ts, rtol=1e-5, atol=1e-7, record_grid=True)
    tight = _mlp_solve(W, y0, ts, rtol=1e-12, atol=1e-12, max_steps=100000)
    n = int(r["stats"][:, 1].max())
    out = dict(zip(MLP_KEYS, W))
    out.update(y0=y0, ts=ts, ys=r["ys"], stats=r["stats"], grid=r["grid"][:, :n],
               ys_tight=tight["ys"])
    return out


def mlp_replay_loss(W, y0, ts, grid, nacc, wloss):
    """sum(ys * wloss) of the float64 replay of the accepted grid (torch tensors W, y0)."""
    import torch
    from oracle import pyoracle
    W1, b1, W2, b2, W3, b3 = W
    sp = torch.nn.functional.softplus
    f = lambda t, y: torch.tanh(sp(sp(y @ W1.T + b1) @ W2.T + b2) @ W3.T + b3)
    loss = 0.0
    for b in range(y0.shape[0]):
        steps = [tuple(grid[b, n]) for n in range(int(nacc[b]))]
        out = pyoracle.replay(f, "tsit5", steps, y0[b], list(ts))
        loss = loss + (torch.stack(out) * torch.as_tensor(wloss[b])).sum()
    return loss


def unit_directions(seed, n_dir, sizes):
    rng = np.random.default_rng(seed)
    dirs = []
    for _ in range(n_dir):
        v = [rng.normal(size=s_) for s_ in sizes]
        nrm = np.sqrt(sum(float((x * x).sum()) for x in v))
        dirs.append([x / nrm for x in v])
    return dirs


def c4(seed=104, B=4, n_ts=8, n_dir=32):
    import torch
    rng = np.random.default_rng(seed)
    W = mlp_weights(rng)
    y0 = rng.normal(size=(B, 64)).astype(np.float32)
    ts = np.linspace(0.0, 1.0, n_ts)
    wloss = rng.normal(size=(B, n_ts, 64))
    r = _mlp_solve(W, y0, ts, rtol=1e-5, atol=1e-7, record_grid=True)
    nacc = r["stats"][:, 1]
    grid = r["grid"][:, :int(nacc.max())]
    Wt = [torch.tensor(w.astype(np.float64), requires_grad=True) for w in W]
    Y0 = torch.tensor(y0.astype(np.float64), requires_grad=True)
    loss = mlp_replay_loss(Wt, Y0, ts, grid, nacc, wloss)
    loss.backward()
    gW = [w.grad.numpy() for w in Wt]
    dirs = unit_directions(seed + 1, n_dir, [w.shape for w in W])
    proj = np.array([sum(float((g * v).sum()) for g, v in zip(gW, d)) for d in dirs])
    out = dict(zip(MLP_KEYS, W))
    out.update(y0=y0, ts=ts, wloss=wloss, grid=grid, nacc=nacc, loss=np.float64(loss.item()),
               y0_bar=Y0.grad.numpy(), dir_seed=np.int64(seed + 1), dir_derivs=proj)
    return out


BUILDERS = {"c1_dopri5_lv_b16": c1, "c2_tsit5_vdp_b16": c2, "c3_mlp_tsit5_b8": c3,
            "c4_mlp_grad_b4": c4, "c5_heun_ou_p8": c5}
EXACT = {"c1_dopri5_lv_b16", "c2_tsit5_vdp_b16", "c3_mlp_tsit5_b8", "c5_heun_ou_p8"}


def main():
    for name, fn in BUILDERS.items():
        np.savez(os.path.join(HERE, name + ".npz"), **fn())
        print("wrote", name)


if __name__ == "__main__":
    main()
