"""GPU tests of SaveAt(fn=...), SaveAt(subs=[...]) and SaveAt(steps=k) through diffeqsolve
(_saveat.py:14-105; _integrate.py:244-275, 456-527, 847-877, 1272-1305).

- fn: rows equal fn applied to the plain solve's rows (bitwise), per-instance args batched
  like the reference's vmap; gradients flow through fn into the discrete adjoint and equal
  those of the plain solve with fn applied by hand;
- subs: a list / dict of SubSaveAts returns ts / ys in the same container, each equal to
  the solve with that SubSaveAt alone;
- steps=k: bit-identical to the FP64 oracle (rows and the t1 rule)."""
import numpy as np
import pytest
import torch

import diffrax_amd as dfx
import oracle

pytestmark = pytest.mark.gpu

Q = (1.5, 1.0, 3.0, 1.0)


def _y0(B=32, seed=0):
    return torch.tensor(np.random.default_rng(seed).uniform(0.5, 2.0, size=(B, 2)),
                        dtype=torch.float64)


def _solve(saveat, y0, args=Q, **kw):
    return dfx.diffeqsolve(dfx.ODETerm(dfx.LotkaVolterra()), dfx.Tsit5(), 0.0, 5.0, 0.1, y0,
                           args, saveat=saveat,
                           stepsize_controller=dfx.PIDController(rtol=1e-6, atol=1e-8), **kw)


def test_fn_rows(cuda):
    y0 = _y0()
    ts = np.linspace(0.0, 5.0, 11)
    fn = lambda t, y, args: torch.stack([y[0] * y[1], t + args[0] * y[0]])
    got = _solve(dfx.SaveAt(t0=True, ts=ts, t1=True, fn=fn), y0)
    plain = _solve(dfx.SaveAt(t0=True, ts=ts, t1=True), y0)
    ys, t = plain.ys, plain.ts
    ref = torch.stack([ys[..., 0] * ys[..., 1], t + Q[0] * ys[..., 0]], -1)
    assert torch.equal(got.ys, ref) and torch.equal(got.ts, plain.ts)


def test_fn_per_instance_args_and_unfilled_rows(cuda):
    y0 = _y0()
    alpha = torch.linspace(1.2, 1.8, 32, dtype=torch.float64)
    args = (alpha, 1.0, 3.0, 1.0)
    got = _solve(dfx.SaveAt(steps=True, fn=lambda t, y, a: a[0] * y.sum()), y0, args,
                 max_steps=512)
    plain = _solve(dfx.SaveAt(steps=True), y0, args, max_steps=512)
    filled = torch.isfinite(plain.ts)
    ref = torch.where(filled, alpha.to(plain.ys.device)[:, None] * plain.ys.sum(-1),
                      torch.full_like(plain.ts, float("inf")))
    assert torch.equal(got.ys, ref)
    assert (~filled).any()


def test_fn_gradient_through_discrete_adjoint(cuda):
    y0 = _y0()
    fn = lambda t, y, args: (y * y).sum()
    ts = np.linspace(0.0, 5.0, 6)
    a = y0.clone().cuda().requires_grad_(True)
    g1, = torch.autograd.grad(_solve(dfx.SaveAt(ts=ts, fn=fn), a).ys.sum(), (a,))
    b = y0.clone().cuda().requires_grad_(True)
    g2, = torch.autograd.grad((_solve(dfx.SaveAt(ts=ts), b).ys ** 2).sum(), (b,))
    torch.testing.assert_close(g1, g2, rtol=1e-13, atol=0)


def test_subs_list_and_dict(cuda):
    y0 = _y0()
    ts = np.linspace(0.0, 5.0, 7)
    subs = [dfx.SubSaveAt(ts=ts), dfx.SubSaveAt(steps=True),
            dfx.SubSaveAt(t1=True, fn=lambda t, y, args: y[0])]
    got = _solve(dfx.SaveAt(subs=subs), y0)
    assert isinstance(got.ys, list) and len(got.ys) == 3
    for sub, ys, t in zip(subs, got.ys, got.ts):
        one = _solve(dfx.SaveAt(subs=sub), y0)
        assert torch.equal(ys, one.ys) and torch.equal(t, one.ts)
    got = _solve(dfx.SaveAt(subs={"a": dfx.SubSaveAt(t0=True), "b": dfx.SubSaveAt(ts=ts)}), y0)
    assert set(got.ys) == {"a", "b"}
    assert torch.equal(got.ys["a"][:, 0].cpu(), y0)


@pytest.mark.parametrize("k", [2, 5])
def test_steps_k_against_oracle(cuda, k):
    y0 = _y0(64, seed=k)
    sol = _solve(dfx.SaveAt(steps=k, t1=True), y0, max_steps=501)
    ref = oracle.solve(y0.numpy(), 5.0, np.array(Q), solver="tsit5", vf="lv", dt0=0.1,
                       rtol=1e-6, atol=1e-8, save_steps=k, save_t1=True, max_steps=501)
    assert np.array_equal(sol.ys.cpu().numpy(), ref["ys"])
    assert np.array_equal(sol.ts.cpu().numpy(), ref["ts"])


def test_device_resident_times_match_host_and_are_checked(cuda):
    """t1 and SaveAt(ts) given as device tensors: the kernel reads them in place and the
    reference's checks (_integrate.py:1107-1112, 1222-1235) run from pinned copies after
    the solve is queued; results equal those of host inputs, bit for bit, and invalid
    times still raise the reference's messages."""
    y0 = _y0(B=16, seed=3)
    t1 = torch.tensor(np.random.default_rng(4).uniform(3.0, 5.0, size=16))
    ts = torch.linspace(0.0, 1.0, 9, dtype=torch.float64)[None] * t1[:, None]
    ctrl = dfx.PIDController(rtol=1e-6, atol=1e-8)
    term = dfx.ODETerm(dfx.LotkaVolterra())

    def solve(t1_, ts_, dt0=0.1):
        return dfx.diffeqsolve(term, dfx.Tsit5(), 0.0, t1_, dt0, y0, Q,
                               saveat=dfx.SaveAt(ts=ts_, t1=True), stepsize_controller=ctrl)
    host = solve(t1, ts)
    dev = solve(t1.to(cuda), ts.to(cuda))
    assert torch.equal(host.ys, dev.ys) and torch.equal(host.ts, dev.ts)
    # one device-resident t1 for every instance
    shared = solve(torch.tensor([4.0], dtype=torch.float64, device=cuda),
                   torch.linspace(0.0, 4.0, 9, dtype=torch.float64, device=cuda))
    ref = solve(4.0, np.linspace(0.0, 4.0, 9))
    assert torch.equal(shared.ys, ref.ys)
    bad = ts.clone()
    bad[3, 4] = bad[3, 6]   # above its successor
    with pytest.raises(RuntimeError, match="increasing or decreasing"):
        solve(t1.to(cuda), bad.to(cuda))
    with pytest.raises(RuntimeError, match="between t0 and t1"):
        solve(t1.to(cuda), (ts * 1.5).to(cuda))
    with pytest.raises(RuntimeError, match=r"\(t1 - t0\) \* dt0 >= 0"):
        solve(t1.to(cuda), ts.to(cuda), dt0=-0.1)
