"""Host-side SaveAt logic (CPU): SaveAt(fn=...) applied to saved rows
(_integrate.py:244-275, 456-527: fn(t, y, args) in the caller's time, rows the solve did
not fill keep the inf preset of 1272-1305), and the SubSaveAt container checks."""
import numpy as np
import pytest
import torch

import diffrax_amd as dfx
from diffrax_amd._integrate import _apply_save_fn


def test_fn_rows_batched_with_per_instance_args():
    B, n, d = 3, 4, 2
    ts = torch.arange(B * n, dtype=torch.float64).reshape(B, n)
    ys = torch.arange(B * n * d, dtype=torch.float64).reshape(B, n, d)
    ts[1, 3] = float("inf")
    ys[1, 3] = float("inf")
    scale = torch.tensor([1.0, 2.0, 3.0], dtype=torch.float64)      # per instance
    shift = 0.5                                                      # shared
    out = _apply_save_fn(lambda t, y, args: args[0] * y[0] + t + args[1], ts, ys,
                         (scale, shift), batched=True)
    ref = scale[:, None] * ys[..., 0] + ts + shift
    ref[1, 3] = float("inf")
    assert torch.equal(out, ref)


def test_fn_pytree_output_unbatched_and_grad():
    ts = torch.tensor([0.0, 1.0, float("inf")], dtype=torch.float64)
    y = torch.tensor([[1.0, 2.0], [3.0, 4.0], [float("inf")] * 2], dtype=torch.float64,
                     requires_grad=True)
    out = _apply_save_fn(lambda t, y, args: {"s": y.sum(), "p": y * t}, ts, y, None,
                         batched=False)
    assert torch.equal(out["s"], torch.tensor([3.0, 7.0, float("inf")], dtype=torch.float64))
    assert torch.isinf(out["p"][2]).all()
    g, = torch.autograd.grad(out["s"][:2].sum() + out["p"][:2].sum(), (y,))
    assert torch.equal(g[:2], torch.tensor([[1.0, 1.0], [2.0, 2.0]], dtype=torch.float64))


def test_subs_container_checks():
    with pytest.raises(ValueError, match="Empty saveat"):
        dfx.SubSaveAt()
    with pytest.raises(ValueError):
        dfx.diffeqsolve(dfx.ODETerm(dfx.LotkaVolterra()), dfx.Tsit5(), 0.0, 1.0, 0.1,
                        torch.ones(2, dtype=torch.float64), saveat=dfx.SaveAt(subs=[1, 2]))
    _ = np
