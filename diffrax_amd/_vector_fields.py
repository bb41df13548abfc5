"""Registered device vector fields.

diffrax evaluates an arbitrary user callable ``vector_field(t, y, args)``
(_term.py:193-211).  On the GPU the callable is one of the functions below, compiled
into the kernels; its parameters travel as a flat per-instance (or shared) array.
Each object is also a plain torch callable with the same formula, so the term can be
inspected on the host.  Parameters come from ``args`` (as a vmapped reference call
would pass them: scalars shared, leading-axis-[B] tensors per instance) or, if args
is None, from the values bound at construction."""
from __future__ import annotations

import math

import torch

from . import _lib


def _as_batched(values, batch, dtype, device):
    """Stack scalars / [B] tensors into a [P] (shared) or [B, P] (per-instance) tensor."""
    cols = [torch.as_tensor(v, dtype=dtype, device=device) for v in values]
    if all(c.ndim == 0 for c in cols):
        return torch.stack(cols).contiguous(), 0
    cols = [c.expand(batch) if c.ndim == 0 else c.reshape(batch) for c in cols]
    out = torch.stack(cols, dim=-1).contiguous()
    return out, out.shape[-1]


class RegisteredVectorField:
    vf_id: int = -1
    name: str = "?"

    def supports(self, d: int) -> bool:
        raise NotImplementedError

    def pack(self, args, batch, d, dtype, device):
        """Return (params tensor, param_stride) for the device kernels."""
        raise NotImplementedError


class LotkaVolterra(RegisteredVectorField):
    """dx = alpha x - beta x y, dy = delta x y - gamma y; args = (alpha, beta, gamma, delta)."""

    vf_id = _lib.VF_LV
    name = "lotka_volterra"

    def __init__(self, alpha=1.5, beta=1.0, gamma=3.0, delta=1.0):
        self.bound = (alpha, beta, gamma, delta)

    def supports(self, d):
        return d == 2

    def _p(self, args):
        return self.bound if args is None else tuple(args)

    def __call__(self, t, y, args=None):
        a, b, g, dl = self._p(args)
        x, v = y[..., 0], y[..., 1]
        return torch.stack([a * x - b * x * v, dl * x * v - g * v], dim=-1)

    def pack(self, args, batch, d, dtype, device):
        return _as_batched(self._p(args), batch, dtype, device)


class VanDerPol(RegisteredVectorField):
    """y'' = mu (1 - y^2) y' - y as (y, y'); args = mu (scalar or [B])."""

    vf_id = _lib.VF_VDP
    name = "van_der_pol"

    def __init__(self, mu=1.0):
        self.bound = mu

    def supports(self, d):
        return d == 2

    def __call__(self, t, y, args=None):
        mu = self.bound if args is None else args
        if isinstance(mu, torch.Tensor) and mu.ndim == 1 and y.ndim == 2:
            mu = mu[:, None][:, 0]
        x, v = y[..., 0], y[..., 1]
        return torch.stack([v, mu * (1 - x * x) * v - x], dim=-1)

    def pack(self, args, batch, d, dtype, device):
        mu = self.bound if args is None else args
        return _as_batched((mu,), batch, dtype, device)


class LinearField(RegisteredVectorField):
    """f = A y; args = A ([d, d] shared or [B, d, d] per instance)."""

    vf_id = _lib.VF_LINEAR
    name = "linear"

    def __init__(self, A=None):
        self.bound = A

    def supports(self, d):
        return d in (1, 2, 3, 4, 10)

    def __call__(self, t, y, args=None):
        A = torch.as_tensor(self.bound if args is None else args, dtype=y.dtype, device=y.device)
        return (A @ y[..., None])[..., 0]

    def pack(self, args, batch, d, dtype, device):
        A = torch.as_tensor(self.bound if args is None else args, dtype=dtype, device=device)
        if A.ndim == 2:
            if A.shape != (d, d):
                raise ValueError(f"A must be [{d}, {d}], got {tuple(A.shape)}")
            return A.reshape(-1).contiguous(), 0
        if A.shape != (batch, d, d):
            raise ValueError(f"A must be [{batch}, {d}, {d}], got {tuple(A.shape)}")
        return A.reshape(batch, d * d).contiguous(), d * d


_ = math


class OUDrift(RegisteredVectorField):
    """Ornstein-Uhlenbeck drift theta * (mu - y); args = (theta, mu, sigma), each [d]."""

    vf_id = _lib.VF_OU
    name = "ou_drift"

    def supports(self, d):
        return True

    def __call__(self, t, y, args):
        theta, mu, _ = args
        return theta * (mu - y)


class DiagonalNoise(RegisteredVectorField):
    """Diagonal diffusion g(t, y) = diag(sigma); args = (theta, mu, sigma).  Used as the
    vector field of a ControlTerm, whose product with dW is sigma * dW."""

    vf_id = _lib.VF_OU
    name = "ou_diagonal_noise"

    def supports(self, d):
        return True

    def __call__(self, t, y, args):
        return args[2] * torch.ones_like(y)


class MLPField(RegisteredVectorField):
    """Neural vector field f(t, y) = tanh(W3 softplus(W2 softplus(W1 y + b1) + b2) + b3)
    with d = 64, width 256, depth 2 -- what eqx.nn.MLP(64, 64, 256, 2, activation=softplus,
    final_activation=tanh) computes.  W1 [256, 64], W2 [256, 256], W3 [64, 256] (torch
    Linear layout: out x in).  ``matmul`` selects the device form: "f32" (float32 MFMA) or
    "bf16" (bfloat16 MFMA operands, float32 accumulate); the RK algebra is float32 in both
    and time / step control float64."""

    vf_id = _lib.VF_MLP
    name = "mlp"

    def __init__(self, W1, b1, W2, b2, W3, b3, matmul="f32"):
        if matmul not in ("f32", "bf16"):
            raise ValueError("matmul must be 'f32' or 'bf16'")
        self.params = [torch.as_tensor(p) for p in (W1, b1, W2, b2, W3, b3)]
        shapes = [(256, 64), (256,), (256, 256), (256,), (64, 256), (64,)]
        for p, s in zip(self.params, shapes):
            if tuple(p.shape) != s:
                raise ValueError(f"MLP parameter of shape {tuple(p.shape)}, expected {s}")
        self.matmul = matmul

    def supports(self, d):
        return d == 64

    def __call__(self, t, y, args=None):
        W1, b1, W2, b2, W3, b3 = (p.to(y) for p in self.params)
        sp = torch.nn.functional.softplus
        return torch.tanh(sp(sp(y @ W1.T + b1) @ W2.T + b2) @ W3.T + b3)
