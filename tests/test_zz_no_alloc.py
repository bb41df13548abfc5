"""The allocation-free SDE adjoint (dfx_sde_heun_adjoint_static, partial sums in the
library's static device array) against the workspace entry point (dfx_sde_heun_adjoint),
through the C ABI: bit for bit in float32 and float64, on two calls in a row with the
second problem smaller than the first (the array keeps the first call's partials), and
the size limit (a problem needing more than DFX_SDE_STATIC_SCRATCH_BYTES is rejected
before any launch)."""
import ctypes
import math

import numpy as np
import pytest
import torch

from diffrax_amd import _lib, _sde

pytestmark = pytest.mark.gpu


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _forward(P, d, N, seed, dtype):
    """Inputs of one OU problem and its forward solve with checkpoints (C ABI)."""
    lib = _lib.load()
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(seed)
    t = lambda a: torch.tensor(a, dtype=dtype, device=dev)
    keep = dict(theta=t(rng.uniform(0.5, 2.0, d)), mu=t(rng.normal(0.0, 1.0, d)),
                sigma=t(rng.uniform(0.1, 0.5, d)), y0=t(rng.normal(0.0, 0.3, (P, d))),
                dW=torch.tensor((rng.standard_normal((N, P, d)) * math.sqrt(1.0 / N))
                                .astype(np.float32), device=dev),
                grid=torch.tensor(_sde.constant_grid(0.0, 1.0, 1.0 / N), device=dev))
    prec = _lib.F32 if dtype == torch.float32 else _lib.F64
    prob = _sde.SdeProblem(precision=prec, d=d, paths=P, num_steps=N, checkpoint_every=32,
                           save_t0=0, save_t1=1)
    n_ck = lib.dfx_sde_n_ckpt(ctypes.byref(prob))
    keep["ckpt"] = torch.empty((max(n_ck, 1), P, d), dtype=dtype, device=dev)
    keep["ys"] = torch.empty((P, 1, d), dtype=dtype, device=dev)
    io = _sde.SdeIO(**{k: v.data_ptr() for k, v in keep.items()})
    assert lib.dfx_sde_heun_solve(ctypes.byref(prob), ctypes.byref(io), _stream()) == 0
    keep["ys_bar"] = t(rng.normal(size=(P, 1, d)))
    return prob, io, keep


def _adjoint(prob, io, keep, static):
    lib = _lib.load()
    P, d = keep["y0"].shape
    y0_bar = torch.empty_like(keep["y0"])
    bars = [torch.empty(d, dtype=torch.float64, device=y0_bar.device) for _ in range(3)]
    g = _sde.SdeGradIO(ys_bar=keep["ys_bar"].data_ptr(), y0_bar=y0_bar.data_ptr(),
                       theta_bar=bars[0].data_ptr(), mu_bar=bars[1].data_ptr(),
                       sigma_bar=bars[2].data_ptr())
    if static:
        rc = lib.dfx_sde_heun_adjoint_static(ctypes.byref(prob), ctypes.byref(io),
                                             ctypes.byref(g), _stream())
    else:
        nbytes = ctypes.c_size_t(0)
        assert lib.dfx_sde_workspace_bytes(ctypes.byref(prob), ctypes.byref(nbytes)) == 0
        ws = torch.empty(nbytes.value, dtype=torch.uint8, device=y0_bar.device)
        rc = lib.dfx_sde_heun_adjoint(ctypes.byref(prob), ctypes.byref(io), ctypes.byref(g),
                                      ctypes.c_void_p(ws.data_ptr()),
                                      ctypes.c_size_t(ws.numel()), _stream())
    assert rc == 0
    torch.cuda.synchronize()
    return [x.cpu() for x in [y0_bar] + bars]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_static_scratch_adjoint_equals_workspace_adjoint(cuda, dtype):
    # the second problem is smaller: its partials must not mix with the first call's
    for P, d, N, seed in ((8192, 16, 256, 0), (1000, 8, 100, 1)):
        prob, io, keep = _forward(P, d, N, seed, dtype)
        got = _adjoint(prob, io, keep, static=True)
        want = _adjoint(prob, io, keep, static=False)
        for name, a, b in zip(("y0", "theta", "mu", "sigma"), got, want):
            assert torch.equal(a, b), (P, name)
        assert all(torch.isfinite(x).all() for x in got)


def test_static_scratch_size_limit(cuda):
    lib = _lib.load()
    big = _sde.SdeProblem(precision=_lib.F32, d=16, paths=1 << 20, num_steps=64,
                          checkpoint_every=32, save_t0=0, save_t1=1)
    nbytes = ctypes.c_size_t(0)
    assert lib.dfx_sde_workspace_bytes(ctypes.byref(big), ctypes.byref(nbytes)) == 0
    assert nbytes.value > _sde.SDE_STATIC_SCRATCH_BYTES
    z = ctypes.c_void_p(8)    # never dereferenced: the size check comes first
    io = _sde.SdeIO(y0=z, theta=z, mu=z, sigma=z, dW=z, grid=z, ys=z, ckpt=z)
    g = _sde.SdeGradIO(ys_bar=z, y0_bar=z, theta_bar=z, mu_bar=z, sigma_bar=z)
    assert lib.dfx_sde_heun_adjoint_static(ctypes.byref(big), ctypes.byref(io), ctypes.byref(g),
                                           _stream()) == -4
