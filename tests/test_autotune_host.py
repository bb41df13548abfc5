"""The launch-shape cache key separates plans that run different kernels or move different
amounts of memory (CPU only: the device name is stubbed)."""
import types

import torch

from diffrax_amd import _autotune, _lib


def _plan(**kw):
    p = _lib.ErkProblem(solver=0, controller=0, vf=1, precision=0, d=2, batch=65536,
                        icoeff=1.0, error_order=5.0)
    io = _lib.ErkIO()
    for k, v in kw.items():
        if k.startswith("io_"):
            setattr(io, k[3:], v)
        else:
            setattr(p if hasattr(p, k) and k not in ("dense_ts", "grid") else io, k, v)
    return types.SimpleNamespace(problem=p, io=io)


def test_key_separates_kernel_variants(monkeypatch):
    monkeypatch.setattr(torch.cuda, "get_device_name", lambda *a: "dev")
    monkeypatch.setattr(torch.cuda, "current_device", lambda: 0)
    base = _autotune._key(_plan())
    variants = [_plan(n_events=1), _plan(n_step_ts=3), _plan(dense_ts=4096),
                _plan(grid=4096), _plan(pcoeff=0.3), _plan(batch=1024), _plan(precision=1),
                _plan(n_ts=64), _plan(ts_stride=64), _plan(save_t1=1), _plan(save_steps=1),
                _plan(max_steps=4096), _plan(rtol=1e-8), _plan(atol=1e-12),
                _plan(io_rtol=4096), _plan(io_atol=4096), _plan(ckpt_every=128)]
    keys = [_autotune._key(v) for v in variants]
    assert base not in keys and len(set(keys)) == len(keys)
    assert _autotune._key(_plan()) == base


def _mlp_plan(**kw):
    from diffrax_amd import _mlp
    p = _mlp.MlpProblem(solver=_lib.TSIT5, precision=_lib.F32, d=64, hidden=256, n_ts=32,
                        max_steps=4096, batch=8192, t0=0.0, t1=1.0, dt0=0.1, rtol=1e-5,
                        atol=1e-7, icoeff=1.0, error_order=float("nan"))
    for k, v in kw.items():
        setattr(p, k, v)
    return types.SimpleNamespace(prob=p)


def test_mlp_key_separates_kernel_variants(monkeypatch):
    """The neural-ODE tuning key: precision, controller, dt0 selection, saves, gradient
    recording, span and batch each select another entry; the chosen wave count does not."""
    monkeypatch.setattr(torch.cuda, "get_device_name", lambda *a: "dev")
    monkeypatch.setattr(torch.cuda, "current_device", lambda: 0)
    base = _autotune._key_mlp(_mlp_plan())
    variants = [_mlp_plan(precision=_lib.BF16_F32), _mlp_plan(pcoeff=0.3),
                _mlp_plan(dt0=float("nan")), _mlp_plan(n_ts=8), _mlp_plan(save_t1=1),
                _mlp_plan(ckpt_every=64), _mlp_plan(t1=2.0), _mlp_plan(batch=1024),
                _mlp_plan(rtol=1e-6), _mlp_plan(max_steps=100), _mlp_plan(ts_stride=32),
                _mlp_plan(save_steps=2)]
    keys = [_autotune._key_mlp(v) for v in variants]
    assert base not in keys and len(set(keys)) == len(keys)
    assert _autotune._key_mlp(_mlp_plan(waves=8)) == base
