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
