"""C ABI and host-side logic (CPU only: no kernel is launched).

- every function declared in include/diffrax_amd.h is exported by libdiffrax_amd.so and
  bound in the Python loader;
- the product's device tableaux (read back through dfx_tableau) are bit-identical to the
  oracle's, so the order-condition pins of the oracle cover them too;
- argument errors are reported as the reference reports them (RuntimeError for the
  error_if checks of _integrate.py:1036-1045, 1222-1235) before any GPU work."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

import diffrax_amd as dfx
import oracle
from diffrax_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_0-9]+\s*\*?\s*(dfx_[a-z0-9_]+)\s*\(",
                            text, flags=re.M)
    return sorted(set(names))


def test_header_symbols_exported_and_bound():
    names = _declared_functions()
    assert len(names) >= 10
    lib = _lib.load()
    for n in names:
        assert hasattr(lib, n), f"{n} not exported by {_lib.LIB_PATH}"
        assert n in _lib._SIGS, f"{n} has no ctypes signature in diffrax_amd._lib"


@pytest.mark.parametrize("solver,sid", [("tsit5", 0), ("dopri5", 1), ("heun", 2)])
def test_device_tableau_matches_oracle(solver, sid):
    lib = _lib.load()
    a = (ctypes.c_double * 49)(); b = (ctypes.c_double * 7)(); be = (ctypes.c_double * 7)()
    c = (ctypes.c_double * 7)(); info = (ctypes.c_int32 * 4)()
    assert lib.dfx_tableau(sid, a, b, be, c, info) == 0
    oa, ob, obe, oc, oinfo = oracle.tableau(solver)
    assert np.array_equal(np.array(a[:]).reshape(7, 7), oa)
    assert np.array_equal(np.array(b[:]), ob) and np.array_equal(np.array(be[:]), obe)
    assert np.array_equal(np.array(c[:]), oc) and tuple(info) == oinfo
    t = getattr(dfx, {"tsit5": "Tsit5", "dopri5": "Dopri5", "heun": "Heun"}[solver])().tableau
    assert t.num_stages == oinfo[0] and t.fsal == bool(oinfo[1])


def test_n_out_and_param_counts():
    lib = _lib.load()
    p = _lib.ErkProblem(n_ts=64, save_t0=1, save_t1=1, save_steps=0, max_steps=100)
    assert lib.dfx_erk_n_out(ctypes.byref(p)) == 66
    p.save_steps = 1
    assert lib.dfx_erk_n_out(ctypes.byref(p)) == 165
    assert lib.dfx_vf_param_count(_lib.VF_MLP, 64, 256) == 98880
    assert lib.dfx_vf_param_count(_lib.VF_OU, 16, 0) == 48


def test_solve_rejects_bad_arguments_without_launching():
    lib = _lib.load()
    p = _lib.ErkProblem(solver=7, controller=0, vf=_lib.VF_LV, precision=0, d=2, batch=4,
                        max_steps=10)
    io = _lib.ErkIO()
    rc = lib.dfx_erk_solve(ctypes.byref(p), ctypes.byref(io), None, 0, None, None)
    assert rc == -1
    p.solver = 0
    p.d = 3   # LV needs d == 2
    assert lib.dfx_erk_solve(ctypes.byref(p), ctypes.byref(io), None, 0, None, None) == -1
    assert lib.dfx_strerror(-3).decode().startswith("workspace")


def test_resume_rejects_bad_arguments_without_launching():
    """dfx_erk_problem.resume: unknown bits, a controller state with step clipping, a
    missing state_in, checkpoints, and the adjoint entries are rejected before any launch."""
    lib = _lib.load()
    z = ctypes.c_void_p(8)    # never dereferenced: every case fails validation first
    ws = ctypes.c_void_p(256)

    def prob(**kw):
        base = dict(solver=_lib.TSIT5, controller=_lib.PID, vf=_lib.VF_LV, precision=_lib.F64,
                    d=2, batch=4, max_steps=10, save_t1=1, dt0=0.1, rtol=1e-6, atol=1e-8,
                    pcoeff=0.0, icoeff=1.0, dcoeff=0.0, dtmin=float("nan"),
                    dtmax=float("nan"), error_order=float("nan"), factormin=0.2,
                    factormax=10.0, safety=0.9)
        base.update(kw)
        return _lib.ErkProblem(**base)

    io = _lib.ErkIO(y0=z, t1=z, params=z, ys=z, ts_out=z, stats=z, result=z, step_ts=z,
                    state_in=z, ckpt=z)
    call = lambda p, i=io: lib.dfx_erk_solve(ctypes.byref(p), ctypes.byref(i), ws, 256, None,
                                             None)
    assert call(prob(resume=8)) == -1
    assert call(prob(resume=_lib.RESUME_CONTROLLER, n_step_ts=1)) == -1
    assert call(prob(resume=_lib.RESUME_SOLVER, ckpt_every=4)) == -1
    no_state = _lib.ErkIO(y0=z, t1=z, params=z, ys=z, ts_out=z, stats=z, result=z)
    assert call(prob(resume=_lib.RESUME_MADE_JUMP), no_state) == -1
    bio = _lib.BacksolveIO(params=z, ys=z, ts_out=z, ys_bar=z, y0_bar=z, params_bar=z, result=z)
    assert lib.dfx_erk_backsolve(ctypes.byref(prob(resume=1)), ctypes.byref(bio), None) == -1
    aio = _lib.AdjointIO(ckpt=z, stats=z, ys_bar=z, y0_bar=z, params_bar=z)
    assert lib.dfx_erk_adjoint(ctypes.byref(prob(resume=1, ckpt_every=4)), ctypes.byref(io),
                               ctypes.byref(aio), ws, 1 << 20, None) == -1


def test_resume_host_validation():
    """Malformed states and unsupported combinations fail on the host (no GPU needed)."""
    with pytest.raises(NotImplementedError, match="PIDController and ConstantStepSize"):
        _solve(stepsize_controller=dfx.ClipStepSizeController(
            dfx.PIDController(rtol=1e-6, atol=1e-8), step_ts=[0.5]),
            saveat=dfx.SaveAt(t1=True, controller_state=True))
    with pytest.raises(NotImplementedError, match="ODE kernels only"):
        dfx.diffeqsolve(dfx.MultiTerm(dfx.ODETerm(dfx.OUDrift()), dfx.ControlTerm(
            dfx.DiagonalNoise(), dfx.IncrementPath(torch.zeros(1, 4, 2)))), dfx.Heun(), 0.0,
            1.0, 0.25, torch.zeros(1, 2), (torch.ones(2), torch.zeros(2), torch.ones(2)),
            saveat=dfx.SaveAt(t1=True), made_jump=True)
    with pytest.raises(ValueError, match="first_step, f0"):
        _solve(solver_state=torch.zeros(2))
    with pytest.raises(ValueError, match="prev_inv"):
        _solve(controller_state=(1.0, 1.0))
    with pytest.raises(ValueError, match="steps_completed"):
        _solve(controller_state=(1.0, 1.0, False), stepsize_controller=dfx.ConstantStepSize())


def _solve(**kw):
    base = dict(terms=dfx.ODETerm(dfx.LotkaVolterra()), solver=dfx.Tsit5(), t0=0.0, t1=1.0,
                dt0=0.1, y0=torch.ones(2, dtype=torch.float64), args=(1.5, 1.0, 3.0, 1.0),
                saveat=dfx.SaveAt(t1=True),
                stepsize_controller=dfx.PIDController(rtol=1e-6, atol=1e-8))
    base.update(kw)
    return dfx.diffeqsolve(**base)


def test_host_validation_errors():
    with pytest.raises(RuntimeError, match="dt0"):
        _solve(dt0=-0.1)
    with pytest.raises(RuntimeError, match="between t0 and t1"):
        _solve(saveat=dfx.SaveAt(ts=[0.5, 3.0]))
    with pytest.raises(RuntimeError, match="increasing"):
        _solve(saveat=dfx.SaveAt(ts=[0.5, 0.2]))
    with pytest.raises(NotImplementedError, match="registered"):
        _solve(terms=dfx.ODETerm(lambda t, y, args: -y))
    with pytest.raises(ValueError):
        dfx.PIDController(rtol=None, atol=1e-6)
    with pytest.raises(ValueError):
        dfx.SaveAt(ts=[0.1], subs=dfx.SubSaveAt(t1=True))
    assert dfx.RESULTS.successful._value == 0
    assert isinstance(dfx.RESULTS[dfx.RESULTS.max_steps_reached], str)


def test_backsolve_rejects_bad_arguments_without_launching():
    lib = _lib.load()
    z = ctypes.c_void_p(8)    # never dereferenced: every case fails validation first
    io = _lib.BacksolveIO(params=z, ys=z, ts_out=z, ys_bar=z, y0_bar=z, params_bar=z, result=z)

    def prob(**kw):
        base = dict(solver=_lib.TSIT5, controller=_lib.PID, vf=_lib.VF_LV, precision=_lib.F64,
                    d=2, batch=4, max_steps=10, save_t1=1, dt0=0.1, rtol=1e-6, atol=1e-8,
                    pcoeff=0.0, icoeff=1.0, dcoeff=0.0, dtmin=float("nan"),
                    dtmax=float("nan"), error_order=float("nan"), factormin=0.2,
                    factormax=10.0, safety=0.9)
        base.update(kw)
        return _lib.ErkProblem(**base)

    call = lambda p, i=io: lib.dfx_erk_backsolve(ctypes.byref(p), ctypes.byref(i), None)
    for bad in (dict(pcoeff=0.3), dict(dcoeff=0.1), dict(dtmin=1e-3), dict(save_steps=1),
                dict(vf=_lib.VF_LINEAR, d=4), dict(vf=_lib.VF_MLP, d=64), dict(n_events=1),
                dict(n_step_ts=1), dict(controller=_lib.CONSTANT, dt0=float("nan")),
                dict(precision=_lib.BF16_F32), dict(solver=9)):
        assert call(prob(**bad)) == -1, bad
    assert call(prob(), _lib.BacksolveIO()) == -1          # missing buffers
    assert call(prob(batch=0), _lib.BacksolveIO()) == 0    # empty batch: nothing to do


def test_adjoint_rejects_bad_arguments_without_launching():
    lib = _lib.load()
    z = ctypes.c_void_p(8)    # never dereferenced: every case fails validation first
    fwd = _lib.ErkIO(y0=z, t1=z, ts=z, params=z, ys=z, ts_out=z, stats=z, result=z)
    io = _lib.AdjointIO(ckpt=z, stats=z, ys_bar=z, y0_bar=z, params_bar=z)

    def prob(**kw):
        base = dict(solver=_lib.TSIT5, controller=_lib.PID, vf=_lib.VF_LV, precision=_lib.F64,
                    d=2, batch=4, max_steps=10, save_t1=1, ckpt_every=4)
        base.update(kw)
        return _lib.ErkProblem(**base)

    def call(p, i=io, f=fwd, ws=z, nbytes=1 << 20):
        return lib.dfx_erk_adjoint(ctypes.byref(p), ctypes.byref(f), ctypes.byref(i), ws,
                                   ctypes.c_size_t(nbytes), None)

    for bad in (dict(vf=_lib.VF_LINEAR, d=10), dict(vf=_lib.VF_MLP, d=64), dict(vf=_lib.VF_VDP, d=3),
                dict(n_events=1), dict(precision=_lib.BF16_F32), dict(solver=9), dict(batch=-1),
                dict(ckpt_every=0), dict(ckpt_every=-2), dict(n_ts=3, ts_stride=0)):
        f = _lib.ErkIO(t1=z, params=z) if "n_ts" in bad else fwd   # n_ts > 0 without ts
        assert call(prob(**bad), f=f) == -1, bad
    assert call(prob(), _lib.AdjointIO(stats=z, ys_bar=z, y0_bar=z, params_bar=z)) == -1  # no ckpt
    assert call(prob(), f=_lib.ErkIO(t1=z)) == -1             # forward inputs missing
    nbytes = ctypes.c_size_t(0)
    assert lib.dfx_erk_adjoint_workspace_bytes(ctypes.byref(prob()), ctypes.byref(nbytes)) == 0
    assert nbytes.value == 512 + 512       # [4, 2 * 4, 2] + [4, 2 * 4, 2] float64, 256-aligned
    assert call(prob(), nbytes=nbytes.value - 1) == -3      # workspace too small
    assert call(prob(), ws=None) == -3
    assert call(prob(batch=0), _lib.AdjointIO(), _lib.ErkIO()) == 0   # empty batch
    n_ck, words = ctypes.c_int32(0), ctypes.c_int32(0)
    assert lib.dfx_erk_ckpt_layout(ctypes.byref(prob(max_steps=10, ckpt_every=4)),
                                   ctypes.byref(n_ck), ctypes.byref(words)) == 0
    assert (n_ck.value, words.value) == (3, 10)             # ceil(10 / 4) records of d + 8
    assert lib.dfx_erk_ckpt_layout(ctypes.byref(prob(ckpt_every=0)), ctypes.byref(n_ck),
                                   ctypes.byref(words)) == -1


def test_solve_needs_checkpoint_buffer():
    """ckpt_every > 0 needs dfx_erk_io.ckpt (rejected before any launch)."""
    lib = _lib.load()
    z = ctypes.c_void_p(8)
    p = _lib.ErkProblem(solver=_lib.TSIT5, controller=_lib.PID, vf=_lib.VF_LV,
                        precision=_lib.F64, d=2, batch=4, max_steps=10, save_t1=1,
                        ckpt_every=4)
    io = _lib.ErkIO(y0=z, t1=z, ts=z, params=z, ys=z, ts_out=z, stats=z, result=z)
    assert lib.dfx_erk_solve(ctypes.byref(p), ctypes.byref(io), z, ctypes.c_size_t(256), None,
                             None) == -1


def test_backsolve_host_guards():
    """BacksolveAdjoint's guards (_adjoint.py:781-858) raise before any GPU work."""
    y0 = torch.ones(2, dtype=torch.float64, requires_grad=True)
    with pytest.raises(ValueError, match="not valid"):
        dfx.BacksolveAdjoint(rtol=1e-3)
    for saveat in (dfx.SaveAt(steps=True), dfx.SaveAt(dense=True),
                   dfx.SaveAt(subs=[dfx.SubSaveAt(t1=True), dfx.SubSaveAt(t0=True)])):
        with pytest.raises(NotImplementedError):
            _solve(y0=y0, saveat=saveat, adjoint=dfx.BacksolveAdjoint())
    with pytest.raises(NotImplementedError, match="events"):
        _solve(y0=y0, adjoint=dfx.BacksolveAdjoint(),
               event=dfx.Event(dfx.AffineCondition(y=[1.0, 0.0], offset=-5.0)))
    with pytest.raises(NotImplementedError, match="integral"):
        _solve(y0=y0, adjoint=dfx.BacksolveAdjoint(),
               stepsize_controller=dfx.PIDController(rtol=1e-6, atol=1e-8, pcoeff=0.3,
                                                     icoeff=0.3))
    with pytest.raises(NotImplementedError, match="integral"):
        _solve(y0=y0, adjoint=dfx.BacksolveAdjoint(
            stepsize_controller=dfx.PIDController(rtol=1e-6, atol=1e-8, dtmin=1e-4)))


def test_no_cpu_fallback():
    """Without a GPU a well-formed solve raises instead of computing on the CPU."""
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _solve()


def test_registered_fields_evaluate_like_formula():
    y = torch.tensor([[1.2, 0.7]], dtype=torch.float64)
    f = dfx.LotkaVolterra()(0.0, y, (1.5, 1.0, 3.0, 1.0))
    assert torch.allclose(f, torch.tensor([[1.5 * 1.2 - 1.2 * 0.7, 1.2 * 0.7 - 3 * 0.7]],
                                          dtype=torch.float64))
    g = dfx.VanDerPol()(0.0, y, 2.0)
    assert torch.allclose(g, torch.tensor([[0.7, 2.0 * (1 - 1.44) * 0.7 - 1.2]], dtype=torch.float64))


def test_sde_sizes_and_argument_checks():
    from diffrax_amd import _sde
    lib = _lib.load()
    p = _sde.SdeProblem(precision=_lib.F32, d=16, paths=1000, num_steps=1000,
                        checkpoint_every=32, save_t0=0, save_t1=1)
    assert lib.dfx_sde_n_ckpt(ctypes.byref(p)) == 31          # ceil(1000 / 32) - 1
    nb = ctypes.c_size_t(0)
    assert lib.dfx_sde_workspace_bytes(ctypes.byref(p), ctypes.byref(nb)) == 0
    assert nb.value == ((1000 * 16 + 255) // 256) * 3 * 16 * 8 + 256
    io, g = _sde.SdeIO(), _sde.SdeGradIO()
    # missing device pointers, a dimension the float32 kernel cannot split (d % 4),
    # an unsupported precision and an adjoint segment length outside {8, 16, 32}
    assert lib.dfx_sde_heun_solve(ctypes.byref(p), ctypes.byref(io), None) == -1
    for field, bad in (("d", 6), ("precision", 7)):
        q = _sde.SdeProblem(precision=p.precision, d=p.d, paths=p.paths, num_steps=p.num_steps,
                            checkpoint_every=p.checkpoint_every, save_t0=0, save_t1=1)
        setattr(q, field, bad)
        assert lib.dfx_sde_heun_solve(ctypes.byref(q), ctypes.byref(io), None) == -1
    # dense output: both buffers, forward only (no checkpoints)
    z = ctypes.c_void_p(8)
    full = dict(y0=z, theta=z, mu=z, sigma=z, dW=z, grid=z, ys=z)
    q = _sde.SdeProblem(precision=_lib.F32, d=16, paths=1000, num_steps=1000, save_t1=1)
    assert lib.dfx_sde_heun_solve(ctypes.byref(q), ctypes.byref(_sde.SdeIO(dense_ys=z, **full)),
                                  None) == -1
    assert lib.dfx_sde_heun_solve(ctypes.byref(p), ctypes.byref(
        _sde.SdeIO(dense_ys=z, dense_k=z, ckpt=z, **full)), None) == -1
    dio = _sde.SdeDenseIO(dense_ys=z, dense_k=z, grid=z, tq=z, n_q=4, tq_stride=2, out=z)
    assert lib.dfx_sde_dense_eval(ctypes.byref(q), ctypes.byref(dio), None) == -1
    p.checkpoint_every = 12
    assert lib.dfx_sde_heun_adjoint(ctypes.byref(p), ctypes.byref(io), ctypes.byref(g), None, 0,
                                    None) == -1


def test_mlp_sizes_and_argument_checks():
    from diffrax_amd import _mlp
    lib = _lib.load()
    nb = ctypes.c_size_t(0)
    assert lib.dfx_mlp_pack_bytes(_lib.F32, ctypes.byref(nb)) == 0
    assert nb.value == 2 * (256 * 64 + 256 * 256 + 64 * 256) * 4   # forward + transposed
    assert lib.dfx_mlp_pack_bytes(_lib.BF16_F32, ctypes.byref(nb)) == 0
    assert nb.value == 3 * 2 * (256 * 64 + 256 * 256 + 64 * 256) * 2   # three bf16 pieces
    assert lib.dfx_mlp_pack_bytes(_lib.F64, ctypes.byref(nb)) == -1
    p = _mlp.MlpProblem(solver=_lib.TSIT5, precision=_lib.F32, d=64, hidden=256, n_ts=32,
                        max_steps=4096, ckpt_every=64, batch=8192)
    assert lib.dfx_mlp_n_ckpt(ctypes.byref(p)) == 64
    io = _mlp.MlpIO()
    z = ctypes.c_void_p(0)
    for field, bad in (("d", 32), ("hidden", 128), ("precision", _lib.F64), ("solver", _lib.DOPRI5),
                       ("waves", 2), ("waves", 16)):
        q = _mlp.MlpProblem(solver=_lib.TSIT5, precision=_lib.F32, d=64, hidden=256, n_ts=32,
                            max_steps=4096, ckpt_every=64, batch=8192)
        setattr(q, field, bad)
        assert lib.dfx_mlp_solve(ctypes.byref(q), z, z, z, z, ctypes.byref(io), None) == -1
    # saved rows: t0 + ts + every k-th step + t1 (no t1 row when k divides max_steps)
    q = _mlp.MlpProblem(d=64, hidden=256, n_ts=5, save_t0=1, save_t1=1, max_steps=100)
    assert lib.dfx_mlp_n_out(ctypes.byref(q)) == 7
    q.save_steps = 3
    assert lib.dfx_mlp_n_out(ctypes.byref(q)) == 7 + 33
    q.save_steps = 4
    assert lib.dfx_mlp_n_out(ctypes.byref(q)) == 6 + 25
    # dense buffers all or none; step saves and dense output are forward only
    zz = ctypes.c_void_p(8)
    base = dict(solver=_lib.TSIT5, precision=_lib.F32, d=64, hidden=256, n_ts=32,
                max_steps=4096, batch=8192)
    part = _mlp.MlpIO(y0=zz, ts=zz, ys=zz, ts_out=zz, stats=zz, result=zz, dense_ts=zz)
    assert lib.dfx_mlp_solve(ctypes.byref(_mlp.MlpProblem(**base)), zz, zz, zz, zz,
                             ctypes.byref(part), None) == -1
    full = _mlp.MlpIO(y0=zz, ts=zz, ys=zz, ts_out=zz, stats=zz, result=zz, grid=zz, ckpt=zz,
                      dense_ts=zz, dense_ys=zz, dense_k=zz)
    assert lib.dfx_mlp_solve(ctypes.byref(_mlp.MlpProblem(ckpt_every=64, **base)), zz, zz, zz,
                             zz, ctypes.byref(full), None) == -1
    steps_io = _mlp.MlpIO(y0=zz, ts=zz, ys=zz, ts_out=zz, stats=zz, result=zz, grid=zz, ckpt=zz)
    assert lib.dfx_mlp_solve(ctypes.byref(_mlp.MlpProblem(ckpt_every=64, save_steps=1, **base)),
                             zz, zz, zz, zz, ctypes.byref(steps_io), None) == -1
    # events: at most 4 conditions, all three event buffers given, forward only
    ev_io = _mlp.MlpIO(y0=zz, ts=zz, ys=zz, ts_out=zz, stats=zz, result=zz, event_coeffs=zz,
                       event_flags=zz, event_mask=zz)
    no_mask = _mlp.MlpIO(y0=zz, ts=zz, ys=zz, ts_out=zz, stats=zz, result=zz, event_coeffs=zz,
                         event_flags=zz)
    for prob, io_ in ((_mlp.MlpProblem(n_events=5, **base), ev_io),
                      (_mlp.MlpProblem(n_events=-1, **base), ev_io),
                      (_mlp.MlpProblem(n_events=1, **base), no_mask),
                      (_mlp.MlpProblem(n_events=1, ckpt_every=64, **base),
                       _mlp.MlpIO(y0=zz, ts=zz, ys=zz, ts_out=zz, stats=zz, result=zz, grid=zz,
                                  ckpt=zz, event_coeffs=zz, event_flags=zz, event_mask=zz))):
        assert lib.dfx_mlp_solve(ctypes.byref(prob), zz, zz, zz, zz, ctypes.byref(io_),
                                 None) == -1
    # step clipping (counts not negative, times given) and resume
    clip_io = _mlp.MlpIO(y0=zz, ts=zz, ys=zz, ts_out=zz, stats=zz, result=zz, step_ts=zz,
                         jump_ts=zz)
    for prob, io_ in ((_mlp.MlpProblem(n_step_ts=-1, **base), clip_io),
                      (_mlp.MlpProblem(n_jump_ts=2, **base), no_mask),
                      # resume: known bits, state_in given, not with checkpoints or clipping
                      (_mlp.MlpProblem(resume=8, **base), _mlp.MlpIO(
                          y0=zz, ts=zz, ys=zz, ts_out=zz, stats=zz, result=zz, state_in=zz)),
                      (_mlp.MlpProblem(resume=2, **base), no_mask),
                      (_mlp.MlpProblem(resume=2, n_step_ts=1, **base), _mlp.MlpIO(
                          y0=zz, ts=zz, ys=zz, ts_out=zz, stats=zz, result=zz, state_in=zz,
                          step_ts=zz)),
                      (_mlp.MlpProblem(resume=1, ckpt_every=64, **base), _mlp.MlpIO(
                          y0=zz, ts=zz, ys=zz, ts_out=zz, stats=zz, result=zz, grid=zz,
                          ckpt=zz, state_in=zz)),
                      (_mlp.MlpProblem(ckpt_every=64, **base), _mlp.MlpIO(
                          y0=zz, ts=zz, ys=zz, ts_out=zz, stats=zz, result=zz, grid=zz,
                          ckpt=zz, state_out=zz))):
        assert lib.dfx_mlp_solve(ctypes.byref(prob), zz, zz, zz, zz, ctypes.byref(io_),
                                 None) == -1
    dio = _lib.DenseIO(dense_ts=zz, dense_ys=zz, dense_k=zz, stats=zz, tq=zz, n_q=4,
                       tq_stride=2, out=zz)
    assert lib.dfx_mlp_dense_eval(ctypes.byref(_mlp.MlpProblem(**base)), ctypes.byref(dio),
                                  None) == -1       # stride shorter than the query count
    assert lib.dfx_mlp_dense_eval(ctypes.byref(_mlp.MlpProblem(**dict(base, d=32))),
                                  ctypes.byref(dio), None) == -1
    adj = _mlp.MlpAdjIO()
    # segment past the last checkpoint, segment longer than the checkpoint spacing
    for seg, steps in ((64, 64), (0, 65)):
        assert lib.dfx_mlp_adj_segment(ctypes.byref(p), z, z, z, z, ctypes.byref(adj), seg, steps,
                                       None) == -1
    zs, rows, part = (ctypes.c_size_t(0) for _ in range(3))
    assert lib.dfx_mlp_adj_workspace(ctypes.byref(p), 4, ctypes.byref(zs), ctypes.byref(rows),
                                     ctypes.byref(part)) == 0
    assert zs.value == 8192 * 4 * 7 * 64 * 4       # f of every stage of the segment
    chunks = -(-(4 * 7 * 8192) // 4096)
    assert rows.value == chunks * 4096 * (256 + 64 + 256 + 256 + 64 + 256) * 4


def test_mlp_backsolve_sizes_and_argument_checks():
    from diffrax_amd import _mlp
    lib = _lib.load()
    p = _mlp.MlpProblem(solver=_lib.TSIT5, precision=_lib.F32, d=64, hidden=256, n_ts=32,
                        max_steps=4096, batch=8192, dt0=-0.1, rtol=1e-6, atol=1e-8, icoeff=1.0,
                        dtmin=float("nan"), dtmax=float("nan"), factormin=0.2, factormax=10.0,
                        safety=0.9, error_order=float("nan"))
    st, rows, part = (ctypes.c_size_t(0) for _ in range(3))
    assert lib.dfx_mlp_backsolve_workspace(ctypes.byref(p), 8, ctypes.byref(st),
                                           ctypes.byref(rows), ctypes.byref(part)) == 0
    nrow = -(-(8 * 7 * 8192) // 4096) * 4096
    assert rows.value == nrow * (256 + 64 + 256 + 256 + 64 + 256) * 4
    # y and a_y per row, the row state, a weight per row-buffer row, two counters
    assert st.value >= 8192 * 128 * 4 + 8192 * 64 + nrow * 4 + 8
    assert lib.dfx_mlp_backsolve_workspace(ctypes.byref(p), 0, ctypes.byref(st),
                                           ctypes.byref(rows), ctypes.byref(part)) == -1
    io = _mlp.MlpBacksolveIO()
    z = ctypes.c_void_p(0)
    assert lib.dfx_mlp_backsolve(ctypes.byref(p), z, z, z, z, ctypes.byref(io), 8, 1, None) == -1
    q = _mlp.MlpProblem()
    ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(p))
    q.dt0 = float("nan")        # the backward solves need dt0
    assert lib.dfx_mlp_backsolve(ctypes.byref(q), z, z, z, z, ctypes.byref(io), 8, 1, None) == -1
    q.dt0, q.precision = -0.1, _lib.BF16_F32   # float32 products on a float32 packing only
    assert lib.dfx_mlp_backsolve(ctypes.byref(q), z, z, z, z, ctypes.byref(io), 8, 1, None) == -1
    cp = ctypes.c_void_p(0)
    assert lib.dfx_mlp_backsolve_status(ctypes.byref(p), 8, ctypes.c_void_p(4096),
                                        ctypes.byref(cp)) == 0
    assert cp.value > 4096 and (cp.value - 4096) % 256 == 0
