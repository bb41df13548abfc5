"""GPU parity of the small-system ERK kernel across solvers, vector fields and options.

Every case runs the same seeded inputs through the GPU kernel and the FP64 C oracle.
Float64 bar: identical (num_steps, accepted, rejected) and result codes per instance, and
saved states / times bit-identical (non-default PID exponents use the math library's pow on
both sides, so those cases are held to 1e-12 relative instead)."""
import numpy as np
import pytest
import torch

import diffrax_amd as dfx
import oracle

pytestmark = pytest.mark.gpu

SOLVERS = {"tsit5": dfx.Tsit5(), "dopri5": dfx.Dopri5(), "heun": dfx.Heun()}


def _vf(name):
    return {"lv": dfx.LotkaVolterra(), "vdp": dfx.VanDerPol(), "linear": dfx.LinearField()}[name]


def _run(solver, vf, y0, params, t0, t1, dt0, ts=None, *, controller="pid", rtol=1e-6,
         atol=1e-8, save_t0=False, save_t1=False, save_steps=False, max_steps=4096,
         dtype=torch.float64, launch=None, **pid):
    y0 = np.asarray(y0, dtype=np.float64)
    B, d = y0.shape
    # device call
    if vf == "lv":
        args = tuple(params) if np.ndim(params) == 1 else tuple(torch.tensor(params[:, i]) for i in range(4))
    elif vf == "vdp":
        args = float(params[0]) if np.ndim(params) == 1 else torch.tensor(params[:, 0])
    else:
        args = torch.tensor(np.asarray(params).reshape((d, d) if np.ndim(params) == 1 else (B, d, d)))
    ctrl = (dfx.PIDController(rtol=rtol, atol=atol, **pid) if controller == "pid"
            else dfx.ConstantStepSize())
    saveat = dfx.SaveAt(t0=save_t0, t1=save_t1, ts=ts, steps=save_steps)
    t1_arg = torch.tensor(t1) if np.ndim(t1) else t1
    sol = dfx.diffeqsolve(dfx.ODETerm(_vf(vf)), SOLVERS[solver], t0, t1_arg, dt0,
                          torch.tensor(y0, dtype=dtype), args, saveat=saveat,
                          stepsize_controller=ctrl, max_steps=max_steps, throw=False,
                          launch=launch)
    torch.cuda.synchronize()
    got = dict(ys=sol.ys.double().cpu().numpy(), ts=sol.ts.cpu().numpy(),
               stats=np.stack([sol.stats[k].cpu().numpy() for k in
                               ("num_steps", "num_accepted_steps", "num_rejected_steps")], -1),
               result=sol.result.cpu().numpy())
    okw = dict(controller=controller, rtol=rtol, atol=atol, save_t0=save_t0, save_t1=save_t1,
               save_steps=save_steps, max_steps=max_steps)
    okw.update({k: v for k, v in pid.items()})
    ref = oracle.solve(y0, t1, params, ts=ts, solver=solver, vf=vf, t0=t0, dt0=dt0, **okw)
    return got, ref


def _assert_exact(got, ref, tol=0.0):
    assert np.array_equal(got["stats"], ref["stats"]), (got["stats"][:4], ref["stats"][:4])
    assert np.array_equal(got["result"], ref["result"])
    if tol == 0.0:
        assert np.array_equal(got["ts"], ref["ts"])
        same = (got["ys"] == ref["ys"]) | (np.isnan(got["ys"]) & np.isnan(ref["ys"]))
        assert same.all(), oracle.max_rel_err(got["ys"], ref["ys"], 1e-12)
    else:
        assert oracle.max_rel_err(got["ys"], ref["ys"], 1e-8) <= tol


RNG = np.random.default_rng(7)
LV_Y0 = RNG.uniform(0.5, 2.0, size=(256, 2))
LV_Q = np.array([1.5, 1.0, 3.0, 1.0])
VDP_Y0 = RNG.uniform(-2.0, 2.0, size=(256, 2))
VDP_MU = np.exp(RNG.uniform(np.log(0.1), np.log(5.0), size=(256, 1)))


@pytest.mark.parametrize("solver", ["tsit5", "dopri5", "heun"])
@pytest.mark.parametrize("vf", ["lv", "vdp"])
def test_solvers_vector_fields(cuda, solver, vf):
    y0, q = (LV_Y0, LV_Q) if vf == "lv" else (VDP_Y0, VDP_MU)
    ts = np.linspace(0.0, 8.0, 33)
    got, ref = _run(solver, vf, y0, q, 0.0, 8.0, 0.05, ts, max_steps=20000)
    _assert_exact(got, ref)


@pytest.mark.parametrize("d", [1, 3, 10])
def test_linear_fields_and_auto_dt0(cuda, d):
    rng = np.random.default_rng(d)
    A = rng.normal(size=(d, d)) * 0.5
    y0 = rng.normal(size=(64, d))
    ts = np.linspace(0.0, 3.0, 7)
    got, ref = _run("tsit5", "linear", y0, A.ravel(), 0.0, 3.0, None, ts, rtol=1e-8, atol=1e-8)
    _assert_exact(got, ref)


def test_reverse_time_and_t0_t1_saves(cuda):
    ts = np.linspace(5.0, 0.5, 10)
    got, ref = _run("dopri5", "lv", LV_Y0, LV_Q, 5.0, 0.5, -0.1, ts, save_t0=True, save_t1=True)
    _assert_exact(got, ref)
    assert got["ts"][0, 0] == 5.0 and got["ts"][0, -1] == 0.5


def test_constant_step_size(cuda):
    got, ref = _run("tsit5", "vdp", VDP_Y0, VDP_MU, 0.0, 3.0, 0.01, np.linspace(0, 3, 5),
                    controller="constant", save_t1=True)
    _assert_exact(got, ref)
    assert (got["stats"][:, 0] == 300).all()


def test_save_steps(cuda):
    got, ref = _run("tsit5", "lv", LV_Y0[:32], LV_Q, 0.0, 4.0, 0.1, None, save_steps=True,
                    max_steps=256)
    _assert_exact(got, ref)


@pytest.mark.parametrize("k,max_steps", [(3, 256), (4, 256), (7, 300)])
def test_save_every_kth_step(cuda, k, max_steps):
    """SaveAt(steps=k) with t0, ts and t1: every k-th accepted step, the t1 row kept when
    the last step was not a multiple of k and max_steps is not (_integrate.py:493-527,
    847-877, 1272-1305)."""
    got, ref = _run("dopri5", "lv", LV_Y0[:64], LV_Q, 0.0, 4.0, 0.1, np.linspace(0, 4, 5),
                    save_t0=True, save_t1=True, save_steps=k, max_steps=max_steps)
    _assert_exact(got, ref)
    assert got["ys"].shape[1] == 1 + 5 + max_steps // k + (max_steps % k != 0)


def test_max_steps_reached_and_per_instance_t1(cuda):
    t1 = np.where(np.arange(256) % 2 == 0, 2.0, 40.0)
    got, ref = _run("tsit5", "vdp", VDP_Y0, VDP_MU, 0.0, t1, 0.1, None, save_t1=True,
                    max_steps=64)
    _assert_exact(got, ref)
    assert (got["result"] == 1).any() and (got["result"] == 0).any()


def test_dtmin_dtmax(cuda):
    got, ref = _run("tsit5", "vdp", VDP_Y0, VDP_MU, 0.0, 10.0, 0.1, None, save_t1=True,
                    dtmax=0.05, dtmin=1e-3, force_dtmin=False)
    _assert_exact(got, ref)


def test_generic_pid_coefficients(cuda):
    got, ref = _run("dopri5", "lv", LV_Y0, LV_Q, 0.0, 10.0, 0.1, np.linspace(0, 10, 11),
                    pcoeff=0.2, icoeff=0.4, dcoeff=0.1)
    _assert_exact(got, ref, tol=1e-12)


def test_t0_equals_t1(cuda):
    got, ref = _run("tsit5", "lv", LV_Y0[:8], LV_Q, 1.0, 1.0, 0.1, np.array([1.0, 1.0]),
                    save_t1=True)
    _assert_exact(got, ref)
    assert (got["stats"][:, 0] == 0).all()


def test_refill_queue_kernel_matches_static(cuda):
    ts = np.linspace(0.0, 8.0, 33)
    a, ref = _run("tsit5", "vdp", VDP_Y0, VDP_MU, 0.0, 8.0, 0.05, ts, launch=(64, 64))
    _assert_exact(a, ref)


def test_autotuned_launch_matches_oracle(cuda):
    """launch='auto' times the candidate shapes (static and refill, three block sizes) and
    solves with the fastest: the results stay bit-identical."""
    ts = np.linspace(0.0, 8.0, 33)
    a, ref = _run("tsit5", "vdp", VDP_Y0, VDP_MU, 0.0, 8.0, 0.05, ts, launch="auto")
    _assert_exact(a, ref)


def test_mixed_precision_form_error_bound(cuda):
    """float32 state / float64 time on Lotka-Volterra: within 100 * rtol of the FP64 oracle."""
    ts = np.linspace(0.0, 10.0, 100)
    got, ref = _run("dopri5", "lv", LV_Y0, LV_Q, 0.0, 10.0, 0.1, ts, dtype=torch.float32)
    err = oracle.max_rel_err(got["ys"], ref["ys"], 1e-8)
    print("mixed-precision LV max rel err", err)
    assert err < 1e-4


@pytest.mark.parametrize("dt0", [0.1, None])
def test_per_instance_tolerances(cuda, dt0):
    """rtol / atol per instance (the reference under jax.vmap over the controller's
    tolerances): bit-identical to the oracle, which equals B separate scalar-tolerance
    solves (tests/test_oracle_saveat.py pins that)."""
    rng = np.random.default_rng(21)
    B = 256
    rt = 10 ** rng.uniform(-9, -4, size=B)
    at = 10 ** rng.uniform(-10, -6, size=B)
    ts = np.linspace(0.0, 10.0, 11)
    sol = dfx.diffeqsolve(dfx.ODETerm(dfx.VanDerPol()), dfx.Tsit5(), 0.0, 10.0, dt0,
                          torch.tensor(VDP_Y0), torch.tensor(VDP_MU[:, 0]),
                          saveat=dfx.SaveAt(ts=ts),
                          stepsize_controller=dfx.PIDController(rtol=rt, atol=at),
                          max_steps=100000)
    ref = oracle.solve(VDP_Y0, 10.0, VDP_MU, ts=ts, solver="tsit5", vf="vdp", dt0=dt0,
                       rtol=rt, atol=at, max_steps=100000)
    assert np.array_equal(sol.ys.cpu().numpy(), ref["ys"])
    st = np.stack([sol.stats[k].cpu().numpy() for k in
                   ("num_steps", "num_accepted_steps", "num_rejected_steps")], -1)
    assert np.array_equal(st, ref["stats"])
    assert len(np.unique(st[:, 0])) > 20          # the tolerances really differ per instance


@pytest.mark.parametrize("dtmax", [None, 0.7])
def test_zero_and_nan_error_norms(cuda, dtmax):
    """The default controller's fifth root at its special inputs: instances whose field is
    zero (scaled error 0, so 1/error = inf and the factor clips to factormax) and instances
    started from NaN (error NaN -> inf, 1/error = 0), next to ordinary ones, with and
    without dtmax: identical to the oracle, NaN rows included."""
    rng = np.random.default_rng(33)
    B = 192
    A = (rng.normal(size=(B, 2, 2)) * 0.5 - np.eye(2)).reshape(B, 4)
    A[::3] = 0.0                         # zero field
    y0 = rng.uniform(-1.0, 1.0, size=(B, 2))
    y0[1::6, 0] = np.nan                 # NaN state
    ts = np.linspace(0.0, 6.0, 13)
    kw = {} if dtmax is None else {"dtmax": dtmax}
    got, ref = _run("tsit5", "linear", y0, A, 0.0, 6.0, 0.05, ts, max_steps=256, **kw)
    _assert_exact(got, ref)
    assert (got["stats"][::3, 2] == 0).all()          # zero field: nothing rejected
    assert (got["result"][1::6] != 0).all()           # NaN: the solve does not succeed


@pytest.mark.parametrize("ts_kind", ["one", "two_equal", "dense", "per_instance_to_t1"])
def test_save_time_prefetch_edges(cuda, ts_kind):
    """The next save time is read one save ahead: a single save time, two equal ones
    (two rows from one step), 1,001 save times (many saves inside one step) and per-instance
    rows that end exactly at t1: identical to the oracle."""
    rng = np.random.default_rng(44)
    B = 128
    y0 = VDP_Y0[:B]
    mu = VDP_MU[:B]
    if ts_kind == "one":
        ts = np.array([3.7])
    elif ts_kind == "two_equal":
        ts = np.array([2.5, 2.5])
    elif ts_kind == "dense":
        ts = np.linspace(0.0, 8.0, 1001)
    else:
        ts = np.sort(rng.uniform(0.0, 8.0, size=(B, 9)), axis=1)
        ts[:, -1] = 8.0
    got, ref = _run("tsit5", "vdp", y0, mu, 0.0, 8.0, 0.05, ts, save_t1=True)
    _assert_exact(got, ref)
    assert np.isfinite(got["ys"]).all()
