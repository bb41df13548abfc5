"""GPU parity of the small-system ERK kernel against the FP64 C oracle.

Bar (FP64 plain form): saved states within 1e-10 relative (max |y - y_ref| /
max(|y_ref|, atol)) and identical (num_steps, accepted, rejected) per trajectory."""
import numpy as np
import pytest
import torch

import diffrax_amd as dfx
import oracle

pytestmark = pytest.mark.gpu

LV_Q = np.array([1.5, 1.0, 3.0, 1.0])


def _gpu_solve(vf, solver, t0, t1, dt0, y0, args, ts, rtol, atol, dtype=torch.float64, **kw):
    saveat = kw.pop("saveat", None) or dfx.SaveAt(ts=ts)
    ctrl = kw.pop("controller", None) or dfx.PIDController(rtol=rtol, atol=atol)
    sol = dfx.diffeqsolve(dfx.ODETerm(vf), solver, t0, t1, dt0,
                          torch.as_tensor(y0, dtype=dtype), args, saveat=saveat,
                          stepsize_controller=ctrl, throw=False, **kw)
    torch.cuda.synchronize()
    stats = np.stack([sol.stats["num_steps"].cpu().numpy(),
                      sol.stats["num_accepted_steps"].cpu().numpy(),
                      sol.stats["num_rejected_steps"].cpu().numpy()], -1)
    return sol.ys.cpu().numpy(), sol.ts.cpu().numpy(), stats, sol.result.cpu().numpy()


def test_c1_dopri5_lotka_volterra_b1024(cuda):
    """Config: Dopri5 + PID(1e-6, 1e-8), LV, B=1024, y0~U[0.5,2]^2, t in [0,10], dt0=0.1,
    SaveAt(ts=linspace(0,10,100))."""
    rng = np.random.default_rng(0)
    y0 = rng.uniform(0.5, 2.0, size=(1024, 2))
    ts = np.linspace(0.0, 10.0, 100)
    ref = oracle.solve(y0, 10.0, LV_Q, ts=ts, solver="dopri5", vf="lv", dt0=0.1,
                       rtol=1e-6, atol=1e-8)
    ys, tso, stats, res = _gpu_solve(dfx.LotkaVolterra(), dfx.Dopri5(), 0.0, 10.0, 0.1, y0,
                                     tuple(LV_Q), ts, 1e-6, 1e-8)
    print("c1 stats mismatches", int((stats != ref["stats"]).any(-1).sum()))
    assert np.array_equal(stats, ref["stats"])
    assert np.array_equal(res, ref["result"])
    assert np.array_equal(tso, ref["ts"])
    err = oracle.max_rel_err(ys, ref["ys"], 1e-8)
    print("c1 max rel err", err)
    assert err <= 1e-10


def test_c2_tsit5_vdp_uneven_full(cuda):
    """Full config: Tsit5 + PID(1e-6, 1e-9), 65,536 VdP oscillators, mu~LogU[0.1,5],
    y0~U[-2,2]^2, t1~U[5,50], 64 saves scaled to [0, t1] per instance."""
    rng = np.random.default_rng(1)
    B = 65536
    mu = np.exp(rng.uniform(np.log(0.1), np.log(5.0), size=B))
    y0 = rng.uniform(-2.0, 2.0, size=(B, 2))
    t1 = rng.uniform(5.0, 50.0, size=B)
    ts = np.linspace(0.0, 1.0, 64)[None, :] * t1[:, None]
    ref = oracle.solve(y0, t1, mu[:, None], ts=ts, solver="tsit5", vf="vdp", dt0=0.1,
                       rtol=1e-6, atol=1e-9, max_steps=16384)
    ys, tso, stats, res = _gpu_solve(dfx.VanDerPol(), dfx.Tsit5(), 0.0, torch.tensor(t1), 0.1,
                                     y0, torch.tensor(mu), ts, 1e-6, 1e-9, max_steps=16384)
    mism = int((stats != ref["stats"]).any(-1).sum())
    print("c2 stats mismatches", mism, "of", B)
    assert mism == 0
    err = oracle.max_rel_err(ys, ref["ys"], 1e-9)
    print("c2 max rel err", err)
    assert err <= 1e-10
