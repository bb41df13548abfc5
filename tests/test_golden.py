"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the
FP64 oracle): the oracle must reproduce them bit for bit (CPU), and the FP64 kernels must
match them bit for bit (GPU)."""
import os

import numpy as np
import pytest
import torch

import diffrax_amd as dfx
from tests.golden import make_golden

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(make_golden.BUILDERS))
def test_oracle_reproduces_fixture(name):
    want = _load(name)
    got = make_golden.BUILDERS[name]()
    for k, v in want.items():
        assert np.array_equal(np.asarray(got[k]), v), (name, k)


@pytest.mark.gpu
def test_kernels_match_fixtures(cuda):
    g = _load("c1_dopri5_lv_b16")
    sol = dfx.diffeqsolve(dfx.ODETerm(dfx.LotkaVolterra()), dfx.Dopri5(), 0.0, 10.0, 0.1,
                          torch.tensor(g["y0"]), tuple(g["q"]), saveat=dfx.SaveAt(ts=g["ts"]),
                          stepsize_controller=dfx.PIDController(rtol=1e-6, atol=1e-8))
    assert np.array_equal(sol.ys.cpu().numpy(), g["ys"])
    assert np.array_equal(sol.stats["num_steps"].cpu().numpy(), g["stats"][:, 0])
    g = _load("c2_tsit5_vdp_b16")
    sol = dfx.diffeqsolve(dfx.ODETerm(dfx.VanDerPol()), dfx.Tsit5(), 0.0, torch.tensor(g["t1"]),
                          0.1, torch.tensor(g["y0"]), torch.tensor(g["mu"]),
                          saveat=dfx.SaveAt(ts=g["ts"]), max_steps=16384,
                          stepsize_controller=dfx.PIDController(rtol=1e-6, atol=1e-9))
    assert np.array_equal(sol.ys.cpu().numpy(), g["ys"])
    assert np.array_equal(sol.stats["num_accepted_steps"].cpu().numpy(), g["stats"][:, 1])
    g = _load("c5_heun_ou_p8")
    terms = dfx.MultiTerm(dfx.ODETerm(dfx.OUDrift()),
                          dfx.ControlTerm(dfx.DiagonalNoise(), dfx.IncrementPath(torch.tensor(g["dW"]))))
    sol = dfx.diffeqsolve(terms, dfx.Heun(), 0.0, 1.0, 1.0 / g["dW"].shape[0], torch.tensor(g["y0"]),
                          tuple(torch.tensor(g[k]) for k in ("theta", "mu", "sigma")),
                          saveat=dfx.SaveAt(t1=True), stepsize_controller=dfx.ConstantStepSize())
    assert np.array_equal(sol.ys.cpu().numpy(), g["ys"])
