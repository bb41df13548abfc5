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
        g = np.asarray(got[k])
        if name in make_golden.EXACT or not np.issubdtype(v.dtype, np.floating):
            assert np.array_equal(g, v, equal_nan=True), (name, k)
        else:
            assert np.allclose(g, v, rtol=1e-12, atol=0.0, equal_nan=True), (name, k)


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


def test_mlp_replay_gradient_vs_central_differences():
    """Pins the neural-ODE gradient reference (torch float64 autograd through the oracle's
    accepted grid, fixture c4): the fixture's directional derivatives along 32 seeded unit
    directions of (W1, b1, W2, b2, W3, b3), and dL/dy0 along 4 directions, against central
    differences of the same replay (h = 1e-6), 1e-6 relative."""
    g = _load("c4_mlp_grad_b4")
    W = [torch.tensor(g[k].astype(np.float64)) for k in make_golden.MLP_KEYS]
    y0 = torch.tensor(g["y0"].astype(np.float64))
    L = lambda Wp, y: float(make_golden.mlp_replay_loss(Wp, y, g["ts"], g["grid"], g["nacc"],
                                                        g["wloss"]))
    h = 1e-6
    dirs = make_golden.unit_directions(int(g["dir_seed"]), len(g["dir_derivs"]),
                                       [w.shape for w in W])
    for d, want in zip(dirs, g["dir_derivs"]):
        plus = [w + h * torch.tensor(v) for w, v in zip(W, d)]
        minus = [w - h * torch.tensor(v) for w, v in zip(W, d)]
        fd = (L(plus, y0) - L(minus, y0)) / (2 * h)
        assert abs(fd - want) <= 1e-6 * max(1.0, abs(want)), (fd, want)
    rng = np.random.default_rng(5)
    for _ in range(4):
        v = rng.normal(size=y0.shape)
        v /= np.linalg.norm(v)
        fd = (L(W, y0 + h * torch.tensor(v)) - L(W, y0 - h * torch.tensor(v))) / (2 * h)
        an = float((g["y0_bar"] * v).sum())
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (fd, an)


@pytest.mark.gpu
def test_mlp_kernels_match_fixtures(cuda):
    """Neural-ODE forward (float32 MFMA form) against fixture c3 at matched tolerance, and
    the device discrete adjoint against fixture c4's directional derivatives (1e-3: the
    device's float32 accepted grid may differ from the oracle's in the last steps)."""
    g = _load("c3_mlp_tsit5_b8")
    vf = dfx.MLPField(*(torch.tensor(g[k]) for k in make_golden.MLP_KEYS), matmul="f32")
    sol = dfx.diffeqsolve(dfx.ODETerm(vf), dfx.Tsit5(), 0.0, 1.0, 0.1, torch.tensor(g["y0"]),
                          saveat=dfx.SaveAt(ts=g["ts"]),
                          stepsize_controller=dfx.PIDController(rtol=1e-5, atol=1e-7))
    ys = sol.ys.double().cpu().numpy()
    scale = 1e-7 + 1e-5 * np.abs(g["ys_tight"])
    e_dev = np.max(np.abs(ys - g["ys_tight"]) / scale)
    e_ora = np.max(np.abs(g["ys"] - g["ys_tight"]) / scale)
    assert e_dev <= e_ora + 1.0, (e_dev, e_ora)
    g = _load("c4_mlp_grad_b4")
    dev = torch.device("cuda", 0)
    P = [torch.tensor(g[k], device=dev, requires_grad=True) for k in make_golden.MLP_KEYS]
    y0 = torch.tensor(g["y0"], device=dev, requires_grad=True)
    sol = dfx.diffeqsolve(dfx.ODETerm(dfx.MLPField(*P, matmul="f32")), dfx.Tsit5(), 0.0, 1.0,
                          0.1, y0, saveat=dfx.SaveAt(ts=g["ts"]),
                          stepsize_controller=dfx.PIDController(rtol=1e-5, atol=1e-7),
                          adjoint=dfx.RecursiveCheckpointAdjoint())
    (sol.ys * torch.tensor(g["wloss"], dtype=torch.float32, device=dev)).sum().backward()
    gW = [p.grad.double().cpu().numpy() for p in P]
    dirs = make_golden.unit_directions(int(g["dir_seed"]), len(g["dir_derivs"]),
                                       [w.shape for w in gW])
    got = np.array([sum(float((a * v).sum()) for a, v in zip(gW, d)) for d in dirs])
    want = g["dir_derivs"]
    assert np.max(np.abs(got - want)) <= 1e-3 * np.max(np.abs(want)), (got, want)
    yb = y0.grad.double().cpu().numpy()
    assert np.max(np.abs(yb - g["y0_bar"])) <= 1e-3 * np.max(np.abs(g["y0_bar"]))
