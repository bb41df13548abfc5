"""Writes the golden fixtures in this folder from the FP64 oracle.

Each fixture is a small seeded subset of one benchmark configuration: inputs plus the
oracle's saved states and step statistics.  The CPU test suite regenerates them and
requires bit equality (drift of the oracle fails), the GPU suite compares the kernels
with them.  Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402


def c1(seed=100, B=16):
    rng = np.random.default_rng(seed)
    y0 = rng.uniform(0.5, 2.0, size=(B, 2))
    q = np.array([1.5, 1.0, 3.0, 1.0])
    ts = np.linspace(0.0, 10.0, 100)
    r = oracle.solve(y0, 10.0, q, ts=ts, solver="dopri5", vf="lv", dt0=0.1, rtol=1e-6, atol=1e-8)
    return dict(y0=y0, q=q, ts=ts, ys=r["ys"], stats=r["stats"], result=r["result"])


def c2(seed=101, B=16):
    rng = np.random.default_rng(seed)
    mu = np.exp(rng.uniform(np.log(0.1), np.log(5.0), size=B))
    y0 = rng.uniform(-2.0, 2.0, size=(B, 2))
    t1 = rng.uniform(5.0, 50.0, size=B)
    ts = np.linspace(0.0, 1.0, 64)[None, :] * t1[:, None]
    r = oracle.solve(y0, t1, mu[:, None], ts=ts, solver="tsit5", vf="vdp", dt0=0.1,
                     rtol=1e-6, atol=1e-9, max_steps=16384)
    return dict(mu=mu, y0=y0, t1=t1, ts=ts, ys=r["ys"], stats=r["stats"], result=r["result"])


def c5(seed=102, P=8, d=16, N=100):
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0.5, 2.0, size=d)
    mu = rng.normal(0.0, 1.0, size=d)
    sigma = rng.uniform(0.1, 0.5, size=d)
    dW = (rng.standard_normal(size=(N, P, d)) * np.sqrt(1.0 / N)).astype(np.float32)
    y0 = np.zeros((P, d))
    q = np.concatenate([theta, mu, sigma])
    r = oracle.solve(y0, 1.0, q, dW=dW.astype(np.float64), solver="heun", vf="ou",
                     controller="constant", dt0=1.0 / N, save_t1=True, max_steps=10 * N)
    return dict(theta=theta, mu=mu, sigma=sigma, dW=dW, y0=y0, ys=r["ys"], stats=r["stats"])


BUILDERS = {"c1_dopri5_lv_b16": c1, "c2_tsit5_vdp_b16": c2, "c5_heun_ou_p8": c5}


def main():
    for name, fn in BUILDERS.items():
        np.savez(os.path.join(HERE, name + ".npz"), **fn())
        print("wrote", name)


if __name__ == "__main__":
    main()
