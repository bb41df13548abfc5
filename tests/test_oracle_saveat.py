"""Pins of the oracle's SaveAt(steps=k) (_integrate.py:493-527, 847-877, 1272-1305) on
the CPU, against the same solve saved at every step (SaveAt(steps=1)):
- rows are the end states of accepted steps k, 2k, ... in order, with their times;
- the t1 row follows them unless the number of accepted steps is a multiple of k (then
  the last step's row already holds y(t1)), and is allocated only when max_steps is not a
  multiple of k;
- unfilled rows are inf."""
import numpy as np
import pytest

import oracle


@pytest.mark.parametrize("k", [1, 2, 3, 7])
def test_every_kth_step(k):
    rng = np.random.default_rng(k)
    y0 = rng.uniform(0.5, 2.0, size=(8, 2))
    q = np.array([1.5, 1.0, 3.0, 1.0])
    kw = dict(solver="tsit5", vf="lv", dt0=0.1, rtol=1e-6, atol=1e-8, max_steps=300)
    every = oracle.solve(y0, 10.0, q, save_steps=1, **kw)
    got = oracle.solve(y0, 10.0, q, save_steps=k, save_t1=True, **kw)
    n_out = 300 // k + (300 % k != 0)
    assert got["ys"].shape == (8, n_out, 2)
    for b in range(8):
        n = int(every["stats"][b, 1])
        m = n // k
        assert np.array_equal(got["ys"][b, :m], every["ys"][b, k - 1:m * k:k])
        assert np.array_equal(got["ts"][b, :m], every["ts"][b, k - 1:m * k:k])
        rows = m
        if n % k != 0 and n_out > m:
            assert np.array_equal(got["ys"][b, m], every["ys"][b, n - 1])
            assert got["ts"][b, m] == 10.0
            rows += 1
        assert np.isinf(got["ys"][b, rows:]).all() and np.isinf(got["ts"][b, rows:]).all()


def test_per_instance_tolerances_equal_separate_solves():
    """rtol / atol arrays [B] (a vmapped PIDController): each instance equals its own
    scalar-tolerance solve, with dt0 given or selected automatically."""
    rng = np.random.default_rng(3)
    B = 12
    y0 = rng.uniform(0.5, 2.0, size=(B, 2))
    q = np.array([1.5, 1.0, 3.0, 1.0])
    rt, at = 10 ** rng.uniform(-9, -3, B), 10 ** rng.uniform(-10, -6, B)
    ts = np.linspace(0.0, 10.0, 21)
    for dt0 in (0.1, None):
        kw = dict(solver="tsit5", vf="lv", dt0=dt0, max_steps=100000)
        both = oracle.solve(y0, 10.0, q, ts=ts, rtol=rt, atol=at, **kw)
        for b in range(B):
            one = oracle.solve(y0[b:b + 1], 10.0, q, ts=ts, rtol=rt[b], atol=at[b], **kw)
            assert np.array_equal(one["ys"][0], both["ys"][b])
            assert np.array_equal(one["stats"][0], both["stats"][b])
