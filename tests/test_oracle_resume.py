"""Pins of the Python oracle's solver / controller state export and resume
(diffeqsolve(solver_state=, controller_state=, made_jump=), _integrate.py:1252-1270, 1309;
SaveAt(solver_state=True, controller_state=True, made_jump=True), 1489-1500).

The defining property: a solve cut short by max_steps and resumed from its final state
(y, solver state, controller state, made_jump, next step size) continues the uncut solve
step for step.  Each piece of state matters for that, so dropping or misplacing any one
of them breaks the equality (checked below for the controller state, whose omission is
the likeliest mistake)."""
import numpy as np
import pytest

from oracle import pyoracle


def vdp(mu):
    return lambda t, y: np.array([y[1], mu * (1 - y[0] * y[0]) * y[1] - y[0]])


CASES = [
    ("tsit5", dict(controller="pid", rtol=1e-6, atol=1e-9)),
    ("tsit5", dict(controller="pid", rtol=1e-6, atol=1e-9, pcoeff=0.3, icoeff=0.5,
                   dcoeff=0.1)),
    ("dopri5", dict(controller="pid", rtol=1e-7, atol=1e-9, pcoeff=0.2, icoeff=0.4)),
    ("heun", dict(controller="pid", rtol=1e-4, atol=1e-6)),
    ("tsit5", dict(controller="constant")),
    ("heun", dict(controller="constant")),
]


def _resume_from(part, f, solver, t1, kw, **over):
    """The resumed solve of a max_steps-interrupted part, from its last state."""
    t0 = part["tprev"]
    dt0 = part["tnext"] - part["tprev"]
    assert t0 + dt0 == part["tnext"]   # exact (Sterbenz), so the first step is the same
    args = dict(solver_state=part["solver_state"], controller_state=part["controller_state"],
                made_jump=part["made_jump"])
    args.update(over)
    return pyoracle.solve(f, solver, t0, t1, dt0, part["ys"][-1], [], save_t1=True,
                          max_steps=100000, **kw, **args)


@pytest.mark.parametrize("solver,kw", CASES)
@pytest.mark.parametrize("m", [1, 7, 40])
def test_interrupted_and_resumed_equals_uncut(solver, kw, m):
    f = vdp(2.0)
    y0 = np.array([1.5, -0.5])
    t1 = 6.0
    dt0 = 0.05
    full = pyoracle.solve(f, solver, 0.0, t1, dt0, y0, [], save_t1=True, max_steps=100000, **kw)
    part = pyoracle.solve(f, solver, 0.0, t1, dt0, y0, [], save_t1=True, max_steps=m, **kw)
    assert part["result"] == 1 and part["stats"][0] == m
    rest = _resume_from(part, f, solver, t1, kw)
    n_acc = part["stats"][1]
    assert rest["grid"] == full["grid"][n_acc:]
    assert np.array_equal(rest["ys"][-1], full["ys"][-1])
    assert tuple(a + b for a, b in zip(part["stats"], rest["stats"])) == tuple(full["stats"])
    # the final states agree too
    assert rest["controller_state"] == full["controller_state"]
    if full["solver_state"] is not None:
        assert rest["solver_state"][0] == full["solver_state"][0]
        assert np.array_equal(rest["solver_state"][1], full["solver_state"][1])


def test_controller_state_is_needed():
    """Resuming without the PID state (prev_inv = prev_prev_inv = 1) takes other steps."""
    f = vdp(2.0)
    kw = dict(controller="pid", rtol=1e-6, atol=1e-9, pcoeff=0.3, icoeff=0.5, dcoeff=0.1)
    y0 = np.array([1.5, -0.5])
    full = pyoracle.solve(f, "tsit5", 0.0, 6.0, 0.05, y0, [], save_t1=True, max_steps=100000,
                          **kw)
    part = pyoracle.solve(f, "tsit5", 0.0, 6.0, 0.05, y0, [], save_t1=True, max_steps=40, **kw)
    fresh = _resume_from(part, f, "tsit5", 6.0, kw, controller_state=None)
    assert fresh["grid"] != full["grid"][part["stats"][1]:]


@pytest.mark.parametrize("solver", ["tsit5", "dopri5"])
def test_saved_fsal_stage_is_the_vector_field_at_the_end(solver):
    """The saved solver state's f0 is f(t, y) at the final state, evaluated afresh (the
    last stage of an FSAL step is taken at c = 1 on the SSAL / b-row state)."""
    f = vdp(1.3)
    for m in (3, 25, 100000):
        r = pyoracle.solve(f, solver, 0.0, 4.0, 0.1, np.array([0.7, 0.2]), [], save_t1=True,
                           max_steps=m, rtol=1e-7, atol=1e-9)
        first, f0 = r["solver_state"]
        assert not first
        assert np.array_equal(f0, f(r["tprev"], r["ys"][-1]))


def test_first_step_and_made_jump_reevaluate_the_same_stage():
    """first_step or made_jump make the resumed solve evaluate its first stage afresh; at
    the saved state that evaluation equals the saved f0, so the solve is unchanged, and
    made_jump is cleared by the first accepted step."""
    f = vdp(2.0)
    kw = dict(controller="pid", rtol=1e-6, atol=1e-9)
    part = pyoracle.solve(f, "tsit5", 0.0, 5.0, 0.05, np.array([1.5, -0.5]), [], save_t1=True,
                          max_steps=30, **kw)
    base = _resume_from(part, f, "tsit5", 5.0, kw)
    zero = np.zeros(2)
    for over in (dict(made_jump=True), dict(solver_state=(True, zero))):
        r = _resume_from(part, f, "tsit5", 5.0, kw, **over)
        assert r["grid"] == base["grid"] and np.array_equal(r["ys"], base["ys"])
        assert r["made_jump"] is False
    # a wrong f0 with first_step False is used as given (and changes the solve)
    r = _resume_from(part, f, "tsit5", 5.0, kw, solver_state=(False, zero))
    assert not np.array_equal(r["ys"][-1], base["ys"][-1])


def test_fresh_state_values_reproduce_a_fresh_solve():
    """The initial states the reference builds (runge_kutta.py:424-442, pid.py:389-392,
    constant.py:47-53) passed explicitly change nothing."""
    f = vdp(0.8)
    y0 = np.array([0.3, 1.1])
    a = pyoracle.solve(f, "tsit5", 0.0, 3.0, 0.1, y0, [0.5, 2.0], rtol=1e-6, atol=1e-9)
    b = pyoracle.solve(f, "tsit5", 0.0, 3.0, 0.1, y0, [0.5, 2.0], rtol=1e-6, atol=1e-9,
                       solver_state=(True, np.zeros(2)), controller_state=(1.0, 1.0, False),
                       made_jump=False)
    assert a["grid"] == b["grid"] and np.array_equal(a["ys"], b["ys"])
    num = int(np.ceil(3.0 / np.nextafter(0.1, np.inf)))
    c = pyoracle.solve(f, "heun", 0.0, 3.0, 0.1, y0, [1.0], controller="constant")
    e = pyoracle.solve(f, "heun", 0.0, 3.0, 0.1, y0, [1.0], controller="constant",
                       controller_state=(1, num, 0.0, 3.0))
    assert c["grid"] == e["grid"] and np.array_equal(c["ys"], e["ys"])
    # steps_completed starts at 1 and counts every step (constant.py:47, 77-79)
    assert c["controller_state"] == (num + 1, num, 0.0, 3.0)
