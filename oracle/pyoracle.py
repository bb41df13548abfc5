"""Per-trajectory Python restatement of diffrax's explicit-RK path (test infrastructure).

Independent of ``erk_oracle.c`` (written separately, same reference lines):

* loop / saves / clip-to-end ...... _integrate.py:278-284, 302-491, 823-883
* ERK step (FSAL/SSAL) ............ _solver/runge_kutta.py:643-701, 847-1052, 1170-1203
* tableaux ........................ _solver/tsit5.py:18-100, dopri5.py:11-33, heun.py:12-17
* interpolation ................... _solver/tsit5.py:116-156, _local_interpolation.py:50-139
* PID / constant controllers ...... _step_size_controller/pid.py:33-81, 316-567,
                                    _step_size_controller/constant.py:30-104
* continuous adjoint (backsolve) .. _adjoint.py:646-735, _term.py:767-917; the adjoint
                                    seminorm _misc.py:85-116

The vector field is any Python callable ``f(t, y) -> array`` over numpy (real or
complex) or torch tensors, so the same code serves DETEST problems, complex-step
derivatives and torch autograd replays of a fixed accepted-step grid (the reference
stops gradients through the step-size factor, pid.py:498-530, so its gradient is the
derivative of exactly such a replay).

Every jnp.tensordot of the reference (the stage, b, b_err, Dopri5-midpoint and Tsit5
interpolant sums, _solver/base.py:37-41) is restated as a fused multiply-add chain in
index order: the first nonzero product, then acc = fma(c_m, k_m, acc), each fma rounded
once (exactly, through fractions.Fraction); y0 is added afterwards.
"""
from __future__ import annotations

import math

from fractions import Fraction

import numpy as np

TSIT5_A = [
    [],
    [161 / 1000],
    [-0.8480655492356988544426874250230774675121177393430391537369234245294192976164141156943e-2,
     0.3354806554923569885444268742502307746751211773934303915373692342452941929761641411569],
    [2.897153057105493432130432594192938764924887287701866490314866693455023795137503079289,
     -6.359448489975074843148159912383825625952700647415626703305928850207288721235210244366,
     4.362295432869581411017727318190886861027813359713760212991062156752264926097707165077],
    [5.325864828439256604428877920840511317836476253097040101202360397727981648835607691791,
     -11.74888356406282787774717033978577296188744178259862899288666928009020615663593781589,
     7.495539342889836208304604784564358155658679161518186721010132816213648793440552049753,
     -0.9249506636175524925650207933207191611349983406029535244034750452930469056411389539635e-1],
    [5.861455442946420028659251486982647890394337666164814434818157239052507339770711679748,
     -12.92096931784710929170611868178335939541780751955743459166312250439928519268343184452,
     8.159367898576158643180400794539253485181918321135053305748355423955009222648673734986,
     -0.7158497328140099722453054252582973869127213147363544882721139659546372402303777878835e-1,
     -0.2826905039406838290900305721271224146717633626879770007617876201276764571291579142206e-1],
    [0.9646076681806522951816731316512876333711995238157997181903319145764851595234062815396e-1,
     1 / 100,
     0.4798896504144995747752495322905965199130404621990332488332634944254542060153074523509,
     1.379008574103741893192274821856872770756462643091360525934940067397245698027561293331,
     -3.290069515436080679901047585711363850115683290894936158531296799594813811049925401677,
     2.324710524099773982415355918398765796109060233222962411944060046314465391054716027841],
]
TSIT5_B = TSIT5_A[6] + [0.0]
TSIT5_BHAT = [
    0.9468075576583945807478876255758922856117527357724631226139574065785592789071067303271e-1,
    0.9183565540343253096776363936645313759813746240984095238905939532922955247253608687270e-2,
    0.4877705284247615707855642599631228241516691959761363774365216240304071651579571959813,
    1.234297566930478985655109673884237654035539930748192848315425833500484878378061439761,
    -2.707712349983525454881109975059321670689605166938197378763992255714444407154902012702,
    1.866628418170587035753719399566211498666255505244122593996591602841258328965767580089,
]
TSIT5_BERR = [b - bh for b, bh in zip(TSIT5_B[:6], TSIT5_BHAT)] + [-1 / 66]
TSIT5_C = [0.0, 161 / 1000, 327 / 1000, 9 / 10,
           0.9800255409045096857298102862870245954942137979563024768854764293221195950761080302604,
           1.0, 1.0]

DOPRI5_A = [
    [],
    [1 / 5],
    [3 / 40, 9 / 40],
    [44 / 45, -56 / 15, 32 / 9],
    [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729],
    [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656],
    [35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84],
]
DOPRI5_B = [35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84, 0]
DOPRI5_BERR = [35 / 384 - 1951 / 21600, 0, 500 / 1113 - 22642 / 50085, 125 / 192 - 451 / 720,
               -2187 / 6784 - -12231 / 42400, 11 / 84 - 649 / 6300, -1.0 / 60.0]
DOPRI5_C = [0.0, 1 / 5, 3 / 10, 4 / 5, 8 / 9, 1.0, 1.0]
DOPRI5_CMID = [6025192743 / 30085553152 / 2, 0, 51252292925 / 65400821598 / 2,
               -2691868925 / 45128329728 / 2, 187940372067 / 1594534317056 / 2,
               -1776094331 / 19743644256 / 2, 11237099 / 235043384 / 2]

HEUN_A = [[], [1.0]]
HEUN_B = [0.5, 0.5]
HEUN_BERR = [0.5, -0.5]
HEUN_C = [0.0, 1.0]


class Tableau:
    def __init__(self, name, a, b, berr, c, order, cmid=None):
        self.name = name
        self.a = a
        self.b = b
        self.berr = berr
        self.c = c
        self.order = order
        self.cmid = cmid
        self.S = len(b)
        # runge_kutta.py:143-165: fsal/ssal iff the last a-row equals b_sol[:-1]
        last = list(a[-1])
        self.fsal = self.ssal = (last == list(b[:-1])) and b[-1] == 0


TABLEAUX = {
    "tsit5": Tableau("tsit5", TSIT5_A, TSIT5_B, TSIT5_BERR, TSIT5_C, 5),
    "dopri5": Tableau("dopri5", DOPRI5_A, DOPRI5_B, DOPRI5_BERR, DOPRI5_C, 5, DOPRI5_CMID),
    "heun": Tableau("heun", HEUN_A, HEUN_B, HEUN_BERR, HEUN_C, 2),
}


def _fma1(a, b, c):
    """a*b + c rounded once to float64 (C fma, v_fma_f64), by exact rational arithmetic."""
    if not (math.isfinite(a) and math.isfinite(b)):
        return a * b + c          # an infinite / NaN factor: IEEE a*b + c is the fma result
    if not math.isfinite(c):
        return c                  # the exact product is finite, so inf / NaN c decides
    exact = Fraction(a) * Fraction(b) + Fraction(c)
    if exact == 0:
        # IEEE sign of an exact zero sum: -0 only when both addends are -0
        prod_neg = (math.copysign(1.0, a) * math.copysign(1.0, b)) < 0
        prod_zero = a == 0 or b == 0
        return -0.0 if (prod_zero and prod_neg and math.copysign(1.0, c) < 0) else 0.0
    try:
        return float(exact)
    except OverflowError:
        return math.inf if exact > 0 else -math.inf


def _fma(a, k, acc):
    """acc + a*k with one rounding for float64 values (the restatement of the fused dot
    product); torch tensors and complex values (the autograd / complex-step replays, which
    are compared at 1e-12) use a*k + acc."""
    if isinstance(k, np.ndarray) and k.dtype == np.float64:
        out = np.empty(np.broadcast(k, acc).shape)
        kk, aa = np.broadcast_arrays(k, acc)
        for i in np.ndindex(out.shape):
            out[i] = _fma1(float(a), float(kk[i]), float(aa[i]))
        return out
    if isinstance(k, (float, np.floating)) and isinstance(acc, (float, np.floating)):
        return _fma1(float(a), float(k), float(acc))
    return a * k + acc


def _lincomb(coeffs, ks):
    acc = None
    for a, k in zip(coeffs, ks):
        if a == 0:
            continue
        acc = a * k if acc is None else _fma(a, k, acc)
    return acc


def _lincomb_plain(coeffs, ks):
    """The same sum in index order with separately rounded products and sums (numpy speed
    for long states, where the float32 kernels it checks round differently anyway)."""
    acc = None
    for a, k in zip(coeffs, ks):
        if a == 0:
            continue
        acc = a * k if acc is None else acc + a * k
    return acc


def rk_step(tab, prod, n, ts0, ts1, direction, y0, f0=None, lincomb=_lincomb):
    """One ERK step.  prod(t, y, control, n) returns the k-value (control * f, plus
    the diffusion increment for SDEs).  f0: FSAL value (k0 = control * f0).
    Returns (y1, y_err, ks, f_last)."""
    dt = ts1 - ts0
    control = direction * dt
    ks = []
    f_last = None
    z = None
    for i in range(tab.S):
        if i == 0 and f0 is not None and tab.fsal:
            ks.append(control * f0)
            continue
        inc = lincomb(tab.a[i], ks) if i > 0 else None
        z = y0 if inc is None else y0 + inc
        ti = ts1 if tab.c[i] == 1 else ts0 + tab.c[i] * dt
        k, f = prod(ti * direction, z, control, n)
        ks.append(k)
        if i == tab.S - 1:
            f_last = f
    if tab.ssal:
        y1 = z
    else:
        y1 = y0 + lincomb(tab.b, ks)
    y_err = lincomb(tab.berr, ks)
    return y1, y_err, ks, f_last


def linear_rescale(t0, t, t1):
    if t0 == t1:
        return 0.0
    return (t - t0) / (t1 - t0)


def interpolate(tab, ts0, ts1, y0, y1, ks, t):
    th = linear_rescale(ts0, t, ts1)
    if tab.name == "tsit5":
        t2 = th * th
        bb = [
            -1.0530884977290216 * th * (th - 1.3299890189751412)
            * (t2 - 1.4364028541716351 * th + 0.7139816917074209),
            0.1017 * t2 * (t2 - 2.1966568338249754 * th + 1.2949852507374631),
            2.490627285651252793 * t2 * (t2 - 2.38535645472061657 * th + 1.57803468208092486),
            -16.54810288924490272 * (th - 1.21712927295533244) * (th - 0.61620406037800089) * t2,
            47.37952196281928122 * (th - 1.203071208372362603) * (th - 0.658047292653547382) * t2,
            -34.87065786149660974 * (th - 1.2) * (th - 0.666666666666666667) * t2,
            2.5 * (th - 1) * (th - 0.6) * t2,
        ]
        acc = bb[0] * ks[0]
        for m in range(1, 7):
            acc = _fma(bb[m], ks[m], acc)
        return y0 + acc
    if tab.name == "dopri5":
        ymid = y0 + _lincomb(tab.cmid, ks)
        f0, f1 = ks[0], ks[-1]
        a = 2 * (f1 - f0) - 8 * (y1 + y0) + 16 * ymid
        b = 5 * f0 - 3 * f1 + 18 * y0 + 14 * y1 - 32 * ymid
        c = f1 - 4 * f0 - 11 * y0 - 5 * y1 + 16 * ymid
        return (((a * th + b) * th + c) * th + f0) * th + y0
    k0, k1 = ks[0], ks[-1]
    a = k0 + k1 + 2 * y0 - 2 * y1
    b = -2 * k0 - k1 - 3 * y0 + 3 * y1
    return ((a * th + b) * th + k0) * th + y0


def pow_ctl(x, c):
    """Controller power, correctly rounded for c = 0.2 via 60-digit decimal arithmetic
    (x^0.5 is sqrt, x^1 is x, other exponents use the platform pow)."""
    if c == 0.2:
        if math.isnan(x) or x < 0:
            return math.nan
        if x == 0 or math.isinf(x):
            return x
        from decimal import Decimal, localcontext
        with localcontext() as ctx:
            ctx.prec = 60
            return float(Decimal(x) ** Decimal(c))
    if c == 1:
        return x
    if c == 0.5:
        return math.sqrt(x) if x >= 0 else math.nan
    return x ** c


def _rms(x):
    x = np.asarray(x, dtype=np.float64).ravel()
    if x.size == 1:
        return abs(float(x[0]))
    return math.sqrt(float(sum(float(v) * float(v) for v in x))) / math.sqrt(x.size)


def _jmax(a, b):
    return math.nan if (math.isnan(a) or math.isnan(b)) else max(a, b)


def _jmin(a, b):
    return math.nan if (math.isnan(a) or math.isnan(b)) else min(a, b)


def select_initial_step(f, t0, y0, direction, rtol, atol, error_order):
    """pid.py:33-81."""
    scale = atol + np.abs(y0) * rtol
    f0 = f(t0 * direction, y0)
    d0 = _rms(y0 / scale)
    d1 = _rms(f0 / scale)
    cond = d0 < 1e-5 or d1 < 1e-5
    _d1 = 1.0 if cond else d1
    h0 = 1e-6 if cond else 0.01 * (d0 / _d1)
    y1 = y0 + h0 * f0
    f1 = f((t0 + h0) * direction, y1)
    d2 = _rms((f1 - f0) / scale) / h0
    max_d = _jmax(d1, d2)
    h1 = _jmax(1e-6, h0 * 1e-3) if max_d <= 1e-15 else pow_ctl(0.01 / max_d, 1 / error_order)
    return _jmin(100 * h0, h1)


def solve(f, solver, t0, t1, dt0, y0, ts=(), *, rtol=1e-6, atol=1e-8, controller="pid",
          save_t0=False, save_t1=False, max_steps=4096, pcoeff=0.0, icoeff=1.0, dcoeff=0.0,
          dtmin=None, dtmax=None, force_dtmin=True, factormin=0.2, factormax=10.0,
          safety=0.9, g_dw=None, norm_n=None, fused=True, solver_state=None,
          controller_state=None, made_jump=None):
    """FP64 numpy solve of one trajectory.  Returns dict(ys, ts, stats, result, grid,
    direction, tprev, tnext, solver_state, controller_state, made_jump).  g_dw(n, t, y)
    adds a diffusion increment for SDEs.  norm_n: the error norm is the RMS over the first
    norm_n components only (adjoint_rms_seminorm, _misc.py:85-116, on a state
    (y, a_y, a_params) with norm_n = 2 d).  fused=False sums the stages with plain
    arithmetic (_lincomb_plain).

    Resuming (_integrate.py:1252-1270, 1309): solver_state = (first_step, f0) of an FSAL
    solver (runge_kutta.py:415-442; None for Heun), controller_state = (prev_inv,
    prev_prev_inv, at_dtmin) for the PID controller (pid.py:84-85) or (steps_completed,
    num_steps, t0_sim, t1_sim) for ConstantStepSize (constant.py:14-17), in normalised
    time; with a controller state and dt0 given the first step ends at t0 + dt0 without
    the PID init's clamps.  The returned states are the final ones (_integrate.py:
    1489-1500)."""
    lincomb = _lincomb if fused else _lincomb_plain
    tab = TABLEAUX[solver]
    y = np.array(y0, dtype=np.float64)
    direction = 1 if t0 < t1 else -1
    T0, T1 = t0 * direction, t1 * direction
    ts_n = [t * direction for t in ts]

    def prod(t, z, control, n):
        fz = f(t, z)
        k = control * fz
        if g_dw is not None:
            k = k + g_dw(n, t, z)
        return k, fz

    error_order = tab.order
    outs_t, outs_y = [], []
    if save_t0:
        outs_t.append(T0)
        outs_y.append(y.copy())
    if controller == "pid":
        dt0v = dt0 * direction if dt0 is not None else select_initial_step(
            f, T0, y, direction, rtol, atol, error_order)
        if dtmax is not None:
            dt0v = _jmin(dt0v, dtmax)
        if dtmin is not None:
            dt0v = _jmax(dt0v, dtmin)
        tnext = T0 + dt0v
        pinv = ppinv = 1.0
        at_dtmin = False
        if controller_state is not None:
            pinv, ppinv, at_dtmin = (float(controller_state[0]), float(controller_state[1]),
                                     bool(controller_state[2]))
            if dt0 is not None:
                tnext = T0 + dt0 * direction
    else:
        DT0 = dt0 * direction
        num = int(math.ceil((T1 - T0) / np.nextafter(DT0, math.inf)))
        completed = 1
        c_t0, c_t1 = T0, T1
        tnext = T0 + DT0
        if controller_state is not None:
            completed, num = int(controller_state[0]), int(controller_state[1])
            c_t0, c_t1 = float(controller_state[2]), float(controller_state[3])
    tnext = _jmin(tnext, T1)
    # solver state: (first_step, f0); the first stage is evaluated while first_step or
    # made_jump holds (runge_kutta.py:684-690)
    first, f0 = True, np.zeros_like(y)
    if solver_state is not None and tab.fsal:
        first, f0 = bool(solver_state[0]), np.array(solver_state[1], dtype=np.float64)
    jumped = bool(made_jump) if made_jump is not None else False
    floor = T1
    for _ in range(100):
        floor = float(np.nextafter(floor, -math.inf))
    tprev = T0
    n_steps = n_acc = n_rej = 0
    result = 0
    grid = []
    ts_index = 0
    while tprev < T1 and result == 0 and n_steps < max_steps:
        fsal = None if (first or jumped or not tab.fsal) else f0
        y1, yerr, ks, flast = rk_step(tab, prod, n_steps, tprev, tnext, direction, y, fsal,
                                      lincomb)
        yerr = np.where(np.isnan(yerr), np.inf, yerr)
        step_result = 0
        if controller == "pid":
            prev_dt = tnext - tprev
            y1c = y if np.isnan(y1).any() else y1
            e = yerr / (atol + np.maximum(np.abs(y), np.abs(y1c)) * rtol)
            scaled = _rms(e if norm_n is None else e[:norm_n])
            keep = scaled < 1
            if dtmin is not None:
                keep = keep or at_dtmin
            inv = 1 / scaled if scaled != 0 else math.inf
            c1 = (icoeff + pcoeff + dcoeff) / error_order
            c2 = -(pcoeff + 2 * dcoeff) / error_order
            c3 = dcoeff / error_order
            f1 = 1.0 if c1 == 0 else pow_ctl(inv, c1)
            f2 = 1.0 if c2 == 0 else pow_ctl(pinv, c2)
            f3 = 1.0 if c3 == 0 else pow_ctl(ppinv, c3)
            fmin = 1.0 if keep else factormin
            fmax = factormax if keep else safety
            factor = _jmin(_jmax(safety * f1 * f2 * f3, fmin), fmax)
            dt = prev_dt * factor
            if inv == 0 or math.isinf(inv):
                inv = 1.0
            if dtmax is not None:
                dt = _jmin(dt, dtmax)
            if dtmin is not None:
                if not force_dtmin and dt < dtmin:
                    step_result = 2
                if at_dtmin and factor == 1:
                    dt = dtmin
                at_dtmin = dt <= dtmin
                dt = _jmax(dt, dtmin)
            ntprev = tnext if keep else tprev
            ntnext = ntprev + dt
            pinv, ppinv = (inv, pinv) if keep else (pinv, ppinv)
        else:
            keep = True
            completed += 1
            tn = c_t0 + (c_t1 - c_t0) * (completed / num)
            if completed == num:
                tn = c_t1
            if not num >= 0:
                tn = tnext + c_t1
            ntprev, ntnext = tnext, tn
        ntprev = _jmin(ntprev, T1)
        if ntnext > floor:
            ntnext = T1 if keep else ntprev + 0.5 * (T1 - ntprev)
        n_steps += 1
        if keep:
            n_acc += 1
        else:
            n_rej += 1
        if result == 0:
            result = step_result
        if keep:
            grid.append((tprev, tnext))
            while ts_index < len(ts_n) and ts_n[ts_index] <= tnext:
                outs_t.append(ts_n[ts_index])
                outs_y.append(interpolate(tab, tprev, tnext, y, y1, ks, ts_n[ts_index]))
                ts_index += 1
            y = y1
            if tab.fsal:
                first, f0 = False, flast
            jumped = False
        tprev, tnext = ntprev, ntnext
    if T0 == T1:
        for _ in ts_n:
            outs_t.append(T0)
            outs_y.append(y.copy())
    if save_t1:
        outs_t.append(tprev)
        outs_y.append(y.copy())
    if tprev < T1 and result == 0:
        result = 1
    n_out = int(save_t0) + len(ts_n) + int(save_t1)
    ys = np.full((n_out, y.shape[0]), np.inf)
    tso = np.full(n_out, direction * np.inf)
    for i, (t, v) in enumerate(zip(outs_t, outs_y)):
        tso[i] = t * direction
        ys[i] = v
    if controller == "pid":
        cstate = (pinv, ppinv, bool(at_dtmin))
    else:
        cstate = (completed, num, c_t0, c_t1)
    return dict(ys=ys, ts=tso, stats=(n_steps, n_acc, n_rej), result=result, grid=grid,
                direction=direction, tprev=tprev, tnext=tnext,
                solver_state=(first, f0.copy()) if tab.fsal else None,
                controller_state=cstate, made_jump=jumped)


def backsolve(f_aug, d, n_q, solver, t0, dt0, ts, ys, ys_bar, *, save_t0, rtol, atol,
              max_steps=4096, seminorm=False, fused=True):
    """The continuous adjoint of one trajectory (_loop_backsolve_bwd, _adjoint.py:646-735):
    from the last saved time back to t0, solve the augmented state z = (y, a_y, a_q) under
    f_aug(t, z) = (f, vjp_y(-a_y), vjp_q(-a_y)) (AdjointTerm, _term.py:767-917) between
    consecutive saves, y restarting from the saved state ys[s] and a_y taking the save's
    cotangent ys_bar[s]; each solve starts from dt0 as given (the negated forward dt0 by
    default) with a fresh first stage.  seminorm: adjoint_rms_seminorm (error norm over
    (y, a_y) only).  Returns (a_y(t0) [+ ys_bar[0] when t0 was saved], a_q(t0), worst
    result code, backward steps per interval)."""
    a_y = np.zeros(d)
    a_q = np.zeros(n_q)
    s_end = 1 if save_t0 else 0
    worst = 0
    steps = []
    for s in range(len(ts) - 1, s_end - 1, -1):
        lo = ts[s - 1] if s > 0 else t0
        z0 = np.concatenate([np.asarray(ys[s], dtype=np.float64), a_y + ys_bar[s], a_q])
        r = solve(f_aug, solver, ts[s], lo, dt0, z0, rtol=rtol, atol=atol, save_t1=True,
                  max_steps=max_steps, norm_n=2 * d if seminorm else None, fused=fused)
        z1 = r["ys"][-1]
        a_y, a_q = z1[d:2 * d], z1[2 * d:]
        worst = worst or r["result"]
        steps.append(r["stats"])
    if save_t0 and len(ts) > 0:
        a_y = a_y + ys_bar[0]
    return a_y, a_q, worst, steps


def replay(f, solver, grid, y0, ts, direction=1, g_dw=None, save_t1=False, save_steps=False):
    """Re-run a fixed accepted-step grid (normalised (tprev, tnext) pairs) and return the
    list of saved states at ts, in the forward's row order (a step's ts rows, then its end
    state if save_steps), then y(t1) if save_t1.  Only arithmetic touches y, so y0
    may be a torch tensor with requires_grad (autograd reference) or complex (complex
    step).  The step index passed to g_dw is the accepted-step index, which equals the
    attempted-step index for ConstantStepSize."""
    tab = TABLEAUX[solver]

    def prod(t, z, control, n):
        fz = f(t, z)
        k = control * fz
        if g_dw is not None:
            k = k + g_dw(n, t, z)
        return k, fz

    ts_n = [t * direction for t in ts]
    y = y0
    fsal = None
    out = []
    ts_index = 0
    for n, (a, b) in enumerate(grid):
        y1, _, ks, flast = rk_step(tab, prod, n, a, b, direction, y, fsal)
        while ts_index < len(ts_n) and ts_n[ts_index] <= b:
            out.append(interpolate(tab, a, b, y, y1, ks, ts_n[ts_index]))
            ts_index += 1
        y = y1
        if save_steps and (n + 1) % int(save_steps) == 0:
            out.append(y)
        if tab.fsal:
            fsal = flast
    if save_t1:
        out.append(y)
    return out
