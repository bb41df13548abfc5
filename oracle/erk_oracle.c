/*
 * erk_oracle.c -- FP64 CPU restatement of diffrax's explicit Runge-Kutta path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (diffrax_amd/) links, loads or
 * calls this file.  It is imported by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, always as the checker, never as the thing measured
 * or shipped.
 *
 * What it restates (file:line in the reference diffrax 0.7.2 tree):
 *   - integration loop, accept/revert, counters, saves ......... _integrate.py:302-491
 *   - _clip_to_end, t1_clip_floor (100 x prevbefore) ........... _integrate.py:278-284, 319-322
 *   - t0 == t1 repeat save, _save_t1, max_steps_reached ........ _integrate.py:823-883
 *   - time normalisation by direction, WrapTerm ................ _integrate.py:1156-1176, _term.py:734-764
 *   - output buffers preset to (direction*)inf ................. _integrate.py:1272-1305
 *   - ERK step, FSAL/SSAL, k = control*f, z = y0 + sum a k ..... _solver/runge_kutta.py:643-701, 847-1052, 1170-1203
 *   - stage times t_i = (c_i == 1 ? t1 : t0 + c_i dt) .......... _solver/runge_kutta.py:1023
 *   - Tsit5 / Dopri5 / Heun tableaux ........................... _solver/tsit5.py:18-100, dopri5.py:11-33, heun.py:12-17
 *   - Tsit5 interpolation polynomials .......................... _solver/tsit5.py:116-156
 *   - quartic (Dopri5) / cubic Hermite (Heun) interpolation .... _local_interpolation.py:50-139, dopri5.py:36-47
 *   - linear_rescale ........................................... _misc.py:71-82
 *   - PIDController.init / adapt_step_size ...................... _step_size_controller/pid.py:316-567
 *   - _select_initial_step (dt0 = None) ........................ _step_size_controller/pid.py:33-81
 *   - ConstantStepSize.init / adapt_step_size ................... _step_size_controller/constant.py:30-104
 *   - SDE increment k = f*dt + g.dW (MultiTerm/ControlTerm) .... _term.py:711-722, 420-427
 *   - ClipStepSizeController (step_ts, jump_ts) ................ _step_size_controller/clip.py:120-428
 *   - affine Event conditions, root find on the interpolant .... _event.py:13-118, _integrate.py:542-633, 686-821
 *   - BacksolveAdjoint with AdjointTerm ........................ _adjoint.py:539-873, _term.py:767-917
 *
 * Each part is pinned by tests/test_oracle_*.py against closed forms, exact rational
 * arithmetic, scipy, finite differences or the reference's own test recipes.
 *
 * Parity notes.  Operation order follows the reference where it is visible in the
 * Python source (k = control*f first, then y0 + (sum_j a_ij k_j) summed in stage order;
 * clip = min(max(x, lo), hi) with NaN propagation as jnp.clip).  Where the reference
 * delegates to an absent third-party routine the restatement states its choice:
 *   - optimistix.rms_norm is restated as sqrt(sum x_i^2) / sqrt(n) (|x| when n == 1);
 *     its exact rounding is not pinned by any reference test (parity unpinned).
 *   - jnp.tensordot (XLA, precision=HIGHEST) is restated as a fused multiply-add chain in
 *     ascending stage order: the first nonzero product, then acc = fma(c_m, k_m, acc)
 *     (one rounding per term, as a fused dot product evaluates it); y0 is added after.
 *     Stage sums, the b and b_err sums, Dopri5's midpoint weights and the Tsit5
 *     interpolant's b(theta) . k (tsit5.py:153-156) all use it, and the
 *     discrete adjoint accumulates its transposes (k_bar += b_m y1_bar, += a_il z_bar) the
 *     same way.  The
 *     exact rounding of XLA's dot is not pinned by any reference test (parity unpinned).
 * Build with -ffp-contract=off, so no other a*b+c is fused.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "erk_oracle.h"

/* ------------------------------------------------------------------------- */
/* Tableaux (published coefficients; same float64 values as the reference).   */
/* ------------------------------------------------------------------------- */

#define MAXS 7

typedef struct {
    int S;                 /* stages */
    int fsal;              /* first-same-as-last */
    int ssal;              /* solution-same-as-last */
    int order;             /* solver order (error_order for ODEs) */
    double a[MAXS][MAXS];  /* strictly lower triangular, zero padded */
    double b[MAXS];
    double berr[MAXS];
    double c[MAXS];        /* c[0] = 0 */
    double cmid[MAXS];     /* Dopri5 only */
} tableau_t;

static tableau_t TAB[3];
static int tab_ready = 0;

static void init_tableaux(void) {
    if (tab_ready) return;
    memset(TAB, 0, sizeof(TAB));
    /* ---- Tsit5 (Tsitouras 2011) ------------------------------------------------ */
    tableau_t *t = &TAB[ORC_TSIT5];
    t->S = 7; t->fsal = 1; t->ssal = 1; t->order = 5;
    t->a[1][0] = 161.0 / 1000.0;
    t->a[2][0] = -0.8480655492356988544426874250230774675121177393430391537369234245294192976164141156943e-2;
    t->a[2][1] = 0.3354806554923569885444268742502307746751211773934303915373692342452941929761641411569;
    t->a[3][0] = 2.897153057105493432130432594192938764924887287701866490314866693455023795137503079289;
    t->a[3][1] = -6.359448489975074843148159912383825625952700647415626703305928850207288721235210244366;
    t->a[3][2] = 4.362295432869581411017727318190886861027813359713760212991062156752264926097707165077;
    t->a[4][0] = 5.325864828439256604428877920840511317836476253097040101202360397727981648835607691791;
    t->a[4][1] = -11.74888356406282787774717033978577296188744178259862899288666928009020615663593781589;
    t->a[4][2] = 7.495539342889836208304604784564358155658679161518186721010132816213648793440552049753;
    t->a[4][3] = -0.9249506636175524925650207933207191611349983406029535244034750452930469056411389539635e-1;
    t->a[5][0] = 5.861455442946420028659251486982647890394337666164814434818157239052507339770711679748;
    t->a[5][1] = -12.92096931784710929170611868178335939541780751955743459166312250439928519268343184452;
    t->a[5][2] = 8.159367898576158643180400794539253485181918321135053305748355423955009222648673734986;
    t->a[5][3] = -0.7158497328140099722453054252582973869127213147363544882721139659546372402303777878835e-1;
    t->a[5][4] = -0.2826905039406838290900305721271224146717633626879770007617876201276764571291579142206e-1;
    t->a[6][0] = 0.9646076681806522951816731316512876333711995238157997181903319145764851595234062815396e-1;
    t->a[6][1] = 1.0 / 100.0;
    t->a[6][2] = 0.4798896504144995747752495322905965199130404621990332488332634944254542060153074523509;
    t->a[6][3] = 1.379008574103741893192274821856872770756462643091360525934940067397245698027561293331;
    t->a[6][4] = -3.290069515436080679901047585711363850115683290894936158531296799594813811049925401677;
    t->a[6][5] = 2.324710524099773982415355918398765796109060233222962411944060046314465391054716027841;
    for (int j = 0; j < 6; ++j) t->b[j] = t->a[6][j];
    t->b[6] = 0.0;
    {
        /* b_error = b_sol - b_hat, each difference formed in float64 as in tsit5.py:75-98 */
        volatile double bs0 = t->a[6][0], bs1 = t->a[6][1], bs2 = t->a[6][2];
        volatile double bs3 = t->a[6][3], bs4 = t->a[6][4], bs5 = t->a[6][5];
        t->berr[0] = bs0 - 0.9468075576583945807478876255758922856117527357724631226139574065785592789071067303271e-1;
        t->berr[1] = bs1 - 0.9183565540343253096776363936645313759813746240984095238905939532922955247253608687270e-2;
        t->berr[2] = bs2 - 0.4877705284247615707855642599631228241516691959761363774365216240304071651579571959813;
        t->berr[3] = bs3 - 1.234297566930478985655109673884237654035539930748192848315425833500484878378061439761;
        t->berr[4] = bs4 + 2.707712349983525454881109975059321670689605166938197378763992255714444407154902012702;
        t->berr[5] = bs5 - 1.866628418170587035753719399566211498666255505244122593996591602841258328965767580089;
        t->berr[6] = -1.0 / 66.0;
    }
    t->c[1] = 161.0 / 1000.0; t->c[2] = 327.0 / 1000.0; t->c[3] = 9.0 / 10.0;
    t->c[4] = 0.9800255409045096857298102862870245954942137979563024768854764293221195950761080302604;
    t->c[5] = 1.0; t->c[6] = 1.0;

    /* ---- Dopri5 (Dormand-Prince / Shampine) ------------------------------------ */
    t = &TAB[ORC_DOPRI5];
    t->S = 7; t->fsal = 1; t->ssal = 1; t->order = 5;
    t->a[1][0] = 1.0 / 5.0;
    t->a[2][0] = 3.0 / 40.0; t->a[2][1] = 9.0 / 40.0;
    t->a[3][0] = 44.0 / 45.0; t->a[3][1] = -56.0 / 15.0; t->a[3][2] = 32.0 / 9.0;
    t->a[4][0] = 19372.0 / 6561.0; t->a[4][1] = -25360.0 / 2187.0; t->a[4][2] = 64448.0 / 6561.0;
    t->a[4][3] = -212.0 / 729.0;
    t->a[5][0] = 9017.0 / 3168.0; t->a[5][1] = -355.0 / 33.0; t->a[5][2] = 46732.0 / 5247.0;
    t->a[5][3] = 49.0 / 176.0; t->a[5][4] = -5103.0 / 18656.0;
    t->a[6][0] = 35.0 / 384.0; t->a[6][1] = 0.0; t->a[6][2] = 500.0 / 1113.0;
    t->a[6][3] = 125.0 / 192.0; t->a[6][4] = -2187.0 / 6784.0; t->a[6][5] = 11.0 / 84.0;
    for (int j = 0; j < 6; ++j) t->b[j] = t->a[6][j];
    t->b[6] = 0.0;
    {
        volatile double x;
        x = 35.0 / 384.0;   t->berr[0] = x - 1951.0 / 21600.0;
        t->berr[1] = 0.0;
        x = 500.0 / 1113.0; t->berr[2] = x - 22642.0 / 50085.0;
        x = 125.0 / 192.0;  t->berr[3] = x - 451.0 / 720.0;
        x = -2187.0 / 6784.0; t->berr[4] = x - (-12231.0 / 42400.0);
        x = 11.0 / 84.0;    t->berr[5] = x - 649.0 / 6300.0;
        t->berr[6] = -1.0 / 60.0;
    }
    t->c[1] = 1.0 / 5.0; t->c[2] = 3.0 / 10.0; t->c[3] = 4.0 / 5.0; t->c[4] = 8.0 / 9.0;
    t->c[5] = 1.0; t->c[6] = 1.0;
    {
        volatile double x;
        x = 6025192743.0 / 30085553152.0;     t->cmid[0] = x / 2.0;
        t->cmid[1] = 0.0;
        x = 51252292925.0 / 65400821598.0;    t->cmid[2] = x / 2.0;
        x = -2691868925.0 / 45128329728.0;    t->cmid[3] = x / 2.0;
        x = 187940372067.0 / 1594534317056.0; t->cmid[4] = x / 2.0;
        x = -1776094331.0 / 19743644256.0;    t->cmid[5] = x / 2.0;
        x = 11237099.0 / 235043384.0;         t->cmid[6] = x / 2.0;
    }

    /* ---- Heun ----------------------------------------------------------------- */
    t = &TAB[ORC_HEUN];
    t->S = 2; t->fsal = 0; t->ssal = 0; t->order = 2;
    t->a[1][0] = 1.0;
    t->b[0] = 0.5; t->b[1] = 0.5;
    t->berr[0] = 0.5; t->berr[1] = -0.5;
    t->c[1] = 1.0;
    tab_ready = 1;
}

int orc_tableau(int solver, double *a, double *b, double *berr, double *c, int *info) {
    init_tableaux();
    if (solver < 0 || solver > 2) return -1;
    const tableau_t *t = &TAB[solver];
    for (int i = 0; i < MAXS; ++i) {
        for (int j = 0; j < MAXS; ++j) a[i * MAXS + j] = t->a[i][j];
        b[i] = t->b[i]; berr[i] = t->berr[i]; c[i] = t->c[i];
    }
    info[0] = t->S; info[1] = t->fsal; info[2] = t->ssal; info[3] = t->order;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* jnp-style helpers                                                           */
/* ------------------------------------------------------------------------- */

/* jnp.maximum / jnp.minimum propagate NaN. */
static inline double jmax(double a, double b) { return (isnan(a) || isnan(b)) ? NAN : (a > b ? a : b); }
static inline double jmin(double a, double b) { return (isnan(a) || isnan(b)) ? NAN : (a < b ? a : b); }
static inline double jclip(double x, double lo, double hi) { return jmin(jmax(x, lo), hi); }

/* eqxi.prevbefore: next representable float64 towards -inf. */
static double prevbefore(double x) { return nextafter(x, -INFINITY); }
/* eqxi.nextafter: next representable float64 towards +inf. */
static double nextup(double x) { return nextafter(x, INFINITY); }

/* ---- ClipStepSizeController (_step_size_controller/clip.py) -----------------------
 * The controller works on the times scaled by the direction and sorted (wrap,
 * clip.py:231-241): element i of that array is ts[i] for dir > 0 and -ts[n-1-i] for
 * dir < 0 (ts ascending). */
typedef struct { const double *ts; int n, dir; } clipts_t;
static double clip_at(const clipts_t *c, int i) {
    return c->dir > 0 ? c->ts[i] : -c->ts[c->n - 1 - i];
}
/* _get_t (clip.py:57-65): ts[i], or inf past the end. */
static double clip_get_t(const clipts_t *c, int i) {
    return (c->n == 0 || i >= c->n) ? INFINITY : clip_at(c, i);
}
/* _clip_t (clip.py:68-79): min(ts[i] (prevbefore it for jumps), t). */
static double clip_t(double t, int i, const clipts_t *c, int prev) {
    double u = clip_get_t(c, i);
    if (prev) u = prevbefore(u);
    return jmin(u, t);
}
static int clip_any_eq(const clipts_t *c, double x) {
    for (int i = 0; i < c->n; ++i) if (clip_at(c, i) == x) return 1;
    return 0;
}
/* _bump_next_t0 (clip.py:82-101): a step that ended prevbefore a listed time restarts
 * two ULPs later, past the time; one that ended on it restarts one ULP later. */
static double clip_bump(double t, const clipts_t *c, int *made_jump) {
    const double na = nextup(t);
    const int mj1 = clip_any_eq(c, na), mj2 = clip_any_eq(c, t);
    double r = mj1 ? nextup(na) : t;
    r = mj2 ? na : r;
    *made_jump = mj1 | mj2;
    return r;
}
/* _find_idx_with_hint (clip.py:104-118): first index with ts[i] > t. */
static int clip_find(double t, const clipts_t *c, int hint);
/* Test hooks: the two helpers above on an ascending array (dir = +1 / -1 as wrap). */
int orc_clip_find(const double *ts, int n, int dir, double t, int hint) {
    clipts_t c = {ts, n, dir};
    return clip_find(t, &c, hint);
}
static double clip_bump(double t, const clipts_t *c, int *made_jump);
double orc_clip_bump(const double *ts, int n, int dir, double t, int *made_jump) {
    clipts_t c = {ts, n, dir};
    return clip_bump(t, &c, made_jump);
}
static int clip_find(double t, const clipts_t *c, int hint) {
    int i = hint;
    while (i < c->n && clip_at(c, i) <= t) ++i;
    while (i > 0 && clip_at(c, i - 1) > t) --i;
    return i;
}

/* Controller powers, correctly rounded.  The PID factor is (1/scaled_error)^0.2 and the
 * automatic initial step raises to 1/error_order (pid.py:79, 511-517).  The oracle uses a
 * correctly rounded x^0.2 (fifth root by Newton, double-double residual, and the
 * x^(0.2 - 1/5) = 1 + (0.2 - 1/5) ln x factor), checked in the tests against 60-digit
 * decimal arithmetic and against the C library's pow; x^0.5 is sqrt and x^1 is x. */
static double bits2d(uint64_t b) { double r; memcpy(&r, &b, 8); return r; }
static uint64_t d2bits(double x) { uint64_t b; memcpy(&b, &x, 8); return b; }

static double cr_pow_fifth(double x) {
    if (!(x > 0.0)) return (x == 0.0) ? 0.0 : NAN;
    if (isinf(x)) return x;
    int adj = 0;
    if (x < 2.2250738585072014e-308) { x = x * 1.2676506002282294e30; adj = -20; }
    uint64_t bx = d2bits(x);
    int E = (int)((bx >> 52) & 0x7ff) - 1023;
    int k = (E >= 0) ? E / 5 : -((4 - E) / 5);
    int j = E - 5 * k;
    double f = bits2d((bx & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    double m = f * (double)(1 << j);
    static const double tab[5] = {1.0, 1.148698354997035, 1.3195079107728942,
                                  1.515716566510398, 1.7411011265922482};
    double u = f - 1.0;
    double r = tab[j] * (1.0 + u * (0.19425 - 0.04555 * u));
    for (int it = 0; it < 3; ++it) {
        double r2 = r * r, r4 = r2 * r2;
        r = (4.0 * r + m / r4) * 0.2;
    }
    double h2 = r * r, l2 = fma(r, r, -h2);
    double h4 = h2 * h2, l4 = fma(h2, h2, -h4) + 2.0 * h2 * l2;
    double h5 = h4 * r, l5 = fma(h4, r, -h5) + l4 * r;
    double res = (m - h5) - l5;
    double delta = res / (5.0 * h4);
    int hi = f > 1.4142135623730951;
    double fl = hi ? f * 0.5 : f;
    double s = (fl - 1.0) / (fl + 1.0), s2 = s * s;
    double lnf = 2.0 * s * (1.0 + s2 * (1.0 / 3.0 + s2 * (0.2 + s2 * (1.0 / 7.0 + s2 * (1.0 / 9.0)))));
    double lnx = lnf + (double)(E + hi + 5 * adj) * 0.6931471805599453;
    double corr = r * (1.1102230246251565e-17 * lnx);
    double out = r + (delta + corr);
    return out * bits2d((uint64_t)(1023 + k + adj) << 52);
}

static double pow_ctl(double x, double c) {
    if (c == 0.2) return cr_pow_fifth(x);
    if (c == 1.0) return x;
    if (c == 0.5) { if (x == 0.0) return 0.0; return x < 0.0 ? NAN : sqrt(x); }
    return pow(x, c);
}

double orc_pow(double x, double c) { return pow_ctl(x, c); }

/* optimistix.rms_norm restated (see header). */
static double rms_norm(const double *x, int n) {
    if (n == 1) return fabs(x[0]);
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += x[i] * x[i];
    return sqrt(s) / sqrt((double)n);
}

/* ------------------------------------------------------------------------- */
/* Vector fields                                                               */
/* ------------------------------------------------------------------------- */

typedef struct {
    const orc_problem *p;
    const double *prm;     /* parameters of this instance */
    const double *dW;      /* SDE: increments of this path, indexed [step * dW_step_stride + j] */
    long dW_step_stride;
    double *scratch;       /* MLP hidden buffers, 2*H */
} vfctx_t;

static inline double softplus(double x) {
    /* jax.nn.softplus(x) = logaddexp(x, 0) = max(x, 0) + log1p(exp(-|x|)) */
    return (x > 0.0 ? x : 0.0) + log1p(exp(-fabs(x)));
}

/* f(t, y) in the caller's time (after the WrapTerm direction multiply). */
static void vf_eval(const vfctx_t *cx, double t, const double *y, double *out);
/* w^T df/dy and w^T df/dq of the registered small-system fields (the same products, in
 * the same order, as the device's VF::vjp). */
static void vf_vjp(int vf, int d, const double *q, const double *y, const double *w,
                   double *gy, double *gq) {
    if (vf == ORC_VF_LV) {
        const double x = y[0], v = y[1];
        gy[0] = w[0] * (q[0] - q[1] * v) + w[1] * (q[3] * v);
        gy[1] = w[0] * (-(q[1] * x)) + w[1] * (q[3] * x - q[2]);
        gq[0] = w[0] * x;
        gq[1] = -(w[0] * (x * v));
        gq[2] = -(w[1] * v);
        gq[3] = w[1] * (x * v);
    } else if (vf == ORC_VF_VDP) {
        const double x = y[0], v = y[1];
        gy[0] = w[1] * (-2.0 * q[0] * x * v - 1.0);
        gy[1] = w[0] + w[1] * (q[0] * (1.0 - x * x));
        gq[0] = w[1] * ((1.0 - x * x) * v);
    } else {   /* linear */
        for (int j = 0; j < d; ++j) {
            double s = 0.0;
            for (int i = 0; i < d; ++i) s = s + q[i * d + j] * w[i];
            gy[j] = s;
        }
        for (int i = 0; i < d; ++i)
            for (int j = 0; j < d; ++j) gq[i * d + j] = w[i] * y[j];
    }
}

static void vf_eval(const vfctx_t *cx, double t, const double *y, double *out) {
    const orc_problem *p = cx->p;
    const double *q = cx->prm;
    int d = p->d;
    if (p->adj_base_d > 0) {
        /* AdjointTerm.vf_prod (_term.py:893-917): (f(y), -a^T df/dy, -a^T df/dq) */
        const int db = p->adj_base_d;
        orc_problem base = *p;
        base.d = db;
        base.adj_base_d = 0;
        vfctx_t bcx = *cx;
        bcx.p = &base;
        double w[ORC_MAXD], gy[ORC_MAXD], gq[ORC_MAXD];
        for (int j = 0; j < db; ++j) w[j] = -y[db + j];
        vf_eval(&bcx, t, y, out);
        vf_vjp(p->vf, db, q, y, w, gy, gq);
        const long P = d - 2 * db;
        for (int j = 0; j < db; ++j) out[db + j] = gy[j];
        for (long m = 0; m < P; ++m) out[2 * db + m] = gq[m];
        return;
    }
    switch (p->vf) {
    case ORC_VF_LV: {
        /* dx = alpha x - beta x y ; dy = delta x y - gamma y ; q = [alpha, beta, gamma, delta] */
        double x = y[0], v = y[1];
        out[0] = q[0] * x - q[1] * x * v;
        out[1] = q[3] * x * v - q[2] * v;
        break;
    }
    case ORC_VF_VDP: {
        /* y'' = mu (1 - y^2) y' - y as a first-order system ; q = [mu] */
        double x = y[0], v = y[1];
        out[0] = v;
        out[1] = q[0] * (1.0 - x * x) * v - x;
        break;
    }
    case ORC_VF_LINEAR: {
        /* f = A y, A row-major d x d */
        for (int i = 0; i < d; ++i) {
            double s = 0.0;
            for (int j = 0; j < d; ++j) s += q[i * d + j] * y[j];
            out[i] = s;
        }
        break;
    }
    case ORC_VF_MLP: {
        /* f = tanh(W3 softplus(W2 softplus(W1 y + b1) + b2) + b3) */
        int H = p->hidden;
        const double *W1 = q, *b1 = W1 + (long)H * d, *W2 = b1 + H, *b2 = W2 + (long)H * H;
        const double *W3 = b2 + H, *b3 = W3 + (long)d * H;
        double *h1 = cx->scratch, *h2 = cx->scratch + H;
        for (int i = 0; i < H; ++i) {
            double s = 0.0;
            for (int j = 0; j < d; ++j) s += W1[(long)i * d + j] * y[j];
            h1[i] = softplus(s + b1[i]);
        }
        for (int i = 0; i < H; ++i) {
            double s = 0.0;
            for (int j = 0; j < H; ++j) s += W2[(long)i * H + j] * h1[j];
            h2[i] = softplus(s + b2[i]);
        }
        for (int i = 0; i < d; ++i) {
            double s = 0.0;
            for (int j = 0; j < H; ++j) s += W3[(long)i * H + j] * h2[j];
            out[i] = tanh(s + b3[i]);
        }
        break;
    }
    case ORC_VF_OU: {
        /* drift theta (mu - y) ; q = [theta(d), mu(d), sigma(d)] */
        for (int i = 0; i < d; ++i) out[i] = q[i] * (q[d + i] - y[i]);
        break;
    }
    default:
        for (int i = 0; i < d; ++i) out[i] = NAN;
    }
    (void)t;
}

/* ------------------------------------------------------------------------- */
/* One explicit RK step (runge_kutta.py:446-1203, ERK branch).                 */
/* ------------------------------------------------------------------------- */

typedef struct {
    double k[MAXS][ORC_MAXD];
    double y1[ORC_MAXD];
    double yerr[ORC_MAXD];
    double flast[ORC_MAXD];
} stepout_t;

/* ts0, ts1: normalised times of the step; dir: +-1; fsal_first: evaluate stage 0.
 * f0: stored FSAL vector field value (used when !fsal_first).
 * step_index: SDE increment index. */
static void erk_step(const tableau_t *tb, const vfctx_t *cx, int dir, double ts0, double ts1,
                     const double *y0, int fsal_first, const double *f0, long step_index,
                     stepout_t *o) {
    const orc_problem *p = cx->p;
    int d = p->d, S = tb->S;
    double dt = ts1 - ts0;
    /* WrapTerm.contr: direction * (t1 - t0) ; ODETerm.prod: control * f */
    double control = (double)dir * dt;
    const double *dWn = NULL;
    if (p->vf == ORC_VF_OU) dWn = cx->dW + step_index * cx->dW_step_stride;
    double z[ORC_MAXD], f[ORC_MAXD];
    int start = 0;
    if (tb->fsal && !fsal_first) {
        for (int j = 0; j < d; ++j) o->k[0][j] = control * f0[j];
        start = 1;
    }
    for (int i = start; i < S; ++i) {
        for (int j = 0; j < d; ++j) {
            double acc = 0.0;
            int first = 1;
            for (int m = 0; m < i; ++m) {
                if (tb->a[i][m] == 0.0) continue;
                if (first) { acc = tb->a[i][m] * o->k[m][j]; first = 0; }
                else acc = fma(tb->a[i][m], o->k[m][j], acc);
            }
            z[j] = y0[j] + acc;
        }
        double ti = (tb->c[i] == 1.0) ? ts1 : ts0 + tb->c[i] * dt;
        vf_eval(cx, ti * (double)dir, z, f);
        if (p->vf == ORC_VF_OU) {
            /* MultiTerm(ODETerm(drift), ControlTerm(diag sigma, path)):
             * k = (dt * f) + (sigma * dW)   (_term.py:711-722, 423-427) */
            const double *sig = cx->prm + 2 * d;
            for (int j = 0; j < d; ++j) o->k[i][j] = control * f[j] + sig[j] * dWn[j];
        } else {
            for (int j = 0; j < d; ++j) o->k[i][j] = control * f[j];
        }
        if (i == S - 1) {
            for (int j = 0; j < d; ++j) o->flast[j] = f[j];
            if (tb->ssal) for (int j = 0; j < d; ++j) o->y1[j] = z[j];
        }
    }
    if (!tb->ssal) {
        for (int j = 0; j < d; ++j) {
            double acc = 0.0; int first = 1;
            for (int m = 0; m < S; ++m) {
                if (tb->b[m] == 0.0) continue;
                if (first) { acc = tb->b[m] * o->k[m][j]; first = 0; }
                else acc = fma(tb->b[m], o->k[m][j], acc);
            }
            o->y1[j] = y0[j] + acc;
        }
    }
    for (int j = 0; j < d; ++j) {
        double acc = 0.0; int first = 1;
        for (int m = 0; m < S; ++m) {
            if (tb->berr[m] == 0.0) continue;
            if (first) { acc = tb->berr[m] * o->k[m][j]; first = 0; }
            else acc = fma(tb->berr[m], o->k[m][j], acc);
        }
        o->yerr[j] = acc;
    }
}

/* ------------------------------------------------------------------------- */
/* Local interpolation (dense output inside one step).                         */
/* ------------------------------------------------------------------------- */

static double linear_rescale(double t0, double t, double t1) {
    int cond = (t0 == t1);
    double num = cond ? 0.0 : t - t0;
    double den = cond ? 1.0 : t1 - t0;
    return num / den;
}

/* Tsit5 dense-output weights b_i(theta) (tsit5.py:116-156), in its factorised form. */
static void tsit5_bb(double th, double *bb) {
    double t2 = th * th;
    bb[0] = -1.0530884977290216 * th * (th - 1.3299890189751412) *
            (t2 - 1.4364028541716351 * th + 0.7139816917074209);
    bb[1] = 0.1017 * t2 * (t2 - 2.1966568338249754 * th + 1.2949852507374631);
    bb[2] = 2.490627285651252793 * t2 * (t2 - 2.38535645472061657 * th + 1.57803468208092486);
    bb[3] = -16.54810288924490272 * (th - 1.21712927295533244) * (th - 0.61620406037800089) * t2;
    bb[4] = 47.37952196281928122 * (th - 1.203071208372362603) * (th - 0.658047292653547382) * t2;
    bb[5] = -34.87065786149660974 * (th - 1.2) * (th - 0.666666666666666667) * t2;
    bb[6] = 2.5 * (th - 1.0) * (th - 0.6) * t2;
}

static void interp_eval(int solver, const tableau_t *tb, int d, double ts0, double ts1,
                        const double *y0, const double *y1, const stepout_t *o, double t,
                        double *out) {
    double th = linear_rescale(ts0, t, ts1);
    if (solver == ORC_TSIT5) {
        /* _Tsit5Interpolation.evaluate (tsit5.py:116-156) */
        double bb[7];
        tsit5_bb(th, bb);
        for (int j = 0; j < d; ++j) {
            double acc = bb[0] * o->k[0][j];
            for (int m = 1; m < 7; ++m) acc = fma(bb[m], o->k[m][j], acc);
            out[j] = y0[j] + acc;
        }
    } else if (solver == ORC_DOPRI5) {
        /* FourthOrderPolynomialInterpolation (_local_interpolation.py:95-139) */
        for (int j = 0; j < d; ++j) {
            double acc = 0.0; int first = 1;
            for (int m = 0; m < 7; ++m) {
                if (tb->cmid[m] == 0.0) continue;
                if (first) { acc = tb->cmid[m] * o->k[m][j]; first = 0; }
                else acc = fma(tb->cmid[m], o->k[m][j], acc);
            }
            double ymid = y0[j] + acc;
            double f0 = o->k[0][j], f1 = o->k[6][j];
            double a = 2.0 * (f1 - f0) - 8.0 * (y1[j] + y0[j]) + 16.0 * ymid;
            double b = 5.0 * f0 - 3.0 * f1 + 18.0 * y0[j] + 14.0 * y1[j] - 32.0 * ymid;
            double c = f1 - 4.0 * f0 - 11.0 * y0[j] - 5.0 * y1[j] + 16.0 * ymid;
            out[j] = (((a * th + b) * th + c) * th + f0) * th + y0[j];
        }
    } else {
        /* ThirdOrderHermitePolynomialInterpolation.from_k (_local_interpolation.py:50-92) */
        for (int j = 0; j < d; ++j) {
            double k0 = o->k[0][j], k1 = o->k[tb->S - 1][j];
            double a = k0 + k1 + 2.0 * y0[j] - 2.0 * y1[j];
            double b = -2.0 * k0 - k1 - 3.0 * y0[j] + 3.0 * y1[j];
            out[j] = ((a * th + b) * th + k0) * th + y0[j];
        }
    }
}

/* ---- Event (_event.py:13-118; loop checks _integrate.py:542-633, root find and unsave
 * 686-821) with affine conditions, see erk_oracle.h ---- */
static double ev_value(const orc_problem *p, int i, double t, const double *y) {
    const double *c = p->event_coeffs + (long)i * (p->d + 2);
    double v = c[0] * t;
    for (int j = 0; j < p->d; ++j) v = v + c[1 + j] * y[j];
    return v + c[1 + p->d];
}
/* jnp.sign */
static double ev_sign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : (x == 0.0 ? 0.0 : x)); }
/* The event test of one condition between its previous value and the new one. */
static int ev_triggered(int flags, double old_v, double new_v) {
    if (flags & 1) return new_v > 0.0;
    const int dirn = (flags >> 1) & 3;
    const double so = ev_sign(old_v), sn = ev_sign(new_v);
    if (dirn == 1) return (so <= 0.0) && (sn > 0.0);
    if (dirn == 2) return (so > 0.0) && (sn <= 0.0);
    return so != sn;
}
/* Time of the event on [a, b] (normalised): the end of the step for a boolean condition,
 * else bisection of the condition along the step's interpolant down to adjacent floats,
 * returning the end point with the smaller |value|. */
static double ev_root(const orc_problem *p, const tableau_t *tb, int i, double a, double b,
                      const double *y0, const double *y1, const stepout_t *o) {
    if (p->event_flags[i] & 1) return b;
    double ya[ORC_MAXD], yb[ORC_MAXD], ym[ORC_MAXD];
    interp_eval(p->solver, tb, p->d, a, b, y0, y1, o, a, ya);
    interp_eval(p->solver, tb, p->d, a, b, y0, y1, o, b, yb);
    double ga = ev_value(p, i, a, ya), gb = ev_value(p, i, b, yb);
    if (!(ev_sign(ga) != ev_sign(gb)) || isnan(ga) || isnan(gb)) return b;
    const double lo0 = a, hi0 = b;
    for (int it = 0; it < 2100; ++it) {
        const double m = a + 0.5 * (b - a);
        if (!(m > a && m < b)) break;
        interp_eval(p->solver, tb, p->d, lo0, hi0, y0, y1, o, m, ym);
        const double gm = ev_value(p, i, m, ym);
        if (gm == 0.0) return m;
        if (ev_sign(gm) == ev_sign(ga)) { a = m; ga = gm; } else { b = m; gb = gm; }
    }
    return (fabs(ga) <= fabs(gb)) ? a : b;
}

/* ------------------------------------------------------------------------- */
/* One trajectory                                                              */
/* ------------------------------------------------------------------------- */

static int is_none(double x) { return isnan(x); }

/* _select_initial_step (pid.py:33-81); times normalised, vf at t*dir. */
static double select_initial_step(const vfctx_t *cx, int dir, double t0, const double *y0,
                                  double error_order) {
    const orc_problem *p = cx->p;
    int d = p->d;
    if (p->vf == ORC_VF_OU) return 0.01;
    double scale[ORC_MAXD], tmp[ORC_MAXD], f0[ORC_MAXD], f1[ORC_MAXD], y1[ORC_MAXD] = {0};
    for (int j = 0; j < d; ++j) scale[j] = p->atol + fabs(y0[j]) * p->rtol;
    vf_eval(cx, t0 * dir, y0, f0);
    for (int j = 0; j < d; ++j) tmp[j] = y0[j] / scale[j];
    double d0 = rms_norm(tmp, d);
    for (int j = 0; j < d; ++j) tmp[j] = f0[j] / scale[j];
    double d1 = rms_norm(tmp, d);
    int cond = (d0 < 1e-5) || (d1 < 1e-5);
    double _d1 = cond ? 1.0 : d1;
    double h0 = cond ? 1e-6 : 0.01 * (d0 / _d1);
    double t1 = t0 + h0;
    for (int j = 0; j < d; ++j) y1[j] = y0[j] + h0 * f0[j];
    vf_eval(cx, t1 * dir, y1, f1);
    for (int j = 0; j < d; ++j) tmp[j] = (f1[j] - f0[j]) / scale[j];
    double d2 = rms_norm(tmp, d) / h0;
    double max_d = jmax(d1, d2);
    double h1 = (max_d <= 1e-15) ? jmax(1e-6, h0 * 1e-3) : pow_ctl(0.01 / max_d, 1.0 / error_order);
    return jmin(100.0 * h0, h1);
}

static void solve_one(const orc_problem *p_in, long bi, const double *y0_in, double t1_in,
                      const double *ts_in, const double *prm, const double *dW,
                      long dW_step_stride, double *ys, double *tso, int32_t *stats,
                      int32_t *result_out, double *grid, double *scratch) {
    init_tableaux();
    /* this instance's tolerances (jax.vmap over PIDController(rtol, atol)) */
    orc_problem pl = *p_in;
    if (p_in->rtol_b) pl.rtol = p_in->rtol_b[bi];
    if (p_in->atol_b) pl.atol = p_in->atol_b[bi];
    const orc_problem *p = &pl;
    const tableau_t *tb = &TAB[p->solver];
    int d = p->d;
    int n_ts = p->n_ts;
    int n_out = orc_n_out(p);
    vfctx_t cx = {p, prm, dW, dW_step_stride, scratch};

    /* ---- diffeqsolve set-up (_integrate.py:1156-1305) ---- */
    double t0r = p->t0, t1r = t1_in;
    int dir = (t0r < t1r) ? 1 : -1;
    double T0 = t0r * dir, T1 = t1r * dir;
    double DT0 = is_none(p->dt0) ? NAN : p->dt0 * dir;
    for (int s = 0; s < n_out; ++s) {
        tso[s] = (double)dir * INFINITY;
        for (int j = 0; j < d; ++j) ys[(long)s * d + j] = INFINITY;
    }
    double error_order = is_none(p->error_order) ? (double)tb->order : p->error_order;

    double y[ORC_MAXD], fsal_f[ORC_MAXD];
    for (int j = 0; j < d; ++j) { y[j] = y0_in[j]; fsal_f[j] = 0.0; }
    int fsal_first = 1;

    double tprev = T0, tnext;
    /* controller state */
    double pinv = 1.0, ppinv = 1.0; int at_dtmin = 0;
    int32_t cs_completed = 1, cs_num = 0; double cs_t0 = T0, cs_t1 = T1;
    if (p->controller == ORC_PID) {
        double dt0 = DT0;
        if (is_none(dt0)) dt0 = select_initial_step(&cx, dir, T0, y, error_order);
        if (!is_none(p->dtmax)) dt0 = jmin(dt0, p->dtmax);
        if (!is_none(p->dtmin)) dt0 = jmax(dt0, p->dtmin);
        tnext = T0 + dt0;
    } else {
        /* ConstantStepSize.init (constant.py:30-54) */
        double q = ceil((T1 - T0) / nextup(DT0));
        cs_num = isfinite(T1) ? (int32_t)q : -1;
        cs_t1 = isfinite(T1) ? T1 : DT0;
        cs_completed = 1; cs_t0 = T0;
        tnext = T0 + DT0;
    }
    /* ClipStepSizeController.init (clip.py:243-299): the first step stops at the next
     * step time, or prevbefore the next jump time. */
    clipts_t cstep = {p->step_ts, p->step_ts ? p->n_step_ts : 0, dir};
    clipts_t cjump = {p->jump_ts, p->jump_ts ? p->n_jump_ts : 0, dir};
    int step_index = 0, jump_index = 0, made_jump = 0;
    if (cstep.n > 0) {
        step_index = clip_find(T0, &cstep, 0);
        tnext = clip_t(tnext, step_index, &cstep, 0);
    }
    if (cjump.n > 0) {
        jump_index = clip_find(T0, &cjump, 0);
        tnext = clip_t(tnext, jump_index, &cjump, 1);
    }
    tnext = jmin(tnext, T1);

    double t1_clip_floor = T1;
    for (int r = 0; r < 100; ++r) t1_clip_floor = prevbefore(t1_clip_floor);

    int save_index = 0, ts_index = 0;
    if (p->save_t0) {
        tso[0] = T0;
        for (int j = 0; j < d; ++j) ys[j] = y[j];
        save_index = 1;
    }

    int32_t num_steps = 0, num_acc = 0, num_rej = 0;
    int32_t result = ORC_SUCCESSFUL;
    stepout_t so;
    double ysave[ORC_MAXD];

    /* Event set-up (_integrate.py:1333-1426): values at (t0, y0); a boolean condition
     * that already holds ends the solve before the first step. */
    const int ne = (p->n_events > 0 && p->event_coeffs && p->event_flags) ? p->n_events : 0;
    double ev_old[8];
    int ev_mask[8] = {0};
    int ev_hit = 0;
    double tfinal = NAN, yfinal[ORC_MAXD];
    for (int i = 0; i < ne; ++i) {
        const double v = ev_value(p, i, T0, y);
        ev_old[i] = (p->event_flags[i] & 1) ? (double)(v > 0.0) : v;
        ev_mask[i] = (p->event_flags[i] & 1) ? (v > 0.0) && !ev_hit : 0;
        ev_hit |= ev_mask[i];
    }
    if (ev_hit) {
        result = ORC_EVENT_OCCURRED;
        tfinal = T0;
        for (int j = 0; j < d; ++j) yfinal[j] = y[j];
    }

    while (tprev < T1 && result == ORC_SUCCESSFUL && num_steps < p->max_steps) {
        /* a step that starts after a jump evaluates its first stage afresh
         * (runge_kutta.py:687, made_jump) */
        erk_step(tb, &cx, dir, tprev, tnext, y, fsal_first || made_jump, fsal_f, num_steps, &so);
        for (int j = 0; j < d; ++j) if (isnan(so.yerr[j])) so.yerr[j] = INFINITY;

        int keep;
        double ntprev, ntnext;
        int32_t step_result = ORC_SUCCESSFUL;
        if (p->controller == ORC_PID) {
            /* PIDController.adapt_step_size (pid.py:394-567) */
            double prev_dt = tnext - tprev;
            int anynan = 0;
            for (int j = 0; j < d; ++j) if (isnan(so.y1[j])) anynan = 1;
            double e[ORC_MAXD];
            for (int j = 0; j < d; ++j) {
                double y1c = anynan ? y[j] : so.y1[j];
                double yy = jmax(fabs(y[j]), fabs(y1c));
                e[j] = so.yerr[j] / (p->atol + yy * p->rtol);
            }
            double scaled = rms_norm(e, d);
            keep = scaled < 1.0;
            if (!is_none(p->dtmin)) keep = keep || at_dtmin;
            double inv = 1.0 / scaled;
            double coeff1 = (p->icoeff + p->pcoeff + p->dcoeff) / error_order;
            double coeff2 = -(p->pcoeff + 2.0 * p->dcoeff) / error_order;
            double coeff3 = p->dcoeff / error_order;
            double factor1 = (coeff1 == 0.0) ? 1.0 : pow_ctl(inv, coeff1);
            double factor2 = (coeff2 == 0.0) ? 1.0 : pow_ctl(pinv, coeff2);
            double factor3 = (coeff3 == 0.0) ? 1.0 : pow_ctl(ppinv, coeff3);
            double fmin = keep ? 1.0 : p->factormin;
            double fmax = keep ? p->factormax : p->safety;
            double factor = jclip(p->safety * factor1 * factor2 * factor3, fmin, fmax);
            double dt = prev_dt * factor;
            if (inv == 0.0 || isinf(inv)) inv = 1.0;
            if (!is_none(p->dtmax)) dt = jmin(dt, p->dtmax);
            if (!is_none(p->dtmin)) {
                if (!p->force_dtmin && dt < p->dtmin) step_result = ORC_DT_MIN_REACHED;
                if (at_dtmin && factor == 1.0) dt = p->dtmin;
                at_dtmin = dt <= p->dtmin;
                dt = jmax(dt, p->dtmin);
            }
            ntprev = keep ? tnext : tprev;
            ntnext = ntprev + dt;
            double new_inv = keep ? inv : pinv;
            double new_pinv = keep ? pinv : ppinv;
            pinv = new_inv; ppinv = new_pinv;
        } else {
            /* ConstantStepSize.adapt_step_size (constant.py:56-104) */
            keep = 1;
            int32_t sc = cs_completed + 1;
            double tn = cs_t0 + (cs_t1 - cs_t0) * ((double)sc / (double)cs_num);
            if (sc == cs_num) tn = cs_t1;
            if (!(cs_num >= 0)) tn = tnext + cs_t1;
            cs_completed = sc;
            ntprev = tnext;
            ntnext = tn;
        }
        if (p->controller == ORC_PID && (cstep.n > 0 || cjump.n > 0)) {
            /* ClipStepSizeController.adapt_step_size (clip.py:301-428) */
            int mj = 0;
            if (cstep.n > 0) {
                int unused;
                double bumped = clip_bump(ntprev, &cstep, &unused);
                step_index = clip_find(bumped, &cstep, step_index);
                ntnext = clip_t(ntnext, step_index, &cstep, 0);
            }
            if (cjump.n > 0) {
                ntprev = clip_bump(ntprev, &cjump, &mj);
                ntnext = jmax(nextup(ntprev), ntnext);
                jump_index = clip_find(ntprev, &cjump, jump_index);
                ntnext = clip_t(ntnext, jump_index, &cjump, 1);
            }
            if (keep) made_jump = mj;   /* _integrate.py:425 */
        }
        ntprev = jmin(ntprev, T1);
        /* _clip_to_end */
        if (ntnext > t1_clip_floor) ntnext = keep ? T1 : ntprev + 0.5 * (T1 - ntprev);

        num_steps += 1;
        if (keep) num_acc += 1; else num_rej += 1;
        if (result == ORC_SUCCESSFUL) result = step_result;

        /* Event check at the new state (_integrate.py:542-633); with a root finder the
         * event time is located on this step's interpolant (686-779) and the saves after
         * it are dropped (781-806). */
        if (ne > 0) {
            const double *ynew = keep ? so.y1 : y;
            int had = 0, fired = -1;
            for (int i = 0; i < ne; ++i) {
                const double v = ev_value(p, i, ntprev, ynew);
                const double nv = (p->event_flags[i] & 1) ? (double)(v > 0.0) : v;
                const int trig = ev_triggered(p->event_flags[i], ev_old[i], v);
                ev_mask[i] = trig && !had;
                if (ev_mask[i]) fired = i;
                had |= trig;
                ev_old[i] = nv;
            }
            if (had) {
                result = ORC_EVENT_OCCURRED;
                if (p->event_root_find) {
                    tfinal = ev_root(p, tb, fired, tprev, tnext, y, so.y1, &so);
                    interp_eval(p->solver, tb, d, tprev, tnext, y, so.y1, &so, tfinal, yfinal);
                } else {
                    tfinal = ntprev;
                    for (int j = 0; j < d; ++j) yfinal[j] = ynew[j];
                }
            }
        }

        if (keep) {
            if (grid) {
                grid[(long)(num_acc - 1) * 2 + 0] = tprev;
                grid[(long)(num_acc - 1) * 2 + 1] = tnext;
                if (p->grid_y) {
                    double *gy = p->grid_y + ((long)bi * p->max_steps + (num_acc - 1)) * d;
                    for (int j = 0; j < d; ++j) gy[j] = y[j];
                }
            }
            /* SaveAt(ts): the interpolant spans the step just taken [tprev, tnext]. */
            while (ts_index < n_ts && ts_in[ts_index] * dir <= tnext &&
                   !(result == ORC_EVENT_OCCURRED && ts_in[ts_index] * dir > tfinal)) {
                double tsv = ts_in[ts_index] * dir;
                interp_eval(p->solver, tb, d, tprev, tnext, y, so.y1, &so, tsv, ysave);
                tso[save_index] = tsv;
                for (int j = 0; j < d; ++j) ys[(long)save_index * d + j] = ysave[j];
                ts_index++; save_index++;
            }
            for (int j = 0; j < d; ++j) y[j] = so.y1[j];
            if (tb->fsal) { fsal_first = 0; for (int j = 0; j < d; ++j) fsal_f[j] = so.flast[j]; }
            /* SaveAt(steps=k): every k-th accepted step (_integrate.py:493-527) */
            if (p->save_steps > 0 && num_acc % p->save_steps == 0 &&
                !(result == ORC_EVENT_OCCURRED && ntprev > tfinal)) {
                tso[save_index] = ntprev;
                for (int j = 0; j < d; ++j) ys[(long)save_index * d + j] = y[j];
                save_index++;
            }
        }
        tprev = ntprev;
        tnext = ntnext;
    }

    /* t0 == t1: every requested ts is y0 (_integrate.py:823-845). */
    if (T0 == T1 && n_ts > 0) {
        for (int r = 0; r < n_ts; ++r) {
            tso[save_index + r] = T0;
            for (int j = 0; j < d; ++j) ys[(long)(save_index + r) * d + j] = y[j];
        }
        save_index += n_ts;
    }
    if (result == ORC_EVENT_OCCURRED) {   /* the solve ends at the event */
        tprev = tfinal;
        for (int j = 0; j < d; ++j) y[j] = yfinal[j];
    }
    /* _save_t1 (_integrate.py:847-877): the t1 row unless the last step was saved via
     * steps; with an event root finder, also when it was (its row after the event time
     * was removed). */
    const int via_steps = p->save_steps > 0 && num_acc % p->save_steps == 0;
    const int root = ne > 0 && p->event_root_find;
    const int t1_pred = root ? (p->save_t1 || via_steps) : (p->save_t1 && !via_steps);
    if (t1_pred && save_index < n_out) {
        tso[save_index] = tprev;
        for (int j = 0; j < d; ++j) ys[(long)save_index * d + j] = y[j];
        save_index++;
    }
    if (ne > 0 && p->event_mask)
        for (int i = 0; i < ne; ++i) p->event_mask[bi * ne + i] = ev_mask[i];
    if (tprev < T1 && result == ORC_SUCCESSFUL) result = ORC_MAX_STEPS_REACHED;
    for (int s = 0; s < n_out; ++s) tso[s] = tso[s] * dir;
    stats[0] = num_steps; stats[1] = num_acc; stats[2] = num_rej;
    *result_out = result;
    (void)bi;
}

int orc_n_out(const orc_problem *p) {
    int n = 0;
    if (p->save_t0) n += 1;
    n += p->n_ts;
    /* SaveAt(steps=k) (_integrate.py:1272-1305) */
    if (p->save_steps > 0) n += p->max_steps / p->save_steps;
    if (p->save_t1 && (p->save_steps <= 0 || p->max_steps % p->save_steps != 0)) n += 1;
    return n;
}

long orc_param_count(const orc_problem *p) {
    if (p->adj_base_d > 0) {
        orc_problem base = *p;
        base.d = p->adj_base_d;
        base.adj_base_d = 0;
        return orc_param_count(&base);
    }
    switch (p->vf) {
    case ORC_VF_LV: return 4;
    case ORC_VF_VDP: return 1;
    case ORC_VF_LINEAR: return (long)p->d * p->d;
    case ORC_VF_MLP: {
        long H = p->hidden, d = p->d;
        return H * d + H + H * H + H + d * H + d;
    }
    case ORC_VF_OU: return 3L * p->d;
    }
    return 0;
}

int orc_solve(const orc_problem *p, long B, const double *y0, const double *t1,
              long t1_stride, const double *ts, long ts_stride, const double *params,
              long param_stride, const double *dW, double *ys, double *ts_out,
              int32_t *stats, int32_t *result, double *grid, int nthreads) {
    init_tableaux();
    if (p->d < 1 || p->d > ORC_MAXD) return -1;
    if (p->solver < 0 || p->solver > 2) return -2;
    if (p->controller == ORC_CONSTANT && isnan(p->dt0)) return -3;
    int n_out = orc_n_out(p);
    long H = p->hidden > 0 ? p->hidden : 1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel
    {
        double *scratch = (double *)malloc(sizeof(double) * 2 * H);
#pragma omp for schedule(dynamic, 16)
        for (long b = 0; b < B; ++b) {
            const double *dWb = dW ? dW + b * p->d : NULL;
            long dW_stride = B * p->d;   /* dW layout [steps, B, d] */
            solve_one(p, b, y0 + b * p->d, t1[b * t1_stride], ts ? ts + b * ts_stride : NULL,
                      params + b * param_stride, dWb, dW_stride, ys + b * (long)n_out * p->d,
                      ts_out + b * (long)n_out, stats + b * 3, result + b,
                      grid ? grid + b * (long)p->max_steps * 2 : NULL, scratch);
        }
        free(scratch);
    }
    return 0;
}

/* Single step (test hook): one ERK step from (ts0, y0) to ts1 with a fresh first stage. */
int orc_single_step(const orc_problem *p, const double *params, double ts0, double ts1,
                    const double *y0, double *y1, double *yerr, double *k) {
    init_tableaux();
    const tableau_t *tb = &TAB[p->solver];
    double scratch[2 * 512];
    if (p->hidden > 512) return -1;
    vfctx_t cx = {p, params, NULL, 0, scratch};
    stepout_t so;
    double zero[ORC_MAXD] = {0};
    erk_step(tb, &cx, 1, ts0, ts1, y0, 1, zero, 0, &so);
    for (int j = 0; j < p->d; ++j) { y1[j] = so.y1[j]; yerr[j] = so.yerr[j]; }
    if (k) for (int m = 0; m < tb->S; ++m) for (int j = 0; j < p->d; ++j) k[m * p->d + j] = so.k[m][j];
    return 0;
}

/* BacksolveAdjoint (_adjoint.py:539-873): for each instance, from the last save back to
 * t0, add the save's cotangent to a_y and solve the augmented system backwards over the
 * interval with the forward's solver and controller (dt0 negated, _adjoint.py:629),
 * y restarting from the saved forward state; a fresh solve per interval reproduces the
 * reference's (its carried controller state is inert for the integral PID controller and
 * the first stage after the jump is re-evaluated). */
int orc_backsolve(const orc_problem *p, long B, const double *params, long param_stride,
                  const double *ys, const double *ts_out, const double *ys_bar,
                  double *y0_bar, double *params_bar, int32_t *result, int nthreads) {
    init_tableaux();
    const int db = p->d;
    const long P = orc_param_count(p);
    const int D = 2 * db + (int)P;
    if (D > ORC_MAXD || p->save_steps) return -1;
    const int n_out = orc_n_out(p);
    orc_problem ap = *p;
    ap.d = D;
    ap.adj_base_d = db;
    ap.n_ts = 0;
    ap.save_t0 = 0;
    ap.save_t1 = 1;
    ap.save_steps = 0;
    ap.dt0 = -p->dt0;
    ap.n_events = 0;
    ap.event_coeffs = NULL;
    ap.event_flags = NULL;
    ap.event_mask = NULL;
    ap.step_ts = ap.jump_ts = NULL;
    ap.n_step_ts = ap.n_jump_ts = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 16)
    for (long b = 0; b < B; ++b) {
        const double *prm = params + b * param_stride;
        const double *yb = ys + b * (long)n_out * db;
        const double *gb = ys_bar + b * (long)n_out * db;
        const double *tb = ts_out + b * (long)n_out;
        double ay[ORC_MAXD], aq[ORC_MAXD], aug[ORC_MAXD], out[ORC_MAXD], tso1;
        int32_t st[3], res_seg, res = ORC_SUCCESSFUL;
        for (int j = 0; j < db; ++j) ay[j] = 0.0;
        for (long m = 0; m < P; ++m) aq[m] = 0.0;
        for (int s = n_out - 1; s >= (p->save_t0 ? 1 : 0); --s) {
            for (int j = 0; j < db; ++j) {
                aug[j] = yb[s * db + j];
                aug[db + j] = ay[j] + gb[s * db + j];
            }
            for (long m = 0; m < P; ++m) aug[2 * db + m] = aq[m];
            orc_problem seg = ap;
            seg.t0 = tb[s];
            const double t_end = s > 0 ? tb[s - 1] : p->t0;
            solve_one(&seg, b, aug, t_end, NULL, prm, NULL, 0, out, &tso1, st, &res_seg, NULL, NULL);
            if (res == ORC_SUCCESSFUL) res = res_seg;
            for (int j = 0; j < db; ++j) ay[j] = out[db + j];
            for (long m = 0; m < P; ++m) aq[m] = out[2 * db + m];
        }
        if (p->save_t0 && n_out > 0)
            for (int j = 0; j < db; ++j) ay[j] = ay[j] + gb[j];
        for (int j = 0; j < db; ++j) y0_bar[b * db + j] = ay[j];
        for (long m = 0; m < P; ++m) params_bar[b * P + m] = aq[m];
        result[b] = res;
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Discrete adjoint (RecursiveCheckpointAdjoint)                               */
/* ------------------------------------------------------------------------- */

/* Transpose of interp_eval at time t: adds the cotangents of y0, y1 and k for the
 * cotangent g of the interpolated value (the interpolants are linear in y0, y1, k). */
static void interp_vjp(int solver, const tableau_t *tb, int d, double ts0, double ts1,
                       double t, const double *g, double *gy0, double *gy1,
                       double gk[MAXS][ORC_MAXD]) {
    double th = linear_rescale(ts0, t, ts1);
    if (solver == ORC_TSIT5) {
        double bb[7];
        tsit5_bb(th, bb);
        for (int j = 0; j < d; ++j) {
            gy0[j] = gy0[j] + g[j];
            for (int m = 0; m < 7; ++m) gk[m][j] = fma(bb[m], g[j], gk[m][j]);
        }
    } else if (solver == ORC_DOPRI5) {
        /* out = a th^4 + b th^3 + c th^2 + k0 th + y0 with ymid = y0 + cmid . k */
        double t2 = th * th, t3 = t2 * th, t4 = t3 * th;
        for (int j = 0; j < d; ++j) {
            double ga = t4 * g[j], gb = t3 * g[j], gc = t2 * g[j];
            double gmid = 16.0 * ga - 32.0 * gb + 16.0 * gc;
            gy0[j] = gy0[j] + ((g[j] + (-8.0 * ga + 18.0 * gb - 11.0 * gc)) + gmid);
            gy1[j] = gy1[j] + (-8.0 * ga + 14.0 * gb - 5.0 * gc);
            gk[0][j] = gk[0][j] + (th * g[j] + (-2.0 * ga + 5.0 * gb - 4.0 * gc));
            gk[6][j] = gk[6][j] + (2.0 * ga - 3.0 * gb + gc);
            for (int m = 0; m < 7; ++m)
                if (tb->cmid[m] != 0.0) gk[m][j] = gk[m][j] + tb->cmid[m] * gmid;
        }
    } else {
        /* Hermite: out = a th^3 + b th^2 + k0 th + y0 */
        double t2 = th * th, t3 = t2 * th;
        int L = tb->S - 1;
        for (int j = 0; j < d; ++j) {
            double ga = t3 * g[j], gb = t2 * g[j];
            gy0[j] = gy0[j] + (g[j] + (2.0 * ga - 3.0 * gb));
            gy1[j] = gy1[j] + (-2.0 * ga + 3.0 * gb);
            gk[0][j] = gk[0][j] + (th * g[j] + (ga - 2.0 * gb));
            gk[L][j] = gk[L][j] + (ga - gb);
        }
    }
}

/* One instance: reverse sweep over the accepted steps recorded by the forward solve.
 * Per step, newest first: the stages are recomputed from the recorded start state (as
 * erk_step computes them, every stage evaluated), the cotangents of the rows saved in the
 * step are pulled back through the interpolant, then y1 = y0 + b . k and the stages are
 * transposed, last stage first (runge_kutta.py:847-1052 read backwards). */
static void adjoint_one(const orc_problem *p, const tableau_t *tb, const double *prm,
                        double t1r, const double *tsb, const double *grid, const double *gy,
                        int n, const double *g, double *y0b, double *qb) {
    const int d = p->d, S = tb->S;
    const long P = orc_param_count(p);
    vfctx_t cx = {p, prm, NULL, 0, NULL};
    const int dir = (p->t0 < t1r) ? 1 : -1;
    const double T0 = p->t0 * dir, T1 = t1r * dir;
    int n_saved = 0;
    if (T0 == T1) {
        n_saved = p->n_ts;
    } else if (n > 0) {
        const double tend = grid[(long)(n - 1) * 2 + 1];
        while (n_saved < p->n_ts && tsb[n_saved] * dir <= tend) ++n_saved;
    }
    double ab[ORC_MAXD], gq[ORC_MAXD * ORC_MAXD], gqi[ORC_MAXD * ORC_MAXD];
    for (int j = 0; j < d; ++j) ab[j] = 0.0;
    for (long m = 0; m < P; ++m) gq[m] = 0.0;
    const int ksave = p->save_steps;
    int r = p->save_t0 + n_saved + (ksave > 0 ? n / ksave : 0);
    if (p->save_t1 && !(ksave > 0 && n % ksave == 0))
        for (int j = 0; j < d; ++j) ab[j] = ab[j] + g[(long)r * d + j];
    r -= 1;
    int ti = (T0 == T1) ? -1 : n_saved - 1;
    for (int s = n - 1; s >= 0; --s) {
        const double ts0 = grid[(long)s * 2], ts1 = grid[(long)s * 2 + 1];
        const double prev_end = s > 0 ? grid[(long)(s - 1) * 2 + 1] : -INFINITY;
        const double dt = ts1 - ts0;
        const double control = (double)dir * dt;
        double y0[ORC_MAXD], z[MAXS][ORC_MAXD], k[MAXS][ORC_MAXD], f[ORC_MAXD];
        for (int j = 0; j < d; ++j) y0[j] = gy[(long)s * d + j];
        for (int i = 0; i < S; ++i) {
            for (int j = 0; j < d; ++j) {
                double acc = 0.0; int first = 1;
                for (int m = 0; m < i; ++m) {
                    if (tb->a[i][m] == 0.0) continue;
                    if (first) { acc = tb->a[i][m] * k[m][j]; first = 0; }
                    else acc = fma(tb->a[i][m], k[m][j], acc);
                }
                z[i][j] = y0[j] + acc;
            }
            double tsi = (tb->c[i] == 1.0) ? ts1 : ts0 + tb->c[i] * dt;
            vf_eval(&cx, tsi * (double)dir, z[i], f);
            for (int j = 0; j < d; ++j) k[i][j] = control * f[j];
        }
        double gy1[ORC_MAXD], gy0[ORC_MAXD], gk[MAXS][ORC_MAXD];
        for (int j = 0; j < d; ++j) {
            gy1[j] = ab[j];
            gy0[j] = 0.0;
            for (int m = 0; m < S; ++m) gk[m][j] = 0.0;
        }
        if (ksave > 0 && (s + 1) % ksave == 0) {
            for (int j = 0; j < d; ++j) gy1[j] = gy1[j] + g[(long)r * d + j];
            --r;
        }
        while (ti >= 0 && tsb[ti] * dir > prev_end) {
            interp_vjp(p->solver, tb, d, ts0, ts1, tsb[ti] * dir, g + (long)r * d, gy0, gy1, gk);
            --ti;
            --r;
        }
        for (int j = 0; j < d; ++j) gy0[j] = gy0[j] + gy1[j];
        for (int m = 0; m < S; ++m) {
            if (tb->b[m] == 0.0) continue;
            for (int j = 0; j < d; ++j) gk[m][j] = fma(tb->b[m], gy1[j], gk[m][j]);
        }
        for (int i = S - 1; i >= 0; --i) {
            double w[ORC_MAXD], gz[ORC_MAXD];
            for (int j = 0; j < d; ++j) w[j] = control * gk[i][j];
            vf_vjp(p->vf, d, prm, z[i], w, gz, gqi);
            for (long m = 0; m < P; ++m) gq[m] = gq[m] + gqi[m];
            for (int j = 0; j < d; ++j) gy0[j] = gy0[j] + gz[j];
            for (int l = 0; l < i; ++l) {
                if (tb->a[i][l] == 0.0) continue;
                for (int j = 0; j < d; ++j) gk[l][j] = fma(tb->a[i][l], gz[j], gk[l][j]);
            }
        }
        for (int j = 0; j < d; ++j) ab[j] = gy0[j];
    }
    if (T0 == T1)
        for (int rr = p->save_t0; rr < p->save_t0 + p->n_ts; ++rr)
            for (int j = 0; j < d; ++j) ab[j] = ab[j] + g[(long)rr * d + j];
    if (p->save_t0)
        for (int j = 0; j < d; ++j) ab[j] = ab[j] + g[j];
    for (int j = 0; j < d; ++j) y0b[j] = ab[j];
    for (long m = 0; m < P; ++m) qb[m] = gq[m];
}

int orc_adjoint(const orc_problem *p, long B, const double *params, long param_stride,
                const double *t1, long t1_stride, const double *ts, long ts_stride,
                const double *grid, const double *grid_y, const int32_t *stats,
                const double *ys_bar, double *y0_bar, double *params_bar, int nthreads) {
    init_tableaux();
    if (p->vf != ORC_VF_LV && p->vf != ORC_VF_VDP && p->vf != ORC_VF_LINEAR) return -1;
    if (p->vf == ORC_VF_LINEAR && (p->d < 1 || p->d > 4)) return -1;
    if (p->n_events > 0) return -1;
    const tableau_t *tb = &TAB[p->solver];
    const long P = orc_param_count(p);
    const int n_out = orc_n_out(p);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 16)
    for (long b = 0; b < B; ++b) {
        adjoint_one(p, tb, params + b * param_stride, t1[b * t1_stride],
                    ts ? ts + b * ts_stride : NULL, grid + b * (long)p->max_steps * 2,
                    grid_y + b * (long)p->max_steps * p->d, stats[b * 3 + 1],
                    ys_bar + b * (long)n_out * p->d, y0_bar + b * p->d, params_bar + b * P);
    }
    return 0;
}
