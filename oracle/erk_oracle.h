/*
 * erk_oracle.h -- interface of the FP64 CPU oracle.  TEST INFRASTRUCTURE ONLY:
 * imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 */
#ifndef ERK_ORACLE_H
#define ERK_ORACLE_H
#include <stdint.h>

#define ORC_MAXD 64

enum { ORC_TSIT5 = 0, ORC_DOPRI5 = 1, ORC_HEUN = 2 };
enum { ORC_PID = 0, ORC_CONSTANT = 1 };
enum { ORC_VF_LV = 0, ORC_VF_VDP = 1, ORC_VF_LINEAR = 2, ORC_VF_MLP = 3, ORC_VF_OU = 4 };
/* Result codes: names follow diffrax RESULTS (_solution.py:13-31); successful == 0. */
enum { ORC_SUCCESSFUL = 0, ORC_MAX_STEPS_REACHED = 1, ORC_DT_MIN_REACHED = 2,
       ORC_EVENT_OCCURRED = 3 };

typedef struct {
    int32_t solver, controller, vf, d, hidden, n_ts;
    int32_t save_t0, save_t1, save_steps, max_steps, force_dtmin, pad_;
    double t0, dt0;                       /* dt0 = NaN means None */
    double rtol, atol, pcoeff, icoeff, dcoeff;
    double dtmin, dtmax;                  /* NaN means None */
    double factormin, factormax, safety;
    double error_order;                   /* NaN means the solver's own */
    /* ClipStepSizeController(step_ts=..., jump_ts=...) around the PID controller:
     * ascending times (unscaled), NULL / 0 when absent. */
    const double *step_ts, *jump_ts;
    int32_t n_step_ts, n_jump_ts;
    /* Event (_event.py:13-118) with affine conditions c_t t + c_y . y + c_0 in normalised
     * time: event_coeffs [n_events, d + 2] = (c_t, c_y[d], c_0); event_flags[i] bit 0 =
     * boolean (value > 0), bits 1-2 = direction (0 either, 1 up, 2 down);
     * event_root_find = 1 locates the time by bisection on the step's interpolant;
     * event_mask [B, n_events] out (NULL to skip). */
    const double *event_coeffs;
    const int32_t *event_flags;
    int32_t *event_mask;
    int32_t n_events, event_root_find;
    /* > 0: the vector field is the continuous-adjoint augmentation (AdjointTerm,
     * _term.py:767-917) of vf with state dimension adj_base_d; d = 2 adj_base_d + P. */
    int32_t adj_base_d, pad2_;
    /* With a grid: [B, max_steps, d] start state of every accepted step (or NULL). */
    double *grid_y;
    /* Per-instance PIDController tolerances [B] (a vmapped rtol / atol), or NULL for the
     * scalars above. */
    const double *rtol_b, *atol_b;
} orc_problem;

int orc_n_out(const orc_problem *p);
long orc_param_count(const orc_problem *p);
int orc_tableau(int solver, double *a49, double *b7, double *berr7, double *c7, int *info4);
/* Batched solve.  y0 [B,d]; t1 [B*t1_stride]; ts [B*ts_stride + n_ts]; params
 * [B*param_stride + count]; dW [steps, B, d] or NULL; ys [B, n_out, d];
 * ts_out [B, n_out]; stats [B,3] = (num_steps, accepted, rejected); result [B];
 * grid [B, max_steps, 2] accepted (tprev, tnext) in normalised time, or NULL. */
int orc_solve(const orc_problem *p, long B, const double *y0, const double *t1,
              long t1_stride, const double *ts, long ts_stride, const double *params,
              long param_stride, const double *dW, double *ys, double *ts_out,
              int32_t *stats, int32_t *result, double *grid, int nthreads);
double orc_pow(double x, double c);
/* BacksolveAdjoint (_adjoint.py:539-873) for the forward problem p (SaveAt t0 / ts / t1):
 * ys, ts_out [B, n_out, d] / [B, n_out] from orc_solve, ys_bar their cotangents; writes
 * y0_bar [B, d], params_bar [B, P], result [B]. */
int orc_backsolve(const orc_problem *p, long B, const double *params, long param_stride,
                  const double *ys, const double *ts_out, const double *ys_bar,
                  double *y0_bar, double *params_bar, int32_t *result, int nthreads);
/* Discrete adjoint (RecursiveCheckpointAdjoint, _adjoint.py:264-312) of the forward
 * problem p: grid / grid_y [B, max_steps, 2] / [B, max_steps, d] and stats [B, 3] recorded
 * by orc_solve, ys_bar [B, n_out, d] the cotangents of its saved rows; writes y0_bar
 * [B, d] and params_bar [B, P]. */
int orc_adjoint(const orc_problem *p, long B, const double *params, long param_stride,
                const double *t1, long t1_stride, const double *ts, long ts_stride,
                const double *grid, const double *grid_y, const int32_t *stats,
                const double *ys_bar, double *y0_bar, double *params_bar, int nthreads);
/* ClipStepSizeController helpers (clip.py:82-118), exposed for their unit tests. */
int orc_clip_find(const double *ts, int n, int dir, double t, int hint);
double orc_clip_bump(const double *ts, int n, int dir, double t, int *made_jump);
int orc_single_step(const orc_problem *p, const double *params, double ts0, double ts1,
                    const double *y0, double *y1, double *yerr, double *k);
#endif
