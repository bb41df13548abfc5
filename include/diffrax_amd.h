/*
 * diffrax_amd.h -- C ABI of the MI355X explicit Runge-Kutta integrator.
 *
 * The library replaces the body of one diffrax call,
 *     diffeqsolve(terms, solver, t0, t1, dt0, y0, args, saveat=SaveAt(...),
 *                 stepsize_controller=PIDController(...)|ConstantStepSize(),
 *                 adjoint=RecursiveCheckpointAdjoint())
 * for a batch of independent instances (what jax.vmap over diffeqsolve computes),
 * with the vector field chosen from a registered set of device functions.
 *
 * Conventions.
 *  - Plain C: pointers are device pointers (hipMalloc / torch CUDA tensors), sizes are
 *    element counts, `stream` is a hipStream_t passed as void*.
 *  - The library allocates no device memory.  Scratch memory the kernels need is a
 *    caller-provided workspace sized by the matching *_workspace_bytes call; the one
 *    exception, dfx_sde_heun_adjoint_static, uses a fixed static device array instead.
 *  - Every call is stream ordered, makes no host synchronisation and may be captured
 *    into a hipGraph.
 *  - Return value: 0 on success, a negative DFX_E* code otherwise (dfx_strerror).
 *    Per-instance numerical outcomes go to int32 result[B] with the codes below,
 *    whose names follow diffrax RESULTS (_solution.py:13-31; successful == 0).
 */
#ifndef DIFFRAX_AMD_H
#define DIFFRAX_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enums ---------------------------------------------------------------------- */
enum { DFX_TSIT5 = 0, DFX_DOPRI5 = 1, DFX_HEUN = 2 };          /* solver              */
enum { DFX_PID = 0, DFX_CONSTANT = 1 };                         /* stepsize controller */
enum { DFX_VF_LV = 0, DFX_VF_VDP = 1, DFX_VF_LINEAR = 2,       /* vector field        */
       DFX_VF_MLP = 3, DFX_VF_OU = 4 };
enum { DFX_F64 = 0,          /* state, stages, time and controller in float64            */
       DFX_F32S_F64T = 1,    /* float32 state and stages; float64 time and controller    */
       DFX_F32 = 2,          /* MLP: float32 matmuls (MFMA f32)                          */
       DFX_BF16_F32 = 3 };   /* MLP: bf16 MFMA matmuls, f32 accumulate, f32 RK algebra   */
enum { DFX_SUCCESSFUL = 0, DFX_MAX_STEPS_REACHED = 1, DFX_DT_MIN_REACHED = 2,
       DFX_EVENT_OCCURRED = 3 };

#define DFX_OK 0
#define DFX_E_ARG (-1)        /* invalid argument / unsupported combination */
#define DFX_E_HIP (-2)        /* a HIP runtime call failed                  */
#define DFX_E_WORKSPACE (-3)  /* workspace missing or too small             */
#define DFX_E_TOO_LARGE (-4)  /* problem needs more than the static scratch */

/* ---- problem description (host memory) ------------------------------------------- */
/* Fields mirror diffeqsolve's arguments (_integrate.py:890-911), SaveAt
 * (_saveat.py:65-105) and PIDController (_step_size_controller/pid.py:299-311).
 * NaN in dt0 / dtmin / dtmax / error_order means None. */
typedef struct dfx_erk_problem {
  int32_t solver, controller, vf, precision;
  int32_t d;            /* state dimension                                          */
  int32_t hidden;       /* MLP width (vf == DFX_VF_MLP), else 0                       */
  int32_t n_ts;         /* SaveAt(ts=...) length                                      */
  int32_t save_t0, save_t1;
  int32_t save_steps;   /* SaveAt(steps=k): k > 0 saves every k-th accepted step, 0 none */
  int32_t max_steps, force_dtmin;
  int64_t batch;        /* B instances                                                */
  int64_t t1_stride;    /* 0: one t1 for all; 1: t1[B]                                */
  int64_t ts_stride;    /* 0: ts[n_ts] shared; n_ts: ts[B, n_ts]                       */
  int64_t param_stride; /* 0: shared parameters; P: params[B, P]                        */
  double t0, dt0;
  double rtol, atol, pcoeff, icoeff, dcoeff, dtmin, dtmax;
  double factormin, factormax, safety, error_order;
  /* ClipStepSizeController(PIDController, step_ts, jump_ts) (clip.py:120-428): counts of
   * the ascending time arrays in dfx_erk_io (0 = none; PID only). */
  int32_t n_step_ts, n_jump_ts;
  /* Event (_event.py:13-118): n_events (<= 4) affine conditions c_t t + c_y . y + c_0 in
   * direction-normalised time, as the reference evaluates cond_fn; event_root_find = 1
   * locates the event time on the triggering step's interpolant. */
  int32_t n_events, event_root_find;
  /* RecursiveCheckpointAdjoint (_adjoint.py:264-329): K > 0 makes dfx_erk_solve store the
   * lane state every K accepted steps in dfx_erk_io.ckpt (the restart points of
   * dfx_erk_adjoint); 0 stores none.  The Python package uses
   * K = ceil(max_steps / checkpoints), checkpoints = ceil(sqrt(max_steps)) by default. */
  int32_t ckpt_every;
  /* Resuming a solve: diffeqsolve(solver_state=, controller_state=, made_jump=)
   * (_integrate.py:906-908, 1252-1270, 1309).  A mask of DFX_RESUME_* naming which parts
   * of dfx_erk_io.state_in replace the fresh start; 0 starts afresh.  With
   * DFX_RESUME_CONTROLLER and dt0 given, the first step ends at t0 + dt0 unclamped (the
   * controller's init is skipped); with dt0 NaN the initial step is selected as usual.
   * Not with step_ts / jump_ts clipping, checkpoints or the adjoints. */
  int32_t resume;
} dfx_erk_problem;

enum { DFX_RESUME_SOLVER = 1, DFX_RESUME_CONTROLLER = 2, DFX_RESUME_MADE_JUMP = 4 };

/* The solver / controller state record of one instance (dfx_erk_io.state_in / state_out),
 * DFX_ERK_STATE_WORDS(d) float64 words, times direction-normalised:
 *   [0, d)  solver_state f0: the FSAL stage f(t, y) (runge_kutta.py:415-442, 1199-1200;
 *           zero for Heun, which keeps no solver state)
 *   d       solver_state first_step (1 or 0)
 *   d + 1   made_jump (1 or 0)
 *   d + 2   PIDController state: prev_inv_scaled_error (pid.py:84-85)
 *   d + 3   prev_prev_inv_scaled_error
 *   d + 4   at_dtmin (1 or 0)
 *   d + 5   ConstantStepSize state: steps_completed (constant.py:14-17, 47-53)
 *   d + 6   num_steps (-1: infinite t1)
 *   d + 7   t0_sim
 *   d + 8   t1_sim_or_dt0
 * Both controllers' words are written; a resume reads those of the problem's controller. */
#define DFX_ERK_STATE_WORDS(d) ((d) + 9)

/* ---- device buffers of one solve --------------------------------------------------- */
typedef struct dfx_erk_io {
  const void* y0;       /* [B, d]  state dtype (double for F64, float otherwise)       */
  const double* t1;     /* [B * t1_stride] (at least one element)                      */
  const double* ts;     /* [B * ts_stride + n_ts] or NULL when n_ts == 0               */
  const void* params;   /* vector-field parameters, state dtype (MLP: see below)      */
  void* ys;             /* out [B, n_out, d] state dtype; unfilled entries = inf       */
  double* ts_out;       /* out [B, n_out]; unfilled entries = +inf (as the reference's
                           direction*inf preset after its rescaling by direction)    */
  int32_t* stats;       /* out [B, 3] = num_steps, num_accepted_steps, num_rejected   */
  int32_t* result;      /* out [B] DFX_SUCCESSFUL / DFX_MAX_STEPS_REACHED / ...      */
  double* grid;         /* optional out [B, max_steps, 2]: accepted (tprev, tnext) in
                           direction-normalised time; NULL to skip                  */
  /* SaveAt(dense=True), all three or none (NULL):                                     */
  double* dense_ts;     /* out [B, max_steps + 1]: t0 then each accepted step's end time,
                           direction-normalised (entries past the accepted count unset) */
  void* dense_ys;       /* out [B, max_steps + 1, d] state dtype: the states at dense_ts */
  void* dense_k;        /* out [B, max_steps, stages, d]: each accepted step's stage
                           increments k_i = dt f(z_i)                                 */
  const double* step_ts; /* [n_step_ts] ascending, shared by the batch, or NULL           */
  const double* jump_ts; /* [n_jump_ts] ascending, or NULL                                */
  const double* event_coeffs; /* [n_events, d + 2]: (c_t, c_y[d], c_0) per condition       */
  const int32_t* event_flags; /* [n_events]: bit 0 boolean (value > 0), bits 1-2 direction
                                 (0 either sign change, 1 upcrossing, 2 downcrossing)      */
  int32_t* event_mask;        /* out [B, n_events]: the condition that ended the solve    */
  void* grid_y;         /* optional out [B, max_steps, d] state dtype: each accepted step's
                           start state (needs grid)                                    */
  const double* rtol;   /* optional [B]: per-instance PIDController rtol (a vmapped
                           tolerance); NULL uses dfx_erk_problem.rtol                  */
  const double* atol;   /* optional [B]: per-instance atol; NULL uses the problem's     */
  double* ckpt;         /* out when ckpt_every > 0: [B, n_ckpt, words] (dfx_erk_ckpt_layout)
                           lane state after 0, K, 2K, ... accepted steps; NULL otherwise */
  const double* state_in; /* [B, DFX_ERK_STATE_WORDS(d)] when problem.resume != 0: the
                             state to resume from; NULL otherwise                      */
  double* state_out;      /* optional out [B, DFX_ERK_STATE_WORDS(d)]: the final solver /
                             controller state and made_jump (SaveAt(solver_state=True,
                             controller_state=True, made_jump=True), _integrate.py:
                             1489-1500); NULL to skip                                   */
} dfx_erk_io;

/* Launch shape (NULL: block 64, one lane per instance).  Fewer lanes than instances
 * select the refill kernel; the Python package picks the shape by timing candidates. */
typedef struct dfx_launch {
  int32_t block;        /* threads per workgroup: a multiple of 64, at most 256       */
  int32_t lanes;        /* total persistent lanes (trajectory slots); 0 = batch      */
} dfx_launch;

/* Library version: major*10000 + minor*100 + patch. */
int dfx_version(void);
const char* dfx_strerror(int code);
/* Number of saved rows per instance: t0 + n_ts + (steps ? max_steps : 0) + t1
 * (the allocation rule of _integrate.py:1272-1305 for one SubSaveAt). */
int dfx_erk_n_out(const dfx_erk_problem* p);
/* Number of vector-field parameters per instance (LV 4, VdP 1, linear d*d,
 * MLP h*d + h + h*h + h + d*h + d, OU 3d). */
int64_t dfx_vf_param_count(int32_t vf, int32_t d, int32_t hidden);
/* Copy a solver's tableau: a[7*7] row-major (zero padded), b[7], b_err[7], c[7],
 * info[4] = (stages, fsal, ssal, order).  Host memory.
 * Replaces: ButcherTableau (_solver/runge_kutta.py:48-166) of Tsit5/Dopri5/Heun. */
int dfx_tableau(int32_t solver, double* a, double* b, double* b_err, double* c,
                int32_t* info);

/* Workspace needed by dfx_erk_solve (bytes, 256-aligned). */
int dfx_erk_workspace_bytes(const dfx_erk_problem* p, size_t* bytes);

/* Batched adaptive ODE solve of small systems (LV, VdP, linear; d <= 16).
 * One persistent kernel: each lane owns one trajectory and steps, rejects, saves and
 * refills independently.
 * Replaces: the whole integration loop that AbstractAdjoint.loop runs
 * (_adjoint.py:118-144 -> _integrate.py:302-885), i.e. AbstractRungeKutta.step
 * (_solver/runge_kutta.py:446-1203), PIDController.adapt_step_size
 * (_step_size_controller/pid.py:394-567) / ConstantStepSize.adapt_step_size
 * (_step_size_controller/constant.py:56-104), the interpolation behind SaveAt(ts)
 * (_solver/tsit5.py:103-156, _local_interpolation.py:50-139) and, when dt0 is NaN,
 * _select_initial_step (_step_size_controller/pid.py:33-81). */
int dfx_erk_solve(const dfx_erk_problem* p, const dfx_erk_io* io, void* workspace,
                  size_t workspace_bytes, const dfx_launch* launch, void* stream);

/* Continuous adjoint (BacksolveAdjoint): buffers of a finished dfx_erk_solve with
 * SaveAt(t0 / ts / t1) and the cotangents of its saved states. */
typedef struct dfx_erk_backsolve_io {
  const void* params;      /* as given to dfx_erk_solve (state dtype)                        */
  const void* ys;          /* [B, n_out, d] saved states of the forward solve                */
  const double* ts_out;    /* [B, n_out] their times                                         */
  const void* ys_bar;      /* [B, n_out, d] cotangents of ys                                 */
  void* y0_bar;            /* out [B, d]: dL/dy0                                             */
  void* params_bar;        /* out [B, P]: dL/dparams per instance (sum them if shared)       */
  int32_t* result;         /* out [B]: result codes of the backward solves                   */
} dfx_erk_backsolve_io;

/* Gradients by the continuous adjoint: from the last save back to t0, the augmented
 * system (y, a_y, a_params) with vector field (f, -a_y^T df/dy, -a_y^T df/dparams) is
 * solved backwards between consecutive save times with the forward's solver and
 * controller, y restarting from the saved state and a_y receiving each save's cotangent.
 * p is the forward problem (LV, VdP, linear with d <= 3; PID with integral control and no
 * dtmin, or constant steps; no steps / clip / events).
 * Replaces: BacksolveAdjoint (_adjoint.py:539-873) with AdjointTerm (_term.py:767-917). */
int dfx_erk_backsolve(const dfx_erk_problem* p, const dfx_erk_backsolve_io* io, void* stream);

/* Discrete adjoint (RecursiveCheckpointAdjoint): the checkpoints and counters of a
 * finished dfx_erk_solve run with ckpt_every > 0, and the cotangents of its saved rows. */
typedef struct dfx_erk_adjoint_io {
  const double* ckpt;      /* [B, n_ckpt, words] written by dfx_erk_solve                    */
  const int32_t* stats;    /* [B, 3] written by dfx_erk_solve                                */
  const void* ys_bar;      /* [B, n_out, d] cotangents of the saved rows                     */
  void* y0_bar;            /* out [B, d]: dL/dy0                                             */
  void* params_bar;        /* out [B, P]: dL/dparams per instance (sum them if shared)       */
} dfx_erk_adjoint_io;

/* Checkpoint layout of a problem with ckpt_every = K > 0: n_ckpt = ceil(max_steps / K)
 * records per instance of `words` float64 words (d + 8: y, t_prev, t_next, the PID state,
 * the end of the previous accepted step and six int32 counters). */
int dfx_erk_ckpt_layout(const dfx_erk_problem* p, int32_t* n_ckpt, int32_t* words);
/* Workspace of dfx_erk_adjoint (bytes, 256-aligned): a 2K-record scratch per instance
 * holding two replayed segments' accepted steps (one being swept while the next is
 * replayed), (t_start, t_end) float64 and the start state. */
int dfx_erk_adjoint_workspace_bytes(const dfx_erk_problem* p, size_t* bytes);

/* Gradients of a loss of the saved rows with respect to y0 and the vector-field
 * parameters, by reverse mode through the forward's accepted steps.  Segment by segment,
 * newest first, the forward step loop is restarted from the segment's checkpoint and
 * repeats its accepted steps (rejected ones included) bit for bit; then, per accepted
 * step, newest first, the stages are recomputed from the replayed start state and
 * transposed, together with the interpolation of the rows saved inside the step.  Step
 * sizes are constants of the gradient and rejected steps contribute nothing, as in the
 * reference.  p and fwd are the forward's problem and buffers (its inputs are read: t1,
 * ts, params, rtol / atol, step_ts / jump_ts; its outputs are not touched); p: LV, VdP,
 * linear with d <= 4, no events, no dense output, ckpt_every > 0.
 * Replaces: RecursiveCheckpointAdjoint.loop (_adjoint.py:264-312) differentiating the
 * accepted steps of _integrate.py:302-662 (AbstractRungeKutta.step,
 * _solver/runge_kutta.py:446-1203, and the save interpolation, _integrate.py:456-491),
 * with the stop-gradient sites of pid.py:374, 498-530. */
int dfx_erk_adjoint(const dfx_erk_problem* p, const dfx_erk_io* fwd,
                    const dfx_erk_adjoint_io* io, void* workspace, size_t workspace_bytes,
                    void* stream);

/* Query side of SaveAt(dense=True): buffers a dfx_erk_solve wrote, and query times. */
typedef struct dfx_erk_dense_io {
  const double* dense_ts;  /* [B, max_steps + 1] from dfx_erk_solve                       */
  const void* dense_ys;    /* [B, max_steps + 1, d]                                       */
  const void* dense_k;     /* [B, max_steps, stages, d]                                   */
  const int32_t* stats;    /* [B, 3] from dfx_erk_solve (accepted-step counts)            */
  const double* t1;        /* [B * t1_stride] as given to dfx_erk_solve                   */
  const double* tq;        /* query times [B * tq_stride + n_q] (unnormalised)           */
  int64_t n_q;             /* queries per instance                                        */
  int64_t tq_stride;       /* 0: the same n_q times for every instance; n_q: per instance */
  int32_t left;            /* 1: a time on a step boundary belongs to the earlier step    */
  int32_t derivative;      /* 0: y(t); 1: dy/dt                                           */
  void* out;               /* out [B, n_q, d] state dtype                                 */
} dfx_erk_dense_io;

/* Evaluate the dense solution (or its time derivative) of every instance at its query
 * times.  p is the problem the buffers were solved with.
 * Replaces: DenseInterpolation.evaluate / derivative (_global_interpolation.py:315-400)
 * behind Solution.evaluate / Solution.derivative (_solution.py:103-132), with the
 * solver's local interpolant (_solver/tsit5.py:103-156, _local_interpolation.py:50-139). */
int dfx_erk_dense_eval(const dfx_erk_problem* p, const dfx_erk_dense_io* io, void* stream);

/* ---- SDE: Heun with diagonal noise on the registered OU terms ----------------------- */
/* dy = theta (mu - y) dt + sigma dW, theta/mu/sigma of length d shared by all paths. */
typedef struct dfx_sde_problem {
  int32_t precision;        /* DFX_F64 (float64 state) or DFX_F32 (float32 state)        */
  int32_t d;                /* components per path (multiple of 4 for F32, of 2 for F64) */
  int64_t paths;            /* P                                                        */
  int32_t num_steps;        /* N steps of the ConstantStepSize grid                      */
  int32_t checkpoint_every; /* CK: states stored every CK steps for the adjoint (8/16/32),
                               0 when no gradient will be taken                          */
  int32_t save_t0, save_t1; /* SaveAt(t0=..., t1=...)                                    */
  int32_t save_steps;       /* SaveAt(steps=k): k > 0 also saves y after every k-th step;
                               the t1 row is dropped when k divides num_steps            */
  int32_t n_ts;             /* SaveAt(ts): dfx_sde_io.ts holds n_ts ascending times in
                               [t0, t1], shared by all paths, interpolated inside their
                               steps by Heun's Hermite polynomial                        */
} dfx_sde_problem;

typedef struct dfx_sde_io {
  const void* y0;           /* [P, d] state dtype                                        */
  const void* theta;        /* [d] state dtype                                           */
  const void* mu;           /* [d]                                                       */
  const void* sigma;        /* [d]                                                       */
  const float* dW;          /* [N, P, d] float32 Brownian increments (step-major)         */
  const double* grid;       /* [N + 1] step times of ConstantStepSize (host-computed)     */
  void* ys;                 /* out [P, dfx_sde_n_out, d]: t0, then the ts and every k-th
                               step in time order (a step's ts rows before its step row),
                               then t1                                                    */
  void* ckpt;               /* out [dfx_sde_n_ckpt, P, d] states at steps CK, 2CK, ...
                               (NULL when no gradient is needed)                          */
  const double* ts;         /* [n_ts] SaveAt(ts) times, or NULL when n_ts == 0            */
  void* dense_ys;           /* SaveAt(dense=True), both or neither, forward only (no
                               checkpoints): out [P, N + 1, d] the step-end states         */
  void* dense_k;            /* out [P, N, 2, d] each step's two stage increments          */
} dfx_sde_io;

/* Query side of the SDE's SaveAt(dense=True): buffers dfx_sde_heun_solve wrote. */
typedef struct dfx_sde_dense_io {
  const void* dense_ys;     /* [P, N + 1, d] from dfx_sde_heun_solve                       */
  const void* dense_k;      /* [P, N, 2, d]                                                */
  const double* grid;       /* [N + 1] as given to dfx_sde_heun_solve                      */
  const double* tq;         /* query times [P * tq_stride + n_q]                           */
  int64_t n_q;              /* queries per path                                            */
  int64_t tq_stride;        /* 0: the same n_q times for every path; n_q: per path         */
  int32_t left;             /* 1: a time on a step boundary belongs to the earlier step    */
  int32_t derivative;       /* 0: y(t); 1: dy/dt                                           */
  void* out;                /* out [P, n_q, d] state dtype                                 */
} dfx_sde_dense_io;

typedef struct dfx_sde_grad_io {
  const void* ys_bar;       /* [P, dfx_sde_n_out, d] cotangent of ys                       */
  void* y0_bar;             /* out [P, d] or NULL                                          */
  double* theta_bar;        /* out [d] float64                                             */
  double* mu_bar;           /* out [d]                                                     */
  double* sigma_bar;        /* out [d]                                                     */
} dfx_sde_grad_io;

/* Saved rows per path: save_t0 + n_ts + num_steps / save_steps (when save_steps > 0) +
 * save_t1 unless save_steps divides num_steps (the allocation rule of
 * _integrate.py:1272-1305). */
int dfx_sde_n_out(const dfx_sde_problem* p);
/* Number of stored checkpoints: ceil(N / CK) - 1. */
int dfx_sde_n_ckpt(const dfx_sde_problem* p);
/* Workspace of dfx_sde_heun_adjoint (per-block partial sums), bytes. */
int dfx_sde_workspace_bytes(const dfx_sde_problem* p, size_t* bytes);
/* Forward Heun solve of P independent paths; writes ys and (if ckpt) the checkpoints.
 * Replaces: _integrate.loop (_integrate.py:302-885) with Heun (_solver/heun.py:12-45,
 * _solver/runge_kutta.py:847-1203), MultiTerm(ODETerm, ControlTerm).vf_prod
 * (_term.py:711-722, 420-427) and ConstantStepSize (_step_size_controller/constant.py:30-104). */
int dfx_sde_heun_solve(const dfx_sde_problem* p, const dfx_sde_io* io, void* stream);
/* Evaluate the dense solution (or its time derivative) of every path at its query times:
 * Heun's Hermite polynomial in the step holding t, NaN outside [t0, t1].
 * Replaces: DenseInterpolation.evaluate / derivative (_global_interpolation.py:315-400)
 * with ThirdOrderHermitePolynomialInterpolation (_local_interpolation.py:50-139). */
int dfx_sde_dense_eval(const dfx_sde_problem* p, const dfx_sde_dense_io* io, void* stream);
/* Discrete adjoint of dfx_sde_heun_solve: from ys_bar, the cotangents of y0, theta, mu,
 * sigma, replaying each CK-step segment from its checkpoint.
 * Replaces: reverse-mode through RecursiveCheckpointAdjoint.loop (_adjoint.py:264-312). */
int dfx_sde_heun_adjoint(const dfx_sde_problem* p, const dfx_sde_io* io,
                         const dfx_sde_grad_io* g, void* workspace, size_t workspace_bytes,
                         void* stream);

/* Allocation-free variant of dfx_sde_heun_adjoint.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The per-block partial sums live in a
 * library-owned static device array of DFX_SDE_STATIC_SCRATCH_BYTES on each GPU (held until
 * the process ends) instead of a caller workspace; the caller still passes every input and
 * output.  A problem whose dfx_sde_workspace_bytes exceeds the array returns
 * DFX_E_TOO_LARGE (262,144 paths x 16 components need 6.3 MB).  The current device must be
 * the stream's device; the array's address is looked up on the first call on a device.
 * All calls on one GPU share the array: calls that overlap on one GPU (two streams, two
 * host threads) race.  Every partial is written before it is read, so nothing is cleared
 * between calls.  Results are those of dfx_sde_heun_adjoint bit for bit. */
#define DFX_SDE_STATIC_SCRATCH_BYTES (8u << 20)
int dfx_sde_heun_adjoint_static(const dfx_sde_problem* p, const dfx_sde_io* io,
                                const dfx_sde_grad_io* g, void* stream);

/* ---- neural ODE: dy/dt = MLP(y), d = 64, width 256, depth 2 ------------------------- */
/* f(y) = tanh(W3 softplus(W2 softplus(W1 y + b1) + b2) + b3), W1 [256,64], W2 [256,256],
 * W3 [64,256] (eqx.nn.MLP(64, 64, 256, 2, softplus, final_activation=tanh)).
 * Weights are packed once per set of values into MFMA fragments (dfx_mlp_pack). */
typedef struct dfx_mlp_problem {
  int32_t solver;           /* DFX_TSIT5                                                 */
  int32_t precision;        /* DFX_F32 or DFX_BF16_F32                                   */
  int32_t d, hidden;        /* 64, 256                                                   */
  int32_t n_ts;             /* SaveAt(ts=...) length (ts shared by all instances)        */
  int32_t save_t0, save_t1;
  int32_t max_steps;
  int32_t ckpt_every;       /* C > 0: record accepted grid and every C-th state (adjoint) */
  int32_t force_dtmin;
  int64_t batch;
  double t0, t1, dt0;
  double rtol, atol, pcoeff, icoeff, dcoeff, dtmin, dtmax;
  double factormin, factormax, safety, error_order;
  int32_t waves;            /* wavefronts per 32-trajectory workgroup of dfx_mlp_solve: 4
                               or 8; 0 picks 4 for DFX_F32 and 8 for DFX_BF16_F32 (no
                               effect on the results, which are the same bits)          */
  int32_t save_steps;       /* SaveAt(steps=k): k > 0 also saves every k-th accepted step
                               (rows in time order with the ts rows; forward only)      */
  int64_t ts_stride;        /* 0: ts[n_ts] shared by the batch; n_ts: ts[B, n_ts] per
                               trajectory (a vmapped SaveAt(ts))                        */
  int32_t n_events;         /* terminating Events (_event.py:13-56; <= 4 affine
                               conditions, as for dfx_erk_problem; forward only)         */
  int32_t event_root_find;  /* 1: locate the event time on the step's interpolant       */
  int32_t n_step_ts;        /* ClipStepSizeController (clip.py:120-428) counts, as for
                               dfx_erk_problem (with checkpoints too)                    */
  int32_t n_jump_ts;
  int32_t resume;           /* DFX_RESUME_* mask of dfx_mlp_io.state_in, as
                               dfx_erk_problem.resume (forward only, not with clipping)  */
} dfx_mlp_problem;

typedef struct dfx_mlp_io {
  const float* y0;          /* [B, 64]                                                   */
  const double* ts;         /* [n_ts], or [B, n_ts] with ts_stride = n_ts                */
  float* ys;                /* out [B, n_out, 64]                                        */
  double* ts_out;           /* out [B, n_out]                                            */
  int32_t* stats;           /* out [B, 3]                                                */
  int32_t* result;          /* out [B]                                                   */
  double* grid;             /* out [B, max_steps, 2] when ckpt_every > 0                  */
  float* ckpt;              /* out [B, dfx_mlp_n_ckpt, 64] when ckpt_every > 0            */
  /* SaveAt(dense=True), all three or none (forward only), the layout of dfx_erk_io's:   */
  double* dense_ts;         /* out [B, max_steps + 1]: t0, then each accepted step's end  */
  float* dense_ys;          /* out [B, max_steps + 1, 64]: the states at dense_ts         */
  float* dense_k;           /* out [B, max_steps, 7, 64]: each accepted step's k_i        */
  const double* event_coeffs; /* [n_events, 66]: (c_t, c_y[64], c_0) per condition         */
  const int32_t* event_flags; /* [n_events]: bit 0 boolean, bits 1-2 direction             */
  int32_t* event_mask;        /* out [B, n_events]: the condition that ended the solve     */
  const double* step_ts;      /* [n_step_ts] ascending, shared by the batch, or NULL       */
  const double* jump_ts;      /* [n_jump_ts] ascending, or NULL                            */
  const double* state_in;     /* [B, DFX_MLP_STATE_WORDS] when problem.resume != 0         */
  double* state_out;          /* optional out [B, DFX_MLP_STATE_WORDS]: the final states   */
} dfx_mlp_io;

/* The neural-ODE solver / controller state record: the DFX_ERK_STATE_WORDS layout at
 * d = 64 (the ConstantStepSize words are written as zero). */
#define DFX_MLP_STATE_WORDS DFX_ERK_STATE_WORDS(64)

/* Bytes of the packed-weight buffer for a precision. */
int dfx_mlp_pack_bytes(int32_t precision, size_t* bytes);
/* Pack float32 row-major W1, W2, W3 (device) into MFMA fragments (forward and transposed
 * products) in `packed` (device, dfx_mlp_pack_bytes). */
int dfx_mlp_pack(int32_t precision, const float* w1, const float* w2, const float* w3,
                 void* packed, void* stream);
/* f(y) for `rows` states y [rows, 64] -> f [rows, 64] (device).
 * Replaces: ODETerm.vf (_term.py:193-211) with the MLP vector field. */
int dfx_mlp_eval(int32_t precision, const void* packed, const float* b1, const float* b2,
                 const float* b3, const float* y, float* f, int64_t rows, void* stream);
/* Checkpoint slots per instance: ceil(max_steps / ckpt_every). */
int dfx_mlp_n_ckpt(const dfx_mlp_problem* p);
/* Batched adaptive Tsit5 solve of the neural ODE: one workgroup per 32 trajectories,
 * MLP stages on MFMA, per-trajectory float64 step control.
 * Replaces: the integration loop (_integrate.py:302-885) with ODETerm(MLP) and Tsit5
 * (_solver/tsit5.py, _solver/runge_kutta.py:446-1203) under PIDController
 * (_step_size_controller/pid.py:394-567) and SaveAt(ts) (_integrate.py:456-491). */
int dfx_mlp_solve(const dfx_mlp_problem* p, const void* packed, const float* b1,
                  const float* b2, const float* b3, const dfx_mlp_io* io, void* stream);
/* Saved rows per trajectory of dfx_mlp_solve: t0 + n_ts + (save_steps ? max_steps /
 * save_steps : 0) + t1 (the t1 row is dropped when max_steps is a multiple of save_steps),
 * the allocation rule of _integrate.py:1272-1305. */
int dfx_mlp_n_out(const dfx_mlp_problem* p);
/* The dense solution of a dfx_mlp_solve with dense buffers (or its time derivative) at
 * every trajectory's query times; io->t1 is not read (the problem's t1 is).
 * Replaces: DenseInterpolation.evaluate / derivative (_global_interpolation.py:315-400)
 * with the Tsit5 interpolant (_solver/tsit5.py:103-156), for the neural ODE. */
int dfx_mlp_dense_eval(const dfx_mlp_problem* p, const dfx_erk_dense_io* io, void* stream);

/* ---- neural ODE adjoint ---------------------------------------------------------- */
typedef struct dfx_mlp_adj_io {
  const double* grid;       /* [B, max_steps, 2] from dfx_mlp_solve (ckpt_every > 0)        */
  const float* ckpt;        /* [B, dfx_mlp_n_ckpt, 64] from dfx_mlp_solve                   */
  const int32_t* stats;     /* [B, 3] from dfx_mlp_solve                                    */
  const double* ts;         /* [n_ts] save times                                           */
  const float* ys_bar;      /* [B, n_out, 64] cotangent of the saved states                 */
  float* ybar;              /* in/out [B, 64]: cotangent of the state at the segment's end
                               (in), at its start (out); set to ys_bar of SaveAt(t1) before
                               the last segment; after segment 0 it is dL/dy0 (minus t0 save) */
  float* zs;                /* scratch, dfx_mlp_adj_workspace zs_bytes                      */
  void* rows;               /* scratch, rows_bytes                                          */
  float* partial;           /* scratch, partial_bytes                                       */
  float* dw;                /* in/out accumulator: dW1 | dW2 | dW3 | db1 | db2 | db3 (float32,
                               row-major, 98880 values); zero it before the first segment   */
} dfx_mlp_adj_io;

/* Scratch sizes for a segment of seg_steps (<= ckpt_every) accepted steps. */
int dfx_mlp_adj_workspace(const dfx_mlp_problem* p, int32_t seg_steps, size_t* zs_bytes,
                          size_t* rows_bytes, size_t* partial_bytes);
/* Reverse sweep over accepted steps [segment*C, segment*C + seg_steps) of every
 * instance (C = ckpt_every): recompute from the checkpoint, pull the cotangents back
 * through the saves, the stages and the MLP, and add the segment's weight gradients to
 * io->dw.  Call for segment = ceil(max accepted steps / C) - 1 down to 0.
 * Replaces: reverse-mode differentiation of diffeqsolve through
 * RecursiveCheckpointAdjoint (_adjoint.py:264-312), whose step-size factors are
 * stop_gradient (_step_size_controller/pid.py:498-530). */
int dfx_mlp_adj_segment(const dfx_mlp_problem* p, const void* packed, const float* b1,
                        const float* b2, const float* b3, const dfx_mlp_adj_io* io,
                        int32_t segment, int32_t seg_steps, void* stream);

/* BacksolveAdjoint of the neural-ODE solve with the adjoint seminorm. */
typedef struct dfx_mlp_backsolve_io {
  const float* ys;          /* [B, n_out, 64] the forward's saved states                   */
  const double* ts_out;     /* [B, n_out] their times                                      */
  const float* ys_bar;      /* [B, n_out, 64] cotangent of the saved states                 */
  void* state;              /* scratch, dfx_mlp_backsolve_workspace state_bytes             */
  void* rows;               /* scratch, rows_bytes; zero it once before the first launch   */
  float* partial;           /* scratch, partial_bytes                                      */
  float* dw;                /* in/out accumulator: dW1 | dW2 | dW3 | db1 | db2 | db3
                               (float32, row-major, 98880 values); zero it first           */
  float* ybar;              /* out [B, 64] dL/dy0                                          */
  int32_t* result;          /* out [B] the first failure of the backward solves            */
} dfx_mlp_backsolve_io;

/* Scratch sizes for launches of chunk_iters lock-step iterations. */
int dfx_mlp_backsolve_workspace(const dfx_mlp_problem* back, int32_t chunk_iters,
                                size_t* state_bytes, size_t* rows_bytes,
                                size_t* partial_bytes);
/* Up to chunk_iters attempted steps of every row of the backward solves (first = 1 on the
 * first launch), then the launch's weight gradients added to io->dw.  The products run in
 * float32: `packed` holds the weights packed for DFX_F32 and back->precision is DFX_F32
 * whatever the forward's form.  `back` is the
 * forward's problem with the backward controller's rtol / atol / max_steps and dt0 (the
 * sign the backward solves use: the negated forward dt0 unless BacksolveAdjoint was given
 * its own).  Repeat while the counter from dfx_mlp_backsolve_status (rows unfinished) is
 * nonzero after a launch; a launch after the last does no work.
 * Replaces: _loop_backsolve_bwd (_adjoint.py:581-735) under PIDController(norm=
 * adjoint_rms_seminorm) (_misc.py:85-116) for the MLP vector field (AdjointTerm,
 * _term.py:767-917). */
int dfx_mlp_backsolve(const dfx_mlp_problem* back, const void* packed, const float* b1,
                      const float* b2, const float* b3, const dfx_mlp_backsolve_io* io,
                      int32_t chunk_iters, int32_t first, void* stream);
/* Device address of the two int32 launch counters in the state buffer: rows unfinished
 * after the last launch, and the lock-step iterations it ran. */
int dfx_mlp_backsolve_status(const dfx_mlp_problem* back, int32_t chunk_iters,
                             const void* state, const int32_t** counters);

#ifdef __cplusplus
}
#endif
#endif /* DIFFRAX_AMD_H */
