// api.hip -- C ABI of libdiffrax_amd (declared in include/diffrax_amd.h).
//
// Host-side argument checking and kernel selection.  No device memory is allocated
// here; every buffer, including scratch, comes from the caller.
#include <math.h>
#include <string.h>

#include "../../include/diffrax_amd.h"
#include "erk_small.hpp"
#include "erk_dense.hpp"
#include "erk_adjoint.hpp"

namespace dfx {
int erk_small_solve_f64(int solver, int vf, int d, const ErkKArgs& a, int block, int64_t lanes,
                        hipStream_t s);
int erk_small_solve_f32(int solver, int vf, int d, const ErkKArgs& a, int block, int64_t lanes,
                        hipStream_t s);
int erk_backsolve_f64(int solver, int vf, int d, const ErkKArgs& a, const BacksolveArgs& b,
                      hipStream_t s);
int erk_backsolve_f32(int solver, int vf, int d, const ErkKArgs& a, const BacksolveArgs& b,
                      hipStream_t s);
}  // namespace dfx

using namespace dfx;

namespace {

int solver_order(int solver) { return solver == DFX_HEUN ? 2 : 5; }

bool small_vf_supported(int vf, int d) {
  if (vf == DFX_VF_LV || vf == DFX_VF_VDP) return d == 2;
  if (vf == DFX_VF_LINEAR) return d == 1 || d == 2 || d == 3 || d == 4 || d == 10;
  return false;
}

// Fill PidParams the way PIDController.adapt_step_size derives its exponents
// (pid.py:505-510): coeff1 = (i + p + d)/order, coeff2 = -(p + 2d)/order, coeff3 = d/order.
PidParams make_pid(const dfx_erk_problem* p, double error_order) {
  PidParams q;
  memset(&q, 0, sizeof(q));
  q.rtol = p->rtol;
  q.atol = p->atol;
  q.coeff1 = (p->icoeff + p->pcoeff + p->dcoeff) / error_order;
  q.coeff2 = -(p->pcoeff + 2.0 * p->dcoeff) / error_order;
  q.coeff3 = p->dcoeff / error_order;
  q.dtmin = p->dtmin;
  q.dtmax = p->dtmax;
  q.has_dtmin = !isnan(p->dtmin);
  q.has_dtmax = !isnan(p->dtmax);
  q.force_dtmin = p->force_dtmin;
  q.factormin = p->factormin;
  q.factormax = p->factormax;
  q.safety = p->safety;
  return q;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

int32_t ckpt_count(const dfx_erk_problem* p) {
  return (int32_t)(((int64_t)p->max_steps + p->ckpt_every - 1) / p->ckpt_every);
}

// Checks a small-system problem and the input buffers of its dfx_erk_io, and fills the
// kernel arguments the forward solve and its discrete adjoint share (no outputs).
int fill_erk_inputs(const dfx_erk_problem* p, const dfx_erk_io* io, ErkKArgs& a) {
  if (!p || !io) return DFX_E_ARG;
  if (p->solver < DFX_TSIT5 || p->solver > DFX_HEUN) return DFX_E_ARG;
  if (p->controller != DFX_PID && p->controller != DFX_CONSTANT) return DFX_E_ARG;
  if (p->precision != DFX_F64 && p->precision != DFX_F32S_F64T) return DFX_E_ARG;
  if (!small_vf_supported(p->vf, p->d)) return DFX_E_ARG;
  if (p->batch < 0 || p->n_ts < 0 || p->max_steps < 0 || p->save_steps < 0) return DFX_E_ARG;
  if (p->t1_stride < 0 || p->ts_stride < 0 || p->param_stride < 0) return DFX_E_ARG;
  if (p->controller == DFX_CONSTANT && isnan(p->dt0)) return DFX_E_ARG;
  if (p->ckpt_every < 0) return DFX_E_ARG;
  if (p->n_step_ts < 0 || p->n_jump_ts < 0) return DFX_E_ARG;
  if ((p->n_step_ts > 0 || p->n_jump_ts > 0) && p->controller != DFX_PID) return DFX_E_ARG;
  if (p->n_events < 0 || p->n_events > kMaxEvents) return DFX_E_ARG;
  if ((p->resume & ~(DFX_RESUME_SOLVER | DFX_RESUME_CONTROLLER | DFX_RESUME_MADE_JUMP)) != 0)
    return DFX_E_ARG;
  if ((p->resume & DFX_RESUME_CONTROLLER) && (p->n_step_ts > 0 || p->n_jump_ts > 0))
    return DFX_E_ARG;
  memset(&a, 0, sizeof(a));
  if (p->batch == 0) return DFX_OK;
  if (!io->t1 || !io->params) return DFX_E_ARG;
  if (p->n_ts > 0 && !io->ts) return DFX_E_ARG;
  if ((p->n_step_ts > 0 && !io->step_ts) || (p->n_jump_ts > 0 && !io->jump_ts)) return DFX_E_ARG;
  if (p->n_events > 0 && (!io->event_coeffs || !io->event_flags)) return DFX_E_ARG;
  const double error_order = isnan(p->error_order) ? (double)solver_order(p->solver) : p->error_order;
  a.t1 = io->t1;
  a.ts = p->n_ts > 0 ? io->ts : nullptr;  // no save times: the kernels read t1 instead
  a.params = io->params;
  a.rtol_b = io->rtol;
  a.atol_b = io->atol;
  a.step_ts = io->step_ts;
  a.jump_ts = io->jump_ts;
  a.n_step_ts = p->n_step_ts;
  a.n_jump_ts = p->n_jump_ts;
  a.event_coeffs = io->event_coeffs;
  a.event_flags = io->event_flags;
  a.n_events = p->n_events;
  a.event_root_find = p->event_root_find ? 1 : 0;
  a.B = p->batch;
  a.t1_stride = p->t1_stride;
  a.ts_stride = p->ts_stride;
  a.param_stride = p->param_stride;
  a.n_ts = p->n_ts;
  a.n_out = dfx_erk_n_out(p);
  a.save_t0 = p->save_t0;
  a.save_t1 = p->save_t1;
  a.save_steps = p->save_steps;
  a.max_steps = p->max_steps;
  a.has_dt0 = !isnan(p->dt0);
  a.controller = p->controller;
  a.t0 = p->t0;
  a.dt0 = p->dt0;
  a.error_order = error_order;
  a.pid = make_pid(p, error_order);
  return DFX_OK;
}

}  // namespace

extern "C" {

int dfx_version(void) { return 100; }

const char* dfx_strerror(int code) {
  switch (code) {
    case DFX_OK: return "success";
    case DFX_E_ARG: return "invalid argument or unsupported solver/vector-field/precision combination";
    case DFX_E_HIP: return "HIP runtime error";
    case DFX_E_WORKSPACE: return "workspace missing or too small";
    case DFX_E_TOO_LARGE: return "problem needs more scratch than the library's static array holds";
    default: return "unknown error";
  }
}

int dfx_erk_n_out(const dfx_erk_problem* p) {
  if (!p) return DFX_E_ARG;
  int n = 0;
  if (p->save_t0) n += 1;
  n += p->n_ts;
  // SaveAt(steps=k): every k-th accepted step; the t1 row is kept unless max_steps is a
  // multiple of k (_integrate.py:1272-1305)
  if (p->save_steps > 0) n += p->max_steps / p->save_steps;
  if (p->save_t1 && (p->save_steps <= 0 || p->max_steps % p->save_steps != 0)) n += 1;
  return n;
}

int64_t dfx_vf_param_count(int32_t vf, int32_t d, int32_t hidden) {
  const int64_t D = d, H = hidden;
  switch (vf) {
    case DFX_VF_LV: return 4;
    case DFX_VF_VDP: return 1;
    case DFX_VF_LINEAR: return D * D;
    case DFX_VF_MLP: return H * D + H + H * H + H + D * H + D;
    case DFX_VF_OU: return 3 * D;
    default: return DFX_E_ARG;
  }
}

int dfx_tableau(int32_t solver, double* a, double* b, double* b_err, double* c, int32_t* info) {
  if (!a || !b || !b_err || !c || !info) return DFX_E_ARG;
  for (int i = 0; i < 49; ++i) a[i] = 0.0;
  for (int i = 0; i < 7; ++i) b[i] = b_err[i] = c[i] = 0.0;
  auto fill = [&](auto tab) {
    using Tab = decltype(tab);
    for (int i = 0; i < Tab::S; ++i) {
      for (int j = 0; j < Tab::S; ++j) a[i * 7 + j] = Tab::a[i][j];
      b[i] = Tab::b[i];
      b_err[i] = Tab::berr[i];
      c[i] = Tab::c[i];
    }
    info[0] = Tab::S;
    info[1] = Tab::fsal;
    info[2] = Tab::ssal;
    info[3] = Tab::order;
  };
  switch (solver) {
    case DFX_TSIT5: fill(Tsit5Tab{}); return DFX_OK;
    case DFX_DOPRI5: fill(Dopri5Tab{}); return DFX_OK;
    case DFX_HEUN: fill(HeunTab{}); return DFX_OK;
    default: return DFX_E_ARG;
  }
}

int dfx_erk_workspace_bytes(const dfx_erk_problem* p, size_t* bytes) {
  if (!p || !bytes) return DFX_E_ARG;
  *bytes = 256;  // work counter of the persistent kernel
  return DFX_OK;
}

int dfx_erk_ckpt_layout(const dfx_erk_problem* p, int32_t* n_ckpt, int32_t* words) {
  if (!p || !n_ckpt || !words || p->ckpt_every <= 0 || p->max_steps < 0 || p->d < 1)
    return DFX_E_ARG;
  *n_ckpt = ckpt_count(p);
  *words = p->d + 8;
  return DFX_OK;
}

int dfx_erk_adjoint_workspace_bytes(const dfx_erk_problem* p, size_t* bytes) {
  if (!p || !bytes || p->ckpt_every <= 0 || p->batch < 0 || p->d < 1) return DFX_E_ARG;
  const size_t recs = (size_t)p->batch * 2 * (size_t)p->ckpt_every;  // two segments each
  const size_t elt = p->precision == DFX_F64 ? 8 : 4;
  *bytes = align256(recs * 2 * sizeof(double)) + align256(recs * (size_t)p->d * elt);
  if (*bytes < 256) *bytes = 256;
  return DFX_OK;
}

int dfx_erk_solve(const dfx_erk_problem* p, const dfx_erk_io* io, void* workspace,
                  size_t workspace_bytes, const dfx_launch* launch, void* stream) {
  ErkKArgs a;
  const int rc = fill_erk_inputs(p, io, a);
  if (rc != DFX_OK) return rc;
  if (p->batch == 0) return DFX_OK;
  if (!io->y0 || !io->stats || !io->result) return DFX_E_ARG;
  if (a.n_out > 0 && (!io->ys || !io->ts_out)) return DFX_E_ARG;
  if (!workspace || workspace_bytes < 256) return DFX_E_WORKSPACE;
  a.y0 = io->y0;
  a.ys = io->ys;
  a.ts_out = io->ts_out;
  a.stats = io->stats;
  a.result = io->result;
  a.grid = io->grid;
  if (io->grid_y && !io->grid) return DFX_E_ARG;
  a.grid_y = io->grid_y;
  a.grid_stride = p->max_steps;
  const int dense = (io->dense_ts != nullptr) + (io->dense_ys != nullptr) + (io->dense_k != nullptr);
  if (dense != 0 && dense != 3) return DFX_E_ARG;
  a.dense_ts = io->dense_ts;
  a.dense_ys = io->dense_ys;
  a.dense_k = io->dense_k;
  a.event_mask = io->event_mask;
  if (p->ckpt_every > 0) {
    if (!io->ckpt) return DFX_E_ARG;
    a.ckpt = io->ckpt;
    a.ckpt_every = p->ckpt_every;
    a.n_ckpt = ckpt_count(p);
  }
  a.counter = reinterpret_cast<unsigned int*>(workspace);
  if (p->resume != 0) {
    if (!io->state_in || p->ckpt_every > 0) return DFX_E_ARG;
    a.state_in = io->state_in;
    a.resume = p->resume;
  }
  a.state_out = io->state_out;

  int block = 64;
  int64_t lanes = p->batch;
  if (launch) {
    if (launch->block > 0) block = launch->block;
    if (launch->lanes > 0) lanes = launch->lanes;
  }
  if (block % 64 != 0 || block > 256) return DFX_E_ARG;
  if (lanes > p->batch) lanes = p->batch;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (p->precision == DFX_F64) return erk_small_solve_f64(p->solver, p->vf, p->d, a, block, lanes, s);
  return erk_small_solve_f32(p->solver, p->vf, p->d, a, block, lanes, s);
}

int dfx_erk_backsolve(const dfx_erk_problem* p, const dfx_erk_backsolve_io* io, void* stream) {
  if (!p || !io) return DFX_E_ARG;
  if (p->solver < DFX_TSIT5 || p->solver > DFX_HEUN) return DFX_E_ARG;
  if (p->precision != DFX_F64 && p->precision != DFX_F32S_F64T) return DFX_E_ARG;
  const bool vf_ok = (p->vf == DFX_VF_LV || p->vf == DFX_VF_VDP) ? p->d == 2
                     : (p->vf == DFX_VF_LINEAR && p->d >= 1 && p->d <= 3);
  if (!vf_ok) return DFX_E_ARG;
  if (p->controller != DFX_PID && p->controller != DFX_CONSTANT) return DFX_E_ARG;
  if (p->controller == DFX_CONSTANT && isnan(p->dt0)) return DFX_E_ARG;
  if (p->save_steps || p->n_step_ts > 0 || p->n_jump_ts > 0 || p->n_events > 0) return DFX_E_ARG;
  if (p->resume != 0) return DFX_E_ARG;
  if (p->batch < 0 || p->n_ts < 0 || p->max_steps < 0 || p->param_stride < 0) return DFX_E_ARG;
  if (p->batch == 0) return DFX_OK;
  if (!io->params || !io->ys_bar || !io->y0_bar || !io->params_bar || !io->result) return DFX_E_ARG;
  const int n_out = dfx_erk_n_out(p);
  if (n_out > 0 && (!io->ys || !io->ts_out)) return DFX_E_ARG;
  const double error_order = isnan(p->error_order) ? (double)solver_order(p->solver) : p->error_order;
  ErkKArgs a;
  memset(&a, 0, sizeof(a));
  a.params = io->params;
  a.B = p->batch;
  a.param_stride = p->param_stride;
  a.max_steps = p->max_steps;
  a.has_dt0 = !isnan(p->dt0);
  a.dt0 = -p->dt0;                  // the backward solves take -dt0 (_adjoint.py:629)
  a.controller = p->controller;
  a.error_order = error_order;
  a.pid = make_pid(p, error_order);
  // the adjoint kernels carry no controller state across intervals: integral PID only
  if (a.controller == DFX_PID && (!pid_is_integral(a.pid) || a.pid.has_dtmin)) return DFX_E_ARG;
  BacksolveArgs b;
  memset(&b, 0, sizeof(b));
  b.ys = io->ys;
  b.ts_out = io->ts_out;
  b.ys_bar = io->ys_bar;
  b.y0_bar = io->y0_bar;
  b.q_bar = io->params_bar;
  b.result = io->result;
  b.n_out = n_out;
  b.save_t0 = p->save_t0;
  b.t0 = p->t0;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (p->precision == DFX_F64) return erk_backsolve_f64(p->solver, p->vf, p->d, a, b, s);
  return erk_backsolve_f32(p->solver, p->vf, p->d, a, b, s);
}

int dfx_erk_adjoint(const dfx_erk_problem* p, const dfx_erk_io* fwd,
                    const dfx_erk_adjoint_io* io, void* workspace, size_t workspace_bytes,
                    void* stream) {
  if (!p || !fwd || !io) return DFX_E_ARG;
  const bool vf_ok = (p->vf == DFX_VF_LV || p->vf == DFX_VF_VDP) ? p->d == 2
                     : (p->vf == DFX_VF_LINEAR && p->d >= 1 && p->d <= 4);
  if (!vf_ok) return DFX_E_ARG;
  if (p->n_events > 0 || p->ckpt_every <= 0 || p->resume != 0) return DFX_E_ARG;
  ErkKArgs a;
  const int rc = fill_erk_inputs(p, fwd, a);
  if (rc != DFX_OK) return rc;
  if (p->batch == 0) return DFX_OK;
  if (!io->ckpt || !io->stats || !io->ys_bar || !io->y0_bar || !io->params_bar) return DFX_E_ARG;
  size_t need = 0;
  dfx_erk_adjoint_workspace_bytes(p, &need);
  if (!workspace || workspace_bytes < need) return DFX_E_WORKSPACE;
  a.ckpt_every = p->ckpt_every;
  a.n_ckpt = ckpt_count(p);
  // The replay's arguments: no saves, no checkpoint writes; accepted steps recorded into
  // the per-instance scratch of two segments (2 ckpt_every records).
  ErkKArgs ra = a;
  ra.ts = nullptr;
  ra.n_ts = 0;
  ra.n_out = 0;
  ra.save_t0 = ra.save_t1 = ra.save_steps = 0;
  ra.ckpt = nullptr;
  const size_t recs = (size_t)p->batch * 2 * (size_t)p->ckpt_every;
  ra.grid = reinterpret_cast<double*>(workspace);
  ra.grid_y = reinterpret_cast<char*>(workspace) + align256(recs * 2 * sizeof(double));
  ra.grid_stride = 2 * (int64_t)p->ckpt_every;
  DiscreteAdjArgs b;
  b.ckpt = io->ckpt;
  b.stats = io->stats;
  b.ys_bar = io->ys_bar;
  b.y0_bar = io->y0_bar;
  b.q_bar = io->params_bar;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (p->precision == DFX_F64) return erk_adjoint_f64(p->solver, p->vf, p->d, a, ra, b, s);
  return erk_adjoint_f32(p->solver, p->vf, p->d, a, ra, b, s);
}

int dfx_erk_dense_eval(const dfx_erk_problem* p, const dfx_erk_dense_io* io, void* stream) {
  if (!p || !io) return DFX_E_ARG;
  if (p->solver < DFX_TSIT5 || p->solver > DFX_HEUN) return DFX_E_ARG;
  if (p->precision != DFX_F64 && p->precision != DFX_F32S_F64T) return DFX_E_ARG;
  if (!small_vf_supported(p->vf, p->d)) return DFX_E_ARG;
  if (p->batch < 0 || p->max_steps < 0 || p->t1_stride < 0) return DFX_E_ARG;
  if (io->n_q < 0 || io->tq_stride < 0) return DFX_E_ARG;
  if (io->tq_stride != 0 && io->tq_stride < io->n_q) return DFX_E_ARG;
  if (p->batch == 0 || io->n_q == 0) return DFX_OK;
  if (!io->dense_ts || !io->dense_ys || !io->dense_k || !io->stats || !io->t1 || !io->tq ||
      !io->out)
    return DFX_E_ARG;
  DenseKArgs a;
  memset(&a, 0, sizeof(a));
  a.dense_ts = io->dense_ts;
  a.dense_ys = io->dense_ys;
  a.dense_k = io->dense_k;
  a.stats = io->stats;
  a.t1 = io->t1;
  a.tq = io->tq;
  a.out = io->out;
  a.B = p->batch;
  a.n_q = io->n_q;
  a.tq_stride = io->tq_stride;
  a.t1_stride = p->t1_stride;
  a.max_steps = p->max_steps;
  a.left = io->left ? 1 : 0;
  a.derivative = io->derivative ? 1 : 0;
  a.t0 = p->t0;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  return p->precision == DFX_F64 ? erk_dense_eval_f64(p->solver, p->d, a, s)
                                 : erk_dense_eval_f32(p->solver, p->d, a, s);
}

}  // extern "C"
