// erk_mlp.hpp -- persistent Runge-Kutta solve of the neural ODE dy/dt = MLP(y) (d = 64,
// hidden 256) for batches of trajectories, and weight packing for the MFMA fragments.
//
// One workgroup of four or eight waves integrates a tile of 32 trajectories: every stage of
// every step is one MLP evaluation of the whole tile on MFMA (mlp_tile.hpp), while time, step size,
// error norm, accept/reject, saves and counters stay per trajectory (owner threads, float64
// controller as in erk_common.hpp).  Reference behaviour: the same loop as the small-system
// kernel (_integrate.py:302-885, _solver/runge_kutta.py:446-1203,
// _step_size_controller/pid.py:394-567, _solver/tsit5.py:103-156) with the vector field an
// eqx.nn.MLP(64, 64, width 256, depth 2, softplus, final tanh).
//
// Gradient recording: with ckpt_every = C > 0 the kernel also writes the accepted step grid
// and the state at every C-th accepted step; the adjoint (mlp_adj.hip) replays from there.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include "../../include/diffrax_amd.h"
#include "erk_small.hpp"
#include "mlp_tile.hpp"

namespace dfx {

constexpr int MLP_STATE_WORDS = DFX_ERK_STATE_WORDS(MLP_D);

struct MlpArgs {
  const float* y0;
  const double* ts;
  float* ys;
  double* ts_out;
  int32_t* stats;
  int32_t* result;
  double* grid;
  float* ckpt;
  double* dense_ts;  // SaveAt(dense=True): [B, max_steps + 1], [B, max_steps + 1, 64],
  float* dense_ys;   // [B, max_steps, 7, 64] (or null)
  float* dense_k;
  const double* event_coeffs;  // Event: [n_events, 66] affine conditions (c_t, c_y[64], c_0)
  const int32_t* event_flags;  // bit 0 boolean, bits 1-2 direction (erk_small.hpp)
  int32_t* event_mask;         // [B, n_events] out
  const double* step_ts;       // ClipStepSizeController: sorted ascending (or null)
  const double* jump_ts;
  const double* state_in;      // [B, MLP_STATE_WORDS] solver / controller state to resume
  double* state_out;           // [B, MLP_STATE_WORDS] out (or null)
  int64_t B, ts_stride;  // ts_stride: 0 (ts shared) or n_ts (ts [B, n_ts])
  int32_t n_ts, n_out, save_t0, save_t1, max_steps, ckpt_every, n_ckpt, has_dt0, save_steps;
  int32_t n_events, event_root_find, n_step_ts, n_jump_ts, resume;
  double t0, t1, dt0, error_order;
  PidParams pid;
};

// Ctl: PID_INTEGRAL (the integral controller, the default) or PID_GENERIC (any p, i, d
// coefficients).  NW: waves per tile (4 or 8; MlpWaves); each trajectory has 64 / CPT owner
// threads of CPT components each.  The results do not depend on NW.  Ext: the extras in a
// build of their own: terminating Events (_event.py:13-118, forward only) with affine
// conditions, ClipStepSizeController's step_ts / jump_ts (clip.py:120-428; the gradient
// build records the clipped grid like any other) and the solver / controller state export
// and resume (forward only), as the small-system kernel's.  A jump's made_jump re-evaluates the next step's first stage in
// the reference (runge_kutta.py:684-690); here the FSAL stage is kept, which is the same
// value bit for bit: the last stage is evaluated at the new state itself and the field is
// autonomous.
template <class P, class Tab, bool Grad, int Ctl, int NW, bool Ext = false>
__global__ void __launch_bounds__(64 * NW, 1) erk_mlp_kernel(const MlpArgs a, const MlpWeights<P> W) {
  extern __shared__ __align__(16) unsigned char smem[];
  MlpLds<P>& L = *reinterpret_cast<MlpLds<P>*>(smem);
  constexpr int S = Tab::S;
  constexpr int CPT = MlpWaves<NW>::CPT;  // components per owner
  constexpr int GRP = MLP_D / CPT;         // owners per trajectory
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int orow = tid / GRP;             // trajectory within the tile
  const int c0 = (tid % GRP) * CPT;       // owner's first component
  const int gbase = lane & ~(GRP - 1);    // first lane of this trajectory's owner group
  const bool lead = (tid % GRP) == 0;
  const int64_t traj = (int64_t)blockIdx.x * MLP_TILE + orow;
  const bool valid = traj < a.B;

  using TL = TabLit<Tab, float>;  // float32 stage algebra: the coefficients as literals

  // ---- set-up (diffeqsolve / PIDController.init) ----
  float y[CPT], fsal[CPT];
#pragma unroll
  for (int j = 0; j < CPT; ++j) y[j] = valid ? a.y0[traj * MLP_D + c0 + j] : 0.0f;
  const int dir = (a.t0 < a.t1) ? 1 : -1;
  const double T0 = a.t0 * dir, T1 = a.t1 * dir;
  double floor_t1 = T1;
  for (int r = 0; r < 100; ++r) floor_t1 = prevbefore(floor_t1);
  double tprev = T0;
  PidState pid;
  pid.inv = 1.0; pid.pinv = 1.0; pid.at_dtmin = 0;
  int32_t num_steps = 0, num_acc = 0, num_rej = 0, result = DFX_SUCCESSFUL;
  int32_t save_index = 0, ts_index = 0;
  // save times: shared (ts_stride 0) or this trajectory's row (padding rows read row 0)
  const double* ts_row = a.ts ? a.ts + (valid ? traj : 0) * a.ts_stride : nullptr;
  double next_ts = (a.n_ts > 0) ? ts_row[0] * dir : __builtin_inf();
  float* ys_row = valid ? a.ys + traj * (int64_t)a.n_out * MLP_D : nullptr;
  double* tso_row = valid ? a.ts_out + traj * (int64_t)a.n_out : nullptr;
  auto write_save = [&](double tn, const float (&v)[CPT]) {
#pragma unroll
    for (int j = 0; j < CPT; ++j) ys_row[save_index * MLP_D + c0 + j] = v[j];
    if (lead) tso_row[save_index] = tn * dir;
    ++save_index;
  };
  if (valid && a.save_t0) write_save(T0, y);
  // ---- Event: affine conditions c_t t + c_y . y + c_0, summed in component order (the
  // owner group passes the partial sum from owner to owner), one fma per component, so
  // that every wave count rounds them alike ----
  [[maybe_unused]] double ev_old[kMaxEvents];
  [[maybe_unused]] int ev_mask = 0;
  [[maybe_unused]] double tfinal = 0.0;
  [[maybe_unused]] float yfinal[CPT];
  auto event_value = [&](int i, double t, const float (&v)[CPT]) {
    const double* c = a.event_coeffs + i * (MLP_D + 2);
    double part = c[0] * t;
    for (int q = 0; q < GRP; ++q) {
      double mine = part;
#pragma unroll
      for (int j = 0; j < CPT; ++j) mine = __builtin_fma(c[1 + c0 + j], (double)v[j], mine);
      part = __shfl(mine, gbase + q);
    }
    return part + c[1 + MLP_D];
  };
  if constexpr (Ext) {
    if (valid) {
      // values at (t0, y0); a boolean condition that holds ends the solve here
      // (_integrate.py:1333-1426)
      for (int i = 0; i < a.n_events; ++i) {
        const double v = event_value(i, T0, y);
        const bool boolean = a.event_flags[i] & 1;
        ev_old[i] = boolean ? (double)(v > 0.0) : v;
        if (boolean && v > 0.0 && ev_mask == 0) ev_mask = 1 << i;
      }
      if (ev_mask) {
        result = DFX_EVENT_OCCURRED;
        tfinal = T0;
#pragma unroll
        for (int j = 0; j < CPT; ++j) yfinal[j] = y[j];
      }
    }
  }
  if (valid && a.dense_ts) {  // the dense output's first knot: (t0, y0)
    const int64_t row = traj * ((int64_t)a.max_steps + 1);
    if (lead) a.dense_ts[row] = T0;
#pragma unroll
    for (int j = 0; j < CPT; ++j) a.dense_ys[row * MLP_D + c0 + j] = y[j];
  }

  // ---- first stage of the first step (FSAL seed): f(t0, y0) ----
#pragma unroll
  for (int j = 0; j < CPT; ++j) store_input<P>(L, orow, c0 + j, y[j]);
  __syncthreads();
  mlp_forward<P, NW>(W, L, wave, lane, orow, c0, fsal);

  // ---- the first step size: dt0, or _select_initial_step (pid.py:33-81) ----
  // RMS over the 64 components of per-component squares t: eight partial sums of eight
  // components each, in component order, then the eight partials in order.  With four
  // components per owner the odd owner of a pair continues its even neighbour's partial,
  // so both wave counts round identically.
  auto group_rms = [&](const double (&t)[CPT]) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < CPT; ++j) s += t[j];
    int stride = 1, first = 0;
    if constexpr (CPT == 4) {
      s = __shfl(s, lane & ~1);
#pragma unroll
      for (int j = 0; j < CPT; ++j) s += t[j];
      stride = 2;
      first = 1;
    }
    double ssum = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) ssum += __shfl(s, gbase + stride * q + first);
    return sqrt(ssum) / 8.0;
  };
  double dt0v;
  if (a.has_dt0) {
    dt0v = a.dt0 * dir;
  } else {
    double scale[CPT], t0[CPT], t1[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      scale[j] = a.pid.atol + fabs((double)y[j]) * a.pid.rtol;
      const double e0 = (double)y[j] / scale[j], e1 = (double)fsal[j] / scale[j];
      t0[j] = e0 * e0;
      t1[j] = e1 * e1;
    }
    const double d0 = group_rms(t0), d1 = group_rms(t1);
    const bool cond = (d0 < 1e-5) || (d1 < 1e-5);
    const double h0 = cond ? 1e-6 : 0.01 * (d0 / (cond ? 1.0 : d1));
    __syncthreads();  // every wave is done with the seed's layer-3 partials
#pragma unroll
    for (int j = 0; j < CPT; ++j) store_input<P>(L, orow, c0 + j, y[j] + (float)h0 * fsal[j]);
    __syncthreads();
    float f1[CPT];
    mlp_forward<P, NW>(W, L, wave, lane, orow, c0, f1);
    double t2[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      const double e2 = ((double)f1[j] - (double)fsal[j]) / scale[j];
      t2[j] = e2 * e2;
    }
    const double d2 = group_rms(t2) / h0;
    const double max_d = jmax(d1, d2);
    const double h1 = (max_d <= 1e-15) ? jmax(1e-6, h0 * 1e-3)
                                       : pow_ctl<true>(0.01 / max_d, 1.0 / a.error_order);
    dt0v = jmin(100.0 * h0, h1);
  }
  if (a.pid.has_dtmax) dt0v = jmin(dt0v, a.pid.dtmax);
  if (a.pid.has_dtmin) dt0v = jmax(dt0v, a.pid.dtmin);
  double tnext = T0 + dt0v;
  if constexpr (Ext) {  // ClipStepSizeController.init (clip.py:247-302)
    if (a.n_step_ts > 0) {
      const ClipTimes c{a.step_ts, a.n_step_ts, dir};
      tnext = c.clip(tnext, c.find(T0), false);
    }
    if (a.n_jump_ts > 0) {
      const ClipTimes c{a.jump_ts, a.n_jump_ts, dir};
      tnext = c.clip(tnext, c.find(T0), true);
    }
  }
  tnext = jmin(tnext, T1);
  // Resumed solver / controller state (_integrate.py:1252-1270; the record layout of
  // DFX_ERK_STATE_WORDS).  The FSAL stage is the one given unless the first stage is
  // evaluated afresh (first_step or made_jump, runge_kutta.py:684-690), which the seed
  // evaluation above already did.
  [[maybe_unused]] bool first = true, made_jump = false;
  if constexpr (Ext) {
    if (valid && a.resume) {
      const double* r = a.state_in + traj * MLP_STATE_WORDS;
      if (a.resume & DFX_RESUME_MADE_JUMP) made_jump = r[MLP_D + 1] != 0.0;
      if (a.resume & DFX_RESUME_SOLVER) {
        first = r[MLP_D] != 0.0;
        if (!first && !made_jump) {
#pragma unroll
          for (int j = 0; j < CPT; ++j) fsal[j] = (float)r[c0 + j];
        }
      }
      if (a.resume & DFX_RESUME_CONTROLLER) {
        pid.inv = r[MLP_D + 2];
        pid.pinv = r[MLP_D + 3];
        pid.at_dtmin = r[MLP_D + 4] != 0.0;
        if (a.has_dt0) tnext = jmin(T0 + a.dt0 * dir, T1);  // the controller's init skipped
      }
    }
  }

  auto running = [&]() {
    return valid && tprev < T1 && result == DFX_SUCCESSFUL && num_steps < a.max_steps;
  };

  while (__syncthreads_or(running())) {
    const bool run = running();
    const double ts0 = tprev, ts1 = tnext;
    const double dt = ts1 - ts0;
    const float control = (float)((double)dir * dt);
    float k[S][CPT];
    float z[CPT], f[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) k[0][j] = control * fsal[j];
    static_for<1, S>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        float kc[S];
#pragma unroll
        for (int m = 0; m < S; ++m) kc[m] = (m < i) ? k[m][j] : 0.0f;
        z[j] = y[j] + comb<Tab, 0, i, 0, i, false>(TL::v.a[i], kc, 0.0f);
        store_input<P>(L, orow, c0 + j, z[j]);
      }
      __syncthreads();
      mlp_forward<P, NW>(W, L, wave, lane, orow, c0, f);
#pragma unroll
      for (int j = 0; j < CPT; ++j) k[i][j] = control * f[j];
    });
    // y1 = last stage (SSAL); embedded error
    float yerr[CPT];
    double sq[CPT];
    bool anynan = false;
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      float kc[S];
#pragma unroll
      for (int m = 0; m < S; ++m) kc[m] = k[m][j];
      yerr[j] = comb<Tab, 2, 0, 0, S, false>(TL::v.berr, kc, 0.0f);
      anynan = anynan || (z[j] != z[j]);
    }
    // group reduction in owner order (identical in all owners of the trajectory)
    unsigned long long nanmask = __ballot(anynan);
    const bool group_nan = ((nanmask >> gbase) & ((1ull << GRP) - 1)) != 0;
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      const double y1c = group_nan ? (double)y[j] : (double)z[j];
      const double yy = jmax(fabs((double)y[j]), fabs(y1c));
      double e = (double)yerr[j];
      if (e != e) e = __builtin_inf();
      e = e / (a.pid.atol + yy * a.pid.rtol);
      sq[j] = e * e;
    }
    const double scaled = group_rms(sq);
    double ntprev, ntnext;
    int code = 0;
    PidState npid = pid;
    bool keep = pid_adapt<Ctl>(a.pid, scaled, ts0, ts1, npid, ntprev, ntnext, code);
    [[maybe_unused]] bool mj = false;  // made_jump of this step: only a jump time sets it
    if constexpr (Ext) {  // ClipStepSizeController.adapt_step_size (clip.py:304-428)
      if (a.n_step_ts > 0) {
        const ClipTimes c{a.step_ts, a.n_step_ts, dir};
        bool unused;
        ntnext = c.clip(ntnext, c.find(c.bump(ntprev, unused)), false);
      }
      if (a.n_jump_ts > 0) {
        const ClipTimes c{a.jump_ts, a.n_jump_ts, dir};
        ntprev = c.bump(ntprev, mj);
        ntnext = jmax(nextup(ntprev), ntnext);
        ntnext = c.clip(ntnext, c.find(ntprev), true);
      }
    }
    ntprev = jmin(ntprev, T1);
    if (ntnext > floor_t1) ntnext = keep ? T1 : ntprev + 0.5 * (T1 - ntprev);
    if (run) {
      pid = npid;
      ++num_steps;
      if (keep) ++num_acc; else ++num_rej;
      if (result == DFX_SUCCESSFUL) result = code;
      bool event_now = false;
      if constexpr (Ext) {
        // Event check at the new state (_integrate.py:542-633); with a root finder the time
        // is located on this step's interpolant by bisection down to adjacent floats, as
        // the small-system kernel does (event_root), and saves after it are dropped
        int mask = 0, fired = -1;
        for (int i = 0; i < a.n_events; ++i) {
          float yn[CPT];
#pragma unroll
          for (int j = 0; j < CPT; ++j) yn[j] = keep ? z[j] : y[j];
          const double v = event_value(i, ntprev, yn);
          const int flags = a.event_flags[i];
          const bool trig = event_triggered(flags, ev_old[i], v);
          if (trig && mask == 0) { mask = 1 << i; fired = i; }
          ev_old[i] = (flags & 1) ? (double)(v > 0.0) : v;
        }
        ev_mask = mask;
        if (mask) {
          event_now = true;
          result = DFX_EVENT_OCCURRED;
          auto interp_at = [&](double t, float (&v)[CPT]) {
            const double th = linear_rescale(ts0, t, ts1);
            double bb[7];
            tsit5_weights(th, bb);
#pragma unroll
            for (int j = 0; j < CPT; ++j) {
              float acc = (float)bb[0] * k[0][j];
#pragma unroll
              for (int m = 1; m < 7; ++m) acc = acc + (float)bb[m] * k[m][j];
              v[j] = y[j] + acc;
            }
          };
          if (a.event_root_find) {
            double troot = ts1;
            if (!(a.event_flags[fired] & 1)) {
              float ya[CPT], yb[CPT], ym[CPT];
              interp_at(ts0, ya);
              interp_at(ts1, yb);
              double ga = event_value(fired, ts0, ya), gb = event_value(fired, ts1, yb);
              if (event_sign(ga) != event_sign(gb) && ga == ga && gb == gb) {
                double lo = ts0, hi = ts1;
                troot = -1.0;
                for (int it = 0; it < 2100; ++it) {
                  const double mid = lo + 0.5 * (hi - lo);
                  if (!(mid > lo && mid < hi)) break;
                  interp_at(mid, ym);
                  const double gm = event_value(fired, mid, ym);
                  if (gm == 0.0) { troot = mid; break; }
                  if (event_sign(gm) == event_sign(ga)) { lo = mid; ga = gm; } else { hi = mid; gb = gm; }
                }
                if (troot < 0.0) troot = (fabs(ga) <= fabs(gb)) ? lo : hi;
              }
            }
            tfinal = troot;
            interp_at(tfinal, yfinal);
          } else {
            tfinal = ntprev;
#pragma unroll
            for (int j = 0; j < CPT; ++j) yfinal[j] = keep ? z[j] : y[j];
          }
        }
      }
      if (keep) {
        if constexpr (Grad) {
          const int n = num_acc - 1;
          if (lead) {
            a.grid[(traj * a.max_steps + n) * 2 + 0] = ts0;
            a.grid[(traj * a.max_steps + n) * 2 + 1] = ts1;
          }
          if (n % a.ckpt_every == 0 && n / a.ckpt_every < a.n_ckpt) {
            float* c = a.ckpt + (traj * a.n_ckpt + n / a.ckpt_every) * MLP_D + c0;
#pragma unroll
            for (int j = 0; j < CPT; ++j) c[j] = y[j];
          }
        }
        if (a.dense_ts) {
          // SaveAt(dense=True): the step's end time, end state and stage increments
          // (_integrate.py:529-540)
          const int64_t row = traj * ((int64_t)a.max_steps + 1) + num_acc;
          if (lead) a.dense_ts[row] = ntprev;
          float* dy = a.dense_ys + row * MLP_D + c0;
          float* dk = a.dense_k + (traj * (int64_t)a.max_steps + (num_acc - 1)) * S * MLP_D + c0;
#pragma unroll
          for (int j = 0; j < CPT; ++j) dy[j] = z[j];
#pragma unroll
          for (int m = 0; m < S; ++m) {
#pragma unroll
            for (int j = 0; j < CPT; ++j) dk[m * MLP_D + j] = k[m][j];
          }
        }
        while (next_ts <= ts1 && !(Ext && event_now && next_ts > tfinal)) {
          const double th = linear_rescale(ts0, next_ts, ts1);
          double bb[7];
          tsit5_weights(th, bb);
          float v[CPT];
#pragma unroll
          for (int j = 0; j < CPT; ++j) {
            float acc = (float)bb[0] * k[0][j];
#pragma unroll
            for (int m = 1; m < 7; ++m) acc = acc + (float)bb[m] * k[m][j];
            v[j] = y[j] + acc;
          }
          write_save(next_ts, v);
          ++ts_index;
          next_ts = (ts_index < a.n_ts) ? ts_row[ts_index] * dir : __builtin_inf();
        }
#pragma unroll
        for (int j = 0; j < CPT; ++j) { y[j] = z[j]; fsal[j] = f[j]; }
        if constexpr (Ext) {  // an accepted step takes the step's made_jump (_integrate.py:425)
          first = false;
          made_jump = mj;
        }
        // SaveAt(steps=k): every k-th accepted step (_integrate.py:493-527)
        if (a.save_steps && num_acc % a.save_steps == 0 && !(Ext && event_now && ntprev > tfinal))
          write_save(ntprev, y);
      }
      tprev = ntprev;
      tnext = ntnext;
    }
  }
  if (!valid) return;
  if constexpr (Ext) {
    if (a.state_out) {  // the final states (_integrate.py:1489-1500), as erk_small.hpp writes them
      double* r = a.state_out + traj * MLP_STATE_WORDS;
      const bool resumed_f0 = (a.resume & DFX_RESUME_SOLVER) != 0;
#pragma unroll
      for (int j = 0; j < CPT; ++j)
        r[c0 + j] = !first ? (double)fsal[j]
                           : (resumed_f0 ? a.state_in[traj * MLP_STATE_WORDS + c0 + j] : 0.0);
      if (lead) {
        r[MLP_D] = first ? 1.0 : 0.0;
        r[MLP_D + 1] = made_jump ? 1.0 : 0.0;
        r[MLP_D + 2] = pid.inv;
        r[MLP_D + 3] = pid.pinv;
        r[MLP_D + 4] = pid.at_dtmin ? 1.0 : 0.0;
        for (int w = MLP_D + 5; w < MLP_STATE_WORDS; ++w) r[w] = 0.0;  // ConstantStepSize's
      }
    }
  }
  if (T0 == T1) {
    for (int r = 0; r < a.n_ts; ++r) write_save(T0, y);
  }
  if (Ext && result == DFX_EVENT_OCCURRED) {  // the solve ends at the event
    tprev = tfinal;
#pragma unroll
    for (int j = 0; j < CPT; ++j) y[j] = yfinal[j];
  }
  // _save_t1 (_integrate.py:847-877): no t1 row when the last step was saved via steps,
  // except with an event root finder (that row, after the event time, was removed)
  const bool via_steps = a.save_steps && num_acc % a.save_steps == 0;
  const bool root = Ext && a.n_events > 0 && a.event_root_find;
  const bool t1_pred = root ? (a.save_t1 || via_steps) : (a.save_t1 && !via_steps);
  if (t1_pred && save_index < a.n_out) write_save(tprev, y);
  if (tprev < T1 && result == DFX_SUCCESSFUL) result = DFX_MAX_STEPS_REACHED;
  for (int s = save_index; s < a.n_out; ++s) {
#pragma unroll
    for (int j = 0; j < CPT; ++j) ys_row[s * MLP_D + c0 + j] = __builtin_inff();
    if (lead) tso_row[s] = __builtin_inf();  // (dir * inf) * dir: _integrate.py:1296, 1481
  }
  if (Ext && lead) {
    for (int i = 0; i < a.n_events; ++i) a.event_mask[traj * a.n_events + i] = (ev_mask >> i) & 1;
  }
  if (lead) {
    a.stats[traj * 3 + 0] = num_steps;
    a.stats[traj * 3 + 1] = num_acc;
    a.stats[traj * 3 + 2] = num_rej;
    a.result[traj] = result;
  }
}

// Launch of erk_mlp_kernel with NW waves per tile for the gradient-recording choice and
// the controller build (the integral controller's own or the generic PID one).
template <class P, int NW>
void launch_erk_mlp(const MlpArgs& a, const MlpWeights<P>& W, bool grad, bool integral,
                    unsigned nblocks, hipStream_t s) {
  auto go = [&](auto kern) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)sizeof(MlpLds<P>));
    hipLaunchKernelGGL(kern, dim3(nblocks), dim3(64 * NW), sizeof(MlpLds<P>), s, a, W);
  };
  if (a.n_events > 0 || a.n_step_ts > 0 || a.n_jump_ts > 0 || a.resume || a.state_out) {
    if (grad) {  // step clipping only: events are forward only
      if (integral) go(&erk_mlp_kernel<P, Tsit5Tab, true, PID_INTEGRAL, NW, true>);
      else go(&erk_mlp_kernel<P, Tsit5Tab, true, PID_GENERIC, NW, true>);
    } else {
      if (integral) go(&erk_mlp_kernel<P, Tsit5Tab, false, PID_INTEGRAL, NW, true>);
      else go(&erk_mlp_kernel<P, Tsit5Tab, false, PID_GENERIC, NW, true>);
    }
  } else if (grad) {
    if (integral) go(&erk_mlp_kernel<P, Tsit5Tab, true, PID_INTEGRAL, NW>);
    else go(&erk_mlp_kernel<P, Tsit5Tab, true, PID_GENERIC, NW>);
  } else {
    if (integral) go(&erk_mlp_kernel<P, Tsit5Tab, false, PID_INTEGRAL, NW>);
    else go(&erk_mlp_kernel<P, Tsit5Tab, false, PID_GENERIC, NW>);
  }
}

// One translation unit per policy (erk_mlp_f32.hip, erk_mlp_bf16.hip) instantiates both
// wave counts; waves is 4 or 8.
void launch_erk_mlp_f32(int waves, const MlpArgs& a, const MlpWeights<PolicyF32>& W, bool grad,
                        bool integral, unsigned nblocks, hipStream_t s);
void launch_erk_mlp_bf16(int waves, const MlpArgs& a, const MlpWeights<PolicyBF16x3>& W,
                         bool grad, bool integral, unsigned nblocks, hipStream_t s);

}  // namespace dfx
