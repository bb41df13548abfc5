// erk_small.hpp -- persistent one-lane-per-trajectory ERK kernel for small systems.
//
// Each lane owns one trajectory and runs the whole diffrax integration loop for it
// (_integrate.py:302-885) in registers: FSAL/SSAL stages, embedded error, PID
// accept/reject with end clipping, SaveAt(ts) dense-output saves and the final
// bookkeeping.  A lane whose trajectory finishes writes its results and claims the next
// trajectory id from a work counter (one atomic per wave), so a launch with fewer lanes
// than instances keeps every wave busy until the queue drains.  Lanes of a wave stay in
// lock-step only at the instruction level: accept/reject, save counts and refills are
// per lane.
#pragma once
#include <type_traits>

#include "erk_common.hpp"
#include "vector_fields.hpp"
#include "../../include/diffrax_amd.h"

namespace dfx {

struct ErkKArgs {
  const void* y0;
  const double* t1;
  const double* ts;
  const void* params;
  void* ys;
  double* ts_out;
  int32_t* stats;
  int32_t* result;
  double* grid;       // [B, grid_stride, 2] normalised (t_start, t_end) of accepted steps
  void* grid_y;       // [B, grid_stride, D] start state of each accepted step (with grid)
  double* ckpt;       // [B, n_ckpt, D + 8] lane state every ckpt_every accepted steps
                      // (ErkLane::write_ckpt), the restart points of the discrete adjoint
  const double* rtol_b;  // per-instance PIDController tolerances [B], or null for pid.rtol
  const double* atol_b;
  double* dense_ts;   // SaveAt(dense=True): [B, max_steps + 1] accepted-step ends (t0 first)
  void* dense_ys;     // [B, max_steps + 1, D] states at those times
  void* dense_k;      // [B, max_steps, S, D] stage increments of each accepted step
  const double* step_ts;  // ClipStepSizeController times, ascending (shared by the batch)
  const double* jump_ts;
  const double* event_coeffs;  // Event: [n_events, D + 2] affine conditions (c_t, c_y, c_0)
  const int32_t* event_flags;  // bit 0 boolean, bits 1-2 direction (0 any, 1 up, 2 down)
  int32_t* event_mask;         // [B, n_events] out
  const double* state_in;      // [B, D + 9] solver / controller state to resume from
  double* state_out;           // [B, D + 9] out: the final state (same layout)
  unsigned int* counter;
  int64_t B, t1_stride, ts_stride, param_stride;
  int64_t grid_stride;  // records per instance in grid / grid_y
  int32_t ckpt_every, n_ckpt, resume;  // resume: DFX_RESUME_* parts of state_in taken
  int32_t n_step_ts, n_jump_ts, n_events, event_root_find;
  int32_t n_ts, n_out, save_t0, save_t1, save_steps, max_steps, has_dt0, controller;
  double t0, dt0, error_order;
  PidParams pid;
};

// A float64 value the compiler cannot constant-fold: x + (a VGPR zero written by inline
// asm).  Tableau coefficients pass through it once per trajectory so they live in VGPRs
// (plentiful at one wave per SIMD) instead of being hoisted into SGPRs, which overflow
// and spill through v_readlane on every use.  x + 0 == x for every nonzero x.
__device__ __forceinline__ double in_vgpr(double x) {
  double z;
  asm volatile("v_mov_b64 %0, 0" : "=v"(z));
  return x + z;
}
__device__ __forceinline__ float in_vgpr(float x) {
  float z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  return x + z;
}

// Per-lane copies of a tableau's nonzero coefficients (zeros stay compile-time zeros), as
// C (float64, or the float32 values of the coefficients for float32 stage algebra).
template <class Tab, class C = double>
struct TabRegs {
  static constexpr int S = Tab::S;
  C a[S][S];
  C b[S];
  C berr[S];
  __device__ __forceinline__ void load() {
#pragma unroll
    for (int i = 0; i < S; ++i) {
#pragma unroll
      for (int m = 0; m < S; ++m) a[i][m] = (Tab::a[i][m] != 0.0) ? in_vgpr(C(Tab::a[i][m])) : C(0);
      b[i] = (Tab::b[i] != 0.0) ? in_vgpr(C(Tab::b[i])) : C(0);
      berr[i] = (Tab::berr[i] != 0.0) ? in_vgpr(C(Tab::berr[i])) : C(0);
    }
  }
};

// The tableau as compile-time constants of type C: with C = float the stage algebra takes
// the coefficients as literal operands (v_fmac_f32 with a 32-bit literal), so they occupy
// no registers, which the MFMA kernels need for their fragments.
template <class Tab, class C>
struct TabLit {
  static constexpr int S = Tab::S;
  struct Arr {
    C a[S][S];
    C b[S];
    C berr[S];
  };
  static constexpr Arr make() {
    Arr r{};
    for (int i = 0; i < S; ++i) {
      for (int m = 0; m < S; ++m) r.a[i][m] = C(Tab::a[i][m]);
      r.b[i] = C(Tab::b[i]);
      r.berr[i] = C(Tab::berr[i]);
    }
    return r;
  }
  static constexpr Arr v = make();
};

// Compile-time tableau access: WHICH 0 = a[ROW], 1 = b, 2 = b_error, 3 = Dopri5 c_mid.
template <class Tab, int WHICH, int ROW, int M>
constexpr double tab_coef() {
  if constexpr (WHICH == 0) return Tab::a[ROW][M];
  else if constexpr (WHICH == 1) return Tab::b[M];
  else if constexpr (WHICH == 2) return Tab::berr[M];
  else return Dopri5Tab::cmid[M];
}

// a * b + c with one rounding, in the precision of the state.
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Sum over m < END of coef[m] * k[m] for the coefficients that are nonzero in the tableau,
// in ascending m: the first nonzero product, then acc = fma(coef[m], k[m], acc) (the fused
// dot product the reference's tensordot, _solver/base.py:37-41, is restated as).  Zero
// entries, and the choice of the first term, are resolved at compile time.  `vals` holds
// the coefficient values (in VGPRs).
template <class Tab, int WHICH, int ROW, int M, int END, bool STARTED, class T, int S, class C>
__device__ __forceinline__ T comb(const C* vals, const T (&k)[S], T acc) {
  if constexpr (M >= END) {
    return acc;
  } else {
    constexpr double c = tab_coef<Tab, WHICH, ROW, M>();
    if constexpr (c != 0.0) {
      if constexpr (STARTED) {
        return comb<Tab, WHICH, ROW, M + 1, END, true>(vals, k,
                                                       fma_t(T(vals[M]), k[M], acc));
      } else {
        return comb<Tab, WHICH, ROW, M + 1, END, true>(vals, k, T(vals[M]) * k[M]);
      }
    } else {
      return comb<Tab, WHICH, ROW, M + 1, END, STARTED>(vals, k, acc);
    }
  }
}

// static_for<B, E>(f): f(integral_constant<int, i>) for i in [B, E), unrolled at compile time.
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// The solver's local interpolant on [ts0, ts1] (normalised time) at t: Tsit5's
// (tsit5.py:103-156), Dopri5's quartic through y0, ymid, y1 (_local_interpolation.py:95-139,
// dopri5.py:36-47) and Heun's cubic Hermite (_local_interpolation.py:50-92).
template <class Tab, class T, int D>
__device__ __forceinline__ void erk_interp(double ts0, double ts1, const T (&y0)[D],
                                           const T (&y1)[D], const T (&k)[Tab::S][D], double t,
                                           T (&out)[D]) {
  constexpr int S = Tab::S;
  const double th = linear_rescale(ts0, t, ts1);
  if constexpr (Tab::id == SOLVER_TSIT5) {
    // Tsit5
    double bb[7];
    tsit5_weights(th, bb);
#pragma unroll
    for (int j = 0; j < D; ++j) {
      T acc = T(bb[0]) * k[0][j];
#pragma unroll
      for (int m = 1; m < 7; ++m) acc = fma_t(T(bb[m]), k[m][j], acc);
      out[j] = y0[j] + acc;
    }
  } else if constexpr (Tab::id == SOLVER_DOPRI5) {
    // Dopri5: quartic through y0, ymid, y1 with end slopes k0, k6.
    const T tt = T(th);
#pragma unroll
    for (int j = 0; j < D; ++j) {
      T kc[S];
#pragma unroll
      for (int m = 0; m < S; ++m) kc[m] = k[m][j];
      const T ymid = y0[j] + comb<Tab, 3, 0, 0, S, false>(Dopri5Tab::cmid, kc, T(0));
      const T f0 = k[0][j], f1 = k[6][j];
      const T ca = T(2) * (f1 - f0) - T(8) * (y1[j] + y0[j]) + T(16) * ymid;
      const T cb = T(5) * f0 - T(3) * f1 + T(18) * y0[j] + T(14) * y1[j] - T(32) * ymid;
      const T cc = f1 - T(4) * f0 - T(11) * y0[j] - T(5) * y1[j] + T(16) * ymid;
      out[j] = (((ca * tt + cb) * tt + cc) * tt + f0) * tt + y0[j];
    }
  } else {
    // Heun: cubic Hermite from k0, k_last.
    const T tt = T(th);
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const T k0 = k[0][j], k1 = k[S - 1][j];
      const T ca = k0 + k1 + T(2) * y0[j] - T(2) * y1[j];
      const T cb = T(-2) * k0 - k1 - T(3) * y0[j] + T(3) * y1[j];
      out[j] = ((ca * tt + cb) * tt + k0) * tt + y0[j];
    }
  }
}

// d/dt of erk_interp (in normalised time): the interpolating polynomial differentiated in
// theta, divided by ts1 - ts0; zero on a zero-length interval (linear_rescale is constant
// there).
template <class Tab, class T, int D>
__device__ __forceinline__ void erk_interp_deriv(double ts0, double ts1, const T (&y0)[D],
                                                 const T (&y1)[D], const T (&k)[Tab::S][D],
                                                 double t, T (&out)[D]) {
  constexpr int S = Tab::S;
  const double th = linear_rescale(ts0, t, ts1);
  const double inv_h = (ts0 == ts1) ? 0.0 : 1.0 / (ts1 - ts0);
  if constexpr (Tab::id == SOLVER_TSIT5) {
    double db[7];
    tsit5_dweights(th, db);
#pragma unroll
    for (int j = 0; j < D; ++j) {
      T acc = T(db[0]) * k[0][j];
#pragma unroll
      for (int m = 1; m < 7; ++m) acc = acc + T(db[m]) * k[m][j];
      out[j] = T((double)acc * inv_h);
    }
  } else if constexpr (Tab::id == SOLVER_DOPRI5) {
    const T tt = T(th);
#pragma unroll
    for (int j = 0; j < D; ++j) {
      T kc[S];
#pragma unroll
      for (int m = 0; m < S; ++m) kc[m] = k[m][j];
      const T ymid = y0[j] + comb<Tab, 3, 0, 0, S, false>(Dopri5Tab::cmid, kc, T(0));
      const T f0 = k[0][j], f1 = k[6][j];
      const T ca = T(2) * (f1 - f0) - T(8) * (y1[j] + y0[j]) + T(16) * ymid;
      const T cb = T(5) * f0 - T(3) * f1 + T(18) * y0[j] + T(14) * y1[j] - T(32) * ymid;
      const T cc = f1 - T(4) * f0 - T(11) * y0[j] - T(5) * y1[j] + T(16) * ymid;
      const T dp = ((T(4) * ca * tt + T(3) * cb) * tt + T(2) * cc) * tt + f0;
      out[j] = T((double)dp * inv_h);
    }
  } else {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const T tt = T(th);
      const T k0 = k[0][j], k1 = k[S - 1][j];
      const T ca = k0 + k1 + T(2) * y0[j] - T(2) * y1[j];
      const T cb = T(-2) * k0 - k1 - T(3) * y0[j] + T(3) * y1[j];
      const T dp = (T(3) * ca * tt + T(2) * cb) * tt + k0;
      out[j] = T((double)dp * inv_h);
    }
  }
}

// ClipStepSizeController's view of step_ts / jump_ts (clip.py): the times scaled by the
// direction and sorted (wrap, clip.py:231-241), i.e. ts[i] for dir > 0 and -ts[n-1-i] for
// dir < 0.  Searches are binary (the reference's linear search from a hint finds the same
// index).
struct ClipTimes {
  const double* ts;
  int n, dir;
  __device__ __forceinline__ double at(int i) const { return dir > 0 ? ts[i] : -ts[n - 1 - i]; }
  // _get_t (clip.py:57-65)
  __device__ __forceinline__ double get_t(int i) const {
    return i >= n ? __builtin_inf() : at(i);
  }
  // _find_idx_with_hint (clip.py:104-118): first index with ts[i] > t
  __device__ __forceinline__ int find(double t) const {
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (at(mid) <= t) lo = mid + 1; else hi = mid;
    }
    return lo;
  }
  __device__ __forceinline__ bool contains(double x) const {
    const int i = find(x);
    return i > 0 && at(i - 1) == x;
  }
  // _clip_t (clip.py:68-79)
  __device__ __forceinline__ double clip(double t, int i, bool prev) const {
    double u = get_t(i);
    if (prev) u = prevbefore(u);
    return jmin(u, t);
  }
  // _bump_next_t0 (clip.py:82-101)
  __device__ __forceinline__ double bump(double t, bool& made_jump) const {
    const double na = nextup(t);
    const bool mj1 = contains(na), mj2 = contains(t);
    double r = mj1 ? nextup(na) : t;
    r = mj2 ? na : r;
    made_jump = mj1 || mj2;
    return r;
  }
};

// ---- Event (_event.py:13-118) with affine conditions c_t t + c_y . y + c_0 ----------------
constexpr int kMaxEvents = 4;

template <class T, int D>
__device__ __forceinline__ double event_value(const ErkKArgs& a, int i, double t, const T (&y)[D]) {
  const double* c = a.event_coeffs + i * (D + 2);
  double v = c[0] * t;
#pragma unroll
  for (int j = 0; j < D; ++j) v = v + c[1 + j] * (double)y[j];
  return v + c[1 + D];
}
// jnp.sign
__device__ __forceinline__ double event_sign(double x) {
  return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : (x == 0.0 ? 0.0 : x));
}
// The event test of one condition (_integrate.py:580-602).
__device__ __forceinline__ bool event_triggered(int flags, double old_v, double new_v) {
  if (flags & 1) return new_v > 0.0;
  const int dirn = (flags >> 1) & 3;
  const double so = event_sign(old_v), sn = event_sign(new_v);
  if (dirn == 1) return (so <= 0.0) && (sn > 0.0);
  if (dirn == 2) return (so > 0.0) && (sn <= 0.0);
  return so != sn;
}
// Event time on the step [ta, tb] (normalised): the step end for a boolean condition, else
// bisection of the condition along the step's interpolant down to adjacent floats, the end
// with the smaller |value| (the reference runs optimistix.Newton on the same interpolant,
// _integrate.py:697-776).
template <class Tab, class T, int D>
__device__ double event_root(const ErkKArgs& a, int i, double ta, double tb, const T (&y0)[D],
                             const T (&y1)[D], const T (&k)[Tab::S][D]) {
  if (a.event_flags[i] & 1) return tb;
  T ya[D], yb[D], ym[D];
  erk_interp<Tab, T, D>(ta, tb, y0, y1, k, ta, ya);
  erk_interp<Tab, T, D>(ta, tb, y0, y1, k, tb, yb);
  double ga = event_value<T, D>(a, i, ta, ya), gb = event_value<T, D>(a, i, tb, yb);
  if (!(event_sign(ga) != event_sign(gb)) || ga != ga || gb != gb) return tb;
  double lo = ta, hi = tb;
  for (int it = 0; it < 2100; ++it) {
    const double m = lo + 0.5 * (hi - lo);
    if (!(m > lo && m < hi)) break;
    erk_interp<Tab, T, D>(ta, tb, y0, y1, k, m, ym);
    const double gm = event_value<T, D>(a, i, m, ym);
    if (gm == 0.0) return m;
    if (event_sign(gm) == event_sign(ga)) { lo = m; ga = gm; } else { hi = m; gb = gm; }
  }
  return (fabs(ga) <= fabs(gb)) ? lo : hi;
}

__device__ __forceinline__ unsigned int lane_rank(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned int)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned int)mask, 0u));
}

// Ctl = the PID controller kind (PID_FIFTH, PID_INTEGRAL or PID_GENERIC; ignored under
// ConstantStepSize).  Extras = the solve uses SaveAt(dense=True) or
// ClipStepSizeController; the plain instantiation carries none of that code in its step
// loop.
// Ck: the step loop writes the discrete adjoint's checkpoints when a.ckpt is set.  The
// plain kernels are built without and with it, so a forward solve that records none runs
// no code of theirs; the extras kernels always carry it.
template <class Tab, class VF, class T, int Ctl, bool Extras, bool Ck = Extras>
struct ErkLane {
  static constexpr int D = VF::D;
  static constexpr int P = VF::P;
  static constexpr int S = Tab::S;

  // ---- per-trajectory state -------------------------------------------------------
  TabRegs<Tab> tr;
  int64_t traj;
  int dir;
  double T1, floor_t1, tprev, tnext;
  T y[D];
  T fsal_f[D];
  T q[P];
  bool first;
  bool made_jump;            // the last accepted step ended at a jump (clip.py:383-384)
  double ev_old[kMaxEvents];   // condition values at the current state
  int ev_mask;                 // bit i: condition i triggered (first one wins)
  double tfinal;               // event time and state (the solve's end after an event)
  T yfinal[D];
  int step_index, jump_index;  // ClipStepSizeController hints
  PidState pid;
  PidParams pv;                // a.pid with the hot float64 fields in VGPRs
  double norm_c, norm_rc;      // sqrt(D) (rounded) and its refined reciprocal (rms_norm_rc)
  int32_t cs_completed, cs_num;
  double cs_t0, cs_t1;
  int32_t num_steps, num_acc, num_rej, result;
  int32_t ck_left;    // accepted steps until the next checkpoint (INT32_MAX: none recorded)
  int32_t save_index, ts_index;
  double next_ts;    // normalised save time ts[ts_index], +inf when none left
  double after_raw;  // ts[min(ts_index + 1, n_ts - 1)] as stored, loaded one save ahead
                     // and used only when the next save is reached, so a save never
                     // waits on memory
  const double* ts_row;

  // Unconditional (no branch for the compiler to sink the previous value's use past): the
  // index is clamped into the row, and without save times ts_row points at t1, which is
  // always readable.
  __device__ __forceinline__ void prefetch_ts(const ErkKArgs& a) {
    after_raw = ts_row[max(min(ts_index + 1, a.n_ts - 1), 0)];
  }
  __device__ __forceinline__ void first_ts(const ErkKArgs& a) {
    next_ts = (a.ts != nullptr && a.n_ts > 0) ? ts_row[0] * (double)dir : __builtin_inf();
    after_raw = 0.0;
    prefetch_ts(a);
  }
  // After ++ts_index: the prefetched time becomes the next one (+inf past the last).
  __device__ __forceinline__ void load_next_ts(const ErkKArgs& a) {
    next_ts = (ts_index < a.n_ts) ? after_raw * (double)dir : __builtin_inf();
    prefetch_ts(a);
  }

  __device__ __forceinline__ void eval_f(double t, const T (&z)[D], T (&f)[D]) const {
    VF::eval(t, z, q, f);
  }

  // _select_initial_step (pid.py:33-81) in normalised time.
  __device__ double initial_step(const ErkKArgs& a, double T0) const {
    const double rtol = pv.rtol, atol = pv.atol;
    double scale[D], e[D];
    T f0[D], f1[D], y1[D];
#pragma unroll
    for (int j = 0; j < D; ++j) scale[j] = atol + fabs((double)y[j]) * rtol;
    eval_f(T0 * dir, y, f0);
#pragma unroll
    for (int j = 0; j < D; ++j) e[j] = (double)y[j] / scale[j];
    const double d0 = rms_norm<D>(e);
#pragma unroll
    for (int j = 0; j < D; ++j) e[j] = (double)f0[j] / scale[j];
    const double d1 = rms_norm<D>(e);
    const bool cond = (d0 < 1e-5) || (d1 < 1e-5);
    const double d1_ = cond ? 1.0 : d1;
    const double h0 = cond ? 1e-6 : 0.01 * (d0 / d1_);
#pragma unroll
    for (int j = 0; j < D; ++j) y1[j] = y[j] + T(h0) * f0[j];
    eval_f((T0 + h0) * dir, y1, f1);
#pragma unroll
    for (int j = 0; j < D; ++j) e[j] = ((double)f1[j] - (double)f0[j]) / scale[j];
    const double d2 = rms_norm<D>(e) / h0;
    const double max_d = jmax(d1, d2);
    const double h1 = (max_d <= 1e-15) ? jmax(1e-6, h0 * 1e-3)
                                       : pow_ctl<true>(0.01 / max_d, 1.0 / a.error_order);
    return jmin(100.0 * h0, h1);
  }

  // diffeqsolve set-up for one instance (_integrate.py:1156-1305, pid.py:316-392,
  // constant.py:30-54).
  __device__ __forceinline__ void load_params(const ErkKArgs& a, int64_t id) {
    const T* qp = reinterpret_cast<const T*>(a.params) + id * a.param_stride;
#pragma unroll
    for (int j = 0; j < P; ++j) q[j] = qp[j];
    // The controller's float64 parameters, per lane in VGPRs (see in_vgpr): kept in SGPRs
    // they overflow the scalar file and are reloaded through v_readlane in every step.
    pv = a.pid;
    pv.rtol = a.rtol_b ? a.rtol_b[id] : in_vgpr(a.pid.rtol);
    pv.atol = a.atol_b ? a.atol_b[id] : in_vgpr(a.pid.atol);
    pv.safety = in_vgpr(a.pid.safety);
    pv.factormin = in_vgpr(a.pid.factormin);
    pv.factormax = in_vgpr(a.pid.factormax);
    norm_c = in_vgpr(sqrt((double)D));
    norm_rc = refined_rcp(norm_c);
  }

  // The parts of a solve's set-up that depend only on its span: direction, normalised
  // end, the 100-ulp end-clipping floor (_integrate.py:319-322) and ConstantStepSize's
  // grid (constant.py:30-54); returns the normalised t0.
  __device__ double setup_span(const ErkKArgs& a, double t0r, double t1r) {
    dir = (t0r < t1r) ? 1 : -1;
    const double T0 = t0r * dir;
    T1 = t1r * dir;
    if (a.controller == DFX_PID) {
      cs_num = 0; cs_t0 = T0; cs_t1 = T1;
    } else {
      const double DT0 = a.dt0 * dir;
      const double qn = ceil((T1 - T0) / nextup(DT0));
      cs_num = __builtin_isfinite(T1) ? (int32_t)qn : -1;
      cs_t1 = __builtin_isfinite(T1) ? T1 : DT0;
      cs_t0 = T0;
    }
    double fl = T1;
    for (int r = 0; r < 100; ++r) fl = prevbefore(fl);
    floor_t1 = fl;
    return T0;
  }

  // Time normalisation and controller initialisation of a solve from t0r to t1r for the
  // state in y (_integrate.py:1156-1265, pid.py:316-392, constant.py:30-54, clip.py:243-299);
  // returns the normalised t0.
  __device__ double begin_times(const ErkKArgs& a, double t0r, double t1r) {
#pragma unroll
    for (int j = 0; j < D; ++j) fsal_f[j] = T(0);
    const double T0 = setup_span(a, t0r, t1r);
    first = true;
    made_jump = false;
    step_index = jump_index = 0;
    pid.inv = 1.0; pid.pinv = 1.0; pid.at_dtmin = 0;
    if (a.controller == DFX_PID) {
      double dt0 = a.has_dt0 ? a.dt0 * dir : initial_step(a, T0);
      if (a.pid.has_dtmax) dt0 = jmin(dt0, a.pid.dtmax);
      if (a.pid.has_dtmin) dt0 = jmax(dt0, a.pid.dtmin);
      tnext = T0 + dt0;
      cs_completed = 0;
      if (Extras && a.n_step_ts > 0) {
        const ClipTimes c{a.step_ts, a.n_step_ts, dir};
        step_index = c.find(T0);
        tnext = c.clip(tnext, step_index, false);
      }
      if (Extras && a.n_jump_ts > 0) {
        const ClipTimes c{a.jump_ts, a.n_jump_ts, dir};
        jump_index = c.find(T0);
        tnext = c.clip(tnext, jump_index, true);
      }
    } else {
      cs_completed = 1;
      tnext = T0 + a.dt0 * dir;
    }
    tnext = jmin(tnext, T1);
    tprev = T0;
    num_steps = num_acc = num_rej = 0;
    ck_left = (Ck && a.ckpt) ? a.ckpt_every : INT32_MAX;
    result = DFX_SUCCESSFUL;
    if constexpr (Tab::fsal && !Extras) {
      // The first step's stage 0 (runge_kutta.py:687: z0 = y0 + 0 at t0 + 0*dt), evaluated
      // here once instead of behind a per-lane branch in every step: later steps reuse the
      // last stage (FSAL), and a rejected first step retries from the same y.
      T z[D];
#pragma unroll
      for (int j = 0; j < D; ++j) z[j] = y[j] + T(0);
      eval_f((T0 + 0.0 * (tnext - T0)) * dir, z, fsal_f);
      first = false;
    }
    return T0;
  }

  // One backward solve between two save times of the continuous adjoint: the state in y
  // is integrated from t0r to t1r with no saves (BacksolveAdjoint._scan_fun,
  // _adjoint.py:645-674).
  __device__ void solve_segment(const ErkKArgs& a, double t0r, double t1r) {
    begin_times(a, t0r, t1r);
    save_index = 0;
    ts_index = 0;
    ts_row = a.t1;  // no save times (next_ts = +inf keeps the save loop from running)
    next_ts = after_raw = __builtin_inf();
    ev_mask = 0;
    while (running(a)) step(a);
    if (tprev < T1 && result == DFX_SUCCESSFUL) result = DFX_MAX_STEPS_REACHED;
  }

  // ---- checkpoints of the discrete adjoint ----------------------------------------
  // Record c (CW = D + 8 float64 words) is the lane state after c * ckpt_every accepted
  // steps: y, tprev, tnext, the PID state, the end time of the last accepted step (tl,
  // -inf for c = 0), then six int32 (at_dtmin | made_jump << 1, num_steps, num_rej,
  // step_index, jump_index, cs_completed).  Together with the instance's inputs it
  // restarts the step loop exactly where the forward was.
  static constexpr int CW = D + 8;
  __device__ void write_ckpt(const ErkKArgs& a, int c, double tl) const {
    if (c >= a.n_ckpt) return;
    double* r = a.ckpt + (traj * a.n_ckpt + c) * CW;
#pragma unroll
    for (int j = 0; j < D; ++j) r[j] = (double)y[j];
    r[D] = tprev;
    r[D + 1] = tnext;
    r[D + 2] = pid.inv;
    r[D + 3] = pid.pinv;
    r[D + 4] = tl;
    int32_t* ri = reinterpret_cast<int32_t*>(r + D + 5);
    ri[0] = pid.at_dtmin | (made_jump ? 2 : 0);
    ri[1] = num_steps;
    ri[2] = num_rej;
    ri[3] = step_index;
    ri[4] = jump_index;
    ri[5] = cs_completed;
  }

  // The per-instance set-up a replay needs (no initial step, no saves, no events).
  __device__ void replay_setup(const ErkKArgs& a, int64_t id) {
    tr.load();
    traj = id;
    load_params(a, id);
    setup_span(a, a.t0, a.t1[id * a.t1_stride]);
    save_index = 0;
    ts_index = 0;
    ts_row = a.t1;
    next_ts = after_raw = __builtin_inf();
    ev_mask = 0;
  }

  // Restore the state of checkpoint c of the forward solve (same Tab, VF, T, controller
  // arithmetic and clipping, so the steps that follow repeat the forward's bit for bit).
  // The accepted-step counter restarts at acc0, so the grid records of the steps that
  // follow go to indices acc0, acc0 + 1, ... of the instance's grid row; nothing else in
  // a replay (no saves, dense output or checkpoints) reads the counter.  Returns the end
  // time of the accepted step before the checkpoint (-inf for c = 0).
  __device__ double resume(const ErkKArgs& a, const double* ckpt, int c, int32_t acc0) {
    const double* r = ckpt + (traj * a.n_ckpt + c) * CW;
#pragma unroll
    for (int j = 0; j < D; ++j) y[j] = T(r[j]);
    tprev = r[D];
    tnext = r[D + 1];
    pid.inv = r[D + 2];
    pid.pinv = r[D + 3];
    const double tl = r[D + 4];
    const int32_t* ri = reinterpret_cast<const int32_t*>(r + D + 5);
    pid.at_dtmin = ri[0] & 1;
    made_jump = (ri[0] >> 1) & 1;
    num_steps = ri[1];
    num_rej = ri[2];
    step_index = ri[3];
    jump_index = ri[4];
    cs_completed = ri[5];
    num_acc = acc0;
    ck_left = INT32_MAX;
    result = DFX_SUCCESSFUL;
    // FSAL: the first stage of the next step.  Without extras the forward evaluated it at
    // set-up (c = 0, begin_times) or as the previous step's last stage, f(t1, y1) with
    // t1 = tprev and y1 = y (SSAL); with extras a first step or one after a jump
    // re-evaluates it inside step().
    first = Extras && c == 0;
    if constexpr (Tab::fsal) {
      if (!(Extras && (first || made_jump))) {
        if (c == 0) {
          T z[D];
#pragma unroll
          for (int j = 0; j < D; ++j) z[j] = y[j] + T(0);
          eval_f((tprev + 0.0 * (tnext - tprev)) * dir, z, fsal_f);
        } else {
          eval_f(tprev * dir, y, fsal_f);
        }
      }
    }
    return tl;
  }

  // ---- solver / controller state of a resumed solve ------------------------------
  // Record layout (D + 9 float64 words): f0[D], first_step, made_jump, the PID state
  // (prev_inv_scaled_error, prev_prev_inv_scaled_error, at_dtmin), the ConstantStepSize
  // state (steps_completed, num_steps, t0_sim, t1_sim_or_dt0).  Resuming replaces what
  // begin_times set (_integrate.py:1252-1270): with the controller state and dt0 given the
  // first step ends at t0 + dt0, without the PID init's dtmin / dtmax clamps; with the
  // solver state the first step takes f0 as its first stage unless first_step or
  // made_jump asks for a fresh evaluation (runge_kutta.py:684-690).
  static constexpr int SW = D + 9;
  __device__ void resume_state(const ErkKArgs& a, double T0) {
    const double* r = a.state_in + traj * SW;
    if (a.resume & DFX_RESUME_SOLVER) {
#pragma unroll
      for (int j = 0; j < D; ++j) fsal_f[j] = Tab::fsal ? T(r[j]) : T(0);
      first = Tab::fsal && r[D] != 0.0;
    }
    if (a.resume & DFX_RESUME_MADE_JUMP) made_jump = r[D + 1] != 0.0;
    if (a.resume & DFX_RESUME_CONTROLLER) {
      if (a.controller == DFX_PID) {
        pid.inv = r[D + 2];
        pid.pinv = r[D + 3];
        pid.at_dtmin = r[D + 4] != 0.0;
        if (a.has_dt0) tnext = jmin(T0 + a.dt0 * dir, T1);
      } else {
        cs_completed = (int32_t)r[D + 5];
        cs_num = (int32_t)r[D + 6];
        cs_t0 = r[D + 7];
        cs_t1 = r[D + 8];
      }
    }
  }
  // The final state (SaveAt(solver_state=True, controller_state=True, made_jump=True),
  // _integrate.py:1489-1500).  Before any accepted step the solver state is still the
  // initial one: f0 as resumed, or zero (runge_kutta.py:424-442), with first_step set.
  __device__ void write_state(const ErkKArgs& a) const {
    double* r = a.state_out + traj * SW;
    const bool resumed_f0 = (a.resume & DFX_RESUME_SOLVER) != 0;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      double v = 0.0;
      if (Tab::fsal) v = !first ? (double)fsal_f[j] : (resumed_f0 ? a.state_in[traj * SW + j] : 0.0);
      r[j] = v;
    }
    r[D] = (Tab::fsal && first) ? 1.0 : 0.0;
    r[D + 1] = made_jump ? 1.0 : 0.0;
    r[D + 2] = pid.inv;
    r[D + 3] = pid.pinv;
    r[D + 4] = pid.at_dtmin ? 1.0 : 0.0;
    r[D + 5] = (double)cs_completed;
    r[D + 6] = (double)cs_num;
    r[D + 7] = cs_t0;
    r[D + 8] = cs_t1;
  }

  __device__ void start(const ErkKArgs& a, int64_t id) {
    tr.load();
    traj = id;
    const T* y0p = reinterpret_cast<const T*>(a.y0) + id * D;
#pragma unroll
    for (int j = 0; j < D; ++j) y[j] = y0p[j];
    load_params(a, id);
    const double T0 = begin_times(a, a.t0, a.t1[id * a.t1_stride]);
    if (Extras && a.resume) resume_state(a, T0);
    if (Ck && a.ckpt) write_ckpt(a, 0, -__builtin_inf());
    save_index = 0;
    ts_index = 0;
    ts_row = a.ts ? a.ts + id * a.ts_stride : a.t1;
    first_ts(a);
    if (a.save_t0) {
      write_save(a, T0, y);
    }
    ev_mask = 0;
    if (Extras && a.n_events > 0) {
      // values at (t0, y0); a boolean condition that holds ends the solve here
      // (_integrate.py:1333-1426)
      for (int i = 0; i < a.n_events; ++i) {
        const double v = event_value<T, D>(a, i, T0, y);
        const bool boolean = a.event_flags[i] & 1;
        ev_old[i] = boolean ? (double)(v > 0.0) : v;
        if (boolean && v > 0.0 && ev_mask == 0) ev_mask = 1 << i;
      }
      if (ev_mask) {
        result = DFX_EVENT_OCCURRED;
        tfinal = T0;
#pragma unroll
        for (int j = 0; j < D; ++j) yfinal[j] = y[j];
      }
    }
    if (Extras && a.dense_ts) {
      const int64_t row = traj * ((int64_t)a.max_steps + 1);
      a.dense_ts[row] = T0;
      T* dy = reinterpret_cast<T*>(a.dense_ys) + row * D;
#pragma unroll
      for (int j = 0; j < D; ++j) dy[j] = y[j];
    }
  }

  __device__ __forceinline__ void write_save(const ErkKArgs& a, double t_norm, const T (&v)[D]) {
    T* row = reinterpret_cast<T*>(a.ys) + (traj * a.n_out + save_index) * D;
#pragma unroll
    for (int j = 0; j < D; ++j) row[j] = v[j];
    a.ts_out[traj * a.n_out + save_index] = t_norm * (double)dir;
    ++save_index;
  }

  __device__ __forceinline__ bool running(const ErkKArgs& a) const {
    return tprev < T1 && result == DFX_SUCCESSFUL && num_steps < a.max_steps;
  }

  // One pass of body_fun_aux (_integrate.py:365-662).
  __device__ void step(const ErkKArgs& a) {
    const double ts0 = tprev, ts1 = tnext;
    const double dt = ts1 - ts0;
    const T control = T((double)dir * dt);
    T k[S][D];
    T z[D], f[D];
    if constexpr (Tab::fsal && Extras) {
      if (first || made_jump) {
        // First step, or the first step after a jump (runge_kutta.py:687): stage 0 is
        // evaluated (z0 = y0 + 0, t = t0 + 0*dt).
#pragma unroll
        for (int j = 0; j < D; ++j) z[j] = y[j] + T(0);
        eval_f((ts0 + 0.0 * dt) * dir, z, fsal_f);
      }
    }
    if constexpr (Tab::fsal) {
#pragma unroll
      for (int j = 0; j < D; ++j) k[0][j] = control * fsal_f[j];
    }
    constexpr int start_stage = Tab::fsal ? 1 : 0;
    static_for<start_stage, S>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
#pragma unroll
      for (int j = 0; j < D; ++j) {
        T kc[S];
#pragma unroll
        for (int m = 0; m < S; ++m) kc[m] = (m < i) ? k[m][j] : T(0);
        z[j] = y[j] + comb<Tab, 0, i, 0, i, false>(tr.a[i], kc, T(0));
      }
      constexpr double ci = Tab::c[i];
      const double ti = (ci == 1.0) ? ts1 : ts0 + ci * dt;
      eval_f(ti * dir, z, f);
#pragma unroll
      for (int j = 0; j < D; ++j) k[i][j] = control * f[j];
    });
    T y1[D], yerr[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      T kc[S];
#pragma unroll
      for (int m = 0; m < S; ++m) kc[m] = k[m][j];
      if constexpr (Tab::ssal) {
        y1[j] = z[j];
      } else {
        y1[j] = y[j] + comb<Tab, 1, 0, 0, S, false>(tr.b, kc, T(0));
      }
      yerr[j] = comb<Tab, 2, 0, 0, S, false>(tr.berr, kc, T(0));
      if (yerr[j] != yerr[j]) yerr[j] = T(__builtin_inf());
    }

    // ---- step-size control ----
    bool keep;
    double ntprev, ntnext;
    int code = 0;
    if (a.controller == DFX_PID) {
      bool anynan = false;
#pragma unroll
      for (int j = 0; j < D; ++j) anynan = anynan || (y1[j] != y1[j]);
      double e[D];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        const double y1c = anynan ? (double)y[j] : (double)y1[j];
        const double yy = jmax_nonneg(fabs((double)y[j]), fabs(y1c));
        e[j] = (double)yerr[j] / (pv.atol + yy * pv.rtol);
      }
      double scaled;
      if constexpr (D >= 2) scaled = rms_norm_rc<D>(e, norm_c, norm_rc);
      else scaled = rms_norm<D>(e);
      keep = pid_adapt<Ctl>(pv, scaled, ts0, ts1, pid, ntprev, ntnext, code);
    } else {
      keep = true;
      const int32_t sc = cs_completed + 1;
      double tn = cs_t0 + (cs_t1 - cs_t0) * ((double)sc / (double)cs_num);
      if (sc == cs_num) tn = cs_t1;
      if (!(cs_num >= 0)) tn = ts1 + cs_t1;
      cs_completed = sc;
      ntprev = ts1;
      ntnext = tn;
    }
    // made_jump of the step: only a jump time sets it (PIDController and ConstantStepSize
    // return False, pid.py:566, constant.py:98-104); an accepted step takes it
    // (_integrate.py:425)
    bool mj = false;
    if (Extras && a.controller == DFX_PID && (a.n_step_ts > 0 || a.n_jump_ts > 0)) {
      // ClipStepSizeController.adapt_step_size (clip.py:301-428)
      if (a.n_step_ts > 0) {
        const ClipTimes c{a.step_ts, a.n_step_ts, dir};
        bool unused;
        step_index = c.find(c.bump(ntprev, unused));
        ntnext = c.clip(ntnext, step_index, false);
      }
      if (a.n_jump_ts > 0) {
        const ClipTimes c{a.jump_ts, a.n_jump_ts, dir};
        ntprev = c.bump(ntprev, mj);
        ntnext = jmax(nextup(ntprev), ntnext);
        jump_index = c.find(ntprev);
        ntnext = c.clip(ntnext, jump_index, true);
      }
    }
    if (Extras && keep) made_jump = mj;
    ntprev = jmin(ntprev, T1);
    if (ntnext > floor_t1) ntnext = keep ? T1 : ntprev + 0.5 * (T1 - ntprev);

    ++num_steps;
    if (keep) ++num_acc; else ++num_rej;
    if (result == DFX_SUCCESSFUL) result = code;

    bool event_now = false;
    if (Extras && a.n_events > 0) {
      // Event check at the new state (_integrate.py:542-633); with a root finder the time
      // is located on this step's interpolant (686-779), and saves after it are dropped
      // (781-806).
      int mask = 0, fired = -1;
      for (int i = 0; i < a.n_events; ++i) {
        T yn[D];
#pragma unroll
        for (int j = 0; j < D; ++j) yn[j] = keep ? y1[j] : y[j];
        const double v = event_value<T, D>(a, i, ntprev, yn);
        const int flags = a.event_flags[i];
        const bool trig = event_triggered(flags, ev_old[i], v);
        if (trig && mask == 0) { mask = 1 << i; fired = i; }
        ev_old[i] = (flags & 1) ? (double)(v > 0.0) : v;
      }
      ev_mask = mask;
      if (mask) {
        event_now = true;
        result = DFX_EVENT_OCCURRED;
        if (a.event_root_find) {
          tfinal = event_root<Tab, T, D>(a, fired, ts0, ts1, y, y1, k);
          erk_interp<Tab, T, D>(ts0, ts1, y, y1, k, tfinal, yfinal);
        } else {
          tfinal = ntprev;
#pragma unroll
          for (int j = 0; j < D; ++j) yfinal[j] = keep ? y1[j] : y[j];
        }
      }
    }

    if (keep) {
      if (a.grid) {
        const int64_t rec = traj * a.grid_stride + (num_acc - 1);
        double* g = a.grid + rec * 2;
        g[0] = ts0;
        g[1] = ts1;
        if (a.grid_y) {  // the accepted step's start state (the discrete adjoint's replay)
          T* gy = reinterpret_cast<T*>(a.grid_y) + rec * D;
#pragma unroll
          for (int j = 0; j < D; ++j) gy[j] = y[j];
        }
      }
      if (Extras && a.dense_ts) {
        // SaveAt(dense=True): the step's end time, end state and stage increments
        // (_integrate.py:529-540; the interpolant is rebuilt from them on evaluation).
        const int64_t row = traj * ((int64_t)a.max_steps + 1) + num_acc;
        a.dense_ts[row] = ntprev;
        T* dy = reinterpret_cast<T*>(a.dense_ys) + row * D;
        T* dk = reinterpret_cast<T*>(a.dense_k) +
                ((traj * (int64_t)a.max_steps + (num_acc - 1)) * S) * D;
#pragma unroll
        for (int j = 0; j < D; ++j) dy[j] = y1[j];
#pragma unroll
        for (int m = 0; m < S; ++m) {
#pragma unroll
          for (int j = 0; j < D; ++j) dk[m * D + j] = k[m][j];
        }
      }
      // SaveAt(ts): interpolate inside [ts0, ts1] (_integrate.py:456-491).
      while (next_ts <= ts1 && !(Extras && event_now && next_ts > tfinal)) {
        T v[D];
        interp(ts0, ts1, y, y1, k, next_ts, v);
        write_save(a, next_ts, v);
        ++ts_index;
        load_next_ts(a);
      }
#pragma unroll
      for (int j = 0; j < D; ++j) y[j] = y1[j];
      if constexpr (Tab::fsal) {
        // The last stage's f becomes the next step's first stage (FSAL).
        first = false;
#pragma unroll
        for (int j = 0; j < D; ++j) fsal_f[j] = f[j];
      }
      // SaveAt(steps=k): every k-th accepted step (_integrate.py:493-527)
      if (a.save_steps && num_acc % a.save_steps == 0 && !(Extras && event_now && ntprev > tfinal))
        write_save(a, ntprev, y);
    }
    tprev = ntprev;
    tnext = ntnext;
    if constexpr (Ck) {
      if (keep && --ck_left == 0) {
        // every ckpt_every accepted steps: the discrete adjoint's restart point (without
        // checkpoints the count starts at INT32_MAX and the test below keeps max_steps =
        // INT32_MAX safe)
        ck_left = a.ckpt_every;
        if (a.ckpt) write_ckpt(a, num_acc / a.ckpt_every, ts1);
      }
    }
  }

  __device__ __forceinline__ void interp(double ts0, double ts1, const T (&y0)[D],
                                         const T (&y1)[D], const T (&k)[S][D], double t,
                                         T (&out)[D]) const {
    erk_interp<Tab, T, D>(ts0, ts1, y0, y1, k, t, out);
  }

  // Loop exit bookkeeping (_integrate.py:823-883) and output of counters.
  __device__ void finish(const ErkKArgs& a) {
    const double T0 = a.t0 * dir;
    if (T0 == T1) {
      for (int r = 0; r < a.n_ts; ++r) write_save(a, T0, y);
    }
    if (Extras && result == DFX_EVENT_OCCURRED) {  // the solve ends at the event
      tprev = tfinal;
#pragma unroll
      for (int j = 0; j < D; ++j) y[j] = yfinal[j];
    }
    // _save_t1 (_integrate.py:847-877): the t1 row unless the last step was saved via
    // steps; with an event root finder, also when it was (its row after the event time
    // was removed).
    const bool via_steps = a.save_steps && num_acc % a.save_steps == 0;
    const bool root = Extras && a.n_events > 0 && a.event_root_find;
    const bool t1_pred = root ? (a.save_t1 || via_steps) : (a.save_t1 && !via_steps);
    if (t1_pred && save_index < a.n_out) write_save(a, tprev, y);
    if (tprev < T1 && result == DFX_SUCCESSFUL) result = DFX_MAX_STEPS_REACHED;
    if (Extras && a.n_events > 0 && a.event_mask) {
      for (int i = 0; i < a.n_events; ++i) a.event_mask[traj * a.n_events + i] = (ev_mask >> i) & 1;
    }
    // Unfilled rows keep the reference's preset values: ys inf, ts direction*inf in
    // normalised time, which the final rescaling by the direction makes +inf
    // (_integrate.py:1296, 1481).
    T* rows = reinterpret_cast<T*>(a.ys) + traj * a.n_out * D;
    for (int s = save_index; s < a.n_out; ++s) {
#pragma unroll
      for (int j = 0; j < D; ++j) rows[s * D + j] = T(__builtin_inf());
      a.ts_out[traj * a.n_out + s] = __builtin_inf();
    }
    a.stats[traj * 3 + 0] = num_steps;
    a.stats[traj * 3 + 1] = num_acc;
    a.stats[traj * 3 + 2] = num_rej;
    a.result[traj] = result;
    if (Extras && a.state_out) write_state(a);
  }
};

// ---- continuous adjoint (BacksolveAdjoint, _adjoint.py:539-873) -------------------------
// Input: the forward solve's saved rows (SaveAt t0 / ts / t1, no steps) and their
// cotangents.  Between consecutive save times, from the last one back to t0, the adjoint
// state a = (a_y, a_q) is integrated backwards together with y (restarted from the saved
// forward state) under AdjointVF, after adding that save's cotangent to a_y.  Every
// backward solve uses the forward's solver and step-size controller (dt0 negated), as
// the reference's per-interval diffeqsolve does; with the integral PID controller the
// carried controller state has no effect, so each interval is a fresh solve whose first
// stage is evaluated (made_jump).
struct BacksolveArgs {
  const void* ys;        // [B, n_out, DB] forward saved states
  const double* ts_out;  // [B, n_out] their times
  const void* ys_bar;    // [B, n_out, DB] cotangents
  void* y0_bar;          // out [B, DB]
  void* q_bar;           // out [B, P] (per instance; summed by the caller when shared)
  int32_t* result;       // out [B]
  int32_t n_out, save_t0;
  double t0;             // forward t0
};

template <class Tab, class VF, class T>
__global__ void __launch_bounds__(256) erk_backsolve_kernel(const ErkKArgs a, const BacksolveArgs b) {
  using AVF = AdjointVF<VF, T>;
  constexpr int DB = VF::D, P = VF::P;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= a.B) return;
  ErkLane<Tab, AVF, T, PID_INTEGRAL, false> L;
  L.tr.load();
  L.traj = id;
  L.load_params(a, id);
  T ay[DB], aq[P];
#pragma unroll
  for (int j = 0; j < DB; ++j) ay[j] = T(0);
#pragma unroll
  for (int m = 0; m < P; ++m) aq[m] = T(0);
  const T* ys = reinterpret_cast<const T*>(b.ys) + id * b.n_out * DB;
  const T* g = reinterpret_cast<const T*>(b.ys_bar) + id * b.n_out * DB;
  const double* tso = b.ts_out + id * b.n_out;
  int32_t res = DFX_SUCCESSFUL;
  const int s_end = b.save_t0 ? 1 : 0;
  for (int s = b.n_out - 1; s >= s_end; --s) {
#pragma unroll
    for (int j = 0; j < DB; ++j) {
      L.y[j] = ys[s * DB + j];
      L.y[DB + j] = ay[j] + g[s * DB + j];
    }
#pragma unroll
    for (int m = 0; m < P; ++m) L.y[2 * DB + m] = aq[m];
    L.solve_segment(a, tso[s], s > 0 ? tso[s - 1] : b.t0);
    if (res == DFX_SUCCESSFUL) res = L.result;
#pragma unroll
    for (int j = 0; j < DB; ++j) ay[j] = L.y[DB + j];
#pragma unroll
    for (int m = 0; m < P; ++m) aq[m] = L.y[2 * DB + m];
  }
  if (b.save_t0 && b.n_out > 0) {
#pragma unroll
    for (int j = 0; j < DB; ++j) ay[j] = ay[j] + g[j];
  }
  T* yb = reinterpret_cast<T*>(b.y0_bar) + id * DB;
  T* qb = reinterpret_cast<T*>(b.q_bar) + id * P;
#pragma unroll
  for (int j = 0; j < DB; ++j) yb[j] = ay[j];
#pragma unroll
  for (int m = 0; m < P; ++m) qb[m] = aq[m];
  b.result[id] = res;
}

// One trajectory per lane, no work queue: used when the launch has a lane for every
// instance.  Lanes whose trajectory has finished idle (masked) until the wave's last lane
// is done.
template <class Tab, class VF, class T, int Ctl, bool Extras, bool Ck = Extras>
__global__ void __launch_bounds__(256) erk_small_kernel(const ErkKArgs a) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= a.B) return;
  ErkLane<Tab, VF, T, Ctl, Extras, Ck> L;
  L.start(a, id);
  while (L.running(a)) L.step(a);
  L.finish(a);
}

// Persistent variant with refill: a lane whose trajectory finishes claims the next
// instance id from the work counter (one atomic per wave per claim round).
template <class Tab, class VF, class T, int Ctl, bool Extras>
__global__ void __launch_bounds__(256) erk_small_queue_kernel(const ErkKArgs a) {
  ErkLane<Tab, VF, T, Ctl, Extras> L;
  bool active = false;
  bool drained = false;  // wave-uniform: the work queue is exhausted
  while (true) {
    const bool need = !active && !drained;
    const unsigned long long need_mask = __ballot(need);
    if (need_mask) {
      const int n = __popcll(need_mask);
      const int leader = __ffsll((long long)need_mask) - 1;
      unsigned int base = 0;
      if ((int)__lane_id() == leader) base = atomicAdd(a.counter, (unsigned int)n);
      base = __shfl(base, leader);
      if ((int64_t)base + n >= a.B) drained = true;
      if (need) {
        const int64_t id = (int64_t)base + lane_rank(need_mask);
        if (id < a.B) {
          L.start(a, id);
          active = true;
        }
      }
    }
    if (!__any(active)) break;
    if (active) {
      if (L.running(a)) {
        L.step(a);
      } else {
        L.finish(a);
        active = false;
      }
    }
  }
}

static __global__ void zero_counter_kernel(unsigned int* counter) { *counter = 0u; }

}  // namespace dfx
