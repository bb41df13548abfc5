// erk_adjoint.hpp -- discrete adjoint of the small-system solve: reverse mode through the
// accepted steps, the gradient RecursiveCheckpointAdjoint takes (_adjoint.py:264-312).
//
// The forward solve stores the whole lane state every K accepted steps (ErkKArgs::ckpt,
// ErkLane::write_ckpt; K = ceil(max_steps / checkpoints), sqrt(max_steps) by default), so
// memory is O(B (max_steps / K + K)) instead of a record per step.  The backward pass
// takes the segments newest first: it restarts the forward step loop from the segment's
// checkpoint (ErkLane::resume; same code and arithmetic, so the accepted steps repeat
// bit for bit, rejected ones included), writes each accepted step's (t_start, t_end) and
// start state to one half of a 2K-record per-instance scratch (ErkKArgs::grid / grid_y of
// the replay arguments) and sweeps those records in reverse, while the next (older)
// segment is replayed into the other half by a second wave.  The step-size controller's factors are
// constants of the gradient (pid.py:374, 498-530) and rejected steps contribute nothing,
// so only the accepted steps enter the sweep.
//
// Per accepted step, newest first (one lane per instance):
//   - its stages are recomputed from the recorded start state, exactly as the forward
//     computes them (z_i = y0 + sum_l a_il k_l, k_i = control * f(z_i));
//   - the cotangents of the rows the step saved are pulled back through the interpolant
//     (erk_interp_vjp) and, for SaveAt(steps), onto the end state;
//   - y1 = y0 + sum_i b_i k_i and the stages are transposed: for i from the last stage
//     down, w = control * k_bar_i, (z_bar_i, q_bar_i) = w^T (df/dy, df/dq) at z_i, and
//     z_bar_i flows to y0 and to k_bar_l with weight a_il.
// FSAL reuse of the previous step's last stage is f at the same point as y1 (a_S = b), so
// its derivative is that of f(y0) of the next step.
#pragma once
#include "erk_small.hpp"

namespace dfx {

struct DiscreteAdjArgs {
  const double* ckpt;     // [B, n_ckpt, D + 8] the forward's checkpoints
  const int32_t* stats;   // [B, 3] (column 1: accepted steps)
  const void* ys_bar;     // [B, n_out, D] cotangents of the saved rows
  void* y0_bar;           // out [B, D]
  void* q_bar;            // out [B, P] per instance
};

// Transpose of erk_interp: for out = interp(y0, y1, k, t) and cotangent g, adds the
// cotangents of y0, y1 and k.  The interpolants are linear in (y0, y1, k) at fixed t.
template <class Tab, class T, int D>
__device__ __forceinline__ void erk_interp_vjp(double ts0, double ts1, double t, const T (&g)[D],
                                               T (&gy0)[D], T (&gy1)[D], T (&gk)[Tab::S][D]) {
  constexpr int S = Tab::S;
  const double th = linear_rescale(ts0, t, ts1);
  if constexpr (Tab::id == SOLVER_TSIT5) {
    double bb[7];
    tsit5_weights(th, bb);
#pragma unroll
    for (int j = 0; j < D; ++j) {
      gy0[j] = gy0[j] + g[j];
#pragma unroll
      for (int m = 0; m < 7; ++m) gk[m][j] = fma_t(T(bb[m]), g[j], gk[m][j]);
    }
  } else if constexpr (Tab::id == SOLVER_DOPRI5) {
    // out = ca t^4 + cb t^3 + cc t^2 + k0 t + y0, ymid = y0 + sum cmid_m k_m
    const T tt = T(th);
    const T t2 = tt * tt, t3 = t2 * tt, t4 = t3 * tt;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const T gca = t4 * g[j], gcb = t3 * g[j], gcc = t2 * g[j];
      const T gmid = T(16) * gca - T(32) * gcb + T(16) * gcc;
      gy0[j] = gy0[j] + ((g[j] + (T(-8) * gca + T(18) * gcb - T(11) * gcc)) + gmid);
      gy1[j] = gy1[j] + (T(-8) * gca + T(14) * gcb - T(5) * gcc);
      gk[0][j] = gk[0][j] + (tt * g[j] + (T(-2) * gca + T(5) * gcb - T(4) * gcc));
      gk[6][j] = gk[6][j] + (T(2) * gca - T(3) * gcb + gcc);
      static_for<0, S>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        if constexpr (Dopri5Tab::cmid[m] != 0.0) gk[m][j] = gk[m][j] + T(Dopri5Tab::cmid[m]) * gmid;
      });
    }
  } else {
    // Heun: out = ca t^3 + cb t^2 + k0 t + y0
    const T tt = T(th);
    const T t2 = tt * tt, t3 = t2 * tt;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const T gca = t3 * g[j], gcb = t2 * g[j];
      gy0[j] = gy0[j] + (g[j] + (T(2) * gca - T(3) * gcb));
      gy1[j] = gy1[j] + (T(-2) * gca + T(3) * gcb);
      gk[0][j] = gk[0][j] + (tt * g[j] + (gca - T(2) * gcb));
      gk[S - 1][j] = gk[S - 1][j] + (gca - gcb);
    }
  }
}

// The reverse sweep of one instance: the cotangent state carried from segment to segment
// (newest first) and the transpose of every accepted step of a segment, read from the
// segment's records.  a: the forward's arguments (save layout, ts, t1, params).
template <class Tab, class VF, class T>
struct AdjSweep {
  static constexpr int S = Tab::S, D = VF::D, P = VF::P;
  static constexpr int CW = D + 8;  // checkpoint record words (ErkLane::write_ckpt)
  TabRegs<Tab> tr;
  T q[P], gq[P];
  int64_t id;
  int dir, n, c_last, ksave;
  double T0, T1;
  const T* g;
  const double* tsr;
  T ab[D];    // cotangent of the state at the end of the step being processed
  T grow[D];  // cotangent of save row r, loaded one row ahead
  int r, ti;
  double tsv;

  __device__ void setup(const ErkKArgs& a, const DiscreteAdjArgs& b, int64_t id_, int n_,
                        int K) {
    tr.load();
    id = id_;
    n = n_;
    c_last = n > 0 ? (n - 1) / K : -1;
    ksave = a.save_steps;
    const T* qp = reinterpret_cast<const T*>(a.params) + id * a.param_stride;
#pragma unroll
    for (int m = 0; m < P; ++m) { q[m] = qp[m]; gq[m] = T(0); }
    const double t1r = a.t1[id * a.t1_stride];
    dir = (a.t0 < t1r) ? 1 : -1;
    T0 = a.t0 * dir;
    T1 = t1r * dir;
    g = reinterpret_cast<const T*>(b.ys_bar) + id * (int64_t)a.n_out * D;
    tsr = a.ts ? a.ts + id * a.ts_stride : nullptr;
#pragma unroll
    for (int j = 0; j < D; ++j) { ab[j] = T(0); grow[j] = T(0); }
    r = -1;
    ti = -1;
    tsv = -__builtin_inf();
    if (n == 0) setup_rows(a, -__builtin_inf());
  }

  // SaveAt(ts) rows the forward wrote: all of them for a zero-length solve, else those at
  // or before the end of the last accepted step (a prefix: ts is sorted); then the t1 row.
  __device__ void setup_rows(const ErkKArgs& a, double tend) {
    int n_saved = 0;
    if (T0 == T1) {
      n_saved = a.n_ts;
    } else if (n > 0) {
      while (n_saved < a.n_ts && tsr[n_saved] * dir <= tend) ++n_saved;
    }
    r = a.save_t0 + n_saved + (ksave > 0 ? n / ksave : 0);  // the SaveAt(t1) row
    if (a.save_t1 && !(ksave > 0 && n % ksave == 0)) {
#pragma unroll
      for (int j = 0; j < D; ++j) ab[j] = ab[j] + g[(int64_t)r * D + j];
    }
    r -= 1;           // the last row a step wrote
    ti = (T0 == T1) ? -1 : n_saved - 1;  // the last ts index saved
    if (r >= 0) {
#pragma unroll
      for (int j = 0; j < D; ++j) grow[j] = g[(int64_t)r * D + j];
    }
    tsv = (ti >= 0) ? tsr[ti] * dir : -__builtin_inf();
  }

  // Accepted steps [c K, min((c + 1) K, n)) of segment c, newest first, from their records
  // sg (t_start, t_end) and sy (start state); tl is the end of the step before the segment.
  __device__ void segment(const ErkKArgs& a, int c, int K, const double* sg, const T* sy,
                          double tl) {
    const int lo = c * K;
    const int hi = min(lo + K, n);
    if (c == c_last) setup_rows(a, sg[(int64_t)(hi - 1 - lo) * 2 + 1]);
    // The records are read one step ahead: a lane walks them serially, so a load issued
    // where it is used would expose its full latency once per step.
    double nts0 = sg[(int64_t)(hi - 1 - lo) * 2];
    double nts1 = sg[(int64_t)(hi - 1 - lo) * 2 + 1];
    T ny0[D];
#pragma unroll
    for (int j = 0; j < D; ++j) ny0[j] = sy[(int64_t)(hi - 1 - lo) * D + j];
    for (int s = hi - 1; s >= lo; --s) {
      const int i = s - lo;
      const double ts0 = nts0, ts1 = nts1;
      T y0[D], zs[S][D], k[S][D];
#pragma unroll
      for (int j = 0; j < D; ++j) y0[j] = ny0[j];
      if (i > 0) {
        nts0 = sg[(int64_t)(i - 1) * 2];
        nts1 = sg[(int64_t)(i - 1) * 2 + 1];
#pragma unroll
        for (int j = 0; j < D; ++j) ny0[j] = sy[(int64_t)(i - 1) * D + j];
      }
      const double prev_end = i > 0 ? nts1 : tl;  // end of the previous accepted step
      const double dt = ts1 - ts0;
      const T control = T((double)dir * dt);
      // the step's stages, as the forward computes them
      static_for<0, S>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
#pragma unroll
        for (int j = 0; j < D; ++j) {
          T kc[S];
#pragma unroll
          for (int m = 0; m < S; ++m) kc[m] = (m < i) ? k[m][j] : T(0);
          zs[i][j] = y0[j] + comb<Tab, 0, i, 0, i, false>(tr.a[i], kc, T(0));
        }
        constexpr double ci = Tab::c[i];
        const double tsi = (ci == 1.0) ? ts1 : ts0 + ci * dt;
        T f[D];
        VF::eval(tsi * dir, zs[i], q, f);
#pragma unroll
        for (int j = 0; j < D; ++j) k[i][j] = control * f[j];
      });
      T gy1[D], gy0[D], gk[S][D];
#pragma unroll
      for (int j = 0; j < D; ++j) {
        gy1[j] = ab[j];
        gy0[j] = T(0);
#pragma unroll
        for (int m = 0; m < S; ++m) gk[m][j] = T(0);
      }
      if (ksave > 0 && (s + 1) % ksave == 0) {  // the step's end state (SaveAt(steps=k))
#pragma unroll
        for (int j = 0; j < D; ++j) gy1[j] = gy1[j] + grow[j];
        --r;
        if (r >= 0) {
#pragma unroll
          for (int j = 0; j < D; ++j) grow[j] = g[(int64_t)r * D + j];
        }
      }
      // SaveAt(ts) rows inside this step: t in (end of the previous step, ts1]
      while (ti >= 0 && tsv > prev_end) {
        erk_interp_vjp<Tab, T, D>(ts0, ts1, tsv, grow, gy0, gy1, gk);
        --ti;
        --r;
        tsv = (ti >= 0) ? tsr[ti] * dir : -__builtin_inf();
        if (r >= 0) {
#pragma unroll
          for (int j = 0; j < D; ++j) grow[j] = g[(int64_t)r * D + j];
        }
      }
      // y1 = y0 + sum_m b_m k_m
#pragma unroll
      for (int j = 0; j < D; ++j) gy0[j] = gy0[j] + gy1[j];
      static_for<0, S>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        if constexpr (Tab::b[m] != 0.0) {
#pragma unroll
          for (int j = 0; j < D; ++j) gk[m][j] = fma_t(T(Tab::b[m]), gy1[j], gk[m][j]);
        }
      });
      // stages, last first
      static_for<0, S>([&](auto ic) {
        constexpr int i = S - 1 - decltype(ic)::value;
        T w[D], gz[D], gqi[P];
#pragma unroll
        for (int j = 0; j < D; ++j) w[j] = control * gk[i][j];
        VF::vjp(zs[i], q, w, gz, gqi);
#pragma unroll
        for (int m = 0; m < P; ++m) gq[m] = gq[m] + gqi[m];
#pragma unroll
        for (int j = 0; j < D; ++j) gy0[j] = gy0[j] + gz[j];
        static_for<0, i>([&](auto lc) {
          constexpr int l = decltype(lc)::value;
          if constexpr (Tab::a[i][l] != 0.0) {
#pragma unroll
            for (int j = 0; j < D; ++j) gk[l][j] = fma_t(T(Tab::a[i][l]), gz[j], gk[l][j]);
          }
        });
      });
#pragma unroll
      for (int j = 0; j < D; ++j) ab[j] = gy0[j];
    }
  }

  __device__ void finish(const ErkKArgs& a, const DiscreteAdjArgs& b) {
    if (T0 == T1) {  // a zero-length solve saves y0 in every ts row
      for (int rr = a.save_t0; rr < a.save_t0 + a.n_ts; ++rr) {
#pragma unroll
        for (int j = 0; j < D; ++j) ab[j] = ab[j] + g[(int64_t)rr * D + j];
      }
    }
    if (a.save_t0) {
#pragma unroll
      for (int j = 0; j < D; ++j) ab[j] = ab[j] + g[j];
    }
    T* yb = reinterpret_cast<T*>(b.y0_bar) + id * D;
    T* qb = reinterpret_cast<T*>(b.q_bar) + id * P;
#pragma unroll
    for (int j = 0; j < D; ++j) yb[j] = ab[j];
#pragma unroll
    for (int m = 0; m < P; ++m) qb[m] = gq[m];
  }
};

// One workgroup of two waves per 64 instances: wave 0 replays, wave 1 sweeps, one segment
// apart.  At stage s the replay wave restarts segment c_last - s of each of its instances
// from the checkpoint and records its accepted steps in half (c & 1) of the instance's
// 2K-record scratch, while the sweep wave transposes segment c_last - s + 1 from the
// other half; a barrier ends every stage.  Each instance's segments are still swept
// newest first and every record is written one stage before it is read, so the result is
// that of replaying and sweeping each segment in turn; the two waves' dependency chains
// share a SIMD's issue slots, which one wave leaves more than half idle.
// ra: the forward's arguments with saves, checkpoint writes and events off and grid /
// grid_y pointing at the scratch (grid_stride = 2K records per instance).
template <class Tab, class VF, class T, int Ctl, bool Extras>
__global__ void __launch_bounds__(128) erk_adjoint_kernel(const ErkKArgs a, const ErkKArgs ra,
                                                          const DiscreteAdjArgs b) {
  constexpr int D = VF::D;
  const int role = threadIdx.x >> 6;  // 0: replay, 1: sweep (wave-uniform)
  const int64_t id = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const bool live = id < a.B;
  const int K = a.ckpt_every;
  const int n = live ? min(max(b.stats[id * 3 + 1], 0), a.max_steps) : 0;
  const int c_last = n > 0 ? (n - 1) / K : -1;
  __shared__ int s_cmax;
  if (threadIdx.x == 0) s_cmax = -1;
  __syncthreads();
  if (role == 0 && c_last >= 0) atomicMax(&s_cmax, c_last);
  __syncthreads();
  const int nstages = s_cmax + 2;  // the replay of the oldest segment, then its sweep
  if (role == 0) {
    ErkLane<Tab, VF, T, Ctl, Extras> L;
    if (live) L.replay_setup(ra, id);
    for (int s = 0; s < nstages; ++s) {
      const int c = c_last - s;
      if (live && c >= 0) {
        const int base = (c & 1) * K;  // records at [base, base + K)
        L.resume(ra, b.ckpt, c, base);
        const int target = base + min(K, n - c * K);
        while (L.num_acc < target && L.running(ra)) L.step(ra);
      }
      __syncthreads();
    }
  } else {
    AdjSweep<Tab, VF, T> W;
    if (live) W.setup(a, b, id, n, K);
    for (int s = 0; s < nstages; ++s) {
      const int c = c_last - s + 1;
      if (live && c >= 0 && c <= c_last) {
        const int64_t base = id * ra.grid_stride + (int64_t)(c & 1) * K;
        const double tl = b.ckpt[(id * a.n_ckpt + c) * (D + 8) + D + 4];
        W.segment(a, c, K, ra.grid + base * 2, reinterpret_cast<const T*>(ra.grid_y) + base * D,
                  tl);
      }
      __syncthreads();
    }
    if (live) W.finish(a, b);
  }
}

template <class Tab, class VF, class T, int Ctl, bool X>
static int launch_adjoint_v(const ErkKArgs& a, const ErkKArgs& ra, const DiscreteAdjArgs& b,
                            hipStream_t s) {
  const int64_t nb = (a.B + 63) / 64;
  hipLaunchKernelGGL((erk_adjoint_kernel<Tab, VF, T, Ctl, X>), dim3((unsigned)nb), dim3(128), 0,
                     s, a, ra, b);
  return hipGetLastError() == hipSuccess ? DFX_OK : DFX_E_HIP;
}

// The replay must repeat the forward's controller arithmetic: the default integral
// controller (and ConstantStepSize) on PID_FIFTH, any other PID on PID_GENERIC, and with
// step clipping the extras instantiation on PID_GENERIC (whose exponent-0.2 power is the
// same correctly rounded fifth root, with factors 1 for the p and d terms, so it gives
// PID_FIFTH's bits too).
template <class Tab, class VF, class T>
static int launch_adjoint(const ErkKArgs& a, const ErkKArgs& ra, const DiscreteAdjArgs& b,
                          hipStream_t s) {
  const bool x = a.n_step_ts > 0 || a.n_jump_ts > 0;
  const bool fifth = a.controller != DFX_PID || pid_is_fifth(a.pid);
  if (x) return launch_adjoint_v<Tab, VF, T, PID_GENERIC, true>(a, ra, b, s);
  if (fifth) return launch_adjoint_v<Tab, VF, T, PID_FIFTH, false>(a, ra, b, s);
  return launch_adjoint_v<Tab, VF, T, PID_GENERIC, false>(a, ra, b, s);
}

template <class Tab, class T>
static int dispatch_adjoint_vf(int vf, int d, const ErkKArgs& a, const ErkKArgs& ra,
                               const DiscreteAdjArgs& b, hipStream_t s) {
  switch (vf) {
    case VF_LV:
      if (d == 2) return launch_adjoint<Tab, LotkaVolterraVF<T>, T>(a, ra, b, s);
      break;
    case VF_VDP:
      if (d == 2) return launch_adjoint<Tab, VanDerPolVF<T>, T>(a, ra, b, s);
      break;
    case VF_LINEAR:
      switch (d) {
        case 1: return launch_adjoint<Tab, LinearVF<T, 1>, T>(a, ra, b, s);
        case 2: return launch_adjoint<Tab, LinearVF<T, 2>, T>(a, ra, b, s);
        case 3: return launch_adjoint<Tab, LinearVF<T, 3>, T>(a, ra, b, s);
        case 4: return launch_adjoint<Tab, LinearVF<T, 4>, T>(a, ra, b, s);
        default: break;
      }
      break;
    default: break;
  }
  return DFX_E_ARG;
}

template <class T>
static int dispatch_adjoint(int solver, int vf, int d, const ErkKArgs& a, const ErkKArgs& ra,
                            const DiscreteAdjArgs& b, hipStream_t s) {
  switch (solver) {
    case SOLVER_TSIT5: return dispatch_adjoint_vf<Tsit5Tab, T>(vf, d, a, ra, b, s);
    case SOLVER_DOPRI5: return dispatch_adjoint_vf<Dopri5Tab, T>(vf, d, a, ra, b, s);
    case SOLVER_HEUN: return dispatch_adjoint_vf<HeunTab, T>(vf, d, a, ra, b, s);
    default: return DFX_E_ARG;
  }
}

int erk_adjoint_f64(int solver, int vf, int d, const ErkKArgs& a, const ErkKArgs& ra,
                    const DiscreteAdjArgs& b, hipStream_t s);
int erk_adjoint_f32(int solver, int vf, int d, const ErkKArgs& a, const ErkKArgs& ra,
                    const DiscreteAdjArgs& b, hipStream_t s);

}  // namespace dfx
