// erk_dense.hpp -- evaluation of the dense output a SaveAt(dense=True) solve recorded:
// Solution.evaluate / Solution.derivative for every instance of a batch at query times.
//
// Replaces DenseInterpolation.evaluate / derivative (_global_interpolation.py:315-400):
// the step holding t is found by searchsorted over the accepted-step end times (side
// "left" or "right"), clipped to the recorded steps (_global_interpolation.py:38-46);
// t outside [ts[0], ts[last]] or a solve with no accepted step gives NaN, except the
// zero-length-solve case that returns y0 (_global_interpolation.py:355-356); the step's
// local interpolant (the same erk_interp the solve kernel saves with, so evaluate at a
// SaveAt(ts) time reproduces that saved value bit for bit) is evaluated at t, and the
// derivative is the interpolant's t-derivative times the direction.
//
// One thread per (instance, query); the dense buffers are read once per query.
#pragma once
#include "erk_small.hpp"

namespace dfx {

struct DenseKArgs {
  const double* dense_ts;  // [B, max_steps + 1]
  const void* dense_ys;    // [B, max_steps + 1, D]
  const void* dense_k;     // [B, max_steps, S, D]
  const int32_t* stats;    // [B, 3]; column 1 = accepted steps
  const double* t1;        // [B * t1_stride]
  const double* tq;        // [B * tq_stride + n_q]
  void* out;               // [B, n_q, D]
  int64_t B, n_q, tq_stride, t1_stride;
  int32_t max_steps, left, derivative;
  double t0;
};

template <class Tab, class T, int D>
__global__ void __launch_bounds__(256) erk_dense_eval_kernel(const DenseKArgs a) {
  constexpr int S = Tab::S;
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= a.B * a.n_q) return;
  const int64_t b = gid / a.n_q;
  const int64_t q = gid - b * a.n_q;
  const double t1 = a.t1[b * a.t1_stride];
  const int dir = (a.t0 < t1) ? 1 : -1;
  const double t = a.tq[b * a.tq_stride + q] * dir;
  const int n = a.stats[b * 3 + 1];          // accepted steps: ts_size = n + 1
  const int64_t row0 = b * ((int64_t)a.max_steps + 1);
  const double* ts = a.dense_ts + row0;
  const T* ys = reinterpret_cast<const T*>(a.dense_ys) + row0 * D;
  T* out = reinterpret_cast<T*>(a.out) + gid * D;
  const double ts_first = ts[0], ts_last = ts[n];
  if (n < 1 || t < ts_first || t > ts_last || t != t) {
    // out of bounds (NaN), or the zero-length solve evaluated at t0 (y0).  As in the
    // reference, the trivial case compares the direction-scaled query with the unscaled
    // t0 (with t0 == t1 the direction is -1).
    const bool trivial = (n == 0) && (t == a.t0) && !a.derivative;
#pragma unroll
    for (int j = 0; j < D; ++j) out[j] = trivial ? ys[j] : T(__builtin_nan(""));
    return;
  }
  // searchsorted over ts[0..n] (the buffer beyond n holds no accepted step)
  int lo = 0, hi = n + 1;  // first index with ts[i] >= t (left) or > t (right)
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const bool before = a.left ? (ts[mid] < t) : (ts[mid] <= t);
    if (before) lo = mid + 1; else hi = mid;
  }
  int idx = lo - 1;
  idx = idx < 0 ? 0 : (idx > n - 1 ? n - 1 : idx);
  T y0[D], y1[D], k[S][D], v[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    y0[j] = ys[(int64_t)idx * D + j];
    y1[j] = ys[(int64_t)(idx + 1) * D + j];
  }
  const T* kp = reinterpret_cast<const T*>(a.dense_k) +
                ((b * (int64_t)a.max_steps + idx) * S) * D;
#pragma unroll
  for (int m = 0; m < S; ++m) {
#pragma unroll
    for (int j = 0; j < D; ++j) k[m][j] = kp[m * D + j];
  }
  if (a.derivative) {
    erk_interp_deriv<Tab, T, D>(ts[idx], ts[idx + 1], y0, y1, k, t, v);
#pragma unroll
    for (int j = 0; j < D; ++j) out[j] = T(dir) * v[j];
  } else {
    erk_interp<Tab, T, D>(ts[idx], ts[idx + 1], y0, y1, k, t, v);
#pragma unroll
    for (int j = 0; j < D; ++j) out[j] = v[j];
  }
}

template <class Tab, class T>
static int dense_dispatch_d(int d, const DenseKArgs& a, hipStream_t s) {
  const int64_t n = a.B * a.n_q;
  const unsigned nb = (unsigned)((n + 255) / 256);
#define DFX_DENSE(DV) \
  hipLaunchKernelGGL((erk_dense_eval_kernel<Tab, T, DV>), dim3(nb), dim3(256), 0, s, a)
  switch (d) {
    case 1: DFX_DENSE(1); break;
    case 2: DFX_DENSE(2); break;
    case 3: DFX_DENSE(3); break;
    case 4: DFX_DENSE(4); break;
    case 10: DFX_DENSE(10); break;
    default: return DFX_E_ARG;
  }
#undef DFX_DENSE
  return hipGetLastError() == hipSuccess ? DFX_OK : DFX_E_HIP;
}

template <class T>
static int dense_dispatch(int solver, int d, const DenseKArgs& a, hipStream_t s) {
  switch (solver) {
    case SOLVER_TSIT5: return dense_dispatch_d<Tsit5Tab, T>(d, a, s);
    case SOLVER_DOPRI5: return dense_dispatch_d<Dopri5Tab, T>(d, a, s);
    case SOLVER_HEUN: return dense_dispatch_d<HeunTab, T>(d, a, s);
    default: return DFX_E_ARG;
  }
}

int erk_dense_eval_f64(int solver, int d, const DenseKArgs& a, hipStream_t s);
int erk_dense_eval_f32(int solver, int d, const DenseKArgs& a, hipStream_t s);

}  // namespace dfx
