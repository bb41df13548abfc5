// erk_small_dispatch.hpp -- instantiates the small-system kernel for one state type and
// maps (solver, vector field, d) to a kernel.  Included once per precision build.
#pragma once
#include "erk_small.hpp"

namespace dfx {

template <class Tab, class VF, class T, int G, bool X>
static void launch_small_g(const ErkKArgs& a, int block, int64_t lanes, hipStream_t s) {
  if (!X && a.ckpt) {  // plain path with checkpoints: its own build, one lane per instance
    const int64_t nb = (a.B + block - 1) / block;
    hipLaunchKernelGGL((erk_small_kernel<Tab, VF, T, G, X, true>), dim3((unsigned)nb), dim3(block),
                       0, s, a);
    return;
  }
  const int64_t nblocks = (lanes + block - 1) / block;
  if (lanes >= a.B) {
    hipLaunchKernelGGL((erk_small_kernel<Tab, VF, T, G, X>), dim3((unsigned)nblocks), dim3(block),
                       0, s, a);
  } else {
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(1), 0, s, a.counter);
    hipLaunchKernelGGL((erk_small_queue_kernel<Tab, VF, T, G, X>), dim3((unsigned)nblocks), dim3(block),
                       0, s, a);
  }
}

// Variants: the controller kind (the default integral controller with exponent 0.2 on
// PID_FIFTH, every other PID on PID_GENERIC; ConstantStepSize takes PID_FIFTH, whose
// controller code it never runs) and dense output / step clipping / events / resumed or
// exported solver state (X).
template <class Tab, class VF, class T>
static int launch_small(const ErkKArgs& a, int block, int64_t lanes, hipStream_t s) {
  const bool g = a.controller == DFX_PID && !pid_is_fifth(a.pid);
  const bool x = a.dense_ts != nullptr || a.n_step_ts > 0 || a.n_jump_ts > 0 || a.n_events > 0 ||
                 a.resume != 0 || a.state_out != nullptr;
  if (g) {
    if (x) launch_small_g<Tab, VF, T, PID_GENERIC, true>(a, block, lanes, s);
    else launch_small_g<Tab, VF, T, PID_GENERIC, false>(a, block, lanes, s);
  } else {
    if (x) launch_small_g<Tab, VF, T, PID_FIFTH, true>(a, block, lanes, s);
    else launch_small_g<Tab, VF, T, PID_FIFTH, false>(a, block, lanes, s);
  }
  return hipGetLastError() == hipSuccess ? DFX_OK : DFX_E_HIP;
}

template <class Tab, class T>
static int dispatch_vf(int vf, int d, const ErkKArgs& a, int block, int64_t lanes, hipStream_t s) {
  switch (vf) {
    case VF_LV:
      if (d == 2) return launch_small<Tab, LotkaVolterraVF<T>, T>(a, block, lanes, s);
      break;
    case VF_VDP:
      if (d == 2) return launch_small<Tab, VanDerPolVF<T>, T>(a, block, lanes, s);
      break;
    case VF_LINEAR:
      switch (d) {
        case 1: return launch_small<Tab, LinearVF<T, 1>, T>(a, block, lanes, s);
        case 2: return launch_small<Tab, LinearVF<T, 2>, T>(a, block, lanes, s);
        case 3: return launch_small<Tab, LinearVF<T, 3>, T>(a, block, lanes, s);
        case 4: return launch_small<Tab, LinearVF<T, 4>, T>(a, block, lanes, s);
        case 10: return launch_small<Tab, LinearVF<T, 10>, T>(a, block, lanes, s);
        default: break;
      }
      break;
    default: break;
  }
  return DFX_E_ARG;
}

// Continuous adjoint (BacksolveAdjoint) kernels: the same (solver, vector field, d) map.
template <class Tab, class VF, class T>
static int launch_backsolve(const ErkKArgs& a, const BacksolveArgs& b, hipStream_t s) {
  const int64_t nb = (a.B + 63) / 64;
  hipLaunchKernelGGL((erk_backsolve_kernel<Tab, VF, T>), dim3((unsigned)nb), dim3(64), 0, s, a, b);
  return hipGetLastError() == hipSuccess ? DFX_OK : DFX_E_HIP;
}

template <class Tab, class T>
static int dispatch_backsolve_vf(int vf, int d, const ErkKArgs& a, const BacksolveArgs& b,
                                 hipStream_t s) {
  switch (vf) {
    case VF_LV:
      if (d == 2) return launch_backsolve<Tab, LotkaVolterraVF<T>, T>(a, b, s);
      break;
    case VF_VDP:
      if (d == 2) return launch_backsolve<Tab, VanDerPolVF<T>, T>(a, b, s);
      break;
    case VF_LINEAR:
      if (d == 1) return launch_backsolve<Tab, LinearVF<T, 1>, T>(a, b, s);
      if (d == 2) return launch_backsolve<Tab, LinearVF<T, 2>, T>(a, b, s);
      if (d == 3) return launch_backsolve<Tab, LinearVF<T, 3>, T>(a, b, s);
      break;
    default: break;
  }
  return DFX_E_ARG;
}

template <class T>
static int dispatch_backsolve(int solver, int vf, int d, const ErkKArgs& a, const BacksolveArgs& b,
                             hipStream_t s) {
  switch (solver) {
    case SOLVER_TSIT5: return dispatch_backsolve_vf<Tsit5Tab, T>(vf, d, a, b, s);
    case SOLVER_DOPRI5: return dispatch_backsolve_vf<Dopri5Tab, T>(vf, d, a, b, s);
    case SOLVER_HEUN: return dispatch_backsolve_vf<HeunTab, T>(vf, d, a, b, s);
    default: return DFX_E_ARG;
  }
}

template <class T>
static int dispatch_small(int solver, int vf, int d, const ErkKArgs& a, int block, int64_t lanes,
                          hipStream_t s) {
  switch (solver) {
    case SOLVER_TSIT5: return dispatch_vf<Tsit5Tab, T>(vf, d, a, block, lanes, s);
    case SOLVER_DOPRI5: return dispatch_vf<Dopri5Tab, T>(vf, d, a, block, lanes, s);
    case SOLVER_HEUN: return dispatch_vf<HeunTab, T>(vf, d, a, block, lanes, s);
    default: return DFX_E_ARG;
  }
}

}  // namespace dfx
