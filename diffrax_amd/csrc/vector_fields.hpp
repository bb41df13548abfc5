// vector_fields.hpp -- registered device vector fields for the small-system kernels.
//
// Each functor evaluates f(t, y) for one trajectory held in registers; it stands in for
// the user callable that ODETerm.vf calls (_term.py:193-211).  Parameters are the
// per-instance (or shared) `args` of the solve, loaded once per trajectory.
#pragma once
#include "erk_common.hpp"

namespace dfx {

enum VfId : int32_t { VF_LV = 0, VF_VDP = 1, VF_LINEAR = 2, VF_MLP = 3, VF_OU = 4 };

// Lotka-Volterra: dx = alpha x - beta x y ; dy = delta x y - gamma y.
// q = [alpha, beta, gamma, delta].
template <class T>
struct LotkaVolterraVF {
  static constexpr int D = 2;
  static constexpr int P = 4;
  __device__ __forceinline__ static void eval(double, const T (&y)[D], const T (&q)[P],
                                              T (&f)[D]) {
    const T x = y[0], v = y[1];
    f[0] = q[0] * x - q[1] * x * v;
    f[1] = q[3] * x * v - q[2] * v;
  }
};

// Van der Pol: y'' = mu (1 - y^2) y' - y, state (y, y').  q = [mu].
template <class T>
struct VanDerPolVF {
  static constexpr int D = 2;
  static constexpr int P = 1;
  __device__ __forceinline__ static void eval(double, const T (&y)[D], const T (&q)[P],
                                              T (&f)[D]) {
    const T x = y[0], v = y[1];
    f[0] = v;
    f[1] = q[0] * (T(1) - x * x) * v - x;
  }
};

// Linear: f = A y with A (row-major D x D) as the parameters.
template <class T, int DD>
struct LinearVF {
  static constexpr int D = DD;
  static constexpr int P = DD * DD;
  __device__ __forceinline__ static void eval(double, const T (&y)[D], const T (&q)[P],
                                              T (&f)[D]) {
#pragma unroll
    for (int i = 0; i < D; ++i) {
      T s = T(0);
#pragma unroll
      for (int j = 0; j < D; ++j) s = s + q[i * D + j] * y[j];
      f[i] = s;
    }
  }
};

}  // namespace dfx
