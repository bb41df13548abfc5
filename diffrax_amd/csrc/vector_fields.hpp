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
  // w^T df/dy and w^T df/dq for a cotangent w
  __device__ __forceinline__ static void vjp(const T (&y)[D], const T (&q)[P], const T (&w)[D],
                                             T (&gy)[D], T (&gq)[P]) {
    const T x = y[0], v = y[1];
    gy[0] = w[0] * (q[0] - q[1] * v) + w[1] * (q[3] * v);
    gy[1] = w[0] * (-(q[1] * x)) + w[1] * (q[3] * x - q[2]);
    gq[0] = w[0] * x;
    gq[1] = -(w[0] * (x * v));
    gq[2] = -(w[1] * v);
    gq[3] = w[1] * (x * v);
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
  __device__ __forceinline__ static void vjp(const T (&y)[D], const T (&q)[P], const T (&w)[D],
                                             T (&gy)[D], T (&gq)[P]) {
    const T x = y[0], v = y[1];
    gy[0] = w[1] * (T(-2) * q[0] * x * v - T(1));
    gy[1] = w[0] + w[1] * (q[0] * (T(1) - x * x));
    gq[0] = w[1] * ((T(1) - x * x) * v);
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
  __device__ __forceinline__ static void vjp(const T (&y)[D], const T (&q)[P], const T (&w)[D],
                                             T (&gy)[D], T (&gq)[P]) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      T s = T(0);
#pragma unroll
      for (int i = 0; i < D; ++i) s = s + q[i * D + j] * w[i];
      gy[j] = s;
    }
#pragma unroll
    for (int i = 0; i < D; ++i) {
#pragma unroll
      for (int j = 0; j < D; ++j) gq[i * D + j] = w[i] * y[j];
    }
  }
};

// The augmented vector field of the continuous adjoint (AdjointTerm.vf_prod,
// _term.py:893-917): state (y, a_y, a_q), f_aug = (f(y), -a_y^T df/dy, -a_y^T df/dq).
template <class VF, class T>
struct AdjointVF {
  static constexpr int DB = VF::D;
  static constexpr int P = VF::P;
  static constexpr int D = 2 * DB + P;
  __device__ __forceinline__ static void eval(double t, const T (&z)[D], const T (&q)[P],
                                              T (&f)[D]) {
    T y[DB], w[DB], fy[DB], gy[DB], gq[P];
#pragma unroll
    for (int j = 0; j < DB; ++j) { y[j] = z[j]; w[j] = -z[DB + j]; }
    VF::eval(t, y, q, fy);
    VF::vjp(y, q, w, gy, gq);
#pragma unroll
    for (int j = 0; j < DB; ++j) { f[j] = fy[j]; f[DB + j] = gy[j]; }
#pragma unroll
    for (int m = 0; m < P; ++m) f[2 * DB + m] = gq[m];
  }
};

}  // namespace dfx
