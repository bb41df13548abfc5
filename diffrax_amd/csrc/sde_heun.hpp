// sde_heun.hpp -- kernels of the Heun (Stratonovich) SDE path with diagonal noise for the
// registered Ornstein-Uhlenbeck terms
//     dy = theta (mu - y) dt + sigma dW      (theta, mu, sigma per component).
//
// Replaces, for MultiTerm(ODETerm(drift), ControlTerm(diag, path)) + Heun +
// ConstantStepSize, the reference loop (_integrate.py:302-885) running
// AbstractRungeKutta.step with the Heun tableau (_solver/heun.py:12-17,
// _solver/runge_kutta.py:847-1203): k = f dt + g dW (MultiTerm.vf_prod, _term.py:711-722;
// diagonal ControlTerm.prod, _term.py:423-427), z = y + k0, y1 = y + (k0/2 + k1/2).
// The ConstantStepSize grid (constant.py:30-104, with _clip_to_end) is computed on the
// host and passed in, so the kernels only read dt_n = t_{n+1} - t_n.
//
// Layout: increments dW[N, P, d] float32 (step-major, so one step of all paths is a
// contiguous slab); states [P, d].  Forward: each thread owns V consecutive components of
// one path and walks all N steps, streaming its dW slice with U steps of loads in flight.
// The solve is HBM-bound: 4 bytes of dW per component-step.
//
// Gradient (RecursiveCheckpointAdjoint semantics, _adjoint.py:264-312): the forward pass
// stores the state every CK steps; the adjoint pass gives one component of one path to a
// thread, walks the segments backwards, recomputes a segment from its checkpoint keeping
// states and increments in registers, then runs the exact reverse-mode recursion of the
// Heun step through it.  The step sizes of the current segment sit in LDS (shared by the
// block).  theta/mu/sigma cotangents are summed per segment in the kernel's precision,
// folded into float64 per thread, per block in LDS, then across blocks in a fixed order
// (deterministic).
//
// The float64 instantiations are compiled without FMA contraction (bit parity with the
// FP64 oracle); the float32 ones with contraction, in their own translation unit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "erk_common.hpp"

namespace dfx {

struct SdeArgs {
  const void* y0;         // [P, d]
  const void* theta;      // [d]
  const void* mu;         // [d]
  const void* sigma;      // [d]
  const float* dW;        // [N, P, d]
  const double* grid;     // [N + 1]
  void* ys;               // [P, n_out, d]
  void* ckpt;             // [n_ckpt, P, d] (states at steps CK, 2CK, ...) or null
  const void* ys_bar;     // [P, n_out, d]
  void* y0_bar;           // [P, d] or null
  double* partial;        // [n_blocks, 3, d]
  int64_t P;
  int32_t d, N, CK, n_out, save_t0, save_t1;
  int32_t save_steps;     // SaveAt(steps=k): y after every k-th step
  int32_t n_ts;           // SaveAt(ts): ascending times in [t0, t1], shared by the paths
  const double* ts;       // [n_ts]
  void* dense_ys;         // SaveAt(dense=True): [P, N + 1, d] step-end states (or null)
  void* dense_k;          // [P, N, 2, d] each step's two stage increments
};

// Query side of the SDE dense output: one thread per (path, query, component).
struct SdeDenseArgs {
  const void* dense_ys;
  const void* dense_k;
  const double* grid;     // [N + 1]
  const double* tq;       // [P * tq_stride + n_q]
  void* out;              // [P, n_q, d]
  int64_t P, n_q, tq_stride;
  int32_t d, N, left, derivative;
};

// Saved rows in time order: t0; then per step n the ts inside it (t_n < ts <= t_{n+1}, or
// ts >= t0 for the first), then y_{n+1} when k divides n + 1; then t1 unless k divides N.
// The ts with index i, inside step n, is row save_t0 + i + n / k; the step row of step n is
// save_t0 + (number of ts <= t_{n+1}) + (n + 1) / k - 1.

// Row of the t1 save, and whether there is one: SaveAt(steps=k) already saved the last
// state when k divides N (the _save_t1 rule, _integrate.py:847-877).
__host__ __device__ inline int sde_t1_row(const SdeArgs& a) {
  return a.save_t0 + a.n_ts + (a.save_steps > 0 ? a.N / a.save_steps : 0);
}
__host__ __device__ inline bool sde_has_t1_row(const SdeArgs& a) {
  return a.save_t1 && !(a.save_steps > 0 && a.N % a.save_steps == 0);
}

template <class T, int V>
struct Vec {
  T v[V];
};

// dW is streamed exactly once per pass: nontemporal loads keep it from displacing
// reusable lines in the caches (measured +14 % on the 16.8 GB forward stream).
typedef float dw_f4 __attribute__((ext_vector_type(4)));
typedef float dw_f2 __attribute__((ext_vector_type(2)));

template <class T, int V>
__device__ __forceinline__ Vec<T, V> load_dw(const float* __restrict__ p) {
  Vec<T, V> r;
  if constexpr (V == 4) {
    const dw_f4 x = __builtin_nontemporal_load(reinterpret_cast<const dw_f4*>(p));
    r.v[0] = (T)x[0]; r.v[1] = (T)x[1]; r.v[2] = (T)x[2]; r.v[3] = (T)x[3];
  } else if constexpr (V == 2) {
    const dw_f2 x = __builtin_nontemporal_load(reinterpret_cast<const dw_f2*>(p));
    r.v[0] = (T)x[0]; r.v[1] = (T)x[1];
  } else {
    r.v[0] = (T)__builtin_nontemporal_load(p);
  }
  return r;
}

// One Heun step of one component: k0 = dt*(th*(mu - y)) + sg*w ; z = y + 1*k0 ;
// k1 = dt*(th*(mu - z)) + sg*w ; y1 = y + (0.5*k0 + 0.5*k1).
template <class T>
__device__ __forceinline__ T heun_ou(T y, T dt, T th, T mu, T sg, T w) {
  const T k0 = dt * (th * (mu - y)) + sg * w;
  const T z = y + k0;
  const T k1 = dt * (th * (mu - z)) + sg * w;
  return y + (T(0.5) * k0 + T(0.5) * k1);
}

// The same step, also returning its two stage increments (the interpolant's end slopes).
template <class T>
__device__ __forceinline__ T heun_ou_k(T y, T dt, T th, T mu, T sg, T w, T& k0, T& k1) {
  k0 = dt * (th * (mu - y)) + sg * w;
  const T z = y + k0;
  k1 = dt * (th * (mu - z)) + sg * w;
  return y + (T(0.5) * k0 + T(0.5) * k1);
}

// Heun's dense output inside a step at theta (the third-order Hermite polynomial from the
// end values and the stage increments, ThirdOrderHermitePolynomialInterpolation.from_k,
// _local_interpolation.py:50-139; the small-system kernel's erk_interp in the same form).
template <class T>
__device__ __forceinline__ T heun_interp(T y0, T y1, T k0, T k1, double th) {
  const T tt = T(th);
  const T ca = k0 + k1 + T(2) * y0 - T(2) * y1;
  const T cb = T(-2) * k0 - k1 - T(3) * y0 + T(3) * y1;
  return ((ca * tt + cb) * tt + k0) * tt + y0;
}

// SAVES: rows between t0 and t1 (SaveAt(ts) or SaveAt(steps)), a build of its own.
template <class T, int V, int U, bool SAVES>
__global__ void __launch_bounds__(256) sde_heun_fwd_kernel(const SdeArgs a) {
  const int groups = a.d / V;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= a.P * groups) return;
  const int64_t p = tid / groups;
  const int j0 = (int)(tid - p * groups) * V;
  const T* th_p = reinterpret_cast<const T*>(a.theta) + j0;
  const T* mu_p = reinterpret_cast<const T*>(a.mu) + j0;
  const T* sg_p = reinterpret_cast<const T*>(a.sigma) + j0;
  T th[V], mu[V], sg[V], y[V];
  const T* y0 = reinterpret_cast<const T*>(a.y0) + p * a.d + j0;
#pragma unroll
  for (int v = 0; v < V; ++v) { th[v] = th_p[v]; mu[v] = mu_p[v]; sg[v] = sg_p[v]; y[v] = y0[v]; }
  T* ys = reinterpret_cast<T*>(a.ys) + p * (int64_t)a.n_out * a.d + j0;
  if (a.save_t0) {
#pragma unroll
    for (int v = 0; v < V; ++v) ys[v] = y[v];
  }
  T* dys = SAVES && a.dense_ys ? reinterpret_cast<T*>(a.dense_ys) + p * (int64_t)(a.N + 1) * a.d + j0
                               : nullptr;
  T* dks = SAVES && a.dense_ys ? reinterpret_cast<T*>(a.dense_k) + p * (int64_t)a.N * 2 * a.d + j0
                               : nullptr;
  if (dys) {  // the dense output's first knot (_integrate.py:529-540)
#pragma unroll
    for (int v = 0; v < V; ++v) dys[v] = y[v];
  }
  const int64_t slab = a.P * a.d;  // elements per step
  const float* dwp = a.dW + p * a.d + j0;
  T* ck = reinterpret_cast<T*>(a.ckpt);
  int next_ck = a.CK;  // step count at which the next checkpoint is due
  int next_save = a.save_steps, save_row = a.save_t0, ts_idx = 0;  // SaveAt(steps=k, ts)
  for (int n0 = 0; n0 < a.N; n0 += U) {
    Vec<T, V> w[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (n0 + u < a.N) w[u] = load_dw<T, V>(dwp + (int64_t)(n0 + u) * slab);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int n = n0 + u;
      if (n < a.N) {
        const T dt = (T)(a.grid[n + 1] - a.grid[n]);
        if constexpr (SAVES) {
          T k0[V], k1[V], y1[V];
#pragma unroll
          for (int v = 0; v < V; ++v)
            y1[v] = heun_ou_k<T>(y[v], dt, th[v], mu[v], sg[v], w[u].v[v], k0[v], k1[v]);
          const double tn = a.grid[n], tn1 = a.grid[n + 1];
          while (ts_idx < a.n_ts && a.ts[ts_idx] <= tn1) {
            const double thv = linear_rescale(tn, a.ts[ts_idx], tn1);
            T* o = ys + (int64_t)save_row * a.d;
#pragma unroll
            for (int v = 0; v < V; ++v) o[v] = heun_interp<T>(y[v], y1[v], k0[v], k1[v], thv);
            ++save_row;
            ++ts_idx;
          }
          if (dys) {  // the step's end state and stage increments
            T* oy = dys + (int64_t)(n + 1) * a.d;
            T* ok = dks + (int64_t)n * 2 * a.d;
#pragma unroll
            for (int v = 0; v < V; ++v) { oy[v] = y1[v]; ok[v] = k0[v]; ok[a.d + v] = k1[v]; }
          }
#pragma unroll
          for (int v = 0; v < V; ++v) y[v] = y1[v];
        } else {
#pragma unroll
          for (int v = 0; v < V; ++v) y[v] = heun_ou<T>(y[v], dt, th[v], mu[v], sg[v], w[u].v[v]);
        }
        if (SAVES && a.save_steps && n + 1 == next_save) {
          T* o = ys + (int64_t)save_row * a.d;
#pragma unroll
          for (int v = 0; v < V; ++v) o[v] = y[v];
          ++save_row;
          next_save += a.save_steps;
        }
        if (ck && n + 1 == next_ck) {
          if (n + 1 < a.N) {
            T* c = ck + ((int64_t)(next_ck / a.CK - 1) * a.P + p) * a.d + j0;
#pragma unroll
            for (int v = 0; v < V; ++v) c[v] = y[v];
          }
          next_ck += a.CK;
        }
      }
    }
  }
  if (sde_has_t1_row(a)) {
    T* o = ys + (int64_t)sde_t1_row(a) * a.d;
#pragma unroll
    for (int v = 0; v < V; ++v) o[v] = y[v];
  }
}

// SaveAt(dense=True) evaluation (DenseInterpolation.evaluate / derivative,
// _global_interpolation.py:315-400): the step holding t on the shared grid (left: a time
// on a boundary belongs to the earlier step), then the step's Hermite polynomial as the
// in-kernel saves evaluate it, or its time derivative.  Outside [t0, t1]: NaN.
template <class T>
__global__ void __launch_bounds__(256) sde_dense_eval_kernel(const SdeDenseArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.P * a.n_q * a.d) return;
  const int j = (int)(i % a.d);
  const int64_t pq = i / a.d;
  const int64_t q = pq % a.n_q, p = pq / a.n_q;
  const double t = a.tq[p * a.tq_stride + q];
  const double* g = a.grid;
  T* o = reinterpret_cast<T*>(a.out) + i;
  if (!(t >= g[0] && t <= g[a.N])) {
    *o = T(__builtin_nan(""));
    return;
  }
  int lo = 0, hi = a.N - 1;  // first step whose end is at or after t (left) / after t
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const bool after = a.left ? (g[mid + 1] >= t) : (g[mid + 1] > t);
    if (after) hi = mid; else lo = mid + 1;
  }
  const int n = lo;
  const T* ys = reinterpret_cast<const T*>(a.dense_ys) + (p * (int64_t)(a.N + 1) + n) * a.d + j;
  const T* ks = reinterpret_cast<const T*>(a.dense_k) + (p * (int64_t)a.N + n) * 2 * a.d + j;
  const double th = linear_rescale(g[n], t, g[n + 1]);
  if (!a.derivative) {
    *o = heun_interp<T>(ys[0], ys[a.d], ks[0], ks[a.d], th);
  } else {
    const T y0 = ys[0], y1 = ys[a.d], k0 = ks[0], k1 = ks[a.d], tt = T(th);
    const T ca = k0 + k1 + T(2) * y0 - T(2) * y1;
    const T cb = T(-2) * k0 - k1 - T(3) * y0 + T(3) * y1;
    *o = ((T(3) * ca * tt + T(2) * cb) * tt + k0) / T(g[n + 1] - g[n]);
  }
}

// Reverse-mode of one Heun step for one component (see heun_ou).  Given the state y at the
// start of the step and the cotangent yb of y1, returns the cotangent of y and adds the
// step's parameter cotangents: sth += d/dtheta, smu += dt (k0b + k1b) (times theta gives
// d/dmu), ssg += d/dsigma.
template <class T>
__device__ __forceinline__ T heun_ou_vjp(T y, T dt, T th, T mu, T sg, T w, T yb, T& sth,
                                         T& smu, T& ssg) {
  const T k0 = dt * (th * (mu - y)) + sg * w;
  const T z = y + k0;
  const T c = dt * th;
  const T k1b = T(0.5) * yb;     // y1 = y + (k0/2 + k1/2)
  const T zb = -(c * k1b);       // k1 = dt*(th*(mu - z)) + sg*w
  const T k0b = k1b + zb;        // z = y + k0
  const T s = k0b + k1b;
  sth += dt * ((mu - z) * k1b + (mu - y) * k0b);
  smu += dt * s;
  ssg += w * s;
  return (yb + zb) - c * k0b;    // k0 = dt*(th*(mu - y)) + sg*w
}

// heun_ou_vjp with further cotangents on the step's start state and stage increments
// (from the SaveAt(ts) rows interpolated inside the step): yb0 on y, kb0 on k0, kb1 on k1.
template <class T>
__device__ __forceinline__ T heun_ou_vjp_k(T y, T dt, T th, T mu, T sg, T w, T yb, T yb0,
                                           T kb0, T kb1, T& sth, T& smu, T& ssg) {
  const T k0 = dt * (th * (mu - y)) + sg * w;
  const T z = y + k0;
  const T c = dt * th;
  const T k1b = T(0.5) * yb + kb1;
  const T zb = -(c * k1b);
  const T k0b = (T(0.5) * yb + kb0) + zb;
  const T s = k0b + k1b;
  sth += dt * ((mu - z) * k1b + (mu - y) * k0b);
  smu += dt * s;
  ssg += w * s;
  return ((yb + yb0) + zb) - c * k0b;
}

// One checkpoint segment of the adjoint for one component: recompute the cnt (<= CK) steps
// from the segment's initial state y keeping every step's start state and increment, then
// walk them backwards.  FULL = (cnt == CK) drops the per-step bounds tests.
// With saved rows between t0 and t1 (SAVES) the cotangent of each step row joins yb
// before the step that produced it is reversed, and each ts row inside the step passes
// its cotangent to the step's end values and stage increments through the interpolant's
// theta polynomials.  ts_hi: number of ts <= the segment's end time (updated to its start).
template <class T, int CK, bool FULL, bool SAVES>
__device__ __forceinline__ void adj_segment(T y, const float* __restrict__ dw, int64_t slab,
                                            int cnt, const T* dts, T th, T mu, T sg, T& yb,
                                            double& thb, double& mub, double& sgb,
                                            const SdeArgs& a, const T* ysb, int nbeg,
                                            int& ts_hi) {
  T yh[CK], w[CK];
#pragma unroll
  for (int c = 0; c < CK; ++c)
    if (FULL || c < cnt) w[c] = (T)__builtin_nontemporal_load(dw + (int64_t)c * slab);
#pragma unroll
  for (int c = 0; c < CK; ++c) {
    yh[c] = y;
    if (FULL || c < cnt) y = heun_ou<T>(y, dts[c], th, mu, sg, w[c]);
  }
  T sth = T(0), smu = T(0), ssg = T(0);
#pragma unroll
  for (int c = CK - 1; c >= 0; --c)
    if (FULL || c < cnt) {
      if constexpr (SAVES) {
        const int n = nbeg + c;  // this step; n + 1 steps are done after it
        const int k = a.save_steps;
        if (k && (n + 1) % k == 0)
          yb = yb + ysb[(int64_t)(a.save_t0 + ts_hi + (n + 1) / k - 1) * a.d];
        const double tn = a.grid[n], tn1 = a.grid[n + 1];
        int lo = ts_hi;
        while (lo > 0 && (n == 0 ? true : a.ts[lo - 1] > tn)) --lo;
        if (lo < ts_hi) {
          T k0, k1;
          (void)heun_ou_k<T>(yh[c], dts[c], th, mu, sg, w[c], k0, k1);
          T yb0 = T(0), kb0 = T(0), kb1 = T(0);
          for (int i = lo; i < ts_hi; ++i) {
            const T g = ysb[(int64_t)(a.save_t0 + i + (k ? n / k : 0)) * a.d];
            const T tt = T(linear_rescale(tn, a.ts[i], tn1));
            const T t2 = tt * tt, t3 = t2 * tt;
            yb = yb + (T(3) * t2 - T(2) * t3) * g;
            yb0 = yb0 + ((T(1) - T(3) * t2) + T(2) * t3) * g;
            kb0 = kb0 + ((tt - T(2) * t2) + t3) * g;
            kb1 = kb1 + (t3 - t2) * g;
          }
          ts_hi = lo;
          yb = heun_ou_vjp_k<T>(yh[c], dts[c], th, mu, sg, w[c], yb, yb0, kb0, kb1, sth, smu,
                                ssg);
        } else {
          yb = heun_ou_vjp<T>(yh[c], dts[c], th, mu, sg, w[c], yb, sth, smu, ssg);
        }
      } else {
        yb = heun_ou_vjp<T>(yh[c], dts[c], th, mu, sg, w[c], yb, sth, smu, ssg);
      }
    }
  thb += (double)sth;
  mub += (double)th * (double)smu;
  sgb += (double)ssg;
}

// SAVES: rows between t0 and t1 (SaveAt(ts) or SaveAt(steps=k)) carry cotangents (a build
// of its own, so the plain adjoint's unrolled segment loop has no per-step test).
template <class T, int CK, bool SAVES>
__global__ void __launch_bounds__(256) sde_heun_adj_kernel(const SdeArgs a) {
  __shared__ T dts[CK];
  extern __shared__ double red[];
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = tid < a.P * a.d;
  const int64_t p = live ? tid / a.d : 0;
  const int j = live ? (int)(tid - p * a.d) : 0;
  double thb = 0.0, mub = 0.0, sgb = 0.0;
  const T th = reinterpret_cast<const T*>(a.theta)[j];
  const T mu = reinterpret_cast<const T*>(a.mu)[j];
  const T sg = reinterpret_cast<const T*>(a.sigma)[j];
  const T* ysb = reinterpret_cast<const T*>(a.ys_bar) + p * (int64_t)a.n_out * a.d + j;
  T yb = (live && sde_has_t1_row(a)) ? ysb[(int64_t)sde_t1_row(a) * a.d] : T(0);
  const int64_t slab = a.P * a.d;
  const float* dwp = a.dW + (live ? tid : 0);
  const T* ck = reinterpret_cast<const T*>(a.ckpt);
  const T* y0 = reinterpret_cast<const T*>(a.y0);
  const int nseg = (a.N + CK - 1) / CK;
  int ts_hi = a.n_ts;  // all ts are <= t1
  for (int s = nseg - 1; s >= 0; --s) {
    const int nbeg = s * CK;
    const int cnt = min(CK, a.N - nbeg);
    __syncthreads();  // the previous segment's reads of dts are done
    if (threadIdx.x < (unsigned)cnt)
      dts[threadIdx.x] = (T)(a.grid[nbeg + threadIdx.x + 1] - a.grid[nbeg + threadIdx.x]);
    __syncthreads();
    if (live) {
      const T y = (s == 0) ? y0[tid] : ck[(int64_t)(s - 1) * slab + tid];
      if (cnt == CK)
        adj_segment<T, CK, true, SAVES>(y, dwp + (int64_t)nbeg * slab, slab, CK, dts, th, mu,
                                        sg, yb, thb, mub, sgb, a, ysb, nbeg, ts_hi);
      else
        adj_segment<T, CK, false, SAVES>(y, dwp + (int64_t)nbeg * slab, slab, cnt, dts, th, mu,
                                         sg, yb, thb, mub, sgb, a, ysb, nbeg, ts_hi);
    }
  }
  if (live) {
    if (a.save_t0) yb = yb + ysb[0];
    if (a.y0_bar) reinterpret_cast<T*>(a.y0_bar)[tid] = yb;
  }
  // Block reduction of the parameter cotangents over the paths of this block.
  // LDS layout: [3][d][paths_per_block] (blockDim is a multiple of d).
  const int ppb = blockDim.x / a.d;
  const int lp = threadIdx.x / a.d;
  const int lj = threadIdx.x % a.d;
  red[(0 * a.d + lj) * ppb + lp] = thb;
  red[(1 * a.d + lj) * ppb + lp] = mub;
  red[(2 * a.d + lj) * ppb + lp] = sgb;
  __syncthreads();
  for (int q = threadIdx.x; q < 3 * a.d; q += blockDim.x) {
    double acc = 0.0;
    for (int i = 0; i < ppb; ++i) acc += red[q * ppb + i];
    a.partial[(int64_t)blockIdx.x * 3 * a.d + q] = acc;
  }
}

// Launchers (one translation unit per precision).
int sde_launch_fwd_f32(const SdeArgs& a, hipStream_t s);
int sde_launch_fwd_f64(const SdeArgs& a, hipStream_t s);
int sde_launch_adj_f32(const SdeArgs& a, int64_t blocks, size_t lds, hipStream_t s);
int sde_launch_adj_f64(const SdeArgs& a, int64_t blocks, size_t lds, hipStream_t s);
int sde_launch_dense_f32(const SdeDenseArgs& a, hipStream_t s);
int sde_launch_dense_f64(const SdeDenseArgs& a, hipStream_t s);
constexpr int kSdeBlock = 256;

}  // namespace dfx
