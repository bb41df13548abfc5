// sde_heun.hip -- Heun (Stratonovich) SDE solve with diagonal noise and its discrete
// adjoint, for the registered Ornstein-Uhlenbeck terms
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
// contiguous slab); states [P, d].  Each thread owns V consecutive components of one
// path and walks all N steps, streaming its dW slice with U steps of loads in flight.
// The solve is HBM-bound: 4 bytes of dW per component-step.
//
// Gradient (RecursiveCheckpointAdjoint semantics, _adjoint.py:264-312): the forward pass
// stores the state every CK steps; the adjoint pass walks segments backwards, recomputes a
// segment from its checkpoint keeping the states and increments in registers, then runs
// the exact reverse-mode recursion of the Heun step through it.  theta/mu/sigma
// cotangents are summed per thread in float64, per block in LDS, then across blocks in a
// fixed order (deterministic).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "../../include/diffrax_amd.h"

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
};

template <class T, int V>
struct Vec {
  T v[V];
};

template <class T, int V>
__device__ __forceinline__ Vec<T, V> load_dw(const float* __restrict__ p) {
  Vec<T, V> r;
  if constexpr (V == 4) {
    const float4 x = *reinterpret_cast<const float4*>(p);
    r.v[0] = (T)x.x; r.v[1] = (T)x.y; r.v[2] = (T)x.z; r.v[3] = (T)x.w;
  } else if constexpr (V == 2) {
    const float2 x = *reinterpret_cast<const float2*>(p);
    r.v[0] = (T)x.x; r.v[1] = (T)x.y;
  } else {
    r.v[0] = (T)p[0];
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

template <class T, int V, int U>
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
  const int64_t slab = a.P * a.d;  // elements per step
  const float* dwp = a.dW + p * a.d + j0;
  T* ck = reinterpret_cast<T*>(a.ckpt);
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
#pragma unroll
        for (int v = 0; v < V; ++v) y[v] = heun_ou<T>(y[v], dt, th[v], mu[v], sg[v], w[u].v[v]);
        if (ck && ((n + 1) % a.CK) == 0 && (n + 1) < a.N) {
          T* c = ck + ((int64_t)((n + 1) / a.CK - 1) * a.P + p) * a.d + j0;
#pragma unroll
          for (int v = 0; v < V; ++v) c[v] = y[v];
        }
      }
    }
  }
  if (a.save_t1) {
    T* o = ys + (a.save_t0 ? a.d : 0);
#pragma unroll
    for (int v = 0; v < V; ++v) o[v] = y[v];
  }
}

// Reverse-mode of one Heun step for one component (see heun_ou).  Given the state y at
// the start of the step and the cotangent yb of y1, returns the cotangent of y and adds
// the parameter cotangents.
template <class T>
__device__ __forceinline__ T heun_ou_vjp(T y, T dt, T th, T mu, T sg, T w, T yb,
                                         double& thb, double& mub, double& sgb) {
  const T k0 = dt * (th * (mu - y)) + sg * w;
  const T z = y + k0;
  T k0b = T(0.5) * yb;
  const T k1b = T(0.5) * yb;
  T out = yb;
  // k1 = dt*(th*(mu - z)) + sg*w
  const T zb = -(dt * th) * k1b;
  thb += (double)(dt * (mu - z) * k1b);
  mub += (double)(dt * th * k1b);
  sgb += (double)(w * k1b);
  // z = y + k0
  out = out + zb;
  k0b = k0b + zb;
  // k0 = dt*(th*(mu - y)) + sg*w
  out = out - (dt * th) * k0b;
  thb += (double)(dt * (mu - y) * k0b);
  mub += (double)(dt * th * k0b);
  sgb += (double)(w * k0b);
  return out;
}

template <class T, int V, int CK>
__global__ void __launch_bounds__(256) sde_heun_adj_kernel(const SdeArgs a) {
  const int groups = a.d / V;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = tid < a.P * groups;
  const int64_t p = live ? tid / groups : 0;
  const int j0 = live ? (int)(tid - p * groups) * V : 0;
  double thb[V], mub[V], sgb[V];
#pragma unroll
  for (int v = 0; v < V; ++v) thb[v] = mub[v] = sgb[v] = 0.0;
  if (live) {
    const T* th_p = reinterpret_cast<const T*>(a.theta) + j0;
    const T* mu_p = reinterpret_cast<const T*>(a.mu) + j0;
    const T* sg_p = reinterpret_cast<const T*>(a.sigma) + j0;
    T th[V], mu[V], sg[V], yb[V];
#pragma unroll
    for (int v = 0; v < V; ++v) { th[v] = th_p[v]; mu[v] = mu_p[v]; sg[v] = sg_p[v]; }
    const T* ysb = reinterpret_cast<const T*>(a.ys_bar) + p * (int64_t)a.n_out * a.d + j0;
#pragma unroll
    for (int v = 0; v < V; ++v) yb[v] = a.save_t1 ? ysb[(a.save_t0 ? a.d : 0) + v] : T(0);
    const int64_t slab = a.P * a.d;
    const float* dwp = a.dW + p * a.d + j0;
    const T* ck = reinterpret_cast<const T*>(a.ckpt);
    const T* y0 = reinterpret_cast<const T*>(a.y0) + p * a.d + j0;
    const int nseg = (a.N + CK - 1) / CK;
    for (int s = nseg - 1; s >= 0; --s) {
      const int nbeg = s * CK;
      T yh[CK][V];
      Vec<T, V> w[CK];
      T y[V];
      const T* src = (s == 0) ? y0 : ck + ((int64_t)(s - 1) * a.P + p) * a.d + j0;
#pragma unroll
      for (int v = 0; v < V; ++v) y[v] = src[v];
#pragma unroll
      for (int c = 0; c < CK; ++c) {
        if (nbeg + c < a.N) w[c] = load_dw<T, V>(dwp + (int64_t)(nbeg + c) * slab);
      }
      // recompute the segment, keeping the state at the start of every step
#pragma unroll
      for (int c = 0; c < CK; ++c) {
        const int n = nbeg + c;
#pragma unroll
        for (int v = 0; v < V; ++v) yh[c][v] = y[v];
        if (n < a.N) {
          const T dt = (T)(a.grid[n + 1] - a.grid[n]);
#pragma unroll
          for (int v = 0; v < V; ++v) y[v] = heun_ou<T>(y[v], dt, th[v], mu[v], sg[v], w[c].v[v]);
        }
      }
      // reverse sweep
#pragma unroll
      for (int c = CK - 1; c >= 0; --c) {
        const int n = nbeg + c;
        if (n < a.N) {
          const T dt = (T)(a.grid[n + 1] - a.grid[n]);
#pragma unroll
          for (int v = 0; v < V; ++v)
            yb[v] = heun_ou_vjp<T>(yh[c][v], dt, th[v], mu[v], sg[v], w[c].v[v], yb[v], thb[v],
                                   mub[v], sgb[v]);
        }
      }
    }
    if (a.save_t0) {
#pragma unroll
      for (int v = 0; v < V; ++v) yb[v] = yb[v] + ysb[v];
    }
    if (a.y0_bar) {
      T* o = reinterpret_cast<T*>(a.y0_bar) + p * a.d + j0;
#pragma unroll
      for (int v = 0; v < V; ++v) o[v] = yb[v];
    }
  }
  // Block reduction of the parameter cotangents over the paths of this block.
  // LDS layout: [3][d][paths_per_block]; every thread writes its V components.
  extern __shared__ double red[];
  const int ppb = blockDim.x / groups;  // paths per block (blockDim is a multiple of groups)
  const int lp = threadIdx.x / groups;
  const int lj = (threadIdx.x % groups) * V;
#pragma unroll
  for (int v = 0; v < V; ++v) {
    red[(0 * a.d + lj + v) * ppb + lp] = thb[v];
    red[(1 * a.d + lj + v) * ppb + lp] = mub[v];
    red[(2 * a.d + lj + v) * ppb + lp] = sgb[v];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < 3 * a.d; q += blockDim.x) {
    double acc = 0.0;
    for (int i = 0; i < ppb; ++i) acc += red[q * ppb + i];
    a.partial[(int64_t)blockIdx.x * 3 * a.d + q] = acc;
  }
}

// Sum the per-block partials in block order (one thread per output, fixed order) and
// write theta/mu/sigma cotangents to their own buffers.
__global__ void sde_reduce_kernel(const double* __restrict__ partial, int nblocks, int d,
                                  double* __restrict__ thb, double* __restrict__ mub,
                                  double* __restrict__ sgb) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int nout = 3 * d;
  if (q >= nout) return;
  double acc = 0.0;
  for (int b = 0; b < nblocks; ++b) acc += partial[(int64_t)b * nout + q];
  double* dst = (q < d) ? thb : (q < 2 * d ? mub : sgb);
  dst[q % d] = acc;
}

}  // namespace dfx

using namespace dfx;

namespace {
constexpr int kBlock = 256;

int sde_check(const dfx_sde_problem* p) {
  if (!p) return DFX_E_ARG;
  if (p->precision != DFX_F64 && p->precision != DFX_F32) return DFX_E_ARG;
  if (p->d <= 0 || p->d > 256 || p->paths < 0 || p->num_steps < 0) return DFX_E_ARG;
  const int V = (p->precision == DFX_F32) ? 4 : 2;
  if (p->d % V != 0) return DFX_E_ARG;
  if (kBlock % (p->d / V) != 0) return DFX_E_ARG;
  return DFX_OK;
}

int64_t sde_threads(const dfx_sde_problem* p, int V) { return p->paths * (p->d / V); }
}  // namespace

extern "C" {

int dfx_sde_n_ckpt(const dfx_sde_problem* p) {
  if (!p || p->checkpoint_every <= 0) return 0;
  return (p->num_steps + p->checkpoint_every - 1) / p->checkpoint_every - 1;
}

int dfx_sde_workspace_bytes(const dfx_sde_problem* p, size_t* bytes) {
  if (sde_check(p) != DFX_OK || !bytes) return DFX_E_ARG;
  const int Va = (p->precision == DFX_F32) ? 2 : 1;
  const int64_t nblocks = (sde_threads(p, Va) + kBlock - 1) / kBlock;
  *bytes = (size_t)(nblocks * 3 * p->d) * sizeof(double) + 256;
  return DFX_OK;
}

int dfx_sde_heun_solve(const dfx_sde_problem* p, const dfx_sde_io* io, void* stream) {
  int rc = sde_check(p);
  if (rc != DFX_OK) return rc;
  if (!io || !io->y0 || !io->theta || !io->mu || !io->sigma || !io->grid) return DFX_E_ARG;
  if (p->num_steps > 0 && !io->dW) return DFX_E_ARG;
  if ((p->save_t0 || p->save_t1) && !io->ys) return DFX_E_ARG;
  if (io->ckpt && p->checkpoint_every <= 0) return DFX_E_ARG;
  if (p->paths == 0) return DFX_OK;
  SdeArgs a;
  memset(&a, 0, sizeof(a));
  a.y0 = io->y0; a.theta = io->theta; a.mu = io->mu; a.sigma = io->sigma;
  a.dW = io->dW; a.grid = io->grid; a.ys = io->ys; a.ckpt = io->ckpt;
  a.P = p->paths; a.d = p->d; a.N = p->num_steps; a.CK = p->checkpoint_every > 0 ? p->checkpoint_every : 1;
  a.n_out = (p->save_t0 ? 1 : 0) + (p->save_t1 ? 1 : 0);
  a.save_t0 = p->save_t0; a.save_t1 = p->save_t1;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (p->precision == DFX_F32) {
    const int64_t nb = (sde_threads(p, 4) + kBlock - 1) / kBlock;
    hipLaunchKernelGGL((sde_heun_fwd_kernel<float, 4, 8>), dim3((unsigned)nb), dim3(kBlock), 0, s, a);
  } else {
    const int64_t nb = (sde_threads(p, 2) + kBlock - 1) / kBlock;
    hipLaunchKernelGGL((sde_heun_fwd_kernel<double, 2, 8>), dim3((unsigned)nb), dim3(kBlock), 0, s, a);
  }
  return hipGetLastError() == hipSuccess ? DFX_OK : DFX_E_HIP;
}

int dfx_sde_heun_adjoint(const dfx_sde_problem* p, const dfx_sde_io* io,
                         const dfx_sde_grad_io* g, void* workspace, size_t workspace_bytes,
                         void* stream) {
  int rc = sde_check(p);
  if (rc != DFX_OK) return rc;
  if (!io || !g || !io->y0 || !io->theta || !io->mu || !io->sigma || !io->grid || !g->ys_bar)
    return DFX_E_ARG;
  if (!g->theta_bar || !g->mu_bar || !g->sigma_bar) return DFX_E_ARG;
  if (p->checkpoint_every != 32 && p->checkpoint_every != 16 && p->checkpoint_every != 8)
    return DFX_E_ARG;
  if (dfx_sde_n_ckpt(p) > 0 && !io->ckpt) return DFX_E_ARG;
  size_t need = 0;
  dfx_sde_workspace_bytes(p, &need);
  if (!workspace || workspace_bytes < need) return DFX_E_WORKSPACE;
  SdeArgs a;
  memset(&a, 0, sizeof(a));
  a.y0 = io->y0; a.theta = io->theta; a.mu = io->mu; a.sigma = io->sigma;
  a.dW = io->dW; a.grid = io->grid; a.ckpt = io->ckpt;
  a.ys_bar = g->ys_bar; a.y0_bar = g->y0_bar;
  a.P = p->paths; a.d = p->d; a.N = p->num_steps; a.CK = p->checkpoint_every;
  a.n_out = (p->save_t0 ? 1 : 0) + (p->save_t1 ? 1 : 0);
  a.save_t0 = p->save_t0; a.save_t1 = p->save_t1;
  const int Va = (p->precision == DFX_F32) ? 2 : 1;
  const int64_t nb = (sde_threads(p, Va) + kBlock - 1) / kBlock;
  a.partial = reinterpret_cast<double*>(workspace);
  const size_t lds = (size_t)3 * p->d * (kBlock / (p->d / Va)) * sizeof(double);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (p->paths > 0) {
#define DFX_SDE_ADJ(T, V, CKV) \
  hipLaunchKernelGGL((sde_heun_adj_kernel<T, V, CKV>), dim3((unsigned)nb), dim3(kBlock), lds, s, a)
    if (p->precision == DFX_F32) {
      if (a.CK == 32) DFX_SDE_ADJ(float, 2, 32);
      else if (a.CK == 16) DFX_SDE_ADJ(float, 2, 16);
      else DFX_SDE_ADJ(float, 2, 8);
    } else {
      if (a.CK == 32) DFX_SDE_ADJ(double, 1, 32);
      else if (a.CK == 16) DFX_SDE_ADJ(double, 1, 16);
      else DFX_SDE_ADJ(double, 1, 8);
    }
#undef DFX_SDE_ADJ
    const int nout = 3 * p->d;
    hipLaunchKernelGGL(sde_reduce_kernel, dim3((nout + 63) / 64), dim3(64), 0, s, a.partial,
                       (int)nb, p->d, g->theta_bar, g->mu_bar, g->sigma_bar);
  }
  return hipGetLastError() == hipSuccess ? DFX_OK : DFX_E_HIP;
}

}  // extern "C"
