// erk_mlp.hip -- the neural-ODE entry points: weight packing for the MFMA fragments, the
// vector field on its own, and the batched solve (kernel in erk_mlp.hpp, instantiated per
// policy in erk_mlp_f32.hip / erk_mlp_bf16.hip).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include "erk_mlp.hpp"

namespace dfx {

// f(y) for R independent rows (the registered MLP vector field on its own; ODETerm.vf).
template <class P>
__global__ void __launch_bounds__(MLP_THREADS, 1) mlp_eval_kernel(const float* __restrict__ Y,
                                                                  float* __restrict__ F, int64_t R,
                                                                  const MlpWeights<P> W) {
  extern __shared__ __align__(16) unsigned char smem[];
  MlpLds<P>& L = *reinterpret_cast<MlpLds<P>*>(smem);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int orow = tid >> 3, c0 = (tid & 7) * 8;
  const int64_t row = (int64_t)blockIdx.x * MLP_TILE + orow;
  const bool valid = row < R;
#pragma unroll
  for (int j = 0; j < 8; ++j) store_input<P>(L, orow, c0 + j, valid ? Y[row * MLP_D + c0 + j] : 0.0f);
  __syncthreads();
  float f[8];
  mlp_forward<P>(W, L, wave, lane, orow, c0, f);
  if (valid) {
#pragma unroll
    for (int j = 0; j < 8; ++j) F[row * MLP_D + c0 + j] = f[j];
  }
}

// Pack a row-major float32 matrix into MFMA B fragments.  forward: B[k][n] = W[n][k]
// (W is [N][K]); otherwise B[k][n] = W[k][n] (W is [K][N]).
template <class P>
__global__ void mlp_pack_kernel(const float* __restrict__ W, typename P::act_t* __restrict__ out,
                                int N, int K, int forward) {
  const int64_t total = (int64_t)N * K;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int NKS = K / P::KSTEP;
  const int j = (int)(e % P::KA);
  const int64_t fr = e / P::KA;
  const int lane = (int)(fr % 64);
  const int64_t f2 = fr / 64;
  const int ks = (int)(f2 % NKS);
  const int nt = (int)(f2 / NKS);
  const int n = nt * 32 + (lane & 31);
  const int k = ks * P::KSTEP + (lane >> 5) * P::KA + j;
  const float v = forward ? W[(int64_t)n * K + k] : W[(int64_t)k * N + n];
  out[e] = P::to_act(v);
}

// Three-piece bf16 packing (PolicyBF16x3): fragment element e of plane p holds piece p of
// split3(v); the planes are `plane` elements apart.
__global__ void mlp_pack3_kernel(const float* __restrict__ W, __bf16* __restrict__ out, int N,
                                 int K, int forward, int64_t plane) {
  using P = PolicyBF16x3;
  const int64_t total = (int64_t)N * K;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int NKS = K / P::KSTEP;
  const int j = (int)(e % P::KA);
  const int64_t fr = e / P::KA;
  const int lane = (int)(fr % 64);
  const int64_t f2 = fr / 64;
  const int ks = (int)(f2 % NKS);
  const int nt = (int)(f2 / NKS);
  const int n = nt * 32 + (lane & 31);
  const int k = ks * P::KSTEP + (lane >> 5) * P::KA + j;
  const float v = forward ? W[(int64_t)n * K + k] : W[(int64_t)k * N + n];
  split3(v, out[e], out[e + plane], out[e + 2 * plane]);
}

}  // namespace dfx

using namespace dfx;

namespace {

size_t act_bytes(int32_t precision) { return precision == DFX_BF16_F32 ? 2 : 4; }

// bf16 buffers hold the three split3 pieces of every weight, float32 buffers one copy
int pieces_of(int32_t precision) { return precision == DFX_BF16_F32 ? 3 : 1; }

int mlp_check(const dfx_mlp_problem* p) {
  if (!p) return DFX_E_ARG;
  if (p->d != MLP_D || p->hidden != MLP_H) return DFX_E_ARG;
  if (p->precision != DFX_F32 && p->precision != DFX_BF16_F32) return DFX_E_ARG;
  if (p->solver != DFX_TSIT5) return DFX_E_ARG;  // dense output implemented for Tsit5
  if (p->batch < 0 || p->n_ts < 0 || p->max_steps < 0 || p->ckpt_every < 0) return DFX_E_ARG;
  if (p->waves != 0 && p->waves != 4 && p->waves != 8) return DFX_E_ARG;
  return DFX_OK;
}

template <class P>
MlpWeights<P> weights_from(const void* packed, const float* b1, const float* b2, const float* b3) {
  return mlp_weights_from<P>(packed, P::PIECES, b1, b2, b3);
}

}  // namespace

extern "C" {

int dfx_mlp_pack_bytes(int32_t precision, size_t* bytes) {
  if (!bytes || (precision != DFX_F32 && precision != DFX_BF16_F32)) return DFX_E_ARG;
  *bytes = (size_t)mlp_pack_layout(pieces_of(precision)).total * act_bytes(precision);
  return DFX_OK;
}

int dfx_mlp_pack(int32_t precision, const float* w1, const float* w2, const float* w3, void* packed,
                 void* stream) {
  if (!w1 || !w2 || !w3 || !packed) return DFX_E_ARG;
  if (precision != DFX_F32 && precision != DFX_BF16_F32) return DFX_E_ARG;
  const MlpPackLayout lay = mlp_pack_layout(pieces_of(precision));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  struct Job { const float* W; int N, K, fwd; };
  const Job jobs[6] = {{w1, MLP_H, MLP_D, 1}, {w2, MLP_H, MLP_H, 1}, {w3, MLP_D, MLP_H, 1},
                       {w1, MLP_D, MLP_H, 0}, {w2, MLP_H, MLP_H, 0}, {w3, MLP_H, MLP_D, 0}};
  for (int i = 0; i < 6; ++i) {
    const int64_t n = (int64_t)jobs[i].N * jobs[i].K;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (precision == DFX_BF16_F32) {
      __bf16* o = reinterpret_cast<__bf16*>(packed) + lay.off[i];
      hipLaunchKernelGGL(mlp_pack3_kernel, dim3(nb), dim3(256), 0, s, jobs[i].W, o, jobs[i].N,
                         jobs[i].K, jobs[i].fwd, lay.size[i]);
    } else {
      float* o = reinterpret_cast<float*>(packed) + lay.off[i];
      hipLaunchKernelGGL(mlp_pack_kernel<PolicyF32>, dim3(nb), dim3(256), 0, s, jobs[i].W, o,
                         jobs[i].N, jobs[i].K, jobs[i].fwd);
    }
  }
  return hipGetLastError() == hipSuccess ? DFX_OK : DFX_E_HIP;
}

int dfx_mlp_eval(int32_t precision, const void* packed, const float* b1, const float* b2,
                 const float* b3, const float* y, float* f, int64_t rows, void* stream) {
  if (precision != DFX_F32 && precision != DFX_BF16_F32) return DFX_E_ARG;
  if (!packed || !b1 || !b2 || !b3 || rows < 0) return DFX_E_ARG;
  if (rows == 0) return DFX_OK;
  if (!y || !f) return DFX_E_ARG;
  const unsigned nb = (unsigned)((rows + MLP_TILE - 1) / MLP_TILE);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (precision == DFX_BF16_F32) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&mlp_eval_kernel<PolicyBF16x3>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)sizeof(MlpLds<PolicyBF16x3>));
    hipLaunchKernelGGL(mlp_eval_kernel<PolicyBF16x3>, dim3(nb), dim3(MLP_THREADS),
                       sizeof(MlpLds<PolicyBF16x3>), s, y, f, rows,
                       weights_from<PolicyBF16x3>(packed, b1, b2, b3));
  } else {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&mlp_eval_kernel<PolicyF32>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)sizeof(MlpLds<PolicyF32>));
    hipLaunchKernelGGL(mlp_eval_kernel<PolicyF32>, dim3(nb), dim3(MLP_THREADS),
                       sizeof(MlpLds<PolicyF32>), s, y, f, rows,
                       weights_from<PolicyF32>(packed, b1, b2, b3));
  }
  return hipGetLastError() == hipSuccess ? DFX_OK : DFX_E_HIP;
}

int dfx_mlp_n_ckpt(const dfx_mlp_problem* p) {
  if (!p || p->ckpt_every <= 0) return 0;
  return (p->max_steps + p->ckpt_every - 1) / p->ckpt_every;
}

int dfx_mlp_solve(const dfx_mlp_problem* p, const void* packed, const float* b1, const float* b2,
                  const float* b3, const dfx_mlp_io* io, void* stream) {
  int rc = mlp_check(p);
  if (rc != DFX_OK) return rc;
  if (!packed || !b1 || !b2 || !b3 || !io) return DFX_E_ARG;
  if (p->batch == 0) return DFX_OK;
  const int n_out = (p->save_t0 ? 1 : 0) + p->n_ts + (p->save_t1 ? 1 : 0);
  if (!io->y0 || !io->stats || !io->result) return DFX_E_ARG;
  if (n_out > 0 && (!io->ys || !io->ts_out)) return DFX_E_ARG;
  if (p->n_ts > 0 && !io->ts) return DFX_E_ARG;
  if (p->ckpt_every > 0 && (!io->grid || !io->ckpt)) return DFX_E_ARG;
  const double error_order = isnan(p->error_order) ? 5.0 : p->error_order;
  MlpArgs a;
  memset(&a, 0, sizeof(a));
  a.y0 = io->y0; a.ts = io->ts; a.ys = io->ys; a.ts_out = io->ts_out; a.stats = io->stats;
  a.result = io->result; a.grid = io->grid; a.ckpt = io->ckpt;
  a.B = p->batch; a.n_ts = p->n_ts; a.n_out = n_out; a.save_t0 = p->save_t0;
  a.save_t1 = p->save_t1; a.max_steps = p->max_steps; a.ckpt_every = p->ckpt_every;
  a.n_ckpt = dfx_mlp_n_ckpt(p);
  a.has_dt0 = !isnan(p->dt0);
  a.t0 = p->t0; a.t1 = p->t1; a.dt0 = p->dt0; a.error_order = error_order;
  // the integral controller (the default) has its own build; any p / d terms take the
  // generic one
  const bool integral = mlp_pid_params(p, a.pid);
  const unsigned nblocks = (unsigned)((p->batch + MLP_TILE - 1) / MLP_TILE);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // waves per tile (MlpWaves): the problem's choice, or by default eight for the
  // three-piece form (its 436 registers per lane at four waves leave one wave per SIMD; at
  // eight, two) and four for float32 (0.49 against 0.52 ms per solve of 8,192
  // trajectories at eight)
  const int waves = p->waves != 0 ? p->waves : (p->precision == DFX_BF16_F32 ? 8 : 4);
  const bool g = p->ckpt_every > 0;
  if (p->precision == DFX_BF16_F32)
    launch_erk_mlp_bf16(waves, a, weights_from<PolicyBF16x3>(packed, b1, b2, b3), g, integral,
                        nblocks, s);
  else
    launch_erk_mlp_f32(waves, a, weights_from<PolicyF32>(packed, b1, b2, b3), g, integral,
                       nblocks, s);
  return hipGetLastError() == hipSuccess ? DFX_OK : DFX_E_HIP;
}

}  // extern "C"
