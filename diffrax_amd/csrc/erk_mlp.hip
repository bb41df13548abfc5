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

// Dense output of the neural ODE at query times (erk_dense.hpp's evaluation for d = 64):
// 64 threads per (trajectory, query), one per component; the step holding t, the
// interpolant's weights and the bounds rule as there, and each component's value with
// the arithmetic of the solve kernel's SaveAt(ts) interpolation, so evaluating at a saved
// time returns that saved row bit for bit.
struct MlpDenseArgs {
  const double* dense_ts;
  const float* dense_ys;
  const float* dense_k;
  const int32_t* stats;
  const double* tq;
  float* out;
  int64_t B, n_q, tq_stride;
  int32_t max_steps, left, derivative;
  double t0, t1;
};

__global__ void __launch_bounds__(256) mlp_dense_eval_kernel(const MlpDenseArgs a) {
  constexpr int S = Tsit5Tab::S;
  const int64_t gid = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int j = threadIdx.x & 63;
  if (gid >= a.B * a.n_q) return;
  const int64_t b = gid / a.n_q;
  const int64_t q = gid - b * a.n_q;
  const int dir = (a.t0 < a.t1) ? 1 : -1;
  const double t = a.tq[b * a.tq_stride + q] * dir;
  const int n = a.stats[b * 3 + 1];
  const int64_t row0 = b * ((int64_t)a.max_steps + 1);
  const double* ts = a.dense_ts + row0;
  const float* ys = a.dense_ys + row0 * MLP_D;
  float* out = a.out + gid * MLP_D;
  const double ts_first = ts[0], ts_last = ts[n];
  if (n < 1 || t < ts_first || t > ts_last || t != t) {
    const bool trivial = (n == 0) && (t == a.t0) && !a.derivative;
    out[j] = trivial ? ys[j] : __builtin_nanf("");
    return;
  }
  int lo = 0, hi = n + 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const bool before = a.left ? (ts[mid] < t) : (ts[mid] <= t);
    if (before) lo = mid + 1; else hi = mid;
  }
  int idx = lo - 1;
  idx = idx < 0 ? 0 : (idx > n - 1 ? n - 1 : idx);
  const double ts0 = ts[idx], ts1 = ts[idx + 1];
  const float* kp = a.dense_k + (b * (int64_t)a.max_steps + idx) * S * MLP_D + j;
  const double th = linear_rescale(ts0, t, ts1);
  if (a.derivative) {
    double db[7];
    tsit5_dweights(th, db);
    const double inv_h = (ts0 == ts1) ? 0.0 : 1.0 / (ts1 - ts0);
    float acc = (float)db[0] * kp[0];
#pragma unroll
    for (int m = 1; m < 7; ++m) acc = acc + (float)db[m] * kp[m * MLP_D];
    out[j] = (float)dir * (float)((double)acc * inv_h);
  } else {
    double bb[7];
    tsit5_weights(th, bb);
    float acc = (float)bb[0] * kp[0];
#pragma unroll
    for (int m = 1; m < 7; ++m) acc = acc + (float)bb[m] * kp[m * MLP_D];
    out[j] = ys[(int64_t)idx * MLP_D + j] + acc;
  }
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
  if (p->ts_stride != 0 && p->ts_stride != p->n_ts) return DFX_E_ARG;
  if (p->n_events < 0 || p->n_events > kMaxEvents) return DFX_E_ARG;
  if (p->n_step_ts < 0 || p->n_jump_ts < 0) return DFX_E_ARG;
  if (p->resume & ~(DFX_RESUME_SOLVER | DFX_RESUME_CONTROLLER | DFX_RESUME_MADE_JUMP))
    return DFX_E_ARG;
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

int dfx_mlp_n_out(const dfx_mlp_problem* p) {
  if (!p || p->save_steps < 0 || p->max_steps < 0) return DFX_E_ARG;
  int n = (p->save_t0 ? 1 : 0) + p->n_ts;
  if (p->save_steps > 0) n += p->max_steps / p->save_steps;
  if (p->save_t1 && (p->save_steps <= 0 || p->max_steps % p->save_steps != 0)) n += 1;
  return n;
}

int dfx_mlp_solve(const dfx_mlp_problem* p, const void* packed, const float* b1, const float* b2,
                  const float* b3, const dfx_mlp_io* io, void* stream) {
  int rc = mlp_check(p);
  if (rc != DFX_OK) return rc;
  if (!packed || !b1 || !b2 || !b3 || !io) return DFX_E_ARG;
  if (p->save_steps < 0) return DFX_E_ARG;
  const int dense = (io->dense_ts != nullptr) + (io->dense_ys != nullptr) + (io->dense_k != nullptr);
  if (dense != 0 && dense != 3) return DFX_E_ARG;
  // the adjoints replay the ts saves only, without events
  if (p->ckpt_every > 0 && (dense || p->save_steps > 0 || p->n_events > 0)) return DFX_E_ARG;
  if ((p->n_step_ts > 0 && !io->step_ts) || (p->n_jump_ts > 0 && !io->jump_ts)) return DFX_E_ARG;
  // resume: state_in given; forward only; the clipped controller state is not carried
  if (p->resume && (!io->state_in || p->ckpt_every > 0 || p->n_step_ts > 0 || p->n_jump_ts > 0))
    return DFX_E_ARG;
  if (io->state_out && p->ckpt_every > 0) return DFX_E_ARG;
  if (p->n_events > 0 && (!io->event_coeffs || !io->event_flags || !io->event_mask))
    return DFX_E_ARG;
  if (p->batch == 0) return DFX_OK;
  const int n_out = dfx_mlp_n_out(p);
  if (!io->y0 || !io->stats || !io->result) return DFX_E_ARG;
  if (n_out > 0 && (!io->ys || !io->ts_out)) return DFX_E_ARG;
  if (p->n_ts > 0 && !io->ts) return DFX_E_ARG;
  if (p->ckpt_every > 0 && (!io->grid || !io->ckpt)) return DFX_E_ARG;
  const double error_order = isnan(p->error_order) ? 5.0 : p->error_order;
  MlpArgs a;
  memset(&a, 0, sizeof(a));
  a.y0 = io->y0; a.ts = io->ts; a.ys = io->ys; a.ts_out = io->ts_out; a.stats = io->stats;
  a.result = io->result; a.grid = io->grid; a.ckpt = io->ckpt;
  a.dense_ts = io->dense_ts; a.dense_ys = io->dense_ys; a.dense_k = io->dense_k;
  a.save_steps = p->save_steps;
  a.ts_stride = p->ts_stride;
  a.n_events = p->n_events;
  a.event_root_find = p->event_root_find ? 1 : 0;
  a.event_coeffs = io->event_coeffs;
  a.event_flags = io->event_flags;
  a.event_mask = io->event_mask;
  a.n_step_ts = p->n_step_ts;
  a.n_jump_ts = p->n_jump_ts;
  a.step_ts = io->step_ts;
  a.jump_ts = io->jump_ts;
  a.resume = p->resume;
  a.state_in = io->state_in;
  a.state_out = io->state_out;
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

int dfx_mlp_dense_eval(const dfx_mlp_problem* p, const dfx_erk_dense_io* io, void* stream) {
  if (!p || !io) return DFX_E_ARG;
  if (p->d != MLP_D || p->solver != DFX_TSIT5 || p->batch < 0 || p->max_steps < 0)
    return DFX_E_ARG;
  if (io->n_q < 0 || io->tq_stride < 0) return DFX_E_ARG;
  if (io->tq_stride != 0 && io->tq_stride < io->n_q) return DFX_E_ARG;
  if (p->batch == 0 || io->n_q == 0) return DFX_OK;
  if (!io->dense_ts || !io->dense_ys || !io->dense_k || !io->stats || !io->tq || !io->out)
    return DFX_E_ARG;
  MlpDenseArgs a;
  memset(&a, 0, sizeof(a));
  a.dense_ts = io->dense_ts;
  a.dense_ys = reinterpret_cast<const float*>(io->dense_ys);
  a.dense_k = reinterpret_cast<const float*>(io->dense_k);
  a.stats = io->stats;
  a.tq = io->tq;
  a.out = reinterpret_cast<float*>(io->out);
  a.B = p->batch;
  a.n_q = io->n_q;
  a.tq_stride = io->tq_stride;
  a.max_steps = p->max_steps;
  a.left = io->left ? 1 : 0;
  a.derivative = io->derivative ? 1 : 0;
  a.t0 = p->t0;
  a.t1 = p->t1;
  const int64_t groups = a.B * a.n_q;
  const unsigned nb = (unsigned)((groups + 3) / 4);
  hipLaunchKernelGGL(mlp_dense_eval_kernel, dim3(nb), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), a);
  return hipGetLastError() == hipSuccess ? DFX_OK : DFX_E_HIP;
}

}  // extern "C"
